"""Learned ringdown extrapolator: a dilated causal convolution network, trained on the GPU by
hand-written HIP kernels, that predicts the rest of a run's port waveforms as pole-residue terms
from its first K steps, with a residual-energy check and the full-length run as fallback."""

from .net import NPARAM, BatchArrays, RingNet, gemm_check, init_params

__all__ = ["NPARAM", "BatchArrays", "RingNet", "gemm_check", "init_params"]
