"""Sub-cell models: copper-box edge classification and the surface-impedance fit."""

from .edges import ConductorEdges, classify_box, outside_samples
from .sibc import SibcFit, fit_surface_impedance, hammerstad_jensen, slab_surface_impedance

__all__ = [
    "ConductorEdges",
    "SibcFit",
    "classify_box",
    "fit_surface_impedance",
    "hammerstad_jensen",
    "outside_samples",
    "slab_surface_impedance",
]
