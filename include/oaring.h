/*
 * oaring.h -- C ABI of the learned ringdown extrapolator (gfx950): a dilated causal 1-D
 * convolution network that reads the first K steps of a run's port voltages and predicts the
 * rest of each port's waveform as 16 complex pole-residue pairs, trained on the device by the
 * library's own forward and backward kernels (BF16 operands, FP32 accumulation on MFMA).
 *
 * The reference has no counterpart: its runs stop on the energy criterion or at max_steps, with
 * the matrix-pencil tail (packages/oa-fdtd/src/oa_fdtd/tail.py:39-98) as the only extrapolation.
 * The network plays the role of that tail fit, earlier in the run.
 *
 * Layout (floats, row-major): Win 64x32, bin 64, then six layers (W 64x128 = [tap0 | tap1]
 * input channels, b 64), W1 256x512, b1 256, W2 256x256, b2 256; oaring_param_count() values.
 * A batch: U (32, samples * T) input channels (channel-major, sample s at columns s*T .. s*T+T-1),
 * amp (samples) amplitude scale of each sample's residues, kap and tgt (samples, 4, F, 2) the
 * complex factor from a port-voltage DFT to S-parameter units and the target tail in those
 * units, t0 (samples) the time of the first tail sample [s], omega (F) angular frequencies,
 * dt [s], sig0 [1/s] decay-rate scale, wmax [rad/s] largest pole frequency.
 * All functions return OARING_OK (0) or a status code; oaring_last_error() gives the message.
 */
#ifndef OARING_H
#define OARING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OARING_OK 0
#define OARING_E_INVALID 1
#define OARING_E_OOM 2
#define OARING_E_DEVICE 3
#define OARING_E_NO_DEVICE 5

typedef struct oaring_net oaring_net;

int64_t oaring_param_count(void);
int oaring_create(int samples, int time_len, int n_freq, oaring_net** out);
int oaring_set_params(oaring_net* e, const float* params, int64_t n);
int oaring_get_params(oaring_net* e, float* params, int64_t n);
/* gradients of the mean loss after oaring_backward */
int oaring_get_grads(oaring_net* e, float* grads, int64_t n);
int oaring_reset_adam(oaring_net* e);
int oaring_set_batch(oaring_net* e, const float* U, const float* amp, const float* kap, const float* tgt,
                     const float* t0, const double* omega, float dt, float sig0, float wmax);
/* loss = mean over (sample, port, frequency) of |kap * P - tgt|^2; pred (samples, 4, F, 2) and
 * raw (256, samples) network outputs are optional (NULL) */
int oaring_forward(oaring_net* e, float* loss, float* pred, float* raw);
int oaring_backward(oaring_net* e);
int oaring_adam(oaring_net* e, float lr, float beta1, float beta2, float eps, int step);
/* C = A B (row-major, M x K times K x N) through the network's BF16 MFMA product kernel */
int oaring_gemm_check(int M, int N, int K, const float* A, const float* B, float* C);
const char* oaring_last_error(void);
void oaring_destroy(oaring_net* e);

#ifdef __cplusplus
}
#endif

#endif
