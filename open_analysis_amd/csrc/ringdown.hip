// Learned ringdown extrapolator of the MI355X oa-fdtd engine: training and inference kernels.
//
// A dilated causal 1-D convolution network reads the first K steps of a run's port voltages
// (decimated) plus a geometry descriptor and predicts, per port, 16 complex pole-residue pairs
// of the waveform that follows. Its loss is the S-parameter error of the predicted tail's DFT.
//
// Network (all FP32 master parameters; GEMM operands rounded to BF16, FP32 accumulation on
// MFMA v_mfma_f32_16x16x32_bf16):
//   X0 = tanh(Win U + bin)                     U: 32 input channels x T, Win 64 x 32
//   Xl = X(l-1) + tanh(Wl [X(l-1)(t); X(l-1)(t - d_l)] + bl),  l = 1..6, d_l = 2^(l-1), Wl 64 x 128
//   F  = the last 8 time positions of X6 (512 features per sample)
//   H  = tanh(W1 F + b1)                       W1 256 x 512
//   O  = W2 H + b2                             W2 256 x 256: per port 16 x (u, v, p, q)
// Activations are channel-major: X[c][s * T + t] for sample s, time t.
//
// Every matrix product goes through one tiled MFMA kernel (gemm_bf16) whose operand loaders and
// epilogue are small functors, so the convolution taps, the transposed weights of the backward
// pass and the tanh derivative are read in place (implicit GEMM, no im2col buffers). Products
// with a long reduction (weight gradients, K = samples x time) are split over the grid's z
// dimension into partial tiles that a second kernel sums in a fixed order: results are
// bitwise reproducible.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/oaring.h"

namespace ring {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, PADK = 8;

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }

// ---- operand loaders: value of element (row, col) of the logical operand, 0 outside.
struct Dense {  // p[r * ldr + c * ldc]
    const float* p;
    long long ldr, ldc;
    int rows, cols;
    __device__ float operator()(int r, int c) const {
        return (r < rows && c < cols) ? p[r * ldr + c * ldc] : 0.f;
    }
};

// Causal two-tap convolution input as a (128 x N) operand: row k = tap * 64 + channel,
// column n = s * T + t; tap 0 reads X(t), tap 1 reads X(t - d) (0 before the start).
struct ConvIn {
    const float* x;  // [64][N]
    int T, d, N;
    __device__ float operator()(int k, int n) const {
        if (n >= N || k >= 128) return 0.f;
        const int c = k & 63, tap = k >> 6;
        const int t = n % T;
        if (tap && t < d) return 0.f;
        return x[(long long)c * N + n - tap * d];
    }
};
// The same operand transposed (N x 128), for the weight gradient.
struct ConvInT {
    ConvIn a;
    __device__ float operator()(int n, int k) const { return a(k, n); }
};

// tanh derivative in place: dZ = dY * (1 - (Y - R)^2), R the residual input (nullptr: none)
struct DZ {
    const float* dy;
    const float* y;
    const float* r;
    int rows, N;
    __device__ float at(int c, int n) const {
        const long long q = (long long)c * N + n;
        const float th = r ? y[q] - r[q] : y[q];
        return dy[q] * (1.f - th * th);
    }
    __device__ float operator()(int c, int n) const { return (c < rows && n < N) ? at(c, n) : 0.f; }
};
struct DZT {  // (N x rows) view
    DZ a;
    __device__ float operator()(int n, int c) const { return a(c, n); }
};
// Backward-data operand of a two-tap convolution (128 x N): row k = tap * 64 + out channel,
// tap 1 reads dZ(t + d) (0 past the end).
struct ConvDZ {
    DZ a;
    int T, d;
    __device__ float operator()(int k, int n) const {
        if (n >= a.N || k >= 128) return 0.f;
        const int c = k & 63, tap = k >> 6;
        if (tap) {
            const int t = n % T;
            if (t + d >= T) return 0.f;
            return a.at(c, n + d);
        }
        return a.at(c, n);
    }
};
// Transposed per-tap weights (64 x 128): element (ci, k) = W[k % 64][(k / 64) * 64 + ci]
struct WTaps {
    const float* w;  // [64][128]
    __device__ float operator()(int ci, int k) const {
        return (ci < 64 && k < 128) ? w[(k & 63) * 128 + (k >> 6) * 64 + ci] : 0.f;
    }
};
// Head features (512 x S): row k = j * 64 + c reads X6[c][s * T + T - 8 + j]
struct Feat {
    const float* x;
    int T, S;
    __device__ float operator()(int k, int s) const {
        if (k >= 512 || s >= S) return 0.f;
        return x[(long long)(k & 63) * ((long long)S * T) + (long long)s * T + T - 8 + (k >> 6)];
    }
};
struct FeatT {
    Feat a;
    __device__ float operator()(int s, int k) const { return a(k, s); }
};

// ---- epilogues
struct StoreAct {  // Y = act(acc + b) (+ R): act 0 = identity, 1 = tanh
    float* y;
    const float* b;
    const float* r;
    int act, rows, cols;
    long long ld;
    __device__ void operator()(int m, int n, float acc) const {
        if (m >= rows || n >= cols) return;
        float v = acc + (b ? b[m] : 0.f);
        if (act == 1) v = tanhf(v);
        const long long q = (long long)m * ld + n;
        if (r) v += r[q];
        y[q] = v;
    }
};
struct StorePartial {  // split-K partial tile: P[z][m][n]
    float* p;
    int rows, cols;
    __device__ void operator()(int m, int n, float acc) const {
        if (m >= rows || n >= cols) return;
        p[((long long)blockIdx.z * rows + m) * cols + n] = acc;
    }
};
struct ScatterFeat {  // dF (512 x S) into dX6[c][s * T + T - 8 + j]
    float* dx;
    int T, S;
    __device__ void operator()(int k, int s, float acc) const {
        if (k >= 512 || s >= S) return;
        dx[(long long)(k & 63) * ((long long)S * T) + (long long)s * T + T - 8 + (k >> 6)] += acc;
    }
};

// C (M x N) = A (M x K) . B (K x N); this block covers K range [k0, k1) of split blockIdx.z.
// 256 threads = 4 waves; wave w owns the 32 x 32 sub-tile (w / 2, w % 2) as 2 x 2 MFMA tiles.
template <typename LA, typename LB, typename EP>
__global__ __launch_bounds__(256) void gemm_bf16(LA la, LB lb, EP ep, int M, int N, int K, int kper) {
    __shared__ __bf16 As[BM][BK + PADK];
    __shared__ __bf16 Bt[BN][BK + PADK];  // B stored column-major: Bt[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kbeg = blockIdx.z * kper, kend = min(K, kbeg + kper);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // loader: row / column lr, 8 consecutive k
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + lk + e;
            const bool kin = k < kend;
            As[lr][lk + e] = to_bf16(kin ? la(m0 + lr, k) : 0.f);
            Bt[lr][lk + e] = to_bf16(kin ? lb(k, n0 + lr) : 0.f);
        }
        __syncthreads();
        const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
        const int kq = (lane >> 4) * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            bf16x8 a;
#pragma unroll
            for (int e = 0; e < 8; ++e) a[e] = As[wr + i * 16 + (lane & 15)][kq + e];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                bf16x8 b;
#pragma unroll
                for (int e = 0; e < 8; ++e) b[e] = Bt[wc + j * 16 + (lane & 15)][kq + e];
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                ep(m0 + wr + i * 16 + (lane >> 4) * 4 + r, n0 + wc + j * 16 + (lane & 15), acc[i][j][r]);
}

// out[m][n] (=|+=) sum_z P[z][m][n] in z order
__global__ __launch_bounds__(256) void sum_splits(const float* p, int splits, long long mn, float* out, int add) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= mn) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += p[(long long)z * mn + q];
    out[q] = add ? out[q] + s : s;
}

// db[c] = sum_n dZ(c, n): one workgroup per row, fixed-order tree
__global__ __launch_bounds__(256) void row_sums(DZ dz, float* db) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int n = threadIdx.x; n < dz.N; n += 256) s += dz.at(c, n);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}
// plain row sums of a dense [rows][N] array
__global__ __launch_bounds__(256) void row_sums_plain(const float* x, int N, float* out) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int n = threadIdx.x; n < N; n += 256) s += x[(long long)c * N + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void fill(float* p, long long n, float v) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) p[q] = v;
}
__global__ void copy_f32(const float* s, float* d, long long n) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) d[q] = s[q];
}

// ---- pole-residue tail layer and loss. Raw outputs O[r][s], r = port * 64 + pole * 4 + (u, v,
// p, q): sigma = -softplus(u) * sig0, omega = w_max * sigmoid(v), c = (p + j q) * amp[s]. The
// predicted tail DFT of port i is
//   P_i(f) = sum_m c g(s_m, f) + conj(c) g(conj(s_m), f),  g(s, f) = e^{-j w t0} dt / (1 - e^{(s - j w) dt})
// and the loss is mean over (s, i, f) of |kappa P_i(f) - target|^2 (kappa maps a port voltage DFT
// to S-parameter units). One workgroup per (sample, port); threads over frequencies.
struct PoleArgs {
    const float* O;   // [256][S]
    const float* amp; // [S]
    const double* omega;  // [F]
    const float* kap;     // [S][4][F][2]
    const float* tgt;     // [S][4][F][2]
    const float* t0;      // [S] first tail sample time
    float dt, sig0, wmax;
    int S, F;
    float* dO;     // [256][S] gradient (written)
    float* loss;   // [S * 4] partial losses
    float* pred;   // [S][4][F][2] predicted kappa P (optional)
    double scale;  // d(mean loss)/d(sum): 1 / (S * 4 * F)
};

// g(s, f) and its s-derivative g' = g dt z / (1 - z), z = e^{(s - j w) dt}, in FP64: |z| is
// within 1e-4 of 1 for slowly decaying poles, where 1 - z loses four digits.
struct GPair {
    double gr, gi, hr, hi;
};
__device__ __forceinline__ GPair gfun(double sre, double sim, double dt, double wd, double er, double ei) {
    const double ea = exp(sre * dt);
    const double ang = sim * dt - wd;
    double sn, cs;
    sincos(ang, &sn, &cs);
    const double zr = ea * cs, zi = ea * sn;
    const double dr = 1.0 - zr, di = -zi;
    const double inv = 1.0 / (dr * dr + di * di);
    const double qr = dr * inv * dt, qi = -di * inv * dt;  // dt / (1 - z)
    GPair o;
    o.gr = er * qr - ei * qi;
    o.gi = er * qi + ei * qr;
    const double tr = (zr * dr + zi * di) * inv * dt, ti = (zi * dr - zr * di) * inv * dt;  // dt z / (1 - z)
    o.hr = o.gr * tr - o.gi * ti;
    o.hi = o.gr * ti + o.gi * tr;
    return o;
}

__global__ __launch_bounds__(256) void pole_loss(PoleArgs A) {
    constexpr int NP = 16;
    const int s = blockIdx.x, port = blockIdx.y;
    const int tid = threadIdx.x;
    __shared__ double sr[NP], si[NP], cr[NP], ci[NP];
    __shared__ double red[4][256];
    if (tid < NP) {
        const int r = port * 64 + tid * 4;
        const double u = A.O[(long long)(r + 0) * A.S + s], v = A.O[(long long)(r + 1) * A.S + s];
        const double p = A.O[(long long)(r + 2) * A.S + s], q = A.O[(long long)(r + 3) * A.S + s];
        sr[tid] = -(u > 30.0 ? u : log1p(exp(u))) * (double)A.sig0;
        si[tid] = (double)A.wmax / (1.0 + exp(-v));
        cr[tid] = p * (double)A.amp[s];
        ci[tid] = q * (double)A.amp[s];
    }
    __syncthreads();
    const int f = tid;
    const bool on = f < A.F;
    const double dt = A.dt;
    double er = 0.0, ei = 0.0, wd = 0.0, e_r = 0.0, e_i = 0.0, lsum = 0.0;
    if (on) {
        const double w = A.omega[f];
        sincos(-w * (double)A.t0[s], &ei, &er);
        wd = w * dt;
        double Pr = 0.0, Pi = 0.0;
        for (int m = 0; m < NP; ++m) {
            const GPair a = gfun(sr[m], si[m], dt, wd, er, ei);   // c g(s)
            const GPair b = gfun(sr[m], -si[m], dt, wd, er, ei);  // conj(c) g(conj s)
            Pr += cr[m] * a.gr - ci[m] * a.gi + cr[m] * b.gr + ci[m] * b.gi;
            Pi += cr[m] * a.gi + ci[m] * a.gr + cr[m] * b.gi - ci[m] * b.gr;
        }
        const long long kq = (((long long)s * 4 + port) * A.F + f) * 2;
        const double kr = A.kap[kq], ki = A.kap[kq + 1];
        const double yr = kr * Pr - ki * Pi, yi = kr * Pi + ki * Pr;
        const double rr = yr - A.tgt[kq], ri = yi - A.tgt[kq + 1];
        if (A.pred) {
            A.pred[kq] = (float)yr;
            A.pred[kq + 1] = (float)yi;
        }
        lsum = rr * rr + ri * ri;
        // L = |kappa P - t|^2: dL/dtheta = Re(conj(e) dP/dtheta) with e = 2 conj(kappa) (kappa P - t)
        e_r = 2.0 * (kr * rr + ki * ri);
        e_i = 2.0 * (kr * ri - ki * rr);
    }
    for (int m = 0; m < NP; ++m) {
        double v4[4] = {0.0, 0.0, 0.0, 0.0};
        if (on) {
            const GPair a = gfun(sr[m], si[m], dt, wd, er, ei);
            const GPair b = gfun(sr[m], -si[m], dt, wd, er, ei);
            // dP/dsigma = c g'(s) + conj(c) g'(conj s) ; dP/domega = j c g'(s) - j conj(c) g'(conj s)
            const double a0r = cr[m] * a.hr - ci[m] * a.hi, a0i = cr[m] * a.hi + ci[m] * a.hr;
            const double a1r = cr[m] * b.hr + ci[m] * b.hi, a1i = cr[m] * b.hi - ci[m] * b.hr;
            // dP/dc_re = g(s) + g(conj s) ; dP/dc_im = j g(s) - j g(conj s)
            v4[0] = e_r * (a0r + a1r) + e_i * (a0i + a1i);
            v4[1] = e_r * (-a0i + a1i) + e_i * (a0r - a1r);
            v4[2] = e_r * (a.gr + b.gr) + e_i * (a.gi + b.gi);
            v4[3] = e_r * (-a.gi + b.gi) + e_i * (a.gr - b.gr);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) red[c][tid] = v4[c];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) {
#pragma unroll
                for (int c = 0; c < 4; ++c) red[c][tid] += red[c][tid + h];
            }
            __syncthreads();
        }
        if (tid == 0) {  // chain to the raw outputs
            const int r = port * 64 + m * 4;
            const double u = A.O[(long long)(r + 0) * A.S + s], v = A.O[(long long)(r + 1) * A.S + s];
            const double su = 1.0 / (1.0 + exp(-u)), sv = 1.0 / (1.0 + exp(-v));
            A.dO[(long long)(r + 0) * A.S + s] = (float)(A.scale * red[0][0] * (-(double)A.sig0 * su));
            A.dO[(long long)(r + 1) * A.S + s] = (float)(A.scale * red[1][0] * ((double)A.wmax * sv * (1.0 - sv)));
            A.dO[(long long)(r + 2) * A.S + s] = (float)(A.scale * red[2][0] * (double)A.amp[s]);
            A.dO[(long long)(r + 3) * A.S + s] = (float)(A.scale * red[3][0] * (double)A.amp[s]);
        }
    }
    __syncthreads();
    red[0][tid] = lsum;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) red[0][tid] += red[0][tid + h];
        __syncthreads();
    }
    if (tid == 0) A.loss[s * 4 + port] = (float)red[0][0];
}

// Adam (FP32 master parameters): m = b1 m + (1 - b1) g ; v = b2 v + (1 - b2) g^2 ;
// w -= lr * mhat / (sqrt(vhat) + eps), mhat = m / (1 - b1^t), vhat = v / (1 - b2^t)
__global__ __launch_bounds__(256) void adam(float* w, const float* g, float* m, float* v, long long n, float lr,
                                            float b1, float b2, float eps, float c1, float c2) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace ring

// ---------------------------------------------------------------- host side
namespace {

std::string g_err;
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}
struct Fail {
    int code;
    std::string msg;
};
#define RTRY(expr)                                                                        \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                      \
            throw Fail{_e == hipErrorOutOfMemory ? OARING_E_OOM : OARING_E_DEVICE,        \
                       std::string(#expr) + ": " + hipGetErrorString(_e)};                \
        }                                                                                 \
    } while (0)

// parameter layout (floats)
constexpr int C = 64, CIN = 32, NL = 6, NF = 512, NH = 256, NO = 256;
constexpr long long OFF_WIN = 0, OFF_BIN = OFF_WIN + C * CIN;
constexpr long long OFF_WL = OFF_BIN + C;  // layer l (0-based): W at OFF_WL + l * LSTRIDE, b after it
constexpr long long LSTRIDE = C * 2 * C + C;
constexpr long long OFF_W1 = OFF_WL + NL * LSTRIDE, OFF_B1 = OFF_W1 + (long long)NH * NF;
constexpr long long OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + (long long)NO * NH;
constexpr long long NPARAM = OFF_B2 + NO;
constexpr int KSPLIT = 512;  // reduction length per split of the weight-gradient products

template <typename LA, typename LB, typename EP>
void gemm(hipStream_t st, LA la, LB lb, EP ep, int M, int N, int K, int splits = 1) {
    int kper = (K + splits - 1) / splits;
    kper = (kper + ring::BK - 1) / ring::BK * ring::BK;
    const dim3 grid((N + ring::BN - 1) / ring::BN, (M + ring::BM - 1) / ring::BM, splits);
    hipLaunchKernelGGL((ring::gemm_bf16<LA, LB, EP>), grid, dim3(256), 0, st, la, lb, ep, M, N, K, kper);
    RTRY(hipGetLastError());
}

}  // namespace

struct oaring_net {
    int S = 0, T = 0, F = 0;
    long long N = 0;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    float* U = nullptr;
    float* X[NL + 1] = {};
    float *H = nullptr, *O = nullptr, *dO = nullptr, *dH = nullptr;
    float* dX[2] = {nullptr, nullptr};
    float* part = nullptr;
    long long part_cap = 0;
    float *amp = nullptr, *kap = nullptr, *tgt = nullptr, *t0 = nullptr, *loss = nullptr, *pred = nullptr;
    double* omega = nullptr;
    float dt = 0.f, sig0 = 0.f, wmax = 0.f;
    bool have_batch = false, have_fwd = false;

    template <typename T_>
    T_* dev(size_t n) {
        void* p = nullptr;
        RTRY(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T_)));
        allocs.push_back(p);
        RTRY(hipMemset(p, 0, std::max<size_t>(n, 1) * sizeof(T_)));
        return static_cast<T_*>(p);
    }
    ~oaring_net() {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : allocs) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
    void init(int S_, int T_, int F_) {
        S = S_;
        T = T_;
        F = F_;
        N = (long long)S * T;
        RTRY(hipStreamCreate(&st));
        P = dev<float>(NPARAM);
        G = dev<float>(NPARAM);
        Mo = dev<float>(NPARAM);
        Vo = dev<float>(NPARAM);
        U = dev<float>((size_t)CIN * N);
        for (int l = 0; l <= NL; ++l) X[l] = dev<float>((size_t)C * N);
        H = dev<float>((size_t)NH * S);
        O = dev<float>((size_t)NO * S);
        dO = dev<float>((size_t)NO * S);
        dH = dev<float>((size_t)NH * S);
        dX[0] = dev<float>((size_t)C * N);
        dX[1] = dev<float>((size_t)C * N);
        const long long splits = (N + KSPLIT - 1) / KSPLIT;
        part_cap = splits * C * 2 * C;
        part = dev<float>((size_t)part_cap);
        amp = dev<float>(S);
        t0 = dev<float>(S);
        kap = dev<float>((size_t)S * 4 * F * 2);
        tgt = dev<float>((size_t)S * 4 * F * 2);
        pred = dev<float>((size_t)S * 4 * F * 2);
        loss = dev<float>((size_t)S * 4);
        omega = dev<double>(F);
    }
    void split_gemm_wgrad_sum(long long mn, float* out, int splits) {
        hipLaunchKernelGGL(ring::sum_splits, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, part, splits, mn,
                           out, 0);
        RTRY(hipGetLastError());
    }

    void forward(float* loss_out, float* pred_out, float* raw_out) {
        using namespace ring;
        const int Ni = (int)N;
        // input layer
        gemm(st, Dense{P + OFF_WIN, CIN, 1, C, CIN}, Dense{U, N, 1, CIN, Ni},
             StoreAct{X[0], P + OFF_BIN, nullptr, 1, C, Ni, N}, C, Ni, CIN);
        for (int l = 0; l < NL; ++l) {
            const float* W = P + OFF_WL + l * LSTRIDE;
            gemm(st, Dense{W, 2 * C, 1, C, 2 * C}, ConvIn{X[l], T, 1 << l, Ni},
                 StoreAct{X[l + 1], W + C * 2 * C, X[l], 1, C, Ni, N}, C, Ni, 2 * C);
        }
        gemm(st, Dense{P + OFF_W1, NF, 1, NH, NF}, Feat{X[NL], T, S}, StoreAct{H, P + OFF_B1, nullptr, 1, NH, S, S},
             NH, S, NF);
        gemm(st, Dense{P + OFF_W2, NH, 1, NO, NH}, Dense{H, S, 1, NH, S}, StoreAct{O, P + OFF_B2, nullptr, 0, NO, S, S},
             NO, S, NH);
        PoleArgs a{O, amp, omega, kap, tgt, t0, dt, sig0, wmax, S, F, dO, loss, pred, 1.0 / ((double)S * 4 * F)};
        hipLaunchKernelGGL(ring::pole_loss, dim3(S, 4), dim3(256), 0, st, a);
        RTRY(hipGetLastError());
        std::vector<float> lp((size_t)S * 4);
        RTRY(hipMemcpyAsync(lp.data(), loss, lp.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        if (pred_out)
            RTRY(hipMemcpyAsync(pred_out, pred, (size_t)S * 4 * F * 2 * sizeof(float), hipMemcpyDeviceToHost, st));
        if (raw_out) RTRY(hipMemcpyAsync(raw_out, O, (size_t)NO * S * sizeof(float), hipMemcpyDeviceToHost, st));
        RTRY(hipStreamSynchronize(st));
        double sum = 0.0;
        for (float v : lp) sum += v;  // fixed order
        if (loss_out) *loss_out = (float)(sum / ((double)S * 4 * F));
        have_fwd = true;
    }

    void backward() {
        using namespace ring;
        const int Ni = (int)N;
        // head: O = W2 H + b2
        gemm(st, Dense{dO, S, 1, NO, S}, Dense{H, 1, S, S, NH}, StoreAct{G + OFF_W2, nullptr, nullptr, 0, NO, NH, NH},
             NO, NH, S);
        hipLaunchKernelGGL(ring::row_sums_plain, dim3(NO), dim3(256), 0, st, dO, S, G + OFF_B2);
        RTRY(hipGetLastError());
        gemm(st, Dense{P + OFF_W2, 1, NH, NH, NO}, Dense{dO, S, 1, NO, S}, StoreAct{dH, nullptr, nullptr, 0, NH, S, S},
             NH, S, NO);
        // H = tanh(W1 F + b1)
        const DZ dz1{dH, H, nullptr, NH, S};
        gemm(st, dz1, FeatT{Feat{X[NL], T, S}}, StoreAct{G + OFF_W1, nullptr, nullptr, 0, NH, NF, NF}, NH, NF, S);
        hipLaunchKernelGGL(ring::row_sums, dim3(NH), dim3(256), 0, st, dz1, G + OFF_B1);
        RTRY(hipGetLastError());
        float* g = dX[NL & 1];
        hipLaunchKernelGGL(ring::fill, dim3((unsigned)((C * N + 255) / 256)), dim3(256), 0, st, g, (long long)C * N,
                           0.f);
        RTRY(hipGetLastError());
        gemm(st, Dense{P + OFF_W1, 1, NF, NF, NH}, dz1, ScatterFeat{g, T, S}, NF, S, NH);
        const int splits = (int)((N + KSPLIT - 1) / KSPLIT);
        for (int l = NL - 1; l >= 0; --l) {
            // X(l+1) = X(l) + tanh(Wl [X(l)(t); X(l)(t - d)] + bl); g = dL/dX(l+1)
            const float* W = P + OFF_WL + l * LSTRIDE;
            float* gW = G + OFF_WL + l * LSTRIDE;
            const int d = 1 << l;
            const DZ dz{g, X[l + 1], X[l], C, Ni};
            gemm(st, dz, ConvInT{ConvIn{X[l], T, d, Ni}}, StorePartial{part, C, 2 * C}, C, 2 * C, Ni, splits);
            split_gemm_wgrad_sum((long long)C * 2 * C, gW, splits);
            hipLaunchKernelGGL(ring::row_sums, dim3(C), dim3(256), 0, st, dz, gW + C * 2 * C);
            RTRY(hipGetLastError());
            float* gn = dX[l & 1];
            gemm(st, WTaps{W}, ConvDZ{dz, T, d}, StoreAct{gn, nullptr, g, 0, C, Ni, N}, C, Ni, 2 * C);
            g = gn;
        }
        // X0 = tanh(Win U + bin)
        const DZ dz0{g, X[0], nullptr, C, Ni};
        gemm(st, dz0, Dense{U, 1, N, Ni, CIN}, StorePartial{part, C, CIN}, C, CIN, Ni, splits);
        split_gemm_wgrad_sum((long long)C * CIN, G + OFF_WIN, splits);
        hipLaunchKernelGGL(ring::row_sums, dim3(C), dim3(256), 0, st, dz0, G + OFF_BIN);
        RTRY(hipGetLastError());
        RTRY(hipStreamSynchronize(st));
    }
};

extern "C" {

int64_t oaring_param_count(void) { return NPARAM; }

int oaring_create(int samples, int time_len, int n_freq, oaring_net** out) {
    if (!out) return fail(OARING_E_INVALID, "null output pointer");
    *out = nullptr;
    if (samples < 1 || samples > 4096 || time_len < 8 || n_freq < 1 || n_freq > 256)
        return fail(OARING_E_INVALID, "need 1..4096 samples, >= 8 time samples, 1..256 frequencies");
    if ((long long)samples * time_len >= (1LL << 30)) return fail(OARING_E_INVALID, "batch too large");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return fail(OARING_E_NO_DEVICE, "no ROCm device visible");
    }
    auto* e = new (std::nothrow) oaring_net();
    if (!e) return fail(OARING_E_OOM, "host allocation failed");
    try {
        e->init(samples, time_len, n_freq);
    } catch (const Fail& f) {
        delete e;
        return fail(f.code, f.msg);
    }
    *out = e;
    return OARING_OK;
}

#define NET_CALL(e, ...)                                          \
    do {                                                          \
        if (!(e)) return fail(OARING_E_INVALID, "null network");  \
        try {                                                     \
            __VA_ARGS__;                                          \
        } catch (const Fail& f) {                                 \
            return fail(f.code, f.msg);                           \
        }                                                         \
        return OARING_OK;                                         \
    } while (0)

int oaring_set_params(oaring_net* e, const float* p, int64_t n) {
    if (n != NPARAM || !p) return fail(OARING_E_INVALID, "parameter count mismatch");
    NET_CALL(e, RTRY(hipMemcpy(e->P, p, NPARAM * sizeof(float), hipMemcpyHostToDevice)));
}
int oaring_get_params(oaring_net* e, float* p, int64_t n) {
    if (n != NPARAM || !p) return fail(OARING_E_INVALID, "parameter count mismatch");
    NET_CALL(e, RTRY(hipMemcpy(p, e->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost)));
}
int oaring_get_grads(oaring_net* e, float* g, int64_t n) {
    if (n != NPARAM || !g) return fail(OARING_E_INVALID, "parameter count mismatch");
    NET_CALL(e, RTRY(hipMemcpy(g, e->G, NPARAM * sizeof(float), hipMemcpyDeviceToHost)));
}
int oaring_reset_adam(oaring_net* e) {
    NET_CALL(e, {
        RTRY(hipMemset(e->Mo, 0, NPARAM * sizeof(float)));
        RTRY(hipMemset(e->Vo, 0, NPARAM * sizeof(float)));
    });
}

int oaring_set_batch(oaring_net* e, const float* U, const float* amp, const float* kap, const float* tgt,
                     const float* t0, const double* omega, float dt, float sig0, float wmax) {
    if (!U || !amp || !kap || !tgt || !t0 || !omega) return fail(OARING_E_INVALID, "null batch array");
    if (!(dt > 0.f) || !(sig0 > 0.f) || !(wmax > 0.f)) return fail(OARING_E_INVALID, "bad time scales");
    NET_CALL(e, {
        const size_t nk = (size_t)e->S * 4 * e->F * 2;
        RTRY(hipMemcpy(e->U, U, (size_t)CIN * e->N * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(e->amp, amp, e->S * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(e->kap, kap, nk * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(e->tgt, tgt, nk * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(e->t0, t0, e->S * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(e->omega, omega, e->F * sizeof(double), hipMemcpyHostToDevice));
        e->dt = dt;
        e->sig0 = sig0;
        e->wmax = wmax;
        e->have_batch = true;
        e->have_fwd = false;
    });
}

int oaring_forward(oaring_net* e, float* loss, float* pred, float* raw) {
    if (e && !e->have_batch) return fail(OARING_E_INVALID, "no batch set");
    NET_CALL(e, e->forward(loss, pred, raw));
}
int oaring_backward(oaring_net* e) {
    if (e && !e->have_fwd) return fail(OARING_E_INVALID, "backward needs a forward pass of the same batch");
    NET_CALL(e, e->backward());
}
int oaring_adam(oaring_net* e, float lr, float b1, float b2, float eps, int step) {
    if (step < 1) return fail(OARING_E_INVALID, "Adam step counts from 1");
    NET_CALL(e, {
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(ring::adam, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, e->st, e->P, e->G, e->Mo,
                           e->Vo, NPARAM, lr, b1, b2, eps, c1, c2);
        RTRY(hipGetLastError());
        RTRY(hipStreamSynchronize(e->st));
    });
}

// One MFMA tile product on the device: C (16 x 16) = A (16 x 32) B (32 x 16) in BF16 with FP32
// accumulation, through the same loaders and kernel as the network (layout check of the
// instruction, host arrays row-major).
int oaring_gemm_check(int M, int N, int K, const float* A, const float* B, float* Cout) {
    if (M < 1 || N < 1 || K < 1 || M > 4096 || N > 4096 || K > 65536 || !A || !B || !Cout)
        return fail(OARING_E_INVALID, "bad gemm check arguments");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return fail(OARING_E_NO_DEVICE, "no ROCm device visible");
    }
    float *a = nullptr, *b = nullptr, *c = nullptr, *p = nullptr;
    int rc = OARING_OK;
    try {
        RTRY(hipMalloc(&a, (size_t)M * K * sizeof(float)));
        RTRY(hipMalloc(&b, (size_t)K * N * sizeof(float)));
        RTRY(hipMalloc(&c, (size_t)M * N * sizeof(float)));
        const int splits = (K + KSPLIT - 1) / KSPLIT;
        RTRY(hipMalloc(&p, (size_t)splits * M * N * sizeof(float)));
        RTRY(hipMemcpy(a, A, (size_t)M * K * sizeof(float), hipMemcpyHostToDevice));
        RTRY(hipMemcpy(b, B, (size_t)K * N * sizeof(float), hipMemcpyHostToDevice));
        gemm(nullptr, ring::Dense{a, K, 1, M, K}, ring::Dense{b, N, 1, K, N}, ring::StorePartial{p, M, N}, M, N, K,
             splits);
        hipLaunchKernelGGL(ring::sum_splits, dim3((unsigned)(((long long)M * N + 255) / 256)), dim3(256), 0, nullptr,
                           p, splits, (long long)M * N, c, 0);
        RTRY(hipGetLastError());
        RTRY(hipMemcpy(Cout, c, (size_t)M * N * sizeof(float), hipMemcpyDeviceToHost));
    } catch (const Fail& f) {
        rc = fail(f.code, f.msg);
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (c) (void)hipFree(c);
    if (p) (void)hipFree(p);
    return rc;
}

const char* oaring_last_error(void) { return g_err.c_str(); }
void oaring_destroy(oaring_net* e) { delete e; }

}  // extern "C"
