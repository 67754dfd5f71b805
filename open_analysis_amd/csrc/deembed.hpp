// Plane-wave de-embedding in FP64 on the device: the propagation constant of a uniform line from
// the voltages at three planes, its low-frequency smooth fit, reference-plane shifts, and the
// S matrix from the incident and outgoing waves of every port under every excitation
// (packages/oa-fdtd/src/oa_fdtd/ports/deembed.py: gamma_from_planes :24-53, smooth_gamma :55-82,
// _unequal_residual / _newton_step :85-101, shift_reference :104-121, wave_sparams :264-338).
//
// The work is tiny (ports x excitations x frequencies); it runs on the device so that the whole
// S-parameter path after the DFT stays in FP64 next to the engine. One thread owns one
// (port, excitation) row for the steps that run along frequency (phase unwrapping, the fit),
// and one thread owns one frequency for the per-frequency matrix solve.

#pragma once
#include <hip/hip_runtime.h>

namespace oafd {

struct cd {
    double re, im;
};
__device__ __forceinline__ cd cmk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.re, -a.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) { return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {  // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__device__ __forceinline__ double cabs_(cd a) { return hypot(a.re, a.im); }
__device__ __forceinline__ cd cexp_(cd a) {
    const double m = exp(a.re);
    return cd{m * cos(a.im), m * sin(a.im)};
}
__device__ __forceinline__ cd clog_(cd a) { return cd{log(cabs_(a)), atan2(a.im, a.re)}; }
__device__ __forceinline__ cd csqrt_(cd a) {  // principal root, Re >= 0
    if (a.re == 0.0 && a.im == 0.0) return cd{0.0, a.im};
    const double t = sqrt(0.5 * (fabs(a.re) + cabs_(a)));
    if (a.re >= 0.0) return cd{t, a.im / (2.0 * t)};
    return cd{fabs(a.im) / (2.0 * t), copysign(t, a.im)};
}
__device__ __forceinline__ cd ccosh_(cd a) { return cd{cosh(a.re) * cos(a.im), sinh(a.re) * sin(a.im)}; }
__device__ __forceinline__ cd csinh_(cd a) { return cd{sinh(a.re) * cos(a.im), cosh(a.re) * sin(a.im)}; }
// principal acosh: log(z + sqrt(z + 1) sqrt(z - 1)), Re >= 0
__device__ __forceinline__ cd cacosh_(cd z) {
    return clog_(z + csqrt_(z + cmk(1.0, 0.0)) * csqrt_(z - cmk(1.0, 0.0)));
}
__device__ __forceinline__ bool cfinite(cd a) { return isfinite(a.re) && isfinite(a.im); }

// Python's float modulo (the sign of the divisor), as np.mod
__device__ __forceinline__ double pymod(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

__device__ __forceinline__ cd unequal_residual(cd g, cd v1, cd v2, cd v3, double da, double db) {
    const cd gda = da * g, gdb = db * g;
    return (v1 - v2 * ccosh_(gda)) * csinh_(gdb) + (v3 - v2 * ccosh_(gdb)) * csinh_(gda);
}

// gamma of one row of nf frequencies from the planes' voltages (strided inputs); unwrapped
__device__ void gamma_row(const cd* v1, const cd* v2, const cd* v3, int nf, double da, double db, cd* g) {
    const double dm = 0.5 * (da + db);
    const bool unequal = fabs(db - da) > 1e-9 * da;
    for (int f = 0; f < nf; ++f) {
        const cd c = cdiv(v1[f] + v3[f], 2.0 * v2[f]);
        cd x = cacosh_(c);
        x = cd{x.re / dm, x.im / dm};
        if (x.re < 0.0) x = -x;
        if (unequal) {
            for (int it = 0; it < 8; ++it) {
                const double h = 1e-6 * fmax(cabs_(x), 1.0);
                const cd f0 = unequal_residual(x, v1[f], v2[f], v3[f], da, db);
                const cd f1 = unequal_residual(x + cmk(h, 0.0), v1[f], v2[f], v3[f], da, db);
                const cd df = cd{(f1.re - f0.re) / h, (f1.im - f0.im) / h};
                if (cabs_(df) > 0.0) x = x - cdiv(f0, df);
            }
        }
        if (x.re < 0.0) x = -x;
        if (x.im < 0.0) x.im = -x.im;  // conj
        g[f] = x;
    }
    // np.unwrap of beta * dm along frequency (period 2 pi, discontinuity pi)
    const double pi = 3.141592653589793, tau = 6.283185307179586;
    double prev = g[0].im * dm, corr = 0.0;
    for (int f = 1; f < nf; ++f) {
        const double p = g[f].im * dm;
        const double dd = p - prev;
        double ddm = pymod(dd + pi, tau) - pi;
        if (ddm == -pi && dd > 0.0) ddm = pi;
        if (fabs(dd) >= pi) corr += ddm - dd;
        prev = p;
        g[f].im = (p + corr) / dm;
    }
}

// least squares y ~ X c for three columns (modified Gram-Schmidt QR over the used rows)
__device__ void lsq3(const double* x0, const double* x1, const double* x2, const double* y, const unsigned char* use,
                     int nf, double* q, double* c) {
    // q: 3 * nf scratch
    const double* xs[3] = {x0, x1, x2};
    double r[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < 3; ++a) {
        for (int f = 0; f < nf; ++f) q[a * nf + f] = use[f] ? xs[a][f] : 0.0;
        for (int p = 0; p < a; ++p) {
            double d = 0.0;
            for (int f = 0; f < nf; ++f) d += q[p * nf + f] * q[a * nf + f];
            r[p][a] = d;
            for (int f = 0; f < nf; ++f) q[a * nf + f] -= d * q[p * nf + f];
        }
        double nrm = 0.0;
        for (int f = 0; f < nf; ++f) nrm += q[a * nf + f] * q[a * nf + f];
        nrm = sqrt(nrm);
        r[a][a] = nrm;
        for (int f = 0; f < nf; ++f) q[a * nf + f] = nrm > 0.0 ? q[a * nf + f] / nrm : 0.0;
    }
    double qty[3];
    for (int a = 0; a < 3; ++a) {
        double d = 0.0;
        for (int f = 0; f < nf; ++f) d += use[f] ? q[a * nf + f] * y[f] : 0.0;
        qty[a] = d;
    }
    for (int a = 2; a >= 0; --a) {
        double t = qty[a];
        for (int p = a + 1; p < 3; ++p) t -= r[a][p] * c[p];
        c[a] = r[a][a] != 0.0 ? t / r[a][a] : 0.0;
    }
}

// smooth_gamma on one row; scratch holds 9 nf doubles and nf bytes
__device__ void smooth_row(cd* g, int nf, const double* freqs, double spacing, double min_phase,
                           const unsigned char* mask, double* scratch, unsigned char* use) {
    int n_use = 0, n_ok = 0;
    double fmax_ = 0.0;
    for (int f = 0; f < nf; ++f) fmax_ = fmax(fmax_, freqs[f]);
    for (int f = 0; f < nf; ++f) {
        const bool ok = fabs(g[f].im) * spacing >= min_phase;
        use[f] = (ok && (!mask || mask[f])) ? 1 : 0;
        n_ok += ok;
        n_use += use[f];
    }
    if (n_use < 6 || n_ok == nf) return;
    double* fn = scratch;
    double* sq = scratch + nf;
    double* f2 = scratch + 2 * nf;
    double* f3 = scratch + 3 * nf;
    double* yr = scratch + 4 * nf;
    double* q = scratch + 5 * nf;  // 3 nf
    double* yi = scratch + 8 * nf;
    for (int f = 0; f < nf; ++f) {
        fn[f] = freqs[f] / fmax_;
        sq[f] = sqrt(fn[f]);
        f2[f] = fn[f] * fn[f];
        f3[f] = f2[f] * fn[f];
        yr[f] = g[f].re;
        yi[f] = g[f].im;
    }
    double ca[3], cb[3];
    lsq3(sq, fn, f2, yr, use, nf, q, ca);
    lsq3(fn, f2, f3, yi, use, nf, q, cb);
    for (int f = 0; f < nf; ++f) {
        if (fabs(g[f].im) * spacing >= min_phase) continue;
        g[f] = cd{sq[f] * ca[0] + fn[f] * ca[1] + f2[f] * ca[2], fn[f] * cb[0] + f2[f] * cb[1] + f3[f] * cb[2]};
    }
}

// gamma_from_planes over m rows: v1/v2/v3 (m, nf); out (m, nf)
__global__ void planes_gamma_k(int m, int nf, const cd* v1, const cd* v2, const cd* v3, double da, double db, cd* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const long long o = (long long)r * nf;
    gamma_row(v1 + o, v2 + o, v3 + o, nf, da, db, out + o);
}

// smooth_gamma over m rows in place; scratch (m, 9 nf) doubles and (m, nf) bytes
__global__ void smooth_gamma_k(int m, int nf, cd* g, const double* freqs, double spacing, double min_phase,
                               const unsigned char* mask, double* scratch, unsigned char* use) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    smooth_row(g + (long long)r * nf, nf, freqs, spacing, min_phase, mask, scratch + (long long)r * 9 * nf,
               use + (long long)r * nf);
}

// S'_ij = S_ij e^{g_i L_i} e^{g_j L_j}; s (n, n, nf), gamma (n, nf)
__global__ void shift_reference_k(int n, int nf, const cd* s, const cd* gamma, const double* len, cd* out) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (t >= (long long)n * n * nf) return;
    const int f = (int)(t % nf), ij = (int)(t / nf), i = ij / n, j = ij % n;
    const cd fi = cexp_(len[i] * gamma[(long long)i * nf + f]);
    const cd fj = cexp_(len[j] * gamma[(long long)j * nf + f]);
    out[t] = s[t] * fi * fj;
}

struct WaveArgs {
    int n, nf;
    const cd* v;          // (n, 3, n, nf): port i, plane p, excitation j
    const double* da;     // (n) spacing between planes 1 and 2
    const double* db;     // (n) spacing between planes 2 and 3
    const double* ref;    // (n) middle plane -> reference plane, positive inwards
    const double* port_r; // (n)
    const double* freqs;  // (nf) or null: no smoothing
    const unsigned char* mask;  // (nf) or null
    double min_phase, floor_phase;
    cd* g;       // (n, n, nf) gamma per (i, j)
    cd* a;       // (n, n, nf)
    cd* b;       // (n, n, nf)
    double* scratch;      // (n n, 9 nf)
    unsigned char* use;   // (n n, nf)
};

// per (port i, excitation j): gamma, optional smoothing, waves at the reference plane
__global__ void wave_rows_k(WaveArgs W) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= W.n * W.n) return;
    const int i = r / W.n, j = r % W.n, nf = W.nf;
    const cd* v1 = W.v + (((long long)i * 3 + 0) * W.n + j) * nf;
    const cd* v2 = W.v + (((long long)i * 3 + 1) * W.n + j) * nf;
    const cd* v3 = W.v + (((long long)i * 3 + 2) * W.n + j) * nf;
    cd* g = W.g + (long long)r * nf;
    gamma_row(v1, v2, v3, nf, W.da[i], W.db[i], g);
    if (W.freqs) smooth_row(g, nf, W.freqs, W.da[i], W.min_phase, W.mask, W.scratch + (long long)r * 9 * nf,
                            W.use + (long long)r * nf);
    const double sr = sqrt(W.port_r[i]);
    for (int f = 0; f < nf; ++f) {
        const cd gs = W.da[i] * g[f];
        const cd d_mid = cdiv(v1[f] - v2[f] * ccosh_(gs), csinh_(gs));
        const cd a_mid = 0.5 * (v2[f] + d_mid), b_mid = 0.5 * (v2[f] - d_mid);
        const cd a = a_mid * cexp_(-(W.ref[i] * g[f])), b = b_mid * cexp_(W.ref[i] * g[f]);
        W.a[(long long)r * nf + f] = cd{a.re / sr, a.im / sr};
        W.b[(long long)r * nf + f] = cd{b.re / sr, b.im / sr};
    }
}

constexpr int MAXP = 16;  // ports of one plane-wave extraction

// per frequency: S = B A^-1 (or the driven-wave normalisation), then the port-S fallbacks
__global__ void wave_solve_k(WaveArgs W, const cd* s_port, cd* s) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= W.nf) return;
    const int n = W.n, nf = W.nf;
    auto at = [&](const cd* x, int i, int j) { return x[((long long)i * n + j) * nf + f]; };
    cd M[MAXP][MAXP], X[MAXP][MAXP];
    double dmin = 1e300, off = 0.0;
    bool fin = true;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            const cd v = at(W.a, i, j);
            fin = fin && cfinite(v);
            if (i == j) dmin = fmin(dmin, cabs_(v));
            else off = fmax(off, cabs_(v));
        }
    if (n == 1 || off > 0.3 * dmin || !fin) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) s[((long long)i * n + j) * nf + f] = cdiv(at(W.b, i, j), at(W.a, j, j));
    } else {
        // S A = B  <=>  A^T S^T = B^T: LU with partial pivoting (largest |re| + |im|) on A^T
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                M[i][j] = at(W.a, j, i);
                X[i][j] = at(W.b, j, i);
            }
        for (int c = 0; c < n; ++c) {
            int p = c;
            double best = fabs(M[c][c].re) + fabs(M[c][c].im);
            for (int q = c + 1; q < n; ++q) {
                const double v = fabs(M[q][c].re) + fabs(M[q][c].im);
                if (v > best) {
                    best = v;
                    p = q;
                }
            }
            if (p != c)
                for (int k = 0; k < n; ++k) {
                    const cd t = M[c][k]; M[c][k] = M[p][k]; M[p][k] = t;
                    const cd u = X[c][k]; X[c][k] = X[p][k]; X[p][k] = u;
                }
            for (int q = c + 1; q < n; ++q) {
                const cd l = cdiv(M[q][c], M[c][c]);
                for (int k = c + 1; k < n; ++k) M[q][k] = M[q][k] - l * M[c][k];
                for (int k = 0; k < n; ++k) X[q][k] = X[q][k] - l * X[c][k];
            }
        }
        for (int c = n - 1; c >= 0; --c)
            for (int k = 0; k < n; ++k) {
                cd t = X[c][k];
                for (int q = c + 1; q < n; ++q) t = t - M[c][q] * X[q][k];
                X[c][k] = cdiv(t, M[c][c]);
            }
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) s[((long long)i * n + j) * nf + f] = X[j][i];
    }
    if (!s_port) return;
    // the ill-conditioned band, judged on excitation 0's gamma of every port
    const double thr = W.freqs ? W.floor_phase : W.min_phase;
    bool weak = false;
    for (int i = 0; i < n; ++i) weak = weak || fabs(W.g[((long long)i * n + 0) * nf + f].im) * W.da[i] < thr;
    if (!W.freqs) {
        if (weak)
            for (int q = 0; q < n * n; ++q) s[(long long)q * nf + f] = s_port[(long long)q * nf + f];
        return;
    }
    if (weak)
        for (int i = 0; i < n; ++i) s[((long long)i * n + i) * nf + f] = s_port[((long long)i * n + i) * nf + f];
    double mx = 0.0;
    for (int q = 0; q < n * n; ++q) mx = fmax(mx, cabs_(s[(long long)q * nf + f]));
    if (mx > 1.02)
        for (int q = 0; q < n * n; ++q) s[(long long)q * nf + f] = s_port[(long long)q * nf + f];
}

}  // namespace oafd
