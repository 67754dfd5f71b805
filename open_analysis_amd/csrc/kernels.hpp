// Device kernels of the MI355X (gfx950) oa-fdtd engine.
//
// Storage: one array per field component, (B, Z, Y, X) with x fastest, Z = nz+2, Y = ny+2,
// X = pitch (the reference layout, packages/oa-fdtd/src/oa_fdtd/grid.py:3-10). Fields are
// double-buffered: step n reads (E^n, H^{n-1/2}) from one set and writes (E^{n+1}, H^{n+1/2})
// to the other, so a single fused pass never reads a value another workgroup has already
// advanced.
//
// The fused step (yee_step) marches a 63x15-cell (x, y) tile up a chunk of z planes. Per plane
// it stages E^n of planes k and k+1 (with a one-cell halo) in LDS, computes H^{n+1/2} of plane k
// for the tile plus its low-x and low-y halo (recomputed, not exchanged), stages that in LDS,
// and then advances E of plane k. H of plane k-1 stays in registers. CPML psi recursions run only
// for cells inside a correction box; the Debye ADE current is read and written once per
// dispersive edge; PMC ghosts are produced by mirroring inside the tile.
//
// Equations follow packages/oa-fdtd/src/oa_fdtd/reference.py:152-266 (step order :251-266) and
// kernels/yee.cu, physics.cu; the Debye post-update is folded into the E update and its
// interaction with the later source injection is restored exactly in inject_step.

#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace oafd {

constexpr int TXH = 64;             // H columns per tile = one wavefront
constexpr int NWAVE = 4;            // wavefronts per workgroup
constexpr int TXE = TXH - 1;        // E columns per tile
constexpr int NTHREADS = TXH * NWAVE;
constexpr int MAX_POLES = 16;
constexpr int MAX_CORR = 24;

struct CorrBox {
    int lo[3];
    int n[3];
    long long off;  // psi offset of batch member 0
    long long vol;  // psi values per batch member
};

template <typename R>
struct StepArgs {
    int nx, ny, nz, X, Y, Z;
    long long bs;
    int B;
    int kchunk, nkc;
    R dtmu;    // (dt/mu0) * hscale
    R inv_hs;  // 1 / hscale (H is stored scaled by hscale)
    const void* Ein[3];
    const void* Hin[3];
    void* Eout[3];
    void* Hout[3];
    const uint16_t* idx[3];
    const R* ca;
    const R* cb;
    const R* ip[3];
    const R* id[3];
    int pmc_lo[3], pmc_hi[3];
    const R* pe[3][3];  // per axis eb, ec, ek (E-type, node indices)
    const R* ph[3][3];  // per axis hb, hc, hk (H-type, cell indices)
    signed char corrE[3][2][2];  // [component][curl term][side] -> box or -1
    signed char corrH[3][2][2];
    CorrBox box[MAX_CORR];
    R* psiE;
    const R* psiHin;
    R* psiHout;
    int npole, ade_first, ade_nrow, jk0, njk;
    R kp[MAX_POLES];
    R kq[MAX_POLES];  // (1 + kp) / 2
    const R* beta;    // (npole, ade_nrow)
    R* J;             // (npole, B, 3, njk, Y, X)
};

// ---- storage conversions
template <typename S, typename R>
__device__ __forceinline__ R ld(const void* base, long long i) {
    return (R)(static_cast<const S*>(base)[i]);
}
template <typename R>
__device__ __forceinline__ R ld_bf(const void* base, long long i) {
    return (R)(float)(static_cast<const __bf16*>(base)[i]);
}
template <typename S, typename R>
__device__ __forceinline__ R ldS(const void* base, long long i) {
    if constexpr (__is_same(S, __bf16)) return (R)(float)(static_cast<const S*>(base)[i]);
    else if constexpr (__is_same(S, _Float16)) return (R)(float)(static_cast<const S*>(base)[i]);
    else return (R)(static_cast<const S*>(base)[i]);
}
template <typename S, typename R>
__device__ __forceinline__ void stS(void* base, long long i, R v) {
    if constexpr (__is_same(S, __bf16) || __is_same(S, _Float16))
        static_cast<S*>(base)[i] = (S)(float)v;
    else
        static_cast<S*>(base)[i] = (S)v;
}

__device__ __forceinline__ bool in_box(const CorrBox& b, int i, int j, int k) {
    return (unsigned)(i - b.lo[0]) < (unsigned)b.n[0] && (unsigned)(j - b.lo[1]) < (unsigned)b.n[1] &&
           (unsigned)(k - b.lo[2]) < (unsigned)b.n[2];
}
__device__ __forceinline__ long long box_index(const CorrBox& b, int bb, int i, int j, int k) {
    return b.off + (long long)bb * b.vol +
           ((long long)(k - b.lo[2]) * b.n[1] + (j - b.lo[1])) * b.n[0] + (i - b.lo[0]);
}

template <typename R, int TY>
using Plane = R[3][TY + 1][TXH + 1];

// Stage E^n of storage plane k (3 components, rows j0-1 .. j0+TYH-1, columns i0-1 .. i0+TXH-1)
// into LDS; cells outside the storage are zero.
template <typename S, typename R, int TYH>
__device__ __forceinline__ void stage_plane(const StepArgs<R>& A, int b, int i0, int j0, int k,
                                            Plane<R, TYH>& dst) {
    const int lane = threadIdx.x, wv = threadIdx.y;
    const long long base = (long long)b * A.bs + (long long)k * A.Y * A.X;
    const bool kin = k < A.Z;
    for (int r = wv; r < 3 * (TYH + 1); r += NWAVE) {
        const int c = r / (TYH + 1), lj = r - c * (TYH + 1);
        const int i = i0 - 1 + lane, j = j0 - 1 + lj;
        R v = (R)0;
        if (kin && j < A.Y && i < A.X) v = ldS<S, R>(A.Ein[c], base + (long long)j * A.X + i);
        dst[c][lj][lane] = v;
    }
    const int t = threadIdx.y * TXH + threadIdx.x;
    if (t < 3 * (TYH + 1)) {
        const int c = t / (TYH + 1), lj = t - c * (TYH + 1);
        const int i = i0 - 1 + TXH, j = j0 - 1 + lj;
        R v = (R)0;
        if (kin && j < A.Y && i < A.X) v = ldS<S, R>(A.Ein[c], base + (long long)j * A.X + i);
        dst[c][lj][TXH] = v;
    }
}

// H^{n+1/2}_C at storage (i, j, k), local (li, lj), from hold = H^{n-1/2}:
//   H_C += dtmu * (d1' - d0'),  d_t = (E_b(p + s_a) - E_b(p)) / d_a,  d' = (1 + hk) d + psi in a
// CPML box (terms: t=0 a=(C+1)%3 b=(C+2)%3, t=1 a=(C+2)%3 b=(C+1)%3; pml.py:34-41 signs).
template <int C, typename R, int TYH>
__device__ __forceinline__ R h_new(const StepArgs<R>& A, int i, int j, int k, int li, int lj,
                                   const Plane<R, TYH>& Ec, const Plane<R, TYH>& En, R hold, int b,
                                   bool wpsi) {
    R d[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int ax = t == 0 ? (C + 1) % 3 : (C + 2) % 3;
        const int sc = t == 0 ? (C + 2) % 3 : (C + 1) % 3;
        const R e0 = Ec[sc][lj][li];
        const R e1 = ax == 0 ? Ec[sc][lj][li + 1] : (ax == 1 ? Ec[sc][lj + 1][li] : En[sc][lj][li]);
        const int I = ax == 0 ? i : (ax == 1 ? j : k);
        R dd = (e1 - e0) * A.ip[ax][I];
        const int nA = ax == 0 ? A.nx : (ax == 1 ? A.ny : A.nz);
        const int bid = A.corrH[C][t][2 * I > nA ? 1 : 0];
        if (bid >= 0) {
            const CorrBox& bx = A.box[bid];
            if (in_box(bx, i, j, k)) {
                const long long q = box_index(bx, b, i, j, k);
                const R psi = A.ph[ax][0][I] * A.psiHin[q] + A.ph[ax][1][I] * dd;
                if (wpsi) A.psiHout[q] = psi;
                dd = dd + A.ph[ax][2][I] * dd + psi;
            }
        }
        d[t] = dd;
    }
    return hold + A.dtmu * (d[1] - d[0]);
}

// Where H_C at (i, j) comes from: mode 0 = computed at (si, sj) times sgn (a valid cell or its
// PMC mirror), mode 1 = zero (outside the component's lattice, no PMC ghost).
template <int C, typename R>
__device__ __forceinline__ int h_source(const StepArgs<R>& A, int i, int j, int& si, int& sj, R& sgn) {
    si = i;
    sj = j;
    sgn = (R)1;
    if (C == 0) {
        if (i < 1 || i > A.nx + 1) return 1;
    } else {
        if (i == 0) {
            if (!A.pmc_lo[0]) return 1;
            si = 1;
            sgn = -sgn;
        } else if (i == A.nx + 1) {
            if (!A.pmc_hi[0]) return 1;
            si = A.nx;
            sgn = -sgn;
        } else if (i > A.nx + 1) {
            return 1;
        }
    }
    if (C == 1) {
        if (j < 1 || j > A.ny + 1) return 1;
    } else {
        if (j == 0) {
            if (!A.pmc_lo[1]) return 1;
            sj = 1;
            sgn = -sgn;
        } else if (j == A.ny + 1) {
            if (!A.pmc_hi[1]) return 1;
            sj = A.ny;
            sgn = -sgn;
        } else if (j > A.ny + 1) {
            return 1;
        }
    }
    return 0;
}

template <typename S, typename R, int RPW>
__global__ __launch_bounds__(NTHREADS) void yee_step(StepArgs<R> A) {
    constexpr int TYH = NWAVE * RPW;  // H rows per tile (E rows: TYH - 1)
    constexpr int TYE = TYH - 1;
    __shared__ R sE[2][3][TYH + 1][TXH + 1];
    __shared__ R sH[3][TYH][TXH];

    const int lane = threadIdx.x, wv = threadIdx.y;
    const int i0 = 1 + blockIdx.x * TXE;
    const int j0 = 1 + blockIdx.y * TYE;
    const int zc = blockIdx.z % A.nkc;
    const int b = blockIdx.z / A.nkc;
    const int kbeg = 1 + zc * A.kchunk;
    const int kend = min(kbeg + A.kchunk, A.nz + 2);
    const int kstart = kbeg > 1 ? kbeg - 1 : kbeg;
    const long long bb = (long long)b * A.bs;
    const long long sxy = (long long)A.Y * A.X;
    const int i = i0 - 1 + lane;
    const bool col_in = i <= A.nx + 1;

    R hprev[RPW][3];
    R hcur[RPW][3];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) hprev[r][c] = (R)0;

    int cur = 0;
    stage_plane<S, R, TYH>(A, b, i0, j0, kstart, sE[0]);
    for (int k = kstart; k < kend; ++k) {
        stage_plane<S, R, TYH>(A, b, i0, j0, k + 1, sE[cur ^ 1]);
        __syncthreads();
        const bool own_k = k >= kbeg;
        // ---- H^{n+1/2} of plane k for the tile and its low halo
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int lj = wv * RPW + r;
            const int j = j0 - 1 + lj;
            const bool in = col_in && j <= A.ny + 1;
            const bool own = in && own_k && (lane >= 1 || i == 0) && (lj >= 1 || j == 0);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                R v = (R)0;
                if (in) {
                    int si, sj;
                    R sgn;
                    int mode = 1;
                    if (c == 0) mode = h_source<0, R>(A, i, j, si, sj, sgn);
                    if (c == 1) mode = h_source<1, R>(A, i, j, si, sj, sgn);
                    if (c == 2) mode = h_source<2, R>(A, i, j, si, sj, sgn);
                    const bool zghost = (c != 2) && k == A.nz + 1;
                    if (mode == 0 && zghost) {
                        v = A.pmc_hi[2] ? -hprev[r][c] : (R)0;
                    } else if (mode == 0) {
                        const int sli = si - (i0 - 1), slj = sj - (j0 - 1);
                        const R hold = ldS<S, R>(A.Hin[c], bb + k * sxy + (long long)sj * A.X + si);
                        const bool wpsi = own && si == i && sj == j;
                        if (c == 0) v = h_new<0, R, TYH>(A, si, sj, k, sli, slj, sE[cur], sE[cur ^ 1], hold, b, wpsi);
                        if (c == 1) v = h_new<1, R, TYH>(A, si, sj, k, sli, slj, sE[cur], sE[cur ^ 1], hold, b, wpsi);
                        if (c == 2) v = h_new<2, R, TYH>(A, si, sj, k, sli, slj, sE[cur], sE[cur ^ 1], hold, b, wpsi);
                        v *= sgn;
                    }
                    if (own && (mode == 0)) stS<S, R>(A.Hout[c], bb + k * sxy + (long long)j * A.X + i, v);
                }
                sH[c][lj][lane] = v;
                hcur[r][c] = v;
            }
            if (k == 1 && own_k) {  // z-low ghost plane k = 0 (tangential H mirrors)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    hprev[r][c] = A.pmc_lo[2] ? -hcur[r][c] : (R)0;
                    if (own && A.pmc_lo[2]) stS<S, R>(A.Hout[c], bb + (long long)j * A.X + i, hprev[r][c]);
                }
                hprev[r][2] = (R)0;
            }
        }
        __syncthreads();
        // ---- E^{n+1} of plane k for the tile
        if (own_k) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int lj = wv * RPW + r;
                const int j = j0 - 1 + lj;
                if (lane < 1 || lj < 1 || !col_in || j > A.ny + 1) continue;
                const long long q = k * sxy + (long long)j * A.X + i;
                const R idxv = A.id[0][i], idyv = A.id[1][j], idzv = A.id[2][k];
                R d[3][2];
                // Ex: + (Hz - Hz[j-1]) / dy  - (Hy - Hy[k-1]) / dz
                d[0][0] = (sH[2][lj][lane] - sH[2][lj - 1][lane]) * idyv;
                d[0][1] = (sH[1][lj][lane] - hprev[r][1]) * idzv;
                // Ey: + (Hx - Hx[k-1]) / dz  - (Hz - Hz[i-1]) / dx
                d[1][0] = (sH[0][lj][lane] - hprev[r][0]) * idzv;
                d[1][1] = (sH[2][lj][lane] - sH[2][lj][lane - 1]) * idxv;
                // Ez: + (Hy - Hy[i-1]) / dx  - (Hx - Hx[j-1]) / dy
                d[2][0] = (sH[1][lj][lane] - sH[1][lj][lane - 1]) * idxv;
                d[2][1] = (sH[0][lj][lane] - sH[0][lj - 1][lane]) * idyv;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    // CPML (E-type coefficients at node indices); psi in place, owner only
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int ax = t == 0 ? (c + 1) % 3 : (c + 2) % 3;
                        const int I = ax == 0 ? i : (ax == 1 ? j : k);
                        const int nA = ax == 0 ? A.nx : (ax == 1 ? A.ny : A.nz);
                        const int bid = A.corrE[c][t][2 * I > nA + 1 ? 1 : 0];
                        if (bid >= 0) {
                            const CorrBox& bx = A.box[bid];
                            if (in_box(bx, i, j, k)) {
                                const long long pq = box_index(bx, b, i, j, k);
                                const R psi = A.pe[ax][0][I] * A.psiE[pq] + A.pe[ax][1][I] * d[c][t];
                                A.psiE[pq] = psi;
                                d[c][t] = d[c][t] + A.pe[ax][2][I] * d[c][t] + psi;
                            }
                        }
                    }
                    const int m = A.idx[c][q];
                    const R eold = sE[cur][c][lj][lane];
                    R en = A.ca[m] * eold + A.cb[m] * ((d[c][0] - d[c][1]) * A.inv_hs);
                    if (m >= A.ade_first) {
                        // Debye ADE (reference.py:115-127): E -= cb * sum_p (1+k_p)/2 J_p;
                        // J_p = k_p J_p + beta_p (E^{n+1} - E^n)
                        const int row = m - A.ade_first;
                        const long long jq = ((long long)(k - A.jk0) * A.Y + j) * A.X + i;
                        const long long jstride = (long long)A.njk * sxy;
                        R jt = (R)0;
                        for (int p = 0; p < A.npole; ++p) {
                            const long long jp = ((long long)(p * A.B + b) * 3 + c) * jstride + jq;
                            jt += A.kq[p] * A.J[jp];
                        }
                        en = en - A.cb[m] * jt;
                        const R de = en - eold;
                        for (int p = 0; p < A.npole; ++p) {
                            const long long jp = ((long long)(p * A.B + b) * 3 + c) * jstride + jq;
                            A.J[jp] = A.kp[p] * A.J[jp] + A.beta[(long long)p * A.ade_nrow + row] * de;
                        }
                    }
                    stS<S, R>(A.Eout[c], bb + q, en);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) hprev[r][c] = hcur[r][c];
        cur ^= 1;
        __syncthreads();
    }
}

// ---- surface impedance (physics.cu:55-86 / reference.py:129-149): one thread per (edge, b)
template <typename R>
struct SibcArgs {
    int n, np, B;
    long long bs;
    R inv_hs;
    void* E[3];
    const void* H[3];
    const int* comp;
    const int* flat;
    const int* hcomp;  // (n, 2)
    const int* hflat;  // (n, 2)
    const R* hsign;    // (n, 2)
    const int* mat;
    const R* r0;
    const R* alpha;  // (nmat, np)
    const R* gp;
    R* jprev;  // (B, n)
    R* v;      // (np, B, n)
};

template <typename S, typename R>
__global__ __launch_bounds__(256) void sibc_step(SibcArgs<R> A) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (e >= A.n) return;
    const long long bb = (long long)b * A.bs;
    R jv = (R)0;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const R sg = A.hsign[2 * e + s];
        if (sg != (R)0) jv += sg * (ldS<S, R>(A.H[A.hcomp[2 * e + s]], bb + A.hflat[2 * e + s]) * A.inv_hs);
    }
    const int m = A.mat[e];
    const long long be = (long long)b * A.n + e;
    const R dj = jv - A.jprev[be];
    R out = A.r0[m] * jv;
    for (int p = 0; p < A.np; ++p) {
        const long long q = (long long)p * A.B * A.n + be;
        const R nv = A.alpha[m * A.np + p] * dj + A.gp[p] * A.v[q];
        A.v[q] = nv;
        out += nv;
    }
    A.jprev[be] = jv;
    stS<S, R>(A.E[A.comp[e]], bb + A.flat[e], out);
}

// ---- sources (yee.cu:96-108): one thread per unique target edge, terms in channel order.
// When the edge is dispersive, J_p += beta_p * dE restores the Debye update's dependence on the
// injected amount (the ADE post-update runs after injection in the reference order).
template <typename R>
struct InjectArgs {
    int n, B, nch;
    long long bs;
    long long max_steps;
    void* E[3];
    const uint16_t* idx[3];
    const int* comp;
    const int* flat;
    const int* start;
    const double* amp;
    const int* chan;
    const double* drive;
    const double* wave;
    const long long* step_base;
    int step_off;
    int Y, X;
    int npole, ade_first, ade_nrow, jk0, njk;
    const R* beta;
    R* J;
};

template <typename S, typename R>
__global__ __launch_bounds__(256) void inject_step(InjectArgs<R> A) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= A.n) return;
    const long long step = *A.step_base + A.step_off;
    const int c = A.comp[t];
    const int f = A.flat[t];
    const int m = A.idx[c][f];
    const bool ade = m >= A.ade_first;
    const long long sxy = (long long)A.Y * A.X;
    for (int b = 0; b < A.B; ++b) {
        const long long p = (long long)b * A.bs + f;
        R e = ldS<S, R>(A.E[c], p);
        R added = (R)0;
        for (int q = A.start[t]; q < A.start[t + 1]; ++q) {
            const int ch = A.chan[q];
            const R amt = (R)(A.drive[b * A.nch + ch] * A.wave[(long long)ch * A.max_steps + step] * A.amp[q]);
            e += amt;
            added += amt;
        }
        stS<S, R>(A.E[c], p, e);
        if (ade) {
            const int k = f / (int)sxy;
            const long long jq = (long long)(k - A.jk0) * sxy + (f - (long long)k * sxy);
            const long long jstride = (long long)A.njk * sxy;
            const int row = m - A.ade_first;
            for (int pp = 0; pp < A.npole; ++pp) {
                const long long jp = ((long long)(pp * A.B + b) * 3 + c) * jstride + jq;
                A.J[jp] += A.beta[(long long)pp * A.ade_nrow + row] * added;
            }
        }
    }
}

// ---- probes (yee.cu:112-131): weighted sums in FP64, fixed-order reduction, grid (P, B)
struct ProbeArgs {
    int nE, nH, B;
    long long bs;
    double inv_hs;
    const void* F[6];
    const int* comp;   // (nE + nH), 0..5
    const long long* off;  // (nE + nH + 1)
    const int* flat;
    const double* w;
    double* ringE;  // (CHUNK, B, nE)
    double* ringH;
    int slot;
};

template <typename S>
__global__ __launch_bounds__(256) void probe_step(ProbeArgs A) {
    __shared__ double red[256];
    const int p = blockIdx.x, b = blockIdx.y;
    const int c = A.comp[p];
    const long long bb = (long long)b * A.bs;
    double acc = 0.0;
    for (long long e = A.off[p] + threadIdx.x; e < A.off[p + 1]; e += 256)
        acc += ldS<S, double>(A.F[c], bb + A.flat[e]) * A.w[e];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (p < A.nE)
            A.ringE[((long long)A.slot * A.B + b) * A.nE + p] = red[0];
        else
            A.ringH[((long long)A.slot * A.B + b) * A.nH + (p - A.nE)] = red[0] * A.inv_hs;
    }
}

// ---- running DFT with Kahan-compensated FP64 accumulation: one thread per (f, b*P + p)
__global__ __launch_bounds__(256) void dft_accum(const double* ring, int BP, int nf, const double* omega,
                                                 double dt, double t_off, const long long* step_base,
                                                 int slot0, int count, double* acc, double* comp) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nf * BP) return;
    const int f = (int)(t / BP), q = (int)(t - (long long)f * BP);
    const double w = omega[f];
    const long long n0 = *step_base + slot0;
    double sr = acc[2 * t], si = acc[2 * t + 1];
    double cr = comp[2 * t], ci = comp[2 * t + 1];
    for (int n = 0; n < count; ++n) {
        const double v = ring[(long long)(slot0 + n) * BP + q] * dt;
        const double ang = w * (((double)(n0 + n)) + t_off) * dt;
        double s, c;
        sincos(ang, &s, &c);
        const double yr = v * c - cr;
        const double tr = sr + yr;
        cr = (tr - sr) - yr;
        sr = tr;
        const double yi = -v * s - ci;
        const double ti = si + yi;
        ci = (ti - si) - yi;
        si = ti;
    }
    acc[2 * t] = sr;
    acc[2 * t + 1] = si;
    comp[2 * t] = cr;
    comp[2 * t + 1] = ci;
}

__global__ void advance_step(long long* step_base, int by) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step_base += by;
}

// ---- field energy (engine.py:440-468): one workgroup per (row j, plane k, b), fixed order
struct EnergyArgs {
    int X, Y, Z, B;
    long long bs;
    double inv_hs;
    const void* F[6];
    const uint16_t* idx[3];
    const double* eps;  // relative permittivity per table row, times eps0
    const double* nw[3];
    const double* cw[3];
    double mu0;
    double* partial;
};

template <typename S>
__global__ __launch_bounds__(256) void energy_rows(EnergyArgs A) {
    __shared__ double red[256];
    const int j = blockIdx.x, k = blockIdx.y, b = blockIdx.z;
    const long long q0 = ((long long)k * A.Y + j) * A.X;
    const long long p0 = (long long)b * A.bs + q0;
    double acc = 0.0;
    for (int i = threadIdx.x; i < A.X; i += 256) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            // E_c: cell-type along c, node-type along the others
            const double wx = c == 0 ? A.cw[0][i] : A.nw[0][i];
            const double wy = c == 1 ? A.cw[1][j] : A.nw[1][j];
            const double wz = c == 2 ? A.cw[2][k] : A.nw[2][k];
            const double e = ldS<S, double>(A.F[c], p0 + i);
            acc += 0.5 * A.eps[A.idx[c][q0 + i]] * e * e * (wx * wy * wz);
            // H_c: node-type along c, cell-type along the others
            const double vx = c == 0 ? A.nw[0][i] : A.cw[0][i];
            const double vy = c == 1 ? A.nw[1][j] : A.cw[1][j];
            const double vz = c == 2 ? A.nw[2][k] : A.cw[2][k];
            const double h = ldS<S, double>(A.F[3 + c], p0 + i) * A.inv_hs;
            acc += 0.5 * A.mu0 * h * h * (vx * vy * vz);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) A.partial[((long long)b * A.Z + k) * A.Y + j] = red[0];
}

__global__ __launch_bounds__(256) void sum_partials(const double* partial, long long n, double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---- host transfers of one component (with the H scale undone / applied)
template <typename S>
__global__ __launch_bounds__(256) void to_f64(const void* src, long long n, double scale, double* dst) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) dst[t] = ldS<S, double>(src, t) * scale;
}
template <typename S>
__global__ __launch_bounds__(256) void from_f64(const double* src, long long n, double scale, void* dst) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t < n) stS<S, double>(dst, t, src[t] * scale);
}

}  // namespace oafd
