// Native core of compile_model (host C++, multi-threaded): the per-edge material averaging, the
// merged (eps, sigma) table with the uint16 index arrays, and the Debye edge lists
// (packages/oa-fdtd/src/oa_fdtd/model.py:300-321, 688-722, 331-367). These are the parts that
// walk every edge of the grid; on a 64M-cell grid they are ~90 % of compile time in NumPy.
//
// Every result has the bits of the NumPy formulation: each element is computed with the same
// operations in the same order (the library is built with -ffp-contract=off, so no multiply-add
// is fused), rounding to 12 decimals is rint(x * 1e12) / 1e12 as numpy.round, the table is
// sorted by (eps, sigma) as numpy.unique sorts complex keys, and list order is C order.

#pragma once
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

namespace oafd_compile {

// threads for `items` items of `cost` elements each: at most 16 (one GPU's share of host cores),
// at least 64K elements per thread
inline int n_threads(long long items, long long cost = 1) {
    const unsigned hw = std::thread::hardware_concurrency();
    const long long want = std::max<long long>(1, items * cost / (1 << 16));
    return (int)std::max<long long>(1, std::min<long long>({(long long)(hw ? hw : 1), 16LL, want, items}));
}

// run f(lo, hi) over [0, n) split into contiguous parts, one per thread; cost = elements per item
template <typename F>
void parallel_for(long long n, long long cost, F&& f) {
    const int nt = n_threads(n, cost);
    if (nt <= 1) {
        f(0LL, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        const long long lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([&f, lo, hi]() { f(lo, hi); });
    }
    for (auto& x : th) x.join();
}

// Area-weighted mean of the (up to) four cells around every edge of component comp, on the
// (nz+1, ny+1, nx+1) lattice; cells are (nz+2, ny+2, nx+2); w[a] are the cell weights of axis a
// (length n[a] + 2, zero on ghosts). Accumulation order: (dp, dq) = (0,0), (0,1), (1,0), (1,1)
// with p < q the transverse axes; num += (wp * wq) * cell, den += wp * wq; out = num / den.
inline void edge_average(const int n[3], int comp, const double* const w[3], int m, const double* const* cells,
                         double* const* out) {
    int p = -1, q = -1;
    for (int a = 0; a < 3; ++a)
        if (a != comp) (p < 0 ? p : q) = a;
    const long long cx = n[0] + 2, cy = n[1] + 2;
    const long long lx = n[0] + 1, ly = n[1] + 1, lz = n[2] + 1;
    parallel_for(lz, lx * ly, [&](long long k0, long long k1) {
        std::vector<double> num(m);
        for (long long k = k0; k < k1; ++k)
            for (long long j = 0; j < ly; ++j)
                for (long long i = 0; i < lx; ++i) {
                    const long long t[3] = {i, j, k};
                    double den = 0.0;
                    for (int c = 0; c < m; ++c) num[c] = 0.0;
                    for (int dp = 0; dp < 2; ++dp)
                        for (int dq = 0; dq < 2; ++dq) {
                            long long s[3] = {1, 1, 1};
                            s[p] -= dp;
                            s[q] -= dq;
                            const double ww = w[p][t[p] + 1 - dp] * w[q][t[q] + 1 - dq];
                            const long long cell = ((k + s[2]) * cy + (j + s[1])) * cx + (i + s[0]);
                            for (int c = 0; c < m; ++c) num[c] = num[c] + ww * cells[c][cell];
                            den = den + ww;
                        }
                    const long long o = (k * ly + j) * lx + i;
                    for (int c = 0; c < m; ++c) out[c][o] = num[c] / den;
                }
    });
}

inline double round12(double x) { return std::nearbyint(x * 1e12) / 1e12; }

struct Key {
    double e, s;
    bool operator==(const Key& o) const { return e == o.e && s == o.s; }
    bool operator<(const Key& o) const { return e < o.e || (e == o.e && s < o.s); }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t a, b;
        const double e = k.e == 0.0 ? 0.0 : k.e, s = k.s == 0.0 ? 0.0 : k.s;  // -0 == +0
        std::memcpy(&a, &e, 8);
        std::memcpy(&b, &s, 8);
        return (size_t)(a * 0x9E3779B97F4A7C15ULL ^ (b + 0x7F4A7C159E3779B9ULL + (a << 6) + (a >> 2)));
    }
};

// Sorted unique (round(eps, 12), round(sigma, 12)) over the non-PEC edges of the three
// components, and the uint16 index arrays (storage shape (nz+2, ny+2, pitch), lattice at +1,
// row 0 = PEC). Returns the table size, or -1 when it exceeds max_rows.
inline long long material_table(const int n[3], int pitch, const double* const eps[3], const double* const sig[3],
                                const uint8_t* const pec[3], int max_rows, std::vector<Key>& table,
                                uint16_t* const idx[3]) {
    const long long lx = n[0] + 1, ly = n[1] + 1, lz = n[2] + 1, nl = lx * ly * lz;
    const long long sx = pitch, sy = n[1] + 2;
    // distinct keys per thread, merged
    std::vector<Key> all;
    {
        const int nt = n_threads(nl);
        std::vector<std::vector<Key>> part(nt);
        std::vector<std::thread> th;
        std::atomic<bool> over{false};
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&, t]() {
                std::unordered_map<Key, char, KeyHash> seen;
                const long long lo = nl * t / nt, hi = nl * (t + 1) / nt;
                for (int c = 0; c < 3 && !over; ++c)
                    for (long long e = lo; e < hi; ++e) {
                        if (pec[c][e]) continue;
                        const Key k{round12(eps[c][e]), round12(sig[c][e])};
                        if (seen.emplace(k, 0).second && (long long)seen.size() > max_rows) {
                            over = true;
                            break;
                        }
                    }
                for (auto& kv : seen) part[t].push_back(kv.first);
            });
        for (auto& x : th) x.join();
        if (over) return -1;
        for (auto& v : part) all.insert(all.end(), v.begin(), v.end());
    }
    std::sort(all.begin(), all.end());
    table.clear();
    for (const Key& k : all)
        if (table.empty() || !(table.back() == k)) table.push_back(k);
    if ((long long)table.size() > max_rows) return -1;
    std::unordered_map<Key, uint16_t, KeyHash> row;
    row.reserve(table.size() * 2);
    for (size_t r = 0; r < table.size(); ++r) row.emplace(table[r], (uint16_t)(r + 1));
    for (int c = 0; c < 3; ++c) {
        std::memset(idx[c], 0, sizeof(uint16_t) * (size_t)(n[2] + 2) * sy * sx);
        parallel_for(lz, lx * ly, [&](long long k0, long long k1) {
            for (long long k = k0; k < k1; ++k)
                for (long long j = 0; j < ly; ++j)
                    for (long long i = 0; i < lx; ++i) {
                        const long long e = (k * ly + j) * lx + i;
                        if (pec[c][e]) continue;
                        idx[c][((k + 1) * sy + (j + 1)) * sx + (i + 1)] =
                            row.find(Key{round12(eps[c][e]), round12(sig[c][e])})->second;
                    }
        });
    }
    return (long long)table.size();
}

// Debye edge list of one component: the edges with sum_p del_p > 1e-12 that are not PEC, in C
// order; flat storage index (x pitch `pitch`), beta (P, ne) = EPS0 del_p / (tau_p + dt / 2),
// cbe = cb[idx].
// count pass (out == nullptr) returns ne; the fill pass writes it.
struct AdeOut {
    int64_t* flat;
    double* beta;
    double* cbe;
};
inline long long ade_edges(const int n[3], int pitch, int np, const double* const* del, const uint8_t* pec,
                           const double* taus, double dt, double eps0, const uint16_t* idx, const double* cb_tab,
                           const AdeOut* out) {
    const long long lx = n[0] + 1, ly = n[1] + 1, lz = n[2] + 1;
    const long long sx = pitch, sy = n[1] + 2;
    auto hit = [&](long long e) {
        if (pec[e]) return false;
        double total = 0.0;
        for (int p = 0; p < np; ++p) total = total + del[p][e];
        return total > 1e-12;
    };
    // per-plane counts give each thread its output offset, so the order is C order
    std::vector<long long> cnt(lz + 1, 0);
    parallel_for(lz, lx * ly, [&](long long k0, long long k1) {
        for (long long k = k0; k < k1; ++k) {
            long long c = 0;
            for (long long e = k * ly * lx; e < (k + 1) * ly * lx; ++e) c += hit(e);
            cnt[k + 1] = c;
        }
    });
    for (long long k = 0; k < lz; ++k) cnt[k + 1] += cnt[k];
    const long long ne = cnt[lz];
    if (!out) return ne;
    std::vector<double> fac(np);
    for (int p = 0; p < np; ++p) fac[p] = taus[p] + 0.5 * dt;
    parallel_for(lz, lx * ly, [&](long long k0, long long k1) {
        for (long long k = k0; k < k1; ++k) {
            long long o = cnt[k];
            for (long long j = 0; j < ly; ++j)
                for (long long i = 0; i < lx; ++i) {
                    const long long e = (k * ly + j) * lx + i;
                    if (!hit(e)) continue;
                    const long long s = ((k + 1) * sy + (j + 1)) * sx + (i + 1);
                    out->flat[o] = s;
                    for (int p = 0; p < np; ++p) out->beta[(long long)p * ne + o] = eps0 * del[p][e] / fac[p];
                    out->cbe[o] = cb_tab[idx[s]];
                    ++o;
                }
        }
    });
    return ne;
}

}  // namespace oafd_compile
