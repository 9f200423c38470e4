// Dense complex LU with partial pivoting, triangular solves and FP64
// iterative refinement.
//
// Right-looking blocked LU on a row-major matrix (row = test function), two
// levels of blocking: outer block nb (the rank of the MFMA trailing update),
// inner panels of width w factored by one cooperative kernel.
//   K4 k_panel     cooperative panel factorisation: each workgroup keeps its
//                  rows of the panel in LDS; one grid barrier per column, at
//                  which every workgroup publishes its candidate pivot row so
//                  the winner's row is known to all without a second barrier
//   k_laswp        row interchanges (coalesced: a row is contiguous)
//   K6 k_trsm      unit-lower / upper block triangular solves in LDS, which
//                  also emit the real-expanded operand of the GEMM
//   K5 gemm        MFMA trailing update (gemm.hip)
// Pivot choice is LAPACK's: first row of maximal |re|+|im| (i?amax), so the
// factor and pivots match ?getrf's (the routine behind arma::solve at
// EMSolver/MFIE.cpp:64) up to rounding.
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "gemm.h"
#include "internal.h"

namespace emfa {

constexpr int kPanelMaxW = 64;
constexpr unsigned kSpinLimit = 1u << 22;

// --------------------------------------------------------------------------
// Grid barrier for the cooperative panel: monotonic counter.  Only wave 0 of
// a workgroup writes the data exchanged through global memory, so only wave 0
// needs the release fence; lane 0 arrives, spins (bounded: a timeout sets
// flags[0] and every later barrier falls through, so the kernel always
// exits) and acquires.  The caller continues in wave 0 and synchronises the
// workgroup afterwards.
// --------------------------------------------------------------------------
__device__ __forceinline__ void grid_arrive_wait_wave0(unsigned *bar, unsigned target, int *flags) {
    // The exchanged data is written and read with agent-scope atomic
    // accesses (coherent across the XCDs' L2s), so the arrival only has to
    // wait for those stores to complete; no L2 write-back is needed.
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if ((threadIdx.x & 63) == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

// coherent (agent-scope, relaxed) element stores / loads
__device__ __forceinline__ void st_co(float *p, float v) {
    __hip_atomic_store((unsigned *)p, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_co(const float *p) {
    return __uint_as_float(__hip_atomic_load((unsigned *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ double ld_co(const double *p) {
    return __longlong_as_double(
        (long long)__hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ int ld_co(const int *p) {
    return __hip_atomic_load((int *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename R>
__device__ __forceinline__ void st_co(cx<R> *p, cx<R> v) {
    st_co(&p->re, v.re);
    st_co(&p->im, v.im);
}
template <typename R>
__device__ __forceinline__ cx<R> ld_co(const cx<R> *p) {
    return {ld_co(&p->re), ld_co(&p->im)};
}

template <typename R>
__device__ __forceinline__ R cabs1(cx<R> a) {
    return fabs(a.re) + fabs(a.im);
}

// better(v1,r1) than (v2,r2): larger value, ties -> smaller row; r = -1 means none
template <typename R>
__device__ __forceinline__ bool better(R v1, int r1, R v2, int r2) {
    if (r1 < 0) return false;
    if (r2 < 0) return true;
    if (v1 > v2) return true;
    if (v1 < v2) return false;
    return r1 < r2;
}

template <typename R>
__device__ __forceinline__ void wave_argmax(R &v, int &r) {
    for (int off = 32; off > 0; off >>= 1) {
        R ov = __shfl_xor(v, off, 64);
        int orr = __shfl_xor(r, off, 64);
        if (better(ov, orr, v, r)) {
            v = ov;
            r = orr;
        }
    }
}

// Panel: rows [j0, j0+M) x columns [j0, j0+w) of A (row-major, lda complex).
// Workgroup g owns rows [j0 + g*RB, j0 + min(M, (g+1)*RB)), kept in LDS with a
// padded row stride.  Per column: local i?amax -> wave 0 publishes (value,
// row, row data) -> grid barrier -> wave 0 of every workgroup picks the same
// winner, applies the interchange to its LDS rows -> scale + rank-1 update.
template <typename R>
__global__ __launch_bounds__(256) void k_panel(cx<R> *__restrict__ A, long long lda, int j0, int M, int w, int RB,
                                               int *__restrict__ ipiv, cx<R> *__restrict__ slot_rows,
                                               R *__restrict__ slot_val, int *__restrict__ slot_idx,
                                               cx<R> *__restrict__ rowc, unsigned *bar, unsigned bar_base,
                                               int *flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<R> *rows = (cx<R> *)smem_raw;  // [RB][ws]
    const int ws = w + 1;             // padded stride
    __shared__ cx<R> urow[kPanelMaxW];
    __shared__ R sv[4];
    __shared__ int sr[4];
    __shared__ int s_piv;
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rbeg = j0 + g * RB;
    const int rend = min(j0 + M, rbeg + RB);
    const int nr = max(0, rend - rbeg);
    for (int i = tid; i < nr * w; i += 256) {
        int rr = i / w, cc = i % w;
        rows[rr * ws + cc] = A[(size_t)(rbeg + rr) * lda + j0 + cc];
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
        const int jc = j0 + c;
        const int par = c & 1;
        // local candidate (all waves), combined in wave 0
        R bv = R(-1);
        int br = -1;
        for (int i = tid; i < nr; i += 256) {
            int r = rbeg + i;
            if (r < jc) continue;
            R v = cabs1(rows[i * ws + c]);
            if (better(v, r, bv, br)) {
                bv = v;
                br = r;
            }
        }
        wave_argmax(bv, br);
        if (lane == 0) {
            sv[wave] = bv;
            sr[wave] = br;
        }
        __syncthreads();
        if (wave == 0) {
            bv = sv[0];
            br = sr[0];
            for (int q = 1; q < 4; ++q)
                if (better(sv[q], sr[q], bv, br)) {
                    bv = sv[q];
                    br = sr[q];
                }
            if (lane == 0) {
                st_co(&slot_val[par * G + g], bv);
                st_co(&slot_idx[par * G + g], br);
            }
            if (br >= 0 && lane < w) st_co(&slot_rows[((size_t)par * G + g) * kPanelMaxW + lane], rows[(br - rbeg) * ws + lane]);
            if (jc >= rbeg && jc < rend && lane < w) st_co(&rowc[par * kPanelMaxW + lane], rows[(jc - rbeg) * ws + lane]);
            grid_arrive_wait_wave0(bar, bar_base + (unsigned)(c + 1) * G, flags);
            // winner: every workgroup computes the same (value, row)
            R wv = R(-1);
            int wr = -1;
            for (int q = lane; q < G; q += 64) {
                R v = ld_co(&slot_val[par * G + q]);
                int r = ld_co(&slot_idx[par * G + q]);
                if (better(v, r, wv, wr)) {
                    wv = v;
                    wr = r;
                }
            }
            wave_argmax(wv, wr);
            const bool singular = !(wv > R(0));
            const int p = singular ? jc : wr;
            const int wg = (p - j0) / RB;
            if (lane < w) {
                cx<R> o = ld_co(&rowc[par * kPanelMaxW + lane]);
                cx<R> u = singular ? o : ld_co(&slot_rows[((size_t)par * G + wg) * kPanelMaxW + lane]);
                urow[lane] = u;
                if (jc >= rbeg && jc < rend) rows[(jc - rbeg) * ws + lane] = u;
                if (p != jc && p >= rbeg && p < rend) rows[(p - rbeg) * ws + lane] = o;
            }
            if (lane == 0) {
                s_piv = p;
                if (g == 0) {
                    ipiv[jc] = p;
                    if (singular && flags[1] == 0) flags[1] = jc + 1;
                }
            }
        }
        __syncthreads();
        // scale the column (one thread per row), then rank-1 update of the rest
        const cx<R> piv = urow[c];
        const R den = piv.re * piv.re + piv.im * piv.im;
        const int first = max(jc + 1, rbeg) - rbeg;
        if (den > R(0)) {
            const cx<R> inv = {piv.re / den, -piv.im / den};
            for (int i = first + tid; i < nr; i += 256) rows[i * ws + c] = cmul(rows[i * ws + c], inv);
        }
        __syncthreads();
        const int ncol = w - c - 1;
        if (ncol > 0 && first < nr) {
            const int nupd = (nr - first) * ncol;
            for (int e = tid; e < nupd; e += 256) {
                int i = first + e / ncol, cc = c + 1 + e % ncol;
                cx<R> l = rows[i * ws + c];
                cx<R> u = urow[cc];
                cx<R> &a = rows[i * ws + cc];
                a.re -= l.re * u.re - l.im * u.im;
                a.im -= l.re * u.im + l.im * u.re;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < nr * w; i += 256) {
        int rr = i / w, cc = i % w;
        A[(size_t)(rbeg + rr) * lda + j0 + cc] = rows[rr * ws + cc];
    }
}

// --------------------------------------------------------------------------
// Panel with a packet exchange (default).  Same arithmetic and pivot choice
// as k_panel; what changes is how workgroups agree on the pivot of a column:
// instead of an arrival counter plus separate value/row reads, every
// workgroup writes one 64-bit packet per column into its own slot (double
// buffered by column parity) and every workgroup polls the G packets, so the
// barrier and the candidate values travel together.
//   complex FP32: packet = value bits | row << 32 | (epoch & 0xFFF) << 52;
//     candidate rows are 64-bit words {float bits, epoch}, two per element,
//     so a reader can tell a fresh word from a stale one and the writer never
//     waits for its stores.
//   complex FP64: packet = row | epoch << 32; the value and the row are
//     written first and ordered before the packet by s_waitcnt.
// The epoch of column jc is jc + 1; the slots are zeroed before each
// factorisation.  Rows live in LDS column-major (stride RB + 1) and one
// thread owns a row in the update, which also yields the next column's local
// candidate, so a column costs two workgroup barriers.
// --------------------------------------------------------------------------
constexpr int kPanelMaxG = 256;

__device__ __forceinline__ void st64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld64(const unsigned long long *p) {
    return __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long tag32(float x, unsigned ep) {
    return (unsigned long long)__float_as_uint(x) | ((unsigned long long)ep << 32);
}

// FP32 self-validating pair load; false if the bounded spin ran out
__device__ __forceinline__ bool ld_tagged(const unsigned long long *p, unsigned ep, cx<float> *out, bool active,
                                          int *flags) {
    unsigned spins = 0;
    bool ok = !active;
    for (;;) {
        if (!ok) {
            unsigned long long a = ld64(p), b = ld64(p + 1);
            if ((unsigned)(a >> 32) == ep && (unsigned)(b >> 32) == ep) {
                *out = {__uint_as_float((unsigned)a), __uint_as_float((unsigned)b)};
                ok = true;
            }
        }
        if (__ballot(!ok) == 0) return true;
        if (++spins > kSpinLimit) {
            __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_panel_x(cx<R> *__restrict__ A, long long lda, int j0, int M, int w, int RB,
                                                 int *__restrict__ ipiv, unsigned long long *pk,
                                                 unsigned long long *rw, unsigned long long *cw, R *sval,
                                                 int *flags) {
    constexpr bool kF32 = sizeof(R) == 4;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<R> *cols = (cx<R> *)smem_raw;  // [w][RBp]
    const int RBp = RB + 1;
    __shared__ cx<R> urow[kPanelMaxW];
    __shared__ R sv[4];
    __shared__ int sr[4];
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rbeg = j0 + g * RB;
    const int rend = min(j0 + M, rbeg + RB);
    const int nr = max(0, rend - rbeg);
    for (int e = tid; e < nr * w; e += 256) {
        int rr = e / w, cc = e - rr * w;
        cols[cc * RBp + rr] = A[(size_t)(rbeg + rr) * lda + j0 + cc];
    }
    __syncthreads();
    R bv = R(-1);
    int br = -1;
    for (int i = tid; i < nr; i += 256) {
        R v = cabs1(cols[i]);
        if (better(v, rbeg + i, bv, br)) {
            bv = v;
            br = rbeg + i;
        }
    }
    bool dead = false;  // a bounded wait ran out: finish without exchanging
    for (int c = 0; c < w; ++c) {
        const int jc = j0 + c, par = c & 1;
        const unsigned ep = (unsigned)jc + 1u;
        wave_argmax(bv, br);
        if (lane == 0) {
            sv[wave] = bv;
            sr[wave] = br;
        }
        __syncthreads();
        if (wave == 0) {
            bv = sv[0];
            br = sr[0];
            for (int q = 1; q < 4; ++q)
                if (better(sv[q], sr[q], bv, br)) {
                    bv = sv[q];
                    br = sr[q];
                }
            const bool own_jc = jc >= rbeg && jc < rend;
            // publish this workgroup's candidate
            if (kF32) {
                if (br >= 0 && lane < w) {
                    cx<R> x = cols[lane * RBp + (br - rbeg)];
                    unsigned long long *d = rw + ((size_t)par * G + g) * (2 * kPanelMaxW) + 2 * lane;
                    st64(d, tag32((float)x.re, ep));
                    st64(d + 1, tag32((float)x.im, ep));
                }
                if (own_jc && lane < w) {
                    cx<R> x = cols[lane * RBp + (jc - rbeg)];
                    unsigned long long *d = cw + (size_t)par * (2 * kPanelMaxW) + 2 * lane;
                    st64(d, tag32((float)x.re, ep));
                    st64(d + 1, tag32((float)x.im, ep));
                }
                if (lane == 0) {
                    unsigned idx = br >= 0 ? (unsigned)br : 0xFFFFFu;
                    float v = br >= 0 ? (float)bv : 0.f;
                    st64(&pk[par * kPanelMaxG + g], (unsigned long long)__float_as_uint(v) |
                                                        ((unsigned long long)idx << 32) |
                                                        ((unsigned long long)(ep & 0xFFFu) << 52));
                }
            } else {
                cx<R> *rwc = (cx<R> *)rw, *cwc = (cx<R> *)cw;
                if (br >= 0 && lane < w)
                    st_co(&rwc[((size_t)par * G + g) * kPanelMaxW + lane], cols[lane * RBp + (br - rbeg)]);
                if (own_jc && lane < w) st_co(&cwc[(size_t)par * kPanelMaxW + lane], cols[lane * RBp + (jc - rbeg)]);
                if (lane == 0) st_co(&sval[par * kPanelMaxG + g], bv);
                __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
                if (lane == 0)
                    st64(&pk[par * kPanelMaxG + g],
                         (unsigned long long)(br >= 0 ? (unsigned)br : 0xFFFFFFFFu) | ((unsigned long long)ep << 32));
            }
            // gather the G packets (up to four per lane)
            unsigned long long pkv[4] = {0, 0, 0, 0};
            int okm = 0;
            unsigned spins = 0;
            while (!dead) {
                bool all = true;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int q = lane + 64 * k;
                    if (q < G && !((okm >> k) & 1)) {
                        unsigned long long v = ld64(&pk[par * kPanelMaxG + q]);
                        bool ok = kF32 ? (unsigned)(v >> 52) == (ep & 0xFFFu) : (unsigned)(v >> 32) == ep;
                        if (ok) {
                            pkv[k] = v;
                            okm |= 1 << k;
                        } else {
                            all = false;
                        }
                    }
                }
                if (__ballot(!all) == 0) break;
                if (++spins > kSpinLimit ||
                    __builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0) {
                    __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    dead = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            R wv = R(-1);
            int wr = -1;
            if (!dead) {
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int q = lane + 64 * k;
                    if (q < G) {
                        R v;
                        int r;
                        if (kF32) {
                            unsigned idx = (unsigned)(pkv[k] >> 32) & 0xFFFFFu;
                            r = idx == 0xFFFFFu ? -1 : (int)idx;
                            v = (R)__uint_as_float((unsigned)pkv[k]);
                        } else {
                            unsigned idx = (unsigned)pkv[k];
                            r = idx == 0xFFFFFFFFu ? -1 : (int)idx;
                            v = r >= 0 ? ld_co(&sval[par * kPanelMaxG + q]) : R(-1);
                        }
                        if (better(v, r, wv, wr)) {
                            wv = v;
                            wr = r;
                        }
                    }
                }
            }
            wave_argmax(wv, wr);
            const bool singular = !(wv > R(0));
            int p = singular ? jc : wr;
            if (p < j0 || p >= j0 + M) p = jc;  // only after a failed wait
            const int wg = (p - j0) / RB;
            cx<R> u = {R(0), R(0)}, o = {R(0), R(0)};
            if (!dead) {
                if (kF32) {
                    cx<float> t;
                    bool ok = ld_tagged(cw + (size_t)par * (2 * kPanelMaxW) + 2 * lane, ep, &t, lane < w, flags);
                    o = {(R)t.re, (R)t.im};
                    if (ok && p != jc) {
                        ok = ld_tagged(rw + ((size_t)par * G + wg) * (2 * kPanelMaxW) + 2 * lane, ep, &t, lane < w,
                                       flags);
                        u = {(R)t.re, (R)t.im};
                    } else {
                        u = o;
                    }
                    if (!ok) dead = true;
                } else if (lane < w) {
                    o = ld_co(&((cx<R> *)cw)[(size_t)par * kPanelMaxW + lane]);
                    u = p != jc ? ld_co(&((cx<R> *)rw)[((size_t)par * G + wg) * kPanelMaxW + lane]) : o;
                }
            }
            if (lane < w) {
                urow[lane] = u;
                if (jc >= rbeg && jc < rend) cols[lane * RBp + (jc - rbeg)] = u;
                if (p != jc && p >= rbeg && p < rend) cols[lane * RBp + (p - rbeg)] = o;
            }
            if (lane == 0 && g == 0) {
                ipiv[jc] = p;
                if (singular && !dead && flags[1] == 0) flags[1] = jc + 1;
            }
        }
        __syncthreads();
        // scale column c, rank-1 update, next column's local candidate: one thread per row
        const cx<R> piv = urow[c];
        const R den = piv.re * piv.re + piv.im * piv.im;
        const bool scale = den > R(0);
        cx<R> inv = {R(0), R(0)};
        if (scale) inv = {piv.re / den, -piv.im / den};
        const int first = max(jc + 1, rbeg) - rbeg;
        // tpr threads share a row (a power of two <= 64, so a row's threads sit
        // in one wave and read the multiplier before its owner writes it back)
        int tpr = 1;
        while (tpr < 64 && tpr * 2 * (nr - first) <= 256) tpr *= 2;
        const int jr = tid & (tpr - 1);
        bv = R(-1);
        br = -1;
        for (int i = first + tid / tpr; i < nr; i += 256 / tpr) {
            cx<R> l = cols[c * RBp + i];
            if (scale) l = cmul(l, inv);
            for (int cc = c + 1 + jr; cc < w; cc += tpr) {
                const cx<R> uu = urow[cc];
                cx<R> &a = cols[cc * RBp + i];
                a.re -= l.re * uu.re - l.im * uu.im;
                a.im -= l.re * uu.im + l.im * uu.re;
            }
            if (jr == 0) {
                if (scale) cols[c * RBp + i] = l;
                if (c + 1 < w) {
                    R v = cabs1(cols[(c + 1) * RBp + i]);
                    if (better(v, rbeg + i, bv, br)) {
                        bv = v;
                        br = rbeg + i;
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < nr * w; e += 256) {
        int rr = e / w, cc = e - rr * w;
        A[(size_t)(rbeg + rr) * lda + j0 + cc] = cols[cc * RBp + rr];
    }
}

// Row interchanges k0..k1-1 (sequential, in order) on columns [c0,c1) and [c2,c3).
template <typename R>
__global__ __launch_bounds__(256) void k_laswp(cx<R> *__restrict__ A, long long lda, const int *__restrict__ ipiv,
                                               int k0, int k1, int c0, int c1, int c2, int c3) {
    int t = blockIdx.x * 256 + threadIdx.x;
    int n1 = c1 - c0;
    int col;
    if (t < n1)
        col = c0 + t;
    else if (t - n1 < c3 - c2)
        col = c2 + (t - n1);
    else
        return;
    for (int k = k0; k < k1; ++k) {
        int p = ipiv[k];
        if (p != k) {
            cx<R> a = A[(size_t)k * lda + col];
            A[(size_t)k * lda + col] = A[(size_t)p * lda + col];
            A[(size_t)p * lda + col] = a;
        }
    }
}

// Triangular solve on a block of wb <= 64 rows starting at row r0 of X
// (columns [c0,c1)), with the wb x wb diagonal block of the factor at (r0, r0):
//   upper == 0:  X <- L^{-1} X, L unit lower
//   upper == 1:  X <- U^{-1} X, U upper (non-unit)
// The diagonal block is staged transposed in LDS; each wavefront solves whole
// columns with one lane per row (the pivot value is broadcast by a shuffle),
// so no workgroup barrier sits inside the recurrence.  Each workgroup covers
// 32 columns.  If Bt is given, the solved rows are also written in
// real-expanded transposed form (see gemm.h) at K offset kofs:
//   Bt[2j][2(kofs+r)] = Re, [2j][..+1] = -Im, [2j+1][..] = Im, [2j+1][..+1] = Re.
// Each wavefront keeps CPW columns in flight so that the dependent
// broadcast/FMA chains of independent columns interleave.
constexpr int kTrsmCPW = 2;
constexpr int kTrsmColsPerPass = 4 * kTrsmCPW;

template <typename R>
__device__ __forceinline__ R bcast(R v, int src);
template <>
__device__ __forceinline__ float bcast<float>(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
template <>
__device__ __forceinline__ double bcast<double>(double v, int src) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// cols_per_wg = 4 CPW passes; workgroup `bx` of the solve.  CPW columns per
// wave: more interleave more independent chains per wave and need fewer
// workgroups (less CU time in all), at a longer time per workgroup.
template <typename R, int CPW = kTrsmCPW>
__device__ __forceinline__ void trsm64_body(const cx<R> *__restrict__ F, long long ldf, cx<R> *__restrict__ X,
                                            long long ldx, int r0, int wb, int c0, int c1, int upper,
                                            R *__restrict__ Bt, long long ldbt, int kofs, int passes, int bx,
                                            cx<R> (*Lt)[65]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < wb * wb; e += 256) {
        int r = e / wb, i = e % wb;
        Lt[i][r] = F[(size_t)(r0 + r) * ldf + r0 + i];
    }
    __syncthreads();
    const int cols_per_wg = (4 * CPW) * passes;
    const int cb = c0 + bx * cols_per_wg;
    for (int ps = 0; ps < passes; ++ps) {
        int col[CPW];
        bool ok[CPW];
        cx<R> x[CPW];
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            col[q] = cb + ps * (4 * CPW) + wave * CPW + q;
            ok[q] = col[q] < c1;
            x[q] = (ok[q] && lane < wb) ? X[(size_t)(r0 + lane) * ldx + col[q]] : cx<R>{R(0), R(0)};
        }
        if (!upper) {
            for (int i = 0; i < wb - 1; ++i) {
                cx<R> l = Lt[i][lane];
                const bool act = lane > i && lane < wb;
#pragma unroll
                for (int q = 0; q < CPW; ++q) {
                    R xr = bcast(x[q].re, i), xi = bcast(x[q].im, i);
                    if (act) {
                        x[q].re -= l.re * xr - l.im * xi;
                        x[q].im -= l.re * xi + l.im * xr;
                    }
                }
            }
        } else {
            for (int i = wb - 1; i >= 0; --i) {
                cx<R> d = Lt[i][i];
                R den = d.re * d.re + d.im * d.im;
                cx<R> inv = den > R(0) ? cx<R>{d.re / den, -d.im / den} : cx<R>{R(0), R(0)};
                cx<R> u = Lt[i][lane];
#pragma unroll
                for (int q = 0; q < CPW; ++q) {
                    if (lane == i) x[q] = cmul(x[q], inv);
                    R xr = bcast(x[q].re, i), xi = bcast(x[q].im, i);
                    if (lane < i) {
                        x[q].re -= u.re * xr - u.im * xi;
                        x[q].im -= u.re * xi + u.im * xr;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < CPW; ++q) {
            if (!ok[q] || lane >= wb) continue;
            X[(size_t)(r0 + lane) * ldx + col[q]] = x[q];
            if (Bt) {
                long long j = (long long)(col[q] - c0);
                R *b0 = Bt + (2 * j) * ldbt + 2 * (kofs + lane);
                R *b1 = Bt + (2 * j + 1) * ldbt + 2 * (kofs + lane);
                b0[0] = x[q].re;
                b0[1] = -x[q].im;
                b1[0] = x[q].im;
                b1[1] = x[q].re;
            }
        }
    }
}

// With a batch, F and X are in the batch's matrices and Bt in its scratch.
template <typename R, int CPW>
__global__ __launch_bounds__(256) void k_trsm64(const cx<R> *__restrict__ F, long long ldf, cx<R> *__restrict__ X,
                                                long long ldx, int r0, int wb, int c0, int c1, int upper,
                                                R *__restrict__ Bt, long long ldbt, int kofs, int passes,
                                                Batch bt = Batch()) {
    __shared__ cx<R> Lt[64][65];
    trsm64_body<R, CPW>(batch_mat(bt, F), ldf, batch_mat(bt, X), ldx, r0, wb, c0, c1, upper,
                        Bt ? batch_scr(bt, Bt) : Bt, ldbt, kofs, passes, blockIdx.x, Lt);
}

// --------------------------------------------------------------------------
// Triangular solves in one launch per direction (look-back), for up to 64
// right-hand sides.  Workgroup w owns 64-row block b (b = w forward, the
// mirror backward) and all columns of the chunk: it folds in the solved
// blocks in dependency order as their flags appear (L tile from the factor,
// x tile through coherent loads), then solves its diagonal block with one
// lane per row (v_readlane broadcast), writes x with coherent stores and
// raises its flag.  Blocks depend only on lower workgroup ids, which the
// dispatcher starts first, so every wait ends; the spin is bounded anyway
// (flags[0] reports a timeout).  Forward reads the right-hand side through
// the factor's row permutation, so no separate interchange pass is needed.
// --------------------------------------------------------------------------
constexpr int kLbCols = 64;

template <typename R, bool UPPER>
__global__ __launch_bounds__(256) void k_trsv_lb(const cx<R> *__restrict__ F, long long ldf, int N,
                                                 const cx<R> *__restrict__ in, long long ldi,
                                                 const int *__restrict__ perm, cx<R> *out, long long ldo, int c0,
                                                 int mc, int *bflags, int flagval, int *flags) {
    __shared__ cx<R> Lt[64][65];  // [row][k]
    __shared__ cx<R> Xs[64][kLbCols];
    __shared__ int s_dead;
    const int nblk = (N + 63) / 64;
    const int wg = blockIdx.x;
    const int blk = UPPER ? nblk - 1 - wg : wg;
    const int r0 = blk * 64, nr = min(64, N - r0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int CPT = kLbCols / 4;  // columns per thread: wave + 4 i
    cx<R> acc[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int c = wave + 4 * i;
        acc[i] = {R(0), R(0)};
        if (lane < nr && c < mc) {
            const int row = perm ? perm[r0 + lane] : r0 + lane;
            acc[i] = in[(size_t)row * ldi + c0 + c];
        }
    }
    if (tid == 0) s_dead = 0;
    for (int jj = 0; jj < wg; ++jj) {
        const int jb = UPPER ? nblk - 1 - jj : jj;
        const int k0 = jb * 64, nk = min(64, N - k0);
        // stage the factor tile while the producer may still be running
        for (int e = tid; e < 64 * 64; e += 256) {
            const int rr = e >> 6, cc = e & 63;
            Lt[rr][cc] = (rr < nr && cc < nk) ? F[(size_t)(r0 + rr) * ldf + k0 + cc] : cx<R>{R(0), R(0)};
        }
        if (tid == 0) {
            unsigned spins = 0;
            while (__hip_atomic_load(&bflags[jb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != flagval) {
                if (__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                    ++spins > kSpinLimit) {
                    __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_dead = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (s_dead) break;
        for (int e = tid; e < 64 * kLbCols; e += 256) {
            const int kk = e / kLbCols, cc = e % kLbCols;
            Xs[kk][cc] = (kk < nk && cc < mc) ? ld_co(&out[(size_t)(k0 + kk) * ldo + c0 + cc]) : cx<R>{R(0), R(0)};
        }
        __syncthreads();
        for (int k = 0; k < nk; ++k) {
            const cx<R> l = Lt[lane][k];
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                const cx<R> x = Xs[k][wave + 4 * i];
                acc[i].re -= l.re * x.re - l.im * x.im;
                acc[i].im -= l.re * x.im + l.im * x.re;
            }
        }
        __syncthreads();
    }
    // diagonal block
    for (int e = tid; e < 64 * 64; e += 256) {
        const int rr = e >> 6, cc = e & 63;
        Lt[rr][cc] = (rr < nr && cc < nr) ? F[(size_t)(r0 + rr) * ldf + r0 + cc] : cx<R>{R(0), R(0)};
    }
    __syncthreads();
    if (!UPPER) {
        for (int k = 0; k < nr - 1; ++k) {
            const cx<R> l = Lt[lane][k];
            const bool act = lane > k && lane < nr;
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                const R xr = bcast(acc[i].re, k), xi = bcast(acc[i].im, k);
                if (act) {
                    acc[i].re -= l.re * xr - l.im * xi;
                    acc[i].im -= l.re * xi + l.im * xr;
                }
            }
        }
    } else {
        for (int k = nr - 1; k >= 0; --k) {
            const cx<R> d = Lt[k][k];
            const R den = d.re * d.re + d.im * d.im;
            const cx<R> inv = den > R(0) ? cx<R>{d.re / den, -d.im / den} : cx<R>{R(0), R(0)};
            const cx<R> u = Lt[lane][k];
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                if (lane == k) acc[i] = cmul(acc[i], inv);
                const R xr = bcast(acc[i].re, k), xi = bcast(acc[i].im, k);
                if (lane < k) {
                    acc[i].re -= u.re * xr - u.im * xi;
                    acc[i].im -= u.re * xi + u.im * xr;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int c = wave + 4 * i;
        if (lane < nr && c < mc) st_co(&out[(size_t)(r0 + lane) * ldo + c0 + c], acc[i]);
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&bflags[blk], flagval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Row interchanges as a permutation gather.  k_perm_build replays the
// interchanges k0..k1-1 (ipiv, LAPACK order) on the index set they touch with
// one wavefront and emits (destination row, source row) pairs; k_perm_gather
// copies the source rows of a column range to a staging buffer and
// k_perm_scatter writes them to their destinations.  Every element moves
// once, with no dependent chain of memory accesses.
constexpr int kPermChunkMax = 256;

__global__ __launch_bounds__(64) void k_perm_build(const int *__restrict__ ipiv, int k0, int k1, int *__restrict__ dst,
                                                   int *__restrict__ src, int *__restrict__ count) {
    __shared__ int in_src[512];     // origin of row k0 + i, i < k1 - k0 (<= 256)
    __shared__ int out_row[256], out_src[256];
    __shared__ int n_out;
    __shared__ int piv[kPermChunkMax];
    const int lane = threadIdx.x;
    const int nb = k1 - k0;
    for (int i = lane; i < nb; i += 64) {
        in_src[i] = k0 + i;
        piv[i] = ipiv[k0 + i];
    }
    if (lane == 0) n_out = 0;
    __syncthreads();
    for (int k = k0; k < k1; ++k) {
        const int p = piv[k - k0];
        if (p == k) continue;
        if (p < k1) {
            if (lane == 0) {
                int t = in_src[k - k0];
                in_src[k - k0] = in_src[p - k0];
                in_src[p - k0] = t;
            }
        } else {
            // find p among the rows outside the block touched so far
            int found = -1;
            for (int j = lane; j < n_out; j += 64)
                if (out_row[j] == p) found = j;
            unsigned long long m = __ballot(found >= 0);
            int idx;
            if (m) {
                int srcl = __ffsll((long long)m) - 1;
                idx = __shfl(found, srcl, 64);
            } else {
                idx = n_out;
            }
            if (lane == 0) {
                if (!m) {
                    out_row[idx] = p;
                    out_src[idx] = p;
                    n_out = idx + 1;
                }
                int t = in_src[k - k0];
                in_src[k - k0] = out_src[idx];
                out_src[idx] = t;
            }
        }
        __syncthreads();
    }
    // emit non-identity moves
    __shared__ int n_emit;
    if (lane == 0) n_emit = 0;
    __syncthreads();
    if (lane == 0) {
        int c = 0;
        for (int i = 0; i < nb; ++i)
            if (in_src[i] != k0 + i) {
                dst[c] = k0 + i;
                src[c] = in_src[i];
                ++c;
            }
        for (int j = 0; j < n_out; ++j)
            if (out_src[j] != out_row[j]) {
                dst[c] = out_row[j];
                src[c] = out_src[j];
                ++c;
            }
        n_emit = c;
        *count = c;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_perm_gather(const cx<R> *__restrict__ A, long long lda,
                                                     const int *__restrict__ src, const int *__restrict__ count,
                                                     int c0, int c1, int c2, int c3, cx<R> *__restrict__ stage) {
    const int i = blockIdx.y;
    if (i >= *count) return;
    const int n1 = c1 - c0, ncols = n1 + (c3 - c2);
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ncols) return;
    const int col = t < n1 ? c0 + t : c2 + (t - n1);
    stage[(size_t)i * ncols + t] = A[(size_t)src[i] * lda + col];
}

template <typename R>
__global__ __launch_bounds__(256) void k_perm_scatter(cx<R> *__restrict__ A, long long lda,
                                                      const int *__restrict__ dst, const int *__restrict__ count,
                                                      int c0, int c1, int c2, int c3,
                                                      const cx<R> *__restrict__ stage) {
    const int i = blockIdx.y;
    if (i >= *count) return;
    const int n1 = c1 - c0, ncols = n1 + (c3 - c2);
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= ncols) return;
    const int col = t < n1 ? c0 + t : c2 + (t - n1);
    A[(size_t)dst[i] * lda + col] = stage[(size_t)i * ncols + t];
}

// Real-expanded transposed copy of a whole complex matrix X (n x m) into
// Bt (2m x 2n), used by the residual product.
template <typename R>
__global__ __launch_bounds__(256) void k_expand(const cx<R> *__restrict__ X, int n, int m, R *__restrict__ Bt) {
    long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long long)n * m) return;
    int j = (int)(id / n), k = (int)(id % n);  // k fastest -> coalesced writes
    cx<R> v = X[(size_t)k * m + j];
    long long ldb = 2LL * n;
    R *b0 = Bt + (2LL * j) * ldb + 2 * k;
    R *b1 = Bt + (2LL * j + 1) * ldb + 2 * k;
    b0[0] = v.re;
    b0[1] = -v.im;
    b1[0] = v.im;
    b1[1] = v.re;
}

__global__ void k_zero_u32(unsigned *p, int n, Batch bt = Batch()) {
    p = batch_scr(bt, p);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Deterministic test matrix for the autotuner: entries uniform in
// [-0.5, 0.5] (integer hash of the index) plus n on the diagonal.
template <typename R>
__global__ __launch_bounds__(256) void k_tune_matrix(cx<R> *__restrict__ A, int n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long long)n * n) return;
    unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    const R re = (R)((double)(h & 0xFFFFFF) / 16777216.0 - 0.5);
    const R im = (R)((double)((h >> 24) & 0xFFFFFF) / 16777216.0 - 0.5);
    const bool diag = (i / n) == (i % n);
    A[i] = {re + (diag ? (R)n : R(0)), im};
}

// ------------------------------------------------------------------ host
template <typename R>
struct LuPlan {
    int nb = 256;  // outer block (GEMM rank)
    int w = 64;    // panel width
};

// C (complex, M x ncols, ldc) -= Aop (M x K, lda) * B, with B given as its
// real-expanded transposed form Bt (ldb reals per row); with negate = 0 and
// accumulate = 0, C = Aop * B.
template <typename R>
static emfa_status gemm_update(emfa_ctx *ctx, const char *fam, int M, int ncols, int K, const cx<R> *Aop,
                               long long lda, const R *Bt, long long ldb, cx<R> *C, long long ldc,
                               hipStream_t stream = nullptr, int negate = 1, int accumulate = 1) {
    hipStream_t s = stream ? stream : ctx->stream;
    if (M <= 0 || ncols <= 0 || K <= 0) return EMFA_OK;
    GemmArgs g;
    g.M = M;
    g.Nr = 2 * ncols;
    g.Kr = 2 * K;
    g.A = Aop;
    g.lda = 2 * lda;
    g.Bt = Bt;
    g.ldb = ldb;
    g.C = C;
    g.ldc = 2 * ldc;
    g.alpha_neg = negate;
    g.beta = accumulate;
    g.bt = ctx->batch;
    ProfScope ps(ctx, fam, 8.0 * (double)M * ncols * K, s);
    const int tag = std::strcmp(fam, "lu_gemm") == 0 ? 1 : 0;
    if constexpr (sizeof(R) == 4)
        return gemm_launch_f32(s, g, tag);
    else
        return gemm_launch_f64(s, g, tag);
}

// Operand planes of the three-term bf16 trailing update (FP32 factorisations):
// the outer panel's L block (M x 2jb reals) and expanded U rows (2(N-next) x
// 2jb reals), split once per outer step and read by both update halves.
struct UpdatePlanes {
    __bf16 *ah = nullptr, *al = nullptr, *bh = nullptr, *bl = nullptr;
    int a_rows = 0, b_rows = 0, kp = 0;  // allocated rows per plane, K stride
    static int rup(int a, int b) { return (a + b - 1) / b * b; }
    static size_t bytes(int N, int nb) {
        const size_t kp = (size_t)rup(2 * nb, 16);
        return 2 * kp * sizeof(__bf16) * ((size_t)rup(N, 256) + (size_t)rup(2 * N, 256) + 256) + 1024;
    }
    void carve(void *p, int N, int nb) {
        kp = rup(2 * nb, 16);
        a_rows = rup(N, 256);
        b_rows = rup(2 * N, 256) + 256;
        char *q = (char *)p;
        auto take = [&](size_t b) {
            char *r = q;
            q += (b + 255) & ~(size_t)255;
            return (__bf16 *)r;
        };
        ah = take((size_t)a_rows * kp * sizeof(__bf16));
        al = take((size_t)a_rows * kp * sizeof(__bf16));
        bh = take((size_t)b_rows * kp * sizeof(__bf16));
        bl = take((size_t)b_rows * kp * sizeof(__bf16));
    }
    // split L (M x K complex at Aop, lda) and the U rows (2*ncols x 2K reals, ldb)
    emfa_status split(emfa_ctx *ctx, int M, int ncols, int K, const cx<float> *Aop, long long lda, const float *Bt,
                      long long ldb) {
        ProfScope ps(ctx, "lu_split", 0);
        const int kr = 2 * K, k16 = rup(kr, 16);
        emfa_status st = split_bf16_launch(ctx->stream, (const float *)Aop, 2 * lda, M, kr, ah, al, kp,
                                           rup(M, 256), k16, ctx->batch, true);
        if (st) return st;
        return split_bf16_launch(ctx->stream, Bt, ldb, 2 * ncols, kr, bh, bl, kp,
                                 std::min(b_rows, rup(2 * ncols, 256) + 256), k16, ctx->batch, false);
    }
    // C (M x ncols complex) -= L * U[:, c0 : c0 + ncols]
    emfa_status update(emfa_ctx *ctx, int M, int c0, int ncols, int K, cx<float> *C, long long ldc,
                       hipStream_t s) const {
        if (M <= 0 || ncols <= 0 || K <= 0) return EMFA_OK;
        ProfScope ps(ctx, "lu_gemm_split", 8.0 * (double)M * ncols * K, s);
        Bf3Planes g;
        g.M = M;
        g.Nr = 2 * ncols;
        g.Kp = rup(2 * K, 16);
        g.Ah = ah;
        g.Al = al;
        g.lda = kp;
        g.Bh = bh + (size_t)2 * c0 * kp;
        g.Bl = bl + (size_t)2 * c0 * kp;
        g.ldb = kp;
        g.C = (float *)C;
        g.ldc = 2 * ldc;
        g.alpha_neg = 1;
        g.beta = 1;
        g.bt = ctx->batch;
        // contexts sharing the device keep each update to a multiple of their
        // CU share: twice the share lets the large factors' updates take the
        // CUs that the other workers' latency-bound kernels leave idle
        // (served 64-request stream, medians of three passes: 116.1 solves/s
        // at one share, 118.1 at two, 117.6 at four)
        if (ctx->cu_share > 0) {
            int cap = 2;
            if (const char *e = getenv("EMFA_GEMM_SHARE_CAP")) cap = atoi(e);
            // a batch of factorisations does the work of that many workers
            const int share = ctx->cu_share;
            if (cap > 0) g.max_wgs = std::min(ctx->num_cu, cap * share * std::max(1, ctx->batch.count));
        }
        return gemm_launch_bf3_planes(s, g);
    }
};

// Blocked triangular solve of rows [r0, r0+wb) of X (columns [c0,c1)) with the
// diagonal block of F: 64-row diagonal solves (k_trsm64) + MFMA updates of
// the remaining rows of the block.  Writes the expanded operand of the whole
// block into Bt (2*(c1-c0) x 2*wb reals, row stride 2*wb).
template <typename R>
static emfa_status trsm_blocked(emfa_ctx *ctx, const cx<R> *F, long long ldf, cx<R> *X, long long ldx, int r0, int wb,
                                int c0, int c1, int upper, R *Bt, const char *fam) {
    if (c1 <= c0 || wb <= 0) return EMFA_OK;
    const long long ldbt = 2LL * wb;
    const int ncols = c1 - c0;
    // columns per wave: 8 for the LU's U rows in contexts that share the
    // device (a solve server: fewer workgroups, less CU time while the others
    // wait for it), else 2 (shortest time per workgroup); EMFA_TRSM_CPW sets it
    int wide = ctx->cu_share > 0 && std::strcmp(fam, "lu_trsm") == 0 ? 8 : 2;
    if (const char *e = getenv("EMFA_TRSM_CPW")) wide = atoi(e) == 8 ? 8 : 2;
    const int cpp = 4 * wide;  // columns per pass of a workgroup
    // enough workgroups to fill the chip; more columns per workgroup when wide
    int passes = std::max(1, std::min(8, ncols / (cpp * 1024)));
    const int cpw = cpp * passes;
    const int blocks = (ncols + cpw - 1) / cpw;
    auto trsm = [&](int r, int b, int up, int kofs) {
        const dim3 g(blocks, 1, ctx->batch.count);
        if (wide == 8)
            k_trsm64<R, 8><<<g, 256, 0, ctx->stream>>>(F, ldf, X, ldx, r, b, c0, c1, up, Bt, ldbt, kofs, passes,
                                                      ctx->batch);
        else
            k_trsm64<R, 2><<<g, 256, 0, ctx->stream>>>(F, ldf, X, ldx, r, b, c0, c1, up, Bt, ldbt, kofs, passes,
                                                      ctx->batch);
    };
    if (!upper) {
        for (int s = 0; s < wb; s += 64) {
            int b = std::min(64, wb - s);
            {
                ProfScope ps(ctx, fam, 0);
                trsm(r0 + s, b, 0, s);
                EMFA_HIP_CHECK(hipGetLastError());
            }
            if (s + b < wb) {
                emfa_status st = gemm_update<R>(ctx, "trsm_gemm", wb - s - b, ncols, b,
                                                F + (size_t)(r0 + s + b) * ldf + r0 + s, ldf, Bt + 2 * s, ldbt,
                                                X + (size_t)(r0 + s + b) * ldx + c0, ldx);
                if (st) return st;
            }
        }
    } else {
        int last = ((wb - 1) / 64) * 64;
        for (int s = last; s >= 0; s -= 64) {
            int b = std::min(64, wb - s);
            {
                ProfScope ps(ctx, fam, 0);
                trsm(r0 + s, b, 1, s);
                EMFA_HIP_CHECK(hipGetLastError());
            }
            if (s > 0) {
                emfa_status st = gemm_update<R>(ctx, "trsm_gemm", s, ncols, b, F + (size_t)r0 * ldf + r0 + s, ldf,
                                                Bt + 2 * s, ldbt, X + (size_t)r0 * ldx + c0, ldx);
                if (st) return st;
            }
        }
    }
    return EMFA_OK;
}

// ---------------------------------------------------------------------------
// Inverses of the 256 x 256 diagonal blocks of a factor, built once with the
// factor, turn each diagonal solve of a triangular solve into a product:
// a solve is then two launches per 256-row block (k_dinv_apply and the MFMA
// update of the remaining rows).
// k_tri_inv (FP64 factors; FP32 factors use k_dinv_diag / k_dinv_off below,
// the same recurrences spread over one workgroup per tile): one workgroup
// per block and triangle.  The block is handled as 4 x 4 tiles of 64: each
// diagonal tile is inverted in LDS (one lane per column, unit diagonal for
// L), the off-diagonal tiles of the inverse follow from
//   L: X_ij = -D_i sum_{k=j}^{i-1} T_ik X_kj   (i > j, i ascending)
//   U: X_ij = -D_i sum_{k=i+1}^{j} T_ik X_kj   (i < j, i descending)
// with 64 x 64 tile products in LDS; the other triangle is zeroed.  Rows and
// columns past the matrix are treated as identity.  Tiles of X are written
// to the output and read back by the same workgroup (fenced).
// ---------------------------------------------------------------------------
constexpr int kDinvB = 256;

template <typename R>
struct TileAcc {  // 4 x 4 complex outputs per thread of a 64 x 64 tile
    cx<R> v[4][4];
};

template <typename R>
__device__ __forceinline__ void tile_mac(TileAcc<R> &acc, const cx<R> (*Ta)[65], const cx<R> (*Tb)[65], int ty,
                                         int tx) {
    for (int k = 0; k < 64; ++k) {
        cx<R> a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Ta[ty * 4 + i][k];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Tb[k][tx + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc.v[i][j].re += a[i].re * b[j].re - a[i].im * b[j].im;
                acc.v[i][j].im += a[i].re * b[j].im + a[i].im * b[j].re;
            }
    }
}

// Inverse of a triangular 64 x 64 FP32 tile held in LDS, by recursive
// doubling: inverses of the 1 x 1 diagonal blocks, then for b = 1, 2, .., 32
// each pair of b-blocks is joined (L: X21 = -X22 L21 X11; U: X12 = -X11 U12
// X22), the product's middle factor staged in C.  T is read in its triangle
// only (unit diagonal taken for L); X receives the inverse, zero in the other
// triangle.  All `nthr` threads of the workgroup call it (it synchronises).
template <bool UPPER>
__device__ void tile_inv64(const cx<float> (*T)[65], cx<float> (*X)[65], cx<float> (*C)[65], int tid, int nthr) {
    using R = float;
    for (int e = tid; e < 64 * 64; e += nthr) {
        const int rr = e >> 6, cc = e & 63;
        cx<R> v = {R(0), R(0)};
        if (rr == cc) {
            if (UPPER) {
                const cx<R> d = T[rr][rr];
                const R den = d.re * d.re + d.im * d.im;
                v = den > R(0) ? cx<R>{d.re / den, -d.im / den} : cx<R>{R(0), R(0)};
            } else {
                v = {R(1), R(0)};
            }
        }
        X[rr][cc] = v;
    }
    __syncthreads();
    for (int b = 1; b < 64; b *= 2) {
        // pairs p: block start s = 2 b p; off-diagonal block rows/cols
        // L: C = L21 X11 (rows s+b.., cols s..), then X21 = -X22 C
        // U: C = U12 X22 (rows s.., cols s+b..), then X12 = -X11 C
        const int per = b * b, tot = (64 / (2 * b)) * per;
        for (int e = tid; e < tot; e += nthr) {
            const int pr = e / per, q = e % per, rr = q / b, cc = q % b, s0 = 2 * b * pr;
            cx<R> acc = {R(0), R(0)};
            if (!UPPER) {
                for (int k = 0; k < b; ++k) {
                    const cx<R> a = T[s0 + b + rr][s0 + k], x = X[s0 + k][s0 + cc];
                    acc.re += a.re * x.re - a.im * x.im;
                    acc.im += a.re * x.im + a.im * x.re;
                }
                C[s0 + b + rr][s0 + cc] = acc;
            } else {
                for (int k = 0; k < b; ++k) {
                    const cx<R> a = T[s0 + rr][s0 + b + k], x = X[s0 + b + k][s0 + b + cc];
                    acc.re += a.re * x.re - a.im * x.im;
                    acc.im += a.re * x.im + a.im * x.re;
                }
                C[s0 + rr][s0 + b + cc] = acc;
            }
        }
        __syncthreads();
        for (int e = tid; e < tot; e += nthr) {
            const int pr = e / per, q = e % per, rr = q / b, cc = q % b, s0 = 2 * b * pr;
            cx<R> acc = {R(0), R(0)};
            if (!UPPER) {
                for (int k = 0; k < b; ++k) {
                    const cx<R> a = X[s0 + b + rr][s0 + b + k], x = C[s0 + b + k][s0 + cc];
                    acc.re -= a.re * x.re - a.im * x.im;
                    acc.im -= a.re * x.im + a.im * x.re;
                }
                X[s0 + b + rr][s0 + cc] = acc;
            } else {
                for (int k = 0; k < b; ++k) {
                    const cx<R> a = X[s0 + rr][s0 + k], x = C[s0 + k][s0 + b + cc];
                    acc.re -= a.re * x.re - a.im * x.im;
                    acc.im -= a.re * x.im + a.im * x.re;
                }
                X[s0 + rr][s0 + b + cc] = acc;
            }
        }
        __syncthreads();
    }
}

template <typename R, bool UPPER>
__device__ void tri_inv_block(const cx<R> *__restrict__ F, long long ldf, int N, cx<R> *__restrict__ Dinv,
                              cx<R> (*Ta)[65], cx<R> (*Tb)[65]) {
    const int blk = blockIdx.x, r0 = blk * kDinvB, tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    cx<R> *X = Dinv + (size_t)blk * kDinvB * kDinvB;  // row-major 256 x 256
    // tile (i, k) of the block's triangle; identity past the matrix
    auto load_T = [&](cx<R>(*T)[65], int i, int k) {
        for (int e = tid; e < 64 * 64; e += 256) {
            const int rr = e >> 6, cc = e & 63;
            const int gr = r0 + 64 * i + rr, gc = r0 + 64 * k + cc;
            cx<R> v = {R(0), R(0)};
            if (gr < N && gc < N)
                v = F[(size_t)gr * ldf + gc];
            else if (gr == gc)
                v = {R(1), R(0)};
            T[rr][cc] = v;
        }
    };
    auto load_X = [&](cx<R>(*T)[65], int i, int j) {
        for (int e = tid; e < 64 * 64; e += 256) {
            const int rr = e >> 6, cc = e & 63;
            T[rr][cc] = X[(size_t)(64 * i + rr) * kDinvB + 64 * j + cc];
        }
    };
    auto store_X = [&](int i, int j, const TileAcc<R> &acc, R sign) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                X[(size_t)(64 * i + ty * 4 + a) * kDinvB + 64 * j + tx + 16 * b] = {sign * acc.v[a][b].re,
                                                                                     sign * acc.v[a][b].im};
    };
    auto zero_acc = [](TileAcc<R> &acc) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc.v[a][b] = {R(0), R(0)};
    };
    auto publish = [&]() {
        __threadfence_block();
        __syncthreads();
    };
    for (int step = 0; step < 4; ++step) {
        const int i = UPPER ? 3 - step : step;
        load_T(Ta, i, i);
        // lane t solves column t of the tile's inverse (LDS holds two FP64 tiles only)
        for (int e = tid; e < 64 * 64; e += 256) Tb[e >> 6][e & 63] = {R(0), R(0)};
        __syncthreads();
        if (tid < 64) {
            const int t = tid;
            for (int q = 0; q < 64; ++q) {
                const int r = UPPER ? 63 - q : q;
                cx<R> acc = {r == t ? R(1) : R(0), R(0)};
                const int kb = UPPER ? r + 1 : 0, ke = UPPER ? 64 : r;
                for (int k = kb; k < ke; ++k) {
                    const cx<R> t_rk = Ta[r][k], x = Tb[k][t];
                    acc.re -= t_rk.re * x.re - t_rk.im * x.im;
                    acc.im -= t_rk.re * x.im + t_rk.im * x.re;
                }
                if (UPPER) {
                    const cx<R> d = Ta[r][r];
                    const R den = d.re * d.re + d.im * d.im;
                    acc = den > R(0) ? cmul(acc, cx<R>{d.re / den, -d.im / den}) : cx<R>{R(0), R(0)};
                }
                Tb[r][t] = acc;
            }
        }
        __syncthreads();
        for (int e = tid; e < 64 * 64; e += 256) X[(size_t)(64 * i + (e >> 6)) * kDinvB + 64 * i + (e & 63)] = Tb[e >> 6][e & 63];
        // the other triangle of row tile i is zero
        for (int j = 0; j < 4; ++j)
            if (UPPER ? j < i : j > i)
                for (int e = tid; e < 64 * 64; e += 256)
                    X[(size_t)(64 * i + (e >> 6)) * kDinvB + 64 * j + (e & 63)] = {R(0), R(0)};
        publish();
        // off-diagonal tiles of row tile i
        const int jb = UPPER ? i + 1 : 0, je = UPPER ? 4 : i;
        for (int j = jb; j < je; ++j) {
            TileAcc<R> acc;
            zero_acc(acc);
            const int kb = UPPER ? i + 1 : j, ke = UPPER ? j + 1 : i;
            for (int k = kb; k < ke; ++k) {
                __syncthreads();
                load_T(Ta, i, k);
                load_X(Tb, k, j);
                __syncthreads();
                tile_mac(acc, Ta, Tb, ty, tx);
            }
            __syncthreads();
            // X_ij = -D_i * acc
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) Tb[ty * 4 + a][tx + 16 * b] = acc.v[a][b];
            load_X(Ta, i, i);
            __syncthreads();
            TileAcc<R> out;
            zero_acc(out);
            tile_mac(out, Ta, Tb, ty, tx);
            store_X(i, j, out, R(-1));
            publish();
        }
    }
}

// Grid (blocks, 2): blockIdx.y 0 builds the unit-lower inverses into Dinv,
// 1 the upper ones into Dinv + blocks * 256 * 256, in one launch.
template <typename R>
__global__ __launch_bounds__(256) void k_tri_inv(const cx<R> *__restrict__ F, long long ldf, int N,
                                                 cx<R> *__restrict__ Dinv) {
    __shared__ cx<R> Ta[64][65], Tb[64][65];
    if (blockIdx.y == 0)
        tri_inv_block<R, false>(F, ldf, N, Dinv, Ta, Tb);
    else
        tri_inv_block<R, true>(F, ldf, N, Dinv + (size_t)gridDim.x * kDinvB * kDinvB, Ta, Tb);
}

// FP32 diagonal-block inverses spread over the chip.  k_dinv_diag inverts
// every 64 x 64 diagonal tile (one workgroup per tile and triangle, grid
// (blocks * 4, 2), blockIdx.y 0: unit lower, 1: upper) and zeroes the other
// triangle of the tile's row; k_dinv_off(d), d = 1, 2, 3, forms the tiles on
// the d-th sub-diagonal (L) or super-diagonal (U) of every 256-block, one
// workgroup per tile (grid (blocks * (4 - d), 2)):
//   L: X_{i,i-d} = -X_ii sum_{k=i-d}^{i-1} L_ik X_{k,i-d}
//   U: X_{i,i+d} = -X_ii sum_{k=i+1}^{i+d} U_ik X_{k,i+d}
// so each launch reads only tiles written by the earlier ones.  Layout as
// k_tri_inv: L inverses at Dinv, U inverses at Dinv + blocks * 256 * 256,
// row-major 256 x 256 per block; rows and columns past N act as identity.
__device__ __forceinline__ void dinv_load_F(cx<float> (*T)[65], const cx<float> *__restrict__ F, long long ldf,
                                            int N, int gr0, int gc0, int tid) {
    for (int e = tid; e < 64 * 64; e += 256) {
        const int rr = e >> 6, cc = e & 63, gr = gr0 + rr, gc = gc0 + cc;
        cx<float> v = {0.f, 0.f};
        if (gr < N && gc < N)
            v = F[(size_t)gr * ldf + gc];
        else if (gr == gc)
            v = {1.f, 0.f};
        T[rr][cc] = v;
    }
}

__device__ __forceinline__ void dinv_load_X(cx<float> (*T)[65], const cx<float> *__restrict__ X, int i, int j,
                                            int tid) {
    for (int e = tid; e < 64 * 64; e += 256) T[e >> 6][e & 63] = X[(size_t)(64 * i + (e >> 6)) * kDinvB + 64 * j + (e & 63)];
}

__global__ __launch_bounds__(256) void k_dinv_diag(const cx<float> *__restrict__ F, long long ldf, int N,
                                                   cx<float> *__restrict__ Dinv, int nblk, Batch bf, Batch bd) {
    __shared__ cx<float> Ta[64][65], Tb[64][65], Tc[64][65];
    F = batch_mat(bf, F);  // a batch: factor z and its inverses on blockIdx.z
    Dinv = batch_mat(bd, Dinv);
    const int blk = blockIdx.x >> 2, i = blockIdx.x & 3, tid = threadIdx.x;
    const bool upper = blockIdx.y != 0;
    cx<float> *X = Dinv + (upper ? (size_t)nblk * kDinvB * kDinvB : 0) + (size_t)blk * kDinvB * kDinvB;
    const int r0 = blk * kDinvB + 64 * i;
    dinv_load_F(Ta, F, ldf, N, r0, r0, tid);
    __syncthreads();
    if (upper)
        tile_inv64<true>(Ta, Tb, Tc, tid, 256);
    else
        tile_inv64<false>(Ta, Tb, Tc, tid, 256);
    for (int e = tid; e < 64 * 64; e += 256) X[(size_t)(64 * i + (e >> 6)) * kDinvB + 64 * i + (e & 63)] = Tb[e >> 6][e & 63];
    for (int j = 0; j < 4; ++j)
        if (upper ? j < i : j > i)
            for (int e = tid; e < 64 * 64; e += 256)
                X[(size_t)(64 * i + (e >> 6)) * kDinvB + 64 * j + (e & 63)] = {0.f, 0.f};
}

__global__ __launch_bounds__(256) void k_dinv_off(const cx<float> *__restrict__ F, long long ldf, int N,
                                                  cx<float> *__restrict__ Dinv, int nblk, int d, Batch bf, Batch bd) {
    __shared__ cx<float> Ta[64][65], Tb[64][65];
    F = batch_mat(bf, F);
    Dinv = batch_mat(bd, Dinv);
    const int per = 4 - d, blk = blockIdx.x / per, t = blockIdx.x % per, tid = threadIdx.x;
    const int ty = tid >> 4, tx = tid & 15;
    const bool upper = blockIdx.y != 0;
    cx<float> *X = Dinv + (upper ? (size_t)nblk * kDinvB * kDinvB : 0) + (size_t)blk * kDinvB * kDinvB;
    const int r0 = blk * kDinvB;
    const int i = upper ? t : d + t, j = upper ? i + d : i - d;
    const int kb = upper ? i + 1 : j, ke = upper ? j : i - 1;  // inclusive
    TileAcc<float> acc;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc.v[a][b] = {0.f, 0.f};
    for (int k = kb; k <= ke; ++k) {
        if (k != kb) __syncthreads();
        dinv_load_F(Ta, F, ldf, N, r0 + 64 * i, r0 + 64 * k, tid);
        dinv_load_X(Tb, X, k, j, tid);
        __syncthreads();
        tile_mac(acc, Ta, Tb, ty, tx);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Tb[ty * 4 + a][tx + 16 * b] = acc.v[a][b];
    dinv_load_X(Ta, X, i, i, tid);
    __syncthreads();
    TileAcc<float> out;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) out.v[a][b] = {0.f, 0.f};
    tile_mac(out, Ta, Tb, ty, tx);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            X[(size_t)(64 * i + ty * 4 + a) * kDinvB + 64 * j + tx + 16 * b] = {-out.v[a][b].re, -out.v[a][b].im};
}

// Inverse-scaled off-diagonal blocks of an FP32 factor, for one-launch solve
// steps: G = L[rows below block b, block b] DL_b^-1 and H = U[rows above
// block b, block b] DU_b^-1, each stored where its L or U block sits in GH
// (N x N, row-major; the diagonal blocks of GH are not used).  Grid (row
// tiles of 64, 8 per block: 4 column tiles x {L, U}, batch): a workgroup
// forms one 64 x 64 tile from four tile products; tiles outside the
// triangle return at once.
__global__ __launch_bounds__(256) void k_dinv_scale(const cx<float> *__restrict__ F, int N,
                                                    const cx<float> *__restrict__ Dinv, int nblk,
                                                    cx<float> *__restrict__ GH, Batch bf, Batch bd, Batch bg) {
    __shared__ cx<float> Ta[64][65], Tb[64][65];
    const int r0 = 64 * blockIdx.x, b = blockIdx.y >> 3, upper = (blockIdx.y >> 2) & 1, ct = blockIdx.y & 3;
    const int c0 = kDinvB * b + 64 * ct, tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
    if (r0 >= N || c0 >= N) return;
    if (upper ? r0 + 64 > kDinvB * b : r0 < kDinvB * (b + 1)) return;
    F = batch_mat(bf, F);
    Dinv = batch_mat(bd, Dinv);
    GH = batch_mat(bg, GH);
    const cx<float> *D = Dinv + (upper ? (size_t)nblk * kDinvB * kDinvB : 0) + (size_t)b * kDinvB * kDinvB;
    TileAcc<float> acc;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc.v[a][q] = {0.f, 0.f};
    for (int kc = 0; kc < 4; ++kc) {
        const int k0 = kDinvB * b + 64 * kc;
        if (k0 >= N) break;
        if (kc) __syncthreads();
        for (int e = tid; e < 64 * 64; e += 256) {
            const int rr = e >> 6, cc = e & 63;
            Ta[rr][cc] = (r0 + rr < N && k0 + cc < N) ? F[(size_t)(r0 + rr) * N + k0 + cc] : cx<float>{0.f, 0.f};
            Tb[rr][cc] = D[(size_t)(64 * kc + rr) * kDinvB + 64 * ct + cc];
        }
        __syncthreads();
        tile_mac(acc, Ta, Tb, ty, tx);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = r0 + ty * 4 + a, c = c0 + tx + 16 * q;
            if (r < N && c < N) GH[(size_t)r * N + c] = acc.v[a][q];
        }
}

// Y_b = Dinv_b W_b for one 256-row block (nb valid rows, m columns): 32 x 32
// outputs per workgroup, K in chunks of 64 through LDS.  Writes Y (plain) and
// the block's real-expanded transposed form Bt (2m x 2nb reals, the operand
// layout of the MFMA update).
template <typename R>
__global__ __launch_bounds__(256) void k_dinv_apply(const cx<R> *__restrict__ Dinv, int nb,
                                                    const cx<R> *__restrict__ W, long long ldw, int m,
                                                    cx<R> *__restrict__ Y, long long ldy, R *__restrict__ Bt) {
    __shared__ cx<R> Ds[32][65], Ws[64][33];
    const int tid = threadIdx.x, tx = tid & 31, ty = tid >> 5;  // column tx, rows ty * 4 + a
    const int rb = blockIdx.x * 32, cb = blockIdx.y * 32;
    cx<R> acc[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] = {R(0), R(0)};
    for (int k0 = 0; k0 < nb; k0 += 64) {
        for (int e = tid; e < 32 * 64; e += 256) {
            const int rr = e >> 6, kk = e & 63;
            Ds[rr][kk] = (rb + rr < nb && k0 + kk < nb) ? Dinv[(size_t)(rb + rr) * kDinvB + k0 + kk]
                                                          : cx<R>{R(0), R(0)};
        }
        for (int e = tid; e < 64 * 32; e += 256) {
            const int kk = e >> 5, cc = e & 31;
            Ws[kk][cc] = (k0 + kk < nb && cb + cc < m) ? W[(size_t)(k0 + kk) * ldw + cb + cc] : cx<R>{R(0), R(0)};
        }
        __syncthreads();
        for (int kk = 0; kk < 64; ++kk) {
            const cx<R> w = Ws[kk][tx];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const cx<R> d = Ds[ty * 4 + a][kk];
                acc[a].re += d.re * w.re - d.im * w.im;
                acc[a].im += d.re * w.im + d.im * w.re;
            }
        }
        __syncthreads();
    }
    const int j = cb + tx;
    if (j >= m) return;
    const long long ldbt = 2LL * nb;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = rb + ty * 4 + a;
        if (r >= nb) continue;
        const cx<R> x = acc[a];
        Y[(size_t)r * ldy + j] = x;
        if (!Bt) continue;
        R *b0 = Bt + (2LL * j) * ldbt + 2 * r;
        R *b1 = Bt + (2LL * j + 1) * ldbt + 2 * r;
        b0[0] = x.re;
        b0[1] = -x.im;
        b1[0] = x.im;
        b1[1] = x.re;
    }
}

constexpr int kSkinnyMaxCols = 64;

// Small products with at most 64 columns, the form every few-RHS solve step
// and the refinement residual take:
//   C (M x m) = beta C + sign A (M x K) X (K x m)      (P == null)
//   plane z of P = A X over K slice [z kc, (z + 1) kc)  (P given)
// The column tile MC (8, 16, 32, 64) follows m, so no lane works on padding
// columns: lane tx owns column tx, rows ty + h RG (h < RPT) of the
// workgroup's RPT RG rows (RG = 256 / MC).  Two rows per lane: more rows per
// workgroup cut the re-reads of X but cost more in parallelism (measured:
// eight rows per lane on tall 32 / 64-column products made a 7680-row solve
// 2.5 -> 4.2 ms and the served stream 223 -> 187 solves/s).
// K goes through LDS in chunks of 64 with the next chunk's operands loaded
// into registers while the current chunk is multiplied.
template <int MC>
constexpr int smallmm_rpt() {
    return 2;
}

template <int MC, int RPT = smallmm_rpt<MC>()>
constexpr int smallmm_rb() {
    return RPT * (256 / MC);
}

// row block bx, K slice z of C (+)= sign A X (see k_smallmm); LDS from the caller
template <typename TA, typename TC, int MC, int RPT = smallmm_rpt<MC>()>
__device__ __forceinline__ void smallmm_body(const cx<TA> *__restrict__ A, long long lda,
                                             const cx<TC> *__restrict__ X, long long ldx, cx<TC> *C, long long ldc,
                                             int M, int K, int m, int kc, TC sign, int beta, cx<TC> *__restrict__ P,
                                             int bx, int z, cx<TA> (*As)[65], cx<TC> (*Xs)[MC]) {
    constexpr int RG = 256 / MC, RB = RPT * RG, KC = 64;
    constexpr int AL = RB * KC / 256, XL = KC * MC / 256;  // chunk elements per thread
    const int tid = threadIdx.x, tx = tid % MC, ty = tid / MC;
    const int rb = bx * RB;
    const int kb = z * kc, ke = min(K, kb + kc);
    cx<TA> ra[AL];
    cx<TC> rx[XL];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < AL; ++i) {
            const int e = tid + 256 * i, r = e / KC, k = e % KC;
            ra[i] = (rb + r < M && k0 + k < ke) ? A[(size_t)(rb + r) * lda + k0 + k] : cx<TA>{TA(0), TA(0)};
        }
#pragma unroll
        for (int i = 0; i < XL; ++i) {
            const int e = tid + 256 * i, k = e / MC, c = e % MC;
            rx[i] = (k0 + k < ke && c < m) ? X[(size_t)(k0 + k) * ldx + c] : cx<TC>{TC(0), TC(0)};
        }
    };
    cx<TC> acc[RPT];
#pragma unroll
    for (int h = 0; h < RPT; ++h) acc[h] = {TC(0), TC(0)};
    if (kb < ke) gload(kb);
    for (int k0 = kb; k0 < ke; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < AL; ++i) {
            const int e = tid + 256 * i;
            As[e / KC][e % KC] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < XL; ++i) {
            const int e = tid + 256 * i;
            Xs[e / MC][e % MC] = rx[i];
        }
        __syncthreads();
        if (k0 + KC < ke) gload(k0 + KC);
#pragma unroll 8
        for (int k = 0; k < KC; ++k) {
            const cx<TC> x = Xs[k][tx];
#pragma unroll
            for (int h = 0; h < RPT; ++h) {
                const cx<TA> a = As[ty + h * RG][k];
                acc[h].re += (TC)a.re * x.re - (TC)a.im * x.im;
                acc[h].im += (TC)a.re * x.im + (TC)a.im * x.re;
            }
        }
    }
    if (tx >= m) return;
#pragma unroll
    for (int h = 0; h < RPT; ++h) {
        const int r = rb + ty + h * RG;
        if (r >= M) continue;
        const cx<TC> v = acc[h];
        if (P) {
            P[(size_t)z * M * m + (size_t)r * m + tx] = v;
        } else {
            cx<TC> c = beta ? C[(size_t)r * ldc + tx] : cx<TC>{TC(0), TC(0)};
            c.re += sign * v.re;
            c.im += sign * v.im;
            C[(size_t)r * ldc + tx] = c;
        }
    }
}

template <typename TA, typename TC, int MC, int RPT>
__global__ __launch_bounds__(256) void k_smallmm(const cx<TA> *__restrict__ A, long long lda,
                                                 const cx<TC> *__restrict__ X, long long ldx, cx<TC> *C,
                                                 long long ldc, int M, int K, int m, int kc, TC sign, int beta,
                                                 cx<TC> *__restrict__ P) {
    __shared__ cx<TA> As[smallmm_rb<MC, RPT>()][65];
    __shared__ cx<TC> Xs[64][MC];
    smallmm_body<TA, TC, MC, RPT>(A, lda, X, ldx, C, ldc, M, K, m, kc, sign, beta, P, blockIdx.x, blockIdx.y, As,
                                  Xs);
}

// One step of a triangular solve through the diagonal-block inverses, in one
// launch: workgroups [0, n0) form Y_b = D_b W_b on the block's rows (C0 =
// D W, K = wb), the rest update the other rows with the same W_b through the
// inverse-scaled off-diagonal block (C1 -= G W, G = L21 D_L^-1 or U12 D_U^-1
// stored with the factor).  Both only read W_b and write disjoint rows.
template <typename R, int MC>
__global__ __launch_bounds__(256) void k_solve_step(const cx<R> *__restrict__ D, const cx<R> *__restrict__ G,
                                                    long long ldg, const cx<R> *__restrict__ Wb, int m,
                                                    cx<R> *__restrict__ Yb, cx<R> *__restrict__ C1, int wb, int M1,
                                                    int n0) {
    __shared__ cx<R> As[smallmm_rb<MC>()][65];
    __shared__ cx<R> Xs[64][MC];
    if ((int)blockIdx.x < n0)
        smallmm_body<R, R, MC>(D, kDinvB, Wb, m, Yb, m, wb, wb, m, wb, R(1), 0, nullptr, blockIdx.x, 0, As, Xs);
    else
        smallmm_body<R, R, MC>(G, ldg, Wb, m, C1, m, M1, wb, m, wb, R(-1), 1, nullptr, blockIdx.x - n0, 0, As, Xs);
}

// launch k_smallmm with the column tile for m (<= 64); slices on gridDim.y.
// rpt: rows per lane, 2 (short products: the most workgroups) or 8 (long-K
// products such as the refinement residual: every workgroup reads its slice
// of X, so fewer, taller workgroups read it fewer times; 32 / 64-column tiles)
template <typename TA, typename TC>
static emfa_status smallmm(hipStream_t s, const cx<TA> *A, long long lda, const cx<TC> *X, long long ldx, cx<TC> *C,
                           long long ldc, int M, int K, int m, int slices, int kc, TC sign, int beta, cx<TC> *P,
                           int rpt = 2) {
    if (M <= 0 || m <= 0) return EMFA_OK;
    EMFA_REQUIRE(m <= kSkinnyMaxCols && slices >= 1 && kc > 0, EMFA_E_INVALID, "smallmm: bad shape m=%d", m);
    auto go = [&](auto mc) {
        constexpr int MC = decltype(mc)::value;
        if constexpr (MC >= 32) {
            if (rpt == 8) {
                constexpr int RB = smallmm_rb<MC, 8>();
                const dim3 grid((M + RB - 1) / RB, slices);
                k_smallmm<TA, TC, MC, 8><<<grid, 256, 0, s>>>(A, lda, X, ldx, C, ldc, M, K, m, kc, sign, beta, P);
                return;
            }
        }
        constexpr int RB = smallmm_rb<MC, 2>();
        const dim3 grid((M + RB - 1) / RB, slices);
        k_smallmm<TA, TC, MC, 2><<<grid, 256, 0, s>>>(A, lda, X, ldx, C, ldc, M, K, m, kc, sign, beta, P);
    };
    if (m <= 8)
        go(std::integral_constant<int, 8>{});
    else if (m <= 16)
        go(std::integral_constant<int, 16>{});
    else if (m <= 32)
        go(std::integral_constant<int, 32>{});
    else
        go(std::integral_constant<int, 64>{});
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

// one solve step (k_solve_step) with the column tile for m (<= 64)
template <typename R>
static emfa_status solve_step(hipStream_t s, const cx<R> *D, const cx<R> *G, long long ldg, const cx<R> *Wb, int m,
                              cx<R> *Yb, cx<R> *C1, int wb, int M1) {
    EMFA_REQUIRE(m >= 1 && m <= kSkinnyMaxCols && wb >= 1 && wb <= kDinvB && M1 >= 0, EMFA_E_INVALID,
                 "solve step: bad shape m=%d wb=%d", m, wb);
    auto go = [&](auto mc) {
        constexpr int MC = decltype(mc)::value, RB = smallmm_rb<MC>();
        const int n0 = (wb + RB - 1) / RB, n1 = (M1 + RB - 1) / RB;
        k_solve_step<R, MC><<<n0 + n1, 256, 0, s>>>(D, G, ldg, Wb, m, Yb, C1, wb, M1, n0);
    };
    if (m <= 8)
        go(std::integral_constant<int, 8>{});
    else if (m <= 16)
        go(std::integral_constant<int, 16>{});
    else if (m <= 32)
        go(std::integral_constant<int, 32>{});
    else
        go(std::integral_constant<int, 64>{});
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

// W = P B: row i of W is row perm[i] of B
template <typename R>
__global__ __launch_bounds__(256) void k_gather_rows(const cx<R> *__restrict__ B, const int *__restrict__ perm, int N,
                                                     int m, cx<R> *__restrict__ W) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * m) return;
    const int i = (int)(e / m), c = (int)(e % m);
    W[e] = B[(size_t)perm[i] * m + c];
}

// Workspace of the row interchanges: pair lists and a staging buffer of
// 2*kPermChunk rows x max_cols.
constexpr int kPermChunk = 256;

template <typename R>
struct PermWs {
    int *dst = nullptr, *src = nullptr, *count = nullptr;
    cx<R> *stage = nullptr;
    // layout: dst[512] ints | src[512] ints | count | ... | stage (from byte 8192)
    static size_t bytes(int max_cols) { return 8192 + (size_t)2 * kPermChunk * max_cols * sizeof(cx<R>); }
    void carve(void *p) {
        char *q = (char *)p;
        dst = (int *)q;
        src = (int *)(q + 2048);
        count = (int *)(q + 4096);
        stage = (cx<R> *)(q + 8192);
    }
};

// ---------------------------------------------------------------------------
// Panel without row interchanges (FP32 factorisations that are refined in
// FP64): the w x w diagonal block is factored in one workgroup, then every
// row below solves x U11 = a on its own (L21 = A21 U11^-1), so a panel needs
// no cross-workgroup exchange at all.  Both kernels raise flags[2] when a
// pivot is zero or a multiplier exceeds sqrt(tau2) in modulus (partial
// pivoting keeps them <= 1); the factorisation is then redone with
// interchanges.
// ---------------------------------------------------------------------------
constexpr int kNoPivW = 64;  // widest pivot-free panel

// Diagonal block (w <= 64): lane i of every wave holds row i, wave v holds
// columns 4 v .. 4 v + 3 in registers (16 waves).  LU by 4-column blocks:
// the owning wave factors its four columns (pivots read with v_readlane) and
// publishes the four multiplier columns through LDS; after one barrier every
// wave to the right forms the four pivot-row entries of its columns (a 4 x 4
// unit-lower solve on wave-uniform values) and applies the rank-4 update.
// Then U11^-1 from a copy of U in LDS by recursive doubling on all 16 waves
// (tile_inv64), written to uinv (64 x 64, row-major).  Rows and columns past
// w act as identity; the LU's LDS buffers alternate by block parity.
// reciprocal for the pivot-free FP32 factor (refined in FP64 afterwards):
// the hardware estimate (1 ulp) instead of a correctly rounded division
template <typename R>
__device__ __forceinline__ R fast_recip(R x) {
    if constexpr (sizeof(R) == 4)
        return __builtin_amdgcn_rcpf(x);
    else
        return R(1) / x;
}

template <typename R>
__device__ __forceinline__ R lane_read(R v, int l) {
    if constexpr (sizeof(R) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}

constexpr int kDiagCpw = 4, kDiagWaves = kNoPivW / kDiagCpw;  // columns per wave, waves

template <typename R>
__device__ __forceinline__ void cfms(cx<R> &acc, const cx<R> &a, const cx<R> &b) {  // acc -= a b
    acc.re -= a.re * b.re - a.im * b.im;
    acc.im -= a.re * b.im + a.im * b.re;
}

template <typename R>
__global__ __launch_bounds__(kDiagWaves * 64) void k_diag_lu(cx<R> *__restrict__ A, long long lda, int js, int w,
                                                            int *__restrict__ ipiv, int *flags, R tau2,
                                                            cx<R> *__restrict__ uinv, Batch bt) {
    // the matrix, its (scratch) pivots, flags and U11^-1 for the batch's matrix z
    A = batch_mat(bt, A);
    ipiv = batch_scr(bt, ipiv);
    flags = batch_scr(bt, flags);
    uinv = batch_scr(bt, uinv);
    static_assert(sizeof(R) == 4, "the pivot-free diagonal block is an FP32 kernel");
    __shared__ cx<R> buf[2][kDiagCpw][kNoPivW];
    __shared__ cx<float> Tu[kNoPivW][kNoPivW + 1], Tx[kNoPivW][kNoPivW + 1], Tc[kNoPivW][kNoPivW + 1];
    __shared__ int bad;
    // wave index as a scalar: block tests below are uniform per wave
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cx<R> a[kDiagCpw];
#pragma unroll
    for (int q = 0; q < kDiagCpw; ++q) {
        const int j = kDiagCpw * wv + q;
        a[q] = (lane < w && j < w) ? A[(size_t)(js + lane) * lda + js + j] : cx<R>{lane == j ? R(1) : R(0), R(0)};
    }
    if (threadIdx.x == 0) bad = 0;
    bool mybad = false;
    const int nbk = (w + kDiagCpw - 1) / kDiagCpw;
    __syncthreads();
    for (int b = 0; b < nbk; ++b) {
        const int p = b & 1, c0 = kDiagCpw * b;
        if (wv == b) {
#pragma unroll
            for (int cc = 0; cc < kDiagCpw; ++cc) {
                const int c = c0 + cc;
                cx<R> l = {R(0), R(0)};
                if (c < w) {
                    const R pr = lane_read(a[cc].re, c), pi = lane_read(a[cc].im, c);
                    const R den = pr * pr + pi * pi;
                    const R rd = den > R(0) ? fast_recip(den) : R(0);
                    if (!(den > R(0))) mybad = true;
                    const bool below = lane > c;
                    l = cmul(a[cc], cx<R>{pr * rd, -pi * rd});
                    if (!below) l = {R(0), R(0)};
                    if (l.re * l.re + l.im * l.im > tau2) mybad = true;
                    if (below) a[cc] = l;
#pragma unroll
                    for (int q = cc + 1; q < kDiagCpw; ++q) {
                        const cx<R> u = {lane_read(a[q].re, c), lane_read(a[q].im, c)};
                        cfms(a[q], l, u);
                    }
                }
                buf[p][cc][lane] = l;
            }
        }
        __syncthreads();
        if (wv > b) {
            cx<R> l[kDiagCpw];
#pragma unroll
            for (int m = 0; m < kDiagCpw; ++m) l[m] = buf[p][m][lane];
#pragma unroll
            for (int q = 0; q < kDiagCpw; ++q) {
                cx<R> u[kDiagCpw];
#pragma unroll
                for (int k = 0; k < kDiagCpw; ++k) {
                    u[k] = {lane_read(a[q].re, c0 + k), lane_read(a[q].im, c0 + k)};
#pragma unroll
                    for (int m = 0; m < k; ++m) cfms(u[k], buf[p][m][c0 + k], u[m]);
                }
#pragma unroll
                for (int m = 0; m < kDiagCpw; ++m) cfms(a[q], l[m], u[m]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kDiagCpw; ++q) {
        const int j = kDiagCpw * wv + q;
        if (lane < w && j < w) A[(size_t)(js + lane) * lda + js + j] = a[q];
    }
    // U11^-1 by recursive doubling on all waves, from a copy of U in LDS
    // (the strictly lower part of the copy is never read)
#pragma unroll
    for (int q = 0; q < kDiagCpw; ++q) Tu[lane][kDiagCpw * wv + q] = a[q];
    __syncthreads();
    tile_inv64<true>(Tu, Tx, Tc, threadIdx.x, kDiagWaves * 64);
    for (int e = threadIdx.x; e < kNoPivW * kNoPivW; e += kDiagWaves * 64) uinv[e] = Tx[e >> 6][e & 63];
    if (threadIdx.x < w) ipiv[js + threadIdx.x] = js + threadIdx.x;
    if (mybad) bad = 1;
    __syncthreads();
    if (threadIdx.x == 0 && bad) flags[2] = 1;
}

// Rows r0..r1-1 of the panel: L21 = A21 U11^-1 with the inverse from
// k_diag_lu.  32 rows x w columns per workgroup (lane tx: column, rows
// ty * 8 + a), in place: a workgroup reads all of its rows before writing.
template <typename R>
__device__ __forceinline__ void l21_body(cx<R> *__restrict__ A, long long lda, int js, int w, int r0, int r1,
                                         const cx<R> *__restrict__ uinv, int *flags, R tau2, int bx,
                                         cx<R> (*Us)[kNoPivW + 1], cx<R> (*As)[kNoPivW + 1]) {
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int base = r0 + bx * 32, nrow = min(32, r1 - base);
    for (int e = tid; e < kNoPivW * kNoPivW; e += 256) {
        const int k = e >> 6, c = e & 63;
        Us[k][c] = (k < w && c < w) ? uinv[k * kNoPivW + c] : cx<R>{R(0), R(0)};
    }
    for (int e = tid; e < 32 * kNoPivW; e += 256) {
        const int r = e >> 6, c = e & 63;
        As[r][c] = (r < nrow && c < w) ? A[(size_t)(base + r) * lda + js + c] : cx<R>{R(0), R(0)};
    }
    __syncthreads();
    cx<R> acc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = {R(0), R(0)};
    for (int k = 0; k < w; ++k) {
        const cx<R> u = Us[k][tx];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const cx<R> v = As[ty * 8 + q][k];
            acc[q].re += v.re * u.re - v.im * u.im;
            acc[q].im += v.re * u.im + v.im * u.re;
        }
    }
    bool bad = false;
    if (tx < w) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int r = ty * 8 + q;
            if (r < nrow) {
                A[(size_t)(base + r) * lda + js + tx] = acc[q];
                if (acc[q].re * acc[q].re + acc[q].im * acc[q].im > tau2) bad = true;
            }
        }
    }
    if (bad) flags[2] = 1;
}

template <typename R>
__global__ __launch_bounds__(256) void k_l21(cx<R> *__restrict__ A, long long lda, int js, int w, int r0, int r1,
                                            const cx<R> *__restrict__ uinv, int *flags, R tau2, Batch bt) {
    __shared__ cx<R> Us[kNoPivW][kNoPivW + 1], As[32][kNoPivW + 1];
    l21_body<R>(batch_mat(bt, A), lda, js, w, r0, r1, batch_scr(bt, uinv), batch_scr(bt, flags), tau2, blockIdx.x,
                Us, As);
}

// The two products that follow a pivot-free diagonal block, in one launch:
// workgroups [0, nl21) form L21 rows r0..r1-1 (as k_l21), the rest solve the
// block's U rows for panel columns c0..c1-1 (as k_trsm64 with a unit-lower
// 64-row block, writing Bt for the in-panel update).  The two touch disjoint
// parts of A and only read the diagonal block, so they run side by side.
template <typename R>
__global__ __launch_bounds__(256) void k_l21_u12(cx<R> *__restrict__ A, long long lda, int js, int w, int r0,
                                                int r1, const cx<R> *__restrict__ uinv, int *flags, R tau2,
                                                int nl21, int c0, int c1, R *__restrict__ Bt, long long ldbt,
                                                int passes, Batch bt) {
    __shared__ cx<R> smem[(kNoPivW + 32) * (kNoPivW + 1)];
    A = batch_mat(bt, A);
    uinv = batch_scr(bt, uinv);
    flags = batch_scr(bt, flags);
    Bt = batch_scr(bt, Bt);
    auto *T0 = (cx<R>(*)[kNoPivW + 1])smem;
    if ((int)blockIdx.x < nl21)
        l21_body<R>(A, lda, js, w, r0, r1, uinv, flags, tau2, blockIdx.x, T0, T0 + kNoPivW);
    else
        trsm64_body<R>(A, lda, A, lda, js, w, c0, c1, 0, Bt, ldbt, 0, passes, blockIdx.x - nl21, T0);
}

// Apply the interchanges k0..k1-1 to columns [c0,c1) and [c2,c3).
template <typename R>
static emfa_status laswp_launch(emfa_ctx *ctx, cx<R> *A, long long lda, const int *ipiv, int k0, int k1, int c0,
                                int c1, int c2, int c3, PermWs<R> &ws) {
    c1 = std::max(c0, c1);
    c3 = std::max(c2, c3);
    const int ncols = (c1 - c0) + (c3 - c2);
    if (ncols <= 0 || k1 <= k0) return EMFA_OK;
    ProfScope ps(ctx, "lu_laswp", 0);
    for (int kb = k0; kb < k1; kb += kPermChunk) {
        int ke = std::min(k1, kb + kPermChunk);
        k_perm_build<<<1, 64, 0, ctx->stream>>>(ipiv, kb, ke, ws.dst, ws.src, ws.count);
        dim3 grid((ncols + 255) / 256, 2 * (ke - kb));
        k_perm_gather<R><<<grid, 256, 0, ctx->stream>>>(A, lda, ws.src, ws.count, c0, c1, c2, c3, ws.stage);
        k_perm_scatter<R><<<grid, 256, 0, ctx->stream>>>(A, lda, ws.dst, ws.count, c0, c1, c2, c3, ws.stage);
        EMFA_HIP_CHECK(hipGetLastError());
    }
    return EMFA_OK;
}

// Cooperative launch of a panel kernel.  False, with *coop cleared so the
// rest of the factorisation launches its panels normally, when the runtime
// refuses it: it does while another process on the device holds the
// cooperative-launch resources (a process that has launched cooperatively
// keeps them), e.g. a solver run started beside a server process.
static bool coop_launch(hipStream_t s, const void *fn, int G, void **args, size_t lds, bool *coop) {
    if (hipLaunchCooperativeKernel(fn, dim3(G), dim3(256), args, lds, s) == hipSuccess) return true;
    (void)hipGetLastError();
    *coop = false;
    return false;
}

// Run `body` (a launch sequence on ctx->stream, with the look-ahead stream
// joined back) as the context's graph for `key`: captured and instantiated
// the first time, replayed afterwards.
static emfa_status lu_graph_run(emfa_ctx *ctx, const std::string &key, const std::function<emfa_status()> &body) {
    auto it = ctx->lu_graph.find(key);
    if (it == ctx->lu_graph.end()) {
        // Captured on two streams of its own: a capture may not depend on
        // work outside it, and the device pool leaves waits on other
        // contexts' events on the context's streams when it hands out their
        // freed blocks.  Relaxed mode: other threads' contexts allocate and
        // launch meanwhile.
        if (!ctx->cap_main) {
            EMFA_HIP_CHECK(hipStreamCreateWithFlags(&ctx->cap_main, hipStreamNonBlocking));
            EMFA_HIP_CHECK(hipStreamCreateWithFlags(&ctx->cap_aux, hipStreamNonBlocking));
        }
        std::swap(ctx->stream, ctx->cap_main);
        std::swap(ctx->aux, ctx->cap_aux);
        emfa_status st = EMFA_OK;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
        if (e == hipSuccess) {
            st = body();
            e = hipStreamEndCapture(ctx->stream, &g);
        }
        std::swap(ctx->stream, ctx->cap_main);
        std::swap(ctx->aux, ctx->cap_aux);
        if (st || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            if (st) return st;
            set_error("lu: capturing the factorisation failed: %s", hipGetErrorString(e));
            return EMFA_E_HIP;
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            set_error("lu: instantiating the factorisation graph failed: %s", hipGetErrorString(ei));
            return EMFA_E_HIP;
        }
        if (ctx->lu_graph.size() >= 64) {  // bounded: drop them all once the stream is idle
            EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            for (auto &o : ctx->lu_graph) (void)hipGraphExecDestroy(o.second);
            ctx->lu_graph.clear();
        }
        it = ctx->lu_graph.emplace(key, x).first;
        ++ctx->graph_captures;
    }
    EMFA_HIP_CHECK(hipGraphLaunch(it->second, ctx->stream));
    ++ctx->graph_launches;
    return EMFA_OK;
}

template <typename R>
// With ctx->batch set (count matrices of Am's size, matrix 0 = Am), the
// pivot-free FP32 path with the split update factors all of them with one
// launch sequence; growth[z] then reports matrix z's growth flag and the
// call itself fails only on errors that concern the whole batch.
static emfa_status lu_impl(emfa_ctx *ctx, emfa_mat *Am, int *ipiv, int *info_out, int *interchanges,
                           int *growth = nullptr) {
    const int N = Am->rows;
    const bool batched = ctx->batch.mat != nullptr;
    cx<R> *A = (cx<R> *)Am->d;
    const long long lda = N;
    LuPlan<R> plan;
    if (ctx->lu_nb > 0) plan.nb = std::max(16, ctx->lu_nb);
    if (const char *s = getenv("EMFA_LU_NB")) plan.nb = std::max(16, atoi(s));
    const int nb = plan.nb;

    // panel geometry: co-resident workgroups, rows per workgroup in LDS
    size_t lds_cap = 150 * 1024;
    if (const char *s = getenv("EMFA_PANEL_LDS_KB")) lds_cap = (size_t)std::max(16, std::min(150, atoi(s))) * 1024;
    const int share = ctx->panel_cap > 0 ? ctx->panel_cap : (int)ctx->cu_share;
    int G_max = std::max(1, share > 0 ? std::min(share, ctx->num_cu) : ctx->num_cu);
    if (ctx->panel_g > 0) G_max = std::max(1, std::min(G_max, ctx->panel_g));
    if (const char *s = getenv("EMFA_PANEL_G")) G_max = std::max(1, std::min(G_max, atoi(s)));
    int panel_rows_min = 16;
    if (const char *s = getenv("EMFA_PANEL_ROWS_MIN")) panel_rows_min = std::max(1, atoi(s));
    // scratch: slots + Bt workspace
    size_t slot_rows_b = 2ull * G_max * kPanelMaxW * sizeof(cx<R>);
    size_t slot_val_b = 2ull * G_max * sizeof(R);
    size_t slot_idx_b = 2ull * G_max * sizeof(int);
    size_t rowc_b = 2ull * kPanelMaxW * sizeof(cx<R>);
    size_t bar_b = 256;
    size_t bt_b = (size_t)2 * N * 2 * nb * sizeof(R);
    size_t perm_b = PermWs<R>::bytes(N);
    // packet-exchange slots: packets, candidate rows, current row (zeroed below), values
    size_t x_pk_b = 2ull * kPanelMaxG * 8, x_rw_b = 2ull * kPanelMaxG * kPanelMaxW * 16,
           x_cw_b = 2ull * kPanelMaxW * 16, x_val_b = 2ull * kPanelMaxG * sizeof(R);
    // FP32 factorisations with the bf16-split trailing update keep split operand planes
    bool split_update = false;
    if constexpr (sizeof(R) == 4) {
        split_update = ctx->lu_update_mode == 1;
        if (const char *e = getenv("EMFA_LU_BF16X3")) split_update = atoi(e) != 0;
    }
    size_t planes_b = split_update ? UpdatePlanes::bytes(N, nb) : 0;
    const size_t uinv_b = (size_t)kNoPivW * kNoPivW * sizeof(cx<R>);
    size_t total = slot_rows_b + slot_val_b + slot_idx_b + rowc_b + bar_b + 2 * bt_b + perm_b + x_pk_b + x_rw_b +
                   x_cw_b + x_val_b + planes_b + uinv_b + 8192;
    // a batch: one slab of this layout per matrix (plus its flags and pivots)
    const size_t slab = batched ? (total + 256 + (size_t)N * sizeof(int) + 4096 + 4095) & ~(size_t)4095 : total;
    if (batched) {
        EMFA_REQUIRE(sizeof(R) == 4 && split_update && ctx->batch.count >= 1, EMFA_E_INVALID,
                     "LU batch: FP32 with the split update only");
    }
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, slab * (batched ? ctx->batch.count : 1), &scr);
    if (st != EMFA_OK) return st;
    if (batched) ctx->batch.scratch = (long long)slab;
    char *p = (char *)scr;
    auto take = [&](size_t b) {
        char *q = p;
        p += (b + 255) & ~(size_t)255;
        return (void *)q;
    };
    cx<R> *slot_rows = (cx<R> *)take(slot_rows_b);
    R *slot_val = (R *)take(slot_val_b);
    int *slot_idx = (int *)take(slot_idx_b);
    cx<R> *rowc = (cx<R> *)take(rowc_b);
    unsigned *bar = (unsigned *)take(bar_b);
    R *Bt_in = (R *)take(bt_b);   // U rows of the inner panels (read by the in-panel updates)
    R *Bt_out = (R *)take(bt_b);  // U rows of the outer panel (read by the trailing update)
    PermWs<R> pws;
    pws.carve(take(perm_b));
    unsigned long long *x_pk = (unsigned long long *)take(x_pk_b + x_rw_b + x_cw_b);
    unsigned long long *x_rw = x_pk + x_pk_b / 8, *x_cw = x_rw + x_rw_b / 8;
    R *x_val = (R *)take(x_val_b);
    UpdatePlanes planes;
    if (split_update) planes.carve(take(planes_b), N, nb);
    cx<R> *uinv = (cx<R> *)take(uinv_b);  // U11^-1 of the current pivot-free panel
    // growth flags and pivots: the context's for one matrix, each slab's for a batch
    int *flagsp = batched ? (int *)take(256) : ctx->flags;
    int *ipivp = batched ? (int *)take((size_t)N * sizeof(int)) : ipiv;
    // slots, barrier and flags start at zero (fill kernels, so that the
    // sequence can be captured in a graph)
    auto zero_state = [&]() -> emfa_status {
        const int nw = (int)((x_pk_b + x_rw_b + x_cw_b) / 4);
        k_zero_u32<<<(nw + 255) / 256, 256, 0, ctx->stream>>>((unsigned *)x_pk, nw);
        k_zero_u32<<<1, 64, 0, ctx->stream>>>(bar, 64);
        EMFA_HIP_CHECK(hipGetLastError());
        k_zero_u32<<<dim3(1, 1, ctx->batch.count), 64, 0, ctx->stream>>>((unsigned *)flagsp, 16, ctx->batch);
        EMFA_HIP_CHECK(hipGetLastError());
        return EMFA_OK;
    };
    unsigned bar_base = 0;
    // Cooperative launches of concurrent contexts are serialised by the
    // runtime, so contexts sharing the device (cu_share > 0) launch the
    // panel normally; its workgroups still become co-resident because the
    // shares sum to at most the CU count and nothing else waits on them.
    bool coop = ctx->cu_share <= 0, lookahead = true, xchg = true;
    if (const char *s = getenv("EMFA_PANEL_COOP")) coop = atoi(s) != 0;
    if (ctx->panel_x >= 0) xchg = ctx->panel_x != 0;
    if (const char *s = getenv("EMFA_PANEL_X")) xchg = atoi(s) != 0;
    EMFA_REQUIRE(G_max <= kPanelMaxG, EMFA_E_INVALID, "panel grid %d exceeds %d", G_max, kPanelMaxG);
    if (sizeof(R) == 4 && N >= 0xFFFFF) xchg = false;  // the FP32 packet holds a 20-bit row index
    if (const char *s = getenv("EMFA_LU_LOOKAHEAD")) lookahead = atoi(s) != 0;
    // FP32 factorisations may run without interchanges (panel variant 2)
    bool nopiv = false;
    if constexpr (sizeof(R) == 4) {
        nopiv = ctx->panel_x == 2 && !ctx->force_pivoting;
        if (const char *s = getenv("EMFA_LU_NOPIV")) nopiv = atoi(s) != 0 && !ctx->force_pivoting;
    }
    if (interchanges) *interchanges = nopiv ? 0 : 1;
    EMFA_REQUIRE(!batched || nopiv, EMFA_E_INVALID, "LU batch: pivot-free panels only");
    const R tau2 = R(kNoPivGrowth) * R(kNoPivGrowth);

    // Factor the outer panel [j0, j0+jb): inner panels of width w with the
    // in-panel row interchanges, triangular solves and updates (main stream).
    auto factor_outer = [&](int j0, int jb) -> emfa_status {
        int js = j0;
        while (nopiv && js < j0 + jb) {
            const int w = std::min(std::min(plan.w, kNoPivW), j0 + jb - js);
            const bool right = js + w < j0 + jb;  // panel columns right of the block
            if constexpr (sizeof(R) == 4) {
                ProfScope ps(ctx, "lu_panel", 0);
                const dim3 g1(1, 1, ctx->batch.count);
                k_diag_lu<R><<<g1, kDiagWaves * 64, 0, ctx->stream>>>(A, lda, js, w, ipivp, flagsp, tau2, uinv,
                                                                      ctx->batch);
                const int nl21 = (N - js - w + 31) / 32;
                if (right) {
                    // L21 and the block's U rows (one 64-row step of trsm_blocked) in one launch
                    const int c0 = js + w, c1 = j0 + jb, ncols = c1 - c0;
                    const int passes = std::max(1, std::min(8, ncols / (kTrsmColsPerPass * 1024)));
                    const int cpw = kTrsmColsPerPass * passes, nt = (ncols + cpw - 1) / cpw;
                    k_l21_u12<R><<<dim3(nl21 + nt, 1, ctx->batch.count), 256, 0, ctx->stream>>>(
                        A, lda, js, w, js + w, N, uinv, flagsp, tau2, nl21, c0, c1, Bt_in, 2LL * w, passes,
                        ctx->batch);
                } else if (nl21 > 0) {
                    k_l21<R><<<dim3(nl21, 1, ctx->batch.count), 256, 0, ctx->stream>>>(A, lda, js, w, js + w, N, uinv,
                                                                                        flagsp, tau2, ctx->batch);
                }
                EMFA_HIP_CHECK(hipGetLastError());
            }
            if (right) {
                emfa_status s2 = gemm_update<R>(ctx, "lu_gemm", N - js - w, j0 + jb - js - w, w,
                                                A + (size_t)(js + w) * lda + js, lda, Bt_in, 2LL * w,
                                                A + (size_t)(js + w) * lda + js + w, lda);
                if (s2) return s2;
            }
            js += w;
        }
        while (js < j0 + jb) {
            const int M = N - js;
            int w = std::min(plan.w, j0 + jb - js);
            int G = std::min(G_max, std::max(1, (M + panel_rows_min - 1) / panel_rows_min));
            int RB = (M + G - 1) / G;
            while ((size_t)(RB + 1) * (w + 1) * sizeof(cx<R>) > lds_cap && w > 8) w /= 2;
            EMFA_REQUIRE((size_t)(RB + 1) * (w + 1) * sizeof(cx<R>) <= lds_cap, EMFA_E_INVALID,
                         "LU panel does not fit LDS (N=%d)", N);
            size_t lds = (size_t)RB * (w + 1) * sizeof(cx<R>);
            size_t lds_x = (size_t)(RB + 1) * w * sizeof(cx<R>);
            {
                ProfScope ps(ctx, "lu_panel", 0);
                if (xchg) {
                    void *args[] = {(void *)&A,    (void *)&lda,  (void *)&js,    (void *)&M,
                                    (void *)&w,    (void *)&RB,   (void *)&ipiv,  (void *)&x_pk,
                                    (void *)&x_rw, (void *)&x_cw, (void *)&x_val, (void *)&ctx->flags};
                    if (!(coop && coop_launch(ctx->stream, (const void *)k_panel_x<R>, G, args, lds_x, &coop))) {
                        k_panel_x<R><<<G, 256, lds_x, ctx->stream>>>(A, lda, js, M, w, RB, ipiv, x_pk, x_rw, x_cw,
                                                                     x_val, ctx->flags);
                        EMFA_HIP_CHECK(hipGetLastError());
                    }
                } else {
                    void *args[] = {(void *)&A,        (void *)&lda,      (void *)&js,   (void *)&M,
                                    (void *)&w,        (void *)&RB,       (void *)&ipiv, (void *)&slot_rows,
                                    (void *)&slot_val, (void *)&slot_idx, (void *)&rowc, (void *)&bar,
                                    (void *)&bar_base, (void *)&ctx->flags};
                    // G workgroups of 256 threads with <= 160 KB LDS each always fit one per CU, and
                    // the concurrent update never waits on the panel, so all of them become
                    // resident; the barrier's bounded spin reports it if not.
                    if (!(coop && coop_launch(ctx->stream, (const void *)k_panel<R>, G, args, lds, &coop))) {
                        k_panel<R><<<G, 256, lds, ctx->stream>>>(A, lda, js, M, w, RB, ipiv, slot_rows, slot_val,
                                                                 slot_idx, rowc, bar, bar_base, ctx->flags);
                        EMFA_HIP_CHECK(hipGetLastError());
                    }
                }
            }
            bar_base += (unsigned)w * G;
            emfa_status s2 = laswp_launch<R>(ctx, A, lda, ipiv, js, js + w, j0, js, js + w, j0 + jb, pws);
            if (s2) return s2;
            if (js + w < j0 + jb) {
                s2 = trsm_blocked<R>(ctx, A, lda, A, lda, js, w, js + w, j0 + jb, 0, Bt_in, "lu_trsm");
                if (s2) return s2;
                s2 = gemm_update<R>(ctx, "lu_gemm", N - js - w, j0 + jb - js - w, w,
                                    A + (size_t)(js + w) * lda + js, lda, Bt_in, 2LL * w,
                                    A + (size_t)(js + w) * lda + js + w, lda);
                if (s2) return s2;
            }
            js += w;
        }
        return EMFA_OK;
    };
    // Apply the outer panel's interchanges outside it and solve for its U rows.
    auto outer_swap_trsm = [&](int j0, int jb) -> emfa_status {
        emfa_status s2 = nopiv ? EMFA_OK : laswp_launch<R>(ctx, A, lda, ipiv, j0, j0 + jb, 0, j0, j0 + jb, N, pws);
        if (s2) return s2;
        if (j0 + jb < N) return trsm_blocked<R>(ctx, A, lda, A, lda, j0, jb, j0 + jb, N, 0, Bt_out, "lu_trsm");
        return EMFA_OK;
    };

    // the whole launch sequence of the factorisation
    auto body = [&]() -> emfa_status {
    if ((st = zero_state())) return st;
    bar_base = 0;
    int jb0 = std::min(nb, N);
    if ((st = factor_outer(0, jb0))) return st;
    if ((st = outer_swap_trsm(0, jb0))) return st;
    for (int j0 = 0; j0 < N; j0 += nb) {
        const int jb = std::min(nb, N - j0);
        const int next = j0 + jb;
        if (next >= N) break;
        const int nbn = std::min(nb, N - next);
        const int M = N - next;
        const long long ldbt = 2LL * jb;
        const cx<R> *Aop = A + (size_t)next * lda + j0;
        // C[:, c0 : c0 + ncols] of the trailing block -= L U[:, c0 : c0 + ncols]
        auto trailing = [&](int c0, int ncols, hipStream_t s) -> emfa_status {
            cx<R> *C = A + (size_t)next * lda + next + c0;
            if constexpr (sizeof(R) == 4)
                if (split_update) return planes.update(ctx, M, c0, ncols, jb, C, lda, s);
            return gemm_update<R>(ctx, "lu_gemm", M, ncols, jb, Aop, lda, Bt_out + 2LL * c0 * ldbt, ldbt, C, lda, s);
        };
        if constexpr (sizeof(R) == 4)
            if (split_update && (st = planes.split(ctx, M, M, jb, Aop, lda, Bt_out, ldbt))) return st;
        if (lookahead && N - next - nbn > 0) {
            // G2 (all columns but the next panel's) on the second stream, G1 and the next
            // panel's factorisation on the main stream meanwhile.
            EMFA_HIP_CHECK(hipEventRecord(ctx->ev[0], ctx->stream));
            EMFA_HIP_CHECK(hipStreamWaitEvent(ctx->aux, ctx->ev[0], 0));
            st = trailing(nbn, N - next - nbn, ctx->aux);
            // The second stream joins the main one on every path, errors
            // included: callers order the context's buffers (and their return
            // to the device pool) through the main stream alone.
            hipError_t je = hipEventRecord(ctx->ev[1], ctx->aux);
            if (!st) st = trailing(0, nbn, ctx->stream);
            if (!st) st = factor_outer(next, nbn);
            if (je == hipSuccess) je = hipStreamWaitEvent(ctx->stream, ctx->ev[1], 0);
            if (je != hipSuccess) {
                (void)hipStreamSynchronize(ctx->aux);
                if (!st) {
                    set_error("lu: joining the look-ahead stream failed: %s", hipGetErrorString(je));
                    st = EMFA_E_HIP;
                }
            }
            if (st) return st;
        } else {
            if ((st = trailing(0, M, ctx->stream))) return st;
            if ((st = factor_outer(next, nbn))) return st;
        }
        if ((st = outer_swap_trsm(next, nbn))) return st;
    }
    return EMFA_OK;
    };
    if (batched && ctx->lu_graphs && !ctx->prof) {
        // everything the launch sequence depends on; the matrices themselves
        // are read through the batch's device array of bases
        // (the CU share enters only as shared / alone: the workgroup caps it
        // sets are baked in at capture, and any cap gives the same results)
        char key[320];
        std::snprintf(key, sizeof(key), "%d %d %d %d %d %d %d %d %d %d %d %d %p %zu %p %d", N, nb, plan.w,
                      ctx->batch.count, G_max, (int)split_update, (int)nopiv, (int)xchg, (int)lookahead, (int)coop,
                      (int)(ctx->cu_share > 0), ctx->num_cu, scr, slab, (const void *)ctx->batch.mat, panel_rows_min);
        st = lu_graph_run(ctx, key, body);
    } else {
        st = body();
    }
    if (st) return st;
    if (batched) {
        // per-matrix growth flags from each slab
        std::vector<int> hf(3 * (size_t)ctx->batch.count, 0);
        for (int z = 0; z < ctx->batch.count; ++z)
            EMFA_HIP_CHECK(hipMemcpyAsync(&hf[3 * (size_t)z], (const char *)flagsp + (size_t)z * slab, 3 * sizeof(int),
                                          hipMemcpyDeviceToHost, ctx->stream));
        EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (int z = 0; z < ctx->batch.count; ++z)
            if (growth) growth[z] = hf[3 * (size_t)z + 2] != 0;
        *info_out = 0;
        return EMFA_OK;
    }
    int flags[3] = {0, 0, 0};
    EMFA_HIP_CHECK(hipMemcpyAsync(flags, ctx->flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (flags[0] != 0) {
        set_error("LU panel grid barrier timed out (workgroups not co-resident)");
        return EMFA_E_HIP;
    }
    *info_out = flags[1];
    if (nopiv && flags[2] != 0) {
        set_error("LU without interchanges: zero pivot or multiplier above %g", kNoPivGrowth);
        return EMFA_E_NOCONV;
    }
    return EMFA_OK;
}

emfa_status lu_factor_inplace(emfa_ctx *ctx, emfa_mat *A, int *ipiv, int *info, int *interchanges) {
    if (A->prec == EMFA_F32) return lu_impl<float>(ctx, A, ipiv, info, interchanges);
    return lu_impl<double>(ctx, A, ipiv, info, interchanges);
}

// Solve with the factor (same precision as B).
template <typename R>
static emfa_status solve_impl(emfa_ctx *ctx, const emfa_mat *LUm, const int *ipiv, const int *perm, emfa_mat *Bm,
                              const void *dinv, int interchanges, const void *gh = nullptr) {
    const int N = LUm->rows, m = Bm->cols;
    const cx<R> *F = (const cx<R> *)LUm->d;
    cx<R> *B = (cx<R> *)Bm->d;
    const long long ldf = N, ldb = m;
    if (dinv && !getenv("EMFA_SOLVE_BLOCKED")) {
        // two launches per 256-row block with the diagonal-block inverses:
        // forward W -> Y (W = P B, or B itself without interchanges), backward Y -> B
        const int nblk = (N + kDinvB - 1) / kDinvB;
        const cx<R> *DL = (const cx<R> *)dinv, *DU = DL + (size_t)nblk * kDinvB * kDinvB;
        const size_t vec_b = (((size_t)N * m * sizeof(cx<R>)) + 255) & ~(size_t)255;
        const size_t bt_b = (((size_t)2 * m * 2 * kDinvB * sizeof(R)) + 255) & ~(size_t)255;
        void *scr = nullptr;
        emfa_status st = ctx_scratch(ctx, 2 * vec_b + 2 * bt_b, &scr);
        if (st) return st;
        cx<R> *Y = (cx<R> *)scr;
        cx<R> *W = interchanges ? (cx<R> *)((char *)scr + vec_b) : B;
        R *Bt = (R *)((char *)scr + 2 * vec_b);
        R *Bw = (R *)((char *)scr + 2 * vec_b + bt_b);  // expanded W_b (MFMA apply)
        ProfScope ps(ctx, "solve", 8.0 * (double)N * N * m);
        if (interchanges) {
            const long long tot = (long long)N * m;
            k_gather_rows<R><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(B, perm, N, m, W);
            EMFA_HIP_CHECK(hipGetLastError());
        }
        const dim3 g_apply(kDinvB / 32, (m + 31) / 32);
        // k_smallmm up to 64 columns; above that the expanded-operand MFMA
        // GEMM.  Contexts that share the device take the GEMM from 32
        // columns: k_smallmm is VALU-bound there and the MFMA product needs
        // about half the CU time (served 64-request stream: 193 -> 201
        // solves/s, median of 14 passes), while a lone solve finishes sooner
        // on k_smallmm's many small workgroups.
        int skinny_max = ctx->cu_share > 0 ? 31 : kSkinnyMaxCols;
        if (const char *e = getenv("EMFA_SKINNY_MAX")) skinny_max = std::min(kSkinnyMaxCols, atoi(e));
        const bool skinny = m <= skinny_max;
        const char *apply_env = getenv("EMFA_APPLY_MFMA");  // 0: k_dinv_apply for every width
        const bool apply_mfma = !(apply_env && atoi(apply_env) == 0);
        // fewest columns for the MFMA applies: above 64 only (from 32 columns in
        // server contexts the served 64-request stream measured 187-192
        // against 195-196 solves/s: three launches instead of one per block)
        int apply_min = kSkinnyMaxCols + 1;
        if (const char *e = getenv("EMFA_APPLY_MFMA_MIN")) apply_min = std::max(1, atoi(e));
        // rows [c_r0, c_r0 + M) of C -= F[rows, cols r0 .. r0 + wb) X_b
        auto update = [&](int M, const cx<R> *Fa, const cx<R> *Xb, cx<R> *C, int wb) -> emfa_status {
            if (M <= 0) return EMFA_OK;
            if (!skinny) return gemm_update<R>(ctx, "solve_gemm", M, m, wb, Fa, ldf, Bt, 2LL * wb, C, m);
            ProfScope ps2(ctx, "solve_gemm", 8.0 * (double)M * m * wb);
            return smallmm<R, R>(ctx->stream, Fa, ldf, Xb, m, C, m, M, wb, m, 1, wb, R(-1), 1, nullptr);
        };
        // Y_b = D_b^-1 W_b
        auto apply = [&](const cx<R> *D, int wb, const cx<R> *Wb, cx<R> *Yb) -> emfa_status {
            if (skinny) return smallmm<R, R>(ctx->stream, D, kDinvB, Wb, m, Yb, m, wb, wb, m, 1, wb, R(1), 0, nullptr);
            if (sizeof(R) == 4 && m >= apply_min && apply_mfma) {
                // many FP32 right-hand sides: Y_b = D_b W_b on the MFMA GEMM
                // (W_b expanded first), then Y_b expanded for the update.
                // k_dinv_apply ran 48 VALU workgroups at ~2 TFLOP/s here
                // (181 columns: 42 us per block)
                const long long tot = (long long)wb * m;
                k_expand<R><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(Wb, wb, m, Bw);
                EMFA_HIP_CHECK(hipGetLastError());
                emfa_status s2 = gemm_update<R>(ctx, "solve_apply", wb, m, wb, D, kDinvB, Bw, 2LL * wb, Yb, m,
                                                nullptr, 0, 0);
                if (s2) return s2;
                k_expand<R><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(Yb, wb, m, Bt);
                EMFA_HIP_CHECK(hipGetLastError());
                return EMFA_OK;
            }
            k_dinv_apply<R><<<g_apply, 256, 0, ctx->stream>>>(D, wb, Wb, m, m, Yb, m, Bt);
            EMFA_HIP_CHECK(hipGetLastError());
            return EMFA_OK;
        };
        // with the inverse-scaled blocks GH a step is one launch: the apply and
        // the other rows' update both read the step's input block
        const cx<R> *GH = skinny ? (const cx<R> *)gh : nullptr;
        for (int b = 0; b < nblk; ++b) {
            const int r0 = b * kDinvB, wb = std::min(kDinvB, N - r0);
            if (GH) {
                st = solve_step<R>(ctx->stream, DL + (size_t)b * kDinvB * kDinvB, GH + (size_t)(r0 + wb) * ldf + r0,
                                   ldf, W + (size_t)r0 * m, m, Y + (size_t)r0 * m, W + (size_t)(r0 + wb) * m, wb,
                                   N - r0 - wb);
                if (st) return st;
                continue;
            }
            st = apply(DL + (size_t)b * kDinvB * kDinvB, wb, W + (size_t)r0 * m, Y + (size_t)r0 * m);
            if (st) return st;
            st = update(N - r0 - wb, F + (size_t)(r0 + wb) * ldf + r0, Y + (size_t)r0 * m, W + (size_t)(r0 + wb) * m,
                        wb);
            if (st) return st;
        }
        for (int b = nblk - 1; b >= 0; --b) {
            const int r0 = b * kDinvB, wb = std::min(kDinvB, N - r0);
            if (GH) {
                st = solve_step<R>(ctx->stream, DU + (size_t)b * kDinvB * kDinvB, GH + r0, ldf, Y + (size_t)r0 * m, m,
                                   B + (size_t)r0 * m, Y, wb, r0);
                if (st) return st;
                continue;
            }
            st = apply(DU + (size_t)b * kDinvB * kDinvB, wb, Y + (size_t)r0 * m, B + (size_t)r0 * m);
            if (st) return st;
            st = update(r0, F + r0, B + (size_t)r0 * m, Y, wb);
            if (st) return st;
        }
        return EMFA_OK;
    }
    // few right-hand sides: one look-back launch per direction and column chunk
    // (opt-in: the coherent tile loads make it slower than the blocked path)
    int lb_max = 0;
    if (const char *s = getenv("EMFA_SOLVE_LB_MAX")) lb_max = std::min(atoi(s), 1 << 20);
    if (perm && m <= lb_max) {
        const int nblk = (N + 63) / 64;
        size_t yb = (((size_t)N * m * sizeof(cx<R>)) + 255) & ~(size_t)255;
        void *scr = nullptr;
        emfa_status st = ctx_scratch(ctx, yb + (size_t)nblk * sizeof(int) + 256, &scr);
        if (st) return st;
        cx<R> *Y = (cx<R> *)scr;
        int *bflags = (int *)((char *)scr + yb);
        ProfScope ps(ctx, "solve_trsm", 8.0 * (double)N * N * m);
        k_zero_u32<<<(nblk + 255) / 256, 256, 0, ctx->stream>>>((unsigned *)bflags, nblk);
        for (int c0 = 0, q = 0; c0 < m; c0 += kLbCols, ++q)
            k_trsv_lb<R, false><<<nblk, 256, 0, ctx->stream>>>(F, ldf, N, B, ldb, perm, Y, ldb, c0,
                                                                std::min(kLbCols, m - c0), bflags, 2 * q + 1,
                                                                ctx->flags);
        for (int c0 = 0, q = 0; c0 < m; c0 += kLbCols, ++q)
            k_trsv_lb<R, true><<<nblk, 256, 0, ctx->stream>>>(F, ldf, N, Y, ldb, nullptr, B, ldb, c0,
                                                               std::min(kLbCols, m - c0), bflags, 2 * q + 2,
                                                               ctx->flags);
        EMFA_HIP_CHECK(hipGetLastError());
        return EMFA_OK;
    }
    int nb = 256;
    if (const char *s = getenv("EMFA_SOLVE_NB")) nb = std::max(16, atoi(s));
    size_t bt_b = (((size_t)2 * m * 2 * nb * sizeof(R)) + 255) & ~(size_t)255;
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, bt_b + PermWs<R>::bytes(m) + 1024, &scr);
    if (st) return st;
    R *Bt = (R *)scr;
    PermWs<R> pws;
    pws.carve((char *)scr + bt_b);
    st = laswp_launch<R>(ctx, B, ldb, ipiv, 0, N, 0, m, 0, 0, pws);
    if (st) return st;
    for (int r0 = 0; r0 < N; r0 += nb) {
        int wb = std::min(nb, N - r0);
        st = trsm_blocked<R>(ctx, F, ldf, B, ldb, r0, wb, 0, m, 0, Bt, "solve_trsm");
        if (st) return st;
        st = gemm_update<R>(ctx, "solve_gemm", N - r0 - wb, m, wb, F + (size_t)(r0 + wb) * ldf + r0, ldf, Bt,
                            2LL * wb, B + (size_t)(r0 + wb) * ldb, ldb);
        if (st) return st;
    }
    int last = ((N - 1) / nb) * nb;
    for (int r0 = last; r0 >= 0; r0 -= nb) {
        int wb = std::min(nb, N - r0);
        st = trsm_blocked<R>(ctx, F, ldf, B, ldb, r0, wb, 0, m, 1, Bt, "solve_trsm");
        if (st) return st;
        st = gemm_update<R>(ctx, "solve_gemm", r0, m, wb, F + r0, ldf, Bt, 2LL * wb, B, ldb);
        if (st) return st;
    }
    return EMFA_OK;
}

emfa_status lu_solve_inplace(emfa_ctx *ctx, const emfa_lu *lu, emfa_mat *B) {
    if (lu->a->prec == EMFA_F32)
        return solve_impl<float>(ctx, lu->a, lu->ipiv, lu->perm, B, lu->dinv, lu->interchanges, lu->gh);
    return solve_impl<double>(ctx, lu->a, lu->ipiv, lu->perm, B, lu->dinv, lu->interchanges);
}

// Inverses of the diagonal blocks of L (unit) and U, both [nblk][256][256]
// FP32 diagonal-block inverses of one factor, or (bf, bd) of a batch of factors
static void dinv_launch_f32(hipStream_t s, const cx<float> *F, int N, cx<float> *D, int nblk, const Batch &bf,
                            const Batch &bd) {
    k_dinv_diag<<<dim3(nblk * 4, 2, bf.count), 256, 0, s>>>(F, N, N, D, nblk, bf, bd);
    for (int d = 1; d < 4; ++d) k_dinv_off<<<dim3(nblk * (4 - d), 2, bf.count), 256, 0, s>>>(F, N, N, D, nblk, d, bf, bd);
}

// the inverse-scaled off-diagonal blocks (GH) after the inverses
static void gh_launch_f32(hipStream_t s, const cx<float> *F, int N, const cx<float> *D, int nblk, cx<float> *GH,
                          const Batch &bf, const Batch &bd, const Batch &bg) {
    k_dinv_scale<<<dim3((N + 63) / 64, nblk * 8, bf.count), 256, 0, s>>>(F, N, D, nblk, GH, bf, bd, bg);
}

// With EMFA_SOLVE_GH=1, FP32 factors also keep L21 DL^-1 / U12 DU^-1 for
// one-launch solve steps: a solve at N = 7680 takes 1.3 ms instead of 2.5,
// but building them costs ~2.5 ms per factor there and n^2 more memory, and
// in a saturated solve server (factors reused ~1.6 times, eight workers) it
// cost 13 % of the throughput, so it is off by default.
static bool gh_wanted() {
    const char *e = getenv("EMFA_SOLVE_GH");
    return e && atoi(e) != 0;
}

static emfa_status build_dinv(emfa_ctx *ctx, emfa_lu *lu) {
    const int N = lu->a->rows, nblk = (N + kDinvB - 1) / kDinvB;
    const size_t e = lu->a->prec == EMFA_F32 ? sizeof(cx<float>) : sizeof(cx<double>);
    lu->dinv_bytes = 2 * (size_t)nblk * kDinvB * kDinvB * e;
    if (ctx_alloc(ctx, &lu->dinv, lu->dinv_bytes) != hipSuccess) {
        (void)hipGetLastError();
        lu->dinv = nullptr;
        set_error("emfa_lu_factor: diagonal-block inverse allocation failed");
        return EMFA_E_OOM;
    }
    ProfScope ps(ctx, "lu_dinv", 0);
    if (lu->a->prec == EMFA_F32) {
        const cx<float> *F = (const cx<float> *)lu->a->d;
        cx<float> *D = (cx<float> *)lu->dinv;
        dinv_launch_f32(ctx->stream, F, N, D, nblk, Batch(), Batch());
        if (gh_wanted()) {
            if (ctx_alloc(ctx, &lu->gh, (size_t)N * N * sizeof(cx<float>)) != hipSuccess) {
                (void)hipGetLastError();
                lu->gh = nullptr;
                set_error("emfa_lu_factor: scaled off-diagonal block allocation failed");
                return EMFA_E_OOM;
            }
            gh_launch_f32(ctx->stream, F, N, D, nblk, (cx<float> *)lu->gh, Batch(), Batch(), Batch());
        }
    } else {
        const cx<double> *F = (const cx<double> *)lu->a->d;
        cx<double> *D = (cx<double> *)lu->dinv;
        k_tri_inv<double><<<dim3(nblk, 2), 256, 0, ctx->stream>>>(F, N, N, D);
    }
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

// C = (acc ? C : 0) + sign * sum_s P_s, slices summed in index order
__global__ __launch_bounds__(256) void k_sum_slices(const double *__restrict__ P, int S, long long n,
                                                    double *__restrict__ C, double sign, int acc) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = 0;
    for (int s = 0; s < S; ++s) v += P[(size_t)s * n + i];
    C[i] = (acc ? C[i] : 0.0) + sign * v;
}

// C = beta C + sign * A X in FP64 (A any precision, X and C FP64).
emfa_status gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *X, emfa_mat *C, int negate,
                          int accumulate) {
    const int N = A->rows, K = A->cols, m = X->cols;
    EMFA_REQUIRE(X->rows == K && C->rows == N && C->cols == m && X->prec == EMFA_F64 && C->prec == EMFA_F64,
                 EMFA_E_INVALID, "gemm_residual: shape/precision mismatch");
    // k_smallmm on the VALU up to 64 columns, the FP64 MFMA GEMM above.
    // Contexts that share the device take the GEMM from 17 columns: the
    // served 64-request stream (1045-14981 unknowns, 1-64 excitations per
    // request) measured 116.3 against 107.2 solves/s (median of 3 passes),
    // the VALU residual having held 18 % of the CUs' time.
    static const char *resid_env = [] {
        const char *e = getenv("EMFA_RESID_SKINNY_MAX");
        return e ? e : getenv("EMFA_SKINNY_MAX");
    }();
    const int resid_skinny_max =
        resid_env ? std::min(kSkinnyMaxCols, atoi(resid_env)) : ctx->cu_share > 0 ? 16 : kSkinnyMaxCols;
    if (m <= resid_skinny_max) {
        // row blocks x K slices (summed in slice order) on the VALU
        const int mc = m <= 8 ? 8 : m <= 16 ? 16 : m <= 32 ? 32 : 64;
        // tall residuals with wide column tiles: 8 rows per lane (X re-read 4x less)
        int rpt = N >= 2048 && mc >= 32 ? 8 : 2;
        if (const char *e = getenv("EMFA_RESID_RPT")) rpt = atoi(e) == 8 && mc >= 32 ? 8 : 2;
        const int rbk = rpt * (256 / mc);  // rows per workgroup of k_smallmm
        const long long wgs = (N + rbk - 1) / rbk;
        int S = (int)std::min<long long>(32, std::max<long long>(1, (4LL * ctx->num_cu + wgs - 1) / wgs));
        if (const char *e = getenv("EMFA_RESID_SPLIT")) S = std::max(1, std::min(64, atoi(e)));
        const int kc = (int)(((long long)K + S - 1) / S + 63) / 64 * 64;
        S = (K + kc - 1) / kc;
        const size_t p_b = S > 1 ? (size_t)S * N * m * sizeof(cx<double>) : 0;
        void *scr = nullptr;
        if (p_b) {
            emfa_status st = ctx_scratch(ctx, p_b + 256, &scr);
            if (st) return st;
        }
        cx<double> *P = (cx<double> *)scr;
        ProfScope ps(ctx, "refine_gemm", 8.0 * (double)N * m * K);
        const double sign = negate ? -1.0 : 1.0;
        emfa_status s2 =
            A->prec == EMFA_F32
                ? smallmm<float, double>(ctx->stream, (const cx<float> *)A->d, K, (const cx<double> *)X->d, m,
                                         (cx<double> *)C->d, m, N, K, m, S, kc, sign, accumulate, P, rpt)
                : smallmm<double, double>(ctx->stream, (const cx<double> *)A->d, K, (const cx<double> *)X->d, m,
                                          (cx<double> *)C->d, m, N, K, m, S, kc, sign, accumulate, P, rpt);
        if (s2) return s2;
        if (S > 1) {
            const long long tot2 = (long long)N * 2 * m;
            k_sum_slices<<<(unsigned)((tot2 + 255) / 256), 256, 0, ctx->stream>>>((const double *)P, S, tot2,
                                                                                  (double *)C->d, sign, accumulate);
            EMFA_HIP_CHECK(hipGetLastError());
        }
        return EMFA_OK;
    }
    size_t bt_b = (((size_t)2 * m * 2 * K * sizeof(double)) + 255) & ~(size_t)255;
    // Few right-hand sides leave most CUs idle: split the inner dimension
    // into S slices (multiples of 64 reals) with partial results summed in a
    // fixed order afterwards.
    const long long tiles = (long long)((N + 63) / 64) * ((2 * m + 63) / 64);
    int S = (int)std::min<long long>(16, std::max<long long>(1, (2LL * ctx->num_cu + tiles - 1) / tiles));
    if (const char *e = getenv("EMFA_RESID_SPLIT")) S = std::max(1, std::min(64, atoi(e)));
    long long kc = ((2LL * K + S - 1) / S + 63) / 64 * 64;
    S = (int)((2LL * K + kc - 1) / kc);
    const size_t p_b = S > 1 ? (size_t)S * N * 2 * m * sizeof(double) : 0;
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, bt_b + p_b + 1024, &scr);
    if (st) return st;
    double *Bt = (double *)scr;
    double *P = (double *)((char *)scr + bt_b);
    long long tot = (long long)K * m;
    k_expand<double><<<(int)((tot + 255) / 256), 256, 0, ctx->stream>>>((const cx<double> *)X->d, K, m, Bt);
    EMFA_HIP_CHECK(hipGetLastError());
    GemmArgs g;
    g.M = N;
    g.Nr = 2 * m;
    g.Kr = 2 * K;
    g.A = A->d;
    g.lda = 2LL * K;
    g.Bt = Bt;
    g.ldb = 2LL * K;
    g.C = C->d;
    g.ldc = 2LL * m;
    g.alpha_neg = negate ? 1 : 0;
    g.beta = accumulate ? 1 : 0;
    ProfScope ps(ctx, "refine_gemm", 8.0 * (double)N * m * K);
    if (S == 1) return A->prec == EMFA_F32 ? gemm_launch_f32_f64(ctx->stream, g) : gemm_launch_f64(ctx->stream, g);
    {
        // one launch, slices on blockIdx.z, each writing its own plane of P
        GemmArgs gs = g;
        gs.kslices = S;
        gs.kc = (int)kc;
        gs.cslice = (long long)N * 2 * m;
        gs.C = P;
        gs.ldc = 2LL * m;
        gs.alpha_neg = 0;
        gs.beta = 0;
        emfa_status s2 = A->prec == EMFA_F32 ? gemm_launch_f32_f64(ctx->stream, gs) : gemm_launch_f64(ctx->stream, gs);
        if (s2) return s2;
    }
    const long long tot2 = (long long)N * 2 * m;
    k_sum_slices<<<(unsigned)((tot2 + 255) / 256), 256, 0, ctx->stream>>>(P, S, tot2, (double *)C->d,
                                                                          negate ? -1.0 : 1.0, accumulate);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

// The configuration that factors a batch with one launch sequence: FP32,
// pivot-free panels, the split update, inverse-block solves.
static bool lu_batch_config(const emfa_ctx *ctx, int prec) {
    return prec == EMFA_F32 && ctx->panel_x == 2 && ctx->lu_update_mode == 1 && !ctx->force_pivoting &&
           !getenv("EMFA_SOLVE_BLOCKED") && !getenv("EMFA_LU_NOPIV") && !getenv("EMFA_LU_BF16X3") &&
           !getenv("EMFA_LU_NOBATCH");
}

extern "C" {

// Compose the interchanges of lu->ipiv into one row permutation (used by the
// solves): row i of P A is row perm[i] of A.
static emfa_status compose_perm(emfa_ctx *ctx, emfa_lu *lu) {
    const int n = lu->a->rows;
    std::vector<int> piv(n), perm(n);
    // Copies run on the context's stream: blocks from the shared device pool
    // are ordered against their previous owner only through that stream.
    if (hipMemcpyAsync(piv.data(), lu->ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess ||
        (!lu->perm && ctx_alloc(ctx, (void **)&lu->perm, sizeof(int) * n) != hipSuccess)) {
        (void)hipGetLastError();
        set_error("emfa_lu_factor: permutation setup failed");
        return EMFA_E_HIP;
    }
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        const int p = piv[i];
        if (p < 0 || p >= n) {
            set_error("emfa_lu_factor: pivot %d out of range at column %d", p, i);
            return EMFA_E_HIP;
        }
        std::swap(perm[i], perm[p]);
    }
    if (hipMemcpyAsync(lu->perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("emfa_lu_factor: permutation upload failed");
        return EMFA_E_HIP;
    }
    return EMFA_OK;
}

emfa_status emfa_lu_factor(emfa_ctx *ctx, const emfa_mat *Z, int prec, emfa_lu **out) {
    EMFA_REQUIRE(ctx && Z && out, EMFA_E_INVALID, "emfa_lu_factor: null argument");
    EMFA_REQUIRE(Z->rows == Z->cols, EMFA_E_INVALID, "emfa_lu_factor: matrix is %dx%d, not square", Z->rows,
                 Z->cols);
    EMFA_REQUIRE(prec == EMFA_F32 || prec == EMFA_F64, EMFA_E_INVALID, "emfa_lu_factor: bad precision");
    // graph-captured factorisations go through the batch path as batches of one
    if (ctx->lu_graphs && lu_batch_config(ctx, prec)) return emfa_lu_factor_batch(ctx, &Z, 1, prec, out);
    *out = nullptr;
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    emfa_lu *lu = new emfa_lu();
    lu->ctx = ctx;
    emfa_status st = emfa_mat_create(ctx, Z->rows, Z->cols, prec, &lu->a);
    if (st) {
        delete lu;
        return st;
    }
    if (ctx_alloc(ctx, (void **)&lu->ipiv, sizeof(int) * Z->rows) != hipSuccess) {
        (void)hipGetLastError();
        emfa_mat_destroy(lu->a);
        delete lu;
        set_error("emfa_lu_factor: pivot allocation failed");
        return EMFA_E_OOM;
    }
    st = convert_copy(ctx, Z, lu->a);
    if (!st) st = lu_factor_inplace(ctx, lu->a, lu->ipiv, &lu->info, &lu->interchanges);
    if (st == EMFA_E_NOCONV && lu->interchanges == 0) {
        // the factorisation without interchanges grew: redo it with partial pivoting
        ctx->force_pivoting = 1;
        st = convert_copy(ctx, Z, lu->a);
        if (!st) st = lu_factor_inplace(ctx, lu->a, lu->ipiv, &lu->info, &lu->interchanges);
        ctx->force_pivoting = 0;
        lu->pivot_fallback = 1;
    }
    if (!st) st = compose_perm(ctx, lu);
    if (!st && !getenv("EMFA_SOLVE_BLOCKED")) st = build_dinv(ctx, lu);
    if (st) {
        emfa_lu_destroy(lu);
        return st;
    }
    *out = lu;
    if (lu->info != 0) {
        set_error("emfa_lu_factor: exactly zero pivot at column %d", lu->info);
        return EMFA_E_SINGULAR;
    }
    return EMFA_OK;
}

emfa_status emfa_lu_factor_batch(emfa_ctx *ctx, const emfa_mat *const *Z, int count, int prec, emfa_lu **out) {
    EMFA_REQUIRE(ctx && Z && out && count >= 1, EMFA_E_INVALID, "emfa_lu_factor_batch: bad arguments");
    for (int z = 0; z < count; ++z) out[z] = nullptr;
    for (int z = 0; z < count; ++z)
        EMFA_REQUIRE(Z[z] && Z[z]->rows == Z[z]->cols && Z[z]->rows == Z[0]->rows, EMFA_E_INVALID,
                     "emfa_lu_factor_batch: matrix %d is not square of the batch's size", z);
    EMFA_REQUIRE(prec == EMFA_F32 || prec == EMFA_F64, EMFA_E_INVALID, "emfa_lu_factor_batch: bad precision");
    const int n = Z[0]->rows;
    // one launch sequence for the batch: FP32, pivot-free panels, split update, inverse-block solves
    // (a batch of one when the context runs factorisations as graphs)
    const bool batchable = (count > 1 || ctx->lu_graphs) && lu_batch_config(ctx, prec);
    if (!batchable) {
        emfa_status worst = EMFA_OK;
        for (int z = 0; z < count; ++z) {
            emfa_status st = emfa_lu_factor(ctx, Z[z], prec, &out[z]);
            if (st == EMFA_E_SINGULAR) {
                worst = st;
            } else if (st) {
                for (int y = 0; y < z; ++y) {
                    emfa_lu_destroy(out[y]);
                    out[y] = nullptr;
                }
                return st;
            }
        }
        return worst;
    }
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    std::vector<emfa_lu *> lus(count, nullptr);
    void *dmat = nullptr, *ddinv = nullptr, *dgh = nullptr;
    auto fail = [&](emfa_status st) {
        (void)hipStreamSynchronize(ctx->stream);
        ctx->batch = Batch();
        if (dmat && dmat != ctx->bmat) ctx_free(ctx, dmat, sizeof(void *) * count);
        if (ddinv) ctx_free(ctx, ddinv, sizeof(void *) * count);
        if (dgh) ctx_free(ctx, dgh, sizeof(void *) * count);
        for (auto l : lus)
            if (l) emfa_lu_destroy(l);
        return st;
    };
    for (int z = 0; z < count; ++z) {
        emfa_lu *lu = new emfa_lu();
        lu->ctx = ctx;
        emfa_status st = emfa_mat_create(ctx, n, n, EMFA_F32, &lu->a);
        if (st) {
            delete lu;
            return fail(st);
        }
        lus[z] = lu;
        if (ctx_alloc(ctx, (void **)&lu->ipiv, sizeof(int) * n) != hipSuccess ||
            ctx_alloc(ctx, (void **)&lu->perm, sizeof(int) * n) != hipSuccess) {
            (void)hipGetLastError();
            set_error("emfa_lu_factor_batch: pivot allocation failed");
            return fail(EMFA_E_OOM);
        }
        if ((st = convert_copy(ctx, Z[z], lu->a))) return fail(st);
    }
    // device array of the factors' bases; host copies stay alive until the stream syncs below
    std::vector<const char *> hmat(count), hdinv(count), hgh(count);
    for (int z = 0; z < count; ++z) hmat[z] = (const char *)lus[z]->a->d;
    if (ctx->lu_graphs) {
        // the context's array of bases: its address is part of the captured graphs
        if (ctx->bmat_cap < (size_t)count) {
            if (ctx->bmat) {
                EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                ctx_free(ctx, ctx->bmat, ctx->bmat_cap * sizeof(void *));
                ctx->bmat = nullptr;
                ctx->bmat_cap = 0;
            }
            const size_t cap = std::max<size_t>(16, (size_t)count);
            if (ctx_alloc(ctx, &ctx->bmat, cap * sizeof(void *)) != hipSuccess) {
                (void)hipGetLastError();
                ctx->bmat = nullptr;
                set_error("emfa_lu_factor_batch: allocation failed");
                return fail(EMFA_E_OOM);
            }
            ctx->bmat_cap = cap;
        }
        dmat = ctx->bmat;
    } else if (ctx_alloc(ctx, &dmat, sizeof(void *) * count) != hipSuccess) {
        (void)hipGetLastError();
        set_error("emfa_lu_factor_batch: allocation failed");
        return fail(EMFA_E_OOM);
    }
    EMFA_HIP_CHECK(hipMemcpyAsync(dmat, hmat.data(), sizeof(void *) * count, hipMemcpyHostToDevice, ctx->stream));
    ctx->batch = Batch();
    ctx->batch.mat = (const char *const *)dmat;
    ctx->batch.mat0 = hmat[0];
    ctx->batch.count = count;
    std::vector<int> growth(count, 0);
    int info = 0, inter = 0;
    emfa_status st = lu_impl<float>(ctx, lus[0]->a, lus[0]->ipiv, &info, &inter, growth.data());
    ctx->batch = Batch();
    if (st) return fail(st);
    // no interchanges: identity pivots and permutation
    std::vector<int> iota(n);
    for (int i = 0; i < n; ++i) iota[i] = i;
    for (int z = 0; z < count; ++z) {
        lus[z]->interchanges = 0;
        lus[z]->info = 0;
        EMFA_HIP_CHECK(hipMemcpyAsync(lus[z]->ipiv, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
        EMFA_HIP_CHECK(hipMemcpyAsync(lus[z]->perm, iota.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // a factor that grew is redone alone with partial pivoting
    for (int z = 0; z < count; ++z) {
        if (!growth[z]) continue;
        emfa_lu *lu = lus[z];
        ctx->force_pivoting = 1;
        st = convert_copy(ctx, Z[z], lu->a);
        if (!st) st = lu_factor_inplace(ctx, lu->a, lu->ipiv, &lu->info, &lu->interchanges);
        ctx->force_pivoting = 0;
        lu->pivot_fallback = 1;
        if (!st) st = compose_perm(ctx, lu);
        if (st) return fail(st);
    }
    // inverses of the diagonal blocks of every factor, one launch sequence
    {
        const int nblk = (n + kDinvB - 1) / kDinvB;
        for (int z = 0; z < count; ++z) {
            emfa_lu *lu = lus[z];
            lu->dinv_bytes = 2 * (size_t)nblk * kDinvB * kDinvB * sizeof(cx<float>);
            if (ctx_alloc(ctx, &lu->dinv, lu->dinv_bytes) != hipSuccess) {
                (void)hipGetLastError();
                lu->dinv = nullptr;
                set_error("emfa_lu_factor_batch: diagonal-block inverse allocation failed");
                return fail(EMFA_E_OOM);
            }
            hdinv[z] = (const char *)lu->dinv;
        }
        if (ctx_alloc(ctx, &ddinv, sizeof(void *) * count) != hipSuccess) {
            (void)hipGetLastError();
            set_error("emfa_lu_factor_batch: allocation failed");
            return fail(EMFA_E_OOM);
        }
        EMFA_HIP_CHECK(hipMemcpyAsync(ddinv, hdinv.data(), sizeof(void *) * count, hipMemcpyHostToDevice, ctx->stream));
        Batch bf, bd;
        bf.mat = (const char *const *)dmat;
        bf.mat0 = hmat[0];
        bf.count = count;
        bd.mat = (const char *const *)ddinv;
        bd.mat0 = hdinv[0];
        bd.count = count;
        ProfScope ps(ctx, "lu_dinv", 0);
        dinv_launch_f32(ctx->stream, (const cx<float> *)lus[0]->a->d, n, (cx<float> *)lus[0]->dinv, nblk, bf, bd);
        EMFA_HIP_CHECK(hipGetLastError());
        if (gh_wanted()) {
            for (int z = 0; z < count; ++z) {
                if (ctx_alloc(ctx, &lus[z]->gh, (size_t)n * n * sizeof(cx<float>)) != hipSuccess) {
                    (void)hipGetLastError();
                    lus[z]->gh = nullptr;
                    set_error("emfa_lu_factor_batch: scaled off-diagonal block allocation failed");
                    return fail(EMFA_E_OOM);
                }
                hgh[z] = (const char *)lus[z]->gh;
            }
            if (ctx_alloc(ctx, &dgh, sizeof(void *) * count) != hipSuccess) {
                (void)hipGetLastError();
                set_error("emfa_lu_factor_batch: allocation failed");
                return fail(EMFA_E_OOM);
            }
            EMFA_HIP_CHECK(hipMemcpyAsync(dgh, hgh.data(), sizeof(void *) * count, hipMemcpyHostToDevice, ctx->stream));
            Batch bg;
            bg.mat = (const char *const *)dgh;
            bg.mat0 = hgh[0];
            bg.count = count;
            gh_launch_f32(ctx->stream, (const cx<float> *)lus[0]->a->d, n, (const cx<float> *)lus[0]->dinv, nblk,
                          (cx<float> *)lus[0]->gh, bf, bd, bg);
            EMFA_HIP_CHECK(hipGetLastError());
        }
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (dmat != ctx->bmat) ctx_free(ctx, dmat, sizeof(void *) * count);
    ctx_free(ctx, ddinv, sizeof(void *) * count);
    if (dgh) ctx_free(ctx, dgh, sizeof(void *) * count);
    emfa_status worst = EMFA_OK;
    for (int z = 0; z < count; ++z) {
        out[z] = lus[z];
        if (lus[z]->info != 0) {
            set_error("emfa_lu_factor_batch: exactly zero pivot in matrix %d at column %d", z, lus[z]->info);
            worst = EMFA_E_SINGULAR;
        }
    }
    return worst;
}

emfa_status emfa_lu_destroy(emfa_lu *lu) {
    if (!lu) return EMFA_OK;
    emfa_ctx *c = lu->ctx;
    (void)hipStreamSynchronize(c->stream);
    ctx_free(c, lu->ipiv, sizeof(int) * lu->a->rows);
    if (lu->perm) ctx_free(c, lu->perm, sizeof(int) * lu->a->rows);
    if (lu->dinv) ctx_free(c, lu->dinv, lu->dinv_bytes);
    if (lu->gh) ctx_free(c, lu->gh, (size_t)lu->a->rows * lu->a->rows * sizeof(cx<float>));
    emfa_mat_destroy(lu->a);
    delete lu;
    return EMFA_OK;
}

emfa_status emfa_lu_solve(emfa_ctx *ctx, const emfa_lu *lu, emfa_mat *B) {
    EMFA_REQUIRE(ctx && lu && B, EMFA_E_INVALID, "emfa_lu_solve: null argument");
    EMFA_REQUIRE(B->rows == lu->a->rows, EMFA_E_INVALID, "emfa_lu_solve: B has %d rows, factor %d", B->rows,
                 lu->a->rows);
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    emfa_status st = EMFA_OK;
    if (B->prec == lu->a->prec) {
        st = lu_solve_inplace(ctx, lu, B);
    } else {
        emfa_mat *T = nullptr;
        st = emfa_mat_create(ctx, B->rows, B->cols, lu->a->prec, &T);
        if (st) return st;
        st = convert_copy(ctx, B, T);
        if (!st) st = lu_solve_inplace(ctx, lu, T);
        if (!st) st = convert_copy(ctx, T, B);
        (void)hipStreamSynchronize(ctx->stream);
        emfa_mat_destroy(T);
    }
    if (st) return st;
    int fl = 0;
    EMFA_HIP_CHECK(hipMemcpyAsync(&fl, ctx->flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (fl != 0) {
        set_error("emfa_lu_solve: a bounded wait of the triangular solve timed out");
        return EMFA_E_HIP;
    }
    return EMFA_OK;
}

emfa_status emfa_lu_pivots(const emfa_lu *lu, int *ipiv_host) {
    EMFA_REQUIRE(lu && ipiv_host, EMFA_E_INVALID, "emfa_lu_pivots: null argument");
    EMFA_HIP_CHECK(hipMemcpyAsync(ipiv_host, lu->ipiv, sizeof(int) * lu->a->rows, hipMemcpyDeviceToHost,
                                  lu->ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(lu->ctx->stream));
    return EMFA_OK;
}

emfa_mat *emfa_lu_matrix(emfa_lu *lu) { return lu ? lu->a : nullptr; }
int emfa_lu_info(const emfa_lu *lu) { return lu ? lu->info : -1; }

emfa_status emfa_lu_set_params(emfa_ctx *ctx, int nb, int panel_workgroups, int panel_variant) {
    EMFA_REQUIRE(ctx && nb >= 0 && panel_workgroups >= 0 && panel_variant >= -1 && panel_variant <= 2,
                 EMFA_E_INVALID, "emfa_lu_set_params: bad argument");
    ctx->lu_nb = nb;
    ctx->panel_g = panel_workgroups;
    ctx->panel_x = panel_variant;
    return EMFA_OK;
}

emfa_status emfa_lu_get_pivoting(const emfa_lu *lu, int *interchanges, int *fallback) {
    EMFA_REQUIRE(lu, EMFA_E_INVALID, "emfa_lu_get_pivoting: null factor");
    if (interchanges) *interchanges = lu->interchanges;
    if (fallback) *fallback = lu->pivot_fallback;
    return EMFA_OK;
}

emfa_status emfa_lu_set_update_mode(emfa_ctx *ctx, int mode) {
    EMFA_REQUIRE(ctx && (mode == 0 || mode == 1), EMFA_E_INVALID, "emfa_lu_set_update_mode: bad argument");
    ctx->lu_update_mode = mode;
    return EMFA_OK;
}

emfa_status emfa_lu_autotune(emfa_ctx *ctx, int n, int prec, int for_refinement, emfa_lu_tuning *out) {
    EMFA_REQUIRE(ctx && n > 0 && (prec == EMFA_F32 || prec == EMFA_F64), EMFA_E_INVALID,
                 "emfa_lu_autotune: bad argument");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    // a well-conditioned test matrix of the requested size, generated on the device
    emfa_mat *A = nullptr;
    emfa_status st = emfa_mat_create(ctx, n, n, prec, &A);
    if (st) return st;
    {
        const long long tot = (long long)n * n;
        if (prec == EMFA_F32)
            k_tune_matrix<float><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>((cx<float> *)A->d, n);
        else
            k_tune_matrix<double><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>((cx<double> *)A->d, n);
        EMFA_HIP_CHECK(hipGetLastError());
    }
    const int save_nb = ctx->lu_nb, save_g = ctx->panel_g, save_x = ctx->panel_x, save_m = ctx->lu_update_mode;
    struct Cand {
        int nb, x, g, mode;
    };
    std::vector<Cand> cands;
    for (int nb : {256, 512})
        if (nb <= std::max(256, n))
            for (int x : {1, 0})
                for (int g : {0, 128}) cands.push_back({nb, x, g, 0});
    if (for_refinement && prec == EMFA_F32) {
        for (int g : {0, 128}) cands.push_back({512, 1, g, 1});
        for (int nb : {256, 512})
            if (nb <= std::max(256, n)) cands.push_back({nb, 2, 0, 1});
    }
    hipEvent_t e0, e1;
    EMFA_HIP_CHECK(hipEventCreate(&e0));
    EMFA_HIP_CHECK(hipEventCreate(&e1));
    float best_ms = 1e30f;
    Cand best = cands[0];
    int tried = 0;
    emfa_lu *lu = nullptr;
    for (size_t i = 0; i < cands.size() && !st; ++i) {
        ctx->lu_nb = cands[i].nb;
        ctx->panel_g = cands[i].g;
        ctx->panel_x = cands[i].x;
        ctx->lu_update_mode = cands[i].mode;
        if (i == 0) {  // warm-up (module load, pool)
            st = emfa_lu_factor(ctx, A, prec, &lu);
            emfa_lu_destroy(lu);
            lu = nullptr;
            if (st) break;
        }
        (void)hipEventRecord(e0, ctx->stream);
        st = emfa_lu_factor(ctx, A, prec, &lu);
        (void)hipEventRecord(e1, ctx->stream);
        (void)hipEventSynchronize(e1);
        emfa_lu_destroy(lu);
        lu = nullptr;
        if (st) break;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (out && i < 16) {
            out->cand_nb[i] = cands[i].nb;
            out->cand_variant[i] = cands[i].x;
            out->cand_workgroups[i] = cands[i].g;
            out->cand_update_mode[i] = cands[i].mode;
            out->cand_ms[i] = ms;
        }
        ++tried;
        if (ms < best_ms) {
            best_ms = ms;
            best = cands[i];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    emfa_mat_destroy(A);
    if (st) {
        ctx->lu_nb = save_nb;
        ctx->panel_g = save_g;
        ctx->panel_x = save_x;
        ctx->lu_update_mode = save_m;
        return st;
    }
    ctx->lu_nb = best.nb;
    ctx->panel_g = best.g;
    ctx->panel_x = best.x;
    ctx->lu_update_mode = best.mode;
    if (out) {
        out->nb = best.nb;
        out->panel_variant = best.x;
        out->panel_workgroups = best.g;
        out->update_mode = best.mode;
        out->best_ms = best_ms;
        out->candidates = tried;
    }
    return EMFA_OK;
}

emfa_status emfa_gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *Xin, emfa_mat *C, int negate,
                               int accumulate) {
    EMFA_REQUIRE(ctx && A && Xin && C, EMFA_E_INVALID, "emfa_gemm_residual: null argument");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return gemm_residual(ctx, A, Xin, C, negate, accumulate);
}

}  // extern "C"
