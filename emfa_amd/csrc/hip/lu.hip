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
#include <cmath>
#include <vector>

#include "gemm.h"
#include "internal.h"

namespace emfa {

constexpr int kPanelMaxW = 64;
constexpr unsigned kSpinLimit = 1u << 22;

// --------------------------------------------------------------------------
// grid barrier: monotonic counter; the spin is bounded and a timeout sets
// flags[0] so that every later barrier falls through (the kernel always exits)
// --------------------------------------------------------------------------
__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned target, int *flags) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(bar, 1u);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&flags[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > kSpinLimit) {
                atomicExch(&flags[0], 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __threadfence();
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
__device__ void block_argmax(R &v, int &r, R *sv, int *sr) {
    // wave reduction
    for (int off = 32; off > 0; off >>= 1) {
        R ov = __shfl_down(v, off, 64);
        int orr = __shfl_down(r, off, 64);
        if (better(ov, orr, v, r)) {
            v = ov;
            r = orr;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wave] = v;
        sr[wave] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (better(sv[w], sr[w], v, r)) {
                v = sv[w];
                r = sr[w];
            }
        sv[0] = v;
        sr[0] = r;
    }
    __syncthreads();
    v = sv[0];
    r = sr[0];
    __syncthreads();
}

// Panel: rows [j0, j0+M) x columns [j0, j0+w) of A (row-major, lda complex).
// Workgroup g owns rows [j0 + g*RB, j0 + min(M, (g+1)*RB)).
template <typename R>
__global__ __launch_bounds__(256) void k_panel(cx<R> *__restrict__ A, long long lda, int j0, int M, int w, int RB,
                                               int *__restrict__ ipiv, cx<R> *__restrict__ slot_rows,
                                               R *__restrict__ slot_val, int *__restrict__ slot_idx,
                                               cx<R> *__restrict__ rowc, unsigned *bar, unsigned bar_base,
                                               int *flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<R> *rows = (cx<R> *)smem_raw;  // [RB][w]
    __shared__ cx<R> urow[kPanelMaxW], orow[kPanelMaxW];
    __shared__ R sv[4];
    __shared__ int sr[4];
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int rbeg = j0 + g * RB;
    const int rend = min(j0 + M, rbeg + RB);
    const int nr = max(0, rend - rbeg);
    for (int i = tid; i < nr * w; i += 256) {
        int rr = i / w, cc = i % w;
        rows[i] = A[(size_t)(rbeg + rr) * lda + j0 + cc];
    }
    __syncthreads();
    for (int c = 0; c < w; ++c) {
        const int jc = j0 + c;
        const int par = c & 1;
        // local candidate
        R bv = R(-1);
        int br = -1;
        for (int i = tid; i < nr; i += 256) {
            int r = rbeg + i;
            if (r < jc) continue;
            R v = cabs1(rows[i * w + c]);
            if (better(v, r, bv, br)) {
                bv = v;
                br = r;
            }
        }
        block_argmax(bv, br, sv, sr);
        if (tid == 0) {
            slot_val[par * G + g] = bv;
            slot_idx[par * G + g] = br;
        }
        if (br >= 0 && tid < w) slot_rows[((size_t)par * G + g) * kPanelMaxW + tid] = rows[(br - rbeg) * w + tid];
        if (jc >= rbeg && jc < rend && tid < w) rowc[par * kPanelMaxW + tid] = rows[(jc - rbeg) * w + tid];
        grid_barrier(bar, bar_base + (unsigned)(c + 1) * G, flags);
        // pick the winner (every workgroup computes the same answer)
        R wv = R(-1);
        int wr = -1, wg = -1;
        for (int q = tid; q < G; q += 256) {
            R v = slot_val[par * G + q];
            int r = slot_idx[par * G + q];
            if (better(v, r, wv, wr)) {
                wv = v;
                wr = r;
                wg = q;
            }
        }
        // reduce (value,row) then recover the group from the row
        block_argmax(wv, wr, sv, sr);
        const bool singular = !(wv > R(0));
        const int p = singular ? jc : wr;
        wg = (p - j0) / RB;
        if (tid < w) {
            orow[tid] = rowc[par * kPanelMaxW + tid];
            urow[tid] = singular ? orow[tid] : slot_rows[((size_t)par * G + wg) * kPanelMaxW + tid];
        }
        __syncthreads();
        if (jc >= rbeg && jc < rend && tid < w) rows[(jc - rbeg) * w + tid] = urow[tid];
        if (p != jc && p >= rbeg && p < rend && tid < w) rows[(p - rbeg) * w + tid] = orow[tid];
        if (g == 0 && tid == 0) {
            ipiv[jc] = p;
            if (singular && flags[1] == 0) flags[1] = jc + 1;
        }
        __syncthreads();
        // scale the column and update the remaining panel columns
        const cx<R> piv = urow[c];
        R den = piv.re * piv.re + piv.im * piv.im;
        cx<R> inv = {R(0), R(0)};
        if (den > R(0)) inv = {piv.re / den, -piv.im / den};
        const int first = max(jc + 1, rbeg) - rbeg;  // first local row to update
        for (int i = first + tid; i < nr; i += 256) {
            cx<R> l = rows[i * w + c];
            if (den > R(0)) l = cmul(l, inv);
            rows[i * w + c] = l;
        }
        __syncthreads();
        const int ncol = w - c - 1;
        if (ncol > 0 && first < nr) {
            const int nupd = (nr - first) * ncol;
            for (int e = tid; e < nupd; e += 256) {
                int i = first + e / ncol, cc = c + 1 + e % ncol;
                cx<R> l = rows[i * w + c];
                cx<R> u = urow[cc];
                cx<R> &a = rows[i * w + cc];
                a.re -= l.re * u.re - l.im * u.im;
                a.im -= l.re * u.im + l.im * u.re;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < nr * w; i += 256) {
        int rr = i / w, cc = i % w;
        A[(size_t)(rbeg + rr) * lda + j0 + cc] = rows[i];
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

// Triangular solve on a block of wb rows starting at row r0 of X (columns
// [c0,c1)), with the wb x wb diagonal block of the factor at (r0, r0):
//   upper == 0:  X <- L^{-1} X, L unit lower
//   upper == 1:  X <- U^{-1} X, U upper (non-unit)
// Each workgroup solves 16 columns in LDS.  If Bt is given, the solved block
// is also written in real-expanded transposed form (see gemm.h) for the
// following update: Bt is (2*(c1-c0)) x (2*wb), row stride 2*wb.
constexpr int kTrsmCols = 16;

template <typename R>
__global__ __launch_bounds__(256) void k_trsm(const cx<R> *__restrict__ F, long long ldf, cx<R> *__restrict__ X,
                                              long long ldx, int r0, int wb, int c0, int c1, int upper,
                                              R *__restrict__ Bt) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<R> *xs = (cx<R> *)smem_raw;  // [wb][16]
    const int tid = threadIdx.x;
    const int cb = c0 + blockIdx.x * kTrsmCols;
    const int nc = min(kTrsmCols, c1 - cb);
    for (int e = tid; e < wb * kTrsmCols; e += 256) {
        int r = e / kTrsmCols, cc = e % kTrsmCols;
        xs[e] = cc < nc ? X[(size_t)(r0 + r) * ldx + cb + cc] : cx<R>{R(0), R(0)};
    }
    __syncthreads();
    if (!upper) {
        for (int i = 0; i < wb - 1; ++i) {
            const int nupd = (wb - 1 - i) * kTrsmCols;
            for (int e = tid; e < nupd; e += 256) {
                int r = i + 1 + e / kTrsmCols, cc = e % kTrsmCols;
                cx<R> l = F[(size_t)(r0 + r) * ldf + r0 + i];
                cx<R> xi = xs[i * kTrsmCols + cc];
                cx<R> &a = xs[r * kTrsmCols + cc];
                a.re -= l.re * xi.re - l.im * xi.im;
                a.im -= l.re * xi.im + l.im * xi.re;
            }
            __syncthreads();
        }
    } else {
        for (int i = wb - 1; i >= 0; --i) {
            if (tid < kTrsmCols) {
                cx<R> d = F[(size_t)(r0 + i) * ldf + r0 + i];
                R den = d.re * d.re + d.im * d.im;
                cx<R> inv = den > R(0) ? cx<R>{d.re / den, -d.im / den} : cx<R>{R(0), R(0)};
                xs[i * kTrsmCols + tid] = cmul(xs[i * kTrsmCols + tid], inv);
            }
            __syncthreads();
            const int nupd = i * kTrsmCols;
            for (int e = tid; e < nupd; e += 256) {
                int r = e / kTrsmCols, cc = e % kTrsmCols;
                cx<R> u = F[(size_t)(r0 + r) * ldf + r0 + i];
                cx<R> xi = xs[i * kTrsmCols + cc];
                cx<R> &a = xs[r * kTrsmCols + cc];
                a.re -= u.re * xi.re - u.im * xi.im;
                a.im -= u.re * xi.im + u.im * xi.re;
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < wb * kTrsmCols; e += 256) {
        int r = e / kTrsmCols, cc = e % kTrsmCols;
        if (cc < nc) X[(size_t)(r0 + r) * ldx + cb + cc] = xs[e];
    }
    if (Bt) {
        const long long ldb = 2LL * wb;
        for (int e = tid; e < wb * kTrsmCols; e += 256) {
            int r = e % wb, cc = e / wb;  // r fastest -> contiguous Bt writes
            if (cc >= nc) continue;
            cx<R> v = xs[r * kTrsmCols + cc];
            long long j = (long long)(cb - c0 + cc);
            R *b0 = Bt + (2 * j) * ldb + 2 * r;
            R *b1 = Bt + (2 * j + 1) * ldb + 2 * r;
            b0[0] = v.re;
            b0[1] = -v.im;
            b1[0] = v.im;
            b1[1] = v.re;
        }
    }
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

__global__ void k_zero_u32(unsigned *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ------------------------------------------------------------------ host
template <typename R>
struct LuPlan {
    int nb = 256;  // outer block (GEMM rank)
    int w = 64;    // panel width
};

template <typename R>
static emfa_status gemm_update(emfa_ctx *ctx, int M, int ncols, int K, const cx<R> *Aop, long long lda,
                               const R *Bt, cx<R> *C, long long ldc) {
    if (M <= 0 || ncols <= 0 || K <= 0) return EMFA_OK;
    GemmArgs g;
    g.M = M;
    g.Nr = 2 * ncols;
    g.Kr = 2 * K;
    g.A = Aop;
    g.lda = 2 * lda;
    g.Bt = Bt;
    g.ldb = 2LL * K;
    g.C = C;
    g.ldc = 2 * ldc;
    g.alpha_neg = 1;
    g.beta = 1;
    ProfScope ps(ctx, "lu_gemm", 8.0 * (double)M * ncols * K);
    if constexpr (sizeof(R) == 4)
        return gemm_launch_f32(ctx->stream, g);
    else
        return gemm_launch_f64(ctx->stream, g);
}

template <typename R>
static emfa_status trsm_launch(emfa_ctx *ctx, const cx<R> *F, long long ldf, cx<R> *X, long long ldx, int r0, int wb,
                               int c0, int c1, int upper, R *Bt, const char *fam) {
    if (c1 <= c0 || wb <= 0) return EMFA_OK;
    int blocks = (c1 - c0 + kTrsmCols - 1) / kTrsmCols;
    size_t lds = (size_t)wb * kTrsmCols * sizeof(cx<R>);
    ProfScope ps(ctx, fam, 0);
    k_trsm<R><<<blocks, 256, lds, ctx->stream>>>(F, ldf, X, ldx, r0, wb, c0, c1, upper, Bt);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

template <typename R>
static emfa_status laswp_launch(emfa_ctx *ctx, cx<R> *A, long long lda, const int *ipiv, int k0, int k1, int c0,
                                int c1, int c2, int c3) {
    int ncols = std::max(0, c1 - c0) + std::max(0, c3 - c2);
    if (ncols <= 0 || k1 <= k0) return EMFA_OK;
    ProfScope ps(ctx, "lu_laswp", 0);
    k_laswp<R><<<(ncols + 255) / 256, 256, 0, ctx->stream>>>(A, lda, ipiv, k0, k1, c0, std::max(c0, c1), c2,
                                                              std::max(c2, c3));
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

template <typename R>
static emfa_status lu_impl(emfa_ctx *ctx, emfa_mat *Am, int *ipiv, int *info_out) {
    const int N = Am->rows;
    cx<R> *A = (cx<R> *)Am->d;
    const long long lda = N;
    LuPlan<R> plan;
    if (const char *s = getenv("EMFA_LU_NB")) plan.nb = std::max(16, atoi(s));
    const int nb = plan.nb;

    // panel geometry: co-resident workgroups, rows per workgroup in LDS
    const size_t lds_cap = 150 * 1024;
    const int G_max = std::max(1, ctx->num_cu);
    // scratch: slots + Bt workspace
    size_t slot_rows_b = 2ull * G_max * kPanelMaxW * sizeof(cx<R>);
    size_t slot_val_b = 2ull * G_max * sizeof(R);
    size_t slot_idx_b = 2ull * G_max * sizeof(int);
    size_t rowc_b = 2ull * kPanelMaxW * sizeof(cx<R>);
    size_t bar_b = 256;
    size_t bt_b = (size_t)2 * N * 2 * nb * sizeof(R);
    size_t total = slot_rows_b + slot_val_b + slot_idx_b + rowc_b + bar_b + bt_b + 1024;
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, total, &scr);
    if (st != EMFA_OK) return st;
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
    R *Bt = (R *)take(bt_b);
    k_zero_u32<<<1, 64, 0, ctx->stream>>>(bar, 64);
    EMFA_HIP_CHECK(hipGetLastError());
    k_zero_u32<<<1, 64, 0, ctx->stream>>>((unsigned *)ctx->flags, 16);
    EMFA_HIP_CHECK(hipGetLastError());
    unsigned bar_base = 0;

    for (int j0 = 0; j0 < N; j0 += nb) {
        const int jb = std::min(nb, N - j0);
        int js = j0;
        while (js < j0 + jb) {
            const int M = N - js;
            // panel width and workgroup count that fit LDS
            int w = std::min(plan.w, j0 + jb - js);
            int G = std::min(G_max, std::max(1, (M + 15) / 16));
            int RB = (M + G - 1) / G;
            while ((size_t)RB * w * sizeof(cx<R>) > lds_cap && w > 8) w /= 2;
            EMFA_REQUIRE((size_t)RB * w * sizeof(cx<R>) <= lds_cap, EMFA_E_INVALID,
                         "LU panel does not fit LDS (N=%d)", N);
            size_t lds = (size_t)RB * w * sizeof(cx<R>);
            {
                ProfScope ps(ctx, "lu_panel", 0);
                void *args[] = {(void *)&A,        (void *)&lda,      (void *)&js,   (void *)&M,
                                (void *)&w,        (void *)&RB,       (void *)&ipiv, (void *)&slot_rows,
                                (void *)&slot_val, (void *)&slot_idx, (void *)&rowc, (void *)&bar,
                                (void *)&bar_base, (void *)&ctx->flags};
                hipError_t e = hipLaunchCooperativeKernel((const void *)k_panel<R>, dim3(G), dim3(256), args, lds,
                                                          ctx->stream);
                EMFA_HIP_CHECK(e);
            }
            bar_base += (unsigned)w * G;
            // swaps on the rest of the outer panel
            st = laswp_launch<R>(ctx, A, lda, ipiv, js, js + w, j0, js, js + w, j0 + jb);
            if (st) return st;
            if (js + w < j0 + jb) {
                st = trsm_launch<R>(ctx, A, lda, A, lda, js, w, js + w, j0 + jb, 0, Bt, "lu_trsm");
                if (st) return st;
                st = gemm_update<R>(ctx, N - js - w, j0 + jb - js - w, w, A + (size_t)(js + w) * lda + js, lda, Bt,
                                    A + (size_t)(js + w) * lda + js + w, lda);
                if (st) return st;
            }
            js += w;
        }
        st = laswp_launch<R>(ctx, A, lda, ipiv, j0, j0 + jb, 0, j0, j0 + jb, N);
        if (st) return st;
        if (j0 + jb < N) {
            st = trsm_launch<R>(ctx, A, lda, A, lda, j0, jb, j0 + jb, N, 0, Bt, "lu_trsm");
            if (st) return st;
            st = gemm_update<R>(ctx, N - j0 - jb, N - j0 - jb, jb, A + (size_t)(j0 + jb) * lda + j0, lda, Bt,
                                A + (size_t)(j0 + jb) * lda + j0 + jb, lda);
            if (st) return st;
        }
    }
    int flags[2] = {0, 0};
    EMFA_HIP_CHECK(hipMemcpyAsync(flags, ctx->flags, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (flags[0] != 0) {
        set_error("LU panel grid barrier timed out (workgroups not co-resident)");
        return EMFA_E_HIP;
    }
    *info_out = flags[1];
    return EMFA_OK;
}

emfa_status lu_factor_inplace(emfa_ctx *ctx, emfa_mat *A, int *ipiv, int *info) {
    if (A->prec == EMFA_F32) return lu_impl<float>(ctx, A, ipiv, info);
    return lu_impl<double>(ctx, A, ipiv, info);
}

// Solve with the factor (same precision as B).
template <typename R>
static emfa_status solve_impl(emfa_ctx *ctx, const emfa_mat *LUm, const int *ipiv, emfa_mat *Bm) {
    const int N = LUm->rows, m = Bm->cols;
    const cx<R> *F = (const cx<R> *)LUm->d;
    cx<R> *B = (cx<R> *)Bm->d;
    const long long ldf = N, ldb = m;
    int nb = 256;
    if (const char *s = getenv("EMFA_SOLVE_NB")) nb = std::max(16, atoi(s));
    size_t bt_b = (size_t)2 * m * 2 * nb * sizeof(R);
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, bt_b + 1024, &scr);
    if (st) return st;
    R *Bt = (R *)scr;
    st = laswp_launch<R>(ctx, B, ldb, ipiv, 0, N, 0, m, 0, 0);
    if (st) return st;
    for (int r0 = 0; r0 < N; r0 += nb) {
        int wb = std::min(nb, N - r0);
        st = trsm_launch<R>(ctx, F, ldf, B, ldb, r0, wb, 0, m, 0, Bt, "solve_trsm");
        if (st) return st;
        if (r0 + wb < N) {
            GemmArgs g;
            g.M = N - r0 - wb;
            g.Nr = 2 * m;
            g.Kr = 2 * wb;
            g.A = F + (size_t)(r0 + wb) * ldf + r0;
            g.lda = 2 * ldf;
            g.Bt = Bt;
            g.ldb = 2LL * wb;
            g.C = B + (size_t)(r0 + wb) * ldb;
            g.ldc = 2 * ldb;
            g.alpha_neg = 1;
            g.beta = 1;
            ProfScope ps(ctx, "solve_gemm", 8.0 * g.M * m * wb);
            st = sizeof(R) == 4 ? gemm_launch_f32(ctx->stream, g) : gemm_launch_f64(ctx->stream, g);
            if (st) return st;
        }
    }
    int last = ((N - 1) / nb) * nb;
    for (int r0 = last; r0 >= 0; r0 -= nb) {
        int wb = std::min(nb, N - r0);
        st = trsm_launch<R>(ctx, F, ldf, B, ldb, r0, wb, 0, m, 1, Bt, "solve_trsm");
        if (st) return st;
        if (r0 > 0) {
            GemmArgs g;
            g.M = r0;
            g.Nr = 2 * m;
            g.Kr = 2 * wb;
            g.A = F + r0;
            g.lda = 2 * ldf;
            g.Bt = Bt;
            g.ldb = 2LL * wb;
            g.C = B;
            g.ldc = 2 * ldb;
            g.alpha_neg = 1;
            g.beta = 1;
            ProfScope ps(ctx, "solve_gemm", 8.0 * g.M * m * wb);
            st = sizeof(R) == 4 ? gemm_launch_f32(ctx->stream, g) : gemm_launch_f64(ctx->stream, g);
            if (st) return st;
        }
    }
    return EMFA_OK;
}

emfa_status lu_solve_inplace(emfa_ctx *ctx, const emfa_mat *LU, const int *ipiv, emfa_mat *B) {
    if (LU->prec == EMFA_F32) return solve_impl<float>(ctx, LU, ipiv, B);
    return solve_impl<double>(ctx, LU, ipiv, B);
}

// C = beta C + sign * A X in FP64 (A any precision, X and C FP64).
emfa_status gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *X, emfa_mat *C, int negate,
                          int accumulate) {
    const int N = A->rows, K = A->cols, m = X->cols;
    EMFA_REQUIRE(X->rows == K && C->rows == N && C->cols == m && X->prec == EMFA_F64 && C->prec == EMFA_F64,
                 EMFA_E_INVALID, "gemm_residual: shape/precision mismatch");
    size_t bt_b = (size_t)2 * m * 2 * K * sizeof(double);
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, bt_b + 1024, &scr);
    if (st) return st;
    double *Bt = (double *)scr;
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
    return A->prec == EMFA_F32 ? gemm_launch_f32_f64(ctx->stream, g) : gemm_launch_f64(ctx->stream, g);
}

}  // namespace emfa

using namespace emfa;

extern "C" {

emfa_status emfa_lu_factor(emfa_ctx *ctx, const emfa_mat *Z, int prec, emfa_lu **out) {
    EMFA_REQUIRE(ctx && Z && out, EMFA_E_INVALID, "emfa_lu_factor: null argument");
    EMFA_REQUIRE(Z->rows == Z->cols, EMFA_E_INVALID, "emfa_lu_factor: matrix is %dx%d, not square", Z->rows,
                 Z->cols);
    EMFA_REQUIRE(prec == EMFA_F32 || prec == EMFA_F64, EMFA_E_INVALID, "emfa_lu_factor: bad precision");
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
    if (!st) st = lu_factor_inplace(ctx, lu->a, lu->ipiv, &lu->info);
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

emfa_status emfa_lu_destroy(emfa_lu *lu) {
    if (!lu) return EMFA_OK;
    emfa_ctx *c = lu->ctx;
    (void)hipStreamSynchronize(c->stream);
    ctx_free(c, lu->ipiv, sizeof(int) * lu->a->rows);
    emfa_mat_destroy(lu->a);
    delete lu;
    return EMFA_OK;
}

emfa_status emfa_lu_solve(emfa_ctx *ctx, const emfa_lu *lu, emfa_mat *B) {
    EMFA_REQUIRE(ctx && lu && B, EMFA_E_INVALID, "emfa_lu_solve: null argument");
    EMFA_REQUIRE(B->rows == lu->a->rows, EMFA_E_INVALID, "emfa_lu_solve: B has %d rows, factor %d", B->rows,
                 lu->a->rows);
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    if (B->prec == lu->a->prec) return lu_solve_inplace(ctx, lu->a, lu->ipiv, B);
    emfa_mat *T = nullptr;
    emfa_status st = emfa_mat_create(ctx, B->rows, B->cols, lu->a->prec, &T);
    if (st) return st;
    st = convert_copy(ctx, B, T);
    if (!st) st = lu_solve_inplace(ctx, lu->a, lu->ipiv, T);
    if (!st) st = convert_copy(ctx, T, B);
    (void)hipStreamSynchronize(ctx->stream);
    emfa_mat_destroy(T);
    return st;
}

emfa_status emfa_lu_pivots(const emfa_lu *lu, int *ipiv_host) {
    EMFA_REQUIRE(lu && ipiv_host, EMFA_E_INVALID, "emfa_lu_pivots: null argument");
    EMFA_HIP_CHECK(hipMemcpy(ipiv_host, lu->ipiv, sizeof(int) * lu->a->rows, hipMemcpyDeviceToHost));
    return EMFA_OK;
}

emfa_mat *emfa_lu_matrix(emfa_lu *lu) { return lu ? lu->a : nullptr; }
int emfa_lu_info(const emfa_lu *lu) { return lu ? lu->info : -1; }

emfa_status emfa_gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *Xin, emfa_mat *C, int negate,
                               int accumulate) {
    EMFA_REQUIRE(ctx && A && Xin && C, EMFA_E_INVALID, "emfa_gemm_residual: null argument");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return gemm_residual(ctx, A, Xin, C, negate, accumulate);
}

}  // extern "C"
