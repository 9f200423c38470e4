// BiCGSTAB on the device over any operator y = A x on vectors of n complex
// FP64 values: the dense matrix (K7 matvec) or the MLFMA operator.
//
// Two forms, as the reference has them:
//   variant 0  EMIterative/iml.cpp:3-39 (the overload EMSolver/CFIE.cpp:76
//              calls with eps 0.01 and 500 steps; MLFMA.cpp:402-452 runs the
//              same steps): x0 = 0, stop when ||s|| / ||b|| < eps after the
//              alpha step; "not converged" when all max_iter steps ran.
//   variant 1  EMIterative/iml.cpp:88-154 (the matrix-free overload): stop
//              when ||r||^2 / ||b||^2 <= eps^2; when |rho| falls below
//              eps_f^2 ||r0||^2 (eps_f = float epsilon) the shadow residual
//              restarts at r and the first restart resets the step count.
// The vector updates are fused elementwise kernels; the inner products of a
// step are formed together in one launch (deterministic chunked sums, FP64)
// and read back with one copy.
#include <cmath>
#include <complex>
#include <functional>
#include <vector>

#include "internal.h"

namespace emfa {

namespace {

constexpr int kBiChunks = 64;
constexpr int kMaxPairs = 5;

struct DotPairs {
    const double2 *a[kMaxPairs];  // conj(a) . b
    const double2 *b[kMaxPairs];
};

// partial sums part[pair][chunk] of conj(a_pair) . b_pair over row chunks
__global__ __launch_bounds__(256) void k_bi_dots(DotPairs d, int n, double2 *__restrict__ part) {
    const int pr = blockIdx.y, c = blockIdx.x, nc = gridDim.x;
    const double2 *a = d.a[pr], *b = d.b[pr];
    const int per = (n + nc - 1) / nc;
    const int t0 = c * per, t1 = min(n, t0 + per);
    double sr = 0, si = 0;
    for (int t = t0 + threadIdx.x; t < t1; t += 256) {
        const double2 x = a[t], y = b[t];
        sr += x.x * y.x + x.y * y.y;
        si += x.x * y.y - x.y * y.x;
    }
    __shared__ double rr[256], ri[256];
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            rr[threadIdx.x] += rr[threadIdx.x + st];
            ri[threadIdx.x] += ri[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)pr * nc + c] = make_double2(rr[0], ri[0]);
}

// the chunks of each pair added in chunk order
__global__ __launch_bounds__(64) void k_bi_dots_sum(const double2 *__restrict__ part, int np, int nc,
                                                    double2 *__restrict__ out) {
    const int i = threadIdx.x;
    if (i >= np) return;
    double sr = 0, si = 0;
    for (int c = 0; c < nc; ++c) {
        sr += part[(size_t)i * nc + c].x;
        si += part[(size_t)i * nc + c].y;
    }
    out[i] = make_double2(sr, si);
}

__device__ __forceinline__ double2 cmad(double2 acc, double2 a, double2 x) {  // acc + a x
    return make_double2(acc.x + a.x * x.x - a.y * x.y, acc.y + a.x * x.y + a.y * x.x);
}

// p = r + beta (p - w v)
__global__ __launch_bounds__(256) void k_bi_p(const double2 *__restrict__ r, const double2 *__restrict__ v,
                                              double2 beta, double2 w, int n, double2 *__restrict__ p) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double2 d = cmad(p[t], make_double2(-w.x, -w.y), v[t]);
    p[t] = cmad(r[t], beta, d);
}

// x += alpha p ; s = r - alpha v
__global__ __launch_bounds__(256) void k_bi_xs(const double2 *__restrict__ p, const double2 *__restrict__ r,
                                               const double2 *__restrict__ v, double2 alpha, int n,
                                               double2 *__restrict__ x, double2 *__restrict__ s) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    x[t] = cmad(x[t], alpha, p[t]);
    s[t] = cmad(r[t], make_double2(-alpha.x, -alpha.y), v[t]);
}

// s = r - alpha v
__global__ __launch_bounds__(256) void k_bi_s(const double2 *__restrict__ r, const double2 *__restrict__ v,
                                              double2 alpha, int n, double2 *__restrict__ s) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) s[t] = cmad(r[t], make_double2(-alpha.x, -alpha.y), v[t]);
}

// x += a p + w s ; r = s - w t   (a = 0 drops the p term)
__global__ __launch_bounds__(256) void k_bi_xr(const double2 *__restrict__ p, double2 a,
                                               const double2 *__restrict__ s, const double2 *__restrict__ tv,
                                               double2 w, int n, double2 *__restrict__ x, double2 *__restrict__ r) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double2 xv = x[t];
    if (a.x != 0.0 || a.y != 0.0) xv = cmad(xv, a, p[t]);
    x[t] = cmad(xv, w, s[t]);
    r[t] = cmad(s[t], make_double2(-w.x, -w.y), tv[t]);
}

__global__ __launch_bounds__(256) void k_bi_col(const double2 *__restrict__ M, int n, int m, int j,
                                                double2 *__restrict__ v, double2 *__restrict__ v2,
                                                double2 *__restrict__ zero) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double2 b = M[(size_t)t * m + j];
    v[t] = b;
    if (v2) v2[t] = b;
    if (zero) zero[t] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_bi_put(const double2 *__restrict__ v, int n, int m, int j,
                                                double2 *__restrict__ M) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) M[(size_t)t * m + j] = v[t];
}

__global__ __launch_bounds__(256) void k_bi_copy(const double2 *__restrict__ a, int n, double2 *__restrict__ b) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) b[t] = a[t];
}

__global__ __launch_bounds__(256) void k_bi_zero2(int n, double2 *__restrict__ a, double2 *__restrict__ b) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) a[t] = b[t] = make_double2(0.0, 0.0);
}

// b - y into y
__global__ __launch_bounds__(256) void k_bi_resid(const double2 *__restrict__ b, int n, double2 *__restrict__ y) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) y[t] = make_double2(b[t].x - y[t].x, b[t].y - y[t].y);
}

}  // namespace

// Device bytes the solver's vectors and partial sums take for n unknowns.
size_t bicgstab_work_bytes(int n) {
    const int nchunk = std::max(1, std::min(kBiChunks, (n + 2047) / 2048));
    const size_t vb = ((size_t)n * sizeof(double2) + 255) & ~(size_t)255;
    return 9 * vb + (size_t)kMaxPairs * (nchunk + 1) * sizeof(double2) + 512;
}

emfa_status bicgstab_impl(emfa_ctx *ctx, int n, const std::function<emfa_status(const double2 *, double2 *)> &matvec,
                          const emfa_mat *B, emfa_mat *X, double tol, int max_iter, int variant,
                          emfa_bicgstab_report *rep, void *work, size_t work_bytes) {
    const int nr = B->cols;
    EMFA_REQUIRE(B->rows == n && X->rows == n && X->cols == nr && B->prec == EMFA_F64 && X->prec == EMFA_F64,
                 EMFA_E_INVALID, "bicgstab: B and X must be F64 n x nrhs");
    EMFA_REQUIRE(max_iter >= 1 && tol > 0 && (variant == 0 || variant == 1), EMFA_E_INVALID,
                 "bicgstab: max_iter >= 1, tol > 0, variant 0 or 1");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    const int nchunk = std::max(1, std::min(kBiChunks, (n + 2047) / 2048));
    const size_t vb = ((size_t)n * sizeof(double2) + 255) & ~(size_t)255;
    void *scr = work;
    emfa_status st = EMFA_OK;
    if (scr) {
        EMFA_REQUIRE(work_bytes >= bicgstab_work_bytes(n), EMFA_E_INVALID, "bicgstab: workspace of %zu bytes, %zu needed",
                     work_bytes, bicgstab_work_bytes(n));
    } else if ((st = ctx_scratch(ctx, bicgstab_work_bytes(n), &scr))) {
        return st;
    }
    char *q = (char *)scr;
    auto take = [&](size_t b) {
        char *o = q;
        q += (b + 255) & ~(size_t)255;
        return (double2 *)o;
    };
    double2 *bv = take(vb), *x = take(vb), *r = take(vb), *rs = take(vb), *p = take(vb), *v = take(vb),
            *s = take(vb), *tv = take(vb), *w = take(vb);
    double2 *part = take((size_t)kMaxPairs * nchunk * sizeof(double2));
    double2 *dsum = take(kMaxPairs * sizeof(double2));
    hipStream_t stm = ctx->stream;
    const int nb = (n + 255) / 256;
    using cd = std::complex<double>;
    double2 hd[kMaxPairs];
    // conj(a_i) . b_i for the given pairs, read back to the host
    auto dots = [&](int np, std::initializer_list<std::pair<const double2 *, const double2 *>> pr,
                    cd *out) -> emfa_status {
        DotPairs d{};
        int i = 0;
        for (auto &ab : pr) {
            d.a[i] = ab.first;
            d.b[i] = ab.second;
            ++i;
        }
        k_bi_dots<<<dim3(nchunk, np), 256, 0, stm>>>(d, n, part);
        k_bi_dots_sum<<<1, 64, 0, stm>>>(part, np, nchunk, dsum);
        EMFA_HIP_CHECK(hipGetLastError());
        EMFA_HIP_CHECK(hipMemcpyAsync(hd, dsum, np * sizeof(double2), hipMemcpyDeviceToHost, stm));
        EMFA_HIP_CHECK(hipStreamSynchronize(stm));
        for (int k = 0; k < np; ++k) out[k] = cd(hd[k].x, hd[k].y);
        return EMFA_OK;
    };
    auto d2 = [](cd z) { return make_double2(z.real(), z.imag()); };
    int total = 0, worst_steps = 0;
    double worst = 0;
    bool all_conv = true;
    cd dv[kMaxPairs];
    for (int col = 0; col < nr; ++col) {
        // b, r = b (x0 = 0), shadow residual = b, x = 0
        k_bi_col<<<nb, 256, 0, stm>>>((const double2 *)B->d, n, nr, col, bv, r, x);
        k_bi_copy<<<nb, 256, 0, stm>>>(bv, n, rs);
        EMFA_HIP_CHECK(hipGetLastError());
        if ((st = dots(1, {{bv, bv}}, dv))) return st;
        const double bn2 = dv[0].real(), bn = std::sqrt(std::max(0.0, bn2));
        int steps = 0;
        bool conv = false;
        if (bn2 == 0) {
            conv = true;  // b = 0: x = 0
        } else if (variant == 0) {
            cd rou0 = 1, alph = 1, omg = 1;
            k_bi_zero2<<<nb, 256, 0, stm>>>(n, v, p);
            if ((st = dots(1, {{rs, r}}, dv))) return st;
            cd rou1 = dv[0];
            for (int i = 0; i < max_iter; ++i) {
                ++steps;
                const cd beta = (rou1 / rou0) * (alph / omg);
                k_bi_p<<<nb, 256, 0, stm>>>(r, v, d2(beta), d2(omg), n, p);
                EMFA_HIP_CHECK(hipGetLastError());
                if ((st = matvec(p, v))) return st;
                if ((st = dots(1, {{rs, v}}, dv))) return st;
                alph = rou1 / dv[0];
                k_bi_xs<<<nb, 256, 0, stm>>>(p, r, v, d2(alph), n, x, s);
                EMFA_HIP_CHECK(hipGetLastError());
                if ((st = dots(1, {{s, s}}, dv))) return st;
                if (std::sqrt(std::max(0.0, dv[0].real())) / bn < tol) {
                    conv = true;
                    break;
                }
                if ((st = matvec(s, tv))) return st;
                // t^H s, t^H t; and for the next step's rho: rs^H s, rs^H t
                if ((st = dots(4, {{tv, s}, {tv, tv}, {rs, s}, {rs, tv}}, dv))) return st;
                omg = dv[0] / dv[1];
                k_bi_xr<<<nb, 256, 0, stm>>>(nullptr, make_double2(0, 0), s, tv, d2(omg), n, x, r);
                EMFA_HIP_CHECK(hipGetLastError());
                // the second test of iml.cpp:33-34 reads the same ||s||
                rou0 = rou1;
                rou1 = dv[2] - omg * dv[3];  // rs^H (s - omega t)
            }
            conv = steps != max_iter;  // iml.cpp:38: success unless every step ran
        } else {
            cd rho = 1, alpha = 1, wv = 1;
            k_bi_zero2<<<nb, 256, 0, stm>>>(n, v, p);
            double r0_sq = bn2;
            int restarts = 0;
            const double epsf = 1.1920928955078125e-07, eps2 = epsf * epsf, tol2 = tol * tol;
            for (;;) {
                if ((st = dots(2, {{r, r}, {rs, r}}, dv))) return st;
                const double rr = dv[0].real();
                if (!(rr / bn2 > tol2 && steps < max_iter)) break;
                const cd rho_old = rho;
                rho = dv[1];
                if (std::abs(rho) < eps2 * r0_sq) {
                    k_bi_copy<<<nb, 256, 0, stm>>>(r, n, rs);
                    EMFA_HIP_CHECK(hipGetLastError());
                    rho = r0_sq = rr;
                    if (restarts++ == 0) steps = 0;
                }
                const cd beta = (rho / rho_old) * (alpha / wv);
                k_bi_p<<<nb, 256, 0, stm>>>(r, v, d2(beta), d2(wv), n, p);
                EMFA_HIP_CHECK(hipGetLastError());
                if ((st = matvec(p, v))) return st;
                if ((st = dots(1, {{rs, v}}, dv))) return st;
                alpha = rho / dv[0];
                k_bi_s<<<nb, 256, 0, stm>>>(r, v, d2(alpha), n, s);
                EMFA_HIP_CHECK(hipGetLastError());
                if ((st = matvec(s, tv))) return st;
                if ((st = dots(2, {{tv, tv}, {tv, s}}, dv))) return st;
                wv = dv[0].real() > 0 ? dv[1] / dv[0].real() : cd(0);
                k_bi_xr<<<nb, 256, 0, stm>>>(p, d2(alpha), s, tv, d2(wv), n, x, r);
                EMFA_HIP_CHECK(hipGetLastError());
                ++steps;
            }
            conv = true;  // the overload always reports success (iml.cpp:147-153)
        }
        // true residual of the returned x
        if ((st = matvec(x, w))) return st;
        k_bi_resid<<<nb, 256, 0, stm>>>(bv, n, w);
        EMFA_HIP_CHECK(hipGetLastError());
        if ((st = dots(1, {{w, w}}, dv))) return st;
        const double rel = bn2 > 0 ? std::sqrt(std::max(0.0, dv[0].real()) / bn2) : 0.0;
        k_bi_put<<<nb, 256, 0, stm>>>(x, n, nr, col, (double2 *)X->d);
        EMFA_HIP_CHECK(hipGetLastError());
        total += steps;
        worst_steps = std::max(worst_steps, steps);
        worst = std::max(worst, rel);
        all_conv = all_conv && conv;
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(stm));
    if (rep) {
        rep->iterations = total;
        rep->max_steps = worst_steps;
        rep->relres = worst;
    }
    if (!all_conv) {
        set_error("bicgstab: a right-hand side used all %d steps (relative residual %.3e)", max_iter, worst);
        return EMFA_E_NOCONV;
    }
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

extern "C" emfa_status emfa_bicgstab(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X, double tol,
                                     int max_iter, int variant, emfa_bicgstab_report *rep) {
    EMFA_REQUIRE(ctx && Z && B && X, EMFA_E_INVALID, "emfa_bicgstab: null argument");
    EMFA_REQUIRE(Z->cols == Z->rows, EMFA_E_INVALID, "emfa_bicgstab: Z must be square");
    return bicgstab_impl(ctx, Z->rows, [&](const double2 *in, double2 *out) { return gemv(ctx, Z, in, out); }, B, X,
                         tol, max_iter, variant, rep);
}
