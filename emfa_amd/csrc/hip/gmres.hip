// Restarted GMRES(m) on the device: dense matvec (K7, HBM-bound, FP64
// accumulation), classical Gram-Schmidt applied twice (K8, FP64), Givens
// rotations of the small Hessenberg system on the host.
//
// Same algorithm as EMIterative/iml.cpp:156-207 (called at
// EMSolver/EFIE.cpp:79) with its defects corrected: the projection is
// v_i^H w (iml.cpp:189 computes w^H v_i), the stop test is relative and the
// number of restarts is bounded.
#include <cmath>
#include <complex>
#include <functional>
#include <vector>

#include "internal.h"

namespace emfa {

// y = Z x, one wavefront per row, FP64 accumulation
template <typename TA>
__global__ __launch_bounds__(256) void k_gemv(const cx<TA> *__restrict__ Z, int n, const double2 *__restrict__ x,
                                              double2 *__restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const cx<TA> *zr = Z + (size_t)row * n;
    double sr = 0, si = 0;
    for (int j = lane; j < n; j += 64) {
        cx<TA> a = zr[j];
        double2 b = x[j];
        sr += (double)a.re * b.x - (double)a.im * b.y;
        si += (double)a.re * b.y + (double)a.im * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_down(sr, off, 64);
        si += __shfl_down(si, off, 64);
    }
    if (lane == 0) y[row] = make_double2(sr, si);
}

// Dot products h[i] = V_i^H w for i < k in two deterministic passes:
// k_dots writes partial sums part[i][c] over chunk c of the rows (grid
// chunks x k), k_dots_sum adds the chunks of each i in order.
constexpr int kDotChunks = 64;

__global__ __launch_bounds__(256) void k_dots(const double2 *__restrict__ V, int n, const double2 *__restrict__ w,
                                              double2 *__restrict__ part) {
    const int i = blockIdx.y, c = blockIdx.x, nc = gridDim.x;
    const double2 *v = V + (size_t)i * n;
    const int per = (n + nc - 1) / nc;
    const int t0 = c * per, t1 = min(n, t0 + per);
    double sr = 0, si = 0;
    for (int t = t0 + threadIdx.x; t < t1; t += 256) {
        double2 a = v[t], b = w[t];
        sr += a.x * b.x + a.y * b.y;  // conj(a) * b
        si += a.x * b.y - a.y * b.x;
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
    if (threadIdx.x == 0) part[(size_t)i * nc + c] = make_double2(rr[0], ri[0]);
}

__global__ __launch_bounds__(64) void k_dots_sum(const double2 *__restrict__ part, int k, int nc,
                                                 double2 *__restrict__ h) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= k) return;
    double sr = 0, si = 0;
    for (int c = 0; c < nc; ++c) {
        sr += part[(size_t)i * nc + c].x;
        si += part[(size_t)i * nc + c].y;
    }
    h[i] = make_double2(sr, si);
}

// w -= sum_{i<k} h_i V_i   (sign = -1)  or  x += sum y_i V_i (sign = +1)
__global__ __launch_bounds__(256) void k_combine(const double2 *__restrict__ V, int n, int k,
                                                 const double2 *__restrict__ h, double sign, double2 *__restrict__ w) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double2 acc = w[t];
    for (int i = 0; i < k; ++i) {
        double2 a = V[(size_t)i * n + t], c = h[i];
        acc.x += sign * (c.x * a.x - c.y * a.y);
        acc.y += sign * (c.x * a.y + c.y * a.x);
    }
    w[t] = acc;
}

__global__ __launch_bounds__(256) void k_scale_copy(const double2 *__restrict__ src, int n, double s,
                                                    double2 *__restrict__ dst) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) dst[t] = make_double2(src[t].x * s, src[t].y * s);
}

// column j of a row-major n x m complex matrix <-> contiguous vector
__global__ __launch_bounds__(256) void k_col_get(const double2 *__restrict__ M, int n, int m, int j,
                                                 double2 *__restrict__ v) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) v[t] = M[(size_t)t * m + j];
}
__global__ __launch_bounds__(256) void k_col_put(const double2 *__restrict__ v, int n, int m, int j,
                                                 double2 *__restrict__ M) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) M[(size_t)t * m + j] = v[t];
}
// r = b - r (in place on r)
__global__ __launch_bounds__(256) void k_sub_from(const double2 *__restrict__ b, int n, double2 *__restrict__ r) {
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) r[t] = make_double2(b[t].x - r[t].x, b[t].y - r[t].y);
}

// GMRES core over any operator y = A x on device vectors of n complex FP64.
emfa_status gmres_impl(emfa_ctx *ctx, int n, const std::function<emfa_status(const double2 *, double2 *)> &matvec,
                       const emfa_mat *B, emfa_mat *X, int m, double tol, int max_restart, emfa_gmres_report *rep) {
    const int nr = B->cols;
    EMFA_REQUIRE(B->rows == n && X->rows == n && X->cols == nr && B->prec == EMFA_F64 && X->prec == EMFA_F64,
                 EMFA_E_INVALID, "gmres: B and X must be F64 n x nrhs");
    EMFA_REQUIRE(m >= 1 && m <= 512 && max_restart >= 1, EMFA_E_INVALID, "gmres: bad m/max_restart");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    size_t vb = (size_t)(m + 1) * n * sizeof(double2);
    const int nchunk = std::max(1, std::min(kDotChunks, (n + 2047) / 2048));
    size_t total = vb + 4 * (size_t)n * sizeof(double2) + 2 * (size_t)(m + 2) * sizeof(double2) +
                   (size_t)(m + 2) * nchunk * sizeof(double2) + 8192;
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, total, &scr);
    if (st) return st;
    char *p = (char *)scr;
    auto take = [&](size_t b) {
        char *q = p;
        p += (b + 255) & ~(size_t)255;
        return (double2 *)q;
    };
    double2 *V = take(vb);
    double2 *w = take(n * sizeof(double2));
    double2 *bv = take(n * sizeof(double2));
    double2 *xv = take(n * sizeof(double2));
    double2 *h = take((m + 2) * sizeof(double2));
    double2 *h2 = take((m + 2) * sizeof(double2));
    double2 *part = take((size_t)(m + 2) * nchunk * sizeof(double2));
    auto dots = [&](const double2 *Vb, int k, const double2 *wv, double2 *out) {
        k_dots<<<dim3(nchunk, k), 256, 0, ctx->stream>>>(Vb, n, wv, part);
        k_dots_sum<<<(k + 63) / 64, 64, 0, ctx->stream>>>(part, k, nchunk, out);
    };
    const int nb = (n + 255) / 256;
    hipStream_t s = ctx->stream;
    using cd = std::complex<double>;
    auto norm = [&](const double2 *v, double *out) -> emfa_status {
        dots(v, 1, v, h2);
        EMFA_HIP_CHECK(hipGetLastError());
        double2 r;
        EMFA_HIP_CHECK(hipMemcpyAsync(&r, h2, sizeof(double2), hipMemcpyDeviceToHost, s));
        EMFA_HIP_CHECK(hipStreamSynchronize(s));
        *out = std::sqrt(std::max(0.0, r.x));
        return EMFA_OK;
    };
    int total_it = 0, total_rs = 0;
    double worst = 0;
    bool all_conv = true;
    std::vector<cd> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    std::vector<double2> hh(m + 2), hh2(m + 2);
    for (int col = 0; col < nr; ++col) {
        k_col_get<<<nb, 256, 0, s>>>((const double2 *)B->d, n, nr, col, bv);
        k_col_get<<<nb, 256, 0, s>>>((const double2 *)X->d, n, nr, col, xv);
        EMFA_HIP_CHECK(hipGetLastError());
        double bn = 0;
        if ((st = norm(bv, &bn))) return st;
        if (bn == 0) bn = 1;
        bool conv = false;
        double rr = 1;
        for (int cyc = 0; cyc < max_restart && !conv; ++cyc) {
            ++total_rs;
            if ((st = matvec(xv, w))) return st;
            k_sub_from<<<nb, 256, 0, s>>>(bv, n, w);
            double beta = 0;
            if ((st = norm(w, &beta))) return st;
            rr = beta / bn;
            if (rr < tol) {
                conv = true;
                break;
            }
            k_scale_copy<<<nb, 256, 0, s>>>(w, n, 1.0 / beta, V);
            std::fill(g.begin(), g.end(), cd(0));
            g[0] = beta;
            int j = 0;
            for (; j < m; ++j) {
                if ((st = matvec(V + (size_t)j * n, w))) return st;
                ++total_it;
                // classical Gram-Schmidt, applied twice
                dots(V, j + 1, w, h);
                k_combine<<<nb, 256, 0, s>>>(V, n, j + 1, h, -1.0, w);
                dots(V, j + 1, w, h2);
                k_combine<<<nb, 256, 0, s>>>(V, n, j + 1, h2, -1.0, w);
                EMFA_HIP_CHECK(hipGetLastError());
                EMFA_HIP_CHECK(hipMemcpyAsync(hh.data(), h, sizeof(double2) * (j + 1), hipMemcpyDeviceToHost, s));
                EMFA_HIP_CHECK(hipMemcpyAsync(hh2.data(), h2, sizeof(double2) * (j + 1), hipMemcpyDeviceToHost, s));
                double hn = 0;
                if ((st = norm(w, &hn))) return st;
                for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = cd(hh[i].x + hh2[i].x, hh[i].y + hh2[i].y);
                if (hn > 0) k_scale_copy<<<nb, 256, 0, s>>>(w, n, 1.0 / hn, V + (size_t)(j + 1) * n);
                for (int i = 0; i < j; ++i) {
                    cd a = H[(size_t)i * m + j], b2 = H[(size_t)(i + 1) * m + j];
                    H[(size_t)i * m + j] = std::conj(cs[i]) * a + std::conj(sn[i]) * b2;
                    H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b2;
                }
                cd a = H[(size_t)j * m + j];
                double den = std::sqrt(std::norm(a) + hn * hn);
                if (den == 0) {
                    cs[j] = 1;
                    sn[j] = 0;
                } else {
                    cs[j] = a / den;
                    sn[j] = hn / den;
                }
                H[(size_t)j * m + j] = std::conj(cs[j]) * a + std::conj(sn[j]) * hn;
                g[j + 1] = -sn[j] * g[j];
                g[j] = std::conj(cs[j]) * g[j];
                rr = std::abs(g[j + 1]) / bn;
                if (rr < tol || hn == 0) {
                    ++j;
                    break;
                }
            }
            int kk = std::min(j, m);
            for (int i = kk - 1; i >= 0; --i) {
                cd sacc = g[i];
                for (int t = i + 1; t < kk; ++t) sacc -= H[(size_t)i * m + t] * y[t];
                y[i] = sacc / H[(size_t)i * m + i];
            }
            for (int i = 0; i < kk; ++i) hh[i] = make_double2(y[i].real(), y[i].imag());
            EMFA_HIP_CHECK(hipMemcpyAsync(h, hh.data(), sizeof(double2) * kk, hipMemcpyHostToDevice, s));
            k_combine<<<nb, 256, 0, s>>>(V, n, kk, h, 1.0, xv);
            EMFA_HIP_CHECK(hipGetLastError());
            EMFA_HIP_CHECK(hipStreamSynchronize(s));
        }
        // true residual
        if ((st = matvec(xv, w))) return st;
        k_sub_from<<<nb, 256, 0, s>>>(bv, n, w);
        double rn = 0;
        if ((st = norm(w, &rn))) return st;
        rr = rn / bn;
        conv = rr < tol;
        all_conv = all_conv && conv;
        worst = std::max(worst, rr);
        k_col_put<<<nb, 256, 0, s>>>(xv, n, nr, col, (double2 *)X->d);
        EMFA_HIP_CHECK(hipGetLastError());
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(s));
    if (rep) {
        rep->iterations = total_it;
        rep->restarts = total_rs;
        rep->relres = worst;
    }
    if (!all_conv) {
        set_error("gmres: relative residual %.3e above tolerance %.3e", worst, tol);
        return EMFA_E_NOCONV;
    }
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

emfa_status emfa::gemv(emfa_ctx *ctx, const emfa_mat *Z, const double2 *in, double2 *out) {
    const int n = Z->rows;
    hipStream_t s = ctx->stream;
    ProfScope ps(ctx, "gmres_matvec", (double)n * n * elem_bytes(Z->prec));
    if (Z->prec == EMFA_F32)
        k_gemv<float><<<(n + 3) / 4, 256, 0, s>>>((const cx<float> *)Z->d, n, in, out);
    else
        k_gemv<double><<<(n + 3) / 4, 256, 0, s>>>((const cx<double> *)Z->d, n, in, out);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

extern "C" emfa_status emfa_gmres(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X, int m, double tol,
                                  int max_restart, emfa_gmres_report *rep) {
    EMFA_REQUIRE(ctx && Z && B && X, EMFA_E_INVALID, "emfa_gmres: null argument");
    const int n = Z->rows;
    EMFA_REQUIRE(Z->cols == n, EMFA_E_INVALID, "emfa_gmres: Z must be square");
    auto matvec = [&](const double2 *in, double2 *out) -> emfa_status { return gemv(ctx, Z, in, out); };
    return gmres_impl(ctx, n, matvec, B, X, m, tol, max_restart, rep);
}
