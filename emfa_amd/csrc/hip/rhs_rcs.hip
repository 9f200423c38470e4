// Excitation columns (K3) and far-field RCS (K9), plus small matrix utilities.
#include <cmath>

#include "internal.h"

namespace emfa {

__constant__ double c_r3[3][4] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0},
                                  {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 3.0},
                                  {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0}};
__constant__ double c_r7[7][4] = {
    {0.333333333333333, 0.333333333333333, 1.0 - 0.333333333333333 - 0.333333333333333, 0.112500000000000 * 2},
    {0.797426985353087, 0.101286507323456, 1.0 - 0.797426985353087 - 0.101286507323456, 0.062969590272413 * 2},
    {0.101286507323456, 0.797426985353087, 1.0 - 0.101286507323456 - 0.797426985353087, 0.062969590272413 * 2},
    {0.101286507323456, 0.101286507323456, 1.0 - 0.101286507323456 - 0.101286507323456, 0.062969590272413 * 2},
    {0.470142064105115, 0.470142064105115, 1.0 - 0.470142064105115 - 0.470142064105115, 0.066197076394253 * 2},
    {0.470142064105115, 0.059715871789770, 1.0 - 0.470142064105115 - 0.059715871789770, 0.066197076394253 * 2},
    {0.059715871789770, 0.470142064105115, 1.0 - 0.059715871789770 - 0.470142064105115, 0.066197076394253 * 2}};

struct tri_abs {
    double v[3][3];
    double n[3];
};

__device__ __forceinline__ void load_tri(const double *__restrict__ tg, int T, int t, tri_abs *o) {
    for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 3; ++c) o->v[i][c] = tg[(size_t)c * T + t] + tg[(size_t)(3 + 3 * i + c) * T + t];
    for (int c = 0; c < 3; ++c) o->n[c] = tg[(size_t)(12 + c) * T + t];
}

// V(u, j) = 1/2 l_u sum_{+,-} sum_a w_a rho . [alpha e + (1-alpha) n x h] exp(-jk k.r)
// EFIE (kind 0): 7-point rule, e only (EMSolver/EFIE.cpp:568-601,698-722).
// CFIE (kind 1): 3-point rule (EMSolver/Miscellaneous.cpp:152-189).
// MFIE (kind 2): 7-point rule, n x h part only (EMSolver/MFIE.cpp:289-322,
// 435-436 without its 1/Z0, matching the Z0-scaled MFIE operator of the fill).
// One thread per (u, j), j fastest so the row-major writes coalesce.
template <typename R>
__global__ __launch_bounds__(256) void k_rhs(const double *__restrict__ tg, const int *__restrict__ rtp,
                                             const int *__restrict__ rtm, const int *__restrict__ rap,
                                             const int *__restrict__ ram, const double *__restrict__ rlen, int T,
                                             int N, int nrhs, const double *__restrict__ inc, int kind, double k,
                                             double alpha, cx<R> *__restrict__ B) {
    long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long long)N * nrhs) return;
    int u = (int)(id / nrhs), j = (int)(id % nrhs);
    const double *c = inc + 9 * j;
    double kv[3] = {c[0], c[1], c[2]}, ev[3] = {c[3], c[4], c[5]}, hv[3] = {c[6], c[7], c[8]};
    double vr = 0, vi = 0, mr = 0, mi = 0;
    for (int side = 0; side < 2; ++side) {
        int t = side == 0 ? rtp[u] : rtm[u];
        int a = side == 0 ? rap[u] : ram[u];
        tri_abs ta;
        load_tri(tg, T, t, &ta);
        double sg = side == 0 ? 1.0 : -1.0;
        double nxh[3] = {ta.n[1] * hv[2] - ta.n[2] * hv[1], ta.n[2] * hv[0] - ta.n[0] * hv[2],
                         ta.n[0] * hv[1] - ta.n[1] * hv[0]};
        int Q = kind == EMFA_CFIE ? 3 : 7;
        for (int p = 0; p < Q; ++p) {
            const double *rw = Q == 7 ? c_r7[p] : c_r3[p];
            double r[3];
            for (int cc = 0; cc < 3; ++cc) r[cc] = rw[0] * ta.v[0][cc] + rw[1] * ta.v[1][cc] + rw[2] * ta.v[2][cc];
            double rho[3];
            for (int cc = 0; cc < 3; ++cc) rho[cc] = sg * (r[cc] - ta.v[a][cc]);
            double ph = k * (r[0] * kv[0] + r[1] * kv[1] + r[2] * kv[2]);
            double s, co;
            sincos(ph, &s, &co);
            double w = rw[3];
            double de = rho[0] * ev[0] + rho[1] * ev[1] + rho[2] * ev[2];
            vr += w * de * co;
            vi -= w * de * s;
            if (kind != EMFA_EFIE) {
                double dm = rho[0] * nxh[0] + rho[1] * nxh[1] + rho[2] * nxh[2];
                mr += w * dm * co;
                mi -= w * dm * s;
            }
        }
    }
    double outr, outi;
    if (kind == EMFA_EFIE) {
        outr = vr;
        outi = vi;
    } else if (kind == EMFA_MFIE) {
        outr = mr;
        outi = mi;
    } else {
        outr = alpha * vr + (1.0 - alpha) * mr;
        outi = alpha * vi + (1.0 - alpha) * mi;
    }
    double h = 0.5 * rlen[u];
    B[(size_t)u * nrhs + j] = {(R)(h * outr), (R)(h * outi)};
}

// RCS per direction: Es = sum_u I_u L_u sum_7 w rho exp(+jk s.r);
// sigma = Z0^2 k^2 |1/2 s x (s x Es)|^2 / 4 pi  (EMSolver/EFIE.cpp:788-832).
// One workgroup per direction, FP64 accumulation, LDS tree reduction.
template <typename R>
__global__ __launch_bounds__(256) void k_rcs(const double *__restrict__ tg, const int *__restrict__ rtp,
                                             const int *__restrict__ rtm, const int *__restrict__ rap,
                                             const int *__restrict__ ram, const double *__restrict__ rlen, int T,
                                             int N, const cx<R> *__restrict__ X, int ldx,
                                             const double *__restrict__ dirs, const int *__restrict__ cols, double k,
                                             double *__restrict__ out) {
    const int d = blockIdx.x;
    const double s[3] = {dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]};
    const int col = cols ? cols[d] : 0;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int u = threadIdx.x; u < N; u += 256) {
        cx<R> xu = X[(size_t)u * ldx + col];
        double L = rlen[u];
        double xr = (double)xu.re * L, xi = (double)xu.im * L;
        double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0};
        for (int side = 0; side < 2; ++side) {
            int t = side == 0 ? rtp[u] : rtm[u];
            int a = side == 0 ? rap[u] : ram[u];
            tri_abs ta;
            load_tri(tg, T, t, &ta);
            double sg = side == 0 ? 1.0 : -1.0;
            for (int p = 0; p < 7; ++p) {
                const double *rw = c_r7[p];
                double r[3];
                for (int cc = 0; cc < 3; ++cc)
                    r[cc] = rw[0] * ta.v[0][cc] + rw[1] * ta.v[1][cc] + rw[2] * ta.v[2][cc];
                double ph = k * (r[0] * s[0] + r[1] * s[1] + r[2] * s[2]);
                double sn, co;
                sincos(ph, &sn, &co);
                double w = rw[3];
                for (int cc = 0; cc < 3; ++cc) {
                    double rho = sg * (r[cc] - ta.v[a][cc]) * w;
                    er[cc] += rho * co;
                    ei[cc] += rho * sn;
                }
            }
        }
        for (int cc = 0; cc < 3; ++cc) {
            acc[2 * cc] += er[cc] * xr - ei[cc] * xi;
            acc[2 * cc + 1] += er[cc] * xi + ei[cc] * xr;
        }
    }
    __shared__ double red[6][256];
    for (int i = 0; i < 6; ++i) red[i][threadIdx.x] = acc[i];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int i = 0; i < 6; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double Er[3] = {red[0][0], red[2][0], red[4][0]}, Ei[3] = {red[1][0], red[3][0], red[5][0]};
        // c = s x E ; f = 0.5 s x c
        double cr[3] = {s[1] * Er[2] - s[2] * Er[1], s[2] * Er[0] - s[0] * Er[2], s[0] * Er[1] - s[1] * Er[0]};
        double ci[3] = {s[1] * Ei[2] - s[2] * Ei[1], s[2] * Ei[0] - s[0] * Ei[2], s[0] * Ei[1] - s[1] * Ei[0]};
        double fr[3] = {0.5 * (s[1] * cr[2] - s[2] * cr[1]), 0.5 * (s[2] * cr[0] - s[0] * cr[2]),
                        0.5 * (s[0] * cr[1] - s[1] * cr[0])};
        double fi[3] = {0.5 * (s[1] * ci[2] - s[2] * ci[1]), 0.5 * (s[2] * ci[0] - s[0] * ci[2]),
                        0.5 * (s[0] * ci[1] - s[1] * ci[0])};
        double nrm = 0;
        for (int cc = 0; cc < 3; ++cc) nrm += fr[cc] * fr[cc] + fi[cc] * fi[cc];
        double rcs = kZ0 * kZ0 * k * k * nrm / (4.0 * kPi);
        out[d] = 10.0 * log10(rcs);
    }
}

// Precision conversion / copy of interleaved complex buffers.
template <typename S, typename D>
__global__ __launch_bounds__(256) void k_convert(const S *__restrict__ src, D *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) dst[i] = (D)src[i];
}

// X += D  (X double, D float or double), element-wise over interleaved reals
template <typename S>
__global__ __launch_bounds__(256) void k_axpy(double *__restrict__ x, const S *__restrict__ d, size_t n) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) x[i] += (double)d[i];
}

emfa_status convert_copy(emfa_ctx *ctx, const emfa_mat *src, emfa_mat *dst) {
    size_t n = 2 * (size_t)src->rows * src->cols;
    if (src->prec == dst->prec) {
        EMFA_HIP_CHECK(hipMemcpyAsync(dst->d, src->d, src->bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return EMFA_OK;
    }
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (src->prec == EMFA_F32)
        k_convert<float, double><<<blocks, 256, 0, ctx->stream>>>((const float *)src->d, (double *)dst->d, n);
    else
        k_convert<double, float><<<blocks, 256, 0, ctx->stream>>>((const double *)src->d, (float *)dst->d, n);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

emfa_status axpy_f64(emfa_ctx *ctx, emfa_mat *X, const emfa_mat *D) {
    size_t n = 2 * (size_t)X->rows * X->cols;
    int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    if (D->prec == EMFA_F32)
        k_axpy<float><<<blocks, 256, 0, ctx->stream>>>((double *)X->d, (const float *)D->d, n);
    else
        k_axpy<double><<<blocks, 256, 0, ctx->stream>>>((double *)X->d, (const double *)D->d, n);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

extern "C" emfa_status emfa_rhs_fill(emfa_ctx *ctx, const emfa_mesh *M, int kind, double k, double alpha, int nrhs,
                                     const double *inc, emfa_mat *B) {
    EMFA_REQUIRE(ctx && M && inc && B, EMFA_E_INVALID, "emfa_rhs_fill: null argument");
    EMFA_REQUIRE(nrhs > 0 && B->rows == M->n && B->cols == nrhs, EMFA_E_INVALID,
                 "emfa_rhs_fill: B is %dx%d, expected %dx%d", B->rows, B->cols, M->n, nrhs);
    EMFA_REQUIRE(kind == EMFA_EFIE || kind == EMFA_CFIE || kind == EMFA_MFIE, EMFA_E_INVALID,
                 "emfa_rhs_fill: bad kind");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    void *dinc = nullptr;
    EMFA_HIP_CHECK(ctx_alloc(ctx, &dinc, 9 * sizeof(double) * nrhs));
    EMFA_HIP_CHECK(hipMemcpyAsync(dinc, inc, 9 * sizeof(double) * nrhs, hipMemcpyHostToDevice, ctx->stream));
    long long total = (long long)M->n * nrhs;
    int blocks = (int)((total + 255) / 256);
    {
        ProfScope ps(ctx, "rhs", 0);
        if (B->prec == EMFA_F32)
            k_rhs<float><<<blocks, 256, 0, ctx->stream>>>(M->tg64, M->rtp, M->rtm, M->rap, M->ram, M->rlen64, M->nt,
                                                          M->n, nrhs, (const double *)dinc, kind, k, alpha,
                                                          (cx<float> *)B->d);
        else
            k_rhs<double><<<blocks, 256, 0, ctx->stream>>>(M->tg64, M->rtp, M->rtm, M->rap, M->ram, M->rlen64,
                                                           M->nt, M->n, nrhs, (const double *)dinc, kind, k, alpha,
                                                           (cx<double> *)B->d);
    }
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free(ctx, dinc, 9 * sizeof(double) * nrhs);
    EMFA_HIP_CHECK(e);
    return EMFA_OK;
}

extern "C" emfa_status emfa_rcs(emfa_ctx *ctx, const emfa_mesh *M, double k, const emfa_mat *X, int ndir,
                                const double *dirs, const int *cols, double *out_db) {
    EMFA_REQUIRE(ctx && M && X && dirs && out_db, EMFA_E_INVALID, "emfa_rcs: null argument");
    EMFA_REQUIRE(X->rows == M->n && ndir > 0, EMFA_E_INVALID, "emfa_rcs: X has %d rows, mesh %d RWGs", X->rows, M->n);
    if (cols)
        for (int d = 0; d < ndir; ++d)
            EMFA_REQUIRE(cols[d] >= 0 && cols[d] < X->cols, EMFA_E_INVALID, "emfa_rcs: column %d out of range",
                         cols[d]);
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    size_t bytes = ndir * 3 * sizeof(double) + ndir * sizeof(double) + (cols ? ndir * sizeof(int) : 0);
    void *buf = nullptr;
    EMFA_HIP_CHECK(ctx_alloc(ctx, &buf, bytes));
    double *ddirs = (double *)buf;
    double *dout = ddirs + 3 * ndir;
    int *dcols = cols ? (int *)(dout + ndir) : nullptr;
    EMFA_HIP_CHECK(hipMemcpyAsync(ddirs, dirs, 3 * sizeof(double) * ndir, hipMemcpyHostToDevice, ctx->stream));
    if (cols) EMFA_HIP_CHECK(hipMemcpyAsync(dcols, cols, sizeof(int) * ndir, hipMemcpyHostToDevice, ctx->stream));
    {
        ProfScope ps(ctx, "rcs", 0);
        if (X->prec == EMFA_F32)
            k_rcs<float><<<ndir, 256, 0, ctx->stream>>>(M->tg64, M->rtp, M->rtm, M->rap, M->ram, M->rlen64, M->nt,
                                                        M->n, (const cx<float> *)X->d, X->cols, ddirs, dcols, k,
                                                        dout);
        else
            k_rcs<double><<<ndir, 256, 0, ctx->stream>>>(M->tg64, M->rtp, M->rtm, M->rap, M->ram, M->rlen64, M->nt,
                                                         M->n, (const cx<double> *)X->d, X->cols, ddirs, dcols, k,
                                                         dout);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out_db, dout, ndir * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    ctx_free(ctx, buf, bytes);
    EMFA_HIP_CHECK(e);
    return EMFA_OK;
}
