// Mixed-precision iterative refinement: the factor (FP32 or FP64) gives the
// corrections, the residual R = B - Z X is formed in FP64 with the MFMA
// FP64 product (Z read in its stored precision).
#include <cmath>
#include <vector>

#include "internal.h"

namespace emfa {

// per-column sum of squares of an FP64 complex n x m row-major matrix
__global__ __launch_bounds__(256) void k_colnorm2(const double *__restrict__ X, int n, int m,
                                                  double *__restrict__ out) {
    const int j = blockIdx.x;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double re = X[2 * ((size_t)i * m + j)], im = X[2 * ((size_t)i * m + j) + 1];
        s += re * re + im * im;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

static emfa_status colnorms(emfa_ctx *ctx, const emfa_mat *X, double *dbuf, std::vector<double> &host) {
    k_colnorm2<<<X->cols, 256, 0, ctx->stream>>>((const double *)X->d, X->rows, X->cols, dbuf);
    EMFA_HIP_CHECK(hipGetLastError());
    host.resize(X->cols);
    EMFA_HIP_CHECK(hipMemcpyAsync(host.data(), dbuf, sizeof(double) * X->cols, hipMemcpyDeviceToHost, ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (auto &v : host) v = std::sqrt(v);
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

extern "C" emfa_status emfa_refine(emfa_ctx *ctx, const emfa_mat *Z, const emfa_lu *lu, const emfa_mat *B,
                                   emfa_mat *X, double tol, int maxit, emfa_refine_report *rep) {
    EMFA_REQUIRE(ctx && Z && lu && B && X, EMFA_E_INVALID, "emfa_refine: null argument");
    const int n = Z->rows, m = B->cols;
    EMFA_REQUIRE(Z->cols == n && B->rows == n && X->rows == n && X->cols == m && X->prec == EMFA_F64 &&
                     lu->a->rows == n,
                 EMFA_E_INVALID, "emfa_refine: shape/precision mismatch (X must be F64 n x nrhs)");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    emfa_mat *Bd = nullptr, *R = nullptr, *D = nullptr;
    double *dn = nullptr;
    emfa_status st = EMFA_OK;
    std::vector<double> bn, rn;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(ctx->stream);
        emfa_mat_destroy(Bd);
        emfa_mat_destroy(R);
        emfa_mat_destroy(D);
        ctx_free(ctx, dn, sizeof(double) * m);
    };
    auto relres = [&](double &mx) -> emfa_status {
        emfa_status s2 = colnorms(ctx, R, dn, rn);
        if (s2) return s2;
        mx = 0;
        for (int j = 0; j < m; ++j) mx = std::max(mx, bn[j] > 0 ? rn[j] / bn[j] : rn[j]);
        return EMFA_OK;
    };
    if (ctx_alloc(ctx, (void **)&dn, sizeof(double) * m) != hipSuccess) {
        (void)hipGetLastError();
        set_error("emfa_refine: allocation failed");
        return EMFA_E_OOM;
    }
    st = emfa_mat_create(ctx, n, m, EMFA_F64, &Bd);
    if (!st) st = emfa_mat_create(ctx, n, m, EMFA_F64, &R);
    if (!st) st = emfa_mat_create(ctx, n, m, lu->a->prec, &D);
    if (!st) st = convert_copy(ctx, B, Bd);
    if (!st) st = colnorms(ctx, Bd, dn, bn);
    // initial solve in the factor's precision
    if (!st) st = convert_copy(ctx, Bd, D);
    if (!st) st = lu_solve_inplace(ctx, lu, D);
    if (!st) st = convert_copy(ctx, D, X);
    double r0 = 0, r = 0;
    int it = 0;
    if (!st) st = convert_copy(ctx, Bd, R);
    if (!st) st = gemm_residual(ctx, Z, X, R, 1, 1);
    if (!st) st = relres(r0);
    r = r0;
    while (!st && it < maxit && r > tol) {
        st = convert_copy(ctx, R, D);
        if (!st) st = lu_solve_inplace(ctx, lu, D);
        if (!st) st = axpy_f64(ctx, X, D);
        if (!st) st = convert_copy(ctx, Bd, R);
        if (!st) st = gemm_residual(ctx, Z, X, R, 1, 1);
        if (!st) st = relres(r);
        ++it;
    }
    cleanup();
    if (rep) {
        rep->iterations = it;
        rep->relres_first = r0;
        rep->relres_final = r;
    }
    if (st) return st;
    if (r > tol) {
        set_error("emfa_refine: residual %.3e above tolerance %.3e after %d steps", r, tol, it);
        return EMFA_E_NOCONV;
    }
    return EMFA_OK;
}
