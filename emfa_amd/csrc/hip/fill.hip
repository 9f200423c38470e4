// Impedance fill on gfx950: triangle-pair formulation.
//
// Every RWG-pair entry of EMFA's fill (EMSolver/EFIE.cpp:603-696,
// EMSolver/CFIE.cpp:485-568) is a sum over the 4 triangle pairs the two RWGs
// span of  s_f s_s l_f l_s * Z_tri(tf, a; ts, b), where a (b) is the local
// index of the RWG's free vertex in tf (ts) and s = +1 on the RWG's plus
// triangle, -1 on its minus triangle: every term of the reference kernels is
// bilinear in rho_f, rho_s, so the RWG signs factor out.  Z_tri is computed
// once per triangle pair (9 values) instead of once per RWG pair (the
// reference evaluates each triangle pair 9 times), then gathered.
//
//   K1 k_fill_pairs     one wavefront per (test triangle, 64 source triangles),
//                       source geometry staged in LDS, one lane per source
//                       triangle, quadrature moments accumulated in registers
//   K2 k_fill_singular  FP64 self terms (tf == ts), EFIE.cpp:231-542 and
//                       Miscellaneous.cpp:32-77
//   K1b k_fill_gather   Z(m,n) += l_m l_n s_m (Ztri(n+) - Ztri(n-)) per side
#include <cmath>

#include "em_device.h"
#include "internal.h"

namespace emfa {

// ---------------------------------------------------------------------------
// K1: regular triangle pairs.
//   KIND 0 (EFIE):  Ztri = cE * sum w (1/4 rho_f.rho_s - 1/k^2) G
//   KIND 1 (CFIE):  Ztri = cE * (same) + cM * sum w rho_f.(n_f x (u x rho_s)) (1+jkR) G/R^2
// with G = exp(-jkR)/R, u = r_p - r_q, rho_f = r_p - v_a, rho_s = r_q - v_b.
// Points are kept relative to each triangle's centroid (x_p, y_q) and the
// integrand is accumulated as moments, from which the 9 (a, b) values follow:
//   rho_f.rho_s = x.y - x.B - A.y + A.B          (A = v_a - c_f, B = v_b - c_s)
//   rho_f.(n x (u x rho_s)) = (rho_f.u)(n.rho_s) - (rho_f.rho_s)(n.u)
// ---------------------------------------------------------------------------
template <typename R, int Q, int KIND>
__global__ __launch_bounds__(256) void k_fill_pairs(const R *__restrict__ tg, int T, int t0, int tcount, R k, R ik2,
                                                    R cE_re, R cE_im, R cM, cx<R> *__restrict__ Ztri) {
    __shared__ R s_src[15][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tsb = blockIdx.x * 64;
    for (int i = threadIdx.x; i < 15 * 64; i += 256) {
        int f = i >> 6, l = i & 63, ts = tsb + l;
        s_src[f][l] = ts < T ? tg[(size_t)f * T + ts] : R(0);
    }
    __syncthreads();
    const int tfl = blockIdx.y * 4 + wave;
    if (tfl >= tcount) return;
    const int tf = __builtin_amdgcn_readfirstlane(t0 + tfl);
    const int ts = tsb + lane;
    if (ts >= T) return;

    R cf[3], Af[3][3], nf[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cf[c] = tg[(size_t)c * T + tf];
        nf[c] = tg[(size_t)(12 + c) * T + tf];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) Af[i][c] = tg[(size_t)(3 + 3 * i + c) * T + tf];
    R Bs[3][3], d0[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) d0[c] = cf[c] - s_src[c][lane];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) Bs[i][c] = s_src[3 + 3 * i + c][lane];

    cx<R> *out = Ztri + (size_t)(tfl * 9) * T + ts;
    if (ts == tf) {
#pragma unroll
        for (int ab = 0; ab < 9; ++ab) out[(size_t)ab * T] = {R(0), R(0)};
        return;
    }

    R y[Q][3], wq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        wq[q] = (R)rule<Q>(q, 3);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            y[q][c] = (R)rule<Q>(q, 0) * Bs[0][c] + (R)rule<Q>(q, 1) * Bs[1][c] + (R)rule<Q>(q, 2) * Bs[2][c];
    }

    // EFIE moments (re, im)
    R S0r = 0, S0i = 0, Sxyr = 0, Sxyi = 0;
    R Sxr[3] = {0, 0, 0}, Sxi[3] = {0, 0, 0}, Syr[3] = {0, 0, 0}, Syi[3] = {0, 0, 0};
    // CFIE moments
    R P1r = 0, P1i = 0, P2r = 0, P2i = 0, T0r = 0, T0i = 0, Txyr = 0, Txyi = 0;
    R Q1r[3] = {0, 0, 0}, Q1i[3] = {0, 0, 0}, Q2r[3] = {0, 0, 0}, Q2i[3] = {0, 0, 0};
    R Txr[3] = {0, 0, 0}, Txi[3] = {0, 0, 0}, Tyr[3] = {0, 0, 0}, Tyi[3] = {0, 0, 0};

#pragma unroll
    for (int p = 0; p < Q; ++p) {
        R x[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            x[c] = (R)rule<Q>(p, 0) * Af[0][c] + (R)rule<Q>(p, 1) * Af[1][c] + (R)rule<Q>(p, 2) * Af[2][c];
        const R wp = (R)rule<Q>(p, 3);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            R u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) u[c] = d0[c] + x[c] - y[q][c];
            R r2 = dot3(u, u);
            R r = sqrt(r2);
            R ir = R(1) / r;
            R sn, cs;
            sincos_r(k * r, &sn, &cs);
            R w = wp * wq[q];
            R gr = w * cs * ir, gi = -w * sn * ir;  // w * exp(-jkR)/R
            R xy = dot3(x, y[q]);
            S0r += gr;
            S0i += gi;
            Sxyr += gr * xy;
            Sxyi += gi * xy;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Sxr[c] += gr * x[c];
                Sxi[c] += gi * x[c];
                Syr[c] += gr * y[q][c];
                Syi[c] += gi * y[q][c];
            }
            if (KIND == 1) {
                // Kc = w (1 + jkR) G / R^2
                R ir2 = ir * ir;
                R kr = k * r;
                R kcr = (gr - kr * gi) * ir2, kci = (gi + kr * gr) * ir2;
                R xu = dot3(x, u), ny = dot3(nf, y[q]), nu = dot3(nf, u);
                P1r += kcr * xu * ny;
                P1i += kci * xu * ny;
                P2r += kcr * xu;
                P2i += kci * xu;
                R knr = kcr * nu, kni = kci * nu;
                T0r += knr;
                T0i += kni;
                Txyr += knr * xy;
                Txyi += kni * xy;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    Q1r[c] += kcr * u[c] * ny;
                    Q1i[c] += kci * u[c] * ny;
                    Q2r[c] += kcr * u[c];
                    Q2i[c] += kci * u[c];
                    Txr[c] += knr * x[c];
                    Txi[c] += kni * x[c];
                    Tyr[c] += knr * y[q][c];
                    Tyi[c] += kni * y[q][c];
                }
            }
        }
    }

#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const R *A = Af[a];
            const R *B = Bs[b];
            R AB = dot3(A, B);
            R Ir = R(0.25) * (Sxyr - dot3(Sxr, B) - dot3(A, Syr) + AB * S0r) - ik2 * S0r;
            R Ii = R(0.25) * (Sxyi - dot3(Sxi, B) - dot3(A, Syi) + AB * S0i) - ik2 * S0i;
            R zr = cE_re * Ir - cE_im * Ii;
            R zi = cE_re * Ii + cE_im * Ir;
            if (KIND == 1) {
                R nB = dot3(nf, B);
                R mr = (P1r - nB * P2r - dot3(A, Q1r) + nB * dot3(A, Q2r)) -
                       (Txyr - dot3(Txr, B) - dot3(A, Tyr) + AB * T0r);
                R mi = (P1i - nB * P2i - dot3(A, Q1i) + nB * dot3(A, Q2i)) -
                       (Txyi - dot3(Txi, B) - dot3(A, Tyi) + AB * T0i);
                zr += cM * mr;
                zi += cM * mi;
            }
            out[(size_t)(a * 3 + b) * T] = {zr, zi};
        }
    }
}

// ---------------------------------------------------------------------------
// K2: self terms in FP64 (self_term in em_device.h).
// ---------------------------------------------------------------------------
template <typename R, int Q>
__global__ __launch_bounds__(64) void k_fill_singular(const double *__restrict__ tg, int T, int t0, int tcount, int kind,
                                                      double k, double alpha, double sig_re, double sig_im,
                                                      int cfie_mode, cx<R> *__restrict__ Ztri,
                                                      double *__restrict__ Zsig) {
    int id = blockIdx.x * 64 + threadIdx.x;
    if (id >= tcount * 9) return;
    int tfl = id / 9, ab = id % 9, a = ab / 3, b = ab % 3;
    int tf = t0 + tfl;
    d3 t[3];
    for (int i = 0; i < 3; ++i)
        t[i] = mk3(tg[(size_t)(3 + 3 * i) * T + tf], tg[(size_t)(4 + 3 * i) * T + tf], tg[(size_t)(5 + 3 * i) * T + tf]);
    double area = tg[(size_t)15 * T + tf];
    double zr, zi, gr, gi;
    self_term<Q>(kind, k, alpha, sig_re, sig_im, cfie_mode, t, area, a, b, &zr, &zi, &gr, &gi);
    Ztri[(size_t)(tfl * 9 + ab) * T + tf] = {(R)zr, (R)zi};
    if (Zsig) {
        Zsig[2 * ((size_t)tfl * 9 + ab)] = gr;
        Zsig[2 * ((size_t)tfl * 9 + ab) + 1] = gi;
    }
}

// ---------------------------------------------------------------------------
// Gather: for rows m whose `side` triangle is in the chunk,
//   Z(m,n) += l_m l_n s_m [Ztri(tf,a_m; n+, b_n+) - Ztri(tf,a_m; n-, b_n-)]
//             + l_m l_n [Zsig(tf,a_m,b) for the source triangle equal to tf]
// Each side is a separate launch, so no two threads update one entry.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_fill_gather(const int *__restrict__ t2r, const int *__restrict__ rtp,
                                                     const int *__restrict__ rtm, const int *__restrict__ rap,
                                                     const int *__restrict__ ram, const R *__restrict__ rlen, int T,
                                                     int N, int t0, int side, const cx<R> *__restrict__ Ztri,
                                                     const double *__restrict__ Zsig, cx<R> *__restrict__ Z) {
    const int e = blockIdx.y;  // (tfl, a)
    const int tfl = e / 3, a = e % 3;
    const int tf = t0 + tfl;
    const int code = t2r[(size_t)a * T + tf];
    if (code == 0) return;
    if ((code > 0) != (side > 0)) return;
    const int m = (code > 0 ? code : -code) - 1;
    const R s = code > 0 ? R(1) : R(-1);
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const int tp = rtp[n], tm = rtm[n], bp = rap[n], bm = ram[n];
    const cx<R> *row = Ztri + (size_t)(tfl * 3 + a) * 3 * T;
    cx<R> zp = row[(size_t)bp * T + tp];
    cx<R> zm = row[(size_t)bm * T + tm];
    R vr = s * (zp.re - zm.re), vi = s * (zp.im - zm.im);
    if (Zsig) {
        if (tp == tf) {
            vr += (R)Zsig[2 * ((size_t)(tfl * 3 + a) * 3 + bp)];
            vi += (R)Zsig[2 * ((size_t)(tfl * 3 + a) * 3 + bp) + 1];
        }
        if (tm == tf) {
            vr += (R)Zsig[2 * ((size_t)(tfl * 3 + a) * 3 + bm)];
            vi += (R)Zsig[2 * ((size_t)(tfl * 3 + a) * 3 + bm) + 1];
        }
    }
    R ll = rlen[m] * rlen[n];
    cx<R> *z = Z + (size_t)m * N + n;
    cx<R> old = *z;
    *z = {old.re + ll * vr, old.im + ll * vi};
}

template <typename R>
static emfa_status fill_impl(emfa_ctx *ctx, const emfa_mesh *M, const emfa_fill_params *p, emfa_mat *Zm) {
    const int T = M->nt, N = M->n;
    const R *tg = nullptr;
    const R *rlen = nullptr;
    if constexpr (sizeof(R) == 4) {
        tg = (const R *)M->tg32;
        rlen = (const R *)M->rlen32;
    } else {
        tg = (const R *)M->tg64;
        rlen = (const R *)M->rlen64;
    }
    const bool sigma = p->kind == EMFA_EFIE && std::hypot(p->sigma_re, p->sigma_im) > 1e-8;
    size_t per_tri = 9 * (size_t)T * sizeof(cx<R>);
    size_t budget = (size_t)2 << 30;
    if (const char *s = getenv("EMFA_FILL_SCRATCH_MB")) budget = (size_t)atol(s) << 20;
    int chunk = (int)std::min<size_t>((size_t)T, std::max<size_t>(4, budget / per_tri));
    chunk = std::min(chunk, 16384);
    chunk = (chunk + 3) & ~3;
    size_t bytes = (size_t)chunk * per_tri + (sigma ? (size_t)chunk * 9 * 2 * sizeof(double) : 0);
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, bytes, &scr);
    if (st != EMFA_OK) return st;
    cx<R> *Ztri = (cx<R> *)scr;
    double *Zsig = sigma ? (double *)((char *)scr + (size_t)chunk * per_tri) : nullptr;
    cx<R> *Z = (cx<R> *)Zm->d;
    EMFA_HIP_CHECK(hipMemsetAsync(Z, 0, Zm->bytes, ctx->stream));

    const double k = p->k;
    const double ce = k * kZ0 / (4.0 * kPi);  // |j k Z0 / 4 pi|
    R cE_re = 0, cE_im, cM = 0;
    if (p->kind == EMFA_EFIE) {
        cE_im = (R)ce;
    } else {
        cE_im = (R)(p->alpha * ce);
        cM = (R)((1.0 - p->alpha) * kZ0 / (16.0 * kPi));
    }
    const R kr = (R)k, ik2 = (R)(1.0 / (k * k));
    for (int t0 = 0; t0 < T; t0 += chunk) {
        int tc = std::min(chunk, T - t0);
        {
            ProfScope ps(ctx, "fill_pairs", (double)tc * T * p->quad * p->quad);
            dim3 grid((T + 63) / 64, (tc + 3) / 4);
#define LAUNCH_PAIRS(Q, KIND) \
    k_fill_pairs<R, Q, KIND><<<grid, 256, 0, ctx->stream>>>(tg, T, t0, tc, kr, ik2, cE_re, cE_im, cM, Ztri)
            if (p->quad == 7) {
                if (p->kind == EMFA_EFIE)
                    LAUNCH_PAIRS(7, 0);
                else
                    LAUNCH_PAIRS(7, 1);
            } else {
                if (p->kind == EMFA_EFIE)
                    LAUNCH_PAIRS(3, 0);
                else
                    LAUNCH_PAIRS(3, 1);
            }
#undef LAUNCH_PAIRS
            EMFA_HIP_CHECK(hipGetLastError());
        }
        {
            dim3 grid((tc * 9 + 63) / 64);
            if (p->quad == 7)
                k_fill_singular<R, 7><<<grid, 64, 0, ctx->stream>>>(M->tg64, T, t0, tc, p->kind, k, p->alpha,
                                                                     p->sigma_re, p->sigma_im, p->cfie_mode, Ztri,
                                                                     Zsig);
            else
                k_fill_singular<R, 3><<<grid, 64, 0, ctx->stream>>>(M->tg64, T, t0, tc, p->kind, k, p->alpha,
                                                                     p->sigma_re, p->sigma_im, p->cfie_mode, Ztri,
                                                                     Zsig);
            EMFA_HIP_CHECK(hipGetLastError());
        }
        {
            ProfScope ps(ctx, "fill_gather", 0);
            dim3 grid((N + 255) / 256, tc * 3);
            for (int side = 1; side >= -1; side -= 2)
                k_fill_gather<R><<<grid, 256, 0, ctx->stream>>>(M->t2r, M->rtp, M->rtm, M->rap, M->ram, rlen, T, N,
                                                                t0, side, Ztri, Zsig, Z);
            EMFA_HIP_CHECK(hipGetLastError());
        }
    }
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

extern "C" emfa_status emfa_fill(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_fill_params *p, emfa_mat *Z) {
    EMFA_REQUIRE(ctx && mesh && p && Z, EMFA_E_INVALID, "emfa_fill: null argument");
    EMFA_REQUIRE(Z->rows == mesh->n && Z->cols == mesh->n, EMFA_E_INVALID, "emfa_fill: Z is %dx%d, mesh has %d RWGs",
                 Z->rows, Z->cols, mesh->n);
    EMFA_REQUIRE(p->kind == EMFA_EFIE || p->kind == EMFA_CFIE, EMFA_E_INVALID, "emfa_fill: bad kind %d", p->kind);
    EMFA_REQUIRE(p->quad == 3 || p->quad == 7, EMFA_E_INVALID, "emfa_fill: quadrature must be 3 or 7 (got %d)",
                 p->quad);
    EMFA_REQUIRE(p->k > 0, EMFA_E_INVALID, "emfa_fill: wave number must be positive");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    if (Z->prec == EMFA_F32) return fill_impl<float>(ctx, mesh, p, Z);
    return fill_impl<double>(ctx, mesh, p, Z);
}
