// Multilevel fast multipole matvec (CFIE/EFIE) on gfx950.
//
// Same algorithm as EMSolver/MLFMA.cpp (setup :688-1422, matvec :454-686):
//   octree        cube = extent + 0.1 lambda (EMComponent/MLFMA_Boundary.cpp:15-35),
//                 halved until the box edge is <= threshold * lambda; RWGs
//                 binned by edge midpoint into octants (x*2+y)*2+z; near =
//                 children of the parent's near boxes closer than 1.75 box
//                 edges (= Chebyshev offset <= 1 on the box grid), far = rest
//   samples       L = int(kd + 6 cbrt(kd)), kd = 1.732051 k a; L Gauss-Legendre
//                 theta nodes, 2L Gauss-Legendre phi nodes, sample k = p*L + t
//   interpolation 3-point Lagrange per angle, 9 weights per coarse sample
//   translation   sum_{l <= min(L, kD)} (2l+1) (-j)^l h_l^(2)(kD) P_l(khat.Dhat),
//                 tabulated per box-grid offset (at most 343 per level) instead
//                 of per box pair
//   rad / rec     3-point radiation / receive patterns of each RWG about its
//                 leaf-box centre; the receive pattern carries the quadrature
//                 coefficient w sin(theta) k^2 Z0 / 32
// Patterns are stored [box or RWG][K][3]: the three components of a sample
// sit together (a component-major layout measured slower on gfx950).
//   near          CFIE (or EFIE) entries of all RWG pairs in neighbouring leaf
//                 boxes, FP32 with FP64 self terms, stored CSR in box order
// Kernels (FP32 complex, one writer per output, no atomics, so repeated
// matvecs are bitwise identical):
//   k_ml_near     one wavefront per row of the near matrix
//   k_ml_agg      leaf aggregation: one workgroup per (leaf box, 256 samples)
//   k_ml_up       interpolation + shift to the parent box, per (parent, samples)
//   k_ml_down     translation from far boxes + anterpolation from the parent
//   k_ml_out      receive (dot with the leaf down-going pattern) + near part,
//                 scattered back to RWG order in FP64
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "em_device.h"
#include "internal.h"

namespace emfa {

using cf = cx<float>;

struct MlLevel {
    int nbox = 0, L = 0, K = 0;
    int offsets_used = 0;  // box-grid offsets that occur among the far pairs
    double a = 0;
    std::vector<double> cen;  // nbox x 3
    std::vector<int> ix;      // nbox x 3 integer box coordinates
    std::vector<int> parent, octant, ubeg, uend;
    std::vector<int> near_ptr, near_idx, far_ptr, far_idx, far_code, ch_ptr, ch_idx;
    std::vector<double> xt, wt, xp, wp, W, kh, theta;  // samples
    // device
    cf *up = nullptr, *down = nullptr;
    cf *trans = nullptr;  // 343 x K
    int *far_ptr_d = nullptr, *far_idx_d = nullptr, *far_code_d = nullptr;
    int *ch_ptr_d = nullptr, *ch_idx_d = nullptr, *oct_d = nullptr, *parent_d = nullptr;
    int *ip_col = nullptr;  // interpolation from the child level: K x 9
    float *ip_w = nullptr;
    cf *ph = nullptr;       // 8 x K: exp(-j k khat.(c_parent - c_child)) per child octant
    int *an_ptr = nullptr, *an_kp = nullptr;  // anterpolation from the parent level
    float *an_w = nullptr;
    int *ubeg_d = nullptr, *uend_d = nullptr;  // leaf level: unknown ranges (box order)
    int *near_ptr_d = nullptr, *near_idx_d = nullptr, *ix_d = nullptr;
};

}  // namespace emfa

struct emfa_mlfma {
    emfa_ctx *ctx = nullptr;
    int n = 0, levels = 0, kind = EMFA_CFIE;
    double k = 0, alpha = 1;
    std::vector<emfa::MlLevel> lv;
    int *perm = nullptr, *leaf_of = nullptr;
    long long *row_ptr = nullptr;
    int *col = nullptr;
    emfa::cf *val = nullptr;
    emfa::cf *rad = nullptr, *rec = nullptr;
    emfa::cf *xb = nullptr, *yn = nullptr;
    double *leaf_cen = nullptr, *kh_leaf = nullptr;
    long long nnz = 0;
    std::vector<std::pair<void *, size_t>> allocs;
    size_t bytes = 0;
    // single-level operator (emfa_fmm_create): its one level of boxes is
    // lv[2] (no interpolation, every non-near box far); grid = boxes per axis
    bool single = false;
    int grid[3] = {0, 0, 0};
    // sparse approximate inverse P (emfa_mlfma_precondition): CSR over the
    // box-order rows; with it the operator is P Z and right-hand sides P B
    long long *sai_ptr = nullptr;
    int *sai_col = nullptr;
    emfa::cf *sai_val = nullptr;
    long long sai_nnz = 0;
    double2 *ybox = nullptr;  // Z x in box order, before P
    // channels per pattern sample: 3 (vector), 4 for the IP-SIE operator
    // (emfa_fastwt_create: a scalar line channel besides the vector ones)
    int nc = 3;
    std::vector<double> mid;  // unknown centres (n x 3, unknown order): the binning points
    // IP-SIE half-RWG tables (unknown order): triangle, local index of the
    // free vertex, FP32 / FP64 length
    int *wt_tri = nullptr, *wt_fv = nullptr;
    float *wt_len32 = nullptr;
    double *wt_len64 = nullptr;
    // edge number * 2 + type (0 conformal, 1 nonconformal) per half-RWG; the
    // contour lines of the nonconformal edges (CSR by nonconformal index:
    // triangle, FP32 length); the stabilization factor beta
    int *wt_edge = nullptr, *wt_lptr = nullptr, *wt_ltri = nullptr;
    float *wt_llen = nullptr;
    double wt_beta = 0;
    int nt = 0;  // triangles of the mesh the operator was built on
};

namespace emfa {

// ------------------------------------------------------------------ host math
// Gauss-Legendre nodes (ascending) and weights on [-1, 1] by Newton iteration.
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0);
    w.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int l = 2; l <= n; ++l) {
                double p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) {
                p1 = z;
                p0 = 1;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
            double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        double p0 = 1, p1 = z;
        for (int l = 2; l <= n; ++l) {
            double p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
    }
}

// Three-point Lagrange weights of every `cur` abscissa on the sorted `nxt`
// grid (EMSolver/MLFMA.cpp:957-1002, same index walk).
static void lagrange3(const std::vector<double> &nxt, const std::vector<double> &cur, std::vector<int> &loc,
                      std::vector<double> &el) {
    const int n = (int)cur.size(), nlast = (int)nxt.size() - 1;
    loc.assign(3 * n, 0);
    el.assign(3 * n, 0);
    int nid = 2;
    const double low = nxt[1], high = nxt[nlast - 1];
    for (int c = 0; c < n; ++c) {
        const double xx = cur[c];
        int p0, p1, p2;
        if (xx < low) {
            p0 = 0, p1 = 1, p2 = 2;
        } else if (xx > high) {
            p0 = nlast - 2, p1 = nlast - 1, p2 = nlast;
        } else {
            nid += nxt[nid] < xx ? 1 : 0;
            int off = (nxt[nid] - xx < xx - nxt[nid - 1]) ? 1 : 0;
            p0 = nid + off - 2, p1 = nid + off - 1, p2 = nid + off;
        }
        const double x0 = nxt[p0], x1 = nxt[p1], x2 = nxt[p2];
        loc[3 * c] = p0, loc[3 * c + 1] = p1, loc[3 * c + 2] = p2;
        el[3 * c] = (xx - x1) * (xx - x2) / ((x0 - x1) * (x0 - x2));
        el[3 * c + 1] = (xx - x0) * (xx - x2) / ((x1 - x0) * (x1 - x2));
        el[3 * c + 2] = (xx - x0) * (xx - x1) / ((x2 - x0) * (x2 - x1));
    }
}

static int multipoles(double k, double a) {
    double kd = 1.732051 * k * a;
    return (int)(kd + 6.0 * std::cbrt(kd));
}

// Translation operator for box offset D (receiving minus source centre).
static void translation(double k, const double *D, const std::vector<double> &kh, int Lt, cf *out) {
    using cd = std::complex<double>;
    const double d = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    const double x = k * d;
    const int lmax = (int)std::min((double)Lt, x);
    std::vector<cd> h(lmax + 1);
    h[0] = cd(std::sin(x), std::cos(x)) / x;
    if (lmax >= 1) h[1] = cd(std::sin(x) / x - std::cos(x), std::cos(x) / x + std::sin(x)) / x;
    for (int l = 2; l <= lmax; ++l) h[l] = ((2 * l - 1) / x) * h[l - 1] - h[l - 2];
    const cd jl[4] = {cd(1, 0), cd(0, -1), cd(-1, 0), cd(0, 1)};
    std::vector<cd> c(lmax + 1);
    for (int l = 0; l <= lmax; ++l) c[l] = double(2 * l + 1) * jl[l % 4] * h[l];
    const int K = (int)kh.size() / 3;
    for (int q = 0; q < K; ++q) {
        double ct = (kh[3 * q] * D[0] + kh[3 * q + 1] * D[1] + kh[3 * q + 2] * D[2]) / d;
        double p0 = 1, p1 = ct;
        cd acc = c[0];
        if (lmax >= 1) acc += c[1] * p1;
        for (int l = 2; l <= lmax; ++l) {
            double p2 = ((2 * l - 1) * ct * p1 - (l - 1) * p0) / l;
            p0 = p1;
            p1 = p2;
            acc += c[l] * p2;
        }
        out[q] = {(float)acc.real(), (float)acc.imag()};
    }
}

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ cf cfma(cf acc, cf a, cf b) {
    return {acc.re + a.re * b.re - a.im * b.im, acc.im + a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cf cscale(cf a, float s) { return {a.re * s, a.im * s}; }

// Regular CFIE/EFIE triangle-pair integral for free vertices a (test) and b
// (source), 3 x 3 points, FP32 (same integrand as the dense fill's K1).
template <int KIND>
__device__ cf pair_regular(const float *__restrict__ tg, int T, int tf, int a, int ts, int b, float k, float ik2,
                           float cE, float cM) {
    float cfv[3], csv[3], Af[3], Bs[3], nf[3], vf[3][3], vs[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cfv[c] = tg[(size_t)c * T + tf];
        csv[c] = tg[(size_t)c * T + ts];
        nf[c] = tg[(size_t)(12 + c) * T + tf];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            vf[i][c] = tg[(size_t)(3 + 3 * i + c) * T + tf];
            vs[i][c] = tg[(size_t)(3 + 3 * i + c) * T + ts];
        }
        Af[c] = vf[a][c];
        Bs[c] = vs[b][c];
    }
    float er = 0, ei = 0, mr = 0, mi = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        float x[3], rf[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[c] = (float)rule<3>(p, 0) * vf[0][c] + (float)rule<3>(p, 1) * vf[1][c] + (float)rule<3>(p, 2) * vf[2][c];
            rf[c] = x[c] - Af[c];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            float y[3], rs[3], u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                y[c] = (float)rule<3>(q, 0) * vs[0][c] + (float)rule<3>(q, 1) * vs[1][c] + (float)rule<3>(q, 2) * vs[2][c];
                rs[c] = y[c] - Bs[c];
                u[c] = cfv[c] - csv[c] + x[c] - y[c];
            }
            float r = sqrtf(dot3(u, u));
            float ir = 1.0f / r;
            float sn, cs;
            sincos_r(k * r, &sn, &cs);
            const float w = (1.0f / 9.0f);
            float gr = w * cs * ir, gi = -w * sn * ir;
            float f = 0.25f * dot3(rf, rs) - ik2;
            er += f * gr;
            ei += f * gi;
            if (KIND == EMFA_CFIE) {
                float ir2 = ir * ir, kr = k * r;
                float kcr = (gr - kr * gi) * ir2, kci = (gi + kr * gr) * ir2;
                float m = dot3(rf, u) * dot3(nf, rs) - dot3(rf, rs) * dot3(nf, u);
                mr += kcr * m;
                mi += kci * m;
            }
        }
    }
    // cE is the imaginary coefficient j * cE
    return {-cE * ei + cM * mr, cE * er + cM * mi};
}

template <int KIND>
__global__ __launch_bounds__(256) void k_ml_near_fill(const float *__restrict__ tg, const double *__restrict__ tg64,
                                                      int T, const int *__restrict__ rtp, const int *__restrict__ rtm,
                                                      const int *__restrict__ rap, const int *__restrict__ ram,
                                                      const float *__restrict__ rlen, const int *__restrict__ perm,
                                                      const long long *__restrict__ row_ptr,
                                                      const int *__restrict__ col, int n, float k, float ik2, float cE,
                                                      float cM, double kd, double alpha, int cfie_mode,
                                                      cf *__restrict__ val) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int f = perm[row];
    const int tfs[2] = {rtp[f], rtm[f]}, afs[2] = {rap[f], ram[f]};
    const float lf = rlen[f];
    for (long long j = row_ptr[row] + lane; j < row_ptr[row + 1]; j += 64) {
        const int s = perm[col[j]];
        const int tss[2] = {rtp[s], rtm[s]}, bss[2] = {rap[s], ram[s]};
        float zr = 0, zi = 0;
        for (int i = 0; i < 2; ++i) {
            for (int q = 0; q < 2; ++q) {
                const float sg = (i == q) ? 1.0f : -1.0f;
                const int tf = tfs[i], ts = tss[q];
                if (tf == ts) {
                    d3 t[3];
                    for (int v = 0; v < 3; ++v)
                        t[v] = mk3(tg64[(size_t)(3 + 3 * v) * T + tf], tg64[(size_t)(4 + 3 * v) * T + tf],
                                   tg64[(size_t)(5 + 3 * v) * T + tf]);
                    double sr, si, gr, gi;
                    self_term<3>(KIND, kd, alpha, 0.0, 0.0, cfie_mode, t, tg64[(size_t)15 * T + tf], afs[i], bss[q],
                                 &sr, &si, &gr, &gi);
                    zr += sg * (float)sr;
                    zi += sg * (float)si;
                } else {
                    cf z = pair_regular<KIND>(tg, T, tf, afs[i], ts, bss[q], k, ik2, cE, cM);
                    zr += sg * z.re;
                    zi += sg * z.im;
                }
            }
        }
        const float ll = lf * rlen[s];
        val[j] = {ll * zr, ll * zi};
    }
}

// Radiation and receive patterns of RWG perm[i] about its leaf-box centre
// (EMSolver/MLFMA.cpp:1244-1292,1393-1422), evaluated in FP64, stored FP32.
__global__ __launch_bounds__(256) void k_ml_radrec(const double *__restrict__ tg, int T, const int *__restrict__ rtp,
                                                   const int *__restrict__ rtm, const int *__restrict__ rap,
                                                   const int *__restrict__ ram, const double *__restrict__ rlen,
                                                   const int *__restrict__ perm, const int *__restrict__ leaf_of,
                                                   const double *__restrict__ leaf_cen,
                                                   const double *__restrict__ kh, const double *__restrict__ coeff,
                                                   int n, int K, double k, double alpha, cf *__restrict__ rad,
                                                   cf *__restrict__ rec) {
    const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long long)n * K) return;
    const int i = (int)(id / K), q = (int)(id % K);
    const int u = perm[i], b = leaf_of[i];
    const double kx = kh[3 * q], ky = kh[3 * q + 1], kz = kh[3 * q + 2];
    double vr[3] = {0, 0, 0}, vi[3] = {0, 0, 0}, hr[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int side = 0; side < 2; ++side) {
        const int t = side == 0 ? rtp[u] : rtm[u];
        const int a = side == 0 ? rap[u] : ram[u];
        const double sg = side == 0 ? 1.0 : -1.0;
        double c[3], v[3][3], nn[3];
        for (int cc = 0; cc < 3; ++cc) {
            c[cc] = tg[(size_t)cc * T + t] - leaf_cen[3 * b + cc];
            nn[cc] = tg[(size_t)(12 + cc) * T + t];
            for (int j = 0; j < 3; ++j) v[j][cc] = tg[(size_t)(3 + 3 * j + cc) * T + t];
        }
        for (int p = 0; p < 3; ++p) {
            double x[3], rho[3];
            for (int cc = 0; cc < 3; ++cc) {
                x[cc] = rule<3>(p, 0) * v[0][cc] + rule<3>(p, 1) * v[1][cc] + rule<3>(p, 2) * v[2][cc];
                rho[cc] = sg * (x[cc] - v[a][cc]);
            }
            const double ph = k * (kx * (c[0] + x[0]) + ky * (c[1] + x[1]) + kz * (c[2] + x[2]));
            double sn, cs;
            sincos(ph, &sn, &cs);
            const double w = 1.0 / 3.0;
            double nr[3] = {nn[1] * rho[2] - nn[2] * rho[1], nn[2] * rho[0] - nn[0] * rho[2],
                            nn[0] * rho[1] - nn[1] * rho[0]};
            for (int cc = 0; cc < 3; ++cc) {
                vr[cc] += w * rho[cc] * cs;  // exp(+j ph)
                vi[cc] += w * rho[cc] * sn;
                hr[cc] += w * nr[cc] * cs;   // exp(-j ph)
                hi[cc] -= w * nr[cc] * sn;
            }
        }
    }
    // transverse projection
    const double pr = kx * vr[0] + ky * vr[1] + kz * vr[2], pi = kx * vi[0] + ky * vi[1] + kz * vi[2];
    const double kv[3] = {kx, ky, kz};
    for (int cc = 0; cc < 3; ++cc) {
        vr[cc] -= pr * kv[cc];
        vi[cc] -= pi * kv[cc];
    }
    // k x h
    double cr[3] = {ky * hr[2] - kz * hr[1], kz * hr[0] - kx * hr[2], kx * hr[1] - ky * hr[0]};
    double ci[3] = {ky * hi[2] - kz * hi[1], kz * hi[0] - kx * hi[2], kx * hi[1] - ky * hi[0]};
    const double hl = 0.5 * rlen[u], cq = coeff[q];
    cf *ro = rad + ((size_t)i * K + q) * 3;
    cf *co = rec + ((size_t)i * K + q) * 3;
    for (int cc = 0; cc < 3; ++cc) {
        ro[cc] = {(float)(hl * vr[cc]), (float)(hl * vi[cc])};
        double er = alpha * vr[cc] + (1 - alpha) * cr[cc];
        double ei = -alpha * vi[cc] + (1 - alpha) * ci[cc];
        co[cc] = {(float)(cq * hl * er), (float)(cq * hl * ei)};
    }
}

__global__ __launch_bounds__(256) void k_ml_gather(const double2 *__restrict__ x, const int *__restrict__ perm, int n,
                                                   cf *__restrict__ xb) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        double2 v = x[perm[i]];
        xb[i] = {(float)v.x, (float)v.y};
    }
}

__global__ __launch_bounds__(256) void k_ml_near(const long long *__restrict__ row_ptr, const int *__restrict__ col,
                                                 const cf *__restrict__ val, const cf *__restrict__ xb, int n,
                                                 cf *__restrict__ yn) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float sr = 0, si = 0;
    for (long long j = row_ptr[row] + lane; j < row_ptr[row + 1]; j += 64) {
        cf a = val[j], b = xb[col[j]];
        sr += a.re * b.re - a.im * b.im;
        si += a.re * b.im + a.im * b.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) yn[row] = {sr, si};
}

// Patterns carry NC channels per sample: the three vector components (and,
// for the IP-SIE operator, a fourth scalar line channel), stored together.
// leaf aggregation: up[b][q] = sum_{i in b} x_i rad[i][q], U unknowns per
// loop step with their loads issued together (sums in unknown order for any U)
template <int NC, int U>
__global__ __launch_bounds__(256) void k_ml_agg(const int *__restrict__ ubeg, const int *__restrict__ uend,
                                                const cf *__restrict__ rad, const cf *__restrict__ xb, int K,
                                                int nbox, cf *__restrict__ up) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const int b = (int)(idx / K), q = (int)(idx % K);
    if (b >= nbox) return;
    cf a[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = {0, 0};
    int i = ubeg[b];
    const int i_end = uend[b];
    for (; i + U <= i_end; i += U) {
        cf x[U], r[U][NC];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            x[j] = xb[i + j];
            const cf *src = rad + ((size_t)(i + j) * K + q) * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) r[j][c] = src[c];
        }
#pragma unroll
        for (int j = 0; j < U; ++j)
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], x[j], r[j][c]);
    }
    for (; i < i_end; ++i) {
        const cf x = xb[i];
        const cf *r = rad + ((size_t)i * K + q) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], x, r[c]);
    }
    cf *o = up + ((size_t)b * K + q) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = a[c];
}

// upward: up_p[q] = sum_children ph[oct][q] * sum_9 w * up_c[col]; U children
// per loop step with their loads issued together (sums in child order)
template <int NC, int U>
__global__ __launch_bounds__(256) void k_ml_up(const int *__restrict__ ch_ptr, const int *__restrict__ ch_idx,
                                               const int *__restrict__ ch_oct, const int *__restrict__ ip_col,
                                               const float *__restrict__ ip_w, const cf *__restrict__ ph, int K,
                                               int Kc, int nbox, const cf *__restrict__ upc, cf *__restrict__ upp) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const int p = (int)(idx / K), q = (int)(idx % K);
    if (p >= nbox) return;
    int cols[9];
    float ws[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        cols[j] = ip_col[q * 9 + j];
        ws[j] = ip_w[q * 9 + j];
    }
    cf a[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = {0, 0};
    int e = ch_ptr[p];
    const int e_end = ch_ptr[p + 1];
    for (; e < e_end; e += U) {
        const int nu = min(U, e_end - e);
        cf sm[U][NC], f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int c = 0; c < NC; ++c) sm[u][c] = {0, 0};
            f[u] = {0, 0};
            if (u < nu) {
                const int ch = ch_idx[e + u];
                const cf *src = upc + (size_t)ch * Kc * NC;
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    const cf *v = src + (size_t)cols[j] * NC;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const cf vc = v[c];
                        sm[u][c] = {sm[u][c].re + ws[j] * vc.re, sm[u][c].im + ws[j] * vc.im};
                    }
                }
                f[u] = ph[ch_oct[ch] * K + q];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (u < nu)
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], f[u], sm[u][c]);
    }
    cf *o = upp + ((size_t)p * K + q) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = a[c];
}

// downward at one level: translation from far boxes + anterpolation from the
// parent.  U > 1 takes the far list and the anterpolation list U entries at a
// time, their indices and operands loaded before the U products, so that U
// gathers are in flight per thread instead of one; the sums run in the same
// order for every U (the results are bitwise those of U = 1).
template <int NC, int U>
__global__ __launch_bounds__(256) void k_ml_down(const int *__restrict__ far_ptr, const int *__restrict__ far_idx,
                                                 const int *__restrict__ far_code, const cf *__restrict__ trans,
                                                 const cf *__restrict__ up, int K, const int *__restrict__ parent,
                                                 const int *__restrict__ oct, const int *__restrict__ an_ptr,
                                                 const int *__restrict__ an_kp, const float *__restrict__ an_w,
                                                 const cf *__restrict__ php, int Kp, const cf *__restrict__ downp,
                                                 int nbox, cf *__restrict__ down) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const int b = (int)(idx / K), q = (int)(idx % K);
    if (b >= nbox) return;
    cf a[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) a[c] = {0, 0};
    int e = far_ptr[b];
    const int e_end = far_ptr[b + 1];
    for (; e + U <= e_end; e += U) {
        int fi[U], fc[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            fi[j] = far_idx[e + j];
            fc[j] = far_code[e + j];
        }
        cf t[U], u[U][NC];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            t[j] = trans[(size_t)fc[j] * K + q];
            const cf *src = up + ((size_t)fi[j] * K + q) * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) u[j][c] = src[c];
        }
#pragma unroll
        for (int j = 0; j < U; ++j)
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], t[j], u[j][c]);
    }
    for (; e < e_end; ++e) {
        const cf t = trans[(size_t)far_code[e] * K + q];
        const cf *u = up + ((size_t)far_idx[e] * K + q) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], t, u[c]);
    }
    if (downp) {
        const int p = parent[b], o = oct[b];
        const cf *dp = downp + (size_t)p * Kp * NC;
        const cf *phase = php + (size_t)o * Kp;
        int g = an_ptr[q];
        const int g_end = an_ptr[q + 1];
        for (; g + U <= g_end; g += U) {
            int kp[U];
            float w[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                kp[j] = an_kp[g + j];
                w[j] = an_w[g + j];
            }
            cf f[U], v[U][NC];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const cf f0 = phase[kp[j]];
                f[j] = {w[j] * f0.re, -w[j] * f0.im};  // conj: exp(+j k khat.(c_p - c_b))
#pragma unroll
                for (int c = 0; c < NC; ++c) v[j][c] = dp[(size_t)kp[j] * NC + c];
            }
#pragma unroll
            for (int j = 0; j < U; ++j)
#pragma unroll
                for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], f[j], v[j][c]);
        }
        for (; g < g_end; ++g) {
            const int kq = an_kp[g];
            const cf f0 = phase[kq];
            const cf f = {an_w[g] * f0.re, -an_w[g] * f0.im};
            const cf *v = dp + (size_t)kq * NC;
#pragma unroll
            for (int c = 0; c < NC; ++c) a[c] = cfma(a[c], f, v[c]);
        }
    }
    cf *o = down + ((size_t)b * K + q) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = a[c];
}

// Downward pass at level l >= 3 grouped by parent: the children of parent P
// share their candidate source boxes (children of P's near boxes), so each
// candidate's pattern is read once and applied to every sibling for which it
// is far.  Per child the far boxes are visited in the same order as its far
// list, so the sums equal k_ml_down's.  Then the anterpolation from P.
__global__ __launch_bounds__(256) void k_ml_down_sib(
    const int *__restrict__ pnear_ptr, const int *__restrict__ pnear_idx, const int *__restrict__ pch_ptr,
    const int *__restrict__ pch_idx, int nparent, const int *__restrict__ ix, const int *__restrict__ oct,
    const cf *__restrict__ trans, const cf *__restrict__ up, int K, const int *__restrict__ an_ptr,
    const int *__restrict__ an_kp, const float *__restrict__ an_w, const cf *__restrict__ php, int Kp,
    const cf *__restrict__ downp, cf *__restrict__ down) {
    const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
    const int P = (int)(idx / K), q = (int)(idx % K);
    if (P >= nparent) return;
    const int c0 = pch_ptr[P], nch = min(8, pch_ptr[P + 1] - c0);
    int bx[8], by[8], bz[8];
    cf a[8][3];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int b = s < nch ? pch_idx[c0 + s] : 0;
        bx[s] = ix[3 * b];
        by[s] = ix[3 * b + 1];
        bz[s] = ix[3 * b + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c) a[s][c] = {0, 0};
    }
    for (int e = pnear_ptr[P]; e < pnear_ptr[P + 1]; ++e) {
        const int pn = pnear_idx[e];
        for (int f = pch_ptr[pn]; f < pch_ptr[pn + 1]; ++f) {
            const int c = pch_idx[f];
            const int cx = ix[3 * c], cy = ix[3 * c + 1], cz = ix[3 * c + 2];
            const cf *u = up + ((size_t)c * K + q) * 3;
            const cf u0 = u[0], u1 = u[1], u2 = u[2];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (s >= nch) break;
                const int dx = bx[s] - cx, dy = by[s] - cy, dz = bz[s] - cz;
                if (max(abs(dx), max(abs(dy), abs(dz))) <= 1) continue;
                const cf t = trans[(size_t)((dx + 3) * 49 + (dy + 3) * 7 + (dz + 3)) * K + q];
                a[s][0] = cfma(a[s][0], t, u0);
                a[s][1] = cfma(a[s][1], t, u1);
                a[s][2] = cfma(a[s][2], t, u2);
            }
        }
    }
    const cf *dp = downp + (size_t)P * Kp * 3;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        if (s >= nch) break;
        const int b = pch_idx[c0 + s], o = oct[b];
        cf a0 = a[s][0], a1 = a[s][1], a2 = a[s][2];
        for (int e = an_ptr[q]; e < an_ptr[q + 1]; ++e) {
            const int kp = an_kp[e];
            const cf f0 = php[o * Kp + kp];
            const cf f = {an_w[e] * f0.re, -an_w[e] * f0.im};
            const cf *v = dp + (size_t)kp * 3;
            a0 = cfma(a0, f, v[0]);
            a1 = cfma(a1, f, v[1]);
            a2 = cfma(a2, f, v[2]);
        }
        cf *o3 = down + ((size_t)b * K + q) * 3;
        o3[0] = a0;
        o3[1] = a1;
        o3[2] = a2;
    }
}

// receive + near part, back to RWG order: y[perm[i]] = yn[i] + sum_q rec[i][q] . down[leaf(i)][q]
template <int NC>
__global__ __launch_bounds__(256) void k_ml_out(const cf *__restrict__ rec, const cf *__restrict__ down,
                                                const int *__restrict__ leaf_of, const int *__restrict__ perm,
                                                const cf *__restrict__ yn, int n, int K, double2 *__restrict__ y,
                                                double2 *__restrict__ ybox) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const cf *d = down + (size_t)leaf_of[i] * K * NC;
    const cf *r = rec + (size_t)i * K * NC;
    float sr = 0, si = 0;
    for (int q = lane; q < K; q += 64) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cf a = r[NC * q + c], b = d[NC * q + c];
            sr += a.re * b.re - a.im * b.im;
            si += a.re * b.im + a.im * b.re;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) {
        cf v = yn[i];
        const double2 o = make_double2((double)v.re + (double)sr, (double)v.im + (double)si);
        if (ybox)
            ybox[i] = o;
        else
            y[perm[i]] = o;
    }
}

// y[perm[i]] = sum_j P(i, j) x[j]: the sparse approximate inverse (CSR over
// box-order rows, FP32 entries) applied to a box-order FP64 vector, one
// wavefront per row, FP64 sums
__global__ __launch_bounds__(256) void k_ml_sai(const long long *__restrict__ ptr, const int *__restrict__ col,
                                                const cf *__restrict__ val, const double2 *__restrict__ x,
                                                const int *__restrict__ perm, int n, double2 *__restrict__ y) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double sr = 0, si = 0;
    for (long long j = ptr[row] + lane; j < ptr[row + 1]; j += 64) {
        const cf a = val[j];
        const double2 b = x[col[j]];
        sr += (double)a.re * b.x - (double)a.im * b.y;
        si += (double)a.re * b.y + (double)a.im * b.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) y[perm[row]] = make_double2(sr, si);
}

// box-order copy of an FP64 vector in RWG order
__global__ __launch_bounds__(256) void k_ml_gather64(const double2 *__restrict__ x, const int *__restrict__ perm,
                                                     int n, double2 *__restrict__ xb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) xb[i] = x[perm[i]];
}


// ------------------------------------------------------------------ IP-SIE
// Half-RWG interior-penalty operator (EMSolver/FastWT.cpp).  A half-RWG u
// lives on one triangle wt_tri[u] with free vertex wt_fv[u] (local index) and
// the edge opposite it; its basis is l_u / (2 A) (r - v_free) there.

// 3-point Gauss-Legendre on [0, 1] (EMComponent/Quadrature.h:17-18)
__device__ __forceinline__ double gl3(int p) { return p == 0 ? 0.1127016654 : p == 1 ? 0.5 : 0.8872983346; }
__device__ __forceinline__ double glw3(int p) { return p == 1 ? 0.444444444444445 : 0.2777777777775; }

// len sum_p sum_q glw3[p] w3[q] e^{-jkr} / r between the three points of the
// edge opposite local vertex b of triangle ts and the three Gauss points of
// triangle tt (FastWT.cpp:1494-1511 without the factor Z0 / (4 pi j k)), FP32
__device__ cf wt_line(const float *__restrict__ tg, int T, int tt, int ts, int b, float len, float k) {
    const int i0 = b == 2 ? 0 : b + 1, i1 = b == 0 ? 2 : b - 1;
    float ct[3], cs[3], vt[3][3], e0[3], e1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ct[c] = tg[(size_t)c * T + tt];
        cs[c] = tg[(size_t)c * T + ts];
#pragma unroll
        for (int i = 0; i < 3; ++i) vt[i][c] = tg[(size_t)(3 + 3 * i + c) * T + tt];
        e0[c] = tg[(size_t)(3 + 3 * i0 + c) * T + ts];
        e1[c] = tg[(size_t)(3 + 3 * i1 + c) * T + ts];
    }
    float ar = 0, ai = 0;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const float g = (float)gl3(p), wp = (float)(glw3(p) / 3.0);
        float ep[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) ep[c] = e0[c] + g * (e1[c] - e0[c]);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            float u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                u[c] = ct[c] - cs[c] +
                       ((float)rule<3>(q, 0) * vt[0][c] + (float)rule<3>(q, 1) * vt[1][c] +
                        (float)rule<3>(q, 2) * vt[2][c]) -
                       ep[c];
            const float r = sqrtf(dot3(u, u)), ir = 1.0f / r;
            float sn, cs_;
            sincos_r(k * r, &sn, &cs_);
            ar += wp * cs_ * ir;
            ai -= wp * sn * ir;
        }
    }
    return {len * ar, len * ai};
}

// Near entries of the IP-SIE operator (FastWT.cpp:734-784), one wavefront per
// row of the near matrix (field f = perm[row]) over its columns (source s):
//   z    = l_f l_s K(f, s): K the CFIE triangle-pair kernel with both signs
//          +1, FP32 regular / FP64 self term on a shared triangle;
//   line = -(l_f C(T_f; edge s) + l_s C(T_s; edge f)), C = Z0 / (4 pi j k) x
//          wt_line;
//   ip   = stab l_f for f = s and for the two conformal halves of one edge on
//          different triangles; stab x (length of f's first contour line on
//          s's triangle) between nonconformal edges; stab = 2 beta Z0 / k;
//   entry = z + alpha (line + ip).
template <int KIND>
__global__ __launch_bounds__(256) void k_wt_near_fill(
    const float *__restrict__ tg, const double *__restrict__ tg64, int T, const int *__restrict__ wtri,
    const int *__restrict__ wfv, const float *__restrict__ wlen, const int *__restrict__ wedge,
    const int *__restrict__ lptr, const int *__restrict__ ltri, const float *__restrict__ llen,
    const int *__restrict__ perm, const long long *__restrict__ row_ptr, const int *__restrict__ col, int n, float k,
    float ik2, float cE, float cM, double kd, double alpha, int cfie_mode, float cl, float stab, cf *__restrict__ val) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const int f = perm[row];
    const int tf = wtri[f], af = wfv[f], ef = wedge[f];
    const float lf = wlen[f], a32 = (float)alpha;
    for (long long j = row_ptr[row] + lane; j < row_ptr[row + 1]; j += 64) {
        const int s = perm[col[j]];
        const int ts = wtri[s], as = wfv[s], es = wedge[s];
        const float ls = wlen[s];
        float zr, zi, ip = 0.0f;
        if (tf == ts) {
            d3 t[3];
            for (int v = 0; v < 3; ++v)
                t[v] = mk3(tg64[(size_t)(3 + 3 * v) * T + tf], tg64[(size_t)(4 + 3 * v) * T + tf],
                           tg64[(size_t)(5 + 3 * v) * T + tf]);
            double sr, si, gr, gi;
            self_term<3>(KIND, kd, alpha, 0.0, 0.0, cfie_mode, t, tg64[(size_t)15 * T + tf], af, as, &sr, &si, &gr,
                         &gi);
            zr = (float)sr;
            zi = (float)si;
            if (f == s) ip = stab * lf;
        } else {
            const cf z = pair_regular<KIND>(tg, T, tf, af, ts, as, k, ik2, cE, cM);
            zr = z.re;
            zi = z.im;
            if (!(ef & 1) && ef == es) {
                ip = stab * lf;
            } else if ((ef & 1) && (es & 1)) {
                const int e = ef >> 1;
                for (int q = lptr[e]; q < lptr[e + 1]; ++q)
                    if (ltri[q] == ts) {
                        ip = stab * llen[q];
                        break;
                    }
            }
        }
        const float ll = lf * ls;
        const cf pa = wt_line(tg, T, tf, ts, as, ls, k), pb = wt_line(tg, T, ts, tf, af, lf, k);
        const float sr = lf * pa.re + ls * pb.re, si = lf * pa.im + ls * pb.im;
        // line = -Z0 / (4 pi j k) (sr + j si) = cl (-si + j sr), cl = Z0 / (4 pi k)
        val[j] = {ll * zr + a32 * (-cl * si + ip), ll * zi + a32 * (cl * sr)};
    }
}

// Radiation / receive patterns of half-RWG perm[i] about its leaf-box centre
// (FastWT.cpp:583-653), FP64, stored FP32 with four channels: the vector
// patterns as the MLFMA's with one triangle, and the line pattern
// prad = l sum_p glw3 e^{+jk khat.(gl_p - c)} on the edge, prec = conj(prad)
// times coeff alpha / k^2 (= pcoeff, FastWT.cpp:986-989).
__global__ __launch_bounds__(256) void k_wt_radrec(const double *__restrict__ tg, int T, const int *__restrict__ wtri,
                                                   const int *__restrict__ wfv, const double *__restrict__ wlen,
                                                   const int *__restrict__ perm, const int *__restrict__ leaf_of,
                                                   const double *__restrict__ leaf_cen,
                                                   const double *__restrict__ kh, const double *__restrict__ coeff,
                                                   int n, int K, double k, double alpha, cf *__restrict__ rad,
                                                   cf *__restrict__ rec) {
    const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long long)n * K) return;
    const int i = (int)(id / K), q = (int)(id % K);
    const int u = perm[i], b = leaf_of[i];
    const int t = wtri[u], a = wfv[u];
    const double kx = kh[3 * q], ky = kh[3 * q + 1], kz = kh[3 * q + 2];
    double c[3], v[3][3], nn[3];
    for (int cc = 0; cc < 3; ++cc) {
        c[cc] = tg[(size_t)cc * T + t] - leaf_cen[3 * b + cc];
        nn[cc] = tg[(size_t)(12 + cc) * T + t];
        for (int j = 0; j < 3; ++j) v[j][cc] = tg[(size_t)(3 + 3 * j + cc) * T + t];
    }
    double vr[3] = {0, 0, 0}, vi[3] = {0, 0, 0}, hr[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int p = 0; p < 3; ++p) {
        double x[3], rho[3];
        for (int cc = 0; cc < 3; ++cc) {
            x[cc] = rule<3>(p, 0) * v[0][cc] + rule<3>(p, 1) * v[1][cc] + rule<3>(p, 2) * v[2][cc];
            rho[cc] = x[cc] - v[a][cc];
        }
        const double ph = k * (kx * (c[0] + x[0]) + ky * (c[1] + x[1]) + kz * (c[2] + x[2]));
        double sn, cs;
        sincos(ph, &sn, &cs);
        const double w = 1.0 / 3.0;
        const double nr[3] = {nn[1] * rho[2] - nn[2] * rho[1], nn[2] * rho[0] - nn[0] * rho[2],
                              nn[0] * rho[1] - nn[1] * rho[0]};
        for (int cc = 0; cc < 3; ++cc) {
            vr[cc] += w * rho[cc] * cs;
            vi[cc] += w * rho[cc] * sn;
            hr[cc] += w * nr[cc] * cs;
            hi[cc] -= w * nr[cc] * sn;
        }
    }
    const double pr = kx * vr[0] + ky * vr[1] + kz * vr[2], pi = kx * vi[0] + ky * vi[1] + kz * vi[2];
    const double kv[3] = {kx, ky, kz};
    for (int cc = 0; cc < 3; ++cc) {
        vr[cc] -= pr * kv[cc];
        vi[cc] -= pi * kv[cc];
    }
    const double cr[3] = {ky * hr[2] - kz * hr[1], kz * hr[0] - kx * hr[2], kx * hr[1] - ky * hr[0]};
    const double ci[3] = {ky * hi[2] - kz * hi[1], kz * hi[0] - kx * hi[2], kx * hi[1] - ky * hi[0]};
    const double len = wlen[u], hl = 0.5 * len, cq = coeff[q];
    cf *ro = rad + ((size_t)i * K + q) * 4;
    cf *co = rec + ((size_t)i * K + q) * 4;
    for (int cc = 0; cc < 3; ++cc) {
        ro[cc] = {(float)(hl * vr[cc]), (float)(hl * vi[cc])};
        const double er = alpha * vr[cc] + (1 - alpha) * cr[cc];
        const double ei = -alpha * vi[cc] + (1 - alpha) * ci[cc];
        co[cc] = {(float)(cq * hl * er), (float)(cq * hl * ei)};
    }
    const int i0 = a == 2 ? 0 : a + 1, i1 = a == 0 ? 2 : a - 1;
    double lr = 0, li = 0;
    for (int p = 0; p < 3; ++p) {
        double e[3];
        for (int cc = 0; cc < 3; ++cc) e[cc] = c[cc] + v[i0][cc] + gl3(p) * (v[i1][cc] - v[i0][cc]);
        double sn, cs;
        sincos(k * (kx * e[0] + ky * e[1] + kz * e[2]), &sn, &cs);
        lr += glw3(p) * cs;
        li += glw3(p) * sn;
    }
    const double sc = cq * alpha / (k * k) * len;
    ro[3] = {(float)(len * lr), (float)(len * li)};
    co[3] = {(float)(sc * lr), (float)(-sc * li)};
}

// Excitation of every half-RWG (FastWT.cpp:786-806 fillVbyBox with cVKernel
// :1475-1492): V = l/2 sum_a w3 (alpha rho.E + (1 - alpha) rho.(n x H))
// e^{-jk g_a.k_inc}, FP64, half-RWG order; inc: nrhs x 9 (k, e, h).
__global__ __launch_bounds__(256) void k_wt_rhs(const double *__restrict__ tg, int T, const int *__restrict__ wtri,
                                                const int *__restrict__ wfv, const double *__restrict__ wlen, int n,
                                                const double *__restrict__ inc, int nrhs, double k, double alpha,
                                                double2 *__restrict__ B) {
    const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
    if (id >= (long long)n * nrhs) return;
    const int u = (int)(id / nrhs), j = (int)(id % nrhs);
    const int t = wtri[u], a = wfv[u];
    double c[3], v[3][3], nn[3];
    for (int cc = 0; cc < 3; ++cc) {
        c[cc] = tg[(size_t)cc * T + t];
        nn[cc] = tg[(size_t)(12 + cc) * T + t];
        for (int i = 0; i < 3; ++i) v[i][cc] = tg[(size_t)(3 + 3 * i + cc) * T + t];
    }
    const double *in = inc + 9 * (size_t)j;
    const double nh[3] = {nn[1] * in[8] - nn[2] * in[7], nn[2] * in[6] - nn[0] * in[8], nn[0] * in[7] - nn[1] * in[6]};
    double er = 0, ei = 0, mr = 0, mi = 0;
    for (int p = 0; p < 3; ++p) {
        double x[3], rho[3];
        for (int cc = 0; cc < 3; ++cc) {
            x[cc] = rule<3>(p, 0) * v[0][cc] + rule<3>(p, 1) * v[1][cc] + rule<3>(p, 2) * v[2][cc];
            rho[cc] = x[cc] - v[a][cc];
        }
        double sn, cs;
        sincos(-k * ((c[0] + x[0]) * in[0] + (c[1] + x[1]) * in[1] + (c[2] + x[2]) * in[2]), &sn, &cs);
        const double de = (rho[0] * in[3] + rho[1] * in[4] + rho[2] * in[5]) / 3.0;
        const double dm = (rho[0] * nh[0] + rho[1] * nh[1] + rho[2] * nh[2]) / 3.0;
        er += de * cs;
        ei += de * sn;
        mr += dm * cs;
        mi += dm * sn;
    }
    const double hl = 0.5 * wlen[u];
    B[(size_t)u * nrhs + j] = make_double2(hl * (alpha * er + (1 - alpha) * mr), hl * (alpha * ei + (1 - alpha) * mi));
}

// Bistatic RCS of column `col` of X (half-RWG order) per direction
// (FastWT.cpp:1393-1420): Es = sum_u I_u l_u sum_g w3 (g - v_free)
// e^{+jk g.s}; sigma = Z0^2 k^2 |s x (s x Es) / 2|^2 / 4 pi, in dB.  One
// workgroup per direction, FP64 sums reduced in a fixed order.
__global__ __launch_bounds__(256) void k_wt_rcs(const double *__restrict__ tg, int T, const int *__restrict__ wtri,
                                                const int *__restrict__ wfv, const double *__restrict__ wlen, int n,
                                                const double2 *__restrict__ X, int ncols, int col,
                                                const double *__restrict__ dirs, double k,
                                                double *__restrict__ out_db) {
    const int d = blockIdx.x;
    const double sx = dirs[3 * d], sy = dirs[3 * d + 1], sz = dirs[3 * d + 2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int u = threadIdx.x; u < n; u += 256) {
        const int t = wtri[u], a = wfv[u];
        double c[3], v[3][3];
        for (int cc = 0; cc < 3; ++cc) {
            c[cc] = tg[(size_t)cc * T + t];
            for (int i = 0; i < 3; ++i) v[i][cc] = tg[(size_t)(3 + 3 * i + cc) * T + t];
        }
        double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0};
        for (int p = 0; p < 3; ++p) {
            double x[3];
            for (int cc = 0; cc < 3; ++cc)
                x[cc] = rule<3>(p, 0) * v[0][cc] + rule<3>(p, 1) * v[1][cc] + rule<3>(p, 2) * v[2][cc];
            double sn, cs;
            sincos(k * ((c[0] + x[0]) * sx + (c[1] + x[1]) * sy + (c[2] + x[2]) * sz), &sn, &cs);
            for (int cc = 0; cc < 3; ++cc) {
                const double r = (x[cc] - v[a][cc]) / 3.0;
                er[cc] += r * cs;
                ei[cc] += r * sn;
            }
        }
        const double2 I = X[(size_t)u * ncols + col];
        const double l = wlen[u], ir = l * I.x, ii = l * I.y;
        for (int cc = 0; cc < 3; ++cc) {
            acc[2 * cc] += er[cc] * ir - ei[cc] * ii;
            acc[2 * cc + 1] += er[cc] * ii + ei[cc] * ir;
        }
    }
    __shared__ double red[6][256];
    for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double Er[3] = {red[0][0], red[2][0], red[4][0]}, Ei[3] = {red[1][0], red[3][0], red[5][0]};
        const double s[3] = {sx, sy, sz};
        double fr[3], fi[3];
        // f = s x (s x E) / 2 = (s (s.E) - E) / 2
        const double pr = s[0] * Er[0] + s[1] * Er[1] + s[2] * Er[2], pi = s[0] * Ei[0] + s[1] * Ei[1] + s[2] * Ei[2];
        for (int c = 0; c < 3; ++c) {
            fr[c] = 0.5 * (s[c] * pr - Er[c]);
            fi[c] = 0.5 * (s[c] * pi - Ei[c]);
        }
        const double nrm = fr[0] * fr[0] + fi[0] * fi[0] + fr[1] * fr[1] + fi[1] * fi[1] + fr[2] * fr[2] + fi[2] * fi[2];
        out_db[d] = 10.0 * log10(kZ0 * kZ0 * k * k * nrm / (4.0 * kPi));
    }
}

// ------------------------------------------------------------------ host
template <typename T>
static emfa_status ml_upload(emfa_mlfma *M, T **dst, const std::vector<T> &src) {
    *dst = nullptr;
    if (src.empty()) return EMFA_OK;
    size_t b = src.size() * sizeof(T);
    void *p = nullptr;
    if (ctx_alloc(M->ctx, &p, b) != hipSuccess) {
        (void)hipGetLastError();
        set_error("mlfma: cannot allocate %zu bytes", b);
        return EMFA_E_OOM;
    }
    M->allocs.push_back({p, b});
    M->bytes += b;
    EMFA_HIP_CHECK(hipMemcpyAsync(p, src.data(), b, hipMemcpyHostToDevice, M->ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(M->ctx->stream));  // src may be a temporary
    *dst = (T *)p;
    return EMFA_OK;
}

template <typename T>
static emfa_status ml_alloc(emfa_mlfma *M, T **dst, size_t count) {
    *dst = nullptr;
    size_t b = std::max<size_t>(count, 1) * sizeof(T);
    void *p = nullptr;
    if (ctx_alloc(M->ctx, &p, b) != hipSuccess) {
        (void)hipGetLastError();
        set_error("mlfma: cannot allocate %zu bytes", b);
        return EMFA_E_OOM;
    }
    M->allocs.push_back({p, b});
    M->bytes += b;
    *dst = (T *)p;
    return EMFA_OK;
}

static void ml_free(emfa_mlfma *M) {
    if (!M) return;
    if (M->ctx) (void)hipStreamSynchronize(M->ctx->stream);
    for (auto &a : M->allocs) ctx_free(M->ctx, a.first, a.second);
    M->allocs.clear();
}

#define ML_TRY(x)                        \
    do {                                 \
        emfa_status _s = (x);            \
        if (_s != EMFA_OK) {             \
            ml_free(M);                  \
            delete M;                    \
            return _s;                   \
        }                                \
    } while (0)

static emfa_status ml_leaf_setup(emfa_ctx *ctx, emfa_mlfma *M, const emfa_mesh *mesh, const emfa_mlfma_params *prm,
                                 const std::vector<int> &order, emfa_mlfma **out);

// Octree, samples and per-level tables of the multilevel operator over n
// unknowns binned by their centres mid (n x 3, unknown order); on success *Mout
// holds the operator without its leaf part and `order` the box order.
static emfa_status ml_tree(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *prm, int n,
                           const std::vector<double> &mid, int nc, emfa_mlfma **Mout, std::vector<int> &order) {
    const int nv = mesh->nv;
    const double k = prm->k, lam = 2 * kPi / k;
    const std::vector<double> &xyz = mesh->xyz_host;
    // octree (MLFMA_Boundary.cpp:15-35)
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nv; ++v)
        for (int c = 0; c < 3; ++c) {
            mn[c] = std::min(mn[c], xyz[3 * v + c]);
            mx[c] = std::max(mx[c], xyz[3 * v + c]);
        }
    double center[3], ext = 0;
    for (int c = 0; c < 3; ++c) {
        center[c] = (mn[c] + mx[c]) / 2;
        ext = std::max(ext, mx[c] - mn[c]);
    }
    const double L0 = ext + 0.1 * lam, delta = prm->box_threshold * lam;
    int levels = 0;
    for (double ml = L0; ml > delta; ml /= 2) ++levels;
    EMFA_REQUIRE(levels >= 2, EMFA_E_INVALID,
                 "mlfma: %d level(s) only; the box threshold (%.3g wavelengths) is too large for this mesh", levels,
                 prm->box_threshold);
    EMFA_REQUIRE(levels <= 14, EMFA_E_INVALID, "mlfma: %d levels (box threshold too small)", levels);
    emfa_mlfma *M = new emfa_mlfma();
    M->ctx = ctx;
    M->n = n;
    M->nc = nc;
    M->levels = levels;
    M->k = k;
    M->alpha = prm->kind == EMFA_EFIE ? 1.0 : prm->alpha;
    M->kind = prm->kind;
    M->lv.resize(levels + 1);
    order.assign(n, 0);
    for (int u = 0; u < n; ++u) order[u] = u;
    {
        MlLevel &r = M->lv[0];
        r.nbox = 1;
        r.a = L0;
        r.cen = {center[0], center[1], center[2]};
        r.ix = {0, 0, 0};
        r.parent = {-1};
        r.octant = {0};
        r.ubeg = {0};
        r.uend = {n};
    }
    for (int l = 1; l <= levels; ++l) {
        MlLevel &P = M->lv[l - 1], &C = M->lv[l];
        C.a = L0 / std::pow(2.0, l);
        P.ch_ptr.assign(P.nbox + 1, 0);
        std::vector<int> tmp;
        for (int p = 0; p < P.nbox; ++p) {
            const double *pc = &P.cen[3 * p];
            std::vector<int> bucket[8];
            for (int i = P.ubeg[p]; i < P.uend[p]; ++i) {
                int u = order[i];
                int o = ((mid[3 * u] < pc[0] ? 0 : 1) * 2 + (mid[3 * u + 1] < pc[1] ? 0 : 1)) * 2 +
                        (mid[3 * u + 2] < pc[2] ? 0 : 1);
                bucket[o].push_back(u);
            }
            int pos = P.ubeg[p];
            for (int o = 0; o < 8; ++o) {
                if (bucket[o].empty()) continue;
                int bx = o >> 2, by = (o >> 1) & 1, bz = o & 1;
                C.cen.push_back(pc[0] + C.a * (bx - 0.5));
                C.cen.push_back(pc[1] + C.a * (by - 0.5));
                C.cen.push_back(pc[2] + C.a * (bz - 0.5));
                C.ix.push_back(2 * P.ix[3 * p] + bx);
                C.ix.push_back(2 * P.ix[3 * p + 1] + by);
                C.ix.push_back(2 * P.ix[3 * p + 2] + bz);
                C.parent.push_back(p);
                C.octant.push_back(o);
                C.ubeg.push_back(pos);
                for (int u : bucket[o]) order[pos++] = u;
                C.uend.push_back(pos);
                P.ch_idx.push_back(C.nbox);
                ++C.nbox;
            }
            P.ch_ptr[p + 1] = (int)P.ch_idx.size();
        }
    }
    // near / far lists (MLFMA.cpp:836-875)
    {
        MlLevel &l1 = M->lv[1];
        l1.near_ptr.assign(l1.nbox + 1, 0);
        for (int b = 0; b < l1.nbox; ++b) {
            for (int c = 0; c < l1.nbox; ++c) l1.near_idx.push_back(c);
            l1.near_ptr[b + 1] = (int)l1.near_idx.size();
        }
        l1.far_ptr.assign(l1.nbox + 1, 0);
    }
    for (int l = 2; l <= levels; ++l) {
        MlLevel &P = M->lv[l - 1], &C = M->lv[l];
        C.near_ptr.assign(C.nbox + 1, 0);
        C.far_ptr.assign(C.nbox + 1, 0);
        for (int b = 0; b < C.nbox; ++b) {
            const int p = C.parent[b];
            for (int e = P.near_ptr[p]; e < P.near_ptr[p + 1]; ++e) {
                const int pn = P.near_idx[e];
                for (int f = P.ch_ptr[pn]; f < P.ch_ptr[pn + 1]; ++f) {
                    const int c = P.ch_idx[f];
                    int dx = C.ix[3 * b] - C.ix[3 * c], dy = C.ix[3 * b + 1] - C.ix[3 * c + 1],
                        dz = C.ix[3 * b + 2] - C.ix[3 * c + 2];
                    if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) <= 1) {
                        C.near_idx.push_back(c);
                    } else {
                        C.far_idx.push_back(c);
                        C.far_code.push_back((dx + 3) * 49 + (dy + 3) * 7 + (dz + 3));
                    }
                }
            }
            C.near_ptr[b + 1] = (int)C.near_idx.size();
            C.far_ptr[b + 1] = (int)C.far_idx.size();
        }
    }
    // samples (MLFMA.cpp:877-915,1133-1166)
    for (int l = 2; l <= levels; ++l) {
        MlLevel &C = M->lv[l];
        C.L = std::max(3, multipoles(k, C.a));
        C.K = 2 * C.L * C.L;
        gauss_legendre(C.L, C.xt, C.wt);
        gauss_legendre(2 * C.L, C.xp, C.wp);
        C.kh.resize(3 * (size_t)C.K);
        C.W.resize(C.K);
        C.theta.resize(C.K);
        for (int p = 0; p < 2 * C.L; ++p)
            for (int t = 0; t < C.L; ++t) {
                const int q = p * C.L + t;
                const double th = kPi / 2 * (C.xt[t] + 1), ph = kPi * (C.xp[p] + 1);
                C.kh[3 * q] = std::sin(th) * std::cos(ph);
                C.kh[3 * q + 1] = std::sin(th) * std::sin(ph);
                C.kh[3 * q + 2] = std::cos(th);
                C.W[q] = C.wp[p] * C.wt[t];
                C.theta[q] = th;
            }
    }
    // device tables per level
    for (int l = 2; l <= levels; ++l) {
        MlLevel &C = M->lv[l];
        ML_TRY(ml_alloc(M, &C.up, (size_t)C.nbox * C.K * nc));
        ML_TRY(ml_alloc(M, &C.down, (size_t)C.nbox * C.K * nc));
        ML_TRY(ml_upload(M, &C.far_ptr_d, C.far_ptr));
        ML_TRY(ml_upload(M, &C.far_idx_d, C.far_idx.empty() ? std::vector<int>{0} : C.far_idx));
        ML_TRY(ml_upload(M, &C.far_code_d, C.far_code.empty() ? std::vector<int>{0} : C.far_code));
        ML_TRY(ml_upload(M, &C.parent_d, C.parent));
        ML_TRY(ml_upload(M, &C.oct_d, C.octant));
        ML_TRY(ml_upload(M, &C.ix_d, C.ix));
        ML_TRY(ml_upload(M, &C.near_ptr_d, C.near_ptr));
        ML_TRY(ml_upload(M, &C.near_idx_d, C.near_idx.empty() ? std::vector<int>{0} : C.near_idx));
        // translation per offset code in use
        std::vector<char> used(343, 0);
        for (int c : C.far_code) used[c] = 1;
        std::vector<int> codes;
        for (int code = 0; code < 343; ++code)
            if (used[code]) codes.push_back(code);
        C.offsets_used = (int)codes.size();
        std::vector<cf> tr((size_t)343 * C.K, cf{0, 0});
        {
            const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            for (int t = 0; t < nth; ++t)
                pool.emplace_back([&, t] {
                    for (size_t i = t; i < codes.size(); i += nth) {
                        const int code = codes[i];
                        double D[3] = {C.a * (code / 49 - 3), C.a * ((code / 7) % 7 - 3), C.a * (code % 7 - 3)};
                        translation(k, D, C.kh, C.L, &tr[(size_t)code * C.K]);
                    }
                });
            for (auto &th : pool) th.join();
        }
        ML_TRY(ml_upload(M, &C.trans, tr));
        if (l < levels) {
            // interpolation from level l+1 and the child-octant phases
            MlLevel &N = M->lv[l + 1];
            std::vector<int> tloc, ploc;
            std::vector<double> tel, pel;
            lagrange3(N.xt, C.xt, tloc, tel);
            lagrange3(N.xp, C.xp, ploc, pel);
            std::vector<int> ipc((size_t)C.K * 9);
            std::vector<float> ipw((size_t)C.K * 9);
            std::vector<std::vector<std::pair<int, double>>> tr_list(N.K);
            for (int p = 0; p < 2 * C.L; ++p)
                for (int t = 0; t < C.L; ++t) {
                    const int q = p * C.L + t;
                    int j = 0;
                    for (int ia = 0; ia < 3; ++ia)
                        for (int ib = 0; ib < 3; ++ib, ++j) {
                            const int colc = ploc[3 * p + ia] * N.L + tloc[3 * t + ib];
                            const double w = pel[3 * p + ia] * tel[3 * t + ib];
                            ipc[(size_t)q * 9 + j] = colc;
                            ipw[(size_t)q * 9 + j] = (float)w;
                            tr_list[colc].push_back({q, w * C.W[q] / N.W[colc]});
                        }
                }
            ML_TRY(ml_upload(M, &C.ip_col, ipc));
            ML_TRY(ml_upload(M, &C.ip_w, ipw));
            std::vector<cf> ph((size_t)8 * C.K);
            for (int o = 0; o < 8; ++o) {
                const double r[3] = {-N.a * ((o >> 2) - 0.5), -N.a * (((o >> 1) & 1) - 0.5), -N.a * ((o & 1) - 0.5)};
                for (int q = 0; q < C.K; ++q) {
                    double arg = -k * (C.kh[3 * q] * r[0] + C.kh[3 * q + 1] * r[1] + C.kh[3 * q + 2] * r[2]);
                    ph[(size_t)o * C.K + q] = {(float)std::cos(arg), (float)std::sin(arg)};
                }
            }
            ML_TRY(ml_upload(M, &C.ph, ph));
            std::vector<int> chp(C.ch_ptr), chi(C.ch_idx);
            ML_TRY(ml_upload(M, &C.ch_ptr_d, chp));
            ML_TRY(ml_upload(M, &C.ch_idx_d, chi.empty() ? std::vector<int>{0} : chi));
            // anterpolation into level l+1 (transpose, weighted by the quadrature weights)
            std::vector<int> anp(N.K + 1, 0), ank;
            std::vector<float> anw;
            for (int qc = 0; qc < N.K; ++qc) {
                for (auto &e : tr_list[qc]) {
                    ank.push_back(e.first);
                    anw.push_back((float)e.second);
                }
                anp[qc + 1] = (int)ank.size();
            }
            ML_TRY(ml_upload(M, &N.an_ptr, anp));
            ML_TRY(ml_upload(M, &N.an_kp, ank));
            ML_TRY(ml_upload(M, &N.an_w, anw));
        }
    }
    *Mout = M;
    return EMFA_OK;
}

// RWG edge midpoints (n x 3)
static std::vector<double> rwg_mid(const emfa_mesh *mesh) {
    const int n = mesh->n;
    const std::vector<double> &xyz = mesh->xyz_host;
    const std::vector<int> &rwg = mesh->rwg_host;
    std::vector<double> mid(3 * (size_t)n);
    for (int u = 0; u < n; ++u)
        for (int c = 0; c < 3; ++c) mid[3 * u + c] = 0.5 * (xyz[3 * rwg[6 * u] + c] + xyz[3 * rwg[6 * u + 1] + c]);
    return mid;
}

// IP-SIE operator (EMSolver/FastWT.cpp:26-70,167-400,583-784): the MLFMA's
// octree over the half-RWG edge midpoints (FastWT.cpp:217-226), four pattern
// channels, the half-RWG near entries.  The half-RWG tables are checked
// against the mesh: the free vertex is a vertex of the half-RWG's triangle and
// v1, v2 are its other two.
static emfa_status wt_build(emfa_ctx *ctx, const emfa_mesh *mesh, int n, const int *hrwg, const double *length,
                            int n_nc, const int *line_ptr, const int *line_tri, const double *line_seg,
                            const emfa_mlfma_params *prm, double beta, emfa_mlfma **out) {
    const std::vector<double> &xyz = mesh->xyz_host;
    const std::vector<int> &tri = mesh->tri_host;
    const int nt = mesh->nt;
    std::vector<int> wtri(n), wfv(n), wedge(n);
    std::vector<float> len32(n);
    std::vector<double> len64(n), mid(3 * (size_t)n);
    for (int u = 0; u < n; ++u) {
        const int *h = hrwg + 6 * (size_t)u;
        const int t = h[3];
        EMFA_REQUIRE(t >= 0 && t < nt, EMFA_E_INVALID, "fastwt: half-RWG %d: triangle %d of %d", u, t, nt);
        int a = -1;
        for (int c = 0; c < 3; ++c)
            if (tri[3 * (size_t)t + c] == h[2]) a = c;
        EMFA_REQUIRE(a >= 0, EMFA_E_INVALID, "fastwt: half-RWG %d: free vertex %d not on triangle %d", u, h[2], t);
        const int o1 = tri[3 * (size_t)t + (a + 1) % 3], o2 = tri[3 * (size_t)t + (a + 2) % 3];
        EMFA_REQUIRE((h[0] == o1 && h[1] == o2) || (h[0] == o2 && h[1] == o1), EMFA_E_INVALID,
                     "fastwt: half-RWG %d: edge (%d, %d) is not the edge of triangle %d opposite vertex %d", u, h[0],
                     h[1], t, h[2]);
        EMFA_REQUIRE(h[4] >= 0 && (h[5] == 0 || (h[5] == 1 && h[4] < n_nc)), EMFA_E_INVALID,
                     "fastwt: half-RWG %d: edge number %d / type %d", u, h[4], h[5]);
        EMFA_REQUIRE(length[u] > 0, EMFA_E_INVALID, "fastwt: half-RWG %d: length %g", u, length[u]);
        wtri[u] = t;
        wfv[u] = a;
        wedge[u] = 2 * h[4] + h[5];
        len32[u] = (float)length[u];
        len64[u] = length[u];
        for (int c = 0; c < 3; ++c) mid[3 * (size_t)u + c] = 0.5 * (xyz[3 * (size_t)h[0] + c] + xyz[3 * (size_t)h[1] + c]);
    }
    std::vector<int> lptr(n_nc + 1, 0), ltri;
    std::vector<float> llen;
    if (n_nc > 0) {
        EMFA_REQUIRE(line_ptr && line_tri && line_seg && line_ptr[0] == 0, EMFA_E_INVALID, "fastwt: contour lines");
        for (int e = 0; e < n_nc; ++e) {
            EMFA_REQUIRE(line_ptr[e + 1] >= line_ptr[e], EMFA_E_INVALID, "fastwt: line_ptr not ascending");
            lptr[e + 1] = line_ptr[e + 1];
        }
        const int nl = line_ptr[n_nc];
        ltri.resize(nl);
        llen.resize(nl);
        for (int q = 0; q < nl; ++q) {
            EMFA_REQUIRE(line_tri[q] >= 0 && line_tri[q] < nt, EMFA_E_INVALID, "fastwt: line %d: triangle", q);
            ltri[q] = line_tri[q];
            // FP32 length, as ContourEdge::getContourLineLength computes it
            const double *g = line_seg + 6 * (size_t)q;
            const float dx = (float)g[3] - (float)g[0], dy = (float)g[4] - (float)g[1], dz = (float)g[5] - (float)g[2];
            llen[q] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    if (ltri.empty()) {
        ltri.push_back(0);
        llen.push_back(0.0f);
    }
    std::vector<int> order;
    emfa_mlfma *M = nullptr;
    emfa_status st = ml_tree(ctx, mesh, prm, n, mid, 4, &M, order);
    if (st) return st;
    M->mid = std::move(mid);
    M->wt_beta = beta;
    ML_TRY(ml_upload(M, &M->wt_tri, wtri));
    ML_TRY(ml_upload(M, &M->wt_fv, wfv));
    ML_TRY(ml_upload(M, &M->wt_edge, wedge));
    ML_TRY(ml_upload(M, &M->wt_len32, len32));
    ML_TRY(ml_upload(M, &M->wt_len64, len64));
    ML_TRY(ml_upload(M, &M->wt_lptr, lptr));
    ML_TRY(ml_upload(M, &M->wt_ltri, ltri));
    ML_TRY(ml_upload(M, &M->wt_llen, llen));
    return ml_leaf_setup(ctx, M, mesh, prm, order, out);
}

static emfa_status ml_build(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *prm, emfa_mlfma **out) {
    std::vector<double> mid = rwg_mid(mesh);
    std::vector<int> order;
    emfa_mlfma *M = nullptr;
    emfa_status st = ml_tree(ctx, mesh, prm, mesh->n, mid, 3, &M, order);
    if (st) return st;
    M->mid = std::move(mid);
    return ml_leaf_setup(ctx, M, mesh, prm, order, out);
}

// Leaf part shared by both operators: permutation to box order, leaf of each
// position, near matrix (CSR, box order), radiation / receive patterns and the
// near entries.  M->lv[M->levels] is the leaf level with its boxes, unknown
// ranges, near lists and samples set.  On failure M is freed.
static emfa_status ml_leaf_setup(emfa_ctx *ctx, emfa_mlfma *M, const emfa_mesh *mesh, const emfa_mlfma_params *prm,
                                 const std::vector<int> &order, emfa_mlfma **out) {
    const int n = M->n, T = mesh->nt, levels = M->levels;
    M->nt = T;
    const double k = prm->k;
    // leaf data: permutation, leaf of each position, centres
    MlLevel &F = M->lv[levels];
    std::vector<int> leaf_of(n);
    for (int b = 0; b < F.nbox; ++b)
        for (int i = F.ubeg[b]; i < F.uend[b]; ++i) leaf_of[i] = b;
    ML_TRY(ml_upload(M, &M->perm, order));
    ML_TRY(ml_upload(M, &M->leaf_of, leaf_of));
    ML_TRY(ml_upload(M, &M->leaf_cen, F.cen));
    ML_TRY(ml_upload(M, &M->kh_leaf, F.kh));
    ML_TRY(ml_upload(M, &F.ubeg_d, F.ubeg));
    ML_TRY(ml_upload(M, &F.uend_d, F.uend));
    // near matrix structure (box order)
    std::vector<long long> rp(n + 1, 0);
    std::vector<int> cols;
    {
        size_t total = 0;
        for (int b = 0; b < F.nbox; ++b) {
            size_t w = 0;
            for (int e = F.near_ptr[b]; e < F.near_ptr[b + 1]; ++e) w += F.uend[F.near_idx[e]] - F.ubeg[F.near_idx[e]];
            total += w * (size_t)(F.uend[b] - F.ubeg[b]);
        }
        cols.reserve(total);
    }
    for (int b = 0; b < F.nbox; ++b) {
        for (int i = F.ubeg[b]; i < F.uend[b]; ++i) {
            for (int e = F.near_ptr[b]; e < F.near_ptr[b + 1]; ++e) {
                const int c = F.near_idx[e];
                for (int j = F.ubeg[c]; j < F.uend[c]; ++j) cols.push_back(j);
            }
            rp[i + 1] = (long long)cols.size();
        }
    }
    M->nnz = (long long)cols.size();
    ML_TRY(ml_upload(M, &M->row_ptr, rp));
    ML_TRY(ml_upload(M, &M->col, cols));
    ML_TRY(ml_alloc(M, &M->val, cols.size()));
    ML_TRY(ml_alloc(M, &M->rad, (size_t)n * F.K * M->nc));
    ML_TRY(ml_alloc(M, &M->rec, (size_t)n * F.K * M->nc));
    ML_TRY(ml_alloc(M, &M->xb, n));
    ML_TRY(ml_alloc(M, &M->yn, n));
    // coefficients (MLFMA.cpp:382-400)
    std::vector<double> coeff(F.K);
    for (int q = 0; q < F.K; ++q) coeff[q] = F.W[q] * std::sin(F.theta[q]) * k * k * kZ0 / 32.0;
    double *coeff_d = nullptr;
    ML_TRY(ml_upload(M, &coeff_d, coeff));
    hipStream_t s = ctx->stream;
    {
        ProfScope ps(ctx, "mlfma_setup", 0);
        const long long tot = (long long)n * F.K;
        const double ce = k * kZ0 / (4.0 * kPi);
        const float cE = (float)(M->alpha * ce), cM = (float)((1.0 - M->alpha) * kZ0 / (16.0 * kPi));
        if (M->nc == 4) {
            k_wt_radrec<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(mesh->tg64, T, M->wt_tri, M->wt_fv,
                                                                       M->wt_len64, M->perm, M->leaf_of, M->leaf_cen,
                                                                       M->kh_leaf, coeff_d, n, F.K, k, M->alpha,
                                                                       M->rad, M->rec);
            const float cl = (float)(kZ0 / (4.0 * kPi * k)), stab = (float)(2.0 * M->wt_beta * kZ0 / k);
            k_wt_near_fill<EMFA_CFIE><<<(n + 3) / 4, 256, 0, s>>>(
                mesh->tg32, mesh->tg64, T, M->wt_tri, M->wt_fv, M->wt_len32, M->wt_edge, M->wt_lptr, M->wt_ltri,
                M->wt_llen, M->perm, M->row_ptr, M->col, n, (float)k, (float)(1.0 / (k * k)), cE, cM, k, M->alpha,
                prm->cfie_mode, cl, stab, M->val);
        } else {
            k_ml_radrec<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(mesh->tg64, T, mesh->rtp, mesh->rtm,
                                                                       mesh->rap, mesh->ram, mesh->rlen64, M->perm,
                                                                       M->leaf_of, M->leaf_cen, M->kh_leaf, coeff_d,
                                                                       n, F.K, k, M->alpha, M->rad, M->rec);
            if (M->kind == EMFA_EFIE)
                k_ml_near_fill<EMFA_EFIE><<<(n + 3) / 4, 256, 0, s>>>(
                    mesh->tg32, mesh->tg64, T, mesh->rtp, mesh->rtm, mesh->rap, mesh->ram, mesh->rlen32, M->perm,
                    M->row_ptr, M->col, n, (float)k, (float)(1.0 / (k * k)), (float)ce, 0.0f, k, 1.0, 0, M->val);
            else
                k_ml_near_fill<EMFA_CFIE><<<(n + 3) / 4, 256, 0, s>>>(
                    mesh->tg32, mesh->tg64, T, mesh->rtp, mesh->rtm, mesh->rap, mesh->ram, mesh->rlen32, M->perm,
                    M->row_ptr, M->col, n, (float)k, (float)(1.0 / (k * k)), cE, cM, k, M->alpha, prm->cfie_mode,
                    M->val);
        }
        EMFA_HIP_CHECK(hipGetLastError());
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(s));
    *out = M;
    return EMFA_OK;
}

// Single-level operator (EMSolver/FMM.cpp).  The box grid follows
// EMComponent/FMM_Boundary.cpp:15-53 in FP32 arithmetic, as the reference
// computes it, so that the grid, the box of every RWG and L agree with it:
//   delta = box length * lambda; num = int(extent / delta + 0.5);
//   delta = extent / num; the vertex bounds widened by 0.1 delta; boxes per
//   axis int((size + 0.9 delta) / delta); RWG u in box
//   ((x * ny + y) * nz + z) of its edge midpoint; box centre
//   ((x + 0.5) delta, (y + 0.5) delta, (z + 0.5) delta).
// That centre leaves out the lower bound (FMM_Boundary.cpp:47): every centre
// is shifted by the same vector, which changes each pattern by a phase that
// cancels between radiation and receive, so the operator is the same; the
// shift is kept so that the patterns are the reference's.
// Non-empty boxes in grid order, unknowns of a box in RWG order
// (FMM.cpp:157-206); near = centres closer than 1.75 delta (the 27 grid
// neighbours), far = every other box; L = int(kd + 6 cbrt(kd)),
// kd = 1.732051 k delta (FMM.cpp:47-50); translation per far pair
// (FMM.cpp:349-381), tabulated here per grid offset in use.
static emfa_status fmm_build(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *prm, emfa_mlfma **out) {
#pragma clang fp contract(off)
    const int n = mesh->n, nv = mesh->nv;
    const double k = prm->k;
    const std::vector<double> &xyz = mesh->xyz_host;
    const std::vector<int> &rwg = mesh->rwg_host;
    float mn[3], mx[3];
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = (float)xyz[c];
    for (int v = 1; v < nv; ++v)
        for (int c = 0; c < 3; ++c) {
            const float x = (float)xyz[3 * v + c];
            mn[c] = std::min(mn[c], x);
            mx[c] = std::max(mx[c], x);
        }
    const float lam = (float)(2 * kPi / k);
    float delta = (float)prm->box_threshold;
    delta *= lam;
    float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    const float max_len = ext[0] < ext[1] ? (ext[1] < ext[2] ? ext[2] : ext[1]) : (ext[0] < ext[2] ? ext[2] : ext[0]);
    const float ratio = max_len / delta;
    const int num = (int)(ratio + 0.5f);
    EMFA_REQUIRE(num >= 1, EMFA_E_INVALID,
                 "fmm: the box length (%.3g wavelengths) exceeds twice the mesh extent; no box grid", prm->box_threshold);
    delta = max_len / (float)num;
    const float off = 0.1f * delta;
    int nb[3];
    for (int c = 0; c < 3; ++c) {
        mn[c] = mn[c] - off;
        mx[c] = mx[c] + off;
        const float size = mx[c] - mn[c];
        const float d9 = 0.9f * delta;
        nb[c] = (int)((size + d9) / delta);
    }
    const long long cells = (long long)nb[0] * nb[1] * nb[2];
    EMFA_REQUIRE(cells > 0 && cells < (1LL << 31), EMFA_E_INVALID, "fmm: box grid %d x %d x %d out of range", nb[0],
                 nb[1], nb[2]);
    // box of every RWG (edge midpoint, FP32)
    std::vector<long long> cell(n);
    for (int u = 0; u < n; ++u) {
        int ixc[3];
        for (int c = 0; c < 3; ++c) {
            const float s = (float)xyz[3 * rwg[6 * u] + c] + (float)xyz[3 * rwg[6 * u + 1] + c];
            const float mid = s / 2.0f;
            ixc[c] = (int)((mid - mn[c]) / delta);
            EMFA_REQUIRE(ixc[c] >= 0 && ixc[c] < nb[c], EMFA_E_INVALID, "fmm: RWG %d outside the box grid", u);
        }
        cell[u] = ((long long)ixc[0] * nb[1] + ixc[1]) * nb[2] + ixc[2];
    }
    std::vector<int> order(n);
    for (int u = 0; u < n; ++u) order[u] = u;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cell[a] < cell[b]; });

    emfa_mlfma *M = new emfa_mlfma();
    M->ctx = ctx;
    M->n = n;
    M->levels = 2;
    M->single = true;
    for (int c = 0; c < 3; ++c) M->grid[c] = nb[c];
    M->k = k;
    M->alpha = prm->kind == EMFA_EFIE ? 1.0 : prm->alpha;
    M->kind = prm->kind;
    M->lv.resize(3);
    MlLevel &C = M->lv[2];
    C.a = (double)delta;
    for (int i = 0; i < n;) {
        const long long id = cell[order[i]];
        int j = i;
        while (j < n && cell[order[j]] == id) ++j;
        const int x = (int)(id / ((long long)nb[1] * nb[2])), y = (int)((id / nb[2]) % nb[1]), z = (int)(id % nb[2]);
        C.ix.push_back(x);
        C.ix.push_back(y);
        C.ix.push_back(z);
        C.cen.push_back((double)(((float)x + 0.5f) * delta));
        C.cen.push_back((double)(((float)y + 0.5f) * delta));
        C.cen.push_back((double)(((float)z + 0.5f) * delta));
        C.parent.push_back(-1);
        C.octant.push_back(0);
        C.ubeg.push_back(i);
        C.uend.push_back(j);
        ++C.nbox;
        i = j;
    }
    const int B = C.nbox;
    // near / far lists in box order; far codes = dense index of the grid offset
    const long long sx = 2LL * nb[0] - 1, sy = 2LL * nb[1] - 1, sz = 2LL * nb[2] - 1;
    const long long noff = sx * sy * sz;
    EMFA_REQUIRE(noff < (1LL << 28), EMFA_E_INVALID, "fmm: %lld grid offsets (box grid too fine)", noff);
    EMFA_REQUIRE((long long)B * B < (1LL << 31), EMFA_E_INVALID, "fmm: %d boxes, too many box pairs", B);
    std::vector<int> dense((size_t)noff, -1);
    std::vector<long long> codes;
    C.near_ptr.assign(B + 1, 0);
    C.far_ptr.assign(B + 1, 0);
    C.far_idx.reserve((size_t)B * B);
    C.far_code.reserve((size_t)B * B);
    for (int b = 0; b < B; ++b) {
        for (int f = 0; f < B; ++f) {
            const int dx = C.ix[3 * b] - C.ix[3 * f], dy = C.ix[3 * b + 1] - C.ix[3 * f + 1],
                      dz = C.ix[3 * b + 2] - C.ix[3 * f + 2];
            if (std::max(std::abs(dx), std::max(std::abs(dy), std::abs(dz))) <= 1) {
                C.near_idx.push_back(f);
                continue;
            }
            const long long code = ((dx + nb[0] - 1) * sy + (dy + nb[1] - 1)) * sz + (dz + nb[2] - 1);
            if (dense[code] < 0) {
                dense[code] = (int)codes.size();
                codes.push_back(code);
            }
            C.far_idx.push_back(f);
            C.far_code.push_back(dense[code]);
        }
        C.near_ptr[b + 1] = (int)C.near_idx.size();
        C.far_ptr[b + 1] = (int)C.far_idx.size();
    }
    // samples (FMM.cpp:47-50,322-347)
    {
        const float kf = (float)k;
        const float kd = 1.732051f * kf * delta;
        C.L = std::max(1, (int)(kd + 6 * cbrtf(kd)));
    }
    C.K = 2 * C.L * C.L;
    gauss_legendre(C.L, C.xt, C.wt);
    gauss_legendre(2 * C.L, C.xp, C.wp);
    C.kh.resize(3 * (size_t)C.K);
    C.W.resize(C.K);
    C.theta.resize(C.K);
    for (int p = 0; p < 2 * C.L; ++p)
        for (int t = 0; t < C.L; ++t) {
            const int q = p * C.L + t;
            const double th = kPi / 2 * (C.xt[t] + 1), ph = kPi * (C.xp[p] + 1);
            C.kh[3 * q] = std::sin(th) * std::cos(ph);
            C.kh[3 * q + 1] = std::sin(th) * std::sin(ph);
            C.kh[3 * q + 2] = std::cos(th);
            C.W[q] = C.wp[p] * C.wt[t];
            C.theta[q] = th;
        }
    ML_TRY(ml_alloc(M, &C.up, (size_t)B * C.K * M->nc));
    ML_TRY(ml_alloc(M, &C.down, (size_t)B * C.K * M->nc));
    ML_TRY(ml_upload(M, &C.far_ptr_d, C.far_ptr));
    ML_TRY(ml_upload(M, &C.far_idx_d, C.far_idx.empty() ? std::vector<int>{0} : C.far_idx));
    ML_TRY(ml_upload(M, &C.far_code_d, C.far_code.empty() ? std::vector<int>{0} : C.far_code));
    ML_TRY(ml_upload(M, &C.parent_d, C.parent));
    ML_TRY(ml_upload(M, &C.oct_d, C.octant));
    ML_TRY(ml_upload(M, &C.ix_d, C.ix));
    ML_TRY(ml_upload(M, &C.near_ptr_d, C.near_ptr));
    ML_TRY(ml_upload(M, &C.near_idx_d, C.near_idx));
    // translation per grid offset in use, D = receiving minus source centre
    C.offsets_used = (int)codes.size();
    std::vector<cf> tr(std::max<size_t>(1, codes.size()) * C.K, cf{0, 0});
    {
        const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t)
            pool.emplace_back([&, t] {
                for (size_t i = t; i < codes.size(); i += nth) {
                    const long long code = codes[i];
                    const long long ox = code / (sy * sz) - (nb[0] - 1), oy = (code / sz) % sy - (nb[1] - 1),
                                    oz = code % sz - (nb[2] - 1);
                    const double D[3] = {C.a * (double)ox, C.a * (double)oy, C.a * (double)oz};
                    translation(k, D, C.kh, C.L, &tr[i * C.K]);
                }
            });
        for (auto &th : pool) th.join();
    }
    ML_TRY(ml_upload(M, &C.trans, tr));
    M->mid = rwg_mid(mesh);
    return ml_leaf_setup(ctx, M, mesh, prm, order, out);
}
#undef ML_TRY

static emfa_status ml_apply(emfa_ctx *ctx, emfa_mlfma *M, const double2 *x, double2 *y) {
    hipStream_t s = ctx->stream;
    // loads in flight per thread: aggregation and translation take four
    // unknowns / far boxes per loop step, the upward pass two children
    // (EMFA_MLFMA_UNROLL=1: one per step; the sums are the same either way)
    const char *unroll_env = getenv("EMFA_MLFMA_UNROLL");
    const int unroll = unroll_env && atoi(unroll_env) == 1 ? 1 : 4;
    const int n = M->n, lvl = M->levels;
    MlLevel &F = M->lv[lvl];
    k_ml_gather<<<(n + 255) / 256, 256, 0, s>>>(x, M->perm, n, M->xb);
    {
        ProfScope ps(ctx, "mlfma_near", (double)M->nnz * 12);
        k_ml_near<<<(n + 3) / 4, 256, 0, s>>>(M->row_ptr, M->col, M->val, M->xb, n, M->yn);
    }
    {
        ProfScope ps(ctx, "mlfma_agg", 0);
        const unsigned g = (unsigned)(((long long)F.nbox * F.K + 255) / 256);
        if (M->nc == 4) {
            if (unroll == 4)
                k_ml_agg<4, 4><<<g, 256, 0, s>>>(F.ubeg_d, F.uend_d, M->rad, M->xb, F.K, F.nbox, F.up);
            else
                k_ml_agg<4, 1><<<g, 256, 0, s>>>(F.ubeg_d, F.uend_d, M->rad, M->xb, F.K, F.nbox, F.up);
        } else {
            if (unroll == 4)
                k_ml_agg<3, 4><<<g, 256, 0, s>>>(F.ubeg_d, F.uend_d, M->rad, M->xb, F.K, F.nbox, F.up);
            else
                k_ml_agg<3, 1><<<g, 256, 0, s>>>(F.ubeg_d, F.uend_d, M->rad, M->xb, F.K, F.nbox, F.up);
        }
    }
    for (int l = lvl - 1; l >= 2; --l) {
        MlLevel &C = M->lv[l], &N = M->lv[l + 1];
        ProfScope ps(ctx, "mlfma_up", 0);
        const unsigned g = (unsigned)(((long long)C.nbox * C.K + 255) / 256);
#define ML_UP(NC, U)                                                                                               \
    k_ml_up<NC, U><<<g, 256, 0, s>>>(C.ch_ptr_d, C.ch_idx_d, N.oct_d, C.ip_col, C.ip_w, C.ph, C.K, N.K, C.nbox, N.up, \
                                     C.up)
        if (M->nc == 4) {
            if (unroll == 4)
                ML_UP(4, 2);
            else
                ML_UP(4, 1);
        } else {
            if (unroll == 4)
                ML_UP(3, 2);
            else
                ML_UP(3, 1);
        }
#undef ML_UP
    }
    // sibling-grouped translation measured slower than one thread per (box, sample): opt-in
    const char *sib_env = getenv("EMFA_MLFMA_SIBLING");
    const bool sib = sib_env && atoi(sib_env) != 0 && M->nc == 3;
    // far-list entries per step of the translation loop (1 or 4).  Four: C5's
    // translations 47.8 -> 32.4 ms per solve, the IP-SIE cube's 358 -> 241 ms;
    // eight measured slower than four.
    const int down_unroll = unroll;
    for (int l = 2; l <= lvl; ++l) {
        MlLevel &C = M->lv[l];
        MlLevel *P = l > 2 ? &M->lv[l - 1] : nullptr;
        // algorithmic bytes: the level's radiation patterns read and receive
        // patterns written once (nc complex FP32 per box and sample), its
        // translation table (one row of K per box-grid offset in use) and
        // the parent's receive patterns with the anterpolation weights
        const double pb = 8.0 * M->nc;
        const double ws = (double)C.nbox * C.K * pb * 2 + (double)C.offsets_used * C.K * 8.0 +
                          (P ? (double)P->nbox * P->K * pb + (double)C.K * 9 * 12.0 : 0.0);
        ProfScope ps(ctx, "mlfma_down", ws);
        if (P && sib) {
            k_ml_down_sib<<<(unsigned)(((long long)P->nbox * C.K + 255) / 256), 256, 0, s>>>(
                P->near_ptr_d, P->near_idx_d, P->ch_ptr_d, P->ch_idx_d, P->nbox, C.ix_d, C.oct_d, C.trans, C.up,
                C.K, C.an_ptr, C.an_kp, C.an_w, P->ph, P->K, P->down, C.down);
            continue;
        }
        const unsigned g = (unsigned)(((long long)C.nbox * C.K + 255) / 256);
#define ML_DOWN(NC, U)                                                                                            \
    k_ml_down<NC, U><<<g, 256, 0, s>>>(C.far_ptr_d, C.far_idx_d, C.far_code_d, C.trans, C.up, C.K, C.parent_d,     \
                                       C.oct_d, C.an_ptr, C.an_kp, C.an_w, P ? P->ph : nullptr, P ? P->K : 0,      \
                                       P ? P->down : nullptr, C.nbox, C.down)
        if (M->nc == 4) {
            if (down_unroll == 4)
                ML_DOWN(4, 4);
            else
                ML_DOWN(4, 1);
        } else {
            if (down_unroll == 4)
                ML_DOWN(3, 4);
            else
                ML_DOWN(3, 1);
        }
#undef ML_DOWN
    }
    {
        ProfScope ps(ctx, "mlfma_out", 0);
        double2 *yb = M->sai_val ? M->ybox : nullptr;
        if (M->nc == 4)
            k_ml_out<4><<<(n + 3) / 4, 256, 0, s>>>(M->rec, F.down, M->leaf_of, M->perm, M->yn, n, F.K, y, yb);
        else
            k_ml_out<3><<<(n + 3) / 4, 256, 0, s>>>(M->rec, F.down, M->leaf_of, M->perm, M->yn, n, F.K, y, yb);
    }
    if (M->sai_val) {
        ProfScope ps(ctx, "mlfma_precond", (double)M->sai_nnz * 12);
        k_ml_sai<<<(n + 3) / 4, 256, 0, s>>>(M->sai_ptr, M->sai_col, M->sai_val, M->ybox, M->perm, n, y);
    }
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

// out = P b (RWG order in and out) for a preconditioned operator
static emfa_status ml_sai_rhs(emfa_ctx *ctx, emfa_mlfma *M, const double2 *b, double2 *out) {
    const int n = M->n;
    k_ml_gather64<<<(n + 255) / 256, 256, 0, ctx->stream>>>(b, M->perm, n, M->ybox);
    k_ml_sai<<<(n + 3) / 4, 256, 0, ctx->stream>>>(M->sai_ptr, M->sai_col, M->sai_val, M->ybox, M->perm, n, out);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

// Householder QR of A (m x k, column-major, m >= k), then for each own row j in
// [own, own + nu): x = R^-1 (Q^H e_j)[0:k], the least-squares solution of
// A x = e_j, written to out (k x nu, column-major).  A zero column of R
// gives zero coefficients for that column.
static void sai_box(std::vector<std::complex<double>> &A, int m, int k, int own, int nu,
                    std::vector<std::complex<double>> &out) {
    using cd = std::complex<double>;
    std::vector<cd> v((size_t)m * k, cd(0, 0)), rdiag(k);
    std::vector<double> vn(k, 0.0);
    for (int j = 0; j < k; ++j) {
        cd *a = &A[(size_t)j * m];
        double nx = 0;
        for (int i = j; i < m; ++i) nx += std::norm(a[i]);
        nx = std::sqrt(nx);
        if (nx == 0) {
            rdiag[j] = 0;
            continue;
        }
        const cd ph = std::abs(a[j]) > 0 ? a[j] / std::abs(a[j]) : cd(1, 0);
        const cd alpha = -ph * nx;
        cd *vj = &v[(size_t)j * m];
        for (int i = j; i < m; ++i) vj[i] = a[i];
        vj[j] -= alpha;
        double vv = 0;
        for (int i = j; i < m; ++i) vv += std::norm(vj[i]);
        vn[j] = vv;
        // apply H = I - 2 v v^H / (v^H v) to the remaining columns
        for (int c = j + 1; c < k; ++c) {
            cd *ac = &A[(size_t)c * m];
            cd d(0, 0);
            for (int i = j; i < m; ++i) d += std::conj(vj[i]) * ac[i];
            d *= 2.0 / vv;
            for (int i = j; i < m; ++i) ac[i] -= d * vj[i];
        }
        rdiag[j] = alpha;
    }
    std::vector<cd> e(m);
    out.assign((size_t)k * nu, cd(0, 0));
    for (int u = 0; u < nu; ++u) {
        std::fill(e.begin(), e.end(), cd(0, 0));
        e[own + u] = cd(1, 0);
        for (int j = 0; j < k; ++j) {  // Q^H e = H_k ... H_1 e
            if (vn[j] == 0) continue;
            const cd *vj = &v[(size_t)j * m];
            cd d(0, 0);
            for (int i = j; i < m; ++i) d += std::conj(vj[i]) * e[i];
            d *= 2.0 / vn[j];
            for (int i = j; i < m; ++i) e[i] -= d * vj[i];
        }
        cd *x = &out[(size_t)u * k];
        for (int j = k - 1; j >= 0; --j) {  // R x = (Q^H e)[0:k]
            cd t = e[j];
            for (int c = j + 1; c < k; ++c) t -= A[(size_t)c * m + j] * x[c];
            x[j] = rdiag[j] == cd(0, 0) ? cd(0, 0) : t / rdiag[j];
        }
    }
}

// MLFMA::calculatePreconditioner (MLFMA.cpp:714-741,1294-1391; FastWT.cpp:
// 815-914 is the same): thresholds in leaf edges; distances of the unknowns'
// centres (RWG or half-RWG edge midpoints) from the leaf-box centre.
static emfa_status ml_precondition(emfa_ctx *ctx, emfa_mlfma *M, const emfa_mesh *mesh, double row_threshold,
                                   double col_threshold, long long *entries) {
    using cd = std::complex<double>;
    const int n = M->n;
    MlLevel &F = M->lv[M->levels];
    const double rt = row_threshold * F.a, ct = col_threshold * F.a;
    hipStream_t s = ctx->stream;
    std::vector<long long> rp(n + 1);
    std::vector<int> col((size_t)M->nnz), perm(n);
    std::vector<cf> val((size_t)M->nnz);
    EMFA_HIP_CHECK(hipMemcpyAsync(rp.data(), M->row_ptr, rp.size() * sizeof(long long), hipMemcpyDeviceToHost, s));
    EMFA_HIP_CHECK(hipMemcpyAsync(col.data(), M->col, col.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    EMFA_HIP_CHECK(hipMemcpyAsync(val.data(), M->val, val.size() * sizeof(cf), hipMemcpyDeviceToHost, s));
    EMFA_HIP_CHECK(hipMemcpyAsync(perm.data(), M->perm, perm.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    EMFA_HIP_CHECK(hipStreamSynchronize(s));
    (void)mesh;
    EMFA_REQUIRE(M->mid.size() == 3 * (size_t)n, EMFA_E_INVALID, "mlfma preconditioner: operator without centres");
    std::vector<double> mid(3 * (size_t)n);  // box order
    for (int i = 0; i < n; ++i) {
        const int u = perm[i];
        for (int c = 0; c < 3; ++c) mid[3 * i + c] = M->mid[3 * (size_t)u + c];
    }
    struct Entry {
        int row, col;
        cf v;
    };
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<Entry>> part(nth);
    std::vector<int> bad(nth, -1);
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t)
            pool.emplace_back([&, t] {
                std::vector<int> pos(n, -1), rows, cols;
                std::vector<cd> A, res;
                for (int b = t; b < F.nbox; b += nth) {
                    rows.clear();
                    cols.clear();
                    int own = -1;
                    const double *c0 = &F.cen[3 * b];
                    for (int e = F.near_ptr[b]; e < F.near_ptr[b + 1]; ++e) {
                        const int nb = F.near_idx[e];
                        if (nb == b) own = (int)rows.size();
                        for (int i = F.ubeg[nb]; i < F.uend[nb]; ++i) {
                            const double dx = mid[3 * i] - c0[0], dy = mid[3 * i + 1] - c0[1], dz = mid[3 * i + 2] - c0[2];
                            const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
                            if (d <= rt) rows.push_back(i);
                            if (d <= ct) cols.push_back(i);
                        }
                    }
                    const int m = (int)rows.size(), k = (int)cols.size(), nu = F.uend[b] - F.ubeg[b];
                    // the box's own unknowns are the rows from `own` on
                    if (own < 0 || k > m || own + nu > m || rows[own] != F.ubeg[b] || rows[own + nu - 1] != F.uend[b] - 1) {
                        bad[t] = b;
                        return;
                    }
                    A.assign((size_t)m * k, cd(0, 0));
                    for (int c = 0; c < k; ++c) pos[cols[c]] = c;
                    for (int f = 0; f < m; ++f)
                        for (long long q = rp[rows[f]]; q < rp[rows[f] + 1]; ++q) {
                            const int c = pos[col[q]];
                            if (c >= 0) A[(size_t)c * m + f] = cd(val[q].re, val[q].im);
                        }
                    for (int c = 0; c < k; ++c) pos[cols[c]] = -1;
                    sai_box(A, m, k, own, nu, res);
                    for (int u = 0; u < nu; ++u)
                        for (int c = 0; c < k; ++c) {
                            const cd x = res[(size_t)u * k + c];
                            part[t].push_back({cols[c], F.ubeg[b] + u, cf{(float)x.real(), (float)x.imag()}});
                        }
                }
            });
        for (auto &th : pool) th.join();
    }
    for (int t = 0; t < nth; ++t)
        EMFA_REQUIRE(bad[t] < 0, EMFA_E_INVALID,
                     "mlfma preconditioner: box %d: thresholds (row %.3g, col %.3g leaf edges) must keep the box's "
                     "own unknowns among the rows and no more columns than rows",
                     bad[t], row_threshold, col_threshold);
    // CSR over rows, columns ascending
    std::vector<long long> ptr(n + 1, 0);
    for (auto &pp : part)
        for (auto &e : pp) ++ptr[e.row + 1];
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    const long long nnz = ptr[n];
    std::vector<int> pc((size_t)nnz);
    std::vector<cf> pv((size_t)nnz);
    {
        std::vector<long long> fill(ptr.begin(), ptr.end() - 1);
        for (auto &pp : part)
            for (auto &e : pp) {
                pc[(size_t)fill[e.row]] = e.col;
                pv[(size_t)fill[e.row]] = e.v;
                ++fill[e.row];
            }
        for (int i = 0; i < n; ++i) {
            std::vector<std::pair<int, cf>> r;
            for (long long q = ptr[i]; q < ptr[i + 1]; ++q) r.push_back({pc[q], pv[q]});
            std::sort(r.begin(), r.end(), [](const std::pair<int, cf> &a, const std::pair<int, cf> &b) {
                return a.first < b.first;
            });
            for (long long q = ptr[i]; q < ptr[i + 1]; ++q) {
                pc[q] = r[q - ptr[i]].first;
                pv[q] = r[q - ptr[i]].second;
            }
        }
    }
    emfa_status st;
    if ((st = ml_upload(M, &M->sai_ptr, ptr))) return st;
    if ((st = ml_upload(M, &M->sai_col, pc))) return st;
    if ((st = ml_upload(M, &M->sai_val, pv))) return st;
    if ((st = ml_alloc(M, &M->ybox, (size_t)n))) return st;
    M->sai_nnz = nnz;
    if (entries) *entries = nnz;
    return EMFA_OK;
}

}  // namespace emfa

using namespace emfa;

extern "C" {

emfa_status emfa_mlfma_create(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *p, emfa_mlfma **out) {
    EMFA_REQUIRE(ctx && mesh && p && out, EMFA_E_INVALID, "emfa_mlfma_create: null argument");
    *out = nullptr;
    EMFA_REQUIRE(p->kind == EMFA_EFIE || p->kind == EMFA_CFIE, EMFA_E_INVALID, "emfa_mlfma_create: kind");
    EMFA_REQUIRE(p->k > 0 && p->box_threshold > 0, EMFA_E_INVALID, "emfa_mlfma_create: k and threshold must be > 0");
    EMFA_REQUIRE(!mesh->rwg_host.empty(), EMFA_E_INVALID, "emfa_mlfma_create: mesh without host copy");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return ml_build(ctx, mesh, p, out);
}

emfa_status emfa_mlfma_destroy(emfa_mlfma *m) {
    if (!m) return EMFA_OK;
    ml_free(m);
    delete m;
    return EMFA_OK;
}

emfa_status emfa_mlfma_get_info(const emfa_mlfma *m, emfa_mlfma_info *info) {
    EMFA_REQUIRE(m && info, EMFA_E_INVALID, "emfa_mlfma_get_info: null argument");
    std::memset(info, 0, sizeof(*info));
    info->levels = m->single ? 1 : m->levels;
    info->near_nnz = m->nnz;
    info->bytes = m->bytes;
    for (int c = 0; c < 3; ++c) info->grid[c] = m->grid[c];
    for (int l = 0; l <= info->levels && l < 16; ++l) {
        const MlLevel &v = m->lv[m->single && l == 1 ? 2 : m->single ? 0 : l];
        if (m->single && l == 0) continue;
        info->boxes[l] = v.nbox;
        info->multipoles[l] = v.L;
        info->box_edge[l] = v.a;
        info->far_pairs[l] = (long long)v.far_idx.size();
    }
    return EMFA_OK;
}

emfa_status emfa_mlfma_boxes(const emfa_mlfma *m, int level, int *nnear, int *nfar, int *nunk) {
    EMFA_REQUIRE(m, EMFA_E_INVALID, "emfa_mlfma_boxes: null operator");
    const int top = m->single ? 1 : m->levels;
    EMFA_REQUIRE(level >= 1 && level <= top, EMFA_E_INVALID, "emfa_mlfma_boxes: level %d outside 1..%d", level, top);
    const MlLevel &v = m->lv[m->single ? 2 : level];
    for (int b = 0; b < v.nbox; ++b) {
        if (nnear) nnear[b] = v.near_ptr[b + 1] - v.near_ptr[b];
        if (nfar) nfar[b] = v.far_ptr.empty() ? 0 : v.far_ptr[b + 1] - v.far_ptr[b];
        if (nunk) nunk[b] = v.ubeg.empty() ? 0 : v.uend[b] - v.ubeg[b];
    }
    return EMFA_OK;
}

emfa_status emfa_fmm_create(emfa_ctx *ctx, const emfa_mesh *mesh, const emfa_mlfma_params *p, emfa_mlfma **out) {
    EMFA_REQUIRE(ctx && mesh && p && out, EMFA_E_INVALID, "emfa_fmm_create: null argument");
    *out = nullptr;
    EMFA_REQUIRE(p->kind == EMFA_EFIE || p->kind == EMFA_CFIE, EMFA_E_INVALID, "emfa_fmm_create: kind");
    EMFA_REQUIRE(p->k > 0 && p->box_threshold > 0, EMFA_E_INVALID, "emfa_fmm_create: k and box length must be > 0");
    EMFA_REQUIRE(!mesh->rwg_host.empty(), EMFA_E_INVALID, "emfa_fmm_create: mesh without host copy");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return fmm_build(ctx, mesh, p, out);
}

emfa_status emfa_fastwt_create(emfa_ctx *ctx, const emfa_mesh *mesh, int n, const int *hrwg, const double *length,
                               int n_nc, const int *line_ptr, const int *line_tri, const double *line_seg,
                               const emfa_mlfma_params *p, double beta, emfa_mlfma **out) {
    EMFA_REQUIRE(ctx && mesh && hrwg && length && p && out, EMFA_E_INVALID, "emfa_fastwt_create: null argument");
    *out = nullptr;
    EMFA_REQUIRE(n > 0 && n_nc >= 0, EMFA_E_INVALID, "emfa_fastwt_create: %d half-RWGs, %d nonconformal", n, n_nc);
    EMFA_REQUIRE(p->kind == EMFA_CFIE && p->alpha >= 0 && p->alpha <= 1, EMFA_E_INVALID,
                 "emfa_fastwt_create: the IP-SIE operator is the combined-field one (kind CFIE, alpha in [0, 1])");
    EMFA_REQUIRE(p->k > 0 && p->box_threshold > 0 && beta >= 0, EMFA_E_INVALID,
                 "emfa_fastwt_create: k, threshold must be > 0 and beta >= 0");
    EMFA_REQUIRE(!mesh->tri_host.empty(), EMFA_E_INVALID, "emfa_fastwt_create: mesh without host copy");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return wt_build(ctx, mesh, n, hrwg, length, n_nc, line_ptr, line_tri, line_seg, p, beta, out);
}

emfa_status emfa_fastwt_rhs(emfa_ctx *ctx, const emfa_mlfma *m, const emfa_mesh *mesh, emfa_mat *B, int nrhs,
                            const double *inc) {
    EMFA_REQUIRE(ctx && m && mesh && B && inc, EMFA_E_INVALID, "emfa_fastwt_rhs: null argument");
    EMFA_REQUIRE(m->nc == 4 && mesh->nt == m->nt, EMFA_E_INVALID,
                 "emfa_fastwt_rhs: needs an IP-SIE operator and the mesh it was built on");
    EMFA_REQUIRE(B->rows == m->n && B->cols == nrhs && nrhs > 0 && B->prec == EMFA_F64, EMFA_E_INVALID,
                 "emfa_fastwt_rhs: B must be F64 n x nrhs");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    void *d = nullptr;
    emfa_status st = ctx_scratch(ctx, 9 * sizeof(double) * (size_t)nrhs, &d);
    if (st) return st;
    EMFA_HIP_CHECK(hipMemcpyAsync(d, inc, 9 * sizeof(double) * (size_t)nrhs, hipMemcpyHostToDevice, ctx->stream));
    const long long tot = (long long)m->n * nrhs;
    k_wt_rhs<<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(mesh->tg64, mesh->nt, m->wt_tri, m->wt_fv,
                                                                     m->wt_len64, m->n, (const double *)d, nrhs,
                                                                     m->k, m->alpha, (double2 *)B->d);
    EMFA_HIP_CHECK(hipGetLastError());
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return EMFA_OK;
}

emfa_status emfa_fastwt_rcs(emfa_ctx *ctx, const emfa_mlfma *m, const emfa_mesh *mesh, const emfa_mat *X, int col,
                            int ndir, const double *dirs, double *rcs_db) {
    EMFA_REQUIRE(ctx && m && mesh && X && dirs && rcs_db, EMFA_E_INVALID, "emfa_fastwt_rcs: null argument");
    EMFA_REQUIRE(m->nc == 4 && mesh->nt == m->nt, EMFA_E_INVALID,
                 "emfa_fastwt_rcs: needs an IP-SIE operator and the mesh it was built on");
    EMFA_REQUIRE(X->rows == m->n && X->prec == EMFA_F64 && col >= 0 && col < X->cols && ndir > 0, EMFA_E_INVALID,
                 "emfa_fastwt_rcs: X must be F64 n x m, 0 <= col < m");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    void *d = nullptr;
    emfa_status st = ctx_scratch(ctx, 4 * sizeof(double) * (size_t)ndir, &d);
    if (st) return st;
    double *dd = (double *)d, *od = dd + 3 * (size_t)ndir;
    EMFA_HIP_CHECK(hipMemcpyAsync(dd, dirs, 3 * sizeof(double) * (size_t)ndir, hipMemcpyHostToDevice, ctx->stream));
    k_wt_rcs<<<ndir, 256, 0, ctx->stream>>>(mesh->tg64, mesh->nt, m->wt_tri, m->wt_fv, m->wt_len64, m->n,
                                            (const double2 *)X->d, X->cols, col, dd, m->k, od);
    EMFA_HIP_CHECK(hipGetLastError());
    EMFA_HIP_CHECK(hipMemcpyAsync(rcs_db, od, sizeof(double) * (size_t)ndir, hipMemcpyDeviceToHost, ctx->stream));
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return EMFA_OK;
}

emfa_status emfa_mlfma_matvec(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *X, emfa_mat *Y) {
    EMFA_REQUIRE(ctx && m && X && Y, EMFA_E_INVALID, "emfa_mlfma_matvec: null argument");
    EMFA_REQUIRE(X->rows == m->n && Y->rows == m->n && X->cols == Y->cols && X->prec == EMFA_F64 &&
                     Y->prec == EMFA_F64 && X != Y,
                 EMFA_E_INVALID, "emfa_mlfma_matvec: X, Y must be distinct F64 n x m");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    const int n = m->n, nc = X->cols;
    if (nc == 1) {
        emfa_status st = ml_apply(ctx, m, (const double2 *)X->d, (double2 *)Y->d);
        if (st) return st;
        EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return EMFA_OK;
    }
    void *scr = nullptr;
    emfa_status st = ctx_scratch(ctx, 2 * (size_t)n * sizeof(double2), &scr);
    if (st) return st;
    double2 *xin = (double2 *)scr, *yout = xin + n;
    for (int c = 0; c < nc; ++c) {
        EMFA_HIP_CHECK(hipMemcpy2DAsync(xin, sizeof(double2), (const double2 *)X->d + c, nc * sizeof(double2),
                                        sizeof(double2), n, hipMemcpyDeviceToDevice, ctx->stream));
        if ((st = ml_apply(ctx, m, xin, yout))) return st;
        EMFA_HIP_CHECK(hipMemcpy2DAsync((double2 *)Y->d + c, nc * sizeof(double2), yout, sizeof(double2),
                                        sizeof(double2), n, hipMemcpyDeviceToDevice, ctx->stream));
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return EMFA_OK;
}

// With a preconditioner the system solved is P Z X = P B (MLFMA.cpp:406-407):
// returns P B in a new matrix (null without one)
static emfa_status ml_rhs(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *B, emfa_mat **Bp) {
    *Bp = nullptr;
    if (!m->sai_val) return EMFA_OK;
    EMFA_REQUIRE(B->rows == m->n && B->prec == EMFA_F64, EMFA_E_INVALID, "mlfma: B must be F64 n x m");
    const int n = m->n, nc = B->cols;
    emfa_status st = emfa_mat_create(ctx, n, nc, EMFA_F64, Bp);
    if (st) return st;
    void *scr = nullptr;
    if ((st = ctx_scratch(ctx, 2 * (size_t)n * sizeof(double2), &scr))) {
        emfa_mat_destroy(*Bp);
        *Bp = nullptr;
        return st;
    }
    double2 *xin = (double2 *)scr, *yout = xin + n;
    for (int c = 0; c < nc && !st; ++c) {
        if (hipMemcpy2DAsync(xin, sizeof(double2), (const double2 *)B->d + c, nc * sizeof(double2), sizeof(double2),
                             n, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            st = EMFA_E_HIP;
        if (!st) st = ml_sai_rhs(ctx, m, xin, yout);
        if (!st && hipMemcpy2DAsync((double2 *)(*Bp)->d + c, nc * sizeof(double2), yout, sizeof(double2),
                                    sizeof(double2), n, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            st = EMFA_E_HIP;
    }
    if (st) {
        (void)hipGetLastError();
        emfa_mat_destroy(*Bp);
        *Bp = nullptr;
        set_error("mlfma: preconditioning the right-hand sides failed");
    }
    return st;
}

emfa_status emfa_mlfma_gmres(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *B, emfa_mat *X, int restart, double tol,
                             int max_restart, emfa_gmres_report *rep) {
    EMFA_REQUIRE(ctx && m && B && X, EMFA_E_INVALID, "emfa_mlfma_gmres: null argument");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    auto op = [&](const double2 *in, double2 *out) -> emfa_status { return ml_apply(ctx, m, in, out); };
    emfa_mat *Bp = nullptr;
    emfa_status st = ml_rhs(ctx, m, B, &Bp);
    if (st) return st;
    st = gmres_impl(ctx, m->n, op, Bp ? Bp : B, X, restart, tol, max_restart, rep);
    if (Bp) emfa_mat_destroy(Bp);
    return st;
}

emfa_status emfa_mlfma_bicgstab(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mat *B, emfa_mat *X, double tol,
                                int max_iter, int variant, emfa_bicgstab_report *rep) {
    EMFA_REQUIRE(ctx && m && B && X, EMFA_E_INVALID, "emfa_mlfma_bicgstab: null argument");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    auto op = [&](const double2 *in, double2 *out) -> emfa_status { return ml_apply(ctx, m, in, out); };
    emfa_mat *Bp = nullptr;
    emfa_status st = ml_rhs(ctx, m, B, &Bp);
    if (st) return st;
    st = bicgstab_impl(ctx, m->n, op, Bp ? Bp : B, X, tol, max_iter, variant, rep);
    if (Bp) emfa_mat_destroy(Bp);
    return st;
}

emfa_status emfa_mlfma_precondition(emfa_ctx *ctx, emfa_mlfma *m, const emfa_mesh *mesh, double row_threshold,
                                    double col_threshold, long long *entries) {
    EMFA_REQUIRE(ctx && m && mesh, EMFA_E_INVALID, "emfa_mlfma_precondition: null argument");
    EMFA_REQUIRE(!m->single, EMFA_E_INVALID, "emfa_mlfma_precondition: the single-level operator has none");
    EMFA_REQUIRE(mesh->nt == m->nt && (m->nc == 4 || mesh->n == m->n), EMFA_E_INVALID,
                 "emfa_mlfma_precondition: not the operator's mesh");
    EMFA_REQUIRE(row_threshold > 0 && col_threshold > 0, EMFA_E_INVALID,
                 "emfa_mlfma_precondition: thresholds must be > 0");
    EMFA_REQUIRE(!m->sai_val, EMFA_E_INVALID, "emfa_mlfma_precondition: the operator is preconditioned already");
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    return ml_precondition(ctx, m, mesh, row_threshold, col_threshold, entries);
}

}  // extern "C"
