// Device helpers shared by the dense fill and the MLFMA near field:
// triangle quadrature rules, small vector algebra, the analytic potentials
// of a triangle and the FP64 self terms.
#pragma once
#include <cmath>

#include "internal.h"

namespace emfa {

// Triangle rules, EMComponent/Quadrature.h:20-27 (3 points) and :42-72 (7 points):
// barycentrics {l1, l2, l3} and weight (normalised to sum 1).
static __constant__ double c_rule3[3][4] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 3.0},
                                     {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0, 1.0 / 3.0},
                                     {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 3.0}};
static __constant__ double c_rule7[7][4] = {
    {0.333333333333333, 0.333333333333333, 1.0 - 0.333333333333333 - 0.333333333333333, 0.112500000000000 * 2},
    {0.797426985353087, 0.101286507323456, 1.0 - 0.797426985353087 - 0.101286507323456, 0.062969590272413 * 2},
    {0.101286507323456, 0.797426985353087, 1.0 - 0.101286507323456 - 0.797426985353087, 0.062969590272413 * 2},
    {0.101286507323456, 0.101286507323456, 1.0 - 0.101286507323456 - 0.101286507323456, 0.062969590272413 * 2},
    {0.470142064105115, 0.470142064105115, 1.0 - 0.470142064105115 - 0.470142064105115, 0.066197076394253 * 2},
    {0.470142064105115, 0.059715871789770, 1.0 - 0.470142064105115 - 0.059715871789770, 0.066197076394253 * 2},
    {0.059715871789770, 0.470142064105115, 1.0 - 0.059715871789770 - 0.470142064105115, 0.066197076394253 * 2}};

template <int Q>
__device__ __forceinline__ double rule(int p, int i) {
    return Q == 7 ? c_rule7[p][i] : c_rule3[p][i];
}

__device__ __forceinline__ void sincos_r(float x, float *s, float *c) { sincosf(x, s, c); }
__device__ __forceinline__ void sincos_r(double x, double *s, double *c) { sincos(x, s, c); }

template <typename R>
__device__ __forceinline__ R dot3(const R *a, const R *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 mk3(double x, double y, double z) { return {x, y, z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dotd(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ d3 crossd(d3 a, d3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double normd(d3 a) { return sqrt(dotd(a, a)); }
__device__ __forceinline__ d3 unitd(d3 a) { return a * (1.0 / normd(a)); }

// Analytic integrals over triangle t seen from in-plane point r (Wilton/Graglia
// edge formulas as used at EMSolver/EFIE.cpp:259-291):
//   Is1 = int (rho' - rho)/R dS' (in-plane vector), Is2 = int 1/R dS'
__device__ inline void potentials(const d3 *t, d3 r, d3 *Is1, double *Is2) {
    d3 ni = unitd(crossd(t[1] - t[0], t[2] - t[0]));
    d3 rou = r - ni * dotd(ni, r);
    d3 s1 = mk3(0, 0, 0);
    double s2 = 0;
    for (int i = 0; i < 3; ++i) {
        d3 a = t[i], b = t[(i + 1) % 3];
        d3 li = b - a;
        d3 ui = unitd(crossd(li, ni));
        d3 l1 = unitd(li);
        d3 rp_ = b - ni * dotd(ni, b);
        double p0 = fabs(dotd(a - r, ui));
        double d = dotd(r - a, ni);
        double lp = dotd(b - r, l1);
        double lm = dotd(a - r, l1);
        double rp = normd(r - b);
        double rm = normd(r - a);
        if (p0 < 1e-10) continue;
        d3 pi = unitd((rp_ - rou) - l1 * lp);
        double r02 = p0 * p0 + d * d;
        double lg = log((rp + lp) / (rm + lm));
        s1 = s1 + ui * (0.5 * (r02 * lg + lp * rp - lm * rm));
        double ad = fabs(d);
        s2 += dotd(pi, ui) * (p0 * lg - ad * (atan((p0 * lp) / (r02 + ad * rp)) - atan((p0 * lm) / (r02 + ad * rm))));
    }
    *Is1 = s1;
    *Is2 = s2;
}

template <int Q>
__device__ void self_term(int kind, double k, double alpha, double sig_re, double sig_im, int cfie_mode, const d3 *t,
                          double area, int a, int b, double *zr, double *zi, double *gr, double *gi) {
    d3 pts[Q];
    double w[Q];
    for (int p = 0; p < Q; ++p) {
        pts[p] = t[0] * rule<Q>(p, 0) + t[1] * rule<Q>(p, 1) + t[2] * rule<Q>(p, 2);
        w[p] = rule<Q>(p, 3);
    }
    d3 vf = t[a], vs = t[b];
    const double ik2 = 1.0 / (k * k);
    // Ze1: Taylor part (exp(-jkR)-1)/R ~ -jk - k^2 R/2 + j k^3 R^2/6 (EFIE.cpp:251-256)
    double e1r = 0, e1i = 0, e2r = 0, zmr = 0;
    for (int p = 0; p < Q; ++p) {
        d3 rf = pts[p] - vf;
        d3 rs_last = mk3(0, 0, 0);
        for (int q = 0; q < Q; ++q) {
            d3 rs = pts[q] - vs;
            rs_last = rs;
            double R = normd(pts[p] - pts[q]);
            double cr = -0.5 * k * k * R;
            double ci = (k / 6.0) * (k * k * R * R - 6.0);
            double f = w[p] * w[q] * (0.25 * dotd(rf, rs) - ik2);
            e1r += f * cr;
            e1i += f * ci;
        }
        if (kind == EMFA_CFIE && cfie_mode == 0) {
            // Miscellaneous.cpp:53-66: in-plane log terms with the last inner rho_s
            for (int i = 0; i < 3; ++i) {
                d3 va = t[i], vb = t[(i + 1) % 3];
                d3 Li = unitd(vb - va);
                double Lp = dotd(vb - pts[p], Li);
                double Lm = dotd(va - pts[p], Li);
                double Rp = normd(vb - pts[p]);
                double Rm = normd(va - pts[p]);
                d3 sj0 = va + Li * dotd(pts[p] - va, Li);
                double P0 = normd(sj0 - pts[p]);
                if (P0 > 1e-10) e2r += w[p] * (0.25 * dotd(rf, rs_last) - ik2) * P0 * log((Rp + Lp) / (Rm + Lm));
            }
        } else {
            d3 Is1;
            double Is2;
            potentials(t, pts[p], &Is1, &Is2);
            d3 vec = Is1 + (pts[p] - vs) * Is2;
            e2r += w[p] * (0.25 * dotd(rf, vec) - Is2 * ik2);
        }
        zmr += w[p] * dotd(rf, pts[p] - vs);
    }
    // coefE = j k Z0 / 4 pi applied to (Ze1 + Ze2 / area)
    double sr = e1r + e2r / area, si = e1i;
    double ce = k * kZ0 / (4.0 * kPi);
    double er = -ce * si, ei = ce * sr;
    if (kind == EMFA_EFIE) {
        *zr = er;
        *zi = ei;
        // Sigma term (EFIE.cpp:297-306): pi sigma / (S j k Z0) * sum w (r-vf).(r-vs),
        // times coefE later in the gather via the caller's scale: coefE * pi/(S j k Z0) = 1/(4 S)
        double acc = 0;
        for (int i = 0; i < Q; ++i) acc += w[i] * dotd(pts[i] - vf, pts[i] - vs);
        double f = acc / (4.0 * area);
        *gr = f * sig_re;
        *gi = f * sig_im;
    } else {
        // alpha * Ze + (1 - alpha) * Z0 * Zm / (8 area)
        *zr = alpha * er + (1.0 - alpha) * kZ0 * zmr / (8.0 * area);
        *zi = alpha * ei;
        *gr = 0;
        *gi = 0;
    }
}

}  // namespace emfa
