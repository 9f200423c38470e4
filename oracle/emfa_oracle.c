/*
 * emfa_oracle.c -- FP64 CPU restatement of EMFA's method-of-moments hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (emfa_amd/, include/)
 * links, loads or calls this file.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, as the checker and as the CPU baseline.
 *
 * Every routine cites the reference function it restates (paths relative to
 * the reference repository root).  The reference computes in complex<float>
 * (EMComponent/Custom.h:67-68); this restatement computes the same formulas
 * in double so that it can serve as the FP64 oracle.
 *
 * Conventions shared with the product:
 *   xyz  : nv x 3 doubles (vertex coordinates, metres)
 *   tri  : nt x 3 ints, 0-based vertex ids, in file order
 *   rwg  : n x 6 ints  {v1, v2, vxp, vxm, tp, tm}   (CommonEdge::SCE fields)
 *   Z    : n x n complex, ROW-major (row = test function), interleaved re/im
 *   B, X : n x nrhs complex, row-major
 *
 * Decisions (the parity contract):
 *   - Z0 = 120*pi (EMComponent/Mathdef.h:33), c0 = 299792458.
 *   - quadrature constants are the literal digits of EMComponent/Quadrature.h
 *     evaluated in double (1/3, 1/6 exact).
 *   - the EDM far-field shortcut of EFIE.cpp:647-651 is OFF (it never reaches
 *     the matrix the reference solves, EFIE.cpp:691).
 *   - singular extraction only for identical triangles (EFIE.cpp:655-689).
 *   - CFIE self term reproduces Miscellaneous.cpp:32-77 including the reuse
 *     of the last inner-loop rho_s ("parity" mode); mode 1 ("accurate")
 *     replaces the EFIE part by the EFIE-style analytic extraction.
 *   - the Sigma (surface impedance) self term is added without the RWG sign
 *     factors, exactly as EFIE.cpp:297-306 (and its pm/mp/mm copies) do.
 *   - surface-current corner magnitudes carry over between triangles as
 *     CFIE.cpp:740 lets them (open-boundary triangles).
 *
 * Pins (what checks each routine; tests/ and tests/golden/):
 *   orc_quad_rule       exactness on monomials of degree 2 / 5, weights sum 1
 *   orc_build_rwg       lists of the reference-compiled builder; brute force
 *   orc_incidence       reference-compiled V of EFIE/CFIE/MFIE (angles != 0)
 *   orc_fill_efie       reference-compiled Z (with and without Sigma, plate),
 *                       Duffy quadrature of the singular potentials, Mie
 *   orc_fill_cfie       reference-compiled Z of CFIE and of MFIE (alpha = 0)
 *   orc_rhs             reference-compiled V of EFIE, CFIE and MFIE
 *   orc_rcs             reference-compiled RCS cuts; Mie series
 *   orc_lu, orc_lu_solve  LAPACK getrf / getrs (scipy) pivots and factors
 *   orc_gmres           reaches the LU solution
 *   orc_bicgstab, orc_bicgstab_mf  reference-compiled iml::BICGSTAB (both
 *                       overloads) on the reference's own CFIE Z and V:
 *                       step counts and solutions; reach the LU solution
 *   orc_current         reference-compiled CFIE::calculateSurfaceCurrent on
 *                       the reference's own I (closed meshes and a plate)
 *   orc_potentials      adaptive Duffy quadrature, off-plane 2-D quadrature
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define ORC_PI 3.14159265358979323846264338327950288
static const double ORC_C0 = 299792458.0;
#define ORC_Z0 (120.0 * ORC_PI)

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double norm3(v3 a) { return sqrt(dot(a, a)); }
static inline v3 unit(v3 a) { return scl(a, 1.0 / norm3(a)); }

static inline v3 vtx(const double *xyz, int i) { return mk(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]); }

int orc_version(void) { return 1; }

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* Quadrature: EMComponent/Quadrature.h:20-27 (3-pt), :42-72 (7-pt);   */
/* Quadrature.cpp:13-18 (Gauss3Point), :44-53 (Gauss7Point).           */
/* Weights are normalised to sum 1 (no area factor).                   */
/* ------------------------------------------------------------------ */
static const double W3[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0};
static const double B3[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                                {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0},
                                {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}};
static const double W7[7] = {0.112500000000000 * 2, 0.062969590272413 * 2, 0.062969590272413 * 2,
                             0.062969590272413 * 2, 0.066197076394253 * 2, 0.066197076394253 * 2,
                             0.066197076394253 * 2};
static const double G7[7][2] = {{0.333333333333333, 0.333333333333333}, {0.797426985353087, 0.101286507323456},
                                {0.101286507323456, 0.797426985353087}, {0.101286507323456, 0.101286507323456},
                                {0.470142064105115, 0.470142064105115}, {0.470142064105115, 0.059715871789770},
                                {0.059715871789770, 0.470142064105115}};

/* points of rule q (3 or 7) on triangle (a,b,c); returns weights pointer */
static const double *quad_pts(int q, v3 a, v3 b, v3 c, v3 *p)
{
    if (q == 7) {
        for (int i = 0; i < 7; ++i) {
            double l1 = G7[i][0], l2 = G7[i][1], l3 = 1.0 - l1 - l2;
            p[i] = add(add(scl(a, l1), scl(b, l2)), scl(c, l3));
        }
        return W7;
    }
    for (int i = 0; i < 3; ++i)
        p[i] = add(add(scl(a, B3[i][0]), scl(b, B3[i][1])), scl(c, B3[i][2]));
    return W3;
}

int orc_quad_rule(int q, double *bary /* q x 3 */, double *w /* q */)
{
    if (q != 3 && q != 7) return -1;
    for (int i = 0; i < q; ++i) {
        if (q == 7) {
            bary[3 * i] = G7[i][0];
            bary[3 * i + 1] = G7[i][1];
            bary[3 * i + 2] = 1.0 - G7[i][0] - G7[i][1];
            w[i] = W7[i];
        } else {
            bary[3 * i] = B3[i][0];
            bary[3 * i + 1] = B3[i][1];
            bary[3 * i + 2] = B3[i][2];
            w[i] = W3[i];
        }
    }
    return q;
}

/* Heron area: EMComponent/VectorR3.h:147-154 */
static double tri_area(v3 a, v3 b, v3 c)
{
    double la = norm3(sub(a, b)), lb = norm3(sub(b, c)), lc = norm3(sub(c, a));
    double s = 0.5 * (la + lb + lc);
    return sqrt(s * (s - la) * (s - lb) * (s - lc));
}

/* Unit normal: EMComponent/Triangle.cpp:47-61 (both branches give AB x BC's direction) */
static v3 tri_normal(v3 a, v3 b, v3 c)
{
    v3 ab = sub(b, a), bc = sub(c, b), ca = sub(a, c);
    v3 n = (dot(ab, bc) < dot(bc, ca)) ? cross(ab, bc) : cross(bc, ca);
    return unit(n);
}

static void tri_vtx(const double *xyz, const int *tri, int t, v3 *v)
{
    v[0] = vtx(xyz, tri[3 * t]);
    v[1] = vtx(xyz, tri[3 * t + 1]);
    v[2] = vtx(xyz, tri[3 * t + 2]);
}

/* ------------------------------------------------------------------ */
/* RWG list: EMComponent/CommonEdge.cpp:74-106 (buildCommonEdge) and   */
/* :704-743 (updateCommonEdge).  O(T*E) linear scan exactly as the     */
/* reference: for triangle (A,B,C) the candidate edges are (A,B),(B,C),*/
/* (C,A); an unmatched stored edge (v1,v2) matches when it appears     */
/* reversed; new edges are appended in that order with vxp = opposite  */
/* vertex, tp = triangle.  Unknown order = first-appearance order of   */
/* the edges that end up shared by two triangles.                      */
/* Returns the RWG count, or -1 if out_cap is too small.               */
/* ------------------------------------------------------------------ */
int orc_build_rwg(int nt, const int *tri, int *rwg_out /* cap x 6 */, int out_cap, int *n_boundary)
{
    int cap = 3 * nt;
    int *e = (int *)malloc(sizeof(int) * 6 * (size_t)cap); /* v1 v2 vxp vxm tp tm */
    int ne = 0;
    for (int t = 0; t < nt; ++t) {
        int A = tri[3 * t], B = tri[3 * t + 1], C = tri[3 * t + 2];
        int new1 = 1, new2 = 1, new3 = 1;
        for (int i = 0; i < ne; ++i) {
            int *s = e + 6 * i;
            if (s[5] != -1) continue; /* isCommonEdgeTemp */
            int v1 = s[0], v2 = s[1];
            if (A == v2 && B == v1) {
                new1 = 0;
                s[3] = C;
                s[5] = t;
            } else if (B == v2 && C == v1) {
                new2 = 0;
                s[3] = A;
                s[5] = t;
            } else if (C == v2 && A == v1) {
                new3 = 0;
                s[3] = B;
                s[5] = t;
            }
        }
        if (new1) { int *s = e + 6 * ne++; s[0] = A; s[1] = B; s[2] = C; s[3] = -1; s[4] = t; s[5] = -1; }
        if (new2) { int *s = e + 6 * ne++; s[0] = B; s[1] = C; s[2] = A; s[3] = -1; s[4] = t; s[5] = -1; }
        if (new3) { int *s = e + 6 * ne++; s[0] = C; s[1] = A; s[2] = B; s[3] = -1; s[4] = t; s[5] = -1; }
    }
    int n = 0, nb = 0;
    for (int i = 0; i < ne; ++i) {
        int *s = e + 6 * i;
        if (s[5] == -1) { ++nb; continue; }
        if (n >= out_cap) { free(e); return -1; }
        memcpy(rwg_out + 6 * n, s, sizeof(int) * 6);
        ++n;
    }
    if (n_boundary) *n_boundary = nb;
    free(e);
    return n;
}

/* ------------------------------------------------------------------ */
/* Incidence vectors: EMSolver/EM.cpp:29-41.  Angles in degrees.       */
/* ------------------------------------------------------------------ */
void orc_incidence(double theta_deg, double phi_deg, double pol_deg, double *kvec, double *evec, double *hvec)
{
    double d2r = ORC_PI / 180.0;
    double th = theta_deg * d2r, ph = phi_deg * d2r, po = pol_deg * d2r;
    v3 k = mk(-sin(th) * cos(ph), -sin(th) * sin(ph), -cos(th));
    v3 e = mk(-cos(po) * cos(th) * cos(ph) - sin(po) * sin(ph), -cos(po) * cos(th) * sin(ph) + sin(po) * cos(ph),
              cos(po) * sin(th));
    v3 h = cross(k, e);
    kvec[0] = k.x; kvec[1] = k.y; kvec[2] = k.z;
    evec[0] = e.x; evec[1] = e.y; evec[2] = e.z;
    hvec[0] = h.x; hvec[1] = h.y; hvec[2] = h.z;
}

/* ------------------------------------------------------------------ */
/* EFIE regular pair integral, EMSolver/EFIE.cpp:139-160 (eZppKernel)  */
/*   sum_pq w_p w_q (1/4 rho_f.rho_s - 1/k^2) exp(-jkR)/R              */
/* with rho_f = r_p - vf, rho_s = r_q - vs.  The other three sign     */
/* combinations are +/- this value (EFIE.cpp:661-688).                 */
/* ------------------------------------------------------------------ */
static cplx efie_regular(int q, double k, const v3 *tf, const v3 *ts, v3 vf, v3 vs)
{
    v3 pf[7], ps[7];
    const double *wf = quad_pts(q, tf[0], tf[1], tf[2], pf);
    const double *ws = quad_pts(q, ts[0], ts[1], ts[2], ps);
    cplx acc = 0;
    double ik2 = 1.0 / (k * k);
    for (int p = 0; p < q; ++p) {
        v3 rf = sub(pf[p], vf);
        for (int s = 0; s < q; ++s) {
            v3 rs = sub(ps[s], vs);
            double R = norm3(sub(pf[p], ps[s]));
            cplx G = cexp(-I * k * R) / R;
            acc += wf[p] * ws[s] * (0.25 * dot(rf, rs) - ik2) * G;
        }
    }
    return acc;
}

/* Analytic in-plane potential integrals over triangle ts seen from r:
 *   Is1 = int (rho' - rho)/R dS'  (in-plane vector part)
 *   Is2 = int 1/R dS'
 * EMSolver/EFIE.cpp:259-291 (edge loop of eZppSingular). */
static void efie_potentials(const v3 *ts, v3 r, v3 *Is1, double *Is2)
{
    v3 n = cross(sub(ts[1], ts[0]), sub(ts[2], ts[0]));
    v3 ni = unit(n);
    v3 rou = sub(r, scl(ni, dot(ni, r)));
    v3 s1 = mk(0, 0, 0);
    double s2 = 0;
    for (int i = 0; i < 3; ++i) {
        v3 a = ts[i], b = ts[(i + 1) % 3];
        v3 li = sub(b, a);
        v3 ui = unit(cross(li, ni));
        v3 l1 = unit(li);
        v3 rou_p = sub(b, scl(ni, dot(ni, b)));
        double p0 = fabs(dot(sub(a, r), ui));
        double d = dot(sub(r, a), ni);
        double lp = dot(sub(b, r), l1);
        double lm = dot(sub(a, r), l1);
        double rp = norm3(sub(r, b));
        double rm = norm3(sub(r, a));
        if (p0 < 1e-10) continue;
        v3 pi = unit(sub(sub(rou_p, rou), scl(l1, lp)));
        double r02 = p0 * p0 + d * d;
        double lg = log((rp + lp) / (rm + lm));
        s1 = add(s1, scl(ui, 0.5 * (r02 * lg + lp * rp - lm * rm)));
        double ad = fabs(d);
        s2 += dot(pi, ui) * (p0 * lg - ad * (atan((p0 * lp) / (r02 + ad * rp)) - atan((p0 * lm) / (r02 + ad * rm))));
    }
    *Is1 = s1;
    *Is2 = s2;
}

/* EFIE self term, EMSolver/EFIE.cpp:231-542 (eZ{pp,pm,mp,mm}Singular),
 * returned WITHOUT the sign product sf*ss (the caller multiplies) and
 * with the Sigma term returned separately (it carries no sign). */
static void efie_singular(int q, double k, const v3 *t, v3 vf, v3 vs, cplx sigma, cplx *zsign, cplx *zsigma)
{
    v3 pts[7];
    const double *w = quad_pts(q, t[0], t[1], t[2], pts);
    double S = tri_area(t[0], t[1], t[2]);
    double ik2 = 1.0 / (k * k);
    cplx Ze1 = 0, Ze2 = 0;
    for (int p = 0; p < q; ++p) {
        v3 rf = sub(pts[p], vf);
        for (int s = 0; s < q; ++s) {
            v3 rs = sub(pts[s], vs);
            double R = norm3(sub(pts[p], pts[s]));
            cplx coff = -0.5 * k * k * R + (I * k / 6.0) * (k * k * R * R - 6.0);
            Ze1 += w[p] * w[s] * (0.25 * dot(rf, rs) - ik2) * coff;
        }
        v3 Is1;
        double Is2;
        efie_potentials(t, pts[p], &Is1, &Is2);
        v3 vec = add(Is1, scl(sub(pts[p], vs), Is2));
        Ze2 += w[p] * (0.25 * dot(rf, vec) - Is2 * ik2);
    }
    *zsign = Ze1 + Ze2 / S;
    cplx z3 = 0;
    if (cabs(sigma) > 1e-8) {
        cplx acc = 0;
        for (int i = 0; i < q; ++i) acc += w[i] * dot(sub(pts[i], vf), sub(pts[i], vs));
        z3 = (ORC_PI * sigma / (S * I * k * ORC_Z0)) * acc;
    }
    *zsigma = z3;
}

/* ------------------------------------------------------------------ */
/* EFIE fill, EMSolver/EFIE.cpp:603-696 (RWG-pair loop, 4 triangle     */
/* pairs per entry, coef = j k Z0 / 4 pi).                             */
/* ------------------------------------------------------------------ */
void orc_fill_efie(const double *xyz, const int *tri, int n, const int *rwg, double k, int q, double sig_re,
                   double sig_im, int row0, int row1, double *Zout /* (row1-row0) x n complex */)
{
    cplx coef = I * k * ORC_Z0 / (4.0 * ORC_PI);
    cplx sigma = sig_re + I * sig_im;
#pragma omp parallel for schedule(dynamic, 4)
    for (int f = row0; f < row1; ++f) {
        const int *rf = rwg + 6 * f;
        v3 tfp[3], tfm[3];
        tri_vtx(xyz, tri, rf[4], tfp);
        tri_vtx(xyz, tri, rf[5], tfm);
        v3 vfp = vtx(xyz, rf[2]), vfm = vtx(xyz, rf[3]);
        double lf = norm3(sub(vtx(xyz, rf[0]), vtx(xyz, rf[1])));
        for (int s = 0; s < n; ++s) {
            const int *rs = rwg + 6 * s;
            v3 tsp[3], tsm[3];
            tri_vtx(xyz, tri, rs[4], tsp);
            tri_vtx(xyz, tri, rs[5], tsm);
            v3 vsp = vtx(xyz, rs[2]), vsm = vtx(xyz, rs[3]);
            double ls = norm3(sub(vtx(xyz, rs[0]), vtx(xyz, rs[1])));
            cplx zz = 0, zs, zg;
            /* (field tri, source tri, field free vertex, source free vertex, sign) */
            int tfi[4] = {rf[4], rf[4], rf[5], rf[5]};
            int tsi[4] = {rs[4], rs[5], rs[4], rs[5]};
            const v3 *TF[4] = {tfp, tfp, tfm, tfm};
            const v3 *TS[4] = {tsp, tsm, tsp, tsm};
            v3 VF[4] = {vfp, vfp, vfm, vfm};
            v3 VS[4] = {vsp, vsm, vsp, vsm};
            double sg[4] = {1, -1, -1, 1};
            for (int c = 0; c < 4; ++c) {
                if (tfi[c] == tsi[c]) {
                    efie_singular(q, k, TF[c], VF[c], VS[c], sigma, &zs, &zg);
                    zz += sg[c] * zs + zg;
                } else {
                    zz += sg[c] * efie_regular(q, k, TF[c], TS[c], VF[c], VS[c]);
                }
            }
            cplx z = coef * lf * ls * zz;
            size_t o = 2 * ((size_t)(f - row0) * n + s);
            Zout[o] = creal(z);
            Zout[o + 1] = cimag(z);
        }
    }
}

/* ------------------------------------------------------------------ */
/* CFIE regular pair, EMSolver/Miscellaneous.cpp:6-30 (cfieZKernel).   */
/* ------------------------------------------------------------------ */
static cplx cfie_regular(int q, double k, double alpha, const v3 *tf, const v3 *ts, v3 vf, v3 vs, v3 nf, double fs,
                         double ss)
{
    v3 pf[7], ps[7];
    const double *wf = quad_pts(q, tf[0], tf[1], tf[2], pf);
    const double *ws = quad_pts(q, ts[0], ts[1], ts[2], ps);
    cplx Ze = 0, Zm = 0;
    for (int p = 0; p < q; ++p) {
        v3 rf = scl(sub(pf[p], vf), fs);
        for (int s = 0; s < q; ++s) {
            v3 rs = scl(sub(ps[s], vs), ss);
            v3 d = sub(pf[p], ps[s]);
            double R = norm3(d);
            cplx G = cexp(-I * k * R) / R;
            Ze += wf[p] * ws[s] * (0.25 * dot(rf, rs) - (fs * ss) / (k * k)) * G;
            Zm += wf[p] * ws[s] * dot(rf, cross(nf, cross(d, rs))) * (1.0 + I * k * R) * G / (R * R);
        }
    }
    Ze *= I * k * ORC_Z0 / (4.0 * ORC_PI);
    Zm /= 16.0 * ORC_PI;
    return alpha * Ze + (1.0 - alpha) * ORC_Z0 * Zm;
}

/* CFIE self term, EMSolver/Miscellaneous.cpp:32-77 (cfieZSingular).
 * mode 0 reproduces the reference, including Ze2's reuse of the rho_s
 * left over from the last inner-loop point (Miscellaneous.cpp:47,53-66).
 * mode 1 replaces the EFIE part by the EFIE-style extraction. */
static cplx cfie_singular(int q, double k, double alpha, const v3 *t, v3 vf, v3 vs, double fs, double ss, int mode)
{
    v3 pts[7];
    const double *w = quad_pts(q, t[0], t[1], t[2], pts);
    double area = tri_area(t[0], t[1], t[2]);
    cplx Ze1 = 0, Ze2 = 0, Zm = 0, Ze;
    double sgn = fs * ss;
    for (int p = 0; p < q; ++p) {
        v3 rf = scl(sub(pts[p], vf), fs);
        v3 rs = mk(0, 0, 0);
        for (int s = 0; s < q; ++s) {
            rs = scl(sub(pts[s], vs), ss);
            double R = norm3(sub(pts[p], pts[s]));
            cplx coff = -0.5 * k * k * R + (I * k / 6.0) * (k * k * R * R - 6.0);
            Ze1 += w[p] * w[s] * (0.25 * dot(rf, rs) - sgn / (k * k)) * coff;
        }
        if (mode == 0) {
            for (int i = 0; i < 3; ++i) {
                v3 a = t[i], b = t[(i + 1) % 3];
                v3 Li = unit(sub(b, a));
                double Lp = dot(sub(b, pts[p]), Li);
                double Lm = dot(sub(a, pts[p]), Li);
                double Rp = norm3(sub(b, pts[p]));
                double Rm = norm3(sub(a, pts[p]));
                v3 sj0 = add(a, scl(Li, dot(sub(pts[p], a), Li)));
                double P0 = norm3(sub(sj0, pts[p]));
                if (P0 > 1e-10) Ze2 += w[p] * (0.25 * dot(rf, rs) - sgn / (k * k)) * P0 * log((Rp + Lp) / (Rm + Lm));
            }
        } else {
            v3 Is1;
            double Is2;
            efie_potentials(t, pts[p], &Is1, &Is2);
            v3 vec = scl(add(Is1, scl(sub(pts[p], vs), Is2)), ss);
            Ze2 += w[p] * (0.25 * dot(rf, vec) - sgn * Is2 / (k * k));
        }
        v3 rsp = scl(sub(pts[p], vs), ss);
        Zm += w[p] * dot(rf, rsp);
    }
    Ze = (I * k * ORC_Z0 / (4.0 * ORC_PI)) * (Ze1 + Ze2 / area);
    Zm = Zm / (8.0 * area);
    return alpha * Ze + (1.0 - alpha) * ORC_Z0 * Zm;
}

/* CFIE fill, EMSolver/CFIE.cpp:485-568: Z(f,s) = l_f l_s (Zpp+Zpm+Zmp+Zmm),
 * the field triangle's normal enters the MFIE part. */
void orc_fill_cfie(const double *xyz, const int *tri, int n, const int *rwg, double k, double alpha, int q, int mode,
                   int row0, int row1, double *Zout)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int f = row0; f < row1; ++f) {
        const int *rf = rwg + 6 * f;
        v3 tfp[3], tfm[3];
        tri_vtx(xyz, tri, rf[4], tfp);
        tri_vtx(xyz, tri, rf[5], tfm);
        v3 nfp = tri_normal(tfp[0], tfp[1], tfp[2]);
        v3 nfm = tri_normal(tfm[0], tfm[1], tfm[2]);
        v3 vfp = vtx(xyz, rf[2]), vfm = vtx(xyz, rf[3]);
        double lf = norm3(sub(vtx(xyz, rf[0]), vtx(xyz, rf[1])));
        for (int s = 0; s < n; ++s) {
            const int *rs = rwg + 6 * s;
            v3 tsp[3], tsm[3];
            tri_vtx(xyz, tri, rs[4], tsp);
            tri_vtx(xyz, tri, rs[5], tsm);
            v3 vsp = vtx(xyz, rs[2]), vsm = vtx(xyz, rs[3]);
            double ls = norm3(sub(vtx(xyz, rs[0]), vtx(xyz, rs[1])));
            int tfi[4] = {rf[4], rf[4], rf[5], rf[5]};
            int tsi[4] = {rs[4], rs[5], rs[4], rs[5]};
            const v3 *TF[4] = {tfp, tfp, tfm, tfm};
            const v3 *TS[4] = {tsp, tsm, tsp, tsm};
            v3 VF[4] = {vfp, vfp, vfm, vfm};
            v3 VS[4] = {vsp, vsm, vsp, vsm};
            v3 NF[4] = {nfp, nfp, nfm, nfm};
            double fsg[4] = {1, 1, -1, -1};
            double ssg[4] = {1, -1, 1, -1};
            cplx zz = 0;
            for (int c = 0; c < 4; ++c) {
                if (tfi[c] == tsi[c])
                    zz += cfie_singular(q, k, alpha, TF[c], VF[c], VS[c], fsg[c], ssg[c], mode);
                else
                    zz += cfie_regular(q, k, alpha, TF[c], TS[c], VF[c], VS[c], NF[c], fsg[c], ssg[c]);
            }
            cplx z = lf * ls * zz;
            size_t o = 2 * ((size_t)(f - row0) * n + s);
            Zout[o] = creal(z);
            Zout[o + 1] = cimag(z);
        }
    }
}

/* Surface part of the IP-SIE near entry between half-RWGs,
 * EMSolver/FastWT.cpp:753-762: z(f,s) = l_f l_s K(f,s), K the CFIE
 * triangle-pair kernel with both signs +1 (cfieZSingular when the two
 * half-RWGs share their triangle).  hrwg: n x 6 ints (v1, v2, vx, triangle,
 * ...), len[n] the edge lengths as the reference holds them. */
void orc_fill_hrwg(const double *xyz, const int *tri, int n, const int *hrwg, const double *len, double k,
                   double alpha, int q, int mode, int row0, int row1, double *Zout)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int f = row0; f < row1; ++f) {
        const int *hf = hrwg + 6 * f;
        v3 tf[3];
        tri_vtx(xyz, tri, hf[3], tf);
        v3 nf = tri_normal(tf[0], tf[1], tf[2]);
        v3 vf = vtx(xyz, hf[2]);
        for (int s = 0; s < n; ++s) {
            const int *hs = hrwg + 6 * s;
            v3 ts[3];
            tri_vtx(xyz, tri, hs[3], ts);
            v3 vs = vtx(xyz, hs[2]);
            cplx zz = hf[3] == hs[3] ? cfie_singular(q, k, alpha, tf, vf, vs, 1.0, 1.0, mode)
                                     : cfie_regular(q, k, alpha, tf, ts, vf, vs, nf, 1.0, 1.0);
            cplx z = len[f] * len[s] * zz;
            size_t o = 2 * ((size_t)(f - row0) * n + s);
            Zout[o] = creal(z);
            Zout[o + 1] = cimag(z);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Excitation.  EFIE: EMSolver/EFIE.cpp:568-601 + 698-722 (7-pt rule). */
/* CFIE: EMSolver/Miscellaneous.cpp:152-189 + CFIE.cpp:570-597 (3-pt). */
/* MFIE: EMSolver/MFIE.cpp:289-322 + 425-437 (7-pt, n x h part), times  */
/* Z0 (the reference divides V by Z0 and leaves Z unscaled; here both  */
/* sides carry Z0 so the MFIE system is CFIE with alpha = 0).          */
/* inc: nrhs x 9 doubles {k(3), e(3), h(3)}; B: n x nrhs complex.      */
/* kind 0 = EFIE, 1 = CFIE (alpha used), 2 = MFIE.                     */
/* ------------------------------------------------------------------ */
void orc_rhs(const double *xyz, const int *tri, int n, const int *rwg, int kind, double k, double alpha, int nrhs,
             const double *inc, double *B)
{
#pragma omp parallel for schedule(static)
    for (int u = 0; u < n; ++u) {
        const int *r = rwg + 6 * u;
        v3 tp[3], tm[3];
        tri_vtx(xyz, tri, r[4], tp);
        tri_vtx(xyz, tri, r[5], tm);
        v3 vp = vtx(xyz, r[2]), vm = vtx(xyz, r[3]);
        double l = norm3(sub(vtx(xyz, r[0]), vtx(xyz, r[1])));
        int q = (kind == 1) ? 3 : 7;
        v3 pp[7], pm[7];
        const double *w = quad_pts(q, tp[0], tp[1], tp[2], pp);
        quad_pts(q, tm[0], tm[1], tm[2], pm);
        v3 np = tri_normal(tp[0], tp[1], tp[2]);
        v3 nm = tri_normal(tm[0], tm[1], tm[2]);
        for (int j = 0; j < nrhs; ++j) {
            const double *c = inc + 9 * j;
            v3 kv = mk(c[0], c[1], c[2]), ev = mk(c[3], c[4], c[5]), hv = mk(c[6], c[7], c[8]);
            cplx Ve = 0, Vm = 0;
            for (int a = 0; a < q; ++a) {
                v3 rp = sub(pp[a], vp);
                cplx G = cexp(-I * k * dot(pp[a], kv));
                Ve += w[a] * dot(rp, ev) * G;
                if (kind != 0) Vm += w[a] * dot(rp, cross(np, hv)) * G;
                v3 rm = sub(vm, pm[a]);
                G = cexp(-I * k * dot(pm[a], kv));
                Ve += w[a] * dot(rm, ev) * G;
                if (kind != 0) Vm += w[a] * dot(rm, cross(nm, hv)) * G;
            }
            cplx v = (kind == 0) ? Ve : (kind == 2) ? Vm : (alpha * Ve + (1.0 - alpha) * Vm);
            v *= 0.5 * l;
            B[2 * ((size_t)u * nrhs + j)] = creal(v);
            B[2 * ((size_t)u * nrhs + j) + 1] = cimag(v);
        }
    }
}

/* ------------------------------------------------------------------ */
/* Bistatic RCS, EMSolver/EFIE.cpp:788-832 (CFIE.cpp:643-685 is the    */
/* same): Es = sum_u I_u L_u sum_7 w rho e^{+jk s.r};                  */
/* sigma = Z0^2 k^2 |1/2 s x (s x Es)|^2 / 4 pi, returned in dB.       */
/* X: n x nrhs complex; col selects the solution column per direction. */
/* ------------------------------------------------------------------ */
void orc_rcs(const double *xyz, const int *tri, int n, const int *rwg, double k, const double *X, int nrhs, int ndir,
             const double *dirs, const int *cols, double *out_db)
{
#pragma omp parallel for schedule(dynamic, 1)
    for (int d = 0; d < ndir; ++d) {
        v3 s = mk(dirs[3 * d], dirs[3 * d + 1], dirs[3 * d + 2]);
        int col = cols ? cols[d] : 0;
        cplx ex = 0, ey = 0, ez = 0;
        for (int u = 0; u < n; ++u) {
            const int *r = rwg + 6 * u;
            v3 tp[3], tm[3], pp[7], pm[7];
            tri_vtx(xyz, tri, r[4], tp);
            tri_vtx(xyz, tri, r[5], tm);
            v3 vp = vtx(xyz, r[2]), vm = vtx(xyz, r[3]);
            double L = norm3(sub(vtx(xyz, r[0]), vtx(xyz, r[1])));
            const double *w = quad_pts(7, tp[0], tp[1], tp[2], pp);
            quad_pts(7, tm[0], tm[1], tm[2], pm);
            cplx sx = 0, sy = 0, sz = 0;
            for (int g = 0; g < 7; ++g) {
                v3 rp = sub(pp[g], vp);
                cplx G = cexp(I * k * dot(pp[g], s)) * w[g];
                sx += rp.x * G; sy += rp.y * G; sz += rp.z * G;
                v3 rm = sub(vm, pm[g]);
                G = cexp(I * k * dot(pm[g], s)) * w[g];
                sx += rm.x * G; sy += rm.y * G; sz += rm.z * G;
            }
            size_t o = 2 * ((size_t)u * nrhs + col);
            cplx cu = (X[o] + I * X[o + 1]) * L;
            ex += sx * cu; ey += sy * cu; ez += sz * cu;
        }
        /* es = 0.5 s x (s x Es) */
        cplx cx = s.y * ez - s.z * ey, cy = s.z * ex - s.x * ez, cz = s.x * ey - s.y * ex;
        cplx fx = 0.5 * (s.y * cz - s.z * cy), fy = 0.5 * (s.z * cx - s.x * cz), fz = 0.5 * (s.x * cy - s.y * cx);
        double nrm = creal(fx * conj(fx) + fy * conj(fy) + fz * conj(fz));
        double rcs = ORC_Z0 * ORC_Z0 * k * k * nrm / (4.0 * ORC_PI);
        out_db[d] = 10.0 * log10(rcs);
    }
}

/* ------------------------------------------------------------------ */
/* Dense LU with partial pivoting, LAPACK ?getrf semantics (the        */
/* routine behind arma::solve in EMSolver/MFIE.cpp:64): pivot = first  */
/* row of maximal |re|+|im| (i?amax), ipiv 0-based.  Row-major A.      */
/* Returns 0, or j+1 for the first exactly-zero pivot (like info).     */
/* ------------------------------------------------------------------ */
int orc_lu(int n, double *Ad, int *ipiv)
{
    cplx *A = (cplx *)Ad;
    int info = 0;
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = -1.0;
        for (int i = j; i < n; ++i) {
            cplx a = A[(size_t)i * n + j];
            double v = fabs(creal(a)) + fabs(cimag(a));
            if (v > best) { best = v; p = i; }
        }
        ipiv[j] = p;
        cplx piv = A[(size_t)p * n + j];
        if (piv != 0) {
            if (p != j) {
                for (int c = 0; c < n; ++c) {
                    cplx t = A[(size_t)j * n + c];
                    A[(size_t)j * n + c] = A[(size_t)p * n + c];
                    A[(size_t)p * n + c] = t;
                }
            }
            cplx r = 1.0 / A[(size_t)j * n + j];
            for (int i = j + 1; i < n; ++i) A[(size_t)i * n + j] *= r;
        } else if (info == 0) {
            info = j + 1;
        }
        const cplx *urow = A + (size_t)j * n;
#pragma omp parallel for schedule(static) if (n - j > 256)
        for (int i = j + 1; i < n; ++i) {
            cplx *row = A + (size_t)i * n;
            cplx l = row[j];
            if (l == 0) continue;
            for (int c = j + 1; c < n; ++c) row[c] -= l * urow[c];
        }
    }
    return info;
}

/* Solve with the factor: apply the row swaps, unit-L forward, U back. */
void orc_lu_solve(int n, const double *LUd, const int *ipiv, int nrhs, double *Bd)
{
    const cplx *A = (const cplx *)LUd;
    cplx *B = (cplx *)Bd;
    for (int j = 0; j < n; ++j) {
        int p = ipiv[j];
        if (p != j)
            for (int c = 0; c < nrhs; ++c) {
                cplx t = B[(size_t)j * nrhs + c];
                B[(size_t)j * nrhs + c] = B[(size_t)p * nrhs + c];
                B[(size_t)p * nrhs + c] = t;
            }
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nrhs; ++c) {
        for (int i = 0; i < n; ++i) {
            cplx s = B[(size_t)i * nrhs + c];
            const cplx *row = A + (size_t)i * n;
            for (int j = 0; j < i; ++j) s -= row[j] * B[(size_t)j * nrhs + c];
            B[(size_t)i * nrhs + c] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            cplx s = B[(size_t)i * nrhs + c];
            const cplx *row = A + (size_t)i * n;
            for (int j = i + 1; j < n; ++j) s -= row[j] * B[(size_t)j * nrhs + c];
            B[(size_t)i * nrhs + c] = s / row[i];
        }
    }
}

/* y = A x (row-major dense, complex) */
static void matvec(int n, const cplx *A, const cplx *x, cplx *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        cplx s = 0;
        const cplx *row = A + (size_t)i * n;
        for (int j = 0; j < n; ++j) s += row[j] * x[j];
        y[i] = s;
    }
}

static double vnorm2(int n, const cplx *x)
{
    double s = 0;
    for (int i = 0; i < n; ++i) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* Restarted GMRES(m), the algorithm of EMIterative/iml.cpp:156-207    */
/* with its two defects corrected (the parity decision): the Arnoldi   */
/* coefficient is h_ij = v_i^H w (iml.cpp:189 uses w^H v_i), the stop  */
/* test is relative ||r||/||b|| and the number of restarts is bounded. */
/* The small least-squares problem is solved with Givens rotations.    */
/* Returns the total number of inner iterations, or -1 if it did not   */
/* converge within max_restart cycles.  *relres gets ||b-Ax||/||b||.   */
/* ------------------------------------------------------------------ */
int orc_gmres(int n, const double *Ad, const double *bd, double *xd, int m, double tol, int max_restart,
              double *relres)
{
    const cplx *A = (const cplx *)Ad;
    const cplx *b = (const cplx *)bd;
    cplx *x = (cplx *)xd;
    cplx *V = (cplx *)calloc((size_t)n * (m + 1), sizeof(cplx));
    cplx *H = (cplx *)calloc((size_t)(m + 1) * m, sizeof(cplx));
    cplx *cs = (cplx *)calloc(m, sizeof(cplx));
    cplx *sn = (cplx *)calloc(m, sizeof(cplx));
    cplx *g = (cplx *)calloc(m + 1, sizeof(cplx));
    cplx *w = (cplx *)calloc(n, sizeof(cplx));
    cplx *y = (cplx *)calloc(m, sizeof(cplx));
    double bn = vnorm2(n, b);
    if (bn == 0) bn = 1;
    int total = 0, conv = 0;
    double rr = 1;
    for (int cyc = 0; cyc < max_restart && !conv; ++cyc) {
        matvec(n, A, x, w);
        for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
        double beta = vnorm2(n, w);
        rr = beta / bn;
        if (rr < tol) { conv = 1; break; }
        for (int i = 0; i < n; ++i) V[i] = w[i] / beta;
        memset(g, 0, sizeof(cplx) * (m + 1));
        g[0] = beta;
        int j;
        for (j = 0; j < m; ++j) {
            matvec(n, A, V + (size_t)j * n, w);
            ++total;
            for (int i = 0; i <= j; ++i) {
                cplx h = 0;
                const cplx *vi = V + (size_t)i * n;
                for (int t = 0; t < n; ++t) h += conj(vi[t]) * w[t];
                H[i * m + j] = h;
                for (int t = 0; t < n; ++t) w[t] -= h * vi[t];
            }
            double hn = vnorm2(n, w);
            H[(j + 1) * m + j] = hn;
            if (hn > 0)
                for (int t = 0; t < n; ++t) V[(size_t)(j + 1) * n + t] = w[t] / hn;
            for (int i = 0; i < j; ++i) {
                cplx a = H[i * m + j], bb = H[(i + 1) * m + j];
                H[i * m + j] = conj(cs[i]) * a + conj(sn[i]) * bb;
                H[(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
            }
            cplx a = H[j * m + j];
            double den = sqrt(creal(a * conj(a)) + hn * hn);
            if (den == 0) { cs[j] = 1; sn[j] = 0; }
            else { cs[j] = a / den; sn[j] = hn / den; }
            H[j * m + j] = conj(cs[j]) * a + conj(sn[j]) * hn;
            H[(j + 1) * m + j] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = conj(cs[j]) * g[j];
            rr = cabs(g[j + 1]) / bn;
            if (rr < tol || hn == 0) { ++j; conv = (rr < tol); break; }
        }
        int kk = (j > m) ? m : j;
        for (int i = kk - 1; i >= 0; --i) {
            cplx s = g[i];
            for (int t = i + 1; t < kk; ++t) s -= H[i * m + t] * y[t];
            y[i] = s / H[i * m + i];
        }
        for (int i = 0; i < kk; ++i)
            for (int t = 0; t < n; ++t) x[t] += y[i] * V[(size_t)i * n + t];
    }
    matvec(n, A, x, w);
    for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
    rr = vnorm2(n, w) / bn;
    if (relres) *relres = rr;
    free(V); free(H); free(cs); free(sn); free(g); free(w); free(y);
    return (rr < tol) ? total : -1;
}

/* ------------------------------------------------------------------ */
/* BiCGSTAB, EMIterative/iml.cpp:3-39 (the Armadillo overload CFIE     */
/* calls at CFIE.cpp:76): x0 = 0, relative ||s||/||b|| test.           */
/* Returns the iteration count; *ok = (count != max_iter).             */
/* ------------------------------------------------------------------ */
int orc_bicgstab(int n, const double *Ad, const double *bd, double *xd, double eps, int max_iter, int *ok)
{
    const cplx *A = (const cplx *)Ad;
    const cplx *b = (const cplx *)bd;
    cplx *x = (cplx *)xd;
    cplx *r0 = malloc(sizeof(cplx) * n), *r1 = malloc(sizeof(cplx) * n), *v = calloc(n, sizeof(cplx));
    cplx *p0 = calloc(n, sizeof(cplx)), *p1 = malloc(sizeof(cplx) * n), *s = malloc(sizeof(cplx) * n);
    cplx *t = malloc(sizeof(cplx) * n);
    for (int i = 0; i < n; ++i) { x[i] = 0; r0[i] = b[i]; r1[i] = b[i]; }
    cplx rou0 = 1, alph = 1, omg = 1, rou1, beta;
    double bn = vnorm2(n, b);
    int it = 0;
    for (int i = 0; i < max_iter; ++i) {
        ++it;
        rou1 = 0;
        for (int j = 0; j < n; ++j) rou1 += conj(r1[j]) * r0[j];
        beta = (rou1 / rou0) * (alph / omg);
        for (int j = 0; j < n; ++j) p1[j] = r0[j] + beta * (p0[j] - omg * v[j]);
        matvec(n, A, p1, v);
        cplx d = 0;
        for (int j = 0; j < n; ++j) d += conj(r1[j]) * v[j];
        alph = rou1 / d;
        for (int j = 0; j < n; ++j) { x[j] += alph * p1[j]; s[j] = r0[j] - alph * v[j]; }
        if (vnorm2(n, s) / bn < eps) break;
        matvec(n, A, s, t);
        cplx ts = 0, tt = 0;
        for (int j = 0; j < n; ++j) { ts += conj(t[j]) * s[j]; tt += conj(t[j]) * t[j]; }
        omg = ts / tt;
        for (int j = 0; j < n; ++j) { x[j] += omg * s[j]; r0[j] = s[j] - omg * t[j]; }
        if (vnorm2(n, s) / bn < eps) break;
        rou0 = rou1;
        memcpy(p0, p1, sizeof(cplx) * n);
    }
    if (ok) *ok = (it != max_iter);
    free(r0); free(r1); free(v); free(p0); free(p1); free(s); free(t);
    return it;
}

/* ------------------------------------------------------------------ */
/* BiCGSTAB, matrix-free overload EMIterative/iml.cpp:88-154: x0 = 0,  */
/* stop when ||r||^2 / ||b||^2 <= eps^2 or after max_iter steps; when  */
/* |rho| < eps_f^2 ||r0||^2 (eps_f = float epsilon, the reference's    */
/* NumTraits<complex<float>>::epsilon()) the shadow residual restarts  */
/* at r, and the first restart resets the step count.  Returns the     */
/* step count; *relres = ||r|| / ||b|| at exit.                        */
/* ------------------------------------------------------------------ */
int orc_bicgstab_mf(int n, const double *Ad, const double *bd, double *xd, double eps, int max_iter, double *relres)
{
    const cplx *A = (const cplx *)Ad;
    const cplx *b = (const cplx *)bd;
    cplx *x = (cplx *)xd;
    cplx *r = malloc(sizeof(cplx) * n), *r0 = malloc(sizeof(cplx) * n), *v = calloc(n, sizeof(cplx));
    cplx *p = calloc(n, sizeof(cplx)), *sv = malloc(sizeof(cplx) * n), *t = malloc(sizeof(cplx) * n);
    for (int i = 0; i < n; ++i) { x[i] = 0; r[i] = b[i]; r0[i] = b[i]; }
    double r0_sq = 0, rhs_sq = 0;
    for (int i = 0; i < n; ++i) { r0_sq += creal(conj(r0[i]) * r0[i]); rhs_sq += creal(conj(b[i]) * b[i]); }
    int it = 0;
    if (rhs_sq == 0) {
        if (relres) *relres = 0;
        free(r); free(r0); free(v); free(p); free(sv); free(t);
        return 0;
    }
    const double epsf = 1.1920928955078125e-07, eps2 = epsf * epsf, tol2 = eps * eps;
    cplx rho = 1, alpha = 1, w = 1;
    int restarts = 0;
    for (;;) {
        double rr = 0;
        for (int i = 0; i < n; ++i) rr += creal(conj(r[i]) * r[i]);
        if (!(rr / rhs_sq > tol2 && it < max_iter)) break;
        cplx rho_old = rho;
        rho = 0;
        for (int i = 0; i < n; ++i) rho += conj(r0[i]) * r[i];
        if (cabs(rho) < eps2 * r0_sq) {
            memcpy(r0, r, sizeof(cplx) * n);
            rho = r0_sq = rr;
            if (restarts++ == 0) it = 0;
        }
        cplx beta = (rho / rho_old) * (alpha / w);
        for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - w * v[i]);
        matvec(n, A, p, v);
        cplx d = 0;
        for (int i = 0; i < n; ++i) d += conj(r0[i]) * v[i];
        alpha = rho / d;
        for (int i = 0; i < n; ++i) sv[i] = r[i] - alpha * v[i];
        matvec(n, A, sv, t);
        double tt = 0;
        cplx ts = 0;
        for (int i = 0; i < n; ++i) { tt += creal(conj(t[i]) * t[i]); ts += conj(t[i]) * sv[i]; }
        w = tt > 0 ? ts / tt : 0;
        for (int i = 0; i < n; ++i) { x[i] += alpha * p[i] + w * sv[i]; r[i] = sv[i] - w * t[i]; }
        ++it;
    }
    double rr = 0;
    for (int i = 0; i < n; ++i) rr += creal(conj(r[i]) * r[i]);
    if (relres) *relres = sqrt(rr / rhs_sq);
    free(r); free(r0); free(v); free(p); free(sv); free(t);
    return it;
}

/* ------------------------------------------------------------------ */
/* Surface current magnitudes per triangle, EMSolver/CFIE.cpp:727-794. */
/* out: nt x 4 doubles {magn1, magn2, magn3, magn_center};            */
/* vout: nt x 9 doubles (the three corner points the reference writes) */
/* ------------------------------------------------------------------ */
void orc_current(const double *xyz, int nt, int n, const int *rwg, const double *Xd, int nrhs, int col, double *out,
                 double *vout)
{
    int *cnt = calloc(nt, sizeof(int));
    int *lst = malloc(sizeof(int) * 3 * (size_t)nt);
    for (int u = 0; u < n; ++u) {
        int tp = rwg[6 * u + 4], tm = rwg[6 * u + 5];
        if (cnt[tp] < 3) lst[3 * tp + cnt[tp]++] = u;
        if (cnt[tm] < 3) lst[3 * tm + cnt[tm]++] = u;
    }
    const cplx *X = (const cplx *)Xd;
    /* The corner magnitudes live outside the triangle loop, as `mag[3]` does
     * at CFIE.cpp:740: a triangle with fewer than three RWGs (an open
     * boundary) keeps the previous triangle's values at its remaining
     * corners.  The reference's current.sc carries those values. */
    double mg[3] = {0, 0, 0};
    for (int t = 0; t < nt; ++t) {
        v3 vx[3] = {mk(0, 0, 0), mk(0, 0, 0), mk(0, 0, 0)};
        double len[3] = {0, 0, 0};
        int m = cnt[t];
        for (int i = 0; i < m; ++i) {
            const int *r = rwg + 6 * lst[3 * t + i];
            vx[i] = vtx(xyz, (t == r[4]) ? r[2] : r[3]);
            len[i] = norm3(sub(vtx(xyz, r[0]), vtx(xyz, r[1])));
        }
        if (m == 1) {
            const int *r = rwg + 6 * lst[3 * t];
            vx[1] = vtx(xyz, r[0]);
            vx[2] = vtx(xyz, r[1]);
        }
        if (m == 2) {
            const int *r1 = rwg + 6 * lst[3 * t], *r2 = rwg + 6 * lst[3 * t + 1];
            if (r1[0] != r2[0] && r1[0] != r2[1]) vx[2] = vtx(xyz, r1[1]);
            else vx[2] = vtx(xyz, r1[0]);
        }
        v3 cen = scl(add(add(vx[0], vx[1]), vx[2]), 1.0 / 3.0);
        double dom = 2.0 * tri_area(vx[0], vx[1], vx[2]);
        cplx cc[3] = {0, 0, 0};
        for (int i = 0; i < m; ++i) {
            const int *r = rwg + 6 * lst[3 * t + i];
            double sg = (t == r[4]) ? 1.0 : -1.0;
            cplx cur = X[(size_t)lst[3 * t + i] * nrhs + col];
            v3 dv = sub(cen, vx[i]);
            cplx f = sg * cur * (len[i] / dom);
            cc[0] += f * dv.x; cc[1] += f * dv.y; cc[2] += f * dv.z;
            cplx nc[3] = {0, 0, 0};
            for (int j = 0; j < m; ++j) {
                if (i == j) continue;
                const int *rj = rwg + 6 * lst[3 * t + j];
                double sj = (t == rj[4]) ? 1.0 : -1.0;
                cplx cj = X[(size_t)lst[3 * t + j] * nrhs + col];
                v3 d2 = sub(vx[i], vx[j]);
                cplx fj = sj * cj * (len[j] / dom);
                nc[0] += fj * d2.x; nc[1] += fj * d2.y; nc[2] += fj * d2.z;
            }
            mg[i] = sqrt(creal(nc[0] * conj(nc[0]) + nc[1] * conj(nc[1]) + nc[2] * conj(nc[2])));
        }
        out[4 * t] = mg[0];
        out[4 * t + 1] = mg[1];
        out[4 * t + 2] = mg[2];
        out[4 * t + 3] = sqrt(creal(cc[0] * conj(cc[0]) + cc[1] * conj(cc[1]) + cc[2] * conj(cc[2])));
        if (vout)
            for (int i = 0; i < 3; ++i) {
                vout[9 * t + 3 * i] = vx[i].x;
                vout[9 * t + 3 * i + 1] = vx[i].y;
                vout[9 * t + 3 * i + 2] = vx[i].z;
            }
    }
    free(cnt);
    free(lst);
}

/* Wave number for a frequency, EMSolver/EFIE.cpp:47 (k = 2 pi f / c). */
double orc_wavenumber(double freq) { return 2.0 * ORC_PI * freq / ORC_C0; }

/* Exported for the tests: the analytic potentials Is1 (vector) and Is2 of
 * triangle t (9 doubles) seen from point r (EMSolver/EFIE.cpp:259-291). */
void orc_potentials(const double *t9, const double *r3, double *is1, double *is2)
{
    v3 t[3] = {mk(t9[0], t9[1], t9[2]), mk(t9[3], t9[4], t9[5]), mk(t9[6], t9[7], t9[8])};
    v3 s1;
    efie_potentials(t, mk(r3[0], r3[1], r3[2]), &s1, is2);
    is1[0] = s1.x;
    is1[1] = s1.y;
    is1[2] = s1.z;
}
