// Driver around the reference's own EFIE/CFIE code (compiled from its
// sources by oracle/ref.mk into oracle/_ref/).  TEST INFRASTRUCTURE ONLY:
// it produces golden vectors for small meshes; nothing in the product links it.
//
// For a NASTRAN mesh it runs, with the reference's classes:
//   Mesh::loadMeshFile + CommonEdge::buildCommonEdge   (mesh as the reference reads it, RWG order)
//   EFIE::fillZ / fillV  or  CFIE::fillZ / fillV        (complex<float> Z and V)
//   a dense solve with Eigen's partial-pivot LU (the alternative left in EMSolver/EFIE.cpp:88)
//   MFIE::fillZ / fillV (EMSolver/MFIE.cpp:324-437) for kind mfie
//   EFIE::getBiRCS / CFIE::getBiRCS / MFIE::getBiRCS on a theta cut at phi = 0
//   CFIE::calculateSurfaceCurrent (EMSolver/CFIE.cpp:727-794) on the solved I
// and writes everything to a binary file (layout in write_output).
//
// MFIE keeps its kernels and members private; the driver reaches them by
// compiling MFIE.h with `private` read as `protected` (the class layout is
// unchanged, only access checking differs).
//
// usage: emfa_ref efie|cfie|mfie mesh.nas freq alpha sigma_re sigma_im theta phi pole out.bin
//        emfa_ref time cfie|efie mesh.nas freq alpha
//        emfa_ref mesh file.nas|.rpt|.stl out.bin
//        emfa_ref formats dir theta_from theta_to theta_delta phi_from phi_to phi_delta
//        emfa_ref contour mesh.nas freq alpha out.bin
// The contour form builds the half-RWG list of the IP-SIE solver with the
// reference's ContourEdge::buildEdge(mesh, HALF_RWG) (EMComponent/
// ContourEdge.cpp:15-95) and writes it with the contour lines of every
// nonconformal edge and, for every pair of half-RWGs, the surface part of
// the near entry that EMSolver/FastWT.cpp:753-762 forms from the reference's
// integral::cfieZKernel / cfieZSingular (Miscellaneous.cpp:6-77); FastWT.cpp
// itself does not compile with this toolchain (its size_t location matrices
// meet Armadillo's 64-bit uword as unsigned long long), so its own loop is
// restated here around the reference's kernels.
// The third form loads a mesh with Mesh::loadMeshFile (EMComponent/Mesh.cpp:
// 22-49, any of the three formats) and writes what the reference holds:
// int32 nv, nt; float xyz[nv*3]; per triangle int32 ids[3] and float
// corners[9], normal[3].
// The second form times the reference's own fillZ, one fillV and the far
// field of 37 directions (getBiRCS) on the mesh, single-threaded as the
// reference runs, and prints "n fillZ_s fillV_s rcs_s_per_direction".
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "CFIE.h"
#include "CommonEdge.h"
#include "EFIE.h"
#include "Mesh.h"
#include "Result.h"
#include "iml.h"
#define private protected
#include "MFIE.h"
#include "ContourEdge.h"
#undef private
#include "Miscellaneous.h"

using component::VectorR3;

namespace {

struct Out {
    std::vector<float> xyz;
    std::vector<int> tri, rwg;
    std::vector<float> Z, V, I, rcs, theta;
    std::vector<float> cur;  // per triangle: v1, v2, v3 (xyz), magn1, magn2, magn3, magnc
    int it_b = -1, it_mf = -1;
    std::vector<float> xb, xmf;  // BiCGSTAB solutions (both overloads)
};

// The reference's BiCGSTAB overloads on (Z, V): the Armadillo one that
// CFIE::solve calls (EMIterative/iml.cpp:3-39, CFIE.cpp:76: eps 0.01, 500
// steps) and the matrix-free one (iml.cpp:88-154) with the product Z x.
void run_bicgstab(const Qcx_mat &Z, const Qcx_vec &V, Out &o) {
    const int n = (int)V.n_elem;
    Qcx_vec xb;
    iml::BICGSTAB(Z, V, xb, 0.01f, 500, o.it_b);
    CMatrix Ze(n, n);
    CVector Ve(n), xmf;
    for (int r = 0; r < n; ++r) {
        Ve(r) = V(r);
        for (int c = 0; c < n; ++c) Ze(r, c) = Z(r, c);
    }
    MulVct mul = [&Ze](const CVector &v) -> CVector { return Ze * v; };
    iml::BICGSTAB(xmf, mul, Ve, 500, 0.01f, o.it_mf);
    for (int r = 0; r < n; ++r) {
        o.xb.insert(o.xb.end(), {xb(r).real(), xb(r).imag()});
        o.xmf.insert(o.xmf.end(), {xmf(r).real(), xmf(r).imag()});
    }
}

// The reference's per-triangle current magnitudes, computed from a given
// solution by the reference's own CFIE::calculateSurfaceCurrent.
class CurrentProbe : public mom::CFIE {
  public:
    void run(std::shared_ptr<component::Mesh> m, std::shared_ptr<component::CommonEdge> ce, const Out &o_in,
             Out &o) {
        mesh_ptr_ = m;
        ce_ptr_ = ce;
        unknowns_ = ce->getCommonEdgeNum();
        I.zeros(unknowns_);
        for (int r = 0; r < unknowns_; ++r) I(r) = Complex(o_in.I[2 * r], o_in.I[2 * r + 1]);
        std::vector<component::CurrentData> cd;
        calculateSurfaceCurrent(&cd);
        for (const auto &d : cd)
            o.cur.insert(o.cur.end(), {d.v1.x, d.v1.y, d.v1.z, d.v2.x, d.v2.y, d.v2.z, d.v3.x, d.v3.y, d.v3.z,
                                       d.magn1, d.magn2, d.magn3, d.magnc});
    }
};

void incidence(float theta_deg, float phi_deg, float pol_deg, VectorR3 &k, VectorR3 &e, VectorR3 &h) {
    // same float arithmetic as EMSolver/EM.cpp:29-41
    value_t pol = pol_deg * component::DegreesToRadians;
    value_t theta = theta_deg * component::DegreesToRadians;
    value_t phi = phi_deg * component::DegreesToRadians;
    k = VectorR3(-sin(theta) * cos(phi), -sin(theta) * sin(phi), -cos(theta));
    e = VectorR3(-cos(pol) * cos(theta) * cos(phi) - sin(pol) * sin(phi),
                 -cos(pol) * cos(theta) * sin(phi) + sin(pol) * cos(phi), cos(pol) * sin(theta));
    h = k * e;
}

template <typename S>
void collect_mesh(S *s, Out &o) {
    auto &m = *s->mesh();
    for (int i = 0; i < m.getNodeNum(); ++i) {
        const VectorR3 &v = m.getVertex(i);
        o.xyz.insert(o.xyz.end(), {v.x, v.y, v.z});
    }
    for (int t = 0; t < m.getTriangleNum(); ++t) {
        int a, b, c;
        m.getTriangleRef(t).getVertex(a, b, c);
        o.tri.insert(o.tri.end(), {a, b, c});
    }
    auto &ce = *s->edges();
    for (int i = 0; i < ce.getCommonEdgeNum(); ++i) {
        const auto &r = ce.getRWGRef(i);
        o.rwg.insert(o.rwg.end(), {r.v1, r.v2, r.vxp, r.vxm, r.tpid, r.tmid});
    }
}

class RefEFIE : public mom::EFIE {
  public:
    void setup(const std::string &nas, float freq, float) {
        mesh_ptr_ = std::make_shared<component::Mesh>();
        ce_ptr_ = std::make_shared<component::CommonEdge>();
        mesh_ptr_->loadMeshFile(nas);
        ce_ptr_->buildCommonEdge(mesh_ptr_);
        unknowns_ = ce_ptr_->getCommonEdgeNum();
        k_ = component::PI2 * freq / component::cc;
        threshold_edm = 3.0f * component::cc / freq;
        Sigma = Complex(0, 0);
        incidence(0, 0, 0, inc_k_, inc_e_, inc_h_);
        I_eigen.setZero(unknowns_);
    }
    int unknowns() const { return unknowns_; }
    void fill_z() {
        Z.zeros(unknowns_, unknowns_);
        Z_eigen.setZero(unknowns_, unknowns_);
        fillZ();
    }
    void fill_v() {
        V.zeros(unknowns_);
        V_eigen.setZero(unknowns_);
        fillV();
    }
    value_t bi_rcs(const VectorR3 &d) const { return getBiRCS(d); }
    void run(const std::string &nas, float freq, Complex sigma, float th, float ph, float po, Out &o) {
        mesh_ptr_ = std::make_shared<component::Mesh>();
        ce_ptr_ = std::make_shared<component::CommonEdge>();
        mesh_ptr_->loadMeshFile(nas);
        ce_ptr_->buildCommonEdge(mesh_ptr_);
        unknowns_ = ce_ptr_->getCommonEdgeNum();
        k_ = component::PI2 * freq / component::cc;
        threshold_edm = 3.0f * component::cc / freq;
        Sigma = sigma;
        incidence(th, ph, po, inc_k_, inc_e_, inc_h_);
        Z.zeros(unknowns_, unknowns_);
        Z_eigen.setZero(unknowns_, unknowns_);
        fillZ();
        V.zeros(unknowns_);
        V_eigen.setZero(unknowns_);
        fillV();
        I_eigen = Z_eigen.partialPivLu().solve(V_eigen);
        collect_mesh(this, o);
        int n = unknowns_;
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) o.Z.insert(o.Z.end(), {Z_eigen(r, c).real(), Z_eigen(r, c).imag()});
        for (int r = 0; r < n; ++r) {
            o.V.insert(o.V.end(), {V_eigen(r).real(), V_eigen(r).imag()});
            o.I.insert(o.I.end(), {I_eigen(r).real(), I_eigen(r).imag()});
        }
        rcs_cut(o);
        CurrentProbe().run(mesh_ptr_, ce_ptr_, o, o);
    }
    std::shared_ptr<component::Mesh> mesh() { return mesh_ptr_; }
    std::shared_ptr<component::CommonEdge> edges() { return ce_ptr_; }

  private:
    void rcs_cut(Out &o) {
        // theta cut at phi = 0 as EMComponent/Result.h:67-79 walks it
        for (value_t thetaDegree = 0; thetaDegree <= 360; thetaDegree += 5) {
            value_t theta = thetaDegree * component::DegreesToRadians;
            VectorR3 sk(sin(theta), 0, cos(theta));
            o.theta.push_back(thetaDegree);
            o.rcs.push_back(getBiRCS(sk));
        }
    }
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Per-phase wall time of the reference's own EFIE/CFIE code on one mesh.
template <typename S>
void time_phases(S &solver, double out[4]) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    solver.fill_z();
    out[1] = seconds_since(t0);
    t0 = clk::now();
    solver.fill_v();
    out[2] = seconds_since(t0);
    t0 = clk::now();
    volatile value_t sink = 0;
    for (int d = 0; d < 37; ++d) {
        value_t theta = value_t(5 * d) * component::DegreesToRadians;
        sink = sink + solver.bi_rcs(VectorR3(sin(theta), 0, cos(theta)));
    }
    out[3] = seconds_since(t0) / 37;
}

class RefCFIE : public mom::CFIE {
  public:
    void setup(const std::string &nas, float freq, float alpha) {
        mesh_ptr_ = std::make_shared<component::Mesh>();
        ce_ptr_ = std::make_shared<component::CommonEdge>();
        mesh_ptr_->loadMeshFile(nas);
        ce_ptr_->buildCommonEdge(mesh_ptr_);
        unknowns_ = ce_ptr_->getCommonEdgeNum();
        k_ = component::PI2 * freq / component::cc;
        alpha_ = alpha;
        incidence(0, 0, 0, inc_k_, inc_e_, inc_h_);
        I.zeros(unknowns_);
    }
    int unknowns() const { return unknowns_; }
    void fill_z() { fillZ(); }
    void fill_v() { fillV(); }
    value_t bi_rcs(const VectorR3 &d) const { return getBiRCS(d); }
    void run(const std::string &nas, float freq, float alpha, float th, float ph, float po, Out &o) {
        mesh_ptr_ = std::make_shared<component::Mesh>();
        ce_ptr_ = std::make_shared<component::CommonEdge>();
        mesh_ptr_->loadMeshFile(nas);
        ce_ptr_->buildCommonEdge(mesh_ptr_);
        unknowns_ = ce_ptr_->getCommonEdgeNum();
        k_ = component::PI2 * freq / component::cc;
        alpha_ = alpha;
        incidence(th, ph, po, inc_k_, inc_e_, inc_h_);
        fillZ();
        fillV();
        int n = unknowns_;
        CMatrix Ze(n, n);
        CVector Ve(n);
        for (int r = 0; r < n; ++r) {
            Ve(r) = V(r);
            for (int c = 0; c < n; ++c) Ze(r, c) = Z(r, c);
        }
        CVector Ie = Ze.partialPivLu().solve(Ve);
        I.zeros(n);
        for (int r = 0; r < n; ++r) I(r) = Ie(r);
        collect_mesh(this, o);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) o.Z.insert(o.Z.end(), {Z(r, c).real(), Z(r, c).imag()});
        for (int r = 0; r < n; ++r) {
            o.V.insert(o.V.end(), {V(r).real(), V(r).imag()});
            o.I.insert(o.I.end(), {I(r).real(), I(r).imag()});
        }
        for (value_t thetaDegree = 0; thetaDegree <= 360; thetaDegree += 5) {
            value_t theta = thetaDegree * component::DegreesToRadians;
            VectorR3 sk(sin(theta), 0, cos(theta));
            o.theta.push_back(thetaDegree);
            o.rcs.push_back(getBiRCS(sk));
        }
        CurrentProbe().run(mesh_ptr_, ce_ptr_, o, o);
        run_bicgstab(Z, V, o);
    }
    std::shared_ptr<component::Mesh> mesh() { return mesh_ptr_; }
    std::shared_ptr<component::CommonEdge> edges() { return ce_ptr_; }
};

class RefMFIE : public mom::MFIE {
  public:
    void run(const std::string &nas, float freq, float th, float ph, float po, Out &o) {
        mesh_ptr_ = std::make_shared<component::Mesh>();
        ce_ptr_ = std::make_shared<component::CommonEdge>();
        mesh_ptr_->loadMeshFile(nas);
        ce_ptr_->buildCommonEdge(mesh_ptr_);
        unknowns_ = ce_ptr_->getCommonEdgeNum();
        k_ = component::PI2 * freq / component::cc;
        incidence(th, ph, po, inc_k_, inc_e_, inc_h_);
        // MFIE::solve (EMSolver/MFIE.cpp:50-73) without its LAPACK call
        Z.zeros(unknowns_, unknowns_);
        I.zeros(unknowns_);
        V.zeros(unknowns_);
        fillZ();
        fillV();
        int n = unknowns_;
        CMatrix Ze(n, n);
        CVector Ve(n);
        for (int r = 0; r < n; ++r) {
            Ve(r) = V(r);
            for (int c = 0; c < n; ++c) Ze(r, c) = Z(r, c);
        }
        CVector Ie = Ze.partialPivLu().solve(Ve);
        for (int r = 0; r < n; ++r) I(r) = Ie(r);
        collect_mesh(this, o);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) o.Z.insert(o.Z.end(), {Z(r, c).real(), Z(r, c).imag()});
        for (int r = 0; r < n; ++r) {
            o.V.insert(o.V.end(), {V(r).real(), V(r).imag()});
            o.I.insert(o.I.end(), {I(r).real(), I(r).imag()});
        }
        for (value_t thetaDegree = 0; thetaDegree <= 360; thetaDegree += 5) {
            value_t theta = thetaDegree * component::DegreesToRadians;
            VectorR3 sk(sin(theta), 0, cos(theta));
            o.theta.push_back(thetaDegree);
            o.rcs.push_back(getBiRCS(sk));
        }
        CurrentProbe().run(mesh_ptr_, ce_ptr_, o, o);
    }
    std::shared_ptr<component::Mesh> mesh() { return mesh_ptr_; }
    std::shared_ptr<component::CommonEdge> edges() { return ce_ptr_; }
};

// layout: int32 nv, nt, n, nang; float xyz[nv*3]; int32 tri[nt*3]; int32 rwg[n*6];
//         float Z[n*n*2]; float V[n*2]; float I[n*2]; float theta[nang]; float rcs[nang];
//         int32 ncur; float cur[ncur*13];
//         int32 it_b; int32 it_mf; if it_b >= 0: float xb[n*2]; float xmf[n*2]
void write_output(const char *path, const Out &o) {
    FILE *f = std::fopen(path, "wb");
    if (!f) std::exit(3);
    int hdr[4] = {(int)o.xyz.size() / 3, (int)o.tri.size() / 3, (int)o.rwg.size() / 6, (int)o.rcs.size()};
    std::fwrite(hdr, sizeof(int), 4, f);
    std::fwrite(o.xyz.data(), sizeof(float), o.xyz.size(), f);
    std::fwrite(o.tri.data(), sizeof(int), o.tri.size(), f);
    std::fwrite(o.rwg.data(), sizeof(int), o.rwg.size(), f);
    std::fwrite(o.Z.data(), sizeof(float), o.Z.size(), f);
    std::fwrite(o.V.data(), sizeof(float), o.V.size(), f);
    std::fwrite(o.I.data(), sizeof(float), o.I.size(), f);
    std::fwrite(o.theta.data(), sizeof(float), o.theta.size(), f);
    std::fwrite(o.rcs.data(), sizeof(float), o.rcs.size(), f);
    int ncur = (int)o.cur.size() / 13;
    std::fwrite(&ncur, sizeof(int), 1, f);
    std::fwrite(o.cur.data(), sizeof(float), o.cur.size(), f);
    std::fwrite(&o.it_b, sizeof(int), 1, f);
    std::fwrite(&o.it_mf, sizeof(int), 1, f);
    std::fwrite(o.xb.data(), sizeof(float), o.xb.size(), f);
    std::fwrite(o.xmf.data(), sizeof(float), o.xmf.size(), f);
    std::fclose(f);
}

}  // namespace

// Result-file format probe: the reference's own writers (EMComponent/
// Result.h:48-104 getBistaticRCS, :234-255 getCurrentDistribution, with the
// prefixes of Result.cpp:26-51) fed with fixed synthetic values.
struct FmtProbe {
    std::string dir_;
    component::Scattering scattering_;
    value_t rcs(const VectorR3 &k) const { return 10.0f * k.z + 5.0f * k.x - 30.0f; }
    void cur(std::vector<component::CurrentData> *d) const {
        d->resize(5);
        for (int t = 0; t < 5; ++t) {
            auto &c = (*d)[t];
            c.v1 = VectorR3(0.1f * t, -0.25f, 1.0f / 3.0f);
            c.v2 = VectorR3(-1.5f + t, 2.0f / 7.0f, 1e-3f * t);
            c.v3 = VectorR3(123.456f, -0.000123f * t, 7.0f);
            c.magn1 = 1e-3f * (t + 1);
            c.magn2 = 0.5f / (t + 1);
            c.magn3 = 3.25f * t;
            c.magnc = 1.0f / 3.0f + t;
        }
    }
};

int write_formats(const char *dir, float th0, float th1, float thd, float ph0, float ph1, float phd) {
    FmtProbe p;
    p.dir_ = dir;
    p.scattering_ = {th0, th1, thd, ph0, ph1, phd};
    component::Result r;
    r.getBistaticRCS(&p, &FmtProbe::rcs);
    r.getCurrentDistribution(&p, &FmtProbe::cur);
    return 0;
}

int dump_mesh(const char *file, const char *out) {
    component::Mesh m;
    m.loadMeshFile(file);
    FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    int hdr[2] = {(int)m.getNodeNum(), (int)m.getTriangleNum()};
    std::fwrite(hdr, sizeof(int), 2, f);
    for (int i = 0; i < hdr[0]; ++i) {
        const VectorR3 &v = m.getVertex(i);
        float x[3] = {v.x, v.y, v.z};
        std::fwrite(x, sizeof(float), 3, f);
    }
    for (int t = 0; t < hdr[1]; ++t) {
        const auto &tr = m.getTriangleRef(t);
        int ids[3];
        tr.getVertex(ids[0], ids[1], ids[2]);
        VectorR3 a, b, c;
        tr.getVertex(a, b, c);
        const VectorR3 &n = tr.getNormal();
        float x[12] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, n.x, n.y, n.z};
        std::fwrite(ids, sizeof(int), 3, f);
        std::fwrite(x, sizeof(float), 12, f);
    }
    std::fclose(f);
    return 0;
}


// IP-SIE contour edges (ContourEdge keeps its lists private; `private` is read
// as `protected` above, which leaves the class layout unchanged).
class RefContour : public component::ContourEdge {
  public:
    const std::vector<HRWG> &hrwgs() const { return ce_list_; }
    const std::vector<ContourLine> &lines() const { return line_list_; }
};

// layout: int32 nv, nt, n, nl; float xyz[nv*3]; int32 tri[nt*3];
//         per half-RWG int32 v1, v2, vx, tid, idx, type and float length;
//         per nonconformal edge int32 count, then count x (int32 tri, float start[3], end[3]);
//         float zfs[n*n*2] (row = field half-RWG, column = source)
int dump_contour(const char *file, float freq, float alpha, const char *out) {
    auto mesh = std::make_shared<component::Mesh>();
    mesh->loadMeshFile(file);
    auto ce = std::make_shared<RefContour>();
    ce->buildEdge(mesh, policy::HALF_RWG);
    const auto &h = ce->hrwgs();
    const auto &ln = ce->lines();
    FILE *f = std::fopen(out, "wb");
    if (!f) return 3;
    int hdr[4] = {(int)mesh->getNodeNum(), (int)mesh->getTriangleNum(), (int)h.size(), (int)ln.size()};
    std::fwrite(hdr, sizeof(int), 4, f);
    for (int i = 0; i < hdr[0]; ++i) {
        const VectorR3 &v = mesh->getVertex(i);
        float x[3] = {v.x, v.y, v.z};
        std::fwrite(x, sizeof(float), 3, f);
    }
    for (int t = 0; t < hdr[1]; ++t) {
        int ids[3];
        mesh->getTriangleRef(t).getVertex(ids[0], ids[1], ids[2]);
        std::fwrite(ids, sizeof(int), 3, f);
    }
    for (const auto &e : h) {
        int r[6] = {e.v1, e.v2, e.vx, e.tid, e.idx, (int)e.type};
        std::fwrite(r, sizeof(int), 6, f);
        std::fwrite(&e.length, sizeof(float), 1, f);
    }
    for (const auto &lst : ln) {
        int c = (int)lst.size();
        std::fwrite(&c, sizeof(int), 1, f);
        for (const auto &p : lst) {
            float x[6] = {p.second.start.x, p.second.start.y, p.second.start.z,
                          p.second.end.x,   p.second.end.y,   p.second.end.z};
            std::fwrite(&p.first, sizeof(int), 1, f);
            std::fwrite(x, sizeof(float), 6, f);
        }
    }
    // k as EMSolver/FastWT.cpp:50-51 computes it
    auto wave_length = component::cc / freq;
    value_t k = component::PI2 / wave_length;
    const int n = (int)h.size();
    std::vector<float> z((size_t)n * n * 2);
    for (int fi = 0; fi < n; ++fi) {
        const auto &fh = h[fi];
        const auto &ft = mesh->getTriangleRef(fh.tid);
        VectorR3 vf3[3];
        ft.getVertex(vf3[0], vf3[1], vf3[2]);
        const VectorR3 vfx = mesh->getVertex(fh.vx), nml = ft.getNormal();
        for (int si = 0; si < n; ++si) {
            const auto &sh = h[si];
            VectorR3 vs3[3];
            mesh->getTriangleRef(sh.tid).getVertex(vs3[0], vs3[1], vs3[2]);
            const VectorR3 vsx = mesh->getVertex(sh.vx);
            Complex zfs = sh.tid == fh.tid
                              ? sh.length * fh.length * mom::integral::cfieZSingular(alpha, k, vf3, vs3, vfx, vsx, 1, 1)
                              : sh.length * fh.length * mom::integral::cfieZKernel(alpha, k, vf3, vs3, vfx, vsx, nml, 1, 1);
            z[2 * ((size_t)fi * n + si)] = zfs.real();
            z[2 * ((size_t)fi * n + si) + 1] = zfs.imag();
        }
    }
    std::fwrite(z.data(), sizeof(float), z.size(), f);
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && std::string(argv[1]) == "mesh") return dump_mesh(argv[2], argv[3]);
    if (argc == 6 && std::string(argv[1]) == "contour")
        return dump_contour(argv[2], std::strtof(argv[3], nullptr), std::strtof(argv[4], nullptr), argv[5]);
    if (argc == 9 && std::string(argv[1]) == "formats")
        return write_formats(argv[2], std::strtof(argv[3], nullptr), std::strtof(argv[4], nullptr),
                             std::strtof(argv[5], nullptr), std::strtof(argv[6], nullptr),
                             std::strtof(argv[7], nullptr), std::strtof(argv[8], nullptr));
    if (argc == 6 && std::string(argv[1]) == "time") {
        std::string kind = argv[2];
        float freq = std::strtof(argv[4], nullptr), alpha = std::strtof(argv[5], nullptr);
        double t[4] = {0, 0, 0, 0};
        if (kind == "cfie") {
            RefCFIE c;
            c.setup(argv[3], freq, alpha);
            t[0] = c.unknowns();
            time_phases(c, t);
        } else if (kind == "efie") {
            RefEFIE e;
            e.setup(argv[3], freq, alpha);
            t[0] = e.unknowns();
            time_phases(e, t);
        } else {
            return 2;
        }
        std::fprintf(stderr, "%d %.6f %.6f %.6e\n", (int)t[0], t[1], t[2], t[3]);
        return 0;
    }
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s efie|cfie|mfie mesh.nas freq alpha sigma_re sigma_im theta phi pole out.bin\n",
                     argv[0]);
        return 2;
    }
    std::string kind = argv[1];
    float freq = std::strtof(argv[3], nullptr), alpha = std::strtof(argv[4], nullptr);
    Complex sigma(std::strtof(argv[5], nullptr), std::strtof(argv[6], nullptr));
    float th = std::strtof(argv[7], nullptr), ph = std::strtof(argv[8], nullptr), po = std::strtof(argv[9], nullptr);
    Out o;
    if (kind == "efie") {
        RefEFIE e;
        e.run(argv[2], freq, sigma, th, ph, po, o);
    } else if (kind == "cfie") {
        RefCFIE c;
        c.run(argv[2], freq, alpha, th, ph, po, o);
    } else if (kind == "mfie") {
        RefMFIE m;
        m.run(argv[2], freq, th, ph, po, o);
    } else {
        return 2;
    }
    write_output(argv[10], o);
    return 0;
}
