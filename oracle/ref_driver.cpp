// Driver around the reference's own EFIE/CFIE code (compiled from its
// sources by oracle/ref.mk into oracle/_ref/).  TEST INFRASTRUCTURE ONLY:
// it produces golden vectors for small meshes; nothing in the product links it.
//
// For a NASTRAN mesh it runs, with the reference's classes:
//   Mesh::loadMeshFile + CommonEdge::buildCommonEdge   (mesh as the reference reads it, RWG order)
//   EFIE::fillZ / fillV  or  CFIE::fillZ / fillV        (complex<float> Z and V)
//   a dense solve with Eigen's partial-pivot LU (the alternative left in EMSolver/EFIE.cpp:88)
//   EFIE::getBiRCS / CFIE::getBiRCS on a theta cut at phi = 0
// and writes everything to a binary file (layout in write_output).
//
// usage: emfa_ref efie|cfie mesh.nas freq alpha sigma_re sigma_im theta phi pole out.bin
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "CFIE.h"
#include "CommonEdge.h"
#include "EFIE.h"
#include "Mesh.h"

using component::VectorR3;

namespace {

struct Out {
    std::vector<float> xyz;
    std::vector<int> tri, rwg;
    std::vector<float> Z, V, I, rcs, theta;
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

class RefCFIE : public mom::CFIE {
  public:
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
    }
    std::shared_ptr<component::Mesh> mesh() { return mesh_ptr_; }
    std::shared_ptr<component::CommonEdge> edges() { return ce_ptr_; }
};

// layout: int32 nv, nt, n, nang; float xyz[nv*3]; int32 tri[nt*3]; int32 rwg[n*6];
//         float Z[n*n*2]; float V[n*2]; float I[n*2]; float theta[nang]; float rcs[nang]
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
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s efie|cfie mesh.nas freq alpha sigma_re sigma_im theta phi pole out.bin\n",
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
    } else {
        return 2;
    }
    write_output(argv[10], o);
    return 0;
}
