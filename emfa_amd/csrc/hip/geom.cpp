// Host-side geometry for the synthetic workloads, the RWG builder and the
// half-RWG (contour edge) builder of the IP-SIE solver.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../../include/emfa_hip.h"

namespace emfa {
void set_error(const char *fmt, ...);
}

namespace {

struct V3 {
    double x, y, z;
};

// Icosahedron with vertices on the unit sphere; faces counter-clockwise seen
// from outside.
const double kPhi = 1.6180339887498948482;

void icosahedron(std::vector<V3> &v, std::vector<int> &f) {
    const double P = kPhi;
    const double base[12][3] = {{-1, P, 0}, {1, P, 0}, {-1, -P, 0}, {1, -P, 0}, {0, -1, P}, {0, 1, P},
                                {0, -1, -P}, {0, 1, -P}, {P, 0, -1}, {P, 0, 1}, {-P, 0, -1}, {-P, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9},  {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8},   {3, 9, 4},  {3, 4, 2},   {3, 2, 6},  {3, 6, 8},
                              {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    v.clear();
    for (auto &b : base) {
        double n = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        v.push_back({b[0] / n, b[1] / n, b[2] / n});
    }
    f.assign(&faces[0][0], &faces[0][0] + 60);
}

}  // namespace

extern "C" emfa_status emfa_geom_icosphere(int subdiv, double ax, double ay, double az, double *xyz, int nv_cap,
                                           int *tri, int nt_cap, int *nv, int *nt) {
    if (subdiv < 0 || subdiv > 9 || !(ax > 0 && ay > 0 && az > 0)) {
        emfa::set_error("emfa_geom_icosphere: bad arguments (subdiv=%d)", subdiv);
        return EMFA_E_INVALID;
    }
    long long T = 20LL << (2 * subdiv), V = 10LL * (1LL << (2 * subdiv)) + 2;
    if (nv) *nv = (int)V;
    if (nt) *nt = (int)T;
    if (!xyz || !tri) return EMFA_OK;  // size query
    if (nv_cap < V || nt_cap < T) {
        emfa::set_error("emfa_geom_icosphere: buffers too small (need %lld vertices, %lld triangles)", V, T);
        return EMFA_E_INVALID;
    }
    std::vector<V3> v;
    std::vector<int> f;
    icosahedron(v, f);
    for (int s = 0; s < subdiv; ++s) {
        std::unordered_map<long long, int> mid;
        mid.reserve(f.size() * 2);
        std::vector<int> nf;
        nf.reserve(f.size() * 4);
        auto m = [&](int a, int b) {
            long long key = a < b ? ((long long)a << 32) | b : ((long long)b << 32) | a;
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            V3 p = {(v[a].x + v[b].x) * 0.5, (v[a].y + v[b].y) * 0.5, (v[a].z + v[b].z) * 0.5};
            double n = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
            v.push_back({p.x / n, p.y / n, p.z / n});
            int id = (int)v.size() - 1;
            mid.emplace(key, id);
            return id;
        };
        for (size_t t = 0; t < f.size(); t += 3) {
            int a = f[t], b = f[t + 1], c = f[t + 2];
            int ab = m(a, b), bc = m(b, c), ca = m(c, a);
            int add[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), add, add + 12);
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); ++i) {
        xyz[3 * i] = v[i].x * ax;
        xyz[3 * i + 1] = v[i].y * ay;
        xyz[3 * i + 2] = v[i].z * az;
    }
    std::memcpy(tri, f.data(), sizeof(int) * f.size());
    return EMFA_OK;
}

extern "C" emfa_status emfa_geom_plate(int cells, double size, double *xyz, int nv_cap, int *tri, int nt_cap, int *nv,
                                       int *nt) {
    if (cells < 1 || !(size > 0)) {
        emfa::set_error("emfa_geom_plate: bad arguments");
        return EMFA_E_INVALID;
    }
    long long V = (long long)(cells + 1) * (cells + 1), T = 2LL * cells * cells;
    if (nv) *nv = (int)V;
    if (nt) *nt = (int)T;
    if (!xyz || !tri) return EMFA_OK;
    if (nv_cap < V || nt_cap < T) {
        emfa::set_error("emfa_geom_plate: buffers too small");
        return EMFA_E_INVALID;
    }
    double h = size / cells;
    for (int j = 0; j <= cells; ++j)
        for (int i = 0; i <= cells; ++i) {
            size_t id = (size_t)j * (cells + 1) + i;
            xyz[3 * id] = i * h - 0.5 * size;
            xyz[3 * id + 1] = j * h - 0.5 * size;
            xyz[3 * id + 2] = 0.0;
        }
    size_t t = 0;
    for (int j = 0; j < cells; ++j)
        for (int i = 0; i < cells; ++i) {
            int v00 = j * (cells + 1) + i, v10 = v00 + 1, v01 = v00 + cells + 1, v11 = v01 + 1;
            int tt[6] = {v00, v10, v11, v00, v11, v01};
            std::memcpy(tri + 3 * t, tt, sizeof(tt));
            t += 2;
        }
    return EMFA_OK;
}

// Edge list with EMComponent/CommonEdge.cpp:704-743 semantics in O(E):
// triangles in order; candidate edges (A,B),(B,C),(C,A); a candidate matches
// every still-unmatched stored edge that is its reverse; unmatched candidates
// are appended with vxp = opposite vertex, tp = triangle.  Boundary edges
// stay in the list with tm = -1.
namespace {
struct Edge {
    int v1, v2, vp, vm, tp, tm;
};

void common_edges(int nt, const int *tri, std::vector<Edge> &e) {
    e.clear();
    e.reserve(3 * (size_t)nt / 2 + 16);
    std::unordered_map<long long, std::vector<int>> open;  // directed (v1,v2) -> unmatched stored edges
    open.reserve(3 * (size_t)nt);
    auto key = [](int a, int b) { return ((long long)(unsigned)a << 32) | (unsigned)b; };
    for (int t = 0; t < nt; ++t) {
        int A = tri[3 * t], B = tri[3 * t + 1], C = tri[3 * t + 2];
        int cand[3][3] = {{A, B, C}, {B, C, A}, {C, A, B}};
        bool fresh[3] = {true, true, true};
        for (int c = 0; c < 3; ++c) {
            auto it = open.find(key(cand[c][1], cand[c][0]));
            if (it == open.end()) continue;
            for (int idx : it->second) {
                e[idx].vm = cand[c][2];
                e[idx].tm = t;
            }
            fresh[c] = false;
            open.erase(it);
        }
        for (int c = 0; c < 3; ++c) {
            if (!fresh[c]) continue;
            e.push_back({cand[c][0], cand[c][1], cand[c][2], -1, t, -1});
            open[key(cand[c][0], cand[c][1])].push_back((int)e.size() - 1);
        }
    }
}
}  // namespace

// RWG list: the shared edges of common_edges in stored order
// (CommonEdge.cpp:86-95).
extern "C" emfa_status emfa_rwg_build(int nt, const int *tri, int *rwg, int cap, int *n, int *n_boundary) {
    if (nt <= 0 || !tri || !n) {
        emfa::set_error("emfa_rwg_build: bad arguments");
        return EMFA_E_INVALID;
    }
    std::vector<Edge> e;
    common_edges(nt, tri, e);
    int cnt = 0, nb = 0;
    for (auto &x : e) {
        if (x.tm == -1) {
            ++nb;
            continue;
        }
        if (rwg) {
            if (cnt >= cap) {
                emfa::set_error("emfa_rwg_build: capacity %d too small", cap);
                return EMFA_E_INVALID;
            }
            int *o = rwg + 6 * (size_t)cnt;
            o[0] = x.v1;
            o[1] = x.v2;
            o[2] = x.vp;
            o[3] = x.vm;
            o[4] = x.tp;
            o[5] = x.tm;
        }
        ++cnt;
    }
    *n = cnt;
    if (n_boundary) *n_boundary = nb;
    return EMFA_OK;
}

// Half-RWG list of the IP-SIE solver (EMComponent/ContourEdge.cpp:15-95,
// buildContourEdge), in FP32 vertex arithmetic as the reference's VectorR3:
//   1. every shared edge of common_edges, in order, gives two conformal
//      half-RWGs, (v1, v2, vp, tp) and (v2, v1, vm, tm), edge number e++;
//   2. a boundary edge whose end points coincide (coordinates, reversed) with
//      those of a LATER boundary edge gives the conformal pair (itself, the
//      first such later edge); a boundary edge without a later partner goes
//      on to step 3 -- the later partner of a pair included, as the
//      reference leaves it in its list;
//   3. a step-3 edge (v1, v2) gets a contour line from every other step-3
//      edge (v3, v4) running against it over an overlap: v1v2.(v3 - v1) > 0,
//      v1v2.(v4 - v2) < 0, and the cosines of (v4 - v1) and (v3 - v2) with
//      v1v2 above 0.95 in modulus (1 where the points coincide); the line
//      runs from (cos(v4 - v1) < 0 ? v1 : v4) to (cos(v3 - v2) < 0 ? v3 : v2)
//      and belongs to the triangle of (v3, v4).  Edges with at least one line
//      are the nonconformal half-RWGs, numbered from 0.
// The pairing of step 2 is found through a hash of the end-point coordinates
// (the first later partner, as the reference's forward scan finds it); step 3
// keeps the reference's all-pairs scan, whose order the line lists follow.
namespace {
#pragma clang fp contract(off)
struct F3 {
    float x, y, z;
};
inline F3 fsub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float fdot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float fnorm(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline bool feq(F3 a, F3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
}  // namespace

extern "C" emfa_status emfa_contour_build(int nv, const double *xyz, int nt, const int *tri, int *hrwg,
                                          double *length, int cap, int *n, int *line_ptr, int *line_tri,
                                          double *line_seg, int line_cap, int *n_nc, int *n_lines) {
    if (nv <= 0 || nt <= 0 || !xyz || !tri || !n || !n_nc || !n_lines) {
        emfa::set_error("emfa_contour_build: bad arguments");
        return EMFA_E_INVALID;
    }
    for (size_t i = 0; i < 3 * (size_t)nt; ++i)
        if (tri[i] < 0 || tri[i] >= nv) {
            emfa::set_error("emfa_contour_build: triangle %zu names vertex %d of %d", i / 3, tri[i], nv);
            return EMFA_E_INVALID;
        }
    std::vector<F3> v(nv);
    for (int i = 0; i < nv; ++i) v[i] = {(float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2]};
    std::vector<Edge> e;
    common_edges(nt, tri, e);
    std::vector<float> elen(e.size());
    for (size_t i = 0; i < e.size(); ++i) elen[i] = fnorm(fsub(v[e[i].v2], v[e[i].v1]));
    struct H {
        int v1, v2, vx, t, idx, type;
        float len;
    };
    std::vector<H> out;
    out.reserve(3 * (size_t)nt);
    std::vector<int> left;  // boundary edges in list order
    int no = 0;
    for (size_t i = 0; i < e.size(); ++i) {
        const Edge &x = e[i];
        if (x.tm != -1) {
            out.push_back({x.v1, x.v2, x.vp, x.tp, no, 0, elen[i]});
            out.push_back({x.v2, x.v1, x.vm, x.tm, no, 0, elen[i]});
            ++no;
        } else {
            left.push_back((int)i);
        }
    }
    // step 2: coordinates of (v1, v2) -> positions in `left` (ascending)
    auto bits = [](float f) {
        f += 0.0f;  // -0 and +0 compare equal
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (uint64_t)u;
    };
    struct KeyHash {
        size_t operator()(const std::array<uint64_t, 6> &k) const {
            uint64_t h = 1469598103934665603ull;
            for (uint64_t x : k) h = (h ^ x) * 1099511628211ull;
            return (size_t)h;
        }
    };
    auto key = [&](F3 a, F3 b) {
        return std::array<uint64_t, 6>{bits(a.x), bits(a.y), bits(a.z), bits(b.x), bits(b.y), bits(b.z)};
    };
    std::unordered_map<std::array<uint64_t, 6>, std::vector<int>, KeyHash> pos;
    pos.reserve(2 * left.size() + 1);
    for (size_t p = 0; p < left.size(); ++p) {
        const Edge &x = e[left[p]];
        pos[key(v[x.v1], v[x.v2])].push_back((int)p);
    }
    std::vector<int> rest;
    for (size_t p = 0; p < left.size(); ++p) {
        const Edge &c = e[left[p]];
        auto it = pos.find(key(v[c.v2], v[c.v1]));
        int q = -1;
        if (it != pos.end()) {
            auto nx = std::upper_bound(it->second.begin(), it->second.end(), (int)p);
            if (nx != it->second.end()) q = *nx;
        }
        if (q >= 0) {
            const Edge &d = e[left[q]];
            out.push_back({c.v1, c.v2, c.vp, c.tp, no, 0, elen[left[p]]});
            out.push_back({d.v1, d.v2, d.vp, d.tp, no, 0, elen[left[q]]});
            ++no;
        } else {
            rest.push_back(left[p]);
        }
    }
    // step 3
    const double threshold = 0.95f;
    std::vector<int> lptr{0}, ltri;
    std::vector<float> lseg;
    int nc = 0;
    for (size_t a = 0; a < rest.size(); ++a) {
        const Edge &c = e[rest[a]];
        const F3 v1 = v[c.v1], v2 = v[c.v2], v1v2 = fsub(v2, v1);
        const double nv1v2 = fnorm(v1v2);
        const size_t before = ltri.size();
        for (size_t b = 0; b < rest.size(); ++b) {
            if (b == a) continue;
            const Edge &d = e[rest[b]];
            const F3 v3 = v[d.v1], v4 = v[d.v2];
            const F3 t1 = fsub(v4, v1), t2 = fsub(v3, v1), t3 = fsub(v4, v2), t4 = fsub(v3, v2);
            if (fdot(v1v2, t2) <= 0 || fdot(v1v2, t3) >= 0) continue;
            const double t1s = feq(v1, v4) ? 1.0 : fdot(v1v2, t1) / (nv1v2 * fnorm(t1));
            const double t4s = feq(v2, v3) ? 1.0 : fdot(v1v2, t4) / (nv1v2 * fnorm(t4));
            if (std::fabs(t1s) > threshold && std::fabs(t4s) > threshold) {
                const F3 st = t1s < 0 ? v1 : v4, en = t4s < 0 ? v3 : v2;
                ltri.push_back(d.tp);
                lseg.insert(lseg.end(), {st.x, st.y, st.z, en.x, en.y, en.z});
            }
        }
        if (ltri.size() > before) {
            out.push_back({c.v1, c.v2, c.vp, c.tp, nc, 1, elen[rest[a]]});
            ++nc;
            lptr.push_back((int)ltri.size());
        }
    }
    *n = (int)out.size();
    *n_nc = nc;
    *n_lines = (int)ltri.size();
    if (hrwg || length) {
        if (!hrwg || !length || cap < *n) {
            emfa::set_error("emfa_contour_build: capacity %d for %d half-RWGs", cap, *n);
            return EMFA_E_INVALID;
        }
        for (size_t i = 0; i < out.size(); ++i) {
            int *o = hrwg + 6 * i;
            o[0] = out[i].v1;
            o[1] = out[i].v2;
            o[2] = out[i].vx;
            o[3] = out[i].t;
            o[4] = out[i].idx;
            o[5] = out[i].type;
            length[i] = out[i].len;
        }
    }
    if (line_ptr || line_tri || line_seg) {
        if (!line_ptr || !line_tri || !line_seg || line_cap < *n_lines) {
            emfa::set_error("emfa_contour_build: line capacity %d for %d lines", line_cap, *n_lines);
            return EMFA_E_INVALID;
        }
        std::memcpy(line_ptr, lptr.data(), lptr.size() * sizeof(int));
        for (size_t i = 0; i < ltri.size(); ++i) line_tri[i] = ltri[i];
        for (size_t i = 0; i < lseg.size(); ++i) line_seg[i] = lseg[i];
    }
    return EMFA_OK;
}
