// Host-side geometry for the synthetic workloads and the RWG builder.
#include <cmath>
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

// RWG list with EMComponent/CommonEdge.cpp:704-743 semantics in O(E):
// triangles in order; candidate edges (A,B),(B,C),(C,A); a candidate matches
// every still-unmatched stored edge that is its reverse; unmatched candidates
// are appended with vxp = opposite vertex, tp = triangle.  Output keeps the
// stored order of the matched edges (CommonEdge.cpp:86-95).
extern "C" emfa_status emfa_rwg_build(int nt, const int *tri, int *rwg, int cap, int *n, int *n_boundary) {
    if (nt <= 0 || !tri || !n) {
        emfa::set_error("emfa_rwg_build: bad arguments");
        return EMFA_E_INVALID;
    }
    struct E {
        int v1, v2, vp, vm, tp, tm;
    };
    std::vector<E> e;
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
