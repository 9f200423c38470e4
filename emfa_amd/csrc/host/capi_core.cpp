// C entry points of the host core (declared in include/emfa_core.h):
// configuration lookup with the loader's semantics, the NAS reader, a whole
// CONFIG.yml run in-process, and the solve server.
#include <cstring>
#include <iostream>
#include <map>
#include <mutex>
#include <string>

#include "../../../include/emfa_core.h"
#include "core.h"
#include "em.h"
#include "serve.h"

struct emfa_server {
    serve::Server *srv = nullptr;
    std::mutex mu;
    std::map<long long, serve::Result> taken;  // waited for, not yet fetched
};

namespace {
thread_local std::string g_core_err;
int fail(int code, const std::string &msg) {
    g_core_err = msg;
    return code;
}
}  // namespace

extern "C" {

const char *emfa_core_last_error(void) { return g_core_err.c_str(); }

// Value of `key` (with its colon) after loading `src` (a path, or the text
// itself when is_text != 0).  Returns 0, 1 (missing key) or 2 (bad file).
int emfa_core_config_get(const char *src, int is_text, const char *key, char *out, int cap) {
    try {
        component::ConfigLoader L;
        std::streambuf *old = std::cout.rdbuf(nullptr);
        if (is_text)
            L.loadFromString(src);
        else
            L.load(src);
        std::cout.rdbuf(old);
        if (!L.has(key)) return fail(1, std::string("missing key ") + key);
        const std::string &v = L.raw(key);
        if ((int)v.size() + 1 > cap) return fail(2, "buffer too small");
        std::memcpy(out, v.c_str(), v.size() + 1);
        return 0;
    } catch (const std::exception &e) {
        return fail(2, e.what());
    }
}

// NAS reader (EMComponent/Mesh.cpp:261-309 dialect).  With null buffers only
// the sizes are returned.  Returns 0 or 3 (file error).
int emfa_core_read_mesh(const char *path, double *xyz, int nv_cap, int *tri, int nt_cap, int *nv, int *nt) {
    try {
        component::Mesh m;
        m.loadMeshFile(path);
        *nv = (int)m.getNodeNum();
        *nt = (int)m.getTriangleNum();
        if (!xyz || !tri) return 0;
        if (nv_cap < *nv || nt_cap < *nt) return fail(1, "buffers too small");
        std::memcpy(xyz, m.xyz().data(), sizeof(double) * m.xyz().size());
        std::memcpy(tri, m.tri().data(), sizeof(int) * m.tri().size());
        return 0;
    } catch (const std::exception &e) {
        return fail(3, e.what());
    }
}

// Test hook for the result-file formats: writes the bistatic files and
// current.sc of fixed synthetic values (an RCS of 10 z + 5 x - 30 dB in
// direction (x, y, z); five triangles of set corners and magnitudes) with the
// writers the solvers use, for comparison with the reference's writers.
int emfa_core_write_formats(const char *dir, float th0, float th1, float thd, float ph0, float ph1, float phd) {
    try {
        component::Scattering sca{th0, th1, thd, ph0, ph1, phd};
        component::result::writeBistatic(dir, sca, [](int nd, const double *d, double *out) {
            for (int i = 0; i < nd; ++i) out[i] = 10.0f * (float)d[3 * i + 2] + 5.0f * (float)d[3 * i] - 30.0f;
        });
        std::vector<component::CurrentData> cd(5);
        for (int t = 0; t < 5; ++t) {
            auto &c = cd[t];
            c.v1 = {0.1f * t, -0.25f, 1.0f / 3.0f};
            c.v2 = {-1.5f + t, 2.0f / 7.0f, 1e-3f * t};
            c.v3 = {123.456f, -0.000123f * t, 7.0f};
            c.magn1 = 1e-3f * (t + 1);
            c.magn2 = 0.5f / (t + 1);
            c.magn3 = 3.25f * t;
            c.magnc = 1.0f / 3.0f + t;
        }
        component::result::writeCurrent(dir, cd);
        return 0;
    } catch (const std::exception &e) {
        return fail(3, e.what());
    }
}

// Runs solver::loadSolver on a config file.  Returns 0, or the launcher's
// exit codes (2 config, 3 file, 4 invalid argument, 5 runtime, 6 other).
int emfa_core_run(const char *config_path) {
    try {
        solver::loadSolver(config_path);
        return 0;
    } catch (const component::ConfigInvalid &e) {
        return fail(2, e.what());
    } catch (const component::FileError &e) {
        return fail(3, e.what());
    } catch (const std::invalid_argument &e) {
        return fail(4, e.what());
    } catch (const std::runtime_error &e) {
        return fail(5, e.what());
    } catch (const std::exception &e) {
        return fail(6, e.what());
    }
}

// ---------------------------------------------------------------- server
int emfa_serve_create(int device, int workers, size_t hbm_budget, int keep_current, emfa_server **out) {
    if (!out || workers < 1 || hbm_budget == 0) return fail(1, "emfa_serve_create: bad argument");
    *out = nullptr;
    try {
        auto *s = new emfa_server;
        try {
            s->srv = new serve::Server(device, workers, hbm_budget, keep_current != 0);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return 0;
    } catch (const std::exception &e) {
        return fail(5, e.what());
    }
}

int emfa_serve_destroy(emfa_server *s) {
    if (!s) return 0;
    delete s->srv;
    delete s;
    return 0;
}

long long emfa_serve_submit(emfa_server *s, const emfa_serve_request *r) {
    if (!s || !r || r->nrhs < 1 || !r->angles) return fail(-1, "emfa_serve_submit: bad argument");
    if (r->mesh_kind != EMFA_MESH_ICOSPHERE && r->mesh_kind != EMFA_MESH_PLATE && r->mesh_kind != EMFA_MESH_USER)
        return fail(-1, "emfa_serve_submit: unknown mesh kind");
    if (r->solver != EMFA_SERVE_DENSE && r->solver != EMFA_SERVE_MLFMA && r->solver != EMFA_SERVE_FASTWT)
        return fail(-1, "emfa_serve_submit: solver must be EMFA_SERVE_DENSE, EMFA_SERVE_MLFMA or EMFA_SERVE_FASTWT");
    if (r->solver == EMFA_SERVE_FASTWT && r->formulation != EMFA_CFIE)
        return fail(-1, "emfa_serve_submit: the IP-SIE operator is the combined-field one (formulation CFIE)");
    if (r->formulation != EMFA_EFIE && r->formulation != EMFA_CFIE)
        return fail(-1, "emfa_serve_submit: formulation must be EFIE or CFIE");
    if (r->precision != EMFA_F32 && r->precision != EMFA_F64) return fail(-1, "emfa_serve_submit: bad precision");
    if (r->quad != 3 && r->quad != 7) return fail(-1, "emfa_serve_submit: quadrature must be 3 or 7");
    if (!(r->freq > 0)) return fail(-1, "emfa_serve_submit: frequency must be positive");
    serve::Request q;
    q.mesh.kind = r->mesh_kind;
    q.mesh.subdiv = r->subdiv;
    q.mesh.ax = r->ax;
    q.mesh.ay = r->ay;
    q.mesh.az = r->az;
    q.formulation = r->formulation;
    q.alpha = r->alpha;
    q.freq = r->freq;
    q.precision = r->precision;
    q.quad = r->quad;
    q.angles.assign(r->angles, r->angles + 3 * (size_t)r->nrhs);
    q.solver = r->solver;
    if (r->solver == EMFA_SERVE_MLFMA || r->solver == EMFA_SERVE_FASTWT) {
        q.mlfma_box = r->mlfma_box > 0 ? r->mlfma_box : 0.25;
        q.iter_tol = r->iter_tol > 0 ? r->iter_tol : 1e-4;
        q.iter_max = r->iter_max > 0 ? r->iter_max : 600;
    }
    if (r->solver == EMFA_SERVE_FASTWT && r->fastwt_beta > 0) q.fastwt_beta = r->fastwt_beta;
    if (r->mesh_kind == EMFA_MESH_USER) {
        if (r->nv < 3 || r->nt < 1 || !r->xyz || !r->tri) return fail(-1, "emfa_serve_submit: empty caller mesh");
        auto xyz = std::make_shared<std::vector<double>>(r->xyz, r->xyz + 3 * (size_t)r->nv);
        auto tri = std::make_shared<std::vector<int>>(r->tri, r->tri + 3 * (size_t)r->nt);
        for (int v : *tri)
            if (v < 0 || v >= r->nv) return fail(-1, "emfa_serve_submit: triangle vertex id out of range");
        int n = 0, nb = 0;
        if (emfa_rwg_build(r->nt, tri->data(), nullptr, 0, &n, &nb) != EMFA_OK || n < 1)
            return fail(-1, "emfa_serve_submit: the caller mesh has no RWG functions");
        q.mesh.hash = serve::mesh_hash(*xyz, *tri);
        q.mesh.nrwg = n;
        q.mesh.xyz = std::move(xyz);
        q.mesh.tri = std::move(tri);
    }
    return s->srv->submit(q);
}

int emfa_serve_wait(emfa_server *s, long long id, emfa_serve_result *res) {
    if (!s || !res) return fail(1, "emfa_serve_wait: bad argument");
    serve::Result r;
    if (!s->srv->wait(id, &r)) return fail(1, "emfa_serve_wait: no such pending request");
    res->id = r.id;
    res->status = r.status;
    res->unknowns = r.unknowns;
    res->nrhs = r.nrhs;
    res->cache_hit = r.cache_hit ? 1 : 0;
    res->refine_iterations = r.refine_iterations;
    res->relres = r.relres;
    res->fp64_fallback = r.fp64_fallback ? 1 : 0;
    res->queue_ms = r.queue_ms;
    res->service_ms = r.service_ms;
    res->latency_ms = r.latency_ms;
    res->batch_size = r.batch_size;
    if (r.status != 0) g_core_err = r.error;
    std::lock_guard<std::mutex> l(s->mu);
    s->taken[r.id] = std::move(r);
    return 0;
}

int emfa_serve_fetch(emfa_server *s, long long id, double *rcs_db, int rcs_cap, double *current,
                     long long current_cap) {
    if (!s) return fail(1, "emfa_serve_fetch: bad argument");
    std::lock_guard<std::mutex> l(s->mu);
    auto it = s->taken.find(id);
    if (it == s->taken.end()) return fail(1, "emfa_serve_fetch: unknown id");
    const serve::Result &r = it->second;
    if (rcs_db && rcs_cap < (int)r.rcs_db.size()) return fail(1, "emfa_serve_fetch: rcs buffer too small");
    if (current && current_cap < (long long)r.current.size())
        return fail(1, "emfa_serve_fetch: current buffer too small");
    if (rcs_db) std::memcpy(rcs_db, r.rcs_db.data(), sizeof(double) * r.rcs_db.size());
    if (current) std::memcpy(current, r.current.data(), sizeof(double) * r.current.size());
    s->taken.erase(it);
    return 0;
}

int emfa_serve_drain(emfa_server *s) {
    if (!s) return fail(1, "emfa_serve_drain: bad argument");
    s->srv->drain();
    return 0;
}

int emfa_serve_hold(emfa_server *s, int on) {
    if (!s) return fail(1, "emfa_serve_hold: bad argument");
    s->srv->hold(on != 0);
    return 0;
}

int emfa_serve_get_stats(emfa_server *s, emfa_serve_stats *out) {
    if (!s || !out) return fail(1, "emfa_serve_get_stats: bad argument");
    serve::Stats t = s->srv->stats();
    out->completed = t.completed;
    out->failed = t.failed;
    out->cache_hits = t.cache_hits;
    out->p50_ms = t.p50_ms;
    out->p99_ms = t.p99_ms;
    out->mean_ms = t.mean_ms;
    out->max_ms = t.max_ms;
    out->wall_s = t.wall_s;
    out->solves_per_s = t.solves_per_s;
    out->rhs_per_s = t.rhs_per_s;
    out->peak_bytes = t.peak_bytes;
    out->pivot_fallbacks = t.pivot_fallbacks;
    out->batched = t.batched;
    out->factor_batches = t.factor_batches;
    out->batch_factors = t.batch_factors;
    return 0;
}

int emfa_serve_reset_stats(emfa_server *s) {
    if (!s) return fail(1, "emfa_serve_reset_stats: bad argument");
    s->srv->reset_stats();
    return 0;
}

int emfa_serve_cached_factors(emfa_server *s) { return s ? (int)s->srv->cached_factors() : 0; }

size_t emfa_serve_footprint(int unknowns, int nrhs, int precision, int factor_cached) {
    return serve::Server::footprint(unknowns, nrhs, precision, factor_cached != 0);
}

}  // extern "C"
