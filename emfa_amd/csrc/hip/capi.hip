// C ABI: errors, contexts, meshes, matrices, profiling.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>

#include "internal.h"

namespace emfa {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// Device memory pool, one per device, shared by every context on it.
// hipFree waits for the whole device, which would serialise contexts that
// work concurrently (a solve server's workers), so freed blocks are kept and
// handed out again.  A freed block carries an event recorded on the freeing
// context's stream; a context that takes it first makes both its streams
// wait on that event, so reuse across contexts stays in stream order.  The
// cache is trimmed, longest-cached blocks first, to three quarters of its
// limit when it exceeds it (EMFA_POOL_CACHE_MB, default half the device's
// memory, so a solve server's whole working set fits), and emptied when
// hipMalloc fails.  It outlives the contexts: the next context in the
// process (a new solve server) takes its blocks instead of allocating again.
// Handing tens of GB back to the driver at each server's end and allocating
// them again for the next one stalled every worker's first allocation of the
// new server for 0.6-2.5 s now and then.
struct PoolBlock {
    void *p;
    size_t bytes;
    hipEvent_t ev;
    unsigned long long seq;  // order of the frees
};

struct DevicePool {
    std::mutex mu;
    std::multimap<size_t, PoolBlock> free_blocks;
    std::map<void *, size_t> live;  // handed-out block -> its size
    std::vector<hipEvent_t> events;
    size_t cached = 0;
    unsigned long long seq = 0;
    size_t limit = 0;  // set with the first allocation on the device
};

static DevicePool &pool(int device) {
    static DevicePool pools[64];
    return pools[device & 63];
}

static size_t pool_limit(int device) {
    if (const char *e = getenv("EMFA_POOL_CACHE_MB")) return (size_t)std::max(0LL, atoll(e)) << 20;
    size_t freeb = 0, total = 0;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) {
        (void)hipGetLastError();
        total = (size_t)64 << 30;
    }
    if (prev >= 0) (void)hipSetDevice(prev);
    return total / 2;
}

static size_t pool_round(size_t bytes) {
    const size_t g = bytes >= (1u << 20) ? (2u << 20) : 4096;
    return (bytes + g - 1) / g * g;
}

// frees cached blocks, longest cached first (blocks of sizes no longer asked
// for go before the ones in use again), until at most `keep` bytes stay
// cached; call with mu held
static void pool_trim_locked(DevicePool &P, size_t keep) {
    while (P.cached > keep && !P.free_blocks.empty()) {
        auto it = P.free_blocks.begin();
        for (auto j = std::next(it); j != P.free_blocks.end(); ++j)
            if (j->second.seq < it->second.seq) it = j;
        (void)hipEventSynchronize(it->second.ev);
        (void)hipFree(it->second.p);
        P.events.push_back(it->second.ev);
        P.cached -= it->second.bytes;
        P.free_blocks.erase(it);
    }
}

hipError_t ctx_alloc(emfa_ctx *ctx, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    const size_t rb = pool_round(bytes);
    DevicePool &P = pool(ctx->device);
    {
        std::lock_guard<std::mutex> l(P.mu);
        if (!P.limit) P.limit = std::max<size_t>(1, pool_limit(ctx->device));
        auto it = P.free_blocks.lower_bound(rb);
        if (it != P.free_blocks.end() && it->first <= rb + rb / 4 + (2u << 20)) {
            PoolBlock b = it->second;
            P.free_blocks.erase(it);
            P.cached -= b.bytes;
            (void)hipStreamWaitEvent(ctx->stream, b.ev, 0);
            (void)hipStreamWaitEvent(ctx->aux, b.ev, 0);
            P.events.push_back(b.ev);
            P.live[b.p] = b.bytes;
            *p = b.p;
            ctx->bytes += b.bytes;
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(p, rb);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> l(P.mu);
        pool_trim_locked(P, 0);
        e = hipMalloc(p, rb);
    }
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> l(P.mu);
        P.live[*p] = rb;
        ctx->bytes += rb;
    }
    return e;
}

void ctx_free(emfa_ctx *ctx, void *p, size_t bytes) {
    if (!p) return;
    DevicePool &P = pool(ctx->device);
    std::lock_guard<std::mutex> l(P.mu);
    size_t rb = pool_round(bytes);
    auto lv = P.live.find(p);
    if (lv != P.live.end()) {
        rb = lv->second;
        P.live.erase(lv);
    }
    hipEvent_t ev = nullptr;
    if (!P.events.empty()) {
        ev = P.events.back();
        P.events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(p);
        ctx->bytes -= rb;
        return;
    }
    (void)hipEventRecord(ev, ctx->stream);
    P.free_blocks.emplace(rb, PoolBlock{p, rb, ev, P.seq++});
    P.cached += rb;
    ctx->bytes -= rb;
    if (P.cached > P.limit) pool_trim_locked(P, P.limit / 4 * 3);
}

emfa_status ctx_scratch(emfa_ctx *ctx, size_t bytes, void **p) {
    if (bytes > ctx->scratch_bytes) {
        if (ctx->scratch) {
            EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            ctx_free(ctx, ctx->scratch, ctx->scratch_bytes);
            ctx->scratch = nullptr;
            ctx->scratch_bytes = 0;
        }
        void *q = nullptr;
        hipError_t e = ctx_alloc(ctx, &q, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("scratch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
            return EMFA_E_OOM;
        }
        ctx->scratch = q;
        ctx->scratch_bytes = bytes;
    }
    *p = ctx->scratch;
    return EMFA_OK;
}

void prof_begin(emfa_ctx *ctx, const char *family, double work, hipEvent_t *a, hipEvent_t *b, hipStream_t s) {
    *a = *b = nullptr;
    if (!ctx->prof) return;
    auto get = [&]() {
        hipEvent_t e = nullptr;
        if (!ctx->event_pool.empty()) {
            e = ctx->event_pool.back();
            ctx->event_pool.pop_back();
        } else {
            (void)hipEventCreate(&e);
        }
        return e;
    };
    *a = get();
    *b = get();
    (void)hipEventRecord(*a, s ? s : ctx->stream);
    ctx->pending.push_back({family, *a, *b, work});
}

void prof_end(emfa_ctx *ctx, hipEvent_t a, hipEvent_t b, hipStream_t s) {
    (void)a;
    if (!ctx->prof || !b) return;
    (void)hipEventRecord(b, s ? s : ctx->stream);
}

void prof_collect(emfa_ctx *ctx) {
    if (ctx->pending.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->aux);
    for (auto &p : ctx->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &f = ctx->fam[p.family];
            f.ms += ms;
            f.launches += 1;
            f.work += p.work;
        }
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    ctx->pending.clear();
}

}  // namespace emfa

using namespace emfa;

extern "C" {

const char *emfa_last_error(void) { return g_err; }

int emfa_version(void) { return 1; }

emfa_status emfa_ctx_create(int device, emfa_ctx **out) {
    EMFA_REQUIRE(out, EMFA_E_INVALID, "emfa_ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return EMFA_E_NODEVICE;
    }
    EMFA_REQUIRE(device >= 0 && device < ndev, EMFA_E_INVALID, "device %d out of range (%d devices)", device, ndev);
    EMFA_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    EMFA_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
        return EMFA_E_NODEVICE;
    }
    emfa_ctx *c = new emfa_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return EMFA_E_HIP;
    }
    for (auto &e : c->ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    void *f = nullptr;
    if (ctx_alloc(c, &f, 64) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        (void)hipStreamDestroy(c->aux);
        delete c;
        set_error("flag allocation failed");
        return EMFA_E_OOM;
    }
    c->flags = (int *)f;
    (void)hipMemsetAsync(c->flags, 0, 64, c->stream);
    *out = c;
    return EMFA_OK;
}

emfa_status emfa_ctx_destroy(emfa_ctx *ctx) {
    if (!ctx) return EMFA_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    prof_collect(ctx);
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto &g : ctx->lu_graph) (void)hipGraphExecDestroy(g.second);
    ctx->lu_graph.clear();
    if (ctx->cap_main) (void)hipStreamDestroy(ctx->cap_main);
    if (ctx->cap_aux) (void)hipStreamDestroy(ctx->cap_aux);
    if (ctx->bmat) ctx_free(ctx, ctx->bmat, ctx->bmat_cap * sizeof(void *));
    if (ctx->scratch) ctx_free(ctx, ctx->scratch, ctx->scratch_bytes);
    ctx_free(ctx, ctx->flags, 64);
    (void)hipStreamSynchronize(ctx->aux);
    for (auto e : ctx->ev) (void)hipEventDestroy(e);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->aux);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EMFA_OK;
}

emfa_status emfa_ctx_sync(emfa_ctx *ctx) {
    EMFA_REQUIRE(ctx, EMFA_E_INVALID, "null ctx");
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return EMFA_OK;
}

void *emfa_ctx_stream(emfa_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

size_t emfa_ctx_bytes_in_use(const emfa_ctx *ctx) { return ctx ? ctx->bytes : 0; }

int emfa_ctx_cu_count(const emfa_ctx *ctx) { return ctx ? ctx->num_cu : 0; }

emfa_status emfa_ctx_set_panel_cap(emfa_ctx *ctx, int wgs) {
    EMFA_REQUIRE(ctx && wgs >= 0, EMFA_E_INVALID, "emfa_ctx_set_panel_cap: bad argument");
    ctx->panel_cap = wgs;
    return EMFA_OK;
}

emfa_status emfa_ctx_set_graphs(emfa_ctx *ctx, int on) {
    EMFA_REQUIRE(ctx, EMFA_E_INVALID, "emfa_ctx_set_graphs: null ctx");
    ctx->lu_graphs = on != 0;
    return EMFA_OK;
}

emfa_status emfa_ctx_graph_stats(const emfa_ctx *ctx, long long *captures, long long *launches) {
    EMFA_REQUIRE(ctx, EMFA_E_INVALID, "emfa_ctx_graph_stats: null ctx");
    if (captures) *captures = ctx->graph_captures;
    if (launches) *launches = ctx->graph_launches;
    return EMFA_OK;
}

emfa_status emfa_ctx_set_cu_share(emfa_ctx *ctx, int cus) {
    EMFA_REQUIRE(ctx && cus >= 0, EMFA_E_INVALID, "emfa_ctx_set_cu_share: bad argument");
    ctx->cu_share = cus;
    return EMFA_OK;
}

// ---------------------------------------------------------------- mesh
emfa_status emfa_mesh_upload(emfa_ctx *ctx, const double *xyz, int nv, const int *tri, int nt, const int *rwg, int n,
                             emfa_mesh **out) {
    EMFA_REQUIRE(ctx && xyz && tri && rwg && out, EMFA_E_INVALID, "emfa_mesh_upload: null argument");
    EMFA_REQUIRE(nv > 0 && nt > 0 && n > 0, EMFA_E_INVALID, "emfa_mesh_upload: empty mesh (nv=%d nt=%d n=%d)", nv, nt,
                 n);
    *out = nullptr;
    for (int t = 0; t < 3 * nt; ++t)
        EMFA_REQUIRE(tri[t] >= 0 && tri[t] < nv, EMFA_E_INVALID, "triangle vertex id %d out of range", tri[t]);
    std::vector<double> tg(16 * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        double v[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) v[i][c] = xyz[3 * tri[3 * t + i] + c];
        double cen[3];
        for (int c = 0; c < 3; ++c) cen[c] = (v[0][c] + v[1][c] + v[2][c]) / 3.0;
        for (int c = 0; c < 3; ++c) tg[(size_t)c * nt + t] = cen[c];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) tg[(size_t)(3 + 3 * i + c) * nt + t] = v[i][c] - cen[c];
        // normal per EMComponent/Triangle.cpp:47-61
        double ab[3], bc[3], ca[3];
        for (int c = 0; c < 3; ++c) {
            ab[c] = v[1][c] - v[0][c];
            bc[c] = v[2][c] - v[1][c];
            ca[c] = v[0][c] - v[2][c];
        }
        auto crs = [](const double *a, const double *b, double *r) {
            r[0] = a[1] * b[2] - a[2] * b[1];
            r[1] = a[2] * b[0] - a[0] * b[2];
            r[2] = a[0] * b[1] - a[1] * b[0];
        };
        double nn[3];
        double d1 = ab[0] * bc[0] + ab[1] * bc[1] + ab[2] * bc[2];
        double d2 = bc[0] * ca[0] + bc[1] * ca[1] + bc[2] * ca[2];
        if (d1 < d2)
            crs(ab, bc, nn);
        else
            crs(bc, ca, nn);
        double mag = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
        EMFA_REQUIRE(mag > 0, EMFA_E_INVALID, "degenerate triangle %d", t);
        for (int c = 0; c < 3; ++c) tg[(size_t)(12 + c) * nt + t] = nn[c] / mag;
        // Heron (EMComponent/VectorR3.h:147-154)
        auto dist = [](const double *a, const double *b) {
            return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                             (a[2] - b[2]) * (a[2] - b[2]));
        };
        double la = dist(v[0], v[1]), lb = dist(v[1], v[2]), lc = dist(v[2], v[0]);
        double s = 0.5 * (la + lb + lc);
        tg[(size_t)15 * nt + t] = std::sqrt(s * (s - la) * (s - lb) * (s - lc));
    }
    std::vector<int> t2r(3 * (size_t)nt, 0), tp(n), tm(n), ap(n), am(n);
    std::vector<double> len(n);
    for (int m = 0; m < n; ++m) {
        const int *r = rwg + 6 * m;
        EMFA_REQUIRE(r[4] >= 0 && r[4] < nt && r[5] >= 0 && r[5] < nt && r[4] != r[5], EMFA_E_INVALID,
                     "RWG %d has invalid triangles (%d,%d)", m, r[4], r[5]);
        int a = -1, b = -1;
        for (int i = 0; i < 3; ++i) {
            if (tri[3 * r[4] + i] == r[2]) a = i;
            if (tri[3 * r[5] + i] == r[3]) b = i;
        }
        EMFA_REQUIRE(a >= 0 && b >= 0, EMFA_E_INVALID, "RWG %d free vertex not in its triangle", m);
        EMFA_REQUIRE(t2r[(size_t)a * nt + r[4]] == 0 && t2r[(size_t)b * nt + r[5]] == 0, EMFA_E_INVALID,
                     "RWG %d: edge shared by more than two triangles (non-manifold)", m);
        t2r[(size_t)a * nt + r[4]] = m + 1;
        t2r[(size_t)b * nt + r[5]] = -(m + 1);
        tp[m] = r[4];
        tm[m] = r[5];
        ap[m] = a;
        am[m] = b;
        double dx = xyz[3 * r[0]] - xyz[3 * r[1]], dy = xyz[3 * r[0] + 1] - xyz[3 * r[1] + 1],
               dz = xyz[3 * r[0] + 2] - xyz[3 * r[1] + 2];
        len[m] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    std::vector<float> tg32(tg.begin(), tg.end()), len32(len.begin(), len.end());
    emfa_mesh *M = new emfa_mesh();
    M->ctx = ctx;
    M->nv = nv;
    M->nt = nt;
    M->n = n;
    M->t2r_host = t2r;
    M->xyz_host.assign(xyz, xyz + 3 * (size_t)nv);
    M->tri_host.assign(tri, tri + 3 * (size_t)nt);
    M->rwg_host.assign(rwg, rwg + 6 * (size_t)n);
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    struct A {
        void **p;
        size_t bytes;
        const void *src;
    } allocs[] = {
        {(void **)&M->tg64, tg.size() * sizeof(double), tg.data()},
        {(void **)&M->tg32, tg32.size() * sizeof(float), tg32.data()},
        {(void **)&M->t2r, t2r.size() * sizeof(int), t2r.data()},
        {(void **)&M->rtp, n * sizeof(int), tp.data()},
        {(void **)&M->rtm, n * sizeof(int), tm.data()},
        {(void **)&M->rap, n * sizeof(int), ap.data()},
        {(void **)&M->ram, n * sizeof(int), am.data()},
        {(void **)&M->rlen64, n * sizeof(double), len.data()},
        {(void **)&M->rlen32, n * sizeof(float), len32.data()},
    };
    for (auto &a : allocs) {
        if (ctx_alloc(ctx, a.p, a.bytes) != hipSuccess) {
            (void)hipGetLastError();
            emfa_mesh_destroy(M);
            set_error("mesh allocation failed");
            return EMFA_E_OOM;
        }
        M->bytes += a.bytes;
        hipError_t e = hipMemcpyAsync(*a.p, a.src, a.bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            emfa_mesh_destroy(M);
            set_error("mesh upload failed: %s", hipGetErrorString(e));
            return EMFA_E_HIP;
        }
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = M;
    return EMFA_OK;
}

emfa_status emfa_mesh_destroy(emfa_mesh *M) {
    if (!M) return EMFA_OK;
    emfa_ctx *c = M->ctx;
    (void)hipStreamSynchronize(c->stream);
    int nt = M->nt, n = M->n;
    ctx_free(c, M->tg64, 16 * (size_t)nt * sizeof(double));
    ctx_free(c, M->tg32, 16 * (size_t)nt * sizeof(float));
    ctx_free(c, M->t2r, 3 * (size_t)nt * sizeof(int));
    ctx_free(c, M->rtp, n * sizeof(int));
    ctx_free(c, M->rtm, n * sizeof(int));
    ctx_free(c, M->rap, n * sizeof(int));
    ctx_free(c, M->ram, n * sizeof(int));
    ctx_free(c, M->rlen64, n * sizeof(double));
    ctx_free(c, M->rlen32, n * sizeof(float));
    delete M;
    return EMFA_OK;
}

int emfa_mesh_unknowns(const emfa_mesh *M) { return M ? M->n : 0; }

// ---------------------------------------------------------------- matrices
emfa_status emfa_mat_create(emfa_ctx *ctx, int rows, int cols, int prec, emfa_mat **out) {
    EMFA_REQUIRE(ctx && out, EMFA_E_INVALID, "emfa_mat_create: null argument");
    EMFA_REQUIRE(rows > 0 && cols > 0, EMFA_E_INVALID, "emfa_mat_create: bad shape %dx%d", rows, cols);
    EMFA_REQUIRE(prec == EMFA_F32 || prec == EMFA_F64, EMFA_E_INVALID, "emfa_mat_create: bad precision %d", prec);
    *out = nullptr;
    emfa_mat *m = new emfa_mat();
    m->ctx = ctx;
    m->rows = rows;
    m->cols = cols;
    m->prec = prec;
    m->bytes = (size_t)rows * cols * elem_bytes(prec);
    EMFA_HIP_CHECK(hipSetDevice(ctx->device));
    if (ctx_alloc(ctx, &m->d, m->bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("emfa_mat_create: cannot allocate %zu bytes", m->bytes);
        delete m;
        return EMFA_E_OOM;
    }
    EMFA_HIP_CHECK(hipMemsetAsync(m->d, 0, m->bytes, ctx->stream));
    *out = m;
    return EMFA_OK;
}

emfa_status emfa_mat_destroy(emfa_mat *m) {
    if (!m) return EMFA_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    ctx_free(m->ctx, m->d, m->bytes);
    delete m;
    return EMFA_OK;
}

int emfa_mat_rows(const emfa_mat *m) { return m ? m->rows : 0; }
int emfa_mat_cols(const emfa_mat *m) { return m ? m->cols : 0; }
int emfa_mat_prec(const emfa_mat *m) { return m ? m->prec : -1; }
void *emfa_mat_device_ptr(emfa_mat *m) { return m ? m->d : nullptr; }

emfa_status emfa_mat_upload(emfa_mat *m, const double *host) {
    EMFA_REQUIRE(m && host, EMFA_E_INVALID, "emfa_mat_upload: null argument");
    size_t ne = 2 * (size_t)m->rows * m->cols;
    if (m->prec == EMFA_F64) {
        EMFA_HIP_CHECK(hipMemcpyAsync(m->d, host, ne * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
    } else {
        std::vector<float> tmp(host, host + ne);
        EMFA_HIP_CHECK(hipMemcpyAsync(m->d, tmp.data(), ne * sizeof(float), hipMemcpyHostToDevice, m->ctx->stream));
        EMFA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
        return EMFA_OK;
    }
    EMFA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    return EMFA_OK;
}

emfa_status emfa_mat_download(const emfa_mat *m, double *host) {
    EMFA_REQUIRE(m && host, EMFA_E_INVALID, "emfa_mat_download: null argument");
    size_t ne = 2 * (size_t)m->rows * m->cols;
    if (m->prec == EMFA_F64) {
        EMFA_HIP_CHECK(hipMemcpyAsync(host, m->d, ne * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        EMFA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    } else {
        std::vector<float> tmp(ne);
        EMFA_HIP_CHECK(hipMemcpyAsync(tmp.data(), m->d, ne * sizeof(float), hipMemcpyDeviceToHost, m->ctx->stream));
        EMFA_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
        for (size_t i = 0; i < ne; ++i) host[i] = tmp[i];
    }
    return EMFA_OK;
}

emfa_status emfa_mat_copy(emfa_mat *dst, const emfa_mat *src) {
    EMFA_REQUIRE(dst && src, EMFA_E_INVALID, "emfa_mat_copy: null argument");
    EMFA_REQUIRE(dst->rows == src->rows && dst->cols == src->cols, EMFA_E_INVALID, "emfa_mat_copy: shape mismatch");
    return convert_copy(dst->ctx, src, dst);
}

// ---------------------------------------------------------------- profiling
emfa_status emfa_prof_enable(emfa_ctx *ctx, int on) {
    EMFA_REQUIRE(ctx, EMFA_E_INVALID, "null ctx");
    if (!on) prof_collect(ctx);
    ctx->prof = on != 0;
    return EMFA_OK;
}

emfa_status emfa_prof_get(emfa_ctx *ctx, const char *family, double *total_ms, long long *launches, double *work) {
    EMFA_REQUIRE(ctx && family, EMFA_E_INVALID, "null argument");
    prof_collect(ctx);
    auto it = ctx->fam.find(family);
    double ms = 0, w = 0;
    long long l = 0;
    if (it != ctx->fam.end()) {
        ms = it->second.ms;
        l = it->second.launches;
        w = it->second.work;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = l;
    if (work) *work = w;
    return EMFA_OK;
}

emfa_status emfa_prof_reset(emfa_ctx *ctx) {
    EMFA_REQUIRE(ctx, EMFA_E_INVALID, "null ctx");
    prof_collect(ctx);
    ctx->fam.clear();
    return EMFA_OK;
}

}  // extern "C"
