#include "serve.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <stdexcept>

namespace serve {

using clock_t_ = std::chrono::steady_clock;

static double ms_between(clock_t_::time_point a, clock_t_::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

std::string MeshSpec::key() const {
    std::ostringstream o;
    o.precision(17);
    if (kind == USER) {
        o << "u:" << std::hex << hash << std::dec << ':' << (xyz ? xyz->size() / 3 : 0) << ':'
          << (tri ? tri->size() / 3 : 0);
        return o.str();
    }
    o << kind << ':' << subdiv << ':' << ax << ':' << ay << ':' << az;
    return o.str();
}

uint64_t mesh_hash(const std::vector<double> &xyz, const std::vector<int> &tri) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t bytes) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < bytes; ++i) {
            h ^= c[i];
            h *= 1099511628211ull;
        }
    };
    mix(xyz.data(), xyz.size() * sizeof(double));
    mix(tri.data(), tri.size() * sizeof(int));
    return h;
}

// What stays resident for a (mesh, frequency, formulation, ...) key: the
// device mesh with Z and its LU, or with the MLFMA operator.
struct Server::Factor {
    emfa_mesh *mesh = nullptr;
    emfa_mat *Z = nullptr;
    emfa_lu *lu = nullptr;
    emfa_mlfma *op = nullptr;
    bool fastwt = false;  // op is the IP-SIE operator; vectors are in half-RWG order
    int n = 0, prec = EMFA_F64;
    size_t bytes = 0;
    double k = 0;
    ~Factor() {
        if (op) emfa_mlfma_destroy(op);
        if (lu) emfa_lu_destroy(lu);
        if (Z) emfa_mat_destroy(Z);
        if (mesh) emfa_mesh_destroy(mesh);
    }
};

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

static void check(emfa_status st, const char *what) {
    if (st != EMFA_OK) throw StatusError(st, std::string(what) + ": " + emfa_last_error());
}

static void build_mesh(const MeshSpec &m, std::vector<double> &xyz, std::vector<int> &tri, std::vector<int> &rwg,
                       int &n) {
    int nv = 0, nt = 0;
    emfa_status st;
    if (m.kind == USER) {
        xyz = *m.xyz;
        tri = *m.tri;
        nt = (int)(tri.size() / 3);
        st = EMFA_OK;
    } else if (m.kind == PLATE) {
        st = emfa_geom_plate(m.subdiv, m.ax, nullptr, 0, nullptr, 0, &nv, &nt);
        if (st) throw StatusError(st, emfa_last_error());
        xyz.resize(3 * (size_t)nv);
        tri.resize(3 * (size_t)nt);
        st = emfa_geom_plate(m.subdiv, m.ax, xyz.data(), nv, tri.data(), nt, &nv, &nt);
    } else {
        st = emfa_geom_icosphere(m.subdiv, m.ax, m.ay, m.az, nullptr, 0, nullptr, 0, &nv, &nt);
        if (st) throw StatusError(st, emfa_last_error());
        xyz.resize(3 * (size_t)nv);
        tri.resize(3 * (size_t)nt);
        st = emfa_geom_icosphere(m.subdiv, m.ax, m.ay, m.az, xyz.data(), nv, tri.data(), nt, &nv, &nt);
    }
    if (st) throw StatusError(st, emfa_last_error());
    int nb = 0;
    st = emfa_rwg_build(nt, tri.data(), nullptr, 0, &n, &nb);
    if (st) throw StatusError(st, emfa_last_error());
    rwg.resize(6 * (size_t)std::max(n, 1));
    st = emfa_rwg_build(nt, tri.data(), rwg.data(), n, &n, &nb);
    if (st) throw StatusError(st, emfa_last_error());
}

static bool env_flag(const char *name, bool dflt) {
    const char *v = std::getenv(name);
    return v ? std::atoi(v) != 0 : dflt;
}

// RWG count of a mesh spec without building it (closed icosphere: 1.5 T =
// 30 4^s edges; plate of c x c cells: 3 c^2 - 2 c interior edges)
static int spec_unknowns(const MeshSpec &m) {
    if (m.kind == USER) return m.nrwg;
    if (m.kind == PLATE) return 3 * m.subdiv * m.subdiv - 2 * m.subdiv;
    return 30 * (1 << (2 * m.subdiv));
}

size_t Server::footprint(int n, int nrhs, int precision, bool factor_cached) {
    const size_t e = precision == EMFA_F32 ? 8 : 16;
    size_t vecs = (size_t)n * nrhs * 16 * 4 + (size_t)n * nrhs * e;  // B, X, R, Bd, D
    if (factor_cached) return vecs;
    // LU workspace at the default outer block of 256: two expanded U-row
    // buffers (2n x 512 reals each) and, for FP32, the bf16 operand planes of
    // the split update (hi and lo, 512 K, L rows + U rows + padding)
    const size_t nbr = 512, rows = (size_t)n + 255;
    size_t ws = 2 * (2 * (size_t)n) * nbr * (e / 2) + (1u << 20);
    if (precision == EMFA_F32) ws += 2 * nbr * 2 * (rows + 2 * rows + 256);
    // the factor keeps the inverses of its 256 x 256 diagonal blocks (L and U)
    // and, in FP32, its off-diagonal blocks scaled by them (n x n)
    size_t dinv = 2 * (((size_t)n + 255) / 256) * 256 * 256 * e;
    if (precision == EMFA_F32 && env_flag("EMFA_SOLVE_GH", false)) dinv += (size_t)n * n * e;
    // solve workspace: two n x nrhs vectors and an expanded block
    ws += 2 * (size_t)n * nrhs * e + 2 * (size_t)nrhs * 512 * e;
    return 2 * (size_t)n * n * e + dinv + vecs + ws;  // Z + factor + workspace
}

Server::Server(int device, int workers, size_t hbm_budget, bool keep_current)
    : device_(device), budget_(hbm_budget), keep_current_(keep_current) {
    workers = std::max(1, workers);
    for (int w = 0; w < workers; ++w) {
        emfa_ctx *c = nullptr;
        emfa_status st = emfa_ctx_create(device, &c);
        if (st) {
            for (auto x : ctxs_) emfa_ctx_destroy(x);
            throw std::runtime_error(std::string("serve: ") + emfa_last_error());
        }
        // concurrent factorisations split the compute units between them; the
        // panels whose workgroups wait on each other keep the static split
        if (workers > 1) {
            emfa_ctx_set_cu_share(c, std::max(1, emfa_ctx_cu_count(c) / workers));
            emfa_ctx_set_panel_cap(c, std::max(1, emfa_ctx_cu_count(c) / workers));
        }
        dynamic_share_ = env_flag("EMFA_SERVE_DYNAMIC_SHARE", true);
        // FP32 factors are only ever used under FP64 refinement: factor them
        // without interchanges (redone with pivoting when the multipliers
        // grow) and with the bf16-split trailing update; FP64 factors pivot
        // outer block 512 for every size (mixed 64-request stream, medians of
        // three passes: 120.1 / 118.1 solves/s with 256, 123.3 / 122.6 with
        // 512, 120.7-120.9 with 512 only from 2048 or 8192 unknowns up; the
        // 256-problem batch of at most 7680 unknowns: 224.3 with 256, 222.5
        // with 512)
        int nb = 512;
        if (const char *e = std::getenv("EMFA_SERVE_LU_NB")) nb = std::max(0, std::atoi(e));
        emfa_lu_set_params(c, nb, 0, env_flag("EMFA_SERVE_NOPIV", true) ? 2 : -1);
        if (const char *e = std::getenv("EMFA_SERVE_REFINE_TOL")) refine_tol_ = std::max(1e-14, std::atof(e));
        if (const char *e = std::getenv("EMFA_SERVE_BATCH_RHS")) batch_max_rhs_ = std::max(0, std::atoi(e));
        if (const char *e = std::getenv("EMFA_SERVE_FACTOR_BATCH")) factor_batch_max_ = std::max(0, std::atoi(e));
        if (env_flag("EMFA_SERVE_SPLIT", true)) emfa_lu_set_update_mode(c, 1);
        largest_first_ = env_flag("EMFA_SERVE_LARGEST_FIRST", true);
        if (const char *e = std::getenv("EMFA_SERVE_MAX_WAIT_MS")) max_wait_ms_ = std::max(0.0, std::atof(e));
        // new FP32 factors (alone or batched) as replayed HIP graphs
        if (env_flag("EMFA_SERVE_GRAPHS", false)) emfa_ctx_set_graphs(c, 1);
        ctxs_.push_back(c);
    }
    for (int w = 0; w < workers; ++w) threads_.emplace_back(&Server::worker, this, w);
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> l(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_mem_.notify_all();
    for (auto &t : threads_) t.join();
    cache_.clear();  // factors are released before their contexts
    lru_.clear();
    for (auto c : ctxs_) emfa_ctx_destroy(c);
}

int64_t Server::submit(const Request &r) {
    std::lock_guard<std::mutex> l(mu_);
    Job j{next_id_++, r, clock_t_::now(), false};
    if (!have_t0_) {
        t0_ = j.arrival;
        have_t0_ = true;
    }
    pending_.insert(j.id);
    queue_.push_back(std::move(j));
    cv_work_.notify_one();
    return next_id_ - 1;
}

bool Server::wait(int64_t id, Result *out) {
    std::unique_lock<std::mutex> l(mu_);
    for (;;) {
        if (id >= 0) {
            if (!pending_.count(id)) return false;  // unknown, or already returned
            auto it = done_.find(id);
            if (it != done_.end()) {
                *out = std::move(it->second);
                done_.erase(it);
                pending_.erase(id);
                return true;
            }
        } else {
            if (!done_.empty()) {
                auto it = done_.begin();
                *out = std::move(it->second);
                pending_.erase(it->first);
                done_.erase(it);
                return true;
            }
            if (pending_.empty()) return false;
        }
        cv_done_.wait(l);
    }
}

// The compute units are split between the workers that are running a
// request: with all of them busy each context gets its static share, and a
// request left alone (the end of a burst, a quiet period) gets the whole
// device for its launches.  A context reads its share at each launch.
void Server::rebalance_locked() {
    if (!dynamic_share_ || ctxs_.size() < 2) return;
    const int ncu = emfa_ctx_cu_count(ctxs_[0]);
    const int share = std::max(1, ncu / std::max(1, busy_workers_));
    for (auto c : ctxs_) emfa_ctx_set_cu_share(c, share);
}

void Server::drain() {
    std::unique_lock<std::mutex> l(mu_);
    if (held_) {
        held_ = false;
        cv_work_.notify_all();
    }
    cv_done_.wait(l, [&] { return queue_.empty() && busy_ == 0; });
}

// A batch submitted while held is ordered as a whole when the hold ends: the
// workers then take the largest queued work first instead of whatever
// arrived first.
void Server::hold(bool on) {
    {
        std::lock_guard<std::mutex> l(mu_);
        held_ = on;
    }
    if (!on) cv_work_.notify_all();
}

size_t Server::cached_factors() {
    std::lock_guard<std::mutex> l(mu_);
    return cache_.size();
}

// Admission: reserve `bytes` within the budget, evicting idle cached factors
// (least recently used first) when needed; waits while other requests hold
// the memory.  `pinned` is what the caller already holds (reservation and
// factors in use); false when bytes + pinned can never fit.  Called with mu_
// locked; returns with it locked.
bool Server::reserve(size_t bytes, size_t pinned) {
    if (bytes + pinned > budget_) return false;
    std::unique_lock<std::mutex> l(mu_, std::adopt_lock);
    for (;;) {
        while (reserved_ + cache_bytes_ + bytes > budget_ && !lru_.empty()) {
            // evict the least recently used factor nobody is using
            bool evicted = false;
            for (auto it = lru_.begin(); it != lru_.end(); ++it) {
                auto c = cache_.find(*it);
                if (c != cache_.end() && c->second.use_count() == 1) {
                    cache_bytes_ -= c->second->bytes;
                    cache_.erase(c);
                    lru_.erase(it);
                    evicted = true;
                    break;
                }
            }
            if (!evicted) break;
        }
        if (reserved_ + cache_bytes_ + bytes <= budget_ || stop_) break;
        cv_mem_.wait(l);
    }
    reserved_ += bytes;
    peak_ = std::max(peak_, reserved_ + cache_bytes_);
    l.release();
    return true;
}

void Server::release(size_t bytes) {
    reserved_ -= bytes;
    cv_mem_.notify_all();
}

std::shared_ptr<Server::Factor> Server::cache_get(const std::string &key) {
    auto it = cache_.find(key);
    if (it == cache_.end()) return nullptr;
    lru_.remove(key);
    lru_.push_back(key);
    return it->second;
}

static std::string factor_key(const Request &r, int prec) {
    std::ostringstream ks;
    ks.precision(17);
    ks << r.mesh.key() << '|' << r.freq << '|' << r.formulation << '|' << r.alpha << '|' << prec << '|' << r.quad;
    if (r.solver == MLFMA_OP) ks << "|mlfma|" << r.mlfma_box;
    if (r.solver == FASTWT_OP) ks << "|fastwt|" << r.mlfma_box << '|' << r.fastwt_beta;
    return ks.str();
}

// Device bytes an MLFMA operator is reserved before it is built (near
// matrix, patterns and tables; the 122880-unknown sphere at 1.5 GHz holds
// 1.09 GB, 9 KB per unknown); the cache then accounts the operator's own
// figure.
static size_t mlfma_estimate(int n) { return (size_t)n * 32 * 1024 + (64u << 20); }

// Continuous batching: a worker that takes a request also takes the queued
// requests that share its factor (mesh, frequency, formulation, alpha,
// precision, quadrature), up to batch_max_rhs_ right-hand sides in total, and
// solves them with one excitation fill, one refinement and one far-field pass
// on stacked right-hand sides.  Requests that arrive while a batch runs join
// the next one.
void Server::worker(int w) {
    emfa_ctx *ctx = ctxs_[w];
    for (;;) {
        // groups[0]: the first queued request and the queued requests that
        // share its factor; groups[1..]: requests for other new FP32 factors
        // of the same size, built together with groups[0]'s (factor batch)
        std::vector<std::vector<Job>> groups(1);
        groups.reserve((size_t)std::max(1, factor_batch_max_) + 1);  // references into groups[0] stay valid
        bool factor_batch = false;
        {
            std::unique_lock<std::mutex> l(mu_);
            cv_work_.wait(l, [&] { return stop_ || (!held_ && !queue_.empty()); });
            if (stop_) return;
            std::vector<Job> &jobs = groups[0];
            // the queued request with the most work first (a new dense factor
            // costs ~n^3, a resident one ~n^2 per right-hand side), unless
            // the oldest one has waited max_wait_ms_
            const bool overdue = ms_between(queue_.front().arrival, clock_t_::now()) >= max_wait_ms_;
            if (largest_first_ && !overdue) {
                auto cost = [&](const Job &j) {
                    const double n = spec_unknowns(j.req.mesh);
                    const std::string k = factor_key(j.req, j.req.precision);
                    const bool fresh = j.req.solver == DENSE && !cache_.count(k) && !building_.count(k);
                    return fresh ? n * n * n : n * n * (double)(j.req.angles.size() / 3 + 1);
                };
                auto best = queue_.begin();
                double bc = cost(*best);
                for (auto it = std::next(queue_.begin()); it != queue_.end(); ++it) {
                    const double c = cost(*it);
                    if (c > bc) {
                        bc = c;
                        best = it;
                    }
                }
                if (best != queue_.begin()) {
                    Job j = std::move(*best);
                    queue_.erase(best);
                    queue_.push_front(std::move(j));
                }
            }
            jobs.push_back(std::move(queue_.front()));
            queue_.pop_front();
            int rhs = (int)(jobs[0].req.angles.size() / 3);
            const std::string key = factor_key(jobs[0].req, jobs[0].req.precision);
            if (batch_max_rhs_ > 0 && rhs < batch_max_rhs_) {
                for (auto it = queue_.begin(); it != queue_.end();) {
                    const int m = (int)(it->req.angles.size() / 3);
                    if (m > 0 && rhs + m <= batch_max_rhs_ && factor_key(it->req, it->req.precision) == key) {
                        rhs += m;
                        jobs.push_back(std::move(*it));
                        it = queue_.erase(it);
                    } else {
                        ++it;
                    }
                }
            }
            const Request &r0 = jobs[0].req;
            if (factor_batch_max_ > 1 && r0.solver == DENSE && r0.precision == EMFA_F32 && !cache_.count(key) &&
                !building_.count(key)) {
                const int n0 = spec_unknowns(r0.mesh);
                std::map<std::string, size_t> group_of{{key, 0}};
                std::vector<int> group_rhs{rhs};
                std::deque<Job> rest;
                for (auto &q : queue_) {
                    const int m = (int)(q.req.angles.size() / 3);
                    bool taken = false;
                    if (q.req.solver == DENSE && q.req.precision == EMFA_F32 && m > 0 &&
                        spec_unknowns(q.req.mesh) == n0) {
                        const std::string k = factor_key(q.req, EMFA_F32);
                        auto g = group_of.find(k);
                        if (g != group_of.end() && g->second > 0 && group_rhs[g->second] + m <= batch_max_rhs_) {
                            group_rhs[g->second] += m;
                            groups[g->second].push_back(std::move(q));
                            taken = true;
                        } else if (g == group_of.end() && (int)groups.size() < factor_batch_max_ && !cache_.count(k) &&
                                   !building_.count(k)) {
                            group_of[k] = groups.size();
                            group_rhs.push_back(m);
                            groups.emplace_back();
                            groups.back().push_back(std::move(q));
                            taken = true;
                        }
                    }
                    if (!taken) rest.push_back(std::move(q));
                }
                queue_.swap(rest);
                if (groups.size() > 1) {
                    factor_batch = true;
                    for (auto &g : group_of) building_.insert(g.first);
                }
            }
            for (auto &g : groups) busy_ += (int)g.size();
            ++busy_workers_;
            rebalance_locked();
        }
        const auto start = clock_t_::now();
        if (factor_batch) {
            prebuild(ctx, groups);
            // the other groups go back to the front of the queue: their
            // factors are resident now, so idle workers answer them meanwhile
            std::lock_guard<std::mutex> l(mu_);
            for (size_t g = groups.size() - 1; g >= 1; --g) {
                busy_ -= (int)groups[g].size();
                for (size_t i = groups[g].size(); i-- > 0;) {
                    groups[g][i].fresh = i == 0;
                    queue_.push_front(std::move(groups[g][i]));
                }
            }
            groups.resize(1);
            cv_work_.notify_all();
        }
        for (auto &jobs : groups) {
            std::vector<Result> res(jobs.size());
            for (size_t i = 0; i < jobs.size(); ++i) {
                res[i].id = jobs[i].id;
                res[i].batch_size = (int)jobs.size();
            }
            try {
                run(ctx, jobs, res, factor_batch || jobs[0].fresh);
            } catch (const StatusError &e) {
                for (auto &x : res) {
                    x.status = e.code;
                    x.error = e.what();
                }
            } catch (const std::exception &e) {
                for (auto &x : res) {
                    x.status = EMFA_E_INVALID;
                    x.error = e.what();
                }
            }
            auto end = clock_t_::now();
            {
                std::lock_guard<std::mutex> l(mu_);
                busy_ -= (int)jobs.size();
                t1_ = end;
                batched_ += (int64_t)jobs.size() - 1;
                for (size_t i = 0; i < jobs.size(); ++i) {
                    Result &x = res[i];
                    x.queue_ms = ms_between(jobs[i].arrival, start);
                    x.service_ms = ms_between(start, end);
                    x.latency_ms = ms_between(jobs[i].arrival, end);
                    if (x.status == EMFA_OK) {
                        lat_.push_back(x.latency_ms);
                        rhs_done_ += x.nrhs;
                        if (x.cache_hit) ++hits_;
                    } else {
                        ++failed_;
                    }
                    done_[x.id] = std::move(x);
                }
            }
            cv_done_.notify_all();
        }
        std::lock_guard<std::mutex> l(mu_);
        --busy_workers_;
        rebalance_locked();
    }
}

// Factor batch: the groups' meshes are built and filled one by one, then all
// matrices are factored with one emfa_lu_factor_batch call and published
// together.  The reservation covers every factor and its LU workspace while
// they are built; the factors' bytes then move to the cache.  On any failure
// the keys are released and run() builds each factor on its own.
void Server::prebuild(emfa_ctx *ctx, const std::vector<std::vector<Job>> &groups) {
    const size_t count = groups.size();
    std::vector<std::string> keys(count);
    for (size_t g = 0; g < count; ++g) keys[g] = factor_key(groups[g][0].req, EMFA_F32);
    size_t held = 0;
    bool published = false;
    try {
        std::vector<std::vector<double>> xyz(count);
        std::vector<std::vector<int>> tri(count), rwg(count);
        std::vector<int> n(count, 0);
        for (size_t g = 0; g < count; ++g) {
            build_mesh(groups[g][0].req.mesh, xyz[g], tri[g], rwg[g], n[g]);
            if (n[g] != n[0]) throw StatusError(EMFA_E_INVALID, "factor batch: sizes differ");
        }
        const size_t need = count * footprint(n[0], 0, EMFA_F32, false);
        {
            mu_.lock();
            bool ok = reserve(need, 0);
            mu_.unlock();
            if (!ok) throw StatusError(EMFA_E_OOM, "factor batch exceeds the HBM budget");
        }
        held = need;
        std::vector<std::shared_ptr<Factor>> fs(count);
        std::vector<const emfa_mat *> Zs(count);
        for (size_t g = 0; g < count; ++g) {
            fs[g] = make_factor(ctx, groups[g][0].req, EMFA_F32, xyz[g], tri[g], rwg[g], n[g]);
            Zs[g] = fs[g]->Z;
        }
        std::vector<emfa_lu *> lus(count, nullptr);
        check(emfa_lu_factor_batch(ctx, Zs.data(), (int)count, EMFA_F32, lus.data()), "LU batch");
        int fallbacks = 0;
        for (size_t g = 0; g < count; ++g) {
            fs[g]->lu = lus[g];
            int inter = 1, fallback = 0;
            check(emfa_lu_get_pivoting(lus[g], &inter, &fallback), "LU");
            fallbacks += fallback;
        }
        // other workers use the factors on their own streams: finish them first
        check(emfa_ctx_sync(ctx), "sync");
        {
            std::lock_guard<std::mutex> l(mu_);
            for (size_t g = 0; g < count; ++g) {
                cache_[keys[g]] = fs[g];
                lru_.push_back(keys[g]);
                cache_bytes_ += fs[g]->bytes;
                reserved_ -= fs[g]->bytes;
                held -= fs[g]->bytes;
                building_.erase(keys[g]);
            }
            release(held);
            held = 0;
            pivot_fallbacks_ += fallbacks;
            ++factor_batches_;
            batch_factors_ += (int64_t)count;
            published = true;
        }
    } catch (const std::exception &) {
        // fall through: run() builds what is missing one factor at a time
    }
    if (!published) {
        std::lock_guard<std::mutex> l(mu_);
        release(held);
        for (auto &k : keys) building_.erase(k);
    }
    cv_build_.notify_all();
}

// Mesh on the device and the filled matrix of a new factor (no LU yet); its
// bytes count Z, the factor and the diagonal-block inverses.
std::shared_ptr<Server::Factor> Server::make_factor(emfa_ctx *ctx, const Request &r, int prec,
                                                    const std::vector<double> &xyz, const std::vector<int> &tri,
                                                    const std::vector<int> &rwg, int n) {
    const double k = 2.0 * 3.14159265358979323846 * r.freq / 299792458.0;
    auto nf = std::make_shared<Factor>();
    nf->n = n;
    nf->prec = prec;
    nf->k = k;
    check(emfa_mesh_upload(ctx, xyz.data(), (int)(xyz.size() / 3), tri.data(), (int)(tri.size() / 3), rwg.data(), n,
                           &nf->mesh),
          "mesh");
    if (r.solver == FASTWT_OP) {
        // FastWT::init .. calculateNearZ (EMSolver/FastWT.cpp:26-70,167-784): the
        // half-RWG list of the mesh, then the IP-SIE operator on the device
        const int nv = (int)(xyz.size() / 3), nt = (int)(tri.size() / 3);
        int nh = 0, n_nc = 0, n_lines = 0;
        check(emfa_contour_build(nv, xyz.data(), nt, tri.data(), nullptr, nullptr, 0, &nh, nullptr, nullptr, nullptr,
                                 0, &n_nc, &n_lines),
              "contour edges");
        std::vector<int> h(6 * (size_t)nh), lptr(n_nc + 1), ltri(std::max(1, n_lines));
        std::vector<double> len(nh), seg(6 * (size_t)std::max(1, n_lines));
        check(emfa_contour_build(nv, xyz.data(), nt, tri.data(), h.data(), len.data(), nh, &nh, lptr.data(),
                                 ltri.data(), seg.data(), (int)ltri.size(), &n_nc, &n_lines),
              "contour edges");
        emfa_mlfma_params mp{EMFA_CFIE, k, r.alpha, r.mlfma_box, 0};
        check(emfa_fastwt_create(ctx, nf->mesh, nh, h.data(), len.data(), n_nc, lptr.data(), ltri.data(), seg.data(),
                                 &mp, r.fastwt_beta, &nf->op),
              "FastWT setup");
        emfa_mlfma_info info{};
        check(emfa_mlfma_get_info(nf->op, &info), "FastWT info");
        nf->bytes = info.bytes;
        nf->n = nh;
        nf->fastwt = true;
        return nf;
    }
    if (r.solver == MLFMA_OP) {
        // MLFMA::init .. calculateNearZ (EMSolver/MLFMA.cpp:23-65, 688-1422) on the device
        emfa_mlfma_params mp{r.formulation, k, r.alpha, r.mlfma_box, 0};
        check(emfa_mlfma_create(ctx, nf->mesh, &mp, &nf->op), "MLFMA setup");
        emfa_mlfma_info info{};
        check(emfa_mlfma_get_info(nf->op, &info), "MLFMA info");
        nf->bytes = info.bytes;
        return nf;
    }
    check(emfa_mat_create(ctx, n, n, prec, &nf->Z), "Z");
    emfa_fill_params fp{r.formulation, r.quad, k, r.alpha, 0.0, 0.0, 0};
    check(emfa_fill(ctx, nf->mesh, &fp, nf->Z), "fill");
    // Z, the factor, the diagonal-block inverses and (FP32, EMFA_SOLVE_GH=1) the scaled off-diagonal blocks
    const bool gh = prec == EMFA_F32 && env_flag("EMFA_SOLVE_GH", false);
    nf->bytes = (gh ? 3 : 2) * (size_t)n * n * (prec == EMFA_F32 ? 8 : 16) +
                2 * (((size_t)n + 255) / 256) * 256 * 256 * (prec == EMFA_F32 ? 8 : 16);
    return nf;
}

// The factor of (request, prec) from the cache, or computed and published.
// A factor another worker is computing right now is waited for, not computed
// twice.  The request's vectors are reserved here too (`vec_bytes`); the
// bytes this call adds to the request's reservation are added to *held.
std::shared_ptr<Server::Factor> Server::acquire(emfa_ctx *ctx, const Request &r, int prec, int nrhs, bool *hit,
                                                size_t *held, size_t pinned) {
    const std::string key = factor_key(r, prec);
    std::shared_ptr<Factor> f;
    {
        std::unique_lock<std::mutex> l(mu_);
        for (;;) {
            f = cache_get(key);
            if (f || !building_.count(key)) break;
            cv_build_.wait(l);
        }
        if (!f) building_.insert(key);
    }
    struct BuildGuard {
        Server *s;
        const std::string &key;
        bool on;
        ~BuildGuard() {
            if (!on) return;
            {
                std::lock_guard<std::mutex> l(s->mu_);
                s->building_.erase(key);
            }
            s->cv_build_.notify_all();
        }
    } bguard{this, key, !f};
    *hit = (bool)f;
    std::vector<double> xyz;
    std::vector<int> tri, rwg;
    int n = 0;
    if (f)
        n = f->n;
    else
        build_mesh(r.mesh, xyz, tri, rwg, n);
    const bool mlfma = r.solver == MLFMA_OP || r.solver == FASTWT_OP;
    // the IP-SIE operator has up to three half-RWGs per triangle
    const int nvec = r.solver == FASTWT_OP && !f ? 3 * (int)(tri.size() / 3) : n;
    const size_t need = mlfma ? (size_t)nvec * nrhs * 16 * 4 + (f ? 0 : mlfma_estimate(nvec))
                              : footprint(n, nrhs, prec, (bool)f);
    {
        mu_.lock();
        bool ok = reserve(need, *held + pinned);
        mu_.unlock();
        if (!ok) throw StatusError(EMFA_E_OOM, "request footprint exceeds the HBM budget");
    }
    *held += need;
    if (f) return f;

    auto nf = make_factor(ctx, r, prec, xyz, tri, rwg, n);
    if (!mlfma) {
        check(emfa_lu_factor(ctx, nf->Z, prec, &nf->lu), "LU");
        int inter = 1, fallback = 0;
        check(emfa_lu_get_pivoting(nf->lu, &inter, &fallback), "LU");
        if (fallback) {
            std::lock_guard<std::mutex> l(mu_);
            ++pivot_fallbacks_;
        }
    }
    // other workers use the factor on their own streams: finish it first
    check(emfa_ctx_sync(ctx), "sync");
    {
        std::lock_guard<std::mutex> l(mu_);
        cache_[key] = nf;
        lru_.push_back(key);
        // the factor's bytes move from this request's reservation to the
        // cache (an MLFMA operator larger than its estimate moves what the
        // request holds and is counted in full by the cache)
        const size_t moved = std::min(nf->bytes, *held);
        cache_bytes_ += nf->bytes;
        reserved_ -= moved;
        *held -= moved;
        building_.erase(key);
        bguard.on = false;
    }
    cv_build_.notify_all();
    return nf;
}

void Server::run(emfa_ctx *ctx, const std::vector<Job> &jobs, std::vector<Result> &res, bool fresh) {
    const Request &r = jobs[0].req;  // the batch shares mesh, frequency, formulation, precision
    std::vector<int> first(jobs.size() + 1, 0);  // column range of each job in the stacked RHS
    for (size_t i = 0; i < jobs.size(); ++i) {
        const int m = (int)(jobs[i].req.angles.size() / 3);
        if (m <= 0) throw StatusError(EMFA_E_INVALID, "request without right-hand sides");
        first[i + 1] = first[i] + m;
    }
    const int nrhs = first.back();
    const double k = 2.0 * 3.14159265358979323846 * r.freq / 299792458.0;

    struct Guard {
        Server *s;
        size_t b = 0;
        ~Guard() {
            std::lock_guard<std::mutex> l(s->mu_);
            s->release(b);
        }
    } guard{this};
    bool hit = false;
    std::shared_ptr<Factor> f = acquire(ctx, r, r.precision, nrhs, &hit, &guard.b, 0);
    const int n = f->n;

    // excitation (EMSolver/EM.cpp:29-41 per incidence), monostatic directions -k
    std::vector<double> inc(9 * (size_t)nrhs), dirs(3 * (size_t)nrhs);
    std::vector<int> cols(nrhs);
    const double d2r = 3.14159265358979323846 / 180.0;
    for (size_t i = 0; i < jobs.size(); ++i)
        for (int jj = 0; jj < first[i + 1] - first[i]; ++jj) {
            const std::vector<double> &ang = jobs[i].req.angles;
            const int j = first[i] + jj;
            double th = ang[3 * jj] * d2r, ph = ang[3 * jj + 1] * d2r, po = ang[3 * jj + 2] * d2r;
            double kv[3] = {-std::sin(th) * std::cos(ph), -std::sin(th) * std::sin(ph), -std::cos(th)};
            double ev[3] = {-std::cos(po) * std::cos(th) * std::cos(ph) - std::sin(po) * std::sin(ph),
                            -std::cos(po) * std::cos(th) * std::sin(ph) + std::sin(po) * std::cos(ph),
                            std::cos(po) * std::sin(th)};
            double hv[3] = {kv[1] * ev[2] - kv[2] * ev[1], kv[2] * ev[0] - kv[0] * ev[2],
                            kv[0] * ev[1] - kv[1] * ev[0]};
            for (int c = 0; c < 3; ++c) {
                inc[9 * j + c] = kv[c];
                inc[9 * j + 3 + c] = ev[c];
                inc[9 * j + 6 + c] = hv[c];
                dirs[3 * j + c] = -kv[c];
            }
            cols[j] = j;
        }
    emfa_mat *B = nullptr, *X = nullptr;
    struct MatGuard {
        emfa_mat **a, **b;
        ~MatGuard() {
            if (*a) emfa_mat_destroy(*a);
            if (*b) emfa_mat_destroy(*b);
        }
    } mg{&B, &X};
    check(emfa_mat_create(ctx, n, nrhs, EMFA_F64, &B), "B");
    if (f->fastwt)
        check(emfa_fastwt_rhs(ctx, f->op, f->mesh, B, nrhs, inc.data()), "rhs");
    else
        check(emfa_rhs_fill(ctx, f->mesh, r.formulation, k, r.alpha, nrhs, inc.data(), B), "rhs");
    check(emfa_mat_create(ctx, n, nrhs, EMFA_F64, &X), "X");
    int iters = 0;
    double relres = 0;
    bool fallback = false;
    if (f->op) {
        // GMRES(30) on the resident operator (the C5 solve), X from zero
        emfa_gmres_report rep{};
        const int restarts = std::max(1, (r.iter_max + 29) / 30);
        emfa_status st = emfa_mlfma_gmres(ctx, f->op, B, X, 30, r.iter_tol, restarts, &rep);
        iters = rep.iterations;
        relres = rep.relres;
        check(st, f->fastwt ? "FastWT GMRES" : "MLFMA GMRES");
    } else if (r.precision == EMFA_F64) {
        check(emfa_mat_copy(X, B), "copy");
        check(emfa_lu_solve(ctx, f->lu, X), "solve");
    } else {
        emfa_refine_report rep{};
        emfa_status st = emfa_refine(ctx, f->Z, f->lu, B, X, refine_tol_, refine_maxit_, &rep);
        iters = rep.iterations;
        relres = rep.relres_final;
        if (st == EMFA_E_NOCONV) {
            // the FP32 factor is too inaccurate for this matrix: solve with the FP64 factor
            bool hit64 = false;
            std::shared_ptr<Factor> f64 = acquire(ctx, r, EMFA_F64, nrhs, &hit64, &guard.b, f->bytes);
            check(emfa_mat_copy(X, B), "copy");
            check(emfa_lu_solve(ctx, f64->lu, X), "solve");
            fallback = true;
        } else {
            check(st, "refine");
        }
    }
    std::vector<double> rcs(nrhs);
    if (f->fastwt) {
        for (int j = 0; j < nrhs; ++j)
            check(emfa_fastwt_rcs(ctx, f->op, f->mesh, X, j, 1, dirs.data() + 3 * (size_t)j, rcs.data() + j), "rcs");
    } else {
        check(emfa_rcs(ctx, f->mesh, k, X, nrhs, dirs.data(), cols.data(), rcs.data()), "rcs");
    }
    std::vector<double> cur;
    if (keep_current_) {
        cur.resize(2 * (size_t)n * nrhs);
        check(emfa_mat_download(X, cur.data()), "download");
    }
    check(emfa_ctx_sync(ctx), "sync");
    for (size_t i = 0; i < jobs.size(); ++i) {
        Result &x = res[i];
        const int c0 = first[i], m = first[i + 1] - first[i];
        // batch mates did not pay for the factor: they count as reuse
        x.cache_hit = (hit && !fresh) || i > 0;
        x.unknowns = n;
        x.nrhs = m;
        x.refine_iterations = iters;
        x.relres = relres;
        x.fp64_fallback = fallback;
        x.rcs_db.assign(rcs.begin() + c0, rcs.begin() + c0 + m);
        if (keep_current_) {
            x.current.resize(2 * (size_t)n * m);
            for (int row = 0; row < n; ++row)
                std::copy(cur.begin() + 2 * ((size_t)row * nrhs + c0), cur.begin() + 2 * ((size_t)row * nrhs + c0 + m),
                          x.current.begin() + 2 * (size_t)row * m);
        }
    }
}

Stats Server::stats() {
    std::lock_guard<std::mutex> l(mu_);
    Stats s;
    s.completed = (int64_t)lat_.size();
    s.failed = failed_;
    s.cache_hits = hits_;
    s.peak_bytes = peak_;
    s.pivot_fallbacks = pivot_fallbacks_;
    s.batched = batched_;
    s.factor_batches = factor_batches_;
    s.batch_factors = batch_factors_;
    if (!lat_.empty()) {
        std::vector<double> v = lat_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            double idx = p * (v.size() - 1);
            size_t lo = (size_t)std::floor(idx), hi = std::min(v.size() - 1, lo + 1);
            return v[lo] + (idx - lo) * (v[hi] - v[lo]);
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_ms = sum / v.size();
        s.max_ms = v.back();
        s.wall_s = std::chrono::duration<double>(t1_ - t0_).count();
        if (s.wall_s > 0) {
            s.solves_per_s = s.completed / s.wall_s;
            s.rhs_per_s = rhs_done_ / s.wall_s;
        }
    }
    return s;
}

void Server::reset_stats() {
    std::lock_guard<std::mutex> l(mu_);
    lat_.clear();
    failed_ = hits_ = rhs_done_ = pivot_fallbacks_ = batched_ = factor_batches_ = batch_factors_ = 0;
    have_t0_ = false;
    peak_ = reserved_ + cache_bytes_;
}

}  // namespace serve
