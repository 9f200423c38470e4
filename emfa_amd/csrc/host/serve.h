// Continuous-batching solve server on one device.
//
// Requests (mesh, frequency, formulation, precision, right-hand sides) are
// queued on arrival; worker threads, each with its own device context and
// stream, take them as HBM budget and a worker slot free up, so the work of
// several requests overlaps on the device.  Factors (Z and its LU) stay
// resident in a cache keyed by (mesh, frequency, formulation, precision,
// quadrature); a repeat request only fills its excitation, refines/solves and
// evaluates the far field.  Every request records its queueing and service
// time; the server reports latency percentiles and throughput.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "emfa_hip.h"

namespace serve {

enum MeshKind { ICOSPHERE = 0, PLATE = 1, USER = 2 };
enum SolverKind { DENSE = 0, MLFMA_OP = 1, FASTWT_OP = 2 };

struct MeshSpec {
    int kind = ICOSPHERE;
    int subdiv = 2;               // icosphere subdivisions, or plate cells per side
    double ax = 1, ay = 1, az = 1;  // semi-axes (icosphere) / side length in ax (plate)
    // USER: the caller's geometry (copied at submit, shared by the requests
    // that name it), its content hash and RWG count
    std::shared_ptr<const std::vector<double>> xyz;
    std::shared_ptr<const std::vector<int>> tri;
    uint64_t hash = 0;
    int nrwg = 0;
    std::string key() const;
};

// 64-bit FNV-1a over the vertex coordinates and triangle ids: the cache key of a caller's mesh
uint64_t mesh_hash(const std::vector<double> &xyz, const std::vector<int> &tri);

struct Request {
    MeshSpec mesh;
    int formulation = EMFA_EFIE;  // EMFA_EFIE or EMFA_CFIE
    double alpha = 1.0;
    double freq = 3e8;
    int precision = EMFA_F64;     // EMFA_F64: FP64 LU; EMFA_F32: FP32 LU + FP64 refinement
    int quad = 3;
    std::vector<double> angles;   // nrhs x 3 (theta, phi, pole) degrees; monostatic RCS per column
    // MLFMA_OP / FASTWT_OP: the multilevel fast multipole operator (or the
    // IP-SIE operator on half-RWGs) stays resident instead of Z and its LU;
    // each request is solved by GMRES(30) on it
    int solver = DENSE;
    double mlfma_box = 0.25;      // finest box edge, wavelengths
    double iter_tol = 1e-4;       // GMRES relative residual
    int iter_max = 600;           // GMRES iterations at most
    // FASTWT_OP: the IP-SIE operator on the mesh's half-RWGs stays resident;
    // beta = stabilization factor x |log10(average size)|
    double fastwt_beta = 0.1 * 1.124764;
};

struct Result {
    int64_t id = -1;
    int status = EMFA_OK;
    std::string error;
    int unknowns = 0, nrhs = 0;
    bool cache_hit = false;
    int refine_iterations = 0;
    double relres = 0;
    bool fp64_fallback = false;  // refinement did not converge; solved with the FP64 factor
    double queue_ms = 0, service_ms = 0, latency_ms = 0;
    int batch_size = 1;           // requests solved together with this one (same factor, one refinement)
    std::vector<double> rcs_db;   // monostatic RCS per right-hand side
    std::vector<double> current;  // optional: solution (interleaved complex, n x nrhs)
};

struct Stats {
    int64_t completed = 0, failed = 0, cache_hits = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
    double wall_s = 0;  // first arrival to last completion
    double solves_per_s = 0, rhs_per_s = 0;
    size_t peak_bytes = 0;
    int64_t pivot_fallbacks = 0;  // FP32 factors redone with interchanges
    int64_t batched = 0;          // requests that joined another request's batch
    int64_t factor_batches = 0;   // factorisations that built several new factors together
    int64_t batch_factors = 0;    // factors built that way
};

class Server {
  public:
    Server(int device, int workers, size_t hbm_budget, bool keep_current = false);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    int64_t submit(const Request &r);
    // blocks until request `id` is done (id < 0: any); false if none pending
    bool wait(int64_t id, Result *out);
    void drain();  // wait until the queue is empty and all workers are idle (releases a hold)
    void hold(bool on);  // while on, workers take no queued request
    Stats stats();
    void reset_stats();
    size_t cached_factors();

    // footprint model used for admission (bytes a request needs while it runs)
    static size_t footprint(int n, int nrhs, int precision, bool factor_cached);

  private:
    struct Factor;
    struct Job {
        int64_t id;
        Request req;
        std::chrono::steady_clock::time_point arrival;
        bool fresh = false;  // its factor was built by a factor batch for it (not a cache hit)
    };
    void worker(int w);
    // solve the jobs of one batch (same factor key) with one refinement;
    // fresh: the factor was just built for them (not a cache hit)
    void run(emfa_ctx *ctx, const std::vector<Job> &jobs, std::vector<Result> &res, bool fresh = false);
    // factor the groups' matrices (distinct keys, one size, FP32) with one
    // batched factorisation and publish them; on failure the groups fall
    // back to building their factors one by one
    void prebuild(emfa_ctx *ctx, const std::vector<std::vector<Job>> &groups);
    std::shared_ptr<Factor> make_factor(emfa_ctx *ctx, const Request &r, int prec, const std::vector<double> &xyz,
                                        const std::vector<int> &tri, const std::vector<int> &rwg, int n);
    std::shared_ptr<Factor> cache_get(const std::string &key);
    std::shared_ptr<Factor> acquire(emfa_ctx *ctx, const Request &r, int prec, int nrhs, bool *hit, size_t *held,
                                    size_t pinned);
    bool reserve(size_t bytes, size_t pinned);  // admission; false when the request can never fit
    void release(size_t bytes);

    int device_;
    size_t budget_;
    // Relative residual the FP64 refinement of an FP32 factor stops at.  The
    // FP32-filled matrix itself differs from the FP64 one by ~6e-8 relative
    // (half an FP32 ulp), so iterating below ~1e-8 cannot bring a result
    // closer to the FP64 answer; EMFA_SERVE_REFINE_TOL overrides it.
    double refine_tol_ = 1e-8;
    int refine_maxit_ = 8;
    bool keep_current_;
    std::vector<std::thread> threads_;
    std::vector<emfa_ctx *> ctxs_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_, cv_mem_, cv_build_;
    std::deque<Job> queue_;
    std::map<int64_t, Result> done_;
    std::set<int64_t> pending_;  // submitted and not yet returned by wait()
    int64_t next_id_ = 0;
    int busy_ = 0;
    int busy_workers_ = 0;   // workers running a request (mu_)
    bool dynamic_share_ = true;
    // EMFA_SERVE_LARGEST_FIRST (default on): take the queued request with the
    // most work, or the oldest once it has waited EMFA_SERVE_MAX_WAIT_MS
    bool largest_first_ = true;
    double max_wait_ms_ = 1000.0;
    void rebalance_locked();  // CU shares of all contexts from busy_workers_
    bool stop_ = false;
    bool held_ = false;
    size_t reserved_ = 0, peak_ = 0;
    // factor cache (LRU)
    std::map<std::string, std::shared_ptr<Factor>> cache_;
    std::list<std::string> lru_;
    size_t cache_bytes_ = 0;
    std::set<std::string> building_;  // factors being computed by some worker
    // stats
    std::vector<double> lat_;
    int64_t failed_ = 0, hits_ = 0, rhs_done_ = 0, pivot_fallbacks_ = 0, batched_ = 0, batch_factors_ = 0;
    int batch_max_rhs_ = 256;  // right-hand sides one batch may stack (EMFA_SERVE_BATCH_RHS, 0: no batching)
    // distinct new FP32 factors of one size built together (EMFA_SERVE_FACTOR_BATCH, <= 1: one at a time)
    int factor_batch_max_ = 8;
    int64_t factor_batches_ = 0;  // batched factorisations run (stats)
    bool have_t0_ = false;
    std::chrono::steady_clock::time_point t0_, t1_;
};

}  // namespace serve
