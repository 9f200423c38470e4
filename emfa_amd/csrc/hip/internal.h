// Internal declarations of the gfx950 MoM library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/emfa_hip.h"

namespace emfa {

void set_error(const char *fmt, ...);

#define EMFA_HIP_CHECK(expr)                                                                 \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) {                                                              \
            ::emfa::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
            return EMFA_E_HIP;                                                               \
        }                                                                                    \
    } while (0)

#define EMFA_REQUIRE(cond, code, ...)         \
    do {                                      \
        if (!(cond)) {                        \
            ::emfa::set_error(__VA_ARGS__);   \
            return (code);                    \
        }                                     \
    } while (0)

// Per-family timing with HIP events on the context stream.
struct ProfFamily {
    double ms = 0;
    long long launches = 0;
    double work = 0;  // algorithmic flops or bytes attributed to the family
};

struct PendingEvent {
    std::string family;
    hipEvent_t a, b;
    double work;
};

// A batch of same-size matrices factored by one launch sequence: kernels take
// matrix z on blockIdx.z.  Pointer arithmetic on the host starts from matrix
// 0 (mat0); a kernel moves a matrix pointer onto matrix z's base (same offset)
// and a scratch pointer by z slabs (every matrix owns one scratch slab of the
// same layout).  The default value is no batch: both moves are identities.
struct Batch {
    const char *const *mat = nullptr;  // device array: base of matrix z
    const char *mat0 = nullptr;        // base of matrix 0
    long long scratch = 0;             // bytes between consecutive scratch slabs
    int count = 1;
};

template <typename T>
__device__ __forceinline__ T *batch_mat(const Batch &b, T *p, int z) {
    if (!b.mat) return p;
    return (T *)(b.mat[z] + ((const char *)p - b.mat0));
}

template <typename T>
__device__ __forceinline__ T *batch_scr(const Batch &b, T *p, int z) {
    return (T *)((char *)p + (long long)z * b.scratch);
}

// matrix / scratch pointers of the matrix on blockIdx.z
template <typename T>
__device__ __forceinline__ T *batch_mat(const Batch &b, T *p) {
    return batch_mat(b, p, (int)blockIdx.z);
}

template <typename T>
__device__ __forceinline__ T *batch_scr(const Batch &b, T *p) {
    return batch_scr(b, p, (int)blockIdx.z);
}

}  // namespace emfa

struct emfa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;  // second stream (LU look-ahead), default priority
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int num_cu = 0;
    // compute units this context may count on while other contexts share
    // the device (0: all); a solve server changes it from its own thread as
    // workers start and finish, so it is read at each launch
    std::atomic<int> cu_share{0};
    // bound on the workgroups of a panel whose workgroups wait on each other
    // (0: cu_share); fixed while contexts share the device, so that such
    // panels of all contexts fit on the device together
    int panel_cap = 0;
    // LU parameters (0 / -1: built-in choice); set by emfa_lu_set_params or the autotuner
    int lu_nb = 0, panel_g = 0, panel_x = -1;  // panel_x 2: FP32 panels without interchanges
    int force_pivoting = 0;                    // set while a grown pivot-free factor is redone
    int lu_update_mode = 0;  // FP32 LU trailing update: 0 FP32 MFMA, 1 three-term bf16 split
    size_t bytes = 0;
    bool prof = false;
    std::map<std::string, emfa::ProfFamily> fam;
    std::vector<emfa::PendingEvent> pending;
    std::vector<hipEvent_t> event_pool;
    // device scratch owned by the context (grown on demand)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int *flags = nullptr;  // [0] spin-timeout flag, [1] singular info, [2] pivot-free growth
    emfa::Batch batch;     // set while a batch of factorisations runs (the LU helpers launch over it)
    // LU launch sequences captured as HIP graphs (emfa_ctx_set_graphs): one
    // executable graph per launch configuration, replayed for every batch of
    // that configuration; the matrices are reached through `bmat`, a device
    // array of their bases whose address stays fixed while its capacity does
    bool lu_graphs = false;
    std::map<std::string, hipGraphExec_t> lu_graph;
    void *bmat = nullptr;
    size_t bmat_cap = 0;  // entries
    long long graph_captures = 0, graph_launches = 0;
    hipStream_t cap_main = nullptr, cap_aux = nullptr;  // streams the graphs are captured on
};

struct emfa_mesh {
    emfa_ctx *ctx = nullptr;
    int nv = 0, nt = 0, n = 0;
    // per-triangle geometry, SoA [field][nt]; fields: 0-2 centroid, 3-11 vertices
    // relative to the centroid (v0xyz, v1xyz, v2xyz), 12-14 unit normal, 15 area
    double *tg64 = nullptr;
    float *tg32 = nullptr;
    // triangle -> RWG: [3][nt], sign*(m+1) for the RWG whose free vertex is local vertex a, 0 if none
    int *t2r = nullptr;
    // RWG tables [n]: tp, tm, ap, am (local index of the free vertex in tp / tm)
    int *rtp = nullptr, *rtm = nullptr, *rap = nullptr, *ram = nullptr;
    double *rlen64 = nullptr;
    float *rlen32 = nullptr;
    std::vector<int> t2r_host;
    std::vector<double> xyz_host;  // vertices (nv x 3)
    std::vector<int> tri_host;     // triangles (nt x 3)
    std::vector<int> rwg_host;     // RWG list (n x 6: v1 v2 vp vm tp tm)
    size_t bytes = 0;
};

struct emfa_mat {
    emfa_ctx *ctx = nullptr;
    int rows = 0, cols = 0, prec = EMFA_F64;
    void *d = nullptr;  // interleaved complex, row-major
    size_t bytes = 0;
};

struct emfa_lu {
    emfa_ctx *ctx = nullptr;
    emfa_mat *a = nullptr;  // factor (L unit-lower below diagonal, U on/above)
    int *ipiv = nullptr;    // device, 0-based absolute row of each pivot
    int *perm = nullptr;    // device: row i of P A is row perm[i] of A (the interchanges composed)
    int info = 0;
    int interchanges = 1;   // 0: factored without row interchanges
    int pivot_fallback = 0; // 1: a pivot-free attempt grew and was redone with interchanges
    void *dinv = nullptr;   // device: inverses of the 256 x 256 diagonal blocks of L, then of U
    size_t dinv_bytes = 0;
    void *gh = nullptr;     // device (FP32 factors): L21 DL^-1 / U12 DU^-1 per 256-block column, N x N
};

namespace emfa {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kZ0 = 120.0 * kPi;

inline size_t elem_bytes(int prec) { return prec == EMFA_F32 ? 8 : 16; }

hipError_t ctx_alloc(emfa_ctx *ctx, void **p, size_t bytes);
void ctx_free(emfa_ctx *ctx, void *p, size_t bytes);
emfa_status ctx_scratch(emfa_ctx *ctx, size_t bytes, void **p);

// profiling helpers (no-ops unless ctx->prof)
void prof_begin(emfa_ctx *ctx, const char *family, double work, hipEvent_t *a, hipEvent_t *b,
                hipStream_t s = nullptr);
void prof_end(emfa_ctx *ctx, hipEvent_t a, hipEvent_t b, hipStream_t s = nullptr);
void prof_collect(emfa_ctx *ctx);

struct ProfScope {
    emfa_ctx *ctx;
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s = nullptr;
    ProfScope(emfa_ctx *c, const char *family, double work, hipStream_t st = nullptr) : ctx(c), s(st) {
        prof_begin(ctx, family, work, &a, &b, s);
    }
    ~ProfScope() { prof_end(ctx, a, b, s); }
};

// complex scalar helpers usable on host and device
template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__host__ __device__ inline cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// LU driver entry (lu.hip)
emfa_status lu_factor_inplace(emfa_ctx *ctx, emfa_mat *A, int *ipiv, int *info, int *interchanges);
// bound on the multipliers of a factorisation without interchanges
constexpr double kNoPivGrowth = 100.0;
emfa_status lu_solve_inplace(emfa_ctx *ctx, const emfa_lu *lu, emfa_mat *B);
emfa_status gemm_residual(emfa_ctx *ctx, const emfa_mat *A, const emfa_mat *X, emfa_mat *C, int negate,
                          int accumulate);
emfa_status convert_copy(emfa_ctx *ctx, const emfa_mat *src, emfa_mat *dst);
emfa_status axpy_f64(emfa_ctx *ctx, emfa_mat *X, const emfa_mat *D);  // X += D (D any prec)
// y = Z x for device vectors of n complex FP64 (K7, FP64 accumulation; gmres.hip)
emfa_status gemv(emfa_ctx *ctx, const emfa_mat *Z, const double2 *x, double2 *y);
// BiCGSTAB over an operator on device vectors (bicgstab.hip)
// work: device workspace of work_bytes >= bicgstab_work_bytes(n), or null
// for the context's scratch
emfa_status bicgstab_impl(emfa_ctx *ctx, int n, const std::function<emfa_status(const double2 *, double2 *)> &matvec,
                          const emfa_mat *B, emfa_mat *X, double tol, int max_iter, int variant,
                          emfa_bicgstab_report *rep, void *work = nullptr, size_t work_bytes = 0);
size_t bicgstab_work_bytes(int n);
// restarted GMRES over an operator on device vectors of n complex FP64 (gmres.hip)
emfa_status gmres_impl(emfa_ctx *ctx, int n, const std::function<emfa_status(const double2 *, double2 *)> &matvec,
                       const emfa_mat *B, emfa_mat *X, int m, double tol, int max_restart, emfa_gmres_report *rep);

}  // namespace emfa
