/* C ABI of the host core (libemfa_core.so): the configuration loader, the
 * NAS mesh reader, a whole CONFIG.yml run, and the continuous-batching solve
 * server.  Plain C types only.  Errors: a non-zero return, with the message
 * in emfa_core_last_error() (per thread). */
#ifndef EMFA_CORE_H
#define EMFA_CORE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *emfa_core_last_error(void);

/* Value of `key` (with its colon, e.g. "freq:") as ConfigLoader stores it
 * (EMComponent/ConfigLoader.cpp:23-50 parse rules).  `src` is a file path,
 * or the text itself when is_text != 0.  Returns 0, 1 (missing key) or 2. */
int emfa_core_config_get(const char *src, int is_text, const char *key, char *out, int cap);

/* Mesh::loadMeshFile (EMComponent/Mesh.cpp:22-49) for .nas (:261-309),
 * .rpt (:178-218) and ASCII .stl (:220-259) files: 0-based triangles, vertex
 * coordinates as the reader stores them (.stl: equal corners joined into one
 * vertex).  With null buffers only nv/nt are returned.  Returns 0, 1
 * (buffers too small), 3. */
int emfa_core_read_mesh(const char *path, double *xyz, int nv_cap, int *tri, int nt_cap, int *nv, int *nt);

/* Test hook of the result-file writers (bistatic_rcs_*.dat of
 * EMComponent/Result.h:48-104 and current.sc of :234-255) on fixed synthetic
 * values, for a byte comparison with the reference's writers.  Returns 0 or 3. */
int emfa_core_write_formats(const char *dir, float theta_from, float theta_to, float theta_delta, float phi_from,
                            float phi_to, float phi_delta);

/* solver::loadSolver(config_path) (EMSolver/Solver.cpp:45-95): reads the
 * config, runs the named solver, writes its result files.  Returns 0 or the
 * launcher's exit codes: 2 config, 3 file, 4 invalid argument, 5 runtime. */
int emfa_core_run(const char *config_path);

/* ---------------------------------------------------------------------
 * Solve server.  No counterpart in the reference (it runs one CONFIG.yml
 * per process).  Requests are queued on submit and taken by worker
 * threads, each with its own device context; a request is admitted when
 * its device footprint fits the HBM budget beside the resident factors,
 * idle factors are evicted least-recently-used first, and a request whose
 * (mesh, frequency, formulation, precision, quadrature) factor is resident
 * only fills its excitation, solves and evaluates the far field.  Results
 * are bitwise the same whatever else runs: every kernel is deterministic
 * and a request runs on one context's streams only. */
typedef struct emfa_server emfa_server;

enum { EMFA_MESH_ICOSPHERE = 0, EMFA_MESH_PLATE = 1, EMFA_MESH_USER = 2 };
enum { EMFA_SERVE_DENSE = 0, EMFA_SERVE_MLFMA = 1, EMFA_SERVE_FASTWT = 2 };

typedef struct {
    int mesh_kind;      /* EMFA_MESH_ICOSPHERE (ellipsoid), EMFA_MESH_PLATE or EMFA_MESH_USER */
    int subdiv;         /* icosphere subdivisions, or plate cells per side */
    double ax, ay, az;  /* semi-axes (m); plate: side length in ax */
    int formulation;    /* EMFA_EFIE or EMFA_CFIE */
    double alpha;       /* CFIE combination factor */
    double freq;        /* Hz */
    int precision;      /* EMFA_F64: FP64 LU; EMFA_F32: FP32 fill + LU, FP64 refinement */
    int quad;           /* 3 or 7 */
    int nrhs;           /* monostatic incidences */
    const double *angles; /* nrhs x (theta, phi, pole), degrees */
    /* EMFA_MESH_USER: the caller's mesh, copied at submit; requests whose
     * vertices and triangles are equal share resident data (the key is a
     * hash of their content) */
    int nv, nt;
    const double *xyz;  /* nv x 3 (m) */
    const int *tri;     /* nt x 3, 0-based vertex ids */
    /* EMFA_SERVE_DENSE: Z and its LU stay resident; EMFA_SERVE_MLFMA: the
     * multilevel fast multipole operator stays resident and each request is
     * solved by GMRES(30) on it (EMSolver/MLFMA.cpp; quad and precision are
     * not used); EMFA_SERVE_FASTWT: the IP-SIE operator on the mesh's
     * half-RWGs (EMSolver/FastWT.cpp; CFIE only, nonconformal meshes
     * allowed) stays resident, solved the same way */
    int solver;
    double mlfma_box;   /* finest box edge in wavelengths (0: 0.25) */
    double iter_tol;    /* GMRES relative residual (0: 1e-4) */
    int iter_max;       /* GMRES iterations at most (0: 600) */
    double fastwt_beta; /* EMFA_SERVE_FASTWT: stabilization factor x |log10(average size)| (0: 0.1 x |log10(0.07503)|) */
} emfa_serve_request;

typedef struct {
    long long id;
    int status;         /* EMFA_OK or an emfa_status code */
    int unknowns, nrhs;
    int cache_hit;
    int refine_iterations;
    double relres;      /* final relative residual of the refinement (FP32 requests) */
    int fp64_fallback;  /* refinement did not converge: solved with the FP64 factor */
    double queue_ms, service_ms, latency_ms;
    int batch_size;     /* requests solved together (same factor, one stacked refinement) */
} emfa_serve_result;

typedef struct {
    long long completed, failed, cache_hits;
    double p50_ms, p99_ms, mean_ms, max_ms; /* arrival-to-completion latency */
    double wall_s;                          /* first arrival to last completion */
    double solves_per_s, rhs_per_s;
    size_t peak_bytes;                      /* admitted footprint + resident factors */
    long long pivot_fallbacks;              /* FP32 factors redone with row interchanges */
    long long batched;                      /* requests that joined another request's batch */
    long long factor_batches;               /* factorisations that built several new FP32 factors together */
    long long batch_factors;                /* factors built that way */
} emfa_serve_stats;

/* workers >= 1 contexts on `device`; hbm_budget in bytes; keep_current != 0
 * keeps each request's solution for emfa_serve_fetch. */
int emfa_serve_create(int device, int workers, size_t hbm_budget, int keep_current, emfa_server **out);
int emfa_serve_destroy(emfa_server *s);
/* Returns the request id (>= 0), or -1 with the message set. */
long long emfa_serve_submit(emfa_server *s, const emfa_serve_request *r);
/* Blocks until request `id` (id < 0: any) is done and fills `res`.
 * Returns 0, or 1 when no such request is pending. */
int emfa_serve_wait(emfa_server *s, long long id, emfa_serve_result *res);
/* Copies a waited-for result's monostatic RCS (dB, nrhs values) and, when
 * kept, its solution (interleaved complex, unknowns x nrhs, row-major) and
 * releases it.  Either buffer may be null.  Returns 0, or 1 (unknown id or
 * buffer too small). */
int emfa_serve_fetch(emfa_server *s, long long id, double *rcs_db, int rcs_cap, double *current, long long current_cap);
int emfa_serve_drain(emfa_server *s);
/* No reference counterpart.  on = 1: workers take no queued request until
 * emfa_serve_hold(s, 0) (or emfa_serve_drain), so that requests submitted
 * together are scheduled as one queue (largest work first) rather than in
 * the order they reached idle workers.  Returns 0, or 1 on a bad argument. */
int emfa_serve_hold(emfa_server *s, int on);
int emfa_serve_get_stats(emfa_server *s, emfa_serve_stats *out);
int emfa_serve_reset_stats(emfa_server *s);
int emfa_serve_cached_factors(emfa_server *s);
/* Footprint model used for admission, in bytes. */
size_t emfa_serve_footprint(int unknowns, int nrhs, int precision, int factor_cached);

#ifdef __cplusplus
}
#endif
#endif
