// Allocation-free BiCGSTAB: emfa_bicgstab with its vectors in a fixed
// device array instead of memory the library allocates, for callers that
// cannot allocate device memory.
#include <map>
#include <mutex>

#include "internal.h"

using namespace emfa;

// ---------------------------------------------------------------------------
// Allocation-free variant: the solver's vectors live in a file-scope device
// array of fixed size (each device the library runs on holds its own copy
// until the process ends).  Host code reaches a device's copy only through
// the address hipGetSymbolAddress returns with that device current, looked up
// once per device and kept.  Calls on one device share the array: calls that
// overlap in time on one device race.  Every vector the solver reads is
// written first by the same call, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 8u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device ordinal -> address of that device's array

emfa_status device_scratch(int device, void **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return EMFA_OK;
    }
    EMFA_HIP_CHECK(hipSetDevice(device));
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        set_error("bicgstab (no allocation): device %d: hipGetSymbolAddress failed: %s", device, hipGetErrorString(e));
        return EMFA_E_HIP;
    }
    g_scratch_ptr[device] = ptr;
    *out = ptr;
    return EMFA_OK;
}
}  // namespace

extern "C" emfa_status emfa_bicgstab_noalloc(emfa_ctx *ctx, const emfa_mat *Z, const emfa_mat *B, emfa_mat *X,
                                             double tol, int max_iter, int variant, emfa_bicgstab_report *rep) {
    EMFA_REQUIRE(ctx && Z && B && X, EMFA_E_INVALID, "emfa_bicgstab_noalloc: null argument");
    EMFA_REQUIRE(Z->cols == Z->rows, EMFA_E_INVALID, "emfa_bicgstab_noalloc: Z must be square");
    const int n = Z->rows;
    EMFA_REQUIRE(n <= EMFA_BICGSTAB_NOALLOC_MAX_N && bicgstab_work_bytes(n) <= kNoAllocScratchBytes, EMFA_E_INVALID,
                 "emfa_bicgstab_noalloc: %d unknowns exceed the fixed workspace (at most %d unknowns, %zu bytes)", n,
                 EMFA_BICGSTAB_NOALLOC_MAX_N, kNoAllocScratchBytes);
    void *ptr = nullptr;
    emfa_status st = device_scratch(ctx->device, &ptr);
    if (st) return st;
    return bicgstab_impl(
        ctx, n, [&](const double2 *in, double2 *out) { return gemv(ctx, Z, in, out); }, B, X, tol, max_iter, variant,
        rep, ptr, kNoAllocScratchBytes);
}

