// MFMA GEMM kernels (see gemm.h for the complex-as-real formulation).
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <type_traits>

#include "gemm.h"

namespace emfa {

// XCD-aware tile order.  The dispatcher hands workgroup L (linear id) to XCD
// L % 8, and each XCD has its own L2.  The tile sequence is cut into eight
// contiguous chunks, one per XCD, and walked in groups of kGroupRows tile
// rows, so the workgroups resident on one XCD share A row panels and B
// column panels in that XCD's L2.  A bijection for any grid, so only speed
// depends on the dispatch order.
constexpr int kGroupRows = 8;

// The same order for linear tile index L of a TX x TY tile grid (persistent
// workgroups: workgroup i takes tiles i, i + G, ..., all on XCD i % 8 when G is
// a multiple of 8).
__device__ __forceinline__ void tile_of_linear(int L, int TX, int TY, int *ty, int *tx) {
    const int nT = TX * TY;
    const int xcd = L & 7, slot = L >> 3;
    const int q = nT >> 3, r = nT & 7;
    const int t = xcd * q + min(xcd, r) + slot;
    const int group = t / (kGroupRows * TX);
    const int first = group * kGroupRows;
    const int rows = min(kGroupRows, TY - first);
    const int local = t - group * kGroupRows * TX;
    *ty = first + local % rows;
    *tx = local / rows;
}

__device__ __forceinline__ void tile_of(int swizzle, int *ty, int *tx) {
    const int TX = gridDim.x, TY = gridDim.y;
    if (!swizzle) {
        *ty = blockIdx.y;
        *tx = blockIdx.x;
        return;
    }
    const int nT = TX * TY;
    const int L = blockIdx.y * TX + blockIdx.x;
    const int xcd = L & 7, slot = L >> 3;
    const int q = nT >> 3, r = nT & 7;
    const int t = xcd * q + min(xcd, r) + slot;
    const int group = t / (kGroupRows * TX);
    const int first = group * kGroupRows;
    const int rows = min(kGroupRows, TY - first);
    const int local = t - group * kGroupRows * TX;
    *ty = first + local % rows;
    *tx = local / rows;
}

// C (+)= sgn * acc for one wave's 64 x 64 tile of 2 x 2 MFMA 32x32 results.
// The loads of C of each 32-row half are issued before its stores: a store
// through C may alias a later load, so a load-add-store per element makes each
// one wait for the previous store's round trip (64 in a row per lane).  Halves
// rather than the whole tile keep the kernel within 256 VGPRs.
__device__ __forceinline__ void epilogue_f32(const GemmArgs &g, float *C, const floatx16 (&acc)[2][2],
                                             long long r0, long long c0, int lane) {
    const float sgn = g.alpha_neg ? -1.f : 1.f;
    const long long j0 = c0 + (lane & 31);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        float old[2][16];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long i = r0 + mt * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), j = j0 + nt * 32;
                old[nt][v] = (g.beta && i < g.M && j < g.Nr) ? C[i * g.ldc + j] : 0.f;
            }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long i = r0 + mt * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), j = j0 + nt * 32;
                if (i < g.M && j < g.Nr) C[i * g.ldc + j] = old[nt][v] + sgn * acc[mt][nt][v];
            }
    }
}

// ---------------------------------------------------------------------------
// FP32: workgroup tile 128 x 128 (reals), 4 waves in 2 x 2, wave tile 64 x 64 =
// 2 x 2 MFMA 32x32 tiles.  K step 32: MFMA t (0..15) takes k = 16*(lane/32) + t,
// so each lane reads 16 consecutive K values of its row from LDS.
// Double-buffered LDS, next tile prefetched into registers during the MFMAs.
// ---------------------------------------------------------------------------
namespace f32k {
constexpr int BM = 128, BN = 128, BK = 32, PAD = 4;
}

// Two waves per SIMD (two 72 KB workgroups per CU): the bound keeps it at 180
// VGPRs without spills; unbounded it took 216 + 64 and ran one.
template <int TAG>
__global__ __launch_bounds__(256, 2) void k_gemm_f32(GemmArgs g) {
    using namespace f32k;
    __shared__ float As[2][BM][BK + PAD];
    __shared__ float Bs[2][BN][BK + PAD];
    const float *__restrict__ A = batch_mat(g.bt, (const float *)g.A);
    const float *__restrict__ Bt = batch_scr(g.bt, (const float *)g.Bt);
    float *__restrict__ C = batch_mat(g.bt, (float *)g.C);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int ty, tx;
    tile_of(g.swizzle, &ty, &tx);
    const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
    const int nk = (g.Kr + BK - 1) / BK;

    float2 ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            int idx = tid + 256 * i;
            int row = idx >> 4, kk = (idx & 15) * 2;
            long long gr = bm + row;
            int gk = k0 + kk;
            ra[i] = (gr < g.M && gk < g.Kr) ? *(const float2 *)(A + gr * g.lda + gk) : make_float2(0.f, 0.f);
            long long gc = bn + row;
            rb[i] = (gc < g.Nr && gk < g.Kr) ? *(const float2 *)(Bt + gc * g.ldb + gk) : make_float2(0.f, 0.f);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            int idx = tid + 256 * i;
            int row = idx >> 4, kk = (idx & 15) * 2;
            As[buf][row][kk] = ra[i].x;
            As[buf][row][kk + 1] = ra[i].y;
            Bs[buf][row][kk] = rb[i].x;
            Bs[buf][row][kk + 1] = rb[i].y;
        }
    };

    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    gload(0);
    sstore(0);
    __syncthreads();
    const int r32 = lane & 31, kg = 16 * (lane >> 5);
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
        float av[2][16], bv[2][16];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const float4 *src = (const float4 *)&As[buf][wm * 64 + mt * 32 + r32][kg];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float4 v = src[q];
                av[mt][4 * q] = v.x;
                av[mt][4 * q + 1] = v.y;
                av[mt][4 * q + 2] = v.z;
                av[mt][4 * q + 3] = v.w;
            }
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const float4 *src = (const float4 *)&Bs[buf][wn * 64 + nt * 32 + r32][kg];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float4 v = src[q];
                bv[nt][4 * q] = v.x;
                bv[nt][4 * q + 1] = v.y;
                bv[nt][4 * q + 2] = v.z;
                bv[nt][4 * q + 3] = v.w;
            }
        }
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int nt = 0; nt < 2; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt][t], bv[nt][t], acc[mt][nt], 0, 0, 0);
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }

    epilogue_f32(g, C, acc, bm + wm * 64, bn + wn * 64, lane);
}

// ---------------------------------------------------------------------------
// FP32, small tiles: workgroup tile 64 x 64 (reals), 4 waves in 2 x 2, one
// MFMA 32x32 tile per wave, K step 32 with the same K mapping as k_gemm_f32
// (identical sums).  For launches whose 128 x 128 grid would leave most CUs
// idle: four times the workgroups and a quarter of each wave's MFMA work
// (a 128 x 128 tile issues 64 MFMAs per K step per wave: ~7 us at K = 128).
// ---------------------------------------------------------------------------
namespace f32s {
constexpr int BM = 64, BN = 64, BK = 32, PAD = 4;
}

template <int TAG>
__global__ __launch_bounds__(256) void k_gemm_f32s(GemmArgs g) {
    using namespace f32s;
    __shared__ float As[2][BM][BK + PAD];
    __shared__ float Bs[2][BN][BK + PAD];
    const float *__restrict__ A = batch_mat(g.bt, (const float *)g.A);
    const float *__restrict__ Bt = batch_scr(g.bt, (const float *)g.Bt);
    float *__restrict__ C = batch_mat(g.bt, (float *)g.C);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int ty, tx;
    tile_of(g.swizzle, &ty, &tx);
    const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
    const int nk = (g.Kr + BK - 1) / BK;

    float2 ra[4], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + 256 * i;
            const int row = idx >> 4, kk = (idx & 15) * 2;
            const long long gr = bm + row, gc = bn + row;
            const int gk = k0 + kk;
            ra[i] = (gr < g.M && gk < g.Kr) ? *(const float2 *)(A + gr * g.lda + gk) : make_float2(0.f, 0.f);
            rb[i] = (gc < g.Nr && gk < g.Kr) ? *(const float2 *)(Bt + gc * g.ldb + gk) : make_float2(0.f, 0.f);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int idx = tid + 256 * i;
            const int row = idx >> 4, kk = (idx & 15) * 2;
            *(float2 *)&As[buf][row][kk] = ra[i];
            *(float2 *)&Bs[buf][row][kk] = rb[i];
        }
    };

    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    // C is read before the K loop: with the short K of the panel and solve
    // updates (a few steps of 32) its load latency would otherwise follow the
    // last MFMA.  Only this workgroup writes its tile of C during the launch.
    const long long j = bn + wn * 32 + (lane & 31);
    float old[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const long long i = bm + wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        old[v] = (g.beta && i < g.M && j < g.Nr) ? C[i * g.ldc + j] : 0.f;
    }
    gload(0);
    sstore(0);
    __syncthreads();
    const int r32 = lane & 31, kg = 16 * (lane >> 5);
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
        float av[16], bv[16];
        const float4 *sa = (const float4 *)&As[buf][wm * 32 + r32][kg];
        const float4 *sb = (const float4 *)&Bs[buf][wn * 32 + r32][kg];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 x = sa[q], y = sb[q];
            av[4 * q] = x.x;
            av[4 * q + 1] = x.y;
            av[4 * q + 2] = x.z;
            av[4 * q + 3] = x.w;
            bv[4 * q] = y.x;
            bv[4 * q + 1] = y.y;
            bv[4 * q + 2] = y.z;
            bv[4 * q + 3] = y.w;
        }
#pragma unroll
        for (int t = 0; t < 16; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], bv[t], acc, 0, 0, 0);
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }

    const float sgn = g.alpha_neg ? -1.f : 1.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const long long i = bm + wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (i < g.M && j < g.Nr) C[i * g.ldc + j] = old[v] + sgn * acc[v];
    }
}

// Variant with 16-byte global loads and a K step of BKV (32 or 64): same
// MFMA mapping; needs lda, ldb, Kr multiples of 4 and 16-byte aligned rows.
template <int TAG, int BKV>
__global__ __launch_bounds__(256) void k_gemm_f32v(GemmArgs g) {
    constexpr int BM = 128, BN = 128, PAD = 4, KP = BKV / 4;   // float4 per row per K step
    constexpr int LOADS = BM * KP / 256;                       // float4 loads per thread per operand
    __shared__ float As[2][BM][BKV + PAD];
    __shared__ float Bs[2][BN][BKV + PAD];
    const float *__restrict__ A = batch_mat(g.bt, (const float *)g.A);
    const float *__restrict__ Bt = batch_scr(g.bt, (const float *)g.Bt);
    float *__restrict__ C = batch_mat(g.bt, (float *)g.C);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int ty, tx;
    tile_of(g.swizzle, &ty, &tx);
    const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
    const int nk = (g.Kr + BKV - 1) / BKV;
    float4 ra[LOADS], rb[LOADS];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            int idx = tid + 256 * i;
            int row = idx / KP, kk = (idx % KP) * 4;
            long long gr = bm + row, gc = bn + row;
            int gk = k0 + kk;
            const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
            ra[i] = (gr < g.M && gk < g.Kr) ? *(const float4 *)(A + gr * g.lda + gk) : z;
            rb[i] = (gc < g.Nr && gk < g.Kr) ? *(const float4 *)(Bt + gc * g.ldb + gk) : z;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            int idx = tid + 256 * i;
            int row = idx / KP, kk = (idx % KP) * 4;
            *(float4 *)&As[buf][row][kk] = ra[i];
            *(float4 *)&Bs[buf][row][kk] = rb[i];
        }
    };
    floatx16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
    gload(0);
    sstore(0);
    __syncthreads();
    const int r32 = lane & 31, half = lane >> 5;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BKV);
#pragma unroll
        for (int ks = 0; ks < BKV / 32; ++ks) {
            const int kg = ks * 32 + 16 * half;
            float av[2][16], bv[2][16];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const float4 *src = (const float4 *)&As[buf][wm * 64 + mt * 32 + r32][kg];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float4 v = src[q];
                    av[mt][4 * q] = v.x;
                    av[mt][4 * q + 1] = v.y;
                    av[mt][4 * q + 2] = v.z;
                    av[mt][4 * q + 3] = v.w;
                }
            }
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const float4 *src = (const float4 *)&Bs[buf][wn * 64 + nt * 32 + r32][kg];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float4 v = src[q];
                    bv[nt][4 * q] = v.x;
                    bv[nt][4 * q + 1] = v.y;
                    bv[nt][4 * q + 2] = v.z;
                    bv[nt][4 * q + 3] = v.w;
                }
            }
#pragma unroll
            for (int t = 0; t < 16; ++t)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt)
                        acc[mt][nt] =
                            __builtin_amdgcn_mfma_f32_32x32x2f32(av[mt][t], bv[nt][t], acc[mt][nt], 0, 0, 0);
        }
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }
    epilogue_f32(g, C, acc, bm + wm * 64, bn + wn * 64, lane);
}

// ---------------------------------------------------------------------------
// FP64 accumulate: workgroup tile 128 x 128, wave tile 64 x 64 = 4 x 4 MFMA
// 16x16 tiles, K step 16: MFMA t (0..3) takes k = 4*(lane/16) + t.
// TA = float converts the matrix operand to double on its way into LDS.
// ---------------------------------------------------------------------------
// FP32 update emulated on bf16 MFMA (opt-in per context, LU trailing update
// only).  Each FP32 operand x is split into hi = bf16(x), lo = bf16(x - hi);
// the product uses hi*hi + hi*lo + lo*hi (three V_MFMA_F32_32X32X16_BF16 per
// block and 16-deep K step) with FP32 accumulation, which keeps about 16
// significant bits of every product; the FP64 refinement on the FP32 matrix
// removes the difference.
// MFMA lane layout: lane l holds row (A) / column (B) l % 32 and the eight
// K values 8 * (l / 32) .. + 7; result element v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Operand split for the plane kernel: one thread per 8 consecutive K of a row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_split_bf16(const float *__restrict__ src, long long ld, int rows, int cols,
                                                    __bf16 *__restrict__ hi, __bf16 *__restrict__ lo, long long ldp,
                                                    int rows_p, int cols_p, Batch bt, int src_mat) {
    src = src_mat ? batch_mat(bt, src) : batch_scr(bt, src);
    hi = batch_scr(bt, hi);
    lo = batch_scr(bt, lo);
    const int chunks = cols_p >> 3;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)rows_p * chunks) return;
    const int r = (int)(t / chunks), k0 = (int)(t % chunks) * 8;
    bf16x8 h, l;
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        float2 v = make_float2(0.f, 0.f);
        if (r < rows) {
            const float *q = src + (long long)r * ld + k0 + j;
            if (k0 + j + 1 < cols)
                v = *(const float2 *)q;
            else if (k0 + j < cols)
                v.x = q[0];
        }
        const __bf16 h0 = (__bf16)v.x, h1 = (__bf16)v.y;
        h[j] = h0;
        h[j + 1] = h1;
        l[j] = (__bf16)(v.x - (float)h0);
        l[j + 1] = (__bf16)(v.y - (float)h1);
    }
    *(bf16x8 *)(hi + (long long)r * ldp + k0) = h;
    *(bf16x8 *)(lo + (long long)r * ldp + k0) = l;
}

// ---------------------------------------------------------------------------
// Three-term bf16 GEMM from pre-split planes (Bf3Planes).  Workgroup tile
// 256 x 256, 8 waves as 4 x 2 with wave tiles 64 x 128 (2 x 4 MFMA 32x32x16).
// The operand tiles go straight from global memory into LDS with 16-byte
// direct loads (global_load_lds_dwordx4: lane l of a wave writes LDS bytes
// base + 16 l), four stages of K = 16 in flight: stage kt + 3 is issued while
// stage kt is multiplied.  No registers stage the operands and no VALU work
// sits in the loop.  An LDS plane row holds 16 bf16 (two 16-byte chunks);
// chunk c of row r is stored at position c ^ swz(r), swz(r) = bit 2 ^ bit 3
// of r: the sixteen rows that one ds_read_b128 cycle serves (sixteen lanes,
// 64 banks of 4 bytes) fall on sixteen distinct 4-bank groups.  (With bit 2
// alone, rows r and r + 8 shared banks: the counters showed four conflict
// cycles per LDS read.)
// Each wave waits (s_waitcnt vmcnt) for its own loads of the stage, then the
// workgroup barrier publishes every wave's loads; the same barrier tells the
// issuing waves that all reads of the buffer being refilled are done.
// ---------------------------------------------------------------------------
namespace bf3p {
constexpr int BM = 256, BN = 256, BK = 16, NS = 4, PLANE = 256 * BK;
}

__device__ __forceinline__ int bf3_swz(int r) { return ((r >> 2) ^ (r >> 3)) & 1; }
__device__ __forceinline__ int bf3p_off(int r, int kg) { return r * bf3p::BK + ((kg ^ bf3_swz(r)) << 3); }

typedef __attribute__((address_space(3))) void lds_void;

__global__ __launch_bounds__(512) void k_gemm_bf3p(Bf3Planes g) {
    using namespace bf3p;
    __shared__ __attribute__((aligned(16))) __bf16 sm[NS][4][PLANE];  // stage, plane (Ah Al Bh Bl)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int TX = (g.Nr + BN - 1) / BN, TY = (g.M + BM - 1) / BM, TT = TX * TY;
    // one tile per workgroup, or (persistent grid) tiles blockIdx.x, + gridDim.x, ...
    // over the tiles of every matrix of the batch (matrix tl / TT)
    const int t0 = g.persistent ? (int)blockIdx.x : (int)((blockIdx.z * TY + blockIdx.y) * gridDim.x + blockIdx.x);
    const int tstep = g.persistent ? (int)gridDim.x : TT * g.bt.count;
    for (int tl = t0; tl < TT * g.bt.count; tl += tstep) {
    if (tl != t0) __syncthreads();  // the previous tile's LDS reads are done
    const int zb = tl / TT;
    int ty, tx;
    tile_of_linear(tl - zb * TT, TX, TY, &ty, &tx);
    const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
    const int nk = g.Kp / BK;

    // this wave's direct loads: plane pl, rows half * 128 + 32 i + lane / 2
    const int pl = wave >> 1, half = wave & 1;
    const __bf16 *src = batch_scr(g.bt, pl == 0 ? g.Ah : pl == 1 ? g.Al : pl == 2 ? g.Bh : g.Bl, zb);
    const long long ld = pl < 2 ? g.lda : g.ldb;
    const long long row0 = pl < 2 ? bm : bn;
    const __bf16 *gp[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = half * 128 + 32 * i + (lane >> 1);
        const int c = (lane & 1) ^ bf3_swz(r);
        gp[i] = src + (row0 + r) * ld + 8 * c;
    }
    auto issue = [&](int kt) {
        const int st = kt & (NS - 1);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            __builtin_amdgcn_global_load_lds((const void *)(gp[i] + (long long)kt * BK),
                                             (lds_void *)&sm[st][pl][(half * 128 + 32 * i) * BK], 16, 0, 0);
    };

    floatx16 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    issue(0);
    if (nk > 1) issue(1);
    if (nk > 2) issue(2);
    const int r32 = lane & 31, kg = lane >> 5;
    for (int kt = 0; kt < nk; ++kt) {
        // loads of stages kt+1, kt+2 (4 per stage) may stay in flight
        if (kt + 2 < nk)
            asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
        else if (kt + 1 < nk)
            asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (kt + 3 < nk) issue(kt + 3);
        const int st = kt & (NS - 1);
        bf16x8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const int o = bf3p_off(wm * 64 + mt * 32 + r32, kg);
            ah[mt] = *(const bf16x8 *)&sm[st][0][o];
            al[mt] = *(const bf16x8 *)&sm[st][1][o];
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int o = bf3p_off(wn * 128 + nt * 32 + r32, kg);
            bh[nt] = *(const bf16x8 *)&sm[st][2][o];
            bl[nt] = *(const bf16x8 *)&sm[st][3][o];
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mt], bh[nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bl[nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bh[nt], acc[mt][nt], 0, 0, 0);
    }

    // epilogue: all C loads of a 32-row slab are issued before its stores;
    // tiles wholly inside C skip the per-element bounds checks
    const float sgn = g.alpha_neg ? -1.f : 1.f;
    float *__restrict__ C = batch_mat(g.bt, g.C, zb) + (bm + wm * 64 + 4 * (lane >> 5)) * g.ldc + bn + wn * 128 + r32;
    const int mrem = (int)min((long long)BM, g.M - bm) - wm * 64 - 4 * (lane >> 5);
    const int nrem = (int)min((long long)BN, (long long)g.Nr - bn) - wn * 128 - r32;
    const bool full = bm + BM <= g.M && bn + BN <= g.Nr;
    auto epilogue = [&](auto guarded) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            float cv[4][16];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                    const bool in = !guarded.value || (i < mrem && j < nrem);
                    cv[nt][v] = (g.beta && in) ? C[i * g.ldc + j] : 0.f;
                }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                    if (!guarded.value || (i < mrem && j < nrem)) C[i * g.ldc + j] = cv[nt][v] + sgn * acc[mt][nt][v];
                }
        }
    };
    if (full)
        epilogue(std::false_type{});
    else
        epilogue(std::true_type{});
    }  // tiles
}

// ---------------------------------------------------------------------------
// The same product with register-staged operand loads and K = 32 per stage.
// Workgroup tile 256 x 256, 8 waves of 64 x 128 as in k_gemm_bf3p; two LDS
// buffers of 4 planes x 256 rows x 32 K (128 KB).  Each thread loads eight
// 16-byte chunks per stage (two per plane; four threads cover the 64 bytes of
// a row's stage) into registers with ordinary global loads, and stores them
// to the other LDS buffer after multiplying the current one; one barrier per
// stage publishes the stores and frees the buffer just read.  A row's stage
// holds 4 chunks; chunk c sits at position c ^ ((r >> 2) & 3), so the
// sixteen rows one ds_read_b128 cycle serves cover all 64 banks.  The K = 16
// halves of a stage are multiplied in order with bf3p's product order, so the
// bits equal k_gemm_bf3p's.  Needs Kp % 32 == 0.
// ---------------------------------------------------------------------------
namespace bf3r {
constexpr int BM = 256, BN = 256, BK = 32, PLANE = 256 * BK;
}

__device__ __forceinline__ int bf3r_off(int r, int c) { return r * bf3r::BK + ((c ^ ((r >> 2) & 3)) << 3); }

__global__ __launch_bounds__(512) void k_gemm_bf3r(Bf3Planes g) {
    using namespace bf3r;
    __shared__ __attribute__((aligned(16))) __bf16 sm[2][4][PLANE];  // buffer, plane (Ah Al Bh Bl)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int TX = (g.Nr + BN - 1) / BN, TY = (g.M + BM - 1) / BM, TT = TX * TY;
    const int t0 = g.persistent ? (int)blockIdx.x : (int)((blockIdx.z * TY + blockIdx.y) * gridDim.x + blockIdx.x);
    const int tstep = g.persistent ? (int)gridDim.x : TT * g.bt.count;
    const int r32 = lane & 31, kg = lane >> 5;
    // this thread's chunks: q = i * 512 + tid, row q >> 2, chunk q & 3 (same in every plane)
    const int q0 = tid, q1 = 512 + tid;
    const int oa0 = (q0 >> 2) * (int)g.lda + 8 * (q0 & 3), oa1 = (q1 >> 2) * (int)g.lda + 8 * (q1 & 3);
    const int ob0 = (q0 >> 2) * (int)g.ldb + 8 * (q0 & 3), ob1 = (q1 >> 2) * (int)g.ldb + 8 * (q1 & 3);
    const int ol0 = bf3r_off(q0 >> 2, q0 & 3), ol1 = bf3r_off(q1 >> 2, q1 & 3);
    for (int tl = t0; tl < TT * g.bt.count; tl += tstep) {
        if (tl != t0) __syncthreads();  // the previous tile's LDS reads are done
        const int zb = tl / TT;
        int ty, tx;
        tile_of_linear(tl - zb * TT, TX, TY, &ty, &tx);
        const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
        const int nk = g.Kp / BK;
        const __bf16 *pah = batch_scr(g.bt, g.Ah, zb) + bm * g.lda, *pal = batch_scr(g.bt, g.Al, zb) + bm * g.lda;
        const __bf16 *pbh = batch_scr(g.bt, g.Bh, zb) + bn * g.ldb, *pbl = batch_scr(g.bt, g.Bl, zb) + bn * g.ldb;
        // staging registers (named, so that they stay in registers)
        uint4 s00, s01, s10, s11, s20, s21, s30, s31;
#define BF3R_GLOAD(kt)                                            \
    do {                                                          \
        const int k0_ = (kt) * BK;                                \
        s00 = *(const uint4 *)(pah + k0_ + oa0);                  \
        s01 = *(const uint4 *)(pah + k0_ + oa1);                  \
        s10 = *(const uint4 *)(pal + k0_ + oa0);                  \
        s11 = *(const uint4 *)(pal + k0_ + oa1);                  \
        s20 = *(const uint4 *)(pbh + k0_ + ob0);                  \
        s21 = *(const uint4 *)(pbh + k0_ + ob1);                  \
        s30 = *(const uint4 *)(pbl + k0_ + ob0);                  \
        s31 = *(const uint4 *)(pbl + k0_ + ob1);                  \
    } while (0)
#define BF3R_LSTORE(b)                                            \
    do {                                                          \
        *(uint4 *)&sm[(b)][0][ol0] = s00;                         \
        *(uint4 *)&sm[(b)][0][ol1] = s01;                         \
        *(uint4 *)&sm[(b)][1][ol0] = s10;                         \
        *(uint4 *)&sm[(b)][1][ol1] = s11;                         \
        *(uint4 *)&sm[(b)][2][ol0] = s20;                         \
        *(uint4 *)&sm[(b)][2][ol1] = s21;                         \
        *(uint4 *)&sm[(b)][3][ol0] = s30;                         \
        *(uint4 *)&sm[(b)][3][ol1] = s31;                         \
    } while (0)
        floatx16 acc[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
        BF3R_GLOAD(0);
        BF3R_LSTORE(0);
        if (nk > 1) BF3R_GLOAD(1);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int buf = kt & 1;
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                bf16x8 ah[2], al[2], bh[4], bl[4];
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) {
                    const int o = bf3r_off(wm * 64 + mt * 32 + r32, 2 * kk + kg);
                    ah[mt] = *(const bf16x8 *)&sm[buf][0][o];
                    al[mt] = *(const bf16x8 *)&sm[buf][1][o];
                }
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const int o = bf3r_off(wn * 128 + nt * 32 + r32, 2 * kk + kg);
                    bh[nt] = *(const bf16x8 *)&sm[buf][2][o];
                    bl[nt] = *(const bf16x8 *)&sm[buf][3][o];
                }
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mt], bh[nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bl[nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mt], bh[nt], acc[mt][nt], 0, 0, 0);
            }
            if (kt + 1 < nk) {
                BF3R_LSTORE(buf ^ 1);  // last read in step kt - 1, before the previous barrier
                if (kt + 2 < nk) BF3R_GLOAD(kt + 2);
            }
            __syncthreads();
        }
#undef BF3R_GLOAD
#undef BF3R_LSTORE

        const float sgn = g.alpha_neg ? -1.f : 1.f;
        float *__restrict__ C =
            batch_mat(g.bt, g.C, zb) + (bm + wm * 64 + 4 * (lane >> 5)) * g.ldc + bn + wn * 128 + r32;
        const int mrem = (int)min((long long)BM, g.M - bm) - wm * 64 - 4 * (lane >> 5);
        const int nrem = (int)min((long long)BN, (long long)g.Nr - bn) - wn * 128 - r32;
        const bool full = bm + BM <= g.M && bn + BN <= g.Nr;
        auto epilogue = [&](auto guarded) {
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                float cv[4][16];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                        const bool in = !guarded.value || (i < mrem && j < nrem);
                        cv[nt][v] = (g.beta && in) ? C[i * g.ldc + j] : 0.f;
                    }
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                        if (!guarded.value || (i < mrem && j < nrem)) C[i * g.ldc + j] = cv[nt][v] + sgn * acc[mt][nt][v];
                    }
            }
        };
        if (full)
            epilogue(std::false_type{});
        else
            epilogue(std::true_type{});
    }  // tiles
}

// ---------------------------------------------------------------------------
// The same product with one wave per SIMD and register-double-buffered
// operand fragments.  Workgroup tile 256 x 256, 4 waves as 2 x 2 with wave
// tiles 128 x 128 (4 x 4 MFMA 32x32x16 tiles: 256 accumulators per lane,
// which the compiler keeps in the accumulation registers), LDS ring of four
// K = 16 stages filled by direct global -> LDS loads (wave w loads plane w:
// Ah, Al, Bh, Bl).  At step kt the wave multiplies the fragments of stage kt,
// already in registers, while it reads the fragments of stage kt + 1 from
// LDS into the other register set (one LDS read per three MFMAs, the order
// pinned with scheduling barriers); the one barrier per step publishes stage
// kt + 1 and frees the slot of stage kt - 1 for the loads of stage kt + 3.  Per
// accumulator the products are added in bf3p's order (Al Bh, Ah Bl, Ah Bh per
// stage, stages in order), so the two kernels give the same bits.
// ---------------------------------------------------------------------------
namespace bf3q {
constexpr int BM = 256, BN = 256, BK = 16, NS = 4, PLANE = 256 * BK;
}

__global__ __launch_bounds__(256, 1) void k_gemm_bf3q(Bf3Planes g) {
    using namespace bf3q;
    __shared__ __attribute__((aligned(16))) __bf16 sm[NS][4][PLANE];  // stage, plane (Ah Al Bh Bl)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int TX = (g.Nr + BN - 1) / BN, TY = (g.M + BM - 1) / BM, TT = TX * TY;
    const int t0 = g.persistent ? (int)blockIdx.x : (int)((blockIdx.z * TY + blockIdx.y) * gridDim.x + blockIdx.x);
    const int tstep = g.persistent ? (int)gridDim.x : TT * g.bt.count;
    const int r32 = lane & 31, kg = lane >> 5;
    for (int tl = t0; tl < TT * g.bt.count; tl += tstep) {
        if (tl != t0) __syncthreads();  // the previous tile's LDS reads are done
        const int zb = tl / TT;
        int ty, tx;
        tile_of_linear(tl - zb * TT, TX, TY, &ty, &tx);
        const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
        const int nk = g.Kp / BK;

        // this wave's direct loads: plane `wave`, rows 32 i + lane / 2 (i < 8),
        // chunk (lane & 1) stored at its swizzled position
        const __bf16 *src = batch_scr(g.bt, wave == 0 ? g.Ah : wave == 1 ? g.Al : wave == 2 ? g.Bh : g.Bl, zb);
        const long long ld = wave < 2 ? g.lda : g.ldb;
        const long long row0 = wave < 2 ? bm : bn;
        const int lr = lane >> 1;
        const __bf16 *gbase = src + (row0 + lr) * ld + 8 * ((lane & 1) ^ bf3_swz(lr));
        const long long rstep = 32 * ld;
        auto issue = [&](int kt) {
            const int st = kt & (NS - 1);
            const __bf16 *gp = gbase + (long long)kt * BK;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                __builtin_amdgcn_global_load_lds((const void *)(gp + i * rstep), (lds_void *)&sm[st][wave][32 * i * BK],
                                                 16, 0, 0);
        };
        int offa[4], offb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            offa[t] = bf3p_off(wm * 128 + t * 32 + r32, kg);
            offb[t] = bf3p_off(wn * 128 + t * 32 + r32, kg);
        }
        bf16x8 ah[2][4], al[2][4], bh[2][4], bl[2][4];
        auto read = [&](int kt, int buf) {
            const int st = kt & (NS - 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                ah[buf][t] = *(const bf16x8 *)&sm[st][0][offa[t]];
                al[buf][t] = *(const bf16x8 *)&sm[st][1][offa[t]];
                bh[buf][t] = *(const bf16x8 *)&sm[st][2][offb[t]];
                bl[buf][t] = *(const bf16x8 *)&sm[st][3][offb[t]];
            }
        };
        floatx16 acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

        issue(0);
        if (nk > 1) issue(1);
        if (nk > 2) issue(2);
        if (nk > 2)
            asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
        else if (nk > 1)
            asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        read(0, 0);
        // one step: fragments of stage kt in set cur, stage kt + 1 read into set nxt
        auto step = [&](int kt, int cur, int nxt) {
            if (kt + 2 < nk)
                asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
            else
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            if (kt + 3 < nk) issue(kt + 3);
            read(kt + 1 < nk ? kt + 1 : kt, nxt);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[cur][mt], bh[cur][nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[cur][mt], bl[cur][nt], acc[mt][nt], 0, 0, 0);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[cur][mt], bh[cur][nt], acc[mt][nt], 0, 0, 0);
            // one LDS read of the next set, then three MFMAs, sixteen times
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 3, 0);
            }
        };
        // set 1 is read while set 0 is multiplied, then becomes set 0
        for (int kt = 0; kt < nk; ++kt) {
            step(kt, 0, 1);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                ah[0][t] = ah[1][t];
                al[0][t] = al[1][t];
                bh[0][t] = bh[1][t];
                bl[0][t] = bl[1][t];
            }
        }

        // epilogue: all C loads of a 32-row slab are issued before its stores
        const float sgn = g.alpha_neg ? -1.f : 1.f;
        float *__restrict__ C =
            batch_mat(g.bt, g.C, zb) + (bm + wm * 128 + 4 * (lane >> 5)) * g.ldc + bn + wn * 128 + r32;
        const int mrem = (int)min((long long)BM, g.M - bm) - wm * 128 - 4 * (lane >> 5);
        const int nrem = (int)min((long long)BN, (long long)g.Nr - bn) - wn * 128 - r32;
        const bool full = bm + BM <= g.M && bn + BN <= g.Nr;
        auto epilogue = [&](auto guarded) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                float cv[4][16];
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                        const bool in = !guarded.value || (i < mrem && j < nrem);
                        cv[nt][v] = (g.beta && in) ? C[i * g.ldc + j] : 0.f;
                    }
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int i = mt * 32 + 8 * (v >> 2) + (v & 3), j = nt * 32;
                        if (!guarded.value || (i < mrem && j < nrem)) C[i * g.ldc + j] = cv[nt][v] + sgn * acc[mt][nt][v];
                    }
            }
        };
        if (full)
            epilogue(std::false_type{});
        else
            epilogue(std::true_type{});
    }  // tiles
}

// ---------------------------------------------------------------------------
namespace f64k {
constexpr int BM = 128, BN = 128, BK = 16, PAD = 2;
}

template <typename TA, int TAG>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(GemmArgs g) {
    using namespace f64k;
    __shared__ double As[2][BM][BK + PAD];
    __shared__ double Bs[2][BN][BK + PAD];
    const TA *__restrict__ A = (const TA *)g.A;
    const double *__restrict__ Bt = (const double *)g.Bt;
    double *C = (double *)g.C;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    int ty, tx;
    tile_of(g.swizzle, &ty, &tx);
    const long long bm = (long long)ty * BM, bn = (long long)tx * BN;
    // split K: slice blockIdx.z covers [kb, kend) and writes its own C plane
    const int kb = g.kslices > 1 ? (int)blockIdx.z * g.kc : 0;
    const int kend = g.kslices > 1 ? min(g.Kr, kb + g.kc) : g.Kr;
    const int nk = (kend - kb + BK - 1) / BK;
    C += (size_t)blockIdx.z * g.cslice;

    // A stays in its own precision in registers and is converted on its way
    // to LDS: converting right after the load made every K step wait for its
    // global loads before its MFMAs (FP32 A: 35 % of the FP64 MFMA peak).
    using A2 = std::conditional_t<sizeof(TA) == 4, float2, double2>;
    A2 ra[4];
    double2 rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int idx = tid + 256 * i;
            int row = idx >> 3, kk = (idx & 7) * 2;
            long long gr = bm + row;
            int gk = kb + k0 + kk;
            if (gr < g.M && gk < kend) {
                ra[i] = *(const A2 *)(A + gr * g.lda + gk);
            } else {
                ra[i].x = TA(0);
                ra[i].y = TA(0);
            }
            long long gc = bn + row;
            rb[i] = (gc < g.Nr && gk < kend) ? *(const double2 *)(Bt + gc * g.ldb + gk) : make_double2(0.0, 0.0);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int idx = tid + 256 * i;
            int row = idx >> 3, kk = (idx & 7) * 2;
            As[buf][row][kk] = (double)ra[i].x;
            As[buf][row][kk + 1] = (double)ra[i].y;
            Bs[buf][row][kk] = rb[i].x;
            Bs[buf][row][kk + 1] = rb[i].y;
        }
    };

    doublex4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;

    gload(0);
    sstore(0);
    __syncthreads();
    const int r16 = lane & 15, kg = 4 * (lane >> 4);
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * BK);
        double av[4][4], bv[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const double2 *src = (const double2 *)&As[buf][wm * 64 + mt * 16 + r16][kg];
            double2 v0 = src[0], v1 = src[1];
            av[mt][0] = v0.x;
            av[mt][1] = v0.y;
            av[mt][2] = v1.x;
            av[mt][3] = v1.y;
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const double2 *src = (const double2 *)&Bs[buf][wn * 64 + nt * 16 + r16][kg];
            double2 v0 = src[0], v1 = src[1];
            bv[nt][0] = v0.x;
            bv[nt][1] = v0.y;
            bv[nt][2] = v1.x;
            bv[nt][3] = v1.y;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mt][t], bv[nt][t], acc[mt][nt], 0, 0, 0);
        if (kt + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }

    // Per 16-row slab, all loads of C before the first store (a store
    // through C may alias a later load, which would otherwise serialise a
    // round trip per element); slab by slab keeps the registers for two
    // waves per SIMD.
    const double sgn = g.alpha_neg ? -1.0 : 1.0;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        double old[4][4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // V_MFMA_F64_16X16X4_F64 result: lane l, element v holds row l/16 + 4v, column l%16
                const long long i = bm + wm * 64 + mt * 16 + (lane >> 4) + 4 * v;
                const long long j = bn + wn * 64 + nt * 16 + r16;
                old[nt][v] = (g.beta && i < g.M && j < g.Nr) ? C[i * g.ldc + j] : 0.0;
            }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long i = bm + wm * 64 + mt * 16 + (lane >> 4) + 4 * v;
                const long long j = bn + wn * 64 + nt * 16 + r16;
                if (i < g.M && j < g.Nr) C[i * g.ldc + j] = old[nt][v] + sgn * acc[mt][nt][v];
            }
    }
}

static bool gemm_args_ok(const GemmArgs &g) {
    return g.M > 0 && g.Nr > 0 && g.Kr > 0 && (g.Kr % 2) == 0 && (g.lda % 2) == 0 && (g.ldb % 2) == 0 &&
           g.Nr <= 65535LL * 128 && g.M <= 65535LL * 128;
}

static int swizzle_default() {
    static const int v = [] {
        const char *e = getenv("EMFA_GEMM_SWIZZLE");
        return e ? atoi(e) : 1;
    }();
    return v;
}

emfa_status gemm_launch_f32(hipStream_t s, const GemmArgs &g0, int tag) {
    GemmArgs g = g0;
    g.swizzle = swizzle_default();
    EMFA_REQUIRE(gemm_args_ok(g) && g.bt.count >= 1 && g.bt.count <= 65535, EMFA_E_INVALID,
                 "gemm f32: bad shape M=%d Nr=%d Kr=%d batch %d", g.M, g.Nr, g.Kr, g.bt.count);
    dim3 grid((g.Nr + f32k::BN - 1) / f32k::BN, (g.M + f32k::BM - 1) / f32k::BM, g.bt.count);
    static const int variant = [] {
        const char *e = getenv("EMFA_GEMM_VARIANT");
        return e ? atoi(e) : 0;
    }();
    const bool vec4 = (g.lda % 4) == 0 && (g.ldb % 4) == 0 && (g.Kr % 4) == 0 &&
                      ((uintptr_t)g.A % 16) == 0 && ((uintptr_t)g.Bt % 16) == 0;
    // launches of fewer than 2048 tiles of 128 x 128 take the 64 x 64 tiles:
    // the in-panel, U-row and solve updates (short K or few columns), where
    // four times the workgroups, four waves per SIMD and C read before the
    // K loop beat the larger tile (C2: in-panel updates 62.5 -> 55.4 ms,
    // solve updates 46.6 -> 40.0 ms); the long-K outer updates keep 128 x
    // 128.  EMFA_GEMM_SMALL sets the tile count (0 disables).
    static const long long small_below = [] {
        const char *e = getenv("EMFA_GEMM_SMALL");
        return e ? std::max(0LL, atoll(e)) : 2048LL;
    }();
    const long long tiles = (long long)grid.x * grid.y * grid.z;
    // tag 1: the LU trailing update, a distinct symbol so profilers can tell it apart
    if (tiles < small_below && variant == 0) {
        dim3 gs((g.Nr + f32s::BN - 1) / f32s::BN, (g.M + f32s::BM - 1) / f32s::BM, g.bt.count);
        if (tag == 1)
            k_gemm_f32s<1><<<gs, 256, 0, s>>>(g);
        else
            k_gemm_f32s<0><<<gs, 256, 0, s>>>(g);
    } else if (vec4 && variant == 1) {
        if (tag == 1)
            k_gemm_f32v<1, 32><<<grid, 256, 0, s>>>(g);
        else
            k_gemm_f32v<0, 32><<<grid, 256, 0, s>>>(g);
    } else if (vec4 && variant == 2) {
        if (tag == 1)
            k_gemm_f32v<1, 64><<<grid, 256, 0, s>>>(g);
        else
            k_gemm_f32v<0, 64><<<grid, 256, 0, s>>>(g);
    } else if (tag == 1) {
        k_gemm_f32<1><<<grid, 256, 0, s>>>(g);
    } else {
        k_gemm_f32<0><<<grid, 256, 0, s>>>(g);
    }
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

emfa_status split_bf16_launch(hipStream_t s, const float *src, long long ld, int rows, int cols, __bf16 *hi,
                              __bf16 *lo, long long ldp, int rows_p, int cols_p, const Batch &bt, bool src_mat) {
    EMFA_REQUIRE(rows >= 0 && cols >= 0 && rows <= rows_p && cols <= cols_p && cols_p % 8 == 0 && ldp >= cols_p &&
                     ldp % 8 == 0 && (ld % 2) == 0 && ((uintptr_t)src % 8) == 0 && ((uintptr_t)hi % 16) == 0 &&
                     ((uintptr_t)lo % 16) == 0,
                 EMFA_E_INVALID, "split_bf16: bad shape rows=%d/%d cols=%d/%d", rows, rows_p, cols, cols_p);
    const long long work = (long long)rows_p * (cols_p / 8);
    if (work == 0) return EMFA_OK;
    k_split_bf16<<<dim3((unsigned)((work + 255) / 256), 1, bt.count), 256, 0, s>>>(src, ld, rows, cols, hi, lo, ldp,
                                                                                  rows_p, cols_p, bt, src_mat ? 1 : 0);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

emfa_status gemm_launch_bf3_planes(hipStream_t s, const Bf3Planes &g) {
    auto al16 = [](const void *p) { return ((uintptr_t)p % 16) == 0; };
    EMFA_REQUIRE(g.M > 0 && g.Nr > 0 && g.Kp > 0 && g.Kp % bf3p::BK == 0 && g.lda >= g.Kp && g.ldb >= g.Kp &&
                     g.lda % 8 == 0 && g.ldb % 8 == 0 && al16(g.Ah) && al16(g.Al) && al16(g.Bh) && al16(g.Bl) &&
                     g.M <= 65535LL * bf3p::BM && g.Nr <= 65535LL * bf3p::BN,
                 EMFA_E_INVALID, "gemm bf16x3 planes: bad shape M=%d Nr=%d Kp=%d", g.M, g.Nr, g.Kp);
    const int TX = (g.Nr + bf3p::BN - 1) / bf3p::BN, TY = (g.M + bf3p::BM - 1) / bf3p::BM;
    EMFA_REQUIRE(g.bt.count >= 1 && g.bt.count <= 65535, EMFA_E_INVALID, "gemm bf16x3 planes: batch %d", g.bt.count);
    Bf3Planes gg = g;
    // the register-staged kernel when K is a multiple of 32, else the
    // direct-load one; EMFA_BF3_KERNEL=p / q select the direct-load or the
    // one-wave-per-SIMD kernel (all give the same bits)
    static const int env_kernel = [] {
        const char *e = getenv("EMFA_BF3_KERNEL");
        return e && e[0] == 'q' ? 2 : e && e[0] == 'p' ? 1 : 4;
    }();
    int kern = g.kernel ? g.kernel : env_kernel;
    if (kern == 4 && g.Kp % bf3r::BK != 0) kern = 1;
    EMFA_REQUIRE(kern == 4 ? (g.lda <= (1LL << 23) && g.ldb <= (1LL << 23)) : true, EMFA_E_INVALID,
                 "gemm bf16x3 planes: plane stride too large");
    const int threads = kern == 2 ? 256 : 512;
    auto launch = [&](dim3 grid) {
        if (kern == 2)
            k_gemm_bf3q<<<grid, threads, 0, s>>>(gg);
        else if (kern == 4)
            k_gemm_bf3r<<<grid, threads, 0, s>>>(gg);
        else
            k_gemm_bf3p<<<grid, threads, 0, s>>>(gg);
    };
    if (g.max_wgs >= 8 && (long long)TX * TY * g.bt.count > g.max_wgs) {
        gg.persistent = 1;
        launch(dim3((g.max_wgs / 8) * 8));
    } else {
        gg.persistent = 0;
        launch(dim3(TX, TY, g.bt.count));
    }
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

emfa_status gemm_launch_f64(hipStream_t s, const GemmArgs &g0, int tag) {
    GemmArgs g = g0;
    g.swizzle = swizzle_default();
    EMFA_REQUIRE(g.bt.count == 1 && !g.bt.mat, EMFA_E_INVALID, "gemm f64: batches are FP32 only");
    EMFA_REQUIRE(gemm_args_ok(g), EMFA_E_INVALID, "gemm f64: bad shape M=%d Nr=%d Kr=%d", g.M, g.Nr, g.Kr);
    dim3 grid((g.Nr + f64k::BN - 1) / f64k::BN, (g.M + f64k::BM - 1) / f64k::BM, std::max(1, g.kslices));
    if (tag == 1)
        k_gemm_f64<double, 1><<<grid, 256, 0, s>>>(g);
    else
        k_gemm_f64<double, 0><<<grid, 256, 0, s>>>(g);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

emfa_status gemm_launch_f32_f64(hipStream_t s, const GemmArgs &g0) {
    GemmArgs g = g0;
    g.swizzle = swizzle_default();
    EMFA_REQUIRE(gemm_args_ok(g) && g.bt.count == 1 && !g.bt.mat, EMFA_E_INVALID,
                 "gemm f32/f64: bad shape M=%d Nr=%d Kr=%d (or a batch)", g.M, g.Nr, g.Kr);
    dim3 grid((g.Nr + f64k::BN - 1) / f64k::BN, (g.M + f64k::BM - 1) / f64k::BM, std::max(1, g.kslices));
    k_gemm_f64<float, 0><<<grid, 256, 0, s>>>(g);
    EMFA_HIP_CHECK(hipGetLastError());
    return EMFA_OK;
}

}  // namespace emfa
