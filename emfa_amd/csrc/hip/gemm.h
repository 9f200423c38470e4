// MFMA GEMM for complex updates written as real products.
//
// A complex product C -= A B (A: M x K, B: K x Nc) on interleaved storage is
// the real product C' -= A' Bt'^T with
//   A'  = A viewed as M x 2K reals (row-major, interleaved re/im),
//   Bt' = 2Nc x 2K reals, Bt'[2j][2k] = Re B(k,j),  Bt'[2j][2k+1] = -Im B(k,j),
//                         Bt'[2j+1][2k] = Im B(k,j), Bt'[2j+1][2k+1] = Re B(k,j),
//   C'  = C viewed as M x 2Nc reals.
// Both operands are K-contiguous, which is the order the MFMA operand fetch
// wants from LDS.  FP32 uses V_MFMA_F32_32X32X2_F32, FP64 (and the mixed
// FP32-matrix / FP64-accumulate residual) uses V_MFMA_F64_16X16X4_F64.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace emfa {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

struct GemmArgs {
    int M, Nr, Kr;          // rows, real columns, real inner dimension
    const void *A;          // TA*, row-major, row stride lda (elements)
    long long lda;
    const void *Bt;         // TC*, row-major 2Nc x Kr, row stride ldb
    long long ldb;
    void *C;                // TC*, row-major, row stride ldc
    long long ldc;
    int alpha_neg;          // 1: C = beta C - A B ; 0: C = beta C + A B
    int beta;               // 0 or 1
    int swizzle = 1;        // XCD-aware tile order (set by the launchers)
    int kslices = 1;        // FP64 launchers: split K over blockIdx.z (C += z * cslice)
    int kc = 0;             // reals of K per slice (multiple of 64)
    long long cslice = 0;   // C elements between slice planes
    Batch bt;               // FP32 launchers: A and C in the batch's matrices, Bt in its scratch slabs
};

// Three-term bf16 product from pre-split operand planes.  Each operand is
// given as two bf16 planes (hi = bf16(x), lo = bf16(x - hi)), row-major with
// ld elements per row; rows are padded with zeros to a multiple of 256
// from the given pointer and K to a multiple of 16 (Kp), so the kernel
// reads whole tiles with no bounds checks.  C (FP32, M x Nr, ldc) gets
// beta C -/+ (Ah Bh^T + Ah Bl^T + Al Bh^T) on rows < M, columns < Nr.
struct Bf3Planes {
    int M, Nr, Kp;
    const __bf16 *Ah, *Al;
    long long lda;
    const __bf16 *Bh, *Bl;
    long long ldb;
    float *C;
    long long ldc;
    int alpha_neg, beta;
    int max_wgs = 0;     // > 0: at most this many workgroups (rounded down to a multiple of 8), looping over tiles
    int persistent = 0;  // set by the launcher
    int kernel = 0;      // 1: k_gemm_bf3p, 2: k_gemm_bf3q, 4: k_gemm_bf3r, 0: EMFA_BF3_KERNEL (p, q) or bf3r
    Batch bt;            // C in the batch's matrices, the planes in its scratch slabs
};

// Split rows x cols FP32 (row stride ld, 8-byte aligned rows) into hi/lo bf16
// planes of rows_p x cols_p (row stride ldp), zero outside rows x cols.
// With a batch, hi / lo are scratch and src is a matrix (src_mat) or scratch.
emfa_status split_bf16_launch(hipStream_t s, const float *src, long long ld, int rows, int cols, __bf16 *hi,
                              __bf16 *lo, long long ldp, int rows_p, int cols_p, const Batch &bt = Batch(),
                              bool src_mat = true);
emfa_status gemm_launch_bf3_planes(hipStream_t s, const Bf3Planes &g);

emfa_status gemm_launch_f32(hipStream_t s, const GemmArgs &g, int tag = 0);  // TA=float,  TC=float
emfa_status gemm_launch_f64(hipStream_t s, const GemmArgs &g, int tag = 0);  // TA=double, TC=double
emfa_status gemm_launch_f32_f64(hipStream_t s, const GemmArgs &g);   // TA=float,  TC=double

}  // namespace emfa
