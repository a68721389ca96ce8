// Grouped, K-concatenated GEMM on CDNA4 matrix cores (MFMA), the contraction
// engine of the DMRG effective-Hamiltonian matvec and environment updates.
//
// One launch covers every output matrix ("problem") of a sweep step for all
// solves of a batch; problem p owns the 64x64 output tiles
// [tile_begin_p, tile_begin_{p+1}).  A problem sums several K-segments, each
// an (optionally transposed) A x B product with its own alpha, so the whole
// sum_k L_k x_k or sum_r U_r R_r^T of one local step is one accumulation in
// registers.  Workgroup: 256 lanes = 4 wavefronts of 64, each wavefront owns
// a 32x32 quarter of the tile as 2x2 MFMA 16x16 blocks.
//   f64 : v_mfma_f64_16x16x4_f64   (K step 4,  LDS chunk 16)
//   f32 : v_mfma_f32_16x16x4_f32   (K step 4,  LDS chunk 16)
//   bf16: v_mfma_f32_16x16x32_bf16 (K step 32, LDS chunk 32, f32 accumulate)
#include <type_traits>

#include "tnx_common.h"

namespace {

template <int MODE> struct Cfg;
template <> struct Cfg<TNX_GEMM_F64> {
  typedef double TI; typedef double TC; typedef double ACC; typedef double CT;
  static constexpr int KB = 16;
};
template <> struct Cfg<TNX_GEMM_F32> {
  typedef float TI; typedef float TC; typedef float ACC; typedef float CT;
  static constexpr int KB = 16;
};
template <> struct Cfg<TNX_GEMM_BF16_F32> {
  typedef __bf16 TI; typedef float TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int KB = 32;
};
template <> struct Cfg<TNX_GEMM_BF16_BF16> {
  typedef __bf16 TI; typedef __bf16 TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int KB = 32;
};

template <typename T> __device__ __forceinline__ double to_d(T v) { return (double)v; }
template <> __device__ __forceinline__ double to_d<__bf16>(__bf16 v) { return (double)(float)v; }
template <typename T> __device__ __forceinline__ T from_d(double v) { return (T)v; }
template <> __device__ __forceinline__ __bf16 from_d<__bf16>(double v) { return (__bf16)(float)v; }

constexpr int TM = 64, TN = 64, NT = 256;

// ---- f64 / f32 path: LDS tiles stored k-major: As[k][i], Bs[k][j] ----------
template <int MODE>
__device__ __forceinline__ void load_tiles_k4(typename Cfg<MODE>::CT (*As)[TM + 4],
                                              typename Cfg<MODE>::CT (*Bs)[TN + 4],
                                              const tnx_gemm_seg& S, int M, int N, int row0,
                                              int col0, int k0) {
  typedef typename Cfg<MODE>::TI TI;
  typedef typename Cfg<MODE>::CT CT;
  const int t = threadIdx.x;
  const TI* A = (const TI*)S.a;
  const TI* B = (const TI*)S.b;
  const double alpha = S.alpha;
  // A tile: 64 rows x 16 k
  if (!S.transA) {
    int i = t >> 2, kq = (t & 3) * 4;
    int gi = row0 + i;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int gk = k0 + kq + e;
      double v = (gi < M && gk < S.K) ? to_d(A[(int64_t)gi * S.lda + gk]) : 0.0;
      As[kq + e][i] = (CT)(alpha * v);
    }
  } else {
    int k = t >> 4, iq = (t & 15) * 4;
    int gk = k0 + k;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int gi = row0 + iq + e;
      double v = (gi < M && gk < S.K) ? to_d(A[(int64_t)gk * S.lda + gi]) : 0.0;
      As[k][iq + e] = (CT)(alpha * v);
    }
  }
  // B tile: 16 k x 64 cols
  if (!S.transB) {
    int k = t >> 4, jq = (t & 15) * 4;
    int gk = k0 + k;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int gj = col0 + jq + e;
      double v = (gj < N && gk < S.K) ? to_d(B[(int64_t)gk * S.ldb + gj]) : 0.0;
      Bs[k][jq + e] = (CT)v;
    }
  } else {
    int j = t >> 2, kq = (t & 3) * 4;
    int gj = col0 + j;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      int gk = k0 + kq + e;
      double v = (gj < N && gk < S.K) ? to_d(B[(int64_t)gj * S.ldb + gk]) : 0.0;
      Bs[kq + e][j] = (CT)v;
    }
  }
}

// ---- bf16 path: LDS tiles stored row-major in k: As[i][k], Bs[j][k] ------
__device__ __forceinline__ void load_tiles_bf16(__bf16 (*As)[32 + 8], __bf16 (*Bs)[32 + 8],
                                                const tnx_gemm_seg& S, int M, int N, int row0,
                                                int col0, int k0) {
  const int t = threadIdx.x;
  const __bf16* A = (const __bf16*)S.a;
  const __bf16* B = (const __bf16*)S.b;
  const float alpha = (float)S.alpha;
  if (!S.transA) {
    int i = t >> 2, kq = (t & 3) * 8;
    int gi = row0 + i;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int gk = k0 + kq + e;
      float v = (gi < M && gk < S.K) ? (float)A[(int64_t)gi * S.lda + gk] : 0.0f;
      As[i][kq + e] = (__bf16)(alpha * v);
    }
  } else {
    int k = t >> 3, iq = (t & 7) * 8;
    int gk = k0 + k;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int gi = row0 + iq + e;
      float v = (gi < M && gk < S.K) ? (float)A[(int64_t)gk * S.lda + gi] : 0.0f;
      As[iq + e][k] = (__bf16)(alpha * v);
    }
  }
  if (!S.transB) {
    int k = t >> 3, jq = (t & 7) * 8;
    int gk = k0 + k;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int gj = col0 + jq + e;
      Bs[jq + e][k] = (gj < N && gk < S.K) ? B[(int64_t)gk * S.ldb + gj] : (__bf16)0.0f;
    }
  } else {
    int j = t >> 2, kq = (t & 3) * 8;
    int gj = col0 + j;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      int gk = k0 + kq + e;
      Bs[j][kq + e] = (gj < N && gk < S.K) ? B[(int64_t)gj * S.ldb + gk] : (__bf16)0.0f;
    }
  }
}

template <int MODE>
__global__ __launch_bounds__(NT) void gemm_grouped_kernel(const tnx_gemm_problem* __restrict__ probs,
                                                          int nprob,
                                                          const tnx_gemm_seg* __restrict__ segs,
                                                          const int32_t* __restrict__ active) {
  typedef Cfg<MODE> C;
  typedef typename C::ACC ACC;
  typedef typename C::TC TC;
  const int bid = tnx_xcd_remap(blockIdx.x, gridDim.x);
  const int pi = tnx_find_owner(probs, sizeof(tnx_gemm_problem),
                                offsetof(tnx_gemm_problem, tile_begin), nprob, bid);
  const tnx_gemm_problem P = probs[pi];
  if (P.active >= 0 && active != nullptr && active[P.active] == 0) return;
  const int tiles_n = (P.N + TN - 1) / TN;
  const int t_local = bid - P.tile_begin;
  const int row0 = (t_local / tiles_n) * TM, col0 = (t_local % tiles_n) * TN;
  if (row0 >= P.M) return;  // stray block past this problem's tiles (cannot happen with exact prefix sums)

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;

  typedef typename std::conditional<sizeof(ACC) == 8, tnx_d4, tnx_f4>::type V4;
  V4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = V4{0, 0, 0, 0};

  if constexpr (MODE == TNX_GEMM_F64 || MODE == TNX_GEMM_F32) {
    typedef typename C::CT CT;
    __shared__ CT As[16][TM + 4];
    __shared__ CT Bs[16][TN + 4];
    for (int s = P.seg_begin; s < P.seg_begin + P.seg_count; ++s) {
      const tnx_gemm_seg S = segs[s];
      for (int k0 = 0; k0 < S.K; k0 += 16) {
        __syncthreads();
        load_tiles_k4<MODE>(As, Bs, S, P.M, P.N, row0, col0, k0);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          CT a0 = As[kk + (lane >> 4)][wr + (lane & 15)];
          CT a1 = As[kk + (lane >> 4)][wr + 16 + (lane & 15)];
          CT b0 = Bs[kk + (lane >> 4)][wc + (lane & 15)];
          CT b1 = Bs[kk + (lane >> 4)][wc + 16 + (lane & 15)];
          if constexpr (MODE == TNX_GEMM_F64) {
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
          } else {
            acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
          }
        }
      }
    }
  } else {
    __shared__ __bf16 As[TM][32 + 8];
    __shared__ __bf16 Bs[TN][32 + 8];
    for (int s = P.seg_begin; s < P.seg_begin + P.seg_count; ++s) {
      const tnx_gemm_seg S = segs[s];
      for (int k0 = 0; k0 < S.K; k0 += 32) {
        __syncthreads();
        load_tiles_bf16(As, Bs, S, P.M, P.N, row0, col0, k0);
        __syncthreads();
        const int kq = (lane >> 4) * 8;
        tnx_bf8 a0 = *(const tnx_bf8*)&As[wr + (lane & 15)][kq];
        tnx_bf8 a1 = *(const tnx_bf8*)&As[wr + 16 + (lane & 15)][kq];
        tnx_bf8 b0 = *(const tnx_bf8*)&Bs[wc + (lane & 15)][kq];
        tnx_bf8 b1 = *(const tnx_bf8*)&Bs[wc + 16 + (lane & 15)][kq];
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
      }
    }
  }

  // epilogue: column lane%16 of each 16x16 block; the row of accumulator
  // element r is 4*(lane/16)+r for the f32-output MFMAs and (lane/16)+4*r for
  // v_mfma_f64_16x16x4_f64 (measured on gfx950)
  TC* Cp = (TC*)P.c;
  const TC* C0 = (const TC*)P.c0;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = (MODE == TNX_GEMM_F64) ? ((lane >> 4) + 4 * r) : ((lane >> 4) * 4 + r);
        int gi = row0 + wr + mi * 16 + rr;
        int gj = col0 + wc + nj * 16 + (lane & 15);
        if (gi < P.M && gj < P.N) {
          double v = (double)acc[mi][nj][r];
          if (C0 != nullptr) v += P.beta * to_d(C0[(int64_t)gi * P.ldc0 + gj]);
          Cp[(int64_t)gi * P.ldc + gj] = from_d<TC>(v);
        }
      }
}

}  // namespace

extern "C" int tnx_gemm_grouped(int mode, const tnx_gemm_problem* prob, int nprob,
                                const tnx_gemm_seg* seg, int ntiles, const int32_t* active,
                                tnx_stream_t stream) {
  if (nprob < 0 || ntiles < 0 || prob == nullptr || seg == nullptr) {
    if (nprob == 0 && ntiles == 0) return TNX_OK;
    tnx_set_error("tnx_gemm_grouped: bad arguments (nprob=%d ntiles=%d)", nprob, ntiles);
    return TNX_EINVAL;
  }
  if (ntiles == 0) return TNX_OK;
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(ntiles), block(NT);
  switch (mode) {
    case TNX_GEMM_F64:
      hipLaunchKernelGGL(gemm_grouped_kernel<TNX_GEMM_F64>, grid, block, 0, s, prob, nprob, seg, active);
      break;
    case TNX_GEMM_F32:
      hipLaunchKernelGGL(gemm_grouped_kernel<TNX_GEMM_F32>, grid, block, 0, s, prob, nprob, seg, active);
      break;
    case TNX_GEMM_BF16_F32:
      hipLaunchKernelGGL(gemm_grouped_kernel<TNX_GEMM_BF16_F32>, grid, block, 0, s, prob, nprob, seg, active);
      break;
    case TNX_GEMM_BF16_BF16:
      hipLaunchKernelGGL(gemm_grouped_kernel<TNX_GEMM_BF16_BF16>, grid, block, 0, s, prob, nprob, seg, active);
      break;
    default:
      tnx_set_error("tnx_gemm_grouped: unknown mode %d", mode);
      return TNX_EINVAL;
  }
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}
