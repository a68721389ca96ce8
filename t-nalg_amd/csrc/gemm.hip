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
//   f64 : v_mfma_f64_16x16x4_f64   (K step 4,  LDS chunk 32, two LDS stages)
//   f32 : v_mfma_f32_16x16x4_f32   (K step 4,  LDS chunk 64, one stage)
//   bf16: v_mfma_f32_16x16x32_bf16 (K step 32, LDS chunk 128, one stage, f32
//         accumulate)
// The K chunks of all segments form one stream; the global loads of the next
// chunk are in flight in registers during the MFMAs of the current one, and
// runs of 16 bytes are loaded as one vector when they are in bounds and
// aligned.
#include <type_traits>

#include "tnx_common.h"

namespace {

// A configuration (kernel template argument CFG) fixes the GEMM mode, the K
// depth KB of one staged chunk, NE (elements of A and of B each lane fetches
// per chunk; 256 x NE = 64 x KB) and DB (LDS double buffering).
// - f64 is MFMA bound at these shapes. Two variants, both with two LDS stages
//   (one barrier per chunk, the commit of chunk c+1 overlapping the MFMAs of
//   chunk c); they differ in depth and occupancy, and neither wins
//   everywhere.  Mid-chain matvec, 64 solves:
//     variant 0, 32-deep chunks, 69.6 KB, 2 workgroups per CU:
//       GEMM 1 (2048 tiles, K 128) 69 us, GEMM 2 (512 tiles, K 512) 64 us;
//     variant 1, 16-deep chunks, 34.8 KB, 4 workgroups per CU:
//       GEMM 1 59 us, GEMM 2 75 us.
//   (One stage of 64: 75 / 69 us.)  The caller picks the variant per launch
//   (the mode's second byte); the engine times both on first use.
// - f32 / bf16 are latency bound at these shapes: one stage of a deep chunk
//   (34.8 KB, four workgroups per CU), two barriers per chunk; a second stage
//   measured slower (f32 GEMM 2: 38 -> 42 us; bf16 GEMM 1: 16 -> 22 us).
// All variants of a mode issue the same MFMAs in the same k order, so they
// give the same bits.
// NT lanes per workgroup as WR x WC wavefronts, each owning a
// (64 / WR) x (64 / WC) part of the tile as 16x16 MFMA blocks.  The 8-wave
// bf16 variants (16x32 per wave) double the waves per workgroup for
// launches of few tiles: mid-chain GEMM 2 (K 512), 16 solves 12.7 -> 11.0
// us, 64 solves 19.9 -> 17.8 us; each output block still sees the same
// MFMAs in the same k order.  Measured and not kept: the same 8-wave split
// for f32 (25.0 -> 25.2 us), 128-deep f32 chunks (25.0 -> 27.7 us) and
// 64-deep f64 chunks with one workgroup per CU (43.1 -> 45.0 us).
template <int CFG> struct Cfg;
template <> struct Cfg<0> {
  typedef double TI; typedef double TC; typedef double ACC; typedef double CT;
  static constexpr int MODE = TNX_GEMM_F64, KB = 32, NE = 8, NT = 256, WR = 2, WC = 2;
  static constexpr bool DB = true;
};
template <> struct Cfg<1> {
  typedef float TI; typedef float TC; typedef float ACC; typedef float CT;
  static constexpr int MODE = TNX_GEMM_F32, KB = 64, NE = 16, NT = 256, WR = 2, WC = 2;
  static constexpr bool DB = false;
};
template <> struct Cfg<2> {
  typedef __bf16 TI; typedef float TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int MODE = TNX_GEMM_BF16_F32, KB = 128, NE = 32, NT = 256, WR = 2, WC = 2;
  static constexpr bool DB = false;
};
template <> struct Cfg<3> {
  typedef __bf16 TI; typedef __bf16 TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int MODE = TNX_GEMM_BF16_BF16, KB = 128, NE = 32, NT = 256, WR = 2, WC = 2;
  static constexpr bool DB = false;
};
template <> struct Cfg<4> {   // f64, variant 1
  typedef double TI; typedef double TC; typedef double ACC; typedef double CT;
  static constexpr int MODE = TNX_GEMM_F64, KB = 16, NE = 4, NT = 256, WR = 2, WC = 2;
  static constexpr bool DB = true;
};
template <> struct Cfg<6> {   // bf16 -> f32, variant 1: 8 waves
  typedef __bf16 TI; typedef float TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int MODE = TNX_GEMM_BF16_F32, KB = 128, NE = 16, NT = 512, WR = 4, WC = 2;
  static constexpr bool DB = false;
};
template <> struct Cfg<7> {   // bf16 -> bf16, variant 1: 8 waves
  typedef __bf16 TI; typedef __bf16 TC; typedef float ACC; typedef __bf16 CT;
  static constexpr int MODE = TNX_GEMM_BF16_BF16, KB = 128, NE = 16, NT = 512, WR = 4, WC = 2;
  static constexpr bool DB = false;
};

template <typename T> __device__ __forceinline__ double to_d(T v) { return (double)v; }
template <> __device__ __forceinline__ double to_d<__bf16>(__bf16 v) { return (double)(float)v; }
template <typename T> __device__ __forceinline__ T from_d(double v) { return (T)v; }
template <> __device__ __forceinline__ __bf16 from_d<__bf16>(double v) { return (__bf16)(float)v; }

constexpr int TM = 64, TN = 64;

// NE contiguous elements starting at p, of which the first `valid` are in
// bounds; out-of-bounds ones read as zero.  Loaded as 16-byte vectors when
// every element is valid and p is 16-byte aligned.
template <typename TI, int NE>
__device__ __forceinline__ void load_run(const TI* p, int valid, TI (&v)[NE]) {
  constexpr int VB = 16 / sizeof(TI);  // elements per 16-byte vector
  if (valid >= NE && (((uintptr_t)p) & 15) == 0) {
#pragma unroll
    for (int q = 0; q < NE / VB; ++q) {
      const uint4 w = *(const uint4*)(p + q * VB);
      TI e[VB];
      __builtin_memcpy(e, &w, 16);
#pragma unroll
      for (int t = 0; t < VB; ++t) v[q * VB + t] = e[t];
    }
  } else {
#pragma unroll
    for (int e = 0; e < NE; ++e) v[e] = (e < valid) ? p[e] : (TI)0.0f;
  }
}

// Per-thread fragment of one K chunk: NE elements of the A tile and NE of the
// B tile, with the addressing mode they were fetched in.
template <int CFG> struct Frag {
  typename Cfg<CFG>::TI a[Cfg<CFG>::NE];
  typename Cfg<CFG>::TI b[Cfg<CFG>::NE];
  double alpha;
  int ta, tb;
};

// Thread t's share of a 64 x KB A tile / KB x 64 B tile (NT threads x NE =
// 64 x KB).
//   transA = 0 (k contiguous): row t/(KB/NE), k = (t%(KB/NE))*NE .. +NE
//   transA = 1 (i contiguous): k  t/(64/NE), i = (t%(64/NE))*NE .. +NE
template <int CFG>
__device__ __forceinline__ void fetch(Frag<CFG>& f, const tnx_gemm_seg& S, int M, int N, int row0,
                                      int col0, int k0) {
  typedef typename Cfg<CFG>::TI TI;
  constexpr int KB = Cfg<CFG>::KB, NE = Cfg<CFG>::NE;
  const int t = threadIdx.x;
  const TI* A = (const TI*)S.a;
  const TI* B = (const TI*)S.b;
  f.alpha = S.alpha;
  f.ta = S.transA;
  f.tb = S.transB;
  if (!S.transA) {
    const int i = t / (KB / NE), kq = (t % (KB / NE)) * NE;
    const int gi = row0 + i, gk = k0 + kq;
    const int valid = (gi < M) ? min(NE, S.K - gk) : 0;
    load_run<TI, NE>(A + (int64_t)(gi < M ? gi : 0) * S.lda + gk, valid, f.a);
  } else {
    const int k = t / (TM / NE), iq = (t % (TM / NE)) * NE;
    const int gi = row0 + iq, gk = k0 + k;
    const int valid = (gk < S.K) ? min(NE, M - gi) : 0;
    load_run<TI, NE>(A + (int64_t)(gk < S.K ? gk : 0) * S.lda + gi, valid, f.a);
  }
  if (!S.transB) {
    const int k = t / (TN / NE), jq = (t % (TN / NE)) * NE;
    const int gj = col0 + jq, gk = k0 + k;
    const int valid = (gk < S.K) ? min(NE, N - gj) : 0;
    load_run<TI, NE>(B + (int64_t)(gk < S.K ? gk : 0) * S.ldb + gj, valid, f.b);
  } else {
    const int j = t / (KB / NE), kq = (t % (KB / NE)) * NE;
    const int gj = col0 + j, gk = k0 + kq;
    const int valid = (gj < N) ? min(NE, S.K - gk) : 0;
    load_run<TI, NE>(B + (int64_t)(gj < N ? gj : 0) * S.ldb + gk, valid, f.b);
  }
}

// f64/f32: LDS tiles k-major, As[k][i], Bs[k][j] (conflict-free MFMA reads)
template <int CFG>
__device__ __forceinline__ void commit_k4(const Frag<CFG>& f, typename Cfg<CFG>::CT (*As)[TM + 4],
                                          typename Cfg<CFG>::CT (*Bs)[TN + 4]) {
  typedef typename Cfg<CFG>::CT CT;
  constexpr int KB = Cfg<CFG>::KB, NE = Cfg<CFG>::NE;
  const int t = threadIdx.x;
  if (!f.ta) {
    const int i = t / (KB / NE), kq = (t % (KB / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) As[kq + e][i] = (CT)(f.alpha * to_d(f.a[e]));
  } else {
    const int k = t / (TM / NE), iq = (t % (TM / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) As[k][iq + e] = (CT)(f.alpha * to_d(f.a[e]));
  }
  if (!f.tb) {
    const int k = t / (TN / NE), jq = (t % (TN / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) Bs[k][jq + e] = (CT)to_d(f.b[e]);
  } else {
    const int j = t / (KB / NE), kq = (t % (KB / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) Bs[kq + e][j] = (CT)to_d(f.b[e]);
  }
}

// bf16: LDS tiles row-major in k, As[i][k], Bs[j][k] (8 consecutive k per lane)
template <int CFG>
__device__ __forceinline__ void commit_bf16(const Frag<CFG>& f, __bf16 (*As)[Cfg<CFG>::KB + 8],
                                            __bf16 (*Bs)[Cfg<CFG>::KB + 8]) {
  constexpr int KB = Cfg<CFG>::KB, NE = Cfg<CFG>::NE;
  const int t = threadIdx.x;
  const float alpha = (float)f.alpha;
  if (!f.ta) {
    const int i = t / (KB / NE), kq = (t % (KB / NE)) * NE;
    if (alpha == 1.0f) {
#pragma unroll
      for (int e = 0; e < NE; ++e) As[i][kq + e] = f.a[e];
    } else {
#pragma unroll
      for (int e = 0; e < NE; ++e) As[i][kq + e] = (__bf16)(alpha * (float)f.a[e]);
    }
  } else {
    const int k = t / (TM / NE), iq = (t % (TM / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) As[iq + e][k] = (__bf16)(alpha * (float)f.a[e]);
  }
  if (!f.tb) {
    const int k = t / (TN / NE), jq = (t % (TN / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) Bs[jq + e][k] = f.b[e];
  } else {
    const int j = t / (KB / NE), kq = (t % (KB / NE)) * NE;
#pragma unroll
    for (int e = 0; e < NE; ++e) Bs[j][kq + e] = f.b[e];
  }
}

template <int CFG>
__global__ __launch_bounds__(Cfg<CFG>::NT) void gemm_grouped_kernel(const tnx_gemm_problem* __restrict__ probs,
                                                          int nprob,
                                                          const tnx_gemm_seg* __restrict__ segs,
                                                          const int32_t* __restrict__ active,
                                                          unsigned long long* ts) {
  tnx_stamp_begin(ts);
  typedef Cfg<CFG> C;
  constexpr int MODE = C::MODE;
  typedef typename C::ACC ACC;
  typedef typename C::TC TC;
  constexpr int KB = C::KB;
  const int bid = tnx_xcd_remap(blockIdx.x, gridDim.x);
  const int pi = tnx_find_owner(probs, sizeof(tnx_gemm_problem),
                                offsetof(tnx_gemm_problem, tile_begin), nprob, bid);
  const tnx_gemm_problem P = probs[pi];
  if (P.active >= 0 && active != nullptr && active[P.active] == 0) return;
  const int tiles_n = (P.N + TN - 1) / TN;
  const int t_local = bid - P.tile_begin;
  const int row0 = (t_local / tiles_n) * TM, col0 = (t_local % tiles_n) * TN;
  if (row0 >= P.M) return;  // cannot happen with exact prefix sums

  static_assert(C::NT * C::NE == TM * KB && C::WR * C::WC * 64 == C::NT, "tile split");
  constexpr int MB = TM / C::WR / 16, NB = TN / C::WC / 16;   // MFMA blocks per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = (wave / C::WC) * (TM / C::WR), wc = (wave % C::WC) * (TN / C::WC);

  typedef typename std::conditional<sizeof(ACC) == 8, tnx_d4, tnx_f4>::type V4;
  V4 acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = V4{0, 0, 0, 0};

  typedef typename std::conditional<(MODE == TNX_GEMM_F64 || MODE == TNX_GEMM_F32),
                                    typename C::CT, __bf16>::type CT;
  constexpr bool K4 = (MODE == TNX_GEMM_F64 || MODE == TNX_GEMM_F32);
  constexpr int AR = K4 ? KB : TM, AC = K4 ? TM + 4 : KB + 8;
  constexpr int BR = K4 ? KB : TN, BC = K4 ? TN + 4 : KB + 8;
  constexpr bool DB = C::DB;
  __shared__ __attribute__((aligned(16))) CT As[DB ? 2 : 1][AR][AC];
  __shared__ __attribute__((aligned(16))) CT Bs[DB ? 2 : 1][BR][BC];

  // Flattened stream of (segment, k0) chunks; (s, k0) is the next chunk to
  // fetch.  Pipeline: chunk c+1 is in registers while chunk c is in LDS;
  // each step runs chunk c's MFMAs, commits chunk c+1 (double buffered: to
  // the other stage, which the barrier of the previous step guarantees every
  // wave has finished reading; single stage: after a barrier), one barrier,
  // then fetches chunk c+2, whose loads fly during the next MFMAs.
  int s = P.seg_begin, k0 = 0;
  const int s_end = P.seg_begin + P.seg_count;
  while (s < s_end && segs[s].K <= 0) ++s;
  Frag<CFG> f;
  tnx_gemm_seg S;
  if (s < s_end) S = segs[s];
  auto advance = [&]() {
    k0 += KB;
    if (k0 >= S.K) {
      k0 = 0;
      ++s;
      while (s < s_end && segs[s].K <= 0) ++s;
      if (s < s_end) S = segs[s];
    }
  };
  bool staged = false;  // chunk c is in LDS
  if (s < s_end) {
    fetch<CFG>(f, S, P.M, P.N, row0, col0, k0);
    advance();
    if constexpr (K4) commit_k4<CFG>(f, As[0], Bs[0]);
    else commit_bf16<CFG>(f, As[0], Bs[0]);
    staged = true;
  }
  __syncthreads();
  bool held = false;  // chunk c+1 is in registers
  if (s < s_end) {
    fetch<CFG>(f, S, P.M, P.N, row0, col0, k0);
    advance();
    held = true;
  }
  int stage = 0;
  while (staged) {
    if constexpr (K4) {
#pragma unroll
      for (int kk = 0; kk < KB; kk += 4) {
        CT a[MB], b[NB];
#pragma unroll
        for (int mi = 0; mi < MB; ++mi) a[mi] = As[stage][kk + (lane >> 4)][wr + 16 * mi + (lane & 15)];
#pragma unroll
        for (int nj = 0; nj < NB; ++nj) b[nj] = Bs[stage][kk + (lane >> 4)][wc + 16 * nj + (lane & 15)];
#pragma unroll
        for (int mi = 0; mi < MB; ++mi)
#pragma unroll
          for (int nj = 0; nj < NB; ++nj) {
            if constexpr (MODE == TNX_GEMM_F64)
              acc[mi][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi], b[nj], acc[mi][nj], 0, 0, 0);
            else
              acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mi], b[nj], acc[mi][nj], 0, 0, 0);
          }
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < KB; kk += 32) {
        const int kq = kk + (lane >> 4) * 8;
        tnx_bf8 a[MB], b[NB];
#pragma unroll
        for (int mi = 0; mi < MB; ++mi) a[mi] = *(const tnx_bf8*)&As[stage][wr + 16 * mi + (lane & 15)][kq];
#pragma unroll
        for (int nj = 0; nj < NB; ++nj) b[nj] = *(const tnx_bf8*)&Bs[stage][wc + 16 * nj + (lane & 15)][kq];
#pragma unroll
        for (int mi = 0; mi < MB; ++mi)
#pragma unroll
          for (int nj = 0; nj < NB; ++nj)
            acc[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mi], b[nj], acc[mi][nj], 0, 0, 0);
      }
    }
    staged = held;
    if constexpr (!DB) __syncthreads();  // every wave has read the single stage
    const int next = DB ? stage ^ 1 : 0;
    if (held) {
      if constexpr (K4) commit_k4<CFG>(f, As[next], Bs[next]);
      else commit_bf16<CFG>(f, As[next], Bs[next]);
    }
    __syncthreads();
    held = false;
    if (staged && s < s_end) {
      fetch<CFG>(f, S, P.M, P.N, row0, col0, k0);
      advance();
      held = true;
    }
    stage = next;
  }

  // epilogue: column lane%16 of each 16x16 block; the row of accumulator
  // element r is 4*(lane/16)+r for the f32-output MFMAs and (lane/16)+4*r for
  // v_mfma_f64_16x16x4_f64 (measured on gfx950)
  TC* Cp = (TC*)P.c;
  const TC* C0 = (const TC*)P.c0;
#pragma unroll
  for (int mi = 0; mi < MB; ++mi)
#pragma unroll
    for (int nj = 0; nj < NB; ++nj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = (MODE == TNX_GEMM_F64) ? ((lane >> 4) + 4 * r) : ((lane >> 4) * 4 + r);
        const int gi = row0 + wr + mi * 16 + rr;
        const int gj = col0 + wc + nj * 16 + (lane & 15);
        if (gi < P.M && gj < P.N) {
          double v = (double)acc[mi][nj][r];
          if (C0 != nullptr) v += P.beta * to_d(C0[(int64_t)gi * P.ldc0 + gj]);
          Cp[(int64_t)gi * P.ldc + gj] = from_d<TC>(v);
        }
      }
  tnx_stamp_end(ts);
}

}  // namespace

extern "C" int tnx_gemm_variants(int mode) {
  switch (mode) {
    case TNX_GEMM_F64: return 2;
    case TNX_GEMM_F32: return 1;
    case TNX_GEMM_BF16_F32:
    case TNX_GEMM_BF16_BF16: return 2;
    default:
      tnx_set_error("tnx_gemm_variants: unknown mode %d", mode);
      return TNX_EINVAL;
  }
}

extern "C" int tnx_gemm_grouped(int mode, const tnx_gemm_problem* prob, int nprob,
                                const tnx_gemm_seg* seg, int ntiles, const int32_t* active,
                                tnx_stream_t stream) {
  return tnx_gemm_grouped_ts(mode, prob, nprob, seg, ntiles, active, stream, nullptr);
}

int tnx_gemm_grouped_ts(int mode, const tnx_gemm_problem* prob, int nprob,
                        const tnx_gemm_seg* seg, int ntiles, const int32_t* active,
                        tnx_stream_t stream, unsigned long long* ts) {
  if (nprob == 0 && ntiles == 0) return TNX_OK;
  if (nprob < 0 || ntiles < 0 || prob == nullptr || seg == nullptr) {
    tnx_set_error("tnx_gemm_grouped: bad arguments (nprob=%d ntiles=%d)", nprob, ntiles);
    return TNX_EINVAL;
  }
  const int base = mode & 0xFF, var = (mode >> 8) & 0xFF;
  const int nvar = tnx_gemm_variants(base);
  if (nvar < 0 || (mode >> 16) != 0 || var >= nvar) {
    tnx_set_error("tnx_gemm_grouped: unknown mode %d (variant %d)", base, var);
    return TNX_EINVAL;
  }
  if (ntiles == 0) return TNX_OK;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(ntiles);
#define TNX_GEMM_LAUNCH(CFG)                                                       \
  hipLaunchKernelGGL(gemm_grouped_kernel<CFG>, grid, dim3(Cfg<CFG>::NT), 0, s, prob, nprob, \
                     seg, active, ts)
  switch (base) {
    case TNX_GEMM_F64:
      if (var == 0) TNX_GEMM_LAUNCH(0);
      else TNX_GEMM_LAUNCH(4);
      break;
    case TNX_GEMM_F32:
      TNX_GEMM_LAUNCH(1);
      break;
    case TNX_GEMM_BF16_F32:
      if (var == 0) TNX_GEMM_LAUNCH(2);
      else TNX_GEMM_LAUNCH(6);
      break;
    default:
      if (var == 0) TNX_GEMM_LAUNCH(3);
      else TNX_GEMM_LAUNCH(7);
      break;
  }
#undef TNX_GEMM_LAUNCH
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}
