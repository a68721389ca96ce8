// Elementwise kernels of the DMRG step: the physical-index mixer
//   dst[a,s,b] (=|+=) sum_terms coef * sum_t O[s,t] src[a,t,b]
// (which is also the linear combination of environment matrices and the
// dtype cast), batched traces / Frobenius products, the matvec composition
// and the library's error/version entry points.
#include <stdarg.h>

#include "tnx_common.h"

static thread_local std::string g_last_error;

void tnx_set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

namespace {

constexpr int MIX_DMAX = 4;   // largest physical dimension (the operator table holds 4 x 4)

// (a, b) pairs per lane and per workgroup of a job of physical dimension d:
// jobs with d <= 2 (spin-1/2 sites) take 4 pairs per lane, larger d 2 (the
// registers of a lane hold PPT x d sources and PPT x d sums either way).
__host__ __device__ constexpr int mix_ppt(int d) { return d <= 2 ? 4 : 2; }

// One lane per (a, b) pair, PPT pairs per lane: the d source values
// src[a, :, b] of a term are loaded once and mixed by the d x d operator in
// registers, so each output element costs one load per term.  Terms are the
// outer loop: a term's record and operator are read once per lane, and the
// loads of the lane's pairs are issued together.
template <int DM>
__device__ __forceinline__ void mix_body(const tnx_mix_job& J, int blk,
                                         const tnx_mix_term* __restrict__ terms,
                                         const double* __restrict__ ops) {
  constexpr int PPT = mix_ppt(DM);
  const int d = J.d, Dr = J.Dr;
  const int npairs = J.Dl * Dr;
  const int base = blk * 256 * PPT;
  int64_t e0[PPT];  // element (a, 0, b) of each pair
  bool ok[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int p = base + q * 256 + threadIdx.x;
    ok[q] = p < npairs;
    const int a = ok[q] ? p / Dr : 0;
    e0[q] = ok[q] ? (int64_t)a * d * Dr + (p - a * Dr) : 0;
  }
  double acc[PPT][DM];
#pragma unroll
  for (int q = 0; q < PPT; ++q)
#pragma unroll
    for (int s = 0; s < DM; ++s) acc[q][s] = 0.0;
  for (int ti = J.term_begin; ti < J.term_begin + J.term_count; ++ti) {
    const tnx_mix_term T = terms[ti];
    double x[PPT][DM];
#pragma unroll
    for (int q = 0; q < PPT; ++q)
#pragma unroll
      for (int t = 0; t < DM; ++t)
        x[q][t] = (ok[q] && t < d) ? tnx_load((const void*)T.src, T.src_dtype,
                                              e0[q] + (int64_t)t * Dr) : 0.0;
    if (T.op < 0) {
#pragma unroll
      for (int q = 0; q < PPT; ++q)
#pragma unroll
        for (int s = 0; s < DM; ++s) acc[q][s] += T.coef * x[q][s];
    } else {
      double O[DM * DM];
      const double* Op = ops + (int64_t)T.op * d * d;
#pragma unroll
      for (int s = 0; s < DM; ++s)
#pragma unroll
        for (int t = 0; t < DM; ++t)
          O[s * DM + t] = (s < d && t < d) ? Op[s * d + t] : 0.0;
#pragma unroll
      for (int q = 0; q < PPT; ++q)
#pragma unroll
        for (int s = 0; s < DM; ++s) {
          if (s < d) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < DM; ++t)
              if (t < d) v += O[s * DM + t] * x[q][t];
            acc[q][s] += T.coef * v;
          }
        }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    if (!ok[q]) continue;
#pragma unroll
    for (int s = 0; s < DM; ++s) {
      if (s < d) {
        const int64_t e = e0[q] + (int64_t)s * Dr;
        double v = acc[q][s];
        if (J.accumulate) v += tnx_load((const void*)J.dst, J.dst_dtype, e);
        tnx_store((void*)J.dst, J.dst_dtype, e, v);
      }
    }
  }
}

// Four consecutive elements [e, e + 4) of a buffer of runtime dtype, as
// double; the caller guarantees that e is a multiple of 4 and the buffer
// 4-element aligned, so the load is one 8-, 16- or two 16-byte accesses.
// The conversions are tnx_load's.
__device__ __forceinline__ void tnx_load4(const void* p, int dtype, int64_t e, double (&v)[4]) {
  if (dtype == TNX_F64) {
    const double2 a = *(const double2*)((const double*)p + e);
    const double2 b = *(const double2*)((const double*)p + e + 2);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
  } else if (dtype == TNX_F32) {
    const float4 a = *(const float4*)((const float*)p + e);
    v[0] = (double)a.x; v[1] = (double)a.y; v[2] = (double)a.z; v[3] = (double)a.w;
  } else {
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    const bf4 a = *(const bf4*)((const __bf16*)p + e);
    v[0] = (double)(float)a[0]; v[1] = (double)(float)a[1];
    v[2] = (double)(float)a[2]; v[3] = (double)(float)a[3];
  }
}

__device__ __forceinline__ void tnx_store4(void* p, int dtype, int64_t e, const double (&v)[4]) {
  if (dtype == TNX_F64) {
    *(double2*)((double*)p + e) = make_double2(v[0], v[1]);
    *(double2*)((double*)p + e + 2) = make_double2(v[2], v[3]);
  } else if (dtype == TNX_F32) {
    *(float4*)((float*)p + e) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
  } else {
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    bf4 a;
    a[0] = (__bf16)(float)v[0]; a[1] = (__bf16)(float)v[1];
    a[2] = (__bf16)(float)v[2]; a[3] = (__bf16)(float)v[3];
    *(bf4*)((__bf16*)p + e) = a;
  }
}

__device__ __forceinline__ bool aligned4(const void* p, int dtype) {
  const int isz = dtype == TNX_F64 ? 8 : dtype == TNX_F32 ? 4 : 2;
  return ((uintptr_t)p & (uintptr_t)(4 * isz - 1)) == 0;
}

// d <= 2 jobs whose rows hold a multiple of 4 pairs (Dr % 4 == 0): lane t
// takes the 4 consecutive pairs [base + 4t, base + 4t + 4), all in one row,
// so each source row of a term is one vector load per lane and a wave reads
// 256 consecutive elements.  Same arithmetic per element as mix_body (terms
// in order, the operator applied to the d sources of the pair), so the same
// bits.  A term or destination that is not 4-element aligned takes scalar
// accesses.
__device__ __forceinline__ void mix_body_vec2(const tnx_mix_job& J, int blk,
                                              const tnx_mix_term* __restrict__ terms,
                                              const double* __restrict__ ops) {
  const int d = J.d, Dr = J.Dr;
  const int npairs = J.Dl * Dr;
  const int p0 = blk * 256 * 4 + 4 * threadIdx.x;
  if (p0 >= npairs) return;
  const int a = p0 / Dr;
  const int64_t e0 = (int64_t)a * d * Dr + (p0 - a * Dr);
  double acc[2][4];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[s][q] = 0.0;
  for (int ti = J.term_begin; ti < J.term_begin + J.term_count; ++ti) {
    const tnx_mix_term T = terms[ti];
    double x[2][4];
    const bool vec = aligned4((const void*)T.src, T.src_dtype);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      if (t < d) {
        if (vec) {
          tnx_load4((const void*)T.src, T.src_dtype, e0 + (int64_t)t * Dr, x[t]);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            x[t][q] = tnx_load((const void*)T.src, T.src_dtype, e0 + (int64_t)t * Dr + q);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) x[t][q] = 0.0;
      }
    }
    if (T.op < 0) {
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[s][q] += T.coef * x[s][q];
    } else {
      double O[4];
      const double* Op = ops + (int64_t)T.op * d * d;
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t) O[s * 2 + t] = (s < d && t < d) ? Op[s * d + t] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          if (s < d) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < 2; ++t)
              if (t < d) v += O[s * 2 + t] * x[t][q];
            acc[s][q] += T.coef * v;
          }
        }
    }
  }
  const bool vec = aligned4((const void*)J.dst, J.dst_dtype);
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    if (s < d) {
      const int64_t e = e0 + (int64_t)s * Dr;
      double v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = acc[s][q];
      if (vec) {
        if (J.accumulate) {
          double o[4];
          tnx_load4((const void*)J.dst, J.dst_dtype, e, o);
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] += o[q];
        }
        tnx_store4((void*)J.dst, J.dst_dtype, e, v);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double w = v[q];
          if (J.accumulate) w += tnx_load((const void*)J.dst, J.dst_dtype, e + q);
          tnx_store((void*)J.dst, J.dst_dtype, e + q, w);
        }
      }
    }
  }
}

// TAG only names the instance: 1 for the matvec's mixer (so profiles can tell
// its launches from the environment, cast and copy mixes), 0 otherwise.
template <int TAG>
__global__ __launch_bounds__(256) void mix_kernel(const tnx_mix_job* __restrict__ jobs, int njob,
                                                  const tnx_mix_term* __restrict__ terms,
                                                  const double* __restrict__ ops,
                                                  const int32_t* __restrict__ active,
                                                  unsigned long long* ts) {
  tnx_stamp_begin(ts);
  const int bid = blockIdx.x;
  const int ji = tnx_find_owner(jobs, sizeof(tnx_mix_job), offsetof(tnx_mix_job, block_begin),
                                njob, bid);
  const tnx_mix_job J = jobs[ji];
  if (J.active >= 0 && active != nullptr && active[J.active] == 0) return;
  if (J.d <= 2 && J.Dr % 4 == 0)
    mix_body_vec2(J, bid - J.block_begin, terms, ops);
  else if (J.d <= 2)
    mix_body<2>(J, bid - J.block_begin, terms, ops);
  else
    mix_body<MIX_DMAX>(J, bid - J.block_begin, terms, ops);
  tnx_stamp_end(ts);
}

__global__ __launch_bounds__(256) void trace_dot_kernel(const tnx_dot_job* __restrict__ jobs,
                                                        double* __restrict__ out) {
  const tnx_dot_job J = jobs[blockIdx.x];
  const double* X = (const double*)J.x;
  const double* Y = (const double*)J.y;
  double acc = 0.0;
  if (Y == nullptr) {
    const int nd = J.rows < J.cols ? J.rows : J.cols;
    for (int i = threadIdx.x; i < nd; i += 256) acc += X[(int64_t)i * J.ldx + i];
  } else {
    const int64_t tot = (int64_t)J.rows * J.cols;
    for (int64_t e = threadIdx.x; e < tot; e += 256) {
      int64_t r = e / J.cols, c = e - r * J.cols;
      acc += X[r * J.ldx + c] * Y[r * J.ldy + c];
    }
  }
  __shared__ double red[4];
  acc = tnx_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// x /= ||x|| for one job per workgroup: the sum of squares in a fixed order
// (lane-strided partials, then the wave and workgroup sums), then the scale.
__global__ __launch_bounds__(1024) void normalize_kernel(const tnx_norm_job* __restrict__ jobs) {
  const tnx_norm_job J = jobs[blockIdx.x];
  double* x = (double*)J.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < J.n; i += 1024) acc += x[i] * x[i];
  __shared__ double red[16];
  __shared__ double scale;
  acc = tnx_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += red[w];
    scale = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
  }
  __syncthreads();
  const double sc = scale;
  for (int64_t i = threadIdx.x; i < J.n; i += 1024) x[i] *= sc;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): ten rounds of two 32 x 32 -> 64 multiplies on
// a 128-bit counter under a 64-bit key bumped by the Weyl constants.
__device__ __forceinline__ void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint32_t)p1;
    c[2] = n2;
    c[3] = (uint32_t)p0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// dst[i] = normal number (offset + i) of the stream of `seed`: pair q =
// (offset + i) / 2 comes from Philox of counter (q, 0) -> two 53-bit
// uniforms in (0, 1) -> Box-Muller (cos, sin).
__global__ __launch_bounds__(256) void fill_normal_kernel(const tnx_fill_job* __restrict__ jobs) {
  const tnx_fill_job J = jobs[blockIdx.x];
  double* x = (double*)J.dst;
  const uint32_t k0 = (uint32_t)J.seed, k1 = (uint32_t)(J.seed >> 32);
  for (int64_t i = threadIdx.x; i < J.n; i += 256) {
    const uint64_t g = (uint64_t)(J.offset + i);
    const uint64_t q = g >> 1;
    uint32_t c[4] = {(uint32_t)q, (uint32_t)(q >> 32), 0u, 0u};
    philox4x32_10(c, k0, k1);
    const uint64_t a = ((uint64_t)c[0] << 32) | c[1];
    const uint64_t b = ((uint64_t)c[2] << 32) | c[3];
    const double u1 = ((double)(a >> 11) + 0.5) * 1.1102230246251565e-16;   // 2^-53
    const double u2 = ((double)(b >> 11) + 0.5) * 1.1102230246251565e-16;
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    x[i] = (g & 1) ? r * sn : r * cs;
  }
}

}  // namespace

extern "C" int tnx_fill_normal_f64(const tnx_fill_job* jobs, int njob, tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0) {
    tnx_set_error("tnx_fill_normal_f64: bad arguments (njob=%d)", njob);
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(fill_normal_kernel, dim3(njob), dim3(256), 0, (hipStream_t)stream, jobs);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_normalize_f64(const tnx_norm_job* jobs, int njob, tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0) {
    tnx_set_error("tnx_normalize_f64: bad arguments (njob=%d)", njob);
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(normalize_kernel, dim3(njob), dim3(1024), 0, (hipStream_t)stream, jobs);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_version(void) { return TNX_VERSION; }

extern "C" const char* tnx_last_error(void) { return g_last_error.c_str(); }

extern "C" int tnx_abi_sizes(int32_t* sizes, int n) {
  const int32_t all[] = {(int32_t)sizeof(tnx_gemm_seg),      (int32_t)sizeof(tnx_gemm_problem),
                         (int32_t)sizeof(tnx_mix_job),       (int32_t)sizeof(tnx_mix_term),
                         (int32_t)sizeof(tnx_matvec_plan),   (int32_t)sizeof(tnx_lanczos_solve),
                         (int32_t)sizeof(tnx_lanczos_result), (int32_t)sizeof(tnx_qr_job),
                         (int32_t)sizeof(tnx_svd_job),       (int32_t)sizeof(tnx_dot_job),
                         (int32_t)sizeof(tnx_chol_job),      (int32_t)sizeof(tnx_svdj_job),
                         (int32_t)sizeof(tnx_prof_stat),     (int32_t)sizeof(tnx_norm_job),
                         (int32_t)sizeof(tnx_fill_job),      (int32_t)sizeof(tnx_kron_term),
                         (int32_t)sizeof(tnx_kron_job),      (int32_t)sizeof(tnx_sym_job),
                         (int32_t)sizeof(tnx_op)};
  const int na = (int)(sizeof(all) / sizeof(all[0]));
  if (sizes == nullptr) return na;
  for (int i = 0; i < n && i < na; ++i) sizes[i] = all[i];
  return na;
}

static int launch_mix(const tnx_mix_job* job, int njob, const tnx_mix_term* term, int nblocks,
                      const double* ops, const int32_t* active, tnx_stream_t stream, bool matvec,
                      unsigned long long* ts = nullptr) {
  if (nblocks == 0 || njob == 0) return TNX_OK;
  if (job == nullptr || term == nullptr || nblocks < 0 || njob < 0) {
    tnx_set_error("tnx_mix: bad arguments (njob=%d nblocks=%d)", njob, nblocks);
    return TNX_EINVAL;
  }
  if (matvec)
    hipLaunchKernelGGL(mix_kernel<1>, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, job, njob,
                       term, ops, active, ts);
  else
    hipLaunchKernelGGL(mix_kernel<0>, dim3(nblocks), dim3(256), 0, (hipStream_t)stream, job, njob,
                       term, ops, active, ts);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_mix(const tnx_mix_job* job, int njob, const tnx_mix_term* term, int nblocks,
                       const double* ops, const int32_t* active, tnx_stream_t stream) {
  return launch_mix(job, njob, term, nblocks, ops, active, stream, false);
}

extern "C" int tnx_trace_dot(const tnx_dot_job* jobs, int njob, double* out, tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || out == nullptr || njob < 0) {
    tnx_set_error("tnx_trace_dot: bad arguments");
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(trace_dot_kernel, dim3(njob), dim3(256), 0, (hipStream_t)stream, jobs, out);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

// Timing class of a matvec GEMM launch (the variant byte of the mode aside).
static int gemm_class(int mode, int second) {
  switch (mode & 0xff) {
    case TNX_GEMM_F64: return TNX_PC_GEMM_F64_1 + second;
    case TNX_GEMM_F32: return TNX_PC_GEMM_F32_1 + second;
    default: return TNX_PC_GEMM_BF16_1 + second;
  }
}

extern "C" int tnx_heff_matvec(const tnx_matvec_plan* plan, const int32_t* active,
                               tnx_stream_t stream) {
  if (plan == nullptr) {
    tnx_set_error("tnx_heff_matvec: null plan");
    return TNX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool prof = tnx_prof_on();
  tnx_prof_count(gemm_class(plan->mode, 0), plan->work[0]);
  if (plan->mix_nblocks > 0) tnx_prof_count(TNX_PC_MIX, plan->work[1]);
  tnx_prof_count(gemm_class(plan->mode2, 1), plan->work[2]);
  int mk = prof ? tnx_prof_mark(gemm_class(plan->mode, 0), s) : -1;
  int rc = tnx_gemm_grouped_ts(plan->mode, plan->g1_prob, plan->g1_nprob, plan->g1_seg,
                               plan->g1_ntiles, active, stream, tnx_prof_ts(mk));
  if (rc) return rc;
  tnx_prof_done(mk, s, plan->work[0]);
  mk = (prof && plan->mix_nblocks > 0) ? tnx_prof_mark(TNX_PC_MIX, s) : -1;
  rc = launch_mix(plan->mix_job, plan->mix_njob, plan->mix_term, plan->mix_nblocks, plan->ops, active,
                  stream, true, tnx_prof_ts(mk));
  if (rc) return rc;
  tnx_prof_done(mk, s, plan->work[1]);
  mk = prof ? tnx_prof_mark(gemm_class(plan->mode2, 1), s) : -1;
  rc = tnx_gemm_grouped_ts(plan->mode2, plan->g2_prob, plan->g2_nprob, plan->g2_seg,
                           plan->g2_ntiles, active, stream, tnx_prof_ts(mk));
  if (rc) return rc;
  tnx_prof_done(mk, s, plan->work[2]);
  return TNX_OK;
}
