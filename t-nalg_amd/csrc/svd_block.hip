// Singular value decomposition of the n x n factor R of an SVD gauge move
// (n <= 1024), by block one-sided Jacobi over many workgroups.
//
// The columns of M = R^T (the rows of R, contiguous in the row-major R) are
// rotated until they are mutually orthogonal: M W = B with W orthogonal, so
// R = W S U^T with S = column norms of B and U = B S^-1.  The outputs avoid
// S^-1: row k of `u` is the k-th left singular vector of R (a column of W)
// and row k of `v` is s_k times the k-th right singular vector (a column of
// B), in descending order of s_k, so R = u^T v exactly.  A gauge move
// A = Q R then keeps A_new = Q u^T (an isometry even when R is rank
// deficient) and hands v to the neighbour.
//
// Columns are grouped into blocks of bw (a power of two, bw * n <= 4096, so
// two blocks of B and of W fill at most 128 KiB of LDS).  One launch is one
// round of the round-robin over block pairs: a workgroup loads its pair into
// LDS, runs one round-robin sweep of plane rotations over the 2 bw columns
// (bw disjoint column pairs at a time, 256 / bw lanes each) and writes the
// columns back.  Rounds repeat until a sweep rotates nothing.  A column
// whose norm is at most n eps ||R||_F is numerical zero (rounding residue of
// a rank-deficient R) and takes part in no rotation: rotating pairs of such
// columns never settles, and leaving them alone keeps R = u^T v exact and u
// orthogonal.
//
// A job with u == 0 asks for the singular values only (the entanglement
// spectrum of a bond whose gauge move was a QR): W is neither formed nor
// rotated, the work area is n^2 + n + 1 doubles (B, column norms, scale),
// and with one block of B per pair there is room for twice the block width
// (bw n <= 8192).  R may then be the work area itself (r == work, ld_r == n):
// forming B is a copy onto itself.
#include "tnx_common.h"

namespace {

constexpr int NT = 256;
constexpr int SJ_NMAX = 1024;
constexpr size_t SJ_LDS = sizeof(double) * 4 * 4096;  // 2 (B, W) x 2 bw x n, bw n <= 4096

// values-only job: no W (see the header)
__device__ __forceinline__ bool values_only(const tnx_svdj_job& J) { return J.u == 0; }
// column norms and the scale ||R||_F^2 follow B (and W, when formed)
__device__ __forceinline__ double* norms_of(const tnx_svdj_job& J) {
  return (double*)J.work + (values_only(J) ? 1 : 2) * (int64_t)J.n * J.n;
}

__device__ __forceinline__ void pair_of(int players, int round, int p, int& a, int& b) {
  if (p == 0) {
    a = players - 1;
    b = round;
  } else {
    a = (round + p) % (players - 1);
    b = (round + players - 1 - p) % (players - 1);
  }
}

// B (column j = row j of R) and W = I, both column-major n x n in the work area.
__global__ __launch_bounds__(NT) void k_svdj_init(const tnx_svdj_job* __restrict__ jobs) {
  const tnx_svdj_job J = jobs[blockIdx.y];
  const int n = J.n;
  const double* R = (const double*)J.r;
  double* B = (double*)J.work;
  double* W = B + (int64_t)n * n;
  const bool vo = values_only(J);
  for (int j = blockIdx.x; j < n; j += gridDim.x)
    for (int i = threadIdx.x; i < n; i += NT) {
      B[(int64_t)j * n + i] = R[(int64_t)j * J.ld_r + i];
      if (!vo) W[(int64_t)j * n + i] = (i == j) ? 1.0 : 0.0;
    }
}

// ||R||_F^2 of each job into work[2 n^2 + n] (one workgroup per job, a fixed
// summation order, so the zero test below does not depend on the batch).
__global__ __launch_bounds__(NT) void k_svdj_scale(const tnx_svdj_job* __restrict__ jobs) {
  const tnx_svdj_job J = jobs[blockIdx.x];
  const int64_t nn = (int64_t)J.n * J.n;
  double* B = (double*)J.work;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < nn; i += NT) {
    const double b = B[i];
    s += b * b;
  }
  s = tnx_wave_sum(s);
  __shared__ double red[NT / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) norms_of(J)[J.n] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_svdj_clear(int32_t* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

__global__ __launch_bounds__(NT) void k_svdj_round(const tnx_svdj_job* __restrict__ jobs, int njob,
                                                   int round, int32_t* flag) {
  extern __shared__ double lds[];
  const int ji = tnx_find_owner(jobs, sizeof(tnx_svdj_job), offsetof(tnx_svdj_job, pair_begin),
                                njob, blockIdx.x);
  const tnx_svdj_job J = jobs[ji];
  const int nb = J.nblk, n = J.n, bw = J.bw, w2 = 2 * J.bw;
  if (round >= nb - 1) return;  // this job has fewer blocks than the largest one
  const int p = blockIdx.x - J.pair_begin;
  int ba, bb;
  pair_of(nb, round, p, ba, bb);
  double* B = (double*)J.work;
  double* W = B + (int64_t)n * n;
  const bool vo = values_only(J);
  double* Bs = lds;
  double* Ws = lds + (int64_t)w2 * n;
  // local column c: block ba for c < bw, block bb after; columns >= n are zero
  for (int c = 0; c < w2; ++c) {
    const int g = (c < bw ? ba * bw + c : bb * bw + (c - bw));
    for (int i = threadIdx.x; i < n; i += NT) {
      Bs[c * n + i] = (g < n) ? B[(int64_t)g * n + i] : 0.0;
      if (!vo) Ws[c * n + i] = (g < n) ? W[(int64_t)g * n + i] : 0.0;
    }
  }
  __shared__ int rotated;
  if (threadIdx.x == 0) rotated = 0;
  __syncthreads();
  const int T = NT / bw;  // lanes per column pair: 8 .. 64, aligned groups of one wavefront
  const int k = threadIdx.x / T, l = threadIdx.x % T;
  const double eps = 2.220446049250313e-16 * (double)(n > 8 ? n : 8);
  const double zero2 = eps * eps * norms_of(J)[n];  // column norm^2 of numerical zero
  for (int ir = 0; ir < w2 - 1; ++ir) {
    int u, v;
    pair_of(w2, ir, k, u, v);
    double* x = Bs + u * n;
    double* y = Bs + v * n;
    double a = 0.0, c = 0.0, g = 0.0;
    // loops unrolled so that several LDS loads are in flight (same operations
    // in the same order per element)
#pragma unroll 4
    for (int i = l; i < n; i += T) {
      const double xi = x[i], yi = y[i];
      a += xi * xi;
      c += yi * yi;
      g += xi * yi;
    }
    for (int o = T >> 1; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      c += __shfl_xor(c, o, 64);
      g += __shfl_xor(g, o, 64);
    }
    if (g != 0.0 && fabs(g) > eps * sqrt(a * c) && a > zero2 && c > zero2) {
      const double zeta = (c - a) / (2.0 * g);
      const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
      const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
      double* wx = Ws + u * n;
      double* wy = Ws + v * n;
#pragma unroll 4
      for (int i = l; i < n; i += T) {
        const double xi = x[i], yi = y[i];
        x[i] = cs * xi - sn * yi;
        y[i] = sn * xi + cs * yi;
        if (!vo) {
          const double pi = wx[i], qi = wy[i];
          wx[i] = cs * pi - sn * qi;
          wy[i] = sn * pi + cs * qi;
        }
      }
      if (l == 0) rotated = 1;
    }
    __syncthreads();
  }
  for (int c = 0; c < w2; ++c) {
    const int g = (c < bw ? ba * bw + c : bb * bw + (c - bw));
    if (g < n)
      for (int i = threadIdx.x; i < n; i += NT) {
        B[(int64_t)g * n + i] = Bs[c * n + i];
        if (!vo) W[(int64_t)g * n + i] = Ws[c * n + i];
      }
  }
  if (threadIdx.x == 0 && rotated) atomicOr(flag, 1);
}

// su[j] = ||B_j|| (one wavefront per column)
__global__ __launch_bounds__(NT) void k_svdj_norms(const tnx_svdj_job* __restrict__ jobs) {
  const tnx_svdj_job J = jobs[blockIdx.y];
  const int n = J.n;
  const int j = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
  if (j >= n) return;
  const double* B = (const double*)J.work;
  double* su = norms_of(J);
  double s = 0.0;
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    const double b = B[(int64_t)j * n + i];
    s += b * b;
  }
  s = tnx_wave_sum(s);
  if ((threadIdx.x & 63) == 0) su[j] = sqrt(s);
}

// Rank of column j in descending order (ties by index); row rank of u and v
// gets column j of W and of B, s[rank] = su[j].
__global__ __launch_bounds__(NT) void k_svdj_sort(const tnx_svdj_job* __restrict__ jobs) {
  const tnx_svdj_job J = jobs[blockIdx.y];
  const int n = J.n, j = blockIdx.x;
  if (j >= n) return;
  const double* B = (const double*)J.work;
  const double* W = B + (int64_t)n * n;
  const double* su = norms_of(J);
  const double sj = su[j];
  int cnt = 0;
  for (int q = threadIdx.x; q < n; q += NT) {
    const double sq = su[q];
    cnt += (sq > sj || (sq == sj && q < j)) ? 1 : 0;
  }
  __shared__ int red[NT / 64];
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  const int rank = red[0] + red[1] + red[2] + red[3];
  if (!values_only(J)) {
    double* u = (double*)J.u;
    double* v = (double*)J.v;
    for (int i = threadIdx.x; i < n; i += NT) {
      u[(int64_t)rank * n + i] = W[(int64_t)j * n + i];
      v[(int64_t)rank * n + i] = B[(int64_t)j * n + i];
    }
  }
  if (threadIdx.x == 0) ((double*)J.s)[rank] = sj;
}

thread_local int32_t* g_flag_pin = nullptr;

}  // namespace

extern "C" int tnx_svd_block_jacobi(const tnx_svdj_job* jobs, int njob, int npairs, int max_nblk,
                                    int max_n, int32_t* flag, int max_sweeps,
                                    tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || flag == nullptr || njob < 0 || npairs <= 0 || max_nblk < 2 ||
      max_n <= 0 || max_n > SJ_NMAX || max_sweeps < 1) {
    tnx_set_error("tnx_svd_block_jacobi: bad arguments (njob=%d npairs=%d nblk=%d n=%d)", njob,
                  npairs, max_nblk, max_n);
    return TNX_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  static bool attr_set = false;
  if (!attr_set) {
    TNX_CHECK_HIP(hipFuncSetAttribute((const void*)k_svdj_round,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)SJ_LDS));
    attr_set = true;
  }
  if (g_flag_pin == nullptr)
    TNX_CHECK_HIP(hipHostMalloc((void**)&g_flag_pin, 64, hipHostMallocDefault));
  hipLaunchKernelGGL(k_svdj_init, dim3(max_n < 1024 ? max_n : 1024, njob), dim3(NT), 0, s, jobs);
  hipLaunchKernelGGL(k_svdj_scale, dim3(njob), dim3(NT), 0, s, jobs);
  TNX_CHECK_LAUNCH();
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps && !converged; ++sweep) {
    hipLaunchKernelGGL(k_svdj_clear, dim3(1), dim3(64), 0, s, flag);
    for (int r = 0; r < max_nblk - 1; ++r)
      hipLaunchKernelGGL(k_svdj_round, dim3(npairs), dim3(NT), SJ_LDS, s, jobs, njob, r, flag);
    TNX_CHECK_LAUNCH();
    TNX_CHECK_HIP(hipMemcpyAsync(g_flag_pin, flag, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TNX_CHECK_HIP(hipStreamSynchronize(s));
    converged = (*(volatile int32_t*)g_flag_pin == 0);
  }
  hipLaunchKernelGGL(k_svdj_norms, dim3((max_n + 3) / 4, njob), dim3(NT), 0, s, jobs);
  hipLaunchKernelGGL(k_svdj_sort, dim3(max_n, njob), dim3(NT), 0, s, jobs);
  TNX_CHECK_LAUNCH();
  return converged ? TNX_OK : TNX_ENOCONV;
}
