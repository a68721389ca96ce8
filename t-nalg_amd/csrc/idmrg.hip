// Kernels of the infinite-size DMRG driver that the finite-size kernels do
// not cover: Kronecker sums (the environment blocks with one cell absorbed)
// and in-place symmetrisation of square blocks.
#include "tnx_common.h"

namespace {

constexpr int KRON_NT = 256;

// dst[(i, k), (j, l)] = sum_t coef_t A_t[i, j] B_t[k, l] for one workgroup's
// KRON_NT consecutive elements of one job (row-major dst, ldd per row); a
// null factor is the identity.  Every term is read by every lane, in term
// order, so the sum is formed in the same order for every element.
__global__ __launch_bounds__(KRON_NT) void kron_kernel(const tnx_kron_job* __restrict__ jobs,
                                                       int njob,
                                                       const tnx_kron_term* __restrict__ terms) {
  const int ji = tnx_find_owner(jobs, sizeof(tnx_kron_job), offsetof(tnx_kron_job, block_begin),
                                njob, (int)blockIdx.x);
  const tnx_kron_job J = jobs[ji];
  const int n = J.p * J.q;
  const int64_t e = (int64_t)(blockIdx.x - J.block_begin) * KRON_NT + threadIdx.x;
  if (e >= (int64_t)n * n) return;
  const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
  const int i = r / J.q, k = r - i * J.q;
  const int j = c / J.q, l = c - j * J.q;
  double acc = 0.0;
  for (int t = 0; t < J.term_count; ++t) {
    const tnx_kron_term T = terms[J.term_begin + t];
    const double a = T.a ? ((const double*)T.a)[(int64_t)i * T.lda + j] : (i == j ? 1.0 : 0.0);
    const double b = T.b ? ((const double*)T.b)[(int64_t)k * T.ldb + l] : (k == l ? 1.0 : 0.0);
    acc += T.coef * (a * b);
  }
  ((double*)J.dst)[(int64_t)r * J.ldd + c] = acc;
}

// X <- (X + X^T) / 2 for `count` blocks of one job; the lane of element
// (i, j), i < j, writes both (i, j) and (j, i), the diagonal stays.
__global__ __launch_bounds__(KRON_NT) void symmetrize_kernel(const tnx_sym_job* __restrict__ jobs,
                                                             int njob) {
  const int ji = tnx_find_owner(jobs, sizeof(tnx_sym_job), offsetof(tnx_sym_job, block_begin),
                                njob, (int)blockIdx.x);
  const tnx_sym_job J = jobs[ji];
  const int64_t per = (int64_t)J.n * J.n;
  const int64_t e = (int64_t)(blockIdx.x - J.block_begin) * KRON_NT + threadIdx.x;
  if (e >= per * J.count) return;
  const int m = (int)(e / per);
  const int64_t w = e - (int64_t)m * per;
  const int i = (int)(w / J.n), j = (int)(w - (int64_t)i * J.n);
  if (i >= j) return;
  double* X = (double*)J.x + (int64_t)m * J.stride;
  const double v = 0.5 * (X[(int64_t)i * J.ld + j] + X[(int64_t)j * J.ld + i]);
  X[(int64_t)i * J.ld + j] = v;
  X[(int64_t)j * J.ld + i] = v;
}

}  // namespace

extern "C" int tnx_kron_f64(const tnx_kron_job* jobs, int njob, const tnx_kron_term* terms,
                            int nblocks, tnx_stream_t stream) {
  if (njob == 0 || nblocks == 0) return TNX_OK;
  if (jobs == nullptr || terms == nullptr || njob < 0 || nblocks < 0) {
    tnx_set_error("tnx_kron_f64: bad arguments (njob=%d nblocks=%d)", njob, nblocks);
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(kron_kernel, dim3(nblocks), dim3(KRON_NT), 0, (hipStream_t)stream, jobs,
                     njob, terms);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_symmetrize_f64(const tnx_sym_job* jobs, int njob, int nblocks,
                                  tnx_stream_t stream) {
  if (njob == 0 || nblocks == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0 || nblocks < 0) {
    tnx_set_error("tnx_symmetrize_f64: bad arguments (njob=%d nblocks=%d)", njob, nblocks);
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(symmetrize_kernel, dim3(nblocks), dim3(KRON_NT), 0, (hipStream_t)stream,
                     jobs, njob);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}
