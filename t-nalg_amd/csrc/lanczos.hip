// Batched Lanczos eigensolver for the lowest eigenpair of every solve's
// effective Hamiltonian (the local problem of one-site DMRG).
//
// All solves of a batch advance in lockstep over Krylov index j; a solve
// that has converged is masked out (its active flag is 0) so the matvec and
// the kernels below skip it.  Vectors are cut into 1024-element chunks, one
// workgroup per chunk.  A per-solve reduction never synchronises blocks
// inside a kernel: each kernel leaves one partial sum per chunk, and every
// block of the next kernel on the stream (ordered after it) sums the
// partials of its own solve in the same fixed order.  A Lanczos step is
//   matvec (3 launches) + dots + update/dots/norm
//   + update/normalise (plus one block per solve: tridiagonal solve,
//     convergence).
// All reductions are f64 (two-pass classical Gram-Schmidt against the full
// basis); the basis is stored in f64, or in f32 for a solve whose matvec
// runs in f32 or bf16 (v_dtype: the basis passes then move half the bytes).
// The matvec may run in f32 or bf16.
// The (j+1)x(j+1) tridiagonal eigenproblem is solved on the device by 257-way
// multisection of the Sturm count (one workgroup) followed by inverse
// iteration with partial pivoting.
// A solve that has not converged after ncv steps is restarted implicitly
// (Sorensen's implicitly restarted Lanczos with exact shifts, as ARPACK's
// dsaupd does for eigsh): the ncv - k largest Ritz values are applied as
// shifts of implicit QR steps on the tridiagonal matrix, which compresses the
// factorisation to k = min(LZ_KEEP, ncv - 1) vectors spanning the k lowest
// Ritz vectors; the next cycle continues the Lanczos recurrence from step k.
#include <mutex>

#include "tnx_common.h"

namespace {

constexpr int NCV_MAX = 32;
constexpr int CHUNK = 1024;  // 256 lanes x 4 elements
constexpr int NT = 256;
constexpr int LZ_KEEP = 4;   // vectors kept by an implicit restart (at most)
constexpr double LZ_REORTH_SKIP = 1e-12;   // second projection left out below this (relative)
constexpr double LZ_REORTH_SKIP_F32 = 1e-6;  // the same for an f32 basis (its rounding level)

struct LzState {
  double alpha[NCV_MAX];
  double beta[NCV_MAX];
  double h1[NCV_MAX + 1];
  double h2[NCV_MAX + 1];
  double y[NCV_MAX];
  // restart: row t of the new basis row u is C[t][u] (t <= ncv, u <= k)
  double C[NCV_MAX + 1][LZ_KEEP + 1];
  double theta, resid, scale, spare;
  int32_t converged, iters, jlast, restarts;
};

// Vectors an implicit restart keeps for a call with ncv Krylov vectors.
__host__ __device__ inline int lz_keep(int ncv) { return ncv - 1 < LZ_KEEP ? ncv - 1 : LZ_KEEP; }

// partA / partB: per-chunk partial sums, [nchunks][NCV_MAX + 1].  Passes that
// follow each other alternate between them (a kernel reads one while it
// writes the other).
struct LzLayout {
  LzState* st;
  int32_t* act;
  double* partA;
  double* partB;
  int32_t* nact;
};

__host__ __device__ inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

__host__ __device__ inline LzLayout lz_layout(void* state, int nsolve, int nchunks) {
  char* p = (char*)state;
  LzLayout L;
  L.st = (LzState*)p;
  p += align256(sizeof(LzState) * (size_t)nsolve);
  L.act = (int32_t*)p;
  p += align256(sizeof(int32_t) * (size_t)nsolve);
  L.partA = (double*)p;
  p += align256(sizeof(double) * (size_t)nchunks * (NCV_MAX + 1));
  L.partB = (double*)p;
  p += align256(sizeof(double) * (size_t)nchunks * (NCV_MAX + 1));
  L.nact = (int32_t*)p;
  return L;
}

struct ChunkInfo {
  int s;        // solve index
  int c;        // chunk within solve
  int nch;      // chunks of this solve
  int i0, i1;   // element range
};

__device__ __forceinline__ ChunkInfo chunk_of(const tnx_lanczos_solve* solves, int nsolve,
                                           int b = -1) {
  if (b < 0) b = blockIdx.x;   // chunk index (the block index unless offset)
  ChunkInfo ci;
  ci.s = tnx_find_owner(solves, sizeof(tnx_lanczos_solve), offsetof(tnx_lanczos_solve, chunk_begin),
                        nsolve, b);
  const int n = solves[ci.s].n;
  ci.c = b - solves[ci.s].chunk_begin;
  ci.nch = (n + CHUNK - 1) / CHUNK;
  ci.i0 = ci.c * CHUNK;
  ci.i1 = min(n, ci.i0 + CHUNK);
  return ci;
}

// Wavefront sum of s into red[wave][k] (a block sum is red[0..3][k] after a
// barrier).
__device__ __forceinline__ void wave_to_lds(double s, int k, double (*red)[NCV_MAX + 1]) {
  s = tnx_wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = s;
}

// Partial dots V[k] . x over this chunk for k = 0..j, wave sums into
// red[wave][k].  Rows are taken four at a time so their loads are in flight
// together (the loop is latency bound otherwise).
template <typename VT>
__device__ __forceinline__ void chunk_dots(const VT* __restrict__ V, int ldv, int j, int i0,
                                           int i1, const double (&x)[CHUNK / NT],
                                           double (*red)[NCV_MAX + 1]) {
  for (int k0 = 0; k0 <= j; k0 += 4) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      if (k <= j) {
        const VT* row = V + (int64_t)k * ldv;
#pragma unroll
        for (int q = 0; q < CHUNK / NT; ++q) {
          const int i = i0 + q * NT + threadIdx.x;
          if (i < i1) s[u] += (double)row[i] * x[q];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k0 + u <= j) wave_to_lds(s[u], k0 + u, red);
  }
}

// x -= sum_{k<=j} h[k] V[k] over this chunk (rows four at a time).
template <typename VT>
__device__ __forceinline__ void chunk_axpy(const VT* __restrict__ V, int ldv, int j, int i0,
                                           int i1, const double* h, double (&x)[CHUNK / NT]) {
  for (int k0 = 0; k0 <= j; k0 += 4) {
    double v[4][CHUNK / NT];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + u;
      const VT* row = V + (int64_t)(k <= j ? k : j) * ldv;
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) {
        const int i = i0 + q * NT + threadIdx.x;
        v[u][q] = (k <= j && i < i1) ? (double)row[i] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = (k0 + u <= j) ? k0 + u : j;
      const double hk = (k0 + u <= j) ? h[kk] : 0.0;
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) x[q] -= hk * v[u][q];
    }
  }
}

// Block-wide sum of one value per thread into red[0..3][0].
__device__ __forceinline__ void block_sums(const double* v, int nv, double (*red)[NCV_MAX + 1]) {
  (void)nv;
  wave_to_lds(v[0], 0, red);
  __syncthreads();
}

// Sum the partials of one solve (chunks [chunk_begin, chunk_begin + nch))
// for k in [0, nk) into dst[k] (all NT lanes; dst is read after a barrier).
// The partials were written by the previous kernel on the stream.  Wave w
// sums the entries k = w, w + 4, ...: its 64 lanes stride over the chunks
// and four entries are read per pass, so the loads are in flight together.
// The summation order depends only on nch, so every block of a solve gets
// the same bits.
__device__ __forceinline__ void reduce_partials(const double* part, int chunk_begin, int nch,
                                                int nk, double* dst) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k0 = wave; k0 < nk; k0 += 16) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = lane; c < nch; c += 64) {
      const double* p = part + (int64_t)(chunk_begin + c) * (NCV_MAX + 1);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k0 + 4 * u < nk) s[u] += p[k0 + 4 * u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double v = tnx_wave_sum(s[u]);
      if (lane == 0 && k0 + 4 * u < nk) dst[k0 + 4 * u] = v;
    }
  }
}

// Number of eigenvalues of the symmetric tridiagonal (a, b) below x.
__device__ inline int sturm_count(const double* a, const double* b, int k, double x) {
  const double tiny = 1e-300;
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0) cnt++;
  for (int i = 1; i < k; ++i) {
    if (fabs(q) < tiny) q = -tiny;
    q = (a[i] - x) - b[i - 1] * b[i - 1] / q;
    if (q < 0) cnt++;
  }
  return cnt;
}

// Lowest eigenpair of the k x k tridiagonal (a, b), all NT lanes of the
// block (a, b in LDS); thread 0 writes y[0..k) and *theta.  `work` is LDS
// scratch of 5 * NCV_MAX doubles for the inverse iteration.  The bracket
// shrinks 257-fold per round: lane t evaluates the Sturm count at the
// (t + 1)-th of 256 interior points and the first point with a count >= 1
// bounds the lowest eigenvalue.
__device__ void tridiag_lowest(const double* a, const double* b, int k, double* y,
                               double* theta, double* work) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (k == 1) {
    if (threadIdx.x == 0) { y[0] = 1.0; *theta = a[0]; }
    return;
  }
  __shared__ unsigned long long masks[NT / 64];
  double lo = a[0], hi = a[0];
  for (int i = 0; i < k; ++i) {
    double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < k - 1 ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  const double span = fmax(hi - lo, 1e-300);
  lo -= 1e-12 * span;
  hi += 1e-12 * span;
  constexpr double NSUB = NT + 1;
  for (int round = 0; round < 16; ++round) {
    const double w = hi - lo;  // identical in every lane
    if (w <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + 1e-300) break;
    const double x = lo + w * (double)(threadIdx.x + 1) / NSUB;
    const int cnt = sturm_count(a, b, k, x);
    const unsigned long long m = __ballot(cnt >= 1);
    if (lane == 0) masks[wave] = m;
    __syncthreads();
    int f = -1;
    for (int q = 0; q < NT / 64; ++q)
      if (masks[q] != 0ull) {
        f = q * 64 + __ffsll((long long)masks[q]) - 1;
        break;
      }
    __syncthreads();  // masks are rewritten next round
    if (f < 0) {
      lo = lo + w * (double)NT / NSUB;
    } else {
      const double xf = lo + w * (double)(f + 1) / NSUB;
      const double xp = lo + w * (double)f / NSUB;
      hi = xf;
      lo = xp;
    }
  }
  if (threadIdx.x != 0) return;
  const double th = 0.5 * (lo + hi);
  *theta = th;
  // inverse iteration with partial pivoting on (T - th I)
  double* z = work;
  double* rhs = work + NCV_MAX;
  double* u0 = work + 2 * NCV_MAX;
  double* u1 = work + 3 * NCV_MAX;
  double* u2 = work + 4 * NCV_MAX;
  const double tiny = 2.220446049250313e-16 * span;
  for (int i = 0; i < k; ++i) z[i] = 1.0 / sqrt((double)k);
  for (int it = 0; it < 2; ++it) {
    for (int i = 0; i < k; ++i) rhs[i] = z[i];
    // current pivot-row candidate (cols i, i+1, i+2)
    double p0 = a[0] - th, p1 = b[0], p2 = 0.0, pr = rhs[0];
    for (int i = 0; i < k - 1; ++i) {
      const double n0 = b[i], n1 = a[i + 1] - th, n2 = (i + 1 < k - 1) ? b[i + 1] : 0.0;
      const double nr = rhs[i + 1];
      if (fabs(p0) >= fabs(n0)) {
        if (fabs(p0) < tiny) p0 = (p0 < 0 ? -tiny : tiny);
        const double f = n0 / p0;
        u0[i] = p0; u1[i] = p1; u2[i] = p2; rhs[i] = pr;
        p0 = n1 - f * p1; p1 = n2 - f * p2; p2 = 0.0; pr = nr - f * pr;
      } else {
        const double f = p0 / n0;
        u0[i] = n0; u1[i] = n1; u2[i] = n2; rhs[i] = nr;
        const double q0 = p1 - f * n1, q1 = p2 - f * n2;
        p0 = q0; p1 = q1; p2 = 0.0; pr = pr - f * nr;
      }
    }
    if (fabs(p0) < tiny) p0 = (p0 < 0 ? -tiny : tiny);
    u0[k - 1] = p0; u1[k - 1] = 0.0; u2[k - 1] = 0.0; rhs[k - 1] = pr;
    for (int i = k - 1; i >= 0; --i) {
      double s = rhs[i];
      if (i + 1 < k) s -= u1[i] * z[i + 1];
      if (i + 2 < k) s -= u2[i] * z[i + 2];
      z[i] = s / u0[i];
    }
    double nrm = 0.0;
    for (int i = 0; i < k; ++i) nrm += z[i] * z[i];
    nrm = 1.0 / sqrt(nrm);
    for (int i = 0; i < k; ++i) z[i] *= nrm;
  }
  for (int i = 0; i < k; ++i) y[i] = z[i];
}

// dtype of a solve's x or w: the call's, or the solve's own field when the
// call passes TNX_DTYPE_PER_SOLVE
__device__ __forceinline__ int dtype_of(int call_dtype, int32_t own) {
  return call_dtype >= 0 ? call_dtype : own;
}

// ---------------------------------------------------------------- kernels

// Block sum of one value per thread (all NT lanes; result in every lane).
__device__ __forceinline__ double block_total(double v, double (*red)[NCV_MAX + 1]) {
  block_sums(&v, 1, red);
  return (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
}

// act[s], read once per block through LDS: block 0 of a solve may clear it
// while the solve's other blocks run k_update_norm, and every lane of a block
// must take the same early return.
__device__ __forceinline__ bool solve_active(const int32_t* act, int s) {
  __shared__ int a;
  if (threadIdx.x == 0) a = __hip_atomic_load(act + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return a != 0;
}

// chunk partial of ||init||^2 -> partA[block][0]
__global__ __launch_bounds__(NT) void k_init_norm(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                                  void* state, int nchunks) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  const tnx_lanczos_solve D = S[ci.s];
  const double* v = (const double*)D.init;
  double acc = 0.0;
  for (int i = ci.i0 + threadIdx.x; i < ci.i1; i += NT) acc += v[i] * v[i];
  __shared__ double red[4][NCV_MAX + 1];
  const double tot = block_total(acc, red);
  if (threadIdx.x == 0) L.partA[(int64_t)blockIdx.x * (NCV_MAX + 1)] = tot;
}

// scale = 1 / ||init|| (summed by every block); V[0] = init * scale (a flat
// vector when init == 0); x = cast(V[0]).  Block 0 of the solve resets the
// solve's state and marks it active.
__global__ __launch_bounds__(NT) void k_init_scale(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                                   void* state, int nchunks, int x_dtype) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  const tnx_lanczos_solve D = S[ci.s];
  __shared__ double tot;
  reduce_partials(L.partA, D.chunk_begin, ci.nch, 1, &tot);
  __syncthreads();
  const double sc = tot > 0.0 ? 1.0 / sqrt(tot) : -1.0;
  if (ci.c == 0 && threadIdx.x == 0) {
    LzState& st = L.st[ci.s];
    st.scale = sc;
    st.converged = 0;
    st.iters = 0;
    st.jlast = 0;
    st.restarts = 0;
    st.theta = 0.0;
    st.resid = 0.0;
    st.y[0] = 1.0;
    for (int k = 1; k < NCV_MAX; ++k) st.y[k] = 0.0;
    L.act[ci.s] = 1;
  }
  const double* v = (const double*)D.init;
  const double flat = 1.0 / sqrt((double)D.n);
  const int xd = dtype_of(x_dtype, D.x_dtype);
  const bool vf = D.v_dtype == TNX_F32;
  for (int i = ci.i0 + threadIdx.x; i < ci.i1; i += NT) {
    double val = sc > 0.0 ? v[i] * sc : flat;
    if (vf) {
      // the basis row as stored: the start vector rounded to f32
      const float vv = (float)val;
      ((float*)D.V)[i] = vv;
      val = (double)vv;
    } else {
      ((double*)D.V)[i] = val;
    }
    tnx_store((void*)D.x, xd, i, val);
  }
}

// pass 1: chunk partials of V[k] . w, k <= j -> partA
__global__ __launch_bounds__(NT) void k_dots(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                             void* state, int nchunks, int j, int w_dtype,
                                             unsigned long long* ts) {
  tnx_stamp_begin(ts);
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  if (L.act[ci.s] == 0) return;
  const tnx_lanczos_solve D = S[ci.s];
  const int wd = dtype_of(w_dtype, D.w_dtype);
  double w[CHUNK / NT];
#pragma unroll
  for (int q = 0; q < CHUNK / NT; ++q) {
    int i = ci.i0 + q * NT + threadIdx.x;
    w[q] = (i < ci.i1) ? tnx_load((const void*)D.w, wd, i) : 0.0;
  }
  __shared__ double red[4][NCV_MAX + 1];
  if (D.v_dtype == TNX_F32) chunk_dots((const float*)D.V, D.ldv, j, ci.i0, ci.i1, w, red);
  else chunk_dots((const double*)D.V, D.ldv, j, ci.i0, ci.i1, w, red);
  __syncthreads();
  for (int k = threadIdx.x; k <= j; k += NT)
    L.partA[(int64_t)blockIdx.x * (NCV_MAX + 1) + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  tnx_stamp_end(ts);
}

// h1 = sum of the pass-1 partials; r = w - sum h1[k] V[k] -> V[j+1];
// pass 2: chunk partials of V[k] . r (k <= j) and of ||r||^2 (slot j+1) -> partB
__global__ __launch_bounds__(NT) void k_update_dots(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                                    void* state, int nchunks, int j, int w_dtype,
                                                    unsigned long long* ts) {
  tnx_stamp_begin(ts);
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  if (L.act[ci.s] == 0) return;
  const tnx_lanczos_solve D = S[ci.s];
  __shared__ double h[NCV_MAX + 1];
  reduce_partials(L.partA, D.chunk_begin, ci.nch, j + 1, h);
  __syncthreads();
  if (ci.c == 0)
    for (int k = threadIdx.x; k <= j; k += NT) L.st[ci.s].h1[k] = h[k];
  const int wd = dtype_of(w_dtype, D.w_dtype);
  double r[CHUNK / NT];
#pragma unroll
  for (int q = 0; q < CHUNK / NT; ++q) {
    int i = ci.i0 + q * NT + threadIdx.x;
    r[q] = (i < ci.i1) ? tnx_load((const void*)D.w, wd, i) : 0.0;
  }
  __shared__ double red[4][NCV_MAX + 1];
  if (D.v_dtype == TNX_F32) {
    // r (kept in f64 registers for its dots and norm) is parked in row j+1
    // rounded to f32; the next pass normalises the parked row
    const float* V = (const float*)D.V;
    chunk_axpy(V, D.ldv, j, ci.i0, ci.i1, h, r);
    float* Vn = (float*)D.V + (int64_t)(j + 1) * D.ldv;
#pragma unroll
    for (int q = 0; q < CHUNK / NT; ++q) {
      int i = ci.i0 + q * NT + threadIdx.x;
      if (i < ci.i1) Vn[i] = (float)r[q];
    }
    chunk_dots(V, D.ldv, j, ci.i0, ci.i1, r, red);
  } else {
    const double* V = (const double*)D.V;
    chunk_axpy(V, D.ldv, j, ci.i0, ci.i1, h, r);
    double* Vn = (double*)D.V + (int64_t)(j + 1) * D.ldv;
#pragma unroll
    for (int q = 0; q < CHUNK / NT; ++q) {
      int i = ci.i0 + q * NT + threadIdx.x;
      if (i < ci.i1) Vn[i] = r[q];
    }
    chunk_dots(V, D.ldv, j, ci.i0, ci.i1, r, red);
  }
  double nn = 0.0;  // ||r||^2 (r is zero past the chunk end)
#pragma unroll
  for (int q = 0; q < CHUNK / NT; ++q) nn += r[q] * r[q];
  wave_to_lds(nn, j + 1, red);
  __syncthreads();
  for (int k = threadIdx.x; k <= j + 1; k += NT)
    L.partB[(int64_t)blockIdx.x * (NCV_MAX + 1) + k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  tnx_stamp_end(ts);
}

// h2 = sum of the pass-2 partials (slots 0..j), ||r||^2 = slot j+1;
// beta_j = sqrt(||r||^2 - ||h2||^2), the norm of r - V h2 (V^T r = h2 and
// V is orthonormal; ||h2|| is at rounding level next to ||r|| except at a
// Lanczos breakdown, which the test below catches).
// Chunk blocks: V[j+1] = (r - V h2) / beta_j and x = cast(V[j+1]) over
// their chunk, when another Krylov vector follows.  When the second
// projection is below LZ_REORTH_SKIP relative to beta_j (the usual case:
// r is already orthogonal to V to that level after the first pass), the
// correction is left out and V is not read: V[j+1] = r / beta_j, whose
// overlap with every earlier basis vector is then at most LZ_REORTH_SKIP,
// far below the semi-orthogonality (sqrt of the unit roundoff) that keeps
// the Ritz values of the tridiagonal matrix accurate to rounding.  Every
// block of a solve takes the same decision from the same sums.
// Blocks [0, nsolve), one per solve: alpha_j = h1[j] + h2[j] and beta_j (the
// same reduction, so the same bits), the tridiagonal eigenproblem and the
// convergence test (clearing act[s]; the solve's chunk blocks, which read
// act[s] when they start, then skip V[j+1] and x, unused after
// convergence).  They come first in the grid, so the serial eigensolve
// starts with the chunks' pass (blocks [nsolve, nsolve + nchunks)) instead
// of after it.
// The active count after a step, written to the host by the last eigensolve
// block of k_update_norm to finish (every block has made its convergence
// decision by then): each block counts itself in nact[1] after a fence; the
// one that completes the count reads nact, writes it to `host` (pinned
// memory the host reads after the launch's event) and clears the count for
// the next launch.  No block waits for another.
__device__ __forceinline__ void publish_nact(const LzLayout& L, int nsolve, int32_t* host) {
  if (threadIdx.x != 0) return;
  __threadfence();
  if (atomicAdd(L.nact + 1, 1) == nsolve - 1) {
    __threadfence();
    const int32_t v = __hip_atomic_load(L.nact, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (host != nullptr) {
      *(volatile int32_t*)host = v;
      __threadfence_system();
    }
    __hip_atomic_store(L.nact + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__global__ __launch_bounds__(NT) void k_update_norm(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                                    void* state, int nchunks, int j, int x_dtype,
                                                    int normalize, int min_iters,
                                                    int32_t* nact_host, unsigned long long* ts) {
  tnx_stamp_begin(ts);
  LzLayout L = lz_layout(state, nsolve, nchunks);
  const bool solver = (int)blockIdx.x < nsolve;
  ChunkInfo ci;
  int s;
  if (solver) {
    s = (int)blockIdx.x;
  } else {
    if (!normalize) return;
    ci = chunk_of(S, nsolve, (int)blockIdx.x - nsolve);
    s = ci.s;
  }
  if (!solve_active(L.act, s)) {
    if (solver) publish_nact(L, nsolve, nact_host);
    return;
  }
  const tnx_lanczos_solve D = S[s];
  __shared__ double h[NCV_MAX + 1];
  __shared__ double beta_s;
  __shared__ int skip_s;
  reduce_partials(L.partB, D.chunk_begin, (D.n + CHUNK - 1) / CHUNK, j + 2, h);
  __syncthreads();
  if (threadIdx.x == 0) {
    double hh = 0.0;
    for (int k = 0; k <= j; ++k) hh += h[k] * h[k];
    const double bb = fmax(h[j + 1] - hh, 0.0);
    beta_s = sqrt(bb);
    const double sk = D.v_dtype == TNX_F32 ? LZ_REORTH_SKIP_F32 : LZ_REORTH_SKIP;
    skip_s = hh <= sk * sk * bb;
  }
  __syncthreads();
  const double beta_j = beta_s;
  if (!solver) {
    const double inv = 1.0 / beta_j;
    const int xd = dtype_of(x_dtype, D.x_dtype);
    double r[CHUNK / NT];
    if (D.v_dtype == TNX_F32) {
      const float* V = (const float*)D.V;
      float* Vn = (float*)D.V + (int64_t)(j + 1) * D.ldv;
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) {
        int i = ci.i0 + q * NT + threadIdx.x;
        r[q] = (i < ci.i1) ? (double)Vn[i] : 0.0;
      }
      if (!skip_s) chunk_axpy(V, D.ldv, j, ci.i0, ci.i1, h, r);
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) {
        int i = ci.i0 + q * NT + threadIdx.x;
        if (i < ci.i1) {
          // x is the basis row as stored, so the next matvec acts on it
          const float v = (float)(r[q] * inv);
          Vn[i] = v;
          tnx_store((void*)D.x, xd, i, (double)v);
        }
      }
    } else {
      const double* V = (const double*)D.V;
      double* Vn = (double*)D.V + (int64_t)(j + 1) * D.ldv;
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) {
        int i = ci.i0 + q * NT + threadIdx.x;
        r[q] = (i < ci.i1) ? Vn[i] : 0.0;
      }
      if (!skip_s) chunk_axpy(V, D.ldv, j, ci.i0, ci.i1, h, r);
#pragma unroll
      for (int q = 0; q < CHUNK / NT; ++q) {
        int i = ci.i0 + q * NT + threadIdx.x;
        if (i < ci.i1) {
          const double v = r[q] * inv;
          Vn[i] = v;
          tnx_store((void*)D.x, xd, i, v);
        }
      }
    }
    tnx_stamp_end(ts);
    return;
  }
  LzState& st = L.st[s];
  __shared__ double a_s[NCV_MAX], b_s[NCV_MAX], y_s[NCV_MAX], th_s;
  __shared__ double work[5 * NCV_MAX];
  // alpha[0..j-1], beta[0..j-1] were written by earlier launches: every lane
  // loads its share (one round trip instead of a serial loop of j loads)
  for (int k = threadIdx.x; k < j; k += NT) {
    a_s[k] = st.alpha[k];
    b_s[k] = st.beta[k];
  }
  for (int k = threadIdx.x; k <= j; k += NT) st.h2[k] = h[k];
  if (threadIdx.x == 0) {
    const double alpha = st.h1[j] + h[j];
    st.alpha[j] = alpha;
    st.beta[j] = beta_j;
    st.iters += 1;
    a_s[j] = alpha;
    b_s[j] = beta_j;
  }
  __syncthreads();
  tridiag_lowest(a_s, b_s, j + 1, y_s, &th_s, work);
  __syncthreads();
  for (int k = threadIdx.x; k <= j; k += NT) st.y[k] = y_s[k];
  if (threadIdx.x == 0) {
    const double beta = b_s[j];
    const double th = th_s;
    const double res = fabs(beta * y_s[j]);
    st.theta = th;
    st.resid = res;
    st.jlast = j;
    const double thr = fmax(D.tol_scale * fabs(1.0 - D.tau * th), D.floor_rel * fabs(th));
    double anorm = 0.0;
    for (int k = 0; k <= j; ++k) anorm = fmax(anorm, fabs(a_s[k]) + (k > 0 ? fabs(b_s[k - 1]) : 0.0));
    const bool breakdown = beta <= 1e-14 * fmax(anorm, 1e-300);
    const bool full = (j + 1 >= D.n);
    if (full || breakdown || ((j + 1 >= min_iters) && res <= thr)) {
      st.converged = 1;
      __hip_atomic_store(L.act + s, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicSub(L.nact, 1);
    }
  }
  publish_nact(L, nsolve, nact_host);
  tnx_stamp_end(ts);
}

// Sturm count (eigenvalues below x) of the m x m tridiagonal held in
// registers: a[i] diagonal, e[i] = b[i]^2 the squared coupling of i, i+1.
__device__ __forceinline__ int sturm_count_reg(const double (&a)[NCV_MAX],
                                               const double (&e)[NCV_MAX], int m, double x) {
  const double tiny = 1e-300;
  int cnt = 0;
  double q = 1.0;
#pragma unroll
  for (int i = 0; i < NCV_MAX; ++i) {
    if (i < m) {
      q = (i == 0) ? (a[0] - x) : (a[i] - x) - e[i - 1] / q;
      if (q < 0) cnt++;
      if (fabs(q) < tiny) q = -tiny;
    }
  }
  return cnt;
}

// Implicit restart, solve part (one workgroup per solve, active solves
// only), after a full cycle of m = ncv steps (V[m] is the normalised next
// Lanczos vector, beta[m-1] its coefficient):
//  1. the m - k largest Ritz values are the shifts: lane i bisects the
//     Sturm count for eigenvalue k + i (a shift needs no more than ~1e-13
//     relative accuracy: any shift gives a valid compressed factorisation,
//     the exact ones remove the unwanted directions);
//  2. lane 0 applies them as implicit QR steps to T = tridiag(alpha, beta)
//     (a Givens bulge chase per shift, within the leading block when a
//     subdiagonal entry has become negligible) and records the rotations;
//  3. lane t < m accumulates row t of Q (the product of the rotations);
//  4. the compressed factorisation: alpha[0..k), beta[0..k-1) from the
//     rotated T, and the new residual direction
//        r = bhat * V Q[:, k] + beta[m-1] Q[m-1, k-1] V[m],  beta[k-1] = |r|,
//     give the coefficients C of the new basis rows 0..k (k_irl_apply).
// The tridiagonal stays in registers (statically indexed unrolled loops) in
// steps 1 and 2.  A residual that vanishes (an invariant subspace) ends the
// solve as converged with its current Ritz pair.
__global__ __launch_bounds__(NT) void k_irl_solve(const tnx_lanczos_solve* __restrict__ S,
                                                  int nsolve, void* state, int nchunks, int m,
                                                  int k) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  const int s = blockIdx.x;
  if (!solve_active(L.act, s)) return;
  LzState& st = L.st[s];
  __shared__ double a_l[NCV_MAX], b_l[NCV_MAX], sh[NCV_MAX];
  __shared__ double rc[NCV_MAX][NCV_MAX], rs[NCV_MAX][NCV_MAX];
  __shared__ int nrot[NCV_MAX];
  __shared__ double qmk, rho_s, bhat_s;
  __shared__ double ak[LZ_KEEP], bk[LZ_KEEP];
  const int t = threadIdx.x;
  if (t < m) {
    a_l[t] = st.alpha[t];
    b_l[t] = st.beta[t];   // b[m-1]: coefficient of V[m]
  }
  __syncthreads();
  const int np = m - k;
  if (t < np) {
    double a[NCV_MAX], e[NCV_MAX];
    double lo = a_l[0], hi = a_l[0];
#pragma unroll
    for (int i = 0; i < NCV_MAX; ++i) {
      a[i] = (i < m) ? a_l[i] : 0.0;
      const double bi = (i < m - 1) ? b_l[i] : 0.0;
      e[i] = bi * bi;
      if (i < m) {
        const double r = fabs(bi) + (i > 0 ? fabs(b_l[i - 1]) : 0.0);
        lo = fmin(lo, a[i] - r);
        hi = fmax(hi, a[i] + r);
      }
    }
    const int idx = k + t;
    for (int it = 0; it < 64; ++it) {
      if (hi - lo <= 1e-13 * fmax(fabs(lo), fabs(hi)) + 1e-300) break;
      const double mid = 0.5 * (lo + hi);
      if (sturm_count_reg(a, e, m, mid) >= idx + 1) hi = mid;
      else lo = mid;
    }
    sh[t] = 0.5 * (lo + hi);
  }
  __syncthreads();
  if (t == 0) {
    double a[NCV_MAX], b[NCV_MAX];
#pragma unroll
    for (int i = 0; i < NCV_MAX; ++i) {
      a[i] = (i < m) ? a_l[i] : 0.0;
      b[i] = (i < m - 1) ? b_l[i] : 0.0;
    }
    for (int p = 0; p < np; ++p) {
      int iend = m - 1;   // chase rows/cols 0..iend
#pragma unroll
      for (int u = NCV_MAX - 2; u >= 0; --u)
        if (u < m - 1 && fabs(b[u]) <= 2.220446049250313e-16 * (fabs(a[u]) + fabs(a[u + 1])))
          iend = u;   // the first negligible coupling from the top
#pragma unroll
      for (int u = 0; u < NCV_MAX - 1; ++u)
        if (u == iend) b[u] = 0.0;
      nrot[p] = iend;
      double x = a[0] - sh[p], z = b[0], bulge = 0.0;
#pragma unroll
      for (int u = 0; u < NCV_MAX - 1; ++u) {
        if (u < iend) {
          // one reciprocal square root per rotation: the chase is a serial
          // chain, and its entries are O(|T|), far from over- or underflow
          double c = 1.0, sn = 0.0;
          if (z != 0.0) {
            const double ri = rsqrt(x * x + z * z);
            c = x * ri;
            sn = z * ri;
          }
          rc[p][u] = c;
          rs[p][u] = sn;
          if (u > 0) b[u - 1] = c * b[u - 1] + sn * bulge;
          const double au = a[u], au1 = a[u + 1], bu = b[u];
          a[u] = c * c * au + 2.0 * c * sn * bu + sn * sn * au1;
          a[u + 1] = sn * sn * au - 2.0 * c * sn * bu + c * c * au1;
          b[u] = c * sn * (au1 - au) + (c * c - sn * sn) * bu;
          if (u + 1 < iend) {
            bulge = sn * b[u + 1];
            b[u + 1] = c * b[u + 1];
            x = b[u];
            z = bulge;
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < LZ_KEEP; ++i) {
      ak[i] = a[i];
      bk[i] = b[i];
    }
  }
  __syncthreads();
  double q[NCV_MAX];   // row t of Q (static indices: kept in registers)
#pragma unroll
  for (int c = 0; c < NCV_MAX; ++c) q[c] = (c == t) ? 1.0 : 0.0;
  if (t < m) {
    for (int p = 0; p < np; ++p) {
      const int n = nrot[p];
#pragma unroll
      for (int u = 0; u < NCV_MAX - 1; ++u) {
        if (u < n) {
          const double c = rc[p][u], sn = rs[p][u];
          const double q0 = q[u], q1 = q[u + 1];
          q[u] = c * q0 + sn * q1;
          q[u + 1] = -sn * q0 + c * q1;
        }
      }
    }
    if (t == m - 1) {
#pragma unroll
      for (int c = 0; c < NCV_MAX; ++c)
        if (c == k - 1) qmk = q[c];
    }
  }
  __syncthreads();
  if (t == 0) {
    const double bhat = bk[k - 1];
    const double sigma = st.beta[m - 1] * qmk;
    const double rho = hypot(bhat, sigma);
    bhat_s = bhat;
    rho_s = rho;
    double anorm = 0.0;
    for (int u = 0; u < m; ++u) anorm = fmax(anorm, fabs(a_l[u]) + fabs(b_l[u]));
    for (int u = 0; u < k; ++u) st.alpha[u] = ak[u];
    for (int u = 0; u + 1 < k; ++u) st.beta[u] = bk[u];
    st.beta[k - 1] = rho;
    st.restarts += 1;
    for (int u = 0; u <= k; ++u) st.C[m][u] = 0.0;
    if (rho > 0.0) st.C[m][k] = sigma / rho;
    if (!(rho > 1e-14 * fmax(anorm, 1e-300))) {
      // invariant subspace: the current Ritz pair is exact
      st.converged = 1;
      __hip_atomic_store(L.act + s, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicSub(L.nact, 1);
    }
  }
  __syncthreads();
  if (t < m) {
    const double f = rho_s > 0.0 ? bhat_s / rho_s : 0.0;
#pragma unroll
    for (int c = 0; c < NCV_MAX; ++c) {
      if (c < k) st.C[t][c] = q[c];
      else if (c == k) st.C[t][c] = f * q[c];
    }
  }
}

// Rows V'[u] = sum_t C[t][u] V[t] (u <= k, t <= m) of this chunk in place,
// x = cast(V'[k]) as stored.
template <typename VT>
__device__ __forceinline__ void irl_apply_rows(VT* V, const tnx_lanczos_solve& D,
                                               const double (*C)[LZ_KEEP + 1],
                                               const ChunkInfo& ci, int m, int k, int xd) {
  for (int i = ci.i0 + threadIdx.x; i < ci.i1; i += NT) {
    double acc[LZ_KEEP + 1];
#pragma unroll
    for (int u = 0; u <= LZ_KEEP; ++u) acc[u] = 0.0;
    for (int t = 0; t <= m; ++t) {
      const double v = (double)V[(int64_t)t * D.ldv + i];
#pragma unroll
      for (int u = 0; u <= LZ_KEEP; ++u)
        if (u <= k) acc[u] += C[t][u] * v;
    }
    double xk = 0.0;
#pragma unroll
    for (int u = 0; u <= LZ_KEEP; ++u)
      if (u <= k) {
        const VT st = (VT)acc[u];
        V[(int64_t)u * D.ldv + i] = st;
        if (u == k) xk = (double)st;
      }
    tnx_store((void*)D.x, xd, i, xk);
  }
}

// Implicit restart, basis part (chunk workgroups of the active solves): the
// new rows V'[u] = sum_t C[t][u] V[t] (u <= k, t <= m) in place (every
// element is read in full before it is written), x = cast(V'[k]).
__global__ __launch_bounds__(NT) void k_irl_apply(const tnx_lanczos_solve* __restrict__ S,
                                                  int nsolve, void* state, int nchunks, int m,
                                                  int k, int x_dtype) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  if (!solve_active(L.act, ci.s)) return;
  const tnx_lanczos_solve D = S[ci.s];
  const LzState& st = L.st[ci.s];
  __shared__ double C[NCV_MAX + 1][LZ_KEEP + 1];
  for (int e = threadIdx.x; e < (m + 1) * (LZ_KEEP + 1); e += NT)
    C[e / (LZ_KEEP + 1)][e % (LZ_KEEP + 1)] = st.C[e / (LZ_KEEP + 1)][e % (LZ_KEEP + 1)];
  __syncthreads();
  const int xd = dtype_of(x_dtype, D.x_dtype);
  if (D.v_dtype == TNX_F32) irl_apply_rows((float*)D.V, D, C, ci, m, k, xd);
  else irl_apply_rows((double*)D.V, D, C, ci, m, k, xd);
}

// out = sum_k y[k] V[k] for every solve; the first chunk's block of each
// solve also writes its result record to `res`
__global__ __launch_bounds__(NT) void k_final(const tnx_lanczos_solve* __restrict__ S, int nsolve,
                                              void* state, int nchunks,
                                              tnx_lanczos_result* res) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  ChunkInfo ci = chunk_of(S, nsolve);
  const tnx_lanczos_solve D = S[ci.s];
  const LzState& st = L.st[ci.s];
  if (ci.c == 0 && threadIdx.x == 0) {
    res[ci.s].theta = st.theta;
    res[ci.s].resid = st.resid;
    res[ci.s].iters = st.iters;
    res[ci.s].converged = st.converged;
  }
  const int kk = st.jlast + 1;
  double* out = (double*)D.out;
  if (D.v_dtype == TNX_F32) {
    const float* V = (const float*)D.V;
    for (int i = ci.i0 + threadIdx.x; i < ci.i1; i += NT) {
      double s = 0.0;
      for (int k = 0; k < kk; ++k) s += st.y[k] * (double)V[(int64_t)k * D.ldv + i];
      out[i] = s;
    }
  } else {
    const double* V = (const double*)D.V;
    for (int i = ci.i0 + threadIdx.x; i < ci.i1; i += NT) {
      double s = 0.0;
      for (int k = 0; k < kk; ++k) s += st.y[k] * V[(int64_t)k * D.ldv + i];
      out[i] = s;
    }
  }
}


// pinned staging for the polled active count and the results
struct Pinned {
  int32_t* nact = nullptr;
  tnx_lanczos_result* res = nullptr;
  int res_cap = 0;
};
thread_local Pinned g_pin;

int ensure_pinned(int nsolve) {
  if (g_pin.nact == nullptr) {
    TNX_CHECK_HIP(hipHostMalloc((void**)&g_pin.nact, 64 * sizeof(int32_t), hipHostMallocDefault));
  }
  if (g_pin.res_cap < nsolve) {
    if (g_pin.res) (void)hipHostFree(g_pin.res);
    g_pin.res = nullptr;
    int cap = nsolve < 256 ? 256 : nsolve;
    TNX_CHECK_HIP(hipHostMalloc((void**)&g_pin.res, sizeof(tnx_lanczos_result) * cap, hipHostMallocDefault));
    g_pin.res_cap = cap;
  }
  return TNX_OK;
}

}  // namespace

extern "C" size_t tnx_lanczos_state_bytes(int nsolve, int nchunks, int ncv) {
  (void)ncv;
  return align256(sizeof(LzState) * (size_t)nsolve) + align256(sizeof(int32_t) * (size_t)nsolve) +
         2 * align256(sizeof(double) * (size_t)nchunks * (NCV_MAX + 1)) + 256;
}

namespace {

// Zeroes n 8-byte words (the solves' state at the start of a call; a kernel
// rather than a memset node, so the launch sequence is plain kernels) and
// sets the active count.
__global__ void k_zero_words(uint64_t* p, size_t n, int32_t* nact, int nsolve) {
  // every solve active; no block of a k_update_norm launch counted done yet
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    nact[0] = nsolve;
    nact[1] = 0;
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

int check_args(const char* who, const tnx_matvec_plan* plans, int nplan,
               const tnx_lanczos_solve* solves, int nsolve, int nchunks, int ncv,
               int max_restarts, int x_dtype, int w_dtype) {
  if (plans == nullptr || nplan <= 0 || solves == nullptr || nsolve <= 0 || nchunks <= 0 ||
      ncv < 2 || ncv > NCV_MAX || max_restarts < 0 || x_dtype < -1 || x_dtype > TNX_BF16 ||
      w_dtype < -1 || w_dtype > TNX_BF16) {
    tnx_set_error("%s: bad arguments (nplan=%d nsolve=%d nchunks=%d ncv=%d)", who, nplan, nsolve,
                  nchunks, ncv);
    return TNX_EINVAL;
  }
  return TNX_OK;
}

// The launches of a call before its first Krylov step: state zeroed, every
// solve active, start vectors normalised into V[0] and x.
int issue_start(const tnx_lanczos_solve* solves, int nsolve, void* state, int nchunks,
                int x_dtype, hipStream_t s) {
  LzLayout L = lz_layout(state, nsolve, nchunks);
  const size_t words = sizeof(LzState) * (size_t)nsolve / sizeof(uint64_t);
  static_assert(sizeof(LzState) % sizeof(uint64_t) == 0, "LzState is whole words");
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((words + 255) / 256 < 1024 ? (words + 255) / 256 : 1024)),
                     dim3(256), 0, s, (uint64_t*)L.st, words, L.nact, nsolve);
  hipLaunchKernelGGL(k_init_norm, dim3(nchunks), dim3(NT), 0, s, solves, nsolve, state, nchunks);
  hipLaunchKernelGGL(k_init_scale, dim3(nchunks), dim3(NT), 0, s, solves, nsolve, state, nchunks,
                     x_dtype);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

// Krylov step j of every active solve: each plan's matvec, then the dots,
// the update with its dots and norm, and the normalisation with the
// tridiagonal solve and the convergence test.
int issue_step(const tnx_matvec_plan* plans, int nplan, const tnx_lanczos_solve* solves,
               int nsolve, void* state, int nchunks, int j, int ncv, int min_iters, int x_dtype,
               int w_dtype, tnx_stream_t stream, int32_t* nact_host = nullptr) {
  hipStream_t s = (hipStream_t)stream;
  LzLayout L = lz_layout(state, nsolve, nchunks);
  for (int p = 0; p < nplan; ++p)
    if (int rc = tnx_heff_matvec(plans + p, L.act, stream)) return rc;
  const dim3 grid(nchunks), block(NT);
  // sampled timing (one plan: its modes give the x and w element sizes, and
  // the basis element size under the policy that a low-precision matvec
  // has an f32 basis); algorithmic bytes over the nchunks * CHUNK
  // elements: basis rows 0..j read once, w / r / x read or written once
  const bool prof = tnx_prof_on() && nplan == 1;
  const double E = (double)nchunks * CHUNK;
  const int m1 = plans[0].mode & 0xff, m2 = plans[0].mode2 & 0xff;
  const double xs = m1 == TNX_GEMM_F64 ? 8.0 : m1 == TNX_GEMM_F32 ? 4.0 : 2.0;
  const double ws = m2 == TNX_GEMM_F64 ? 8.0 : m2 == TNX_GEMM_BF16_BF16 ? 2.0 : 4.0;
  const double vs = m1 == TNX_GEMM_F64 ? 8.0 : 4.0;
  const double vrows = vs * (j + 1);
  tnx_prof_count(TNX_PC_DOTS, E * (vrows + ws));
  tnx_prof_count(TNX_PC_UPDATE_DOTS, E * (vrows + ws + vs));
  tnx_prof_count(TNX_PC_UPDATE_NORM, E * (2.0 * vs + xs));
  int mk = prof ? tnx_prof_mark(TNX_PC_DOTS, s) : -1;
  hipLaunchKernelGGL(k_dots, grid, block, 0, s, solves, nsolve, state, nchunks, j, w_dtype,
                     tnx_prof_ts(mk));
  tnx_prof_done(mk, s, E * (vrows + ws));
  mk = prof ? tnx_prof_mark(TNX_PC_UPDATE_DOTS, s) : -1;
  hipLaunchKernelGGL(k_update_dots, grid, block, 0, s, solves, nsolve, state, nchunks, j, w_dtype,
                     tnx_prof_ts(mk));
  tnx_prof_done(mk, s, E * (vrows + ws + vs));
  // the chunk blocks normalise V[j+1] at every step, the last of a cycle
  // included (an implicit restart combines V[ncv] into the new basis)
  (void)ncv;
  mk = prof ? tnx_prof_mark(TNX_PC_UPDATE_NORM, s) : -1;
  hipLaunchKernelGGL(k_update_norm, dim3(nchunks + nsolve), block, 0, s, solves, nsolve, state,
                     nchunks, j, x_dtype, 1, min_iters, nact_host, tnx_prof_ts(mk));
  // (r read, V[j+1] and x written; the basis rows of the second correction,
  // read only when it is not negligible, are not counted)
  tnx_prof_done(mk, s, E * (2.0 * vs + xs));
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

// Implicit restart of the still-active solves after a full cycle: the next
// cycle continues from Krylov step lz_keep(ncv).
int issue_restart(const tnx_lanczos_solve* solves, int nsolve, void* state, int nchunks, int ncv,
                  int x_dtype, hipStream_t s) {
  const int k = lz_keep(ncv);
  hipLaunchKernelGGL(k_irl_solve, dim3(nsolve), dim3(NT), 0, s, solves, nsolve, state, nchunks,
                     ncv, k);
  hipLaunchKernelGGL(k_irl_apply, dim3(nchunks), dim3(NT), 0, s, solves, nsolve, state, nchunks,
                     ncv, k, x_dtype);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

// Eigenvectors into every solve's `out`, results into `res` (device-visible).
int issue_finish(const tnx_lanczos_solve* solves, int nsolve, void* state, int nchunks,
                 tnx_lanczos_result* res, hipStream_t s) {
  hipLaunchKernelGGL(k_final, dim3(nchunks), dim3(NT), 0, s, solves, nsolve, state, nchunks, res);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

// The Lanczos loop of the entry points below: each Krylov step applies the
// nplan matvec plans in turn (one per precision of the batch), then the
// shared orthogonalisation kernels over all solves.  The host polls the
// active count one step behind the device, so one step stays queued.
// With `next` (nnext recorded launches), those launches are queued behind
// the result copies and the call returns once the results are on the host,
// while the queued launches may still run; *next_rc receives their
// tnx_run_ops code.
int lanczos_run(const tnx_matvec_plan* plans, int nplan, const tnx_lanczos_solve* solves,
                int nsolve, int nchunks, int ncv, int max_restarts, int min_iters, int x_dtype,
                int w_dtype, void* state, tnx_lanczos_result* results, tnx_stream_t stream,
                const tnx_op* next = nullptr, int nnext = 0, int32_t* next_rc = nullptr) {
  if (int rc = check_args("tnx_lanczos_run", plans, nplan, solves, nsolve, nchunks, ncv,
                          max_restarts, x_dtype, w_dtype))
    return rc;
  if (state == nullptr) {
    tnx_set_error("tnx_lanczos_run: null state");
    return TNX_EINVAL;
  }
  if (int rc = ensure_pinned(nsolve)) return rc;
  hipStream_t s = (hipStream_t)stream;
  LzLayout L = lz_layout(state, nsolve, nchunks);
  if (int rc = issue_start(solves, nsolve, state, nchunks, x_dtype, s)) return rc;

  hipEvent_t ev[2];
  TNX_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  TNX_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  volatile int32_t* pin = g_pin.nact;
  pin[0] = nsolve;
  pin[1] = nsolve;
  bool done = false;
  int rc = TNX_OK;
  int issued = 0;  // events recorded so far
  const bool sampled = tnx_prof_call_begin();
  for (int r = 0; r <= max_restarts && !done && !rc; ++r) {
    if (r > 0 && (rc = issue_restart(solves, nsolve, state, nchunks, ncv, x_dtype, s))) break;
    for (int j = (r > 0 ? lz_keep(ncv) : 0); j < ncv && !done; ++j) {
      tnx_prof_set(sampled && r == 0 && j == 0);
      // lagged poll: the step writes its active count to pin[slot], then
      // wait for the previous step (the device keeps one step queued)
      const int slot = issued & 1;
      rc = issue_step(plans, nplan, solves, nsolve, state, nchunks, j, ncv, min_iters, x_dtype,
                      w_dtype, stream, (int32_t*)(pin + slot));
      tnx_prof_set(false);
      if (rc) break;
      TNX_CHECK_HIP(hipEventRecord(ev[slot], s));
      ++issued;
      if (issued >= 2) {
        const int prev = (issued - 2) & 1;
        TNX_CHECK_HIP(hipEventSynchronize(ev[prev]));
        if (pin[prev] <= 0) done = true;
      }
    }
    if (rc) break;
    if (!done) {
      // end of a Krylov cycle: make the restart decision on fresh data
      const int slot = (issued - 1) & 1;
      TNX_CHECK_HIP(hipEventSynchronize(ev[slot]));
      if (pin[slot] <= 0) done = true;
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  if ((rc = issue_finish(solves, nsolve, state, nchunks, g_pin.res, s))) return rc;
  int orc = TNX_OK;
  if (nnext > 0) {
    hipEvent_t fin;
    TNX_CHECK_HIP(hipEventCreateWithFlags(&fin, hipEventDisableTiming));
    TNX_CHECK_HIP(hipEventRecord(fin, s));
    orc = tnx_run_ops(next, nnext, stream);   // sets the message of a failure
    if (next_rc) *next_rc = orc;
    const hipError_t e = hipEventSynchronize(fin);
    (void)hipEventDestroy(fin);
    if (e != hipSuccess) {
      tnx_set_error("tnx_lanczos_run_multi_then: %s", hipGetErrorString(e));
      return TNX_EHIP;
    }
  } else {
    TNX_CHECK_HIP(hipStreamSynchronize(s));
  }
  int nconv = 0;
  for (int i = 0; i < nsolve; ++i) {
    if (results) results[i] = g_pin.res[i];
    nconv += g_pin.res[i].converged;
  }
  if (orc != TNX_OK && orc != TNX_ENOCONV) return orc;
  return nconv == nsolve ? TNX_OK : TNX_ENOCONV;
}

}  // namespace

extern "C" int tnx_lanczos_run_ex(const tnx_matvec_plan* plan, const tnx_lanczos_solve* solves,
                                  int nsolve, int nchunks, int ncv, int max_restarts,
                                  int min_iters, int x_dtype, int w_dtype, void* state,
                                  tnx_lanczos_result* results, tnx_stream_t stream) {
  return lanczos_run(plan, 1, solves, nsolve, nchunks, ncv, max_restarts, min_iters, x_dtype,
                     w_dtype, state, results, stream);
}

extern "C" int tnx_lanczos_run_multi(const tnx_matvec_plan* plans, int nplan,
                                     const tnx_lanczos_solve* solves, int nsolve, int nchunks,
                                     int ncv, int max_restarts, int min_iters, void* state,
                                     tnx_lanczos_result* results, tnx_stream_t stream) {
  return lanczos_run(plans, nplan, solves, nsolve, nchunks, ncv, max_restarts, min_iters,
                     TNX_DTYPE_PER_SOLVE, TNX_DTYPE_PER_SOLVE, state, results, stream);
}

extern "C" int tnx_lanczos_run_multi_then(const tnx_matvec_plan* plans, int nplan,
                                          const tnx_lanczos_solve* solves, int nsolve,
                                          int nchunks, int ncv, int max_restarts, int min_iters,
                                          void* state, tnx_lanczos_result* results,
                                          const tnx_op* next, int nnext, int32_t* next_rc,
                                          tnx_stream_t stream) {
  if (next_rc) *next_rc = TNX_OK;
  if (nnext < 0 || (nnext > 0 && next == nullptr)) {
    tnx_set_error("tnx_lanczos_run_multi_then: bad launch list (nnext=%d)", nnext);
    return TNX_EINVAL;
  }
  return lanczos_run(plans, nplan, solves, nsolve, nchunks, ncv, max_restarts, min_iters,
                     TNX_DTYPE_PER_SOLVE, TNX_DTYPE_PER_SOLVE, state, results, stream, next,
                     nnext, next_rc);
}

extern "C" int tnx_lanczos_run(const tnx_matvec_plan* plan, const tnx_lanczos_solve* solves,
                               int nsolve, int nchunks, int ncv, int max_restarts, int x_dtype,
                               int w_dtype, void* state, tnx_lanczos_result* results,
                               tnx_stream_t stream) {
  return tnx_lanczos_run_ex(plan, solves, nsolve, nchunks, ncv, max_restarts, 1, x_dtype, w_dtype,
                            state, results, stream);
}

// ------------------------------------------------- no-allocation entry point
//
// tnx_lanczos_run_multi_static: the same Lanczos loop with its state in a
// library-owned device array instead of a caller workspace, and no host
// staging memory: it allocates nothing on any call.  Each GPU the library
// runs on holds one copy of `scratch` until the process ends; every call on
// that GPU shares it, so calls that overlap on one GPU race.  The state is
// re-initialised by every call (nothing carries over from the last call).
namespace {

constexpr size_t LZ_STATIC_BYTES = TNX_LANCZOS_STATIC_BYTES;
static __device__ __align__(256) unsigned char scratch[LZ_STATIC_BYTES];

constexpr int MAX_DEVICES = 64;
void* g_scratch_ptr[MAX_DEVICES];   // device address of `scratch` per GPU
std::mutex g_scratch_mu;

// Device address of the current GPU's `scratch` (looked up once per GPU,
// with that GPU current).
int scratch_of_current_device(void** out) {
  int dev = -1;
  TNX_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) {
    tnx_set_error("tnx_lanczos_run_multi_static: device %d out of range", dev);
    return TNX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    TNX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TNX_OK;
}

}  // namespace

extern "C" int tnx_lanczos_run_multi_static(const tnx_matvec_plan* plans, int nplan,
                                            const tnx_lanczos_solve* solves, int nsolve,
                                            int nchunks, int ncv, int max_restarts, int min_iters,
                                            tnx_lanczos_result* results, tnx_stream_t stream) {
  const char* who = "tnx_lanczos_run_multi_static";
  if (int rc = check_args(who, plans, nplan, solves, nsolve, nchunks, ncv, max_restarts,
                          TNX_DTYPE_PER_SOLVE, TNX_DTYPE_PER_SOLVE))
    return rc;
  const size_t state_bytes = tnx_lanczos_state_bytes(nsolve, nchunks, ncv);
  const size_t need = state_bytes + align256(sizeof(tnx_lanczos_result) * (size_t)nsolve);
  if (need > LZ_STATIC_BYTES) {
    tnx_set_error("%s: %d solves over %d chunks need %zu bytes of state, more than the %zu of "
                  "the library's scratch (use tnx_lanczos_run_multi with a workspace)",
                  who, nsolve, nchunks, need, LZ_STATIC_BYTES);
    return TNX_EINVAL;
  }
  void* base = nullptr;
  if (int rc = scratch_of_current_device(&base)) return rc;
  void* state = base;
  tnx_lanczos_result* dres = (tnx_lanczos_result*)((char*)base + state_bytes);
  hipStream_t s = (hipStream_t)stream;
  LzLayout L = lz_layout(state, nsolve, nchunks);
  if (int rc = issue_start(solves, nsolve, state, nchunks, TNX_DTYPE_PER_SOLVE, s)) return rc;
  // Every Krylov step of a cycle is queued without polling: a solve that has
  // converged is masked on the device, so its arithmetic, and the result,
  // are those of tnx_lanczos_run_multi.  The active count is read at the end
  // of each cycle (the restart decision).
  bool done = false;
  const bool sampled = tnx_prof_call_begin();
  for (int r = 0; r <= max_restarts && !done; ++r) {
    if (r > 0)
      if (int rc = issue_restart(solves, nsolve, state, nchunks, ncv, TNX_DTYPE_PER_SOLVE, s))
        return rc;
    for (int j = (r > 0 ? lz_keep(ncv) : 0); j < ncv; ++j) {
      tnx_prof_set(sampled && r == 0 && j == 0);
      const int rc = issue_step(plans, nplan, solves, nsolve, state, nchunks, j, ncv, min_iters,
                                TNX_DTYPE_PER_SOLVE, TNX_DTYPE_PER_SOLVE, stream);
      tnx_prof_set(false);
      if (rc) return rc;
    }
    int32_t nact = 0;
    TNX_CHECK_HIP(hipMemcpyAsync(&nact, L.nact, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    TNX_CHECK_HIP(hipStreamSynchronize(s));
    done = nact <= 0;
  }
  if (int rc = issue_finish(solves, nsolve, state, nchunks, dres, s)) return rc;
  tnx_lanczos_result* host = results;
  int nconv = 0;
  for (int i0 = 0; i0 < nsolve; i0 += 64) {   // results through a bounded stack buffer
    tnx_lanczos_result buf[64];
    const int k = nsolve - i0 < 64 ? nsolve - i0 : 64;
    TNX_CHECK_HIP(hipMemcpyAsync(buf, dres + i0, sizeof(tnx_lanczos_result) * (size_t)k,
                                 hipMemcpyDeviceToHost, s));
    TNX_CHECK_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < k; ++i) {
      if (host) host[i0 + i] = buf[i];
      nconv += buf[i].converged;
    }
  }
  return nconv == nsolve ? TNX_OK : TNX_ENOCONV;
}
