// f64 orthogonalisations of MPS tensors: Householder QR (gauge moves) and
// one-sided Jacobi (Hestenes) SVD (entanglement spectrum moves).
//
// One workgroup (256 lanes) per matrix; the working copy lives column-major
// in the job's work buffer so every reflector / rotation touches contiguous
// columns.  Columns are distributed over the 4 wavefronts, rows over the 64
// lanes of a wavefront, and column dot products are wavefront reductions.
#include "tnx_common.h"

namespace {

constexpr int NT = 256;
constexpr int KMAX = 4096;  // max columns of a gauge-move matrix

__device__ __forceinline__ double src_at(const double* src, int ld, int trans, int r, int c) {
  return trans ? src[(int64_t)c * ld + r] : src[(int64_t)r * ld + c];
}

// column-major m x n work matrix W (ld = m); Q scratch m x k after it.
__global__ __launch_bounds__(NT) void qr_kernel(const tnx_qr_job* __restrict__ jobs) {
  const tnx_qr_job J = jobs[blockIdx.x];
  const int m = J.m, n = J.n, k = m < n ? m : n;
  if (k > KMAX || m <= 0 || n <= 0) return;  // rejected on the host
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* W = (double*)J.work;
  double* Q = W + (int64_t)m * n;
  const double* src = (const double*)J.src;
  __shared__ double tau[KMAX];
  __shared__ double bcast[2];

  for (int64_t e = threadIdx.x; e < (int64_t)m * n; e += NT) {
    int r, c;
    if (J.trans) { c = (int)(e / m); r = (int)(e - (int64_t)c * m); }   // src[c*ld+r] contiguous in r
    else { r = (int)(e / n); c = (int)(e - (int64_t)r * n); }           // src[r*ld+c] contiguous in c
    W[(int64_t)c * m + r] = src_at(src, J.ld, J.trans, r, c);
  }
  __syncthreads();

  for (int j = 0; j < k; ++j) {
    double* wj = W + (int64_t)j * m;
    if (wave == 0) {
      double s = 0.0;
      for (int r = j + 1 + lane; r < m; r += 64) s += wj[r] * wj[r];
      s = tnx_wave_sum(s);
      const double x0 = wj[j];
      double beta, t, scal;
      if (s == 0.0) {
        beta = x0; t = 0.0; scal = 0.0;
      } else {
        double nrm = sqrt(x0 * x0 + s);
        beta = x0 >= 0.0 ? -nrm : nrm;
        t = (beta - x0) / beta;
        scal = 1.0 / (x0 - beta);
      }
      for (int r = j + 1 + lane; r < m; r += 64) wj[r] *= scal;
      if (lane == 0) {
        wj[j] = beta;
        tau[j] = t;
      }
    }
    __syncthreads();
    const double tj = tau[j];
    if (tj != 0.0) {
      for (int c = j + 1 + wave; c < n; c += 4) {
        double* wc = W + (int64_t)c * m;
        double s = (lane == 0) ? wc[j] : 0.0;
        for (int r = j + 1 + lane; r < m; r += 64) s += wj[r] * wc[r];
        s = tnx_wave_sum(s) * tj;
        if (lane == 0) wc[j] -= s;
        for (int r = j + 1 + lane; r < m; r += 64) wc[r] -= s * wj[r];
      }
    }
    __syncthreads();
  }

  // R (k x n, row-major): upper triangle of W
  double* R = (double*)J.r;
  for (int64_t e = threadIdx.x; e < (int64_t)k * n; e += NT) {
    int i = (int)(e / n), c = (int)(e - (int64_t)i * n);
    R[e] = (i <= c) ? W[(int64_t)c * m + i] : 0.0;
  }
  // Q = H_0 ... H_{k-1} [I_k; 0], column-major m x k, backward accumulation
  for (int64_t e = threadIdx.x; e < (int64_t)m * k; e += NT) {
    int c = (int)(e / m), r = (int)(e - (int64_t)c * m);
    Q[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj != 0.0) {
      const double* v = W + (int64_t)j * m;  // v[j] = 1 implicit, v[r>j] stored
      for (int c = j + wave; c < k; c += 4) {
        double* qc = Q + (int64_t)c * m;
        double s = (lane == 0) ? qc[j] : 0.0;
        for (int r = j + 1 + lane; r < m; r += 64) s += v[r] * qc[r];
        s = tnx_wave_sum(s) * tj;
        if (lane == 0) qc[j] -= s;
        for (int r = j + 1 + lane; r < m; r += 64) qc[r] -= s * v[r];
      }
    }
    __syncthreads();
  }
  double* q = (double*)J.q;
  if (J.q_trans) {
    for (int64_t e = threadIdx.x; e < (int64_t)m * k; e += NT) q[e] = Q[e];
  } else {
    for (int64_t e = threadIdx.x; e < (int64_t)m * k; e += NT) {
      int r = (int)(e / k), c = (int)(e - (int64_t)r * k);
      q[e] = Q[(int64_t)c * m + r];
    }
  }
  (void)bcast;
}

// One-sided Jacobi SVD; work: W (m x n col-major), V (n x n col-major).
__global__ __launch_bounds__(NT) void svd_kernel(const tnx_svd_job* __restrict__ jobs, int max_sweeps) {
  const tnx_svd_job J = jobs[blockIdx.x];
  const int m = J.m, n = J.n;
  if (n > KMAX || m < n || n <= 0) return;  // rejected on the host
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double* W = (double*)J.work;
  double* V = W + (int64_t)m * n;
  const double* src = (const double*)J.src;
  __shared__ int rotated;
  __shared__ double sv[KMAX];
  __shared__ int order[KMAX];

  for (int64_t e = threadIdx.x; e < (int64_t)m * n; e += NT) {
    int r, c;
    if (J.trans) { c = (int)(e / m); r = (int)(e - (int64_t)c * m); }
    else { r = (int)(e / n); c = (int)(e - (int64_t)r * n); }
    W[(int64_t)c * m + r] = src_at(src, J.ld, J.trans, r, c);
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    V[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();

  const int np = (n + 1) & ~1;  // players in the round-robin tournament
  // rotate while |a.b| > m*eps*|a||b| (columns orthogonal to working accuracy)
  const double eps = 2.220446049250313e-16 * (double)(m > 8 ? m : 8);
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      for (int p = wave; p < np / 2; p += 4) {
        // circle method: player np-1 fixed, others rotate
        int a = (p == 0) ? np - 1 : (round + p) % (np - 1);
        int b = (round + np - 1 - p) % (np - 1);
        if (p == 0) b = round % (np - 1);
        if (a >= n || b >= n || a == b) continue;
        if (a > b) { int t = a; a = b; b = t; }
        double* wa = W + (int64_t)a * m;
        double* wb = W + (int64_t)b * m;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int r = lane; r < m; r += 64) {
          const double x = wa[r], y = wb[r];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        al = tnx_wave_sum(al);
        be = tnx_wave_sum(be);
        ga = tnx_wave_sum(ga);
        if (ga != 0.0 && fabs(ga) > eps * sqrt(al * be)) {
          if (lane == 0) rotated = 1;
          const double zeta = (be - al) / (2.0 * ga);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
          for (int r = lane; r < m; r += 64) {
            const double x = wa[r], y = wb[r];
            wa[r] = cs * x - sn * y;
            wb[r] = sn * x + cs * y;
          }
          double* va = V + (int64_t)a * n;
          double* vb = V + (int64_t)b * n;
          for (int r = lane; r < n; r += 64) {
            const double x = va[r], y = vb[r];
            va[r] = cs * x - sn * y;
            vb[r] = sn * x + cs * y;
          }
        }
      }
      __syncthreads();
    }
    if (rotated == 0) break;
    __syncthreads();
  }
  // singular values and descending order
  for (int c = wave; c < n; c += 4) {
    const double* wc = W + (int64_t)c * m;
    double s = 0.0;
    for (int r = lane; r < m; r += 64) s += wc[r] * wc[r];
    s = tnx_wave_sum(s);
    if (lane == 0) sv[c] = sqrt(s);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += NT) {
    int rank = 0;
    const double x = sv[c];
    for (int o = 0; o < n; ++o) rank += (sv[o] > x) || (sv[o] == x && o < c);
    order[rank] = c;
  }
  __syncthreads();
  double* s_out = (double*)J.s;
  double* u = (double*)J.u;
  double* svt = (double*)J.svt;
  for (int i = threadIdx.x; i < n; i += NT) s_out[i] = sv[order[i]];
  // U columns (normalised), sorted; zero columns are completed afterwards
  for (int i = wave; i < n; i += 4) {
    const int c = order[i];
    const double sc = sv[c] > 0.0 ? 1.0 / sv[c] : 0.0;
    const double* wc = W + (int64_t)c * m;
    for (int r = lane; r < m; r += 64) {
      const double v = wc[r] * sc;
      if (J.u_trans) u[(int64_t)i * m + r] = v;
      else u[(int64_t)r * n + i] = v;
    }
  }
  // svt[i][col] = s_i * V[col][order i]
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    int i = (int)(e / n), col = (int)(e - (int64_t)i * n);
    const int c = order[i];
    svt[e] = sv[c] * V[(int64_t)c * n + col];
  }
  __syncthreads();
  // complete null columns of U (singular value exactly 0) by Gram-Schmidt
  if (wave == 0) {
    int next_unit = 0;
    for (int i = 0; i < n; ++i) {
      if (sv[order[i]] > 0.0) continue;
      for (; next_unit < m; ++next_unit) {
        // candidate e_{next_unit}, orthogonalised twice against columns != i
        for (int r = lane; r < m; r += 64) {
          const double v = (r == next_unit) ? 1.0 : 0.0;
          if (J.u_trans) u[(int64_t)i * m + r] = v; else u[(int64_t)r * n + i] = v;
        }
        for (int pass = 0; pass < 2; ++pass) {
          for (int o = 0; o < n; ++o) {
            if (o == i || (sv[order[o]] == 0.0 && o > i)) continue;
            double d = 0.0;
            for (int r = lane; r < m; r += 64) {
              const double uo = J.u_trans ? u[(int64_t)o * m + r] : u[(int64_t)r * n + o];
              const double ui = J.u_trans ? u[(int64_t)i * m + r] : u[(int64_t)r * n + i];
              d += uo * ui;
            }
            d = tnx_wave_sum(d);
            for (int r = lane; r < m; r += 64) {
              const double uo = J.u_trans ? u[(int64_t)o * m + r] : u[(int64_t)r * n + o];
              if (J.u_trans) u[(int64_t)i * m + r] -= d * uo; else u[(int64_t)r * n + i] -= d * uo;
            }
          }
        }
        double nn = 0.0;
        for (int r = lane; r < m; r += 64) {
          const double ui = J.u_trans ? u[(int64_t)i * m + r] : u[(int64_t)r * n + i];
          nn += ui * ui;
        }
        nn = tnx_wave_sum(nn);
        if (nn > 0.25) {
          const double sc = 1.0 / sqrt(nn);
          for (int r = lane; r < m; r += 64) {
            if (J.u_trans) u[(int64_t)i * m + r] *= sc; else u[(int64_t)r * n + i] *= sc;
          }
          ++next_unit;
          break;
        }
      }
    }
  }
}

// Cholesky factor + inverse of one n x n SPD matrix (n <= 128), 256 lanes.
// Lane (tr, tc) = (tid / 16, tid % 16) holds the 8 x 8 register tile of rows
// {tr + 16a} and columns {tc + 16b}; the cyclic split keeps the shrinking
// trailing matrix spread over every lane.  Right-looking Cholesky of
// S = G + shift I = R^T R: steps k and k+1 broadcast rows k and k+1 through
// LDS (row pairs double-buffered, so one barrier per two steps) and each lane
// applies the two rank-1 updates to its tile.  R is then parked in LDS and
// R^{-1} is formed from the bottom row up, rows k and k-1 per barrier:
// X[k,:] /= R[k,k], X[i,:] -= R[i,k] X[k,:] for i < k.  Every lane forms the
// second row of a pair from the broadcast rows with the same operations, in
// the same order, as the step that would have followed a barrier, so the
// results are those of one step per barrier.  The factorisation keeps
// 1 / R[k,k] (the 1 / sqrt it divides by) in LDS, so the inverse divides
// nothing.
constexpr int CHOL_NMAX = 128;
constexpr double CHOL_NEAR_ID = 1e-10;  // largest |S - I| entry of the first-order path
constexpr int CHOL_LD = CHOL_NMAX + 1;  // LDS row stride of R: column reads hit distinct banks
constexpr size_t CHOL_LDS = sizeof(double) * (CHOL_NMAX * CHOL_LD + 5 * CHOL_NMAX);

// Cholesky steps k in [16 KA, 16 KA + 16) (k < n), in pairs (k, k+1) that
// stay inside the block (16 KA is even).  With the tile-row block KA of
// column k a compile-time constant, every register index is static, and the
// blocks a < KA / b < KA (rows and columns already factored, whose update
// factors are zero) drop out of the rank-1 updates at compile time, as do
// the blocks a > b: they lie strictly below the diagonal (row tr + 16a >
// column tc + 16b), an entry no broadcast or output reads.  `it` counts
// barriers (buffer parity).
template <int KA>
__device__ __forceinline__ void chol_steps(double (&s)[8][8], double* rowbuf, double* rdiag, int n,
                                           int tr, int tc, int& bad, int& it) {
  const int kend = min(16 * KA + 16, n);
  for (int k = 16 * KA; k < kend; k += 2) {
    double* r0 = rowbuf + (it & 1) * 2 * CHOL_NMAX;
    double* r1 = r0 + CHOL_NMAX;
    ++it;
    const bool two = k + 1 < kend;
    if (tr == (k & 15)) {
#pragma unroll
      for (int b = 0; b < 8; ++b) r0[tc + 16 * b] = s[KA][b];
    }
    if (two && tr == ((k + 1) & 15)) {
#pragma unroll
      for (int b = 0; b < 8; ++b) r1[tc + 16 * b] = s[KA][b];
    }
    __syncthreads();
    double d0 = r0[k];
    if (!(d0 > 0.0)) {
      d0 = 1e-300;
      bad = 1;
    }
    // one reciprocal square root on the serial chain instead of sqrt + divide
    const double ri0 = rsqrt(d0), sq0 = d0 * ri0;
    if (threadIdx.x == 0) rdiag[k] = ri0;
    double ci0[8], cj0[8];
#pragma unroll
    for (int a = KA; a < 8; ++a) ci0[a] = (a > KA || tr + 16 * a > k) ? r0[tr + 16 * a] * ri0 : 0.0;
#pragma unroll
    for (int b = KA; b < 8; ++b) cj0[b] = (b > KA || tc + 16 * b > k) ? r0[tc + 16 * b] * ri0 : 0.0;
#pragma unroll
    for (int a = KA; a < 8; ++a)
#pragma unroll
      for (int b = a; b < 8; ++b) s[a][b] = fma(-ci0[a], cj0[b], s[a][b]);
    double sq1 = 0.0, ri1 = 0.0;
    if (two) {
      // row k+1 after step k: fma(-c0[k+1], c0[x], S[k+1][x]), as its owner
      // would have broadcast it
      const double c0k1 = r0[k + 1] * ri0;
      double d1 = fma(-c0k1, c0k1, r1[k + 1]);
      if (!(d1 > 0.0)) {
        d1 = 1e-300;
        bad = 1;
      }
      ri1 = rsqrt(d1);
      sq1 = d1 * ri1;
      if (threadIdx.x == 0) rdiag[k + 1] = ri1;
      double ci1[8], cj1[8];
#pragma unroll
      for (int a = KA; a < 8; ++a) {
        const int i = tr + 16 * a;
        ci1[a] = (a > KA || i > k + 1) ? fma(-c0k1, r0[i] * ri0, r1[i]) * ri1 : 0.0;
      }
#pragma unroll
      for (int b = KA; b < 8; ++b) {
        const int j = tc + 16 * b;
        cj1[b] = (b > KA || j > k + 1) ? fma(-c0k1, r0[j] * ri0, r1[j]) * ri1 : 0.0;
      }
#pragma unroll
      for (int a = KA; a < 8; ++a)
#pragma unroll
        for (int b = a; b < 8; ++b) s[a][b] = fma(-ci1[a], cj1[b], s[a][b]);
    }
    if (tr == (k & 15)) {
#pragma unroll
      for (int b = KA; b < 8; ++b) {
        const int j = tc + 16 * b;
        s[KA][b] = (j > k) ? s[KA][b] * ri0 : (j == k ? sq0 : s[KA][b]);
      }
    }
    if (two && tr == ((k + 1) & 15)) {
#pragma unroll
      for (int b = KA; b < 8; ++b) {
        const int j = tc + 16 * b;
        s[KA][b] = (j > k + 1) ? s[KA][b] * ri1 : (j == k + 1 ? sq1 : s[KA][b]);
      }
    }
  }
}

// Inverse steps k in [16 KA, 16 KA + 16) from the top down (k < n), in pairs
// (k, k-1) inside the block: rows of blocks a > KA lie below k (no update),
// columns of blocks b < KA lie left of k, where row k of X = R^{-1} is zero.
template <int KA>
__device__ __forceinline__ void inv_steps(double (&s)[8][8], const double* Rl, const double* rdiag,
                                          double* rowbuf, int n, int tr, int tc, int& it) {
  const int khi = min(16 * KA + 15, n - 1);
  for (int k = khi; k >= 16 * KA; k -= 2) {
    double* x0 = rowbuf + (it & 1) * 2 * CHOL_NMAX;
    double* x1 = x0 + CHOL_NMAX;
    ++it;
    const bool two = k - 1 >= 16 * KA;
    if (tr == (k & 15)) {
      const double rkk = rdiag[k];
#pragma unroll
      for (int b = KA; b < 8; ++b) {
        s[KA][b] *= rkk;
        x0[tc + 16 * b] = s[KA][b];
      }
    }
    if (two && tr == ((k - 1) & 15)) {   // row k-1 before step k
#pragma unroll
      for (int b = KA; b < 8; ++b) x1[tc + 16 * b] = s[KA][b];
    }
    __syncthreads();
    double ra[8], xv[8];
#pragma unroll
    for (int a = 0; a <= KA; ++a)
      ra[a] = (a < KA || tr + 16 * a < k) ? Rl[(tr + 16 * a) * CHOL_LD + k] : 0.0;
#pragma unroll
    for (int b = KA; b < 8; ++b) xv[b] = x0[tc + 16 * b];
#pragma unroll
    for (int a = 0; a <= KA; ++a)
#pragma unroll
      for (int b = KA; b < 8; ++b) s[a][b] = fma(-ra[a], xv[b], s[a][b]);
    if (two) {
      const int k1 = k - 1;
      const double rk1 = rdiag[k1];
      const double rk1k = Rl[k1 * CHOL_LD + k];
      if (tr == (k1 & 15)) {
#pragma unroll
        for (int b = KA; b < 8; ++b) s[KA][b] *= rk1;
      }
      // X[k-1,:] = (row k-1 after step k) / R[k-1,k-1], as its owner holds it
      double xw[8];
#pragma unroll
      for (int b = KA; b < 8; ++b) xw[b] = fma(-rk1k, xv[b], x1[tc + 16 * b]) * rk1;
#pragma unroll
      for (int a = 0; a <= KA; ++a)
        ra[a] = (a < KA || tr + 16 * a < k1) ? Rl[(tr + 16 * a) * CHOL_LD + k1] : 0.0;
#pragma unroll
      for (int a = 0; a <= KA; ++a)
#pragma unroll
        for (int b = KA; b < 8; ++b) s[a][b] = fma(-ra[a], xw[b], s[a][b]);
    }
  }
}

__global__ __launch_bounds__(NT) void chol_inv_kernel(const tnx_chol_job* __restrict__ jobs) {
  extern __shared__ double lds[];
  double* Rl = lds;                         // R, row stride CHOL_LD
  double* rowbuf = lds + CHOL_NMAX * CHOL_LD;  // two broadcast row pairs
  double* rdiag = rowbuf + 4 * CHOL_NMAX;      // 1 / R[k,k]
  const tnx_chol_job J = jobs[blockIdx.x];
  const int n = J.n;
  if (n <= 0 || n > CHOL_NMAX) return;  // rejected on the host
  const double* G = (const double*)J.g;
  const int ld = J.ld > 0 ? J.ld : n;
  __shared__ double red[4];
  __shared__ double shift;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double s[8][8];
  double trace = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int i = tr + 16 * a;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int j = tc + 16 * b;
      s[a][b] = (i < n && j < n) ? G[(int64_t)i * ld + j] : 0.0;
      if (i == j) trace += s[a][b];
    }
  }
  trace = tnx_wave_sum(trace);
  if (lane == 0) red[wave] = trace;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = J.shift_ref ? *(const double*)J.shift_ref
                                 : ((red[0] + red[1]) + (red[2] + red[3]));
    shift = J.shift_scale * t;
  }
  __syncthreads();
  // Near-identity Gram matrix (the last pass of CholeskyQR3, whose panel is
  // already orthonormal to working accuracy): with S = I + E, ||E||_max <=
  // CHOL_NEAR_ID and no shift, R = I + U (U = strict upper part of E plus
  // half its diagonal) and R^{-1} = I - U hold up to O(E^2) <= 1e-20, below
  // the rounding of the sequential factorisation, and need no sequential
  // steps.
  {
    double emax = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int i = tr + 16 * a, j = tc + 16 * b;
        if (i < n && j < n) emax = fmax(emax, fabs(s[a][b] - (i == j ? 1.0 : 0.0)));
      }
    for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
    __shared__ double redm[4];
    if (lane == 0) redm[wave] = emax;
    __syncthreads();
    const double em = fmax(fmax(redm[0], redm[1]), fmax(redm[2], redm[3]));
    if (shift == 0.0 && em <= CHOL_NEAR_ID) {
      double* R = (double*)J.r;
      double* Ri = (double*)J.rinv;
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const int i = tr + 16 * a, j = tc + 16 * b;
          if (i < n && j < n) {
            const double u = j > i ? s[a][b] : (j == i ? 0.5 * (s[a][b] - 1.0) : 0.0);
            const double id = (i == j) ? 1.0 : 0.0;
            R[(int64_t)i * ld + j] = id + u;
            Ri[(int64_t)i * ld + j] = id - u;
          }
        }
      if (threadIdx.x == 0 && J.status != 0) *(int32_t*)J.status = 0;
      return;
    }
  }
  {
    const double sh = shift;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
      for (int b = 0; b < 8; ++b)
        if (tr + 16 * a == tc + 16 * b && tr + 16 * a < n) s[a][b] += sh;
  }
  int bad = 0, it = 0;
  // Cholesky: row k of the tiles becomes row k of R at step k.  The trailing
  // block is updated on and above the diagonal blocks, so the part of row k
  // right of column k is current when its owners broadcast it.
  chol_steps<0>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<1>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<2>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<3>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<4>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<5>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<6>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  chol_steps<7>(s, rowbuf, rdiag, n, tr, tc, bad, it);
  double* R = (double*)J.r;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int i = tr + 16 * a;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int j = tc + 16 * b;
      if (i < n && j < n) R[(int64_t)i * ld + j] = (j >= i) ? s[a][b] : 0.0;
      Rl[i * CHOL_LD + j] = s[a][b];
    }
  }
  __syncthreads();
  // X = R^{-1}: the tiles restart as the identity; row k is final once the
  // rows below it have been subtracted.
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      s[a][b] = (tr + 16 * a == tc + 16 * b && tr + 16 * a < n) ? 1.0 : 0.0;
  inv_steps<7>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<6>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<5>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<4>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<3>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<2>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<1>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  inv_steps<0>(s, Rl, rdiag, rowbuf, n, tr, tc, it);
  double* Ri = (double*)J.rinv;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int i = tr + 16 * a;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int j = tc + 16 * b;
      if (i < n && j < n) Ri[(int64_t)i * ld + j] = (j >= i) ? s[a][b] : 0.0;
    }
  }
  if (threadIdx.x == 0 && J.status != 0) *(int32_t*)J.status = bad;
}

}  // namespace

extern "C" int tnx_chol_inv_f64(const tnx_chol_job* jobs, int njob, tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0) {
    tnx_set_error("tnx_chol_inv_f64: bad arguments");
    return TNX_EINVAL;
  }
  static bool attr_set = false;
  if (!attr_set) {
    TNX_CHECK_HIP(hipFuncSetAttribute((const void*)chol_inv_kernel,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)CHOL_LDS));
    attr_set = true;
  }
  hipLaunchKernelGGL(chol_inv_kernel, dim3(njob), dim3(NT), CHOL_LDS, (hipStream_t)stream, jobs);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_qr_f64(const tnx_qr_job* jobs, int njob, tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0) {
    tnx_set_error("tnx_qr_f64: bad arguments");
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(qr_kernel, dim3(njob), dim3(NT), 0, (hipStream_t)stream, jobs);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}

extern "C" int tnx_svd_jacobi_f64(const tnx_svd_job* jobs, int njob, int max_sweeps,
                                  tnx_stream_t stream) {
  if (njob == 0) return TNX_OK;
  if (jobs == nullptr || njob < 0 || max_sweeps < 1) {
    tnx_set_error("tnx_svd_jacobi_f64: bad arguments");
    return TNX_EINVAL;
  }
  hipLaunchKernelGGL(svd_kernel, dim3(njob), dim3(NT), 0, (hipStream_t)stream, jobs, max_sweeps);
  TNX_CHECK_LAUNCH();
  return TNX_OK;
}
