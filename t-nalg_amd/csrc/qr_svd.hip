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

// Cholesky factor + inverse of one n x n SPD matrix (n <= 128), 256 lanes,
// in 16 x 16 blocks (nb = ceil(n / 16) block rows; the padding is the
// identity).  Right-looking, with the inverse carried along: eliminating
// [S | I] = R^T [R | F] gives R and F = R^{-T}, so R^{-1} = F^T comes out of
// the same steps.  Per panel p:
//   1. the owner wave of the diagonal block factors S_pp = R_pp^T R_pp in
//      registers (the block in the v_mfma_f64_16x16x4_f64 output layout,
//      row k read by lane shuffles, one reciprocal square root per step),
//      F_pp = R_pp^{-T} from the same steps;
//   2. row panel p on MFMA: R_pb = F_pp S_pb (b > p), F_pe = F_pp E_pe (e < p);
//   3. trailing update on MFMA: S_ab -= R_pa^T R_pb (p < a <= b) and
//      E_ae -= R_pa^T F_pe (a > p, e <= p).
// Wave w owns block columns w and 7 - w of the upper part (S -> R) and of
// the strictly lower part (E -> F): 16 blocks of 4 accumulator registers per
// lane, about the same work per wave in every phase.  The owner of the next
// diagonal block updates it first and factors it while the other waves
// finish the trailing update; two barriers per panel.  LDS holds the row
// panels as they are formed (R above the diagonal, F below it, F_pp in its
// own buffer), which are the MFMA operands of the later phases and the
// output at the end.
constexpr int CHOL_NMAX = 128;
constexpr double CHOL_NEAR_ID = 1e-10;  // largest |S - I| entry of the first-order path
constexpr int CB_LD = CHOL_NMAX + 1;    // LDS row stride: row and column reads hit distinct banks
constexpr int CB_XLD = 17;              // row stride of the F_pp blocks
constexpr size_t CHOL_LDS = sizeof(double) * (CHOL_NMAX * CB_LD + 8 * 16 * CB_XLD);

// block (a, b) of accumulator slot s of wave w: slots 0, 1 the diagonal
// blocks (w, w), (7-w, 7-w); 2..8 the blocks above them in those columns;
// 9..15 the blocks below them (the F part)
__device__ __forceinline__ void chol_slot(int w, int s, int& a, int& b) {
  const int c0 = w, c1 = 7 - w;
  if (s == 0) {
    a = c0; b = c0;
  } else if (s == 1) {
    a = c1; b = c1;
  } else if (s <= 8) {
    const int t = s - 2;
    if (t < w) { a = t; b = c0; } else { a = t - w; b = c1; }
  } else {
    const int t = s - 9;
    if (t < 7 - w) { a = c0 + 1 + t; b = c0; } else { a = 8 - w + (t - (7 - w)); b = c1; }
  }
}

// 16 x 16 block in the MFMA output layout: row (lane / 16) + 4 r, column lane % 16
__device__ __forceinline__ tnx_d4 chol_ld(const double* p, int ld, int lane) {
  tnx_d4 v;
#pragma unroll
  for (int r = 0; r < 4; ++r) v[r] = p[((lane >> 4) + 4 * r) * ld + (lane & 15)];
  return v;
}
__device__ __forceinline__ void chol_st(double* p, int ld, int lane, tnx_d4 v) {
#pragma unroll
  for (int r = 0; r < 4; ++r) p[((lane >> 4) + 4 * r) * ld + (lane & 15)] = v[r];
}

// acc += sgn * A B over K = 16, A[m][k] = pa[m * sam + k * sak], B[k][n] = pb[k * sbk + n]
__device__ __forceinline__ tnx_d4 chol_mma(tnx_d4 acc, const double* pa, int sam, int sak,
                                           const double* pb, int sbk, int lane, double sgn) {
  const int mn = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 16; kk += 4) {
    const double av = sgn * pa[mn * sam + (kk + kq) * sak];
    const double bv = pb[(kk + kq) * sbk + mn];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  return acc;
}

__device__ __forceinline__ double chol_readlane(double v, int l) {
  const long long x = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)x, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(x >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// d (the diagonal block, output layout) -> R_pp with zeros below the
// diagonal; f -> F_pp = R_pp^{-T}.  Step k: R[k][j] = D[k][j] / sqrt(D[k][k])
// (j > k), F[k][j] = E[k][j] / sqrt(D[k][k]); rows i > k lose R[k][i] times
// row k of R and of F.  Row k sits in lanes 16 (k % 4) .. +15, register k / 4.
__device__ __forceinline__ void chol_diag(tnx_d4& d, tnx_d4& f, int lane, int& bad) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) f[r] = (g + 4 * r == j) ? 1.0 : 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int kr = k >> 2, src = 16 * (k & 3);
    double piv = chol_readlane(d[kr], src + k);
    const double dkj = __shfl(d[kr], src + j, 64);
    const double ekj = __shfl(f[kr], src + j, 64);
    double dki[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) dki[r] = __shfl(d[kr], src + g + 4 * r, 64);
    if (!(piv > 0.0)) {
      piv = 1e-300;
      bad = 1;
    }
    // one reciprocal square root on the serial chain instead of sqrt + divide
    const double rk = rsqrt(piv), sq = piv * rk;
    const double rkj = (j > k) ? dkj * rk : (j == k ? sq : 0.0);
    const double fkj = ekj * rk;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (g + 4 * r > k) {
        const double rki = dki[r] * rk;
        d[r] = fma(-rki, rkj, d[r]);
        f[r] = fma(-rki, fkj, f[r]);
      }
    }
    if (g == (k & 3)) {
      d[kr] = rkj;
      f[kr] = fkj;
    }
  }
}

__global__ __launch_bounds__(NT) void chol_inv_kernel(const tnx_chol_job* __restrict__ jobs) {
  extern __shared__ double lds[];
  double* Sl = lds;                          // CHOL_NMAX rows of CB_LD
  double* Xd = lds + CHOL_NMAX * CB_LD;      // F_pp, 16 rows of CB_XLD per panel
  const tnx_chol_job J = jobs[blockIdx.x];
  const int n = J.n;
  if (n <= 0 || n > CHOL_NMAX) return;  // rejected on the host
  const double* G = (const double*)J.g;
  const int ld = J.ld > 0 ? J.ld : n;
  const int nb = (n + 15) >> 4, np = 16 * nb;
  __shared__ double red[8];
  __shared__ int bad_s;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // G padded with the identity; trace (shift) and largest |G - I| entry
  double trace = 0.0, emax = 0.0;
  for (int i = wave; i < np; i += 4)
    for (int j = lane; j < np; j += 64) {
      double v = (i == j) ? 1.0 : 0.0;
      if (i < n && j < n) {
        const double gv = G[(int64_t)i * ld + j];
        if (i == j) trace += gv;
        emax = fmax(emax, fabs(gv - v));
        v = gv;
      }
      Sl[i * CB_LD + j] = v;
    }
  trace = tnx_wave_sum(trace);
  for (int o = 32; o > 0; o >>= 1) emax = fmax(emax, __shfl_xor(emax, o, 64));
  if (lane == 0) {
    red[wave] = trace;
    red[4 + wave] = emax;
  }
  if (threadIdx.x == 0) bad_s = 0;
  __syncthreads();
  const double shift = J.shift_scale * (J.shift_ref ? *(const double*)J.shift_ref
                                                    : ((red[0] + red[1]) + (red[2] + red[3])));
  const double em = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
  double* R = (double*)J.r;
  double* Ri = (double*)J.rinv;
  // Near-identity Gram matrix (the last pass of CholeskyQR3, whose panel is
  // already orthonormal to working accuracy): with S = I + E, ||E||_max <=
  // CHOL_NEAR_ID and no shift, R = I + U (U = strict upper part of E plus
  // half its diagonal) and R^{-1} = I - U hold up to O(E^2) <= 1e-20, below
  // the rounding of the factorisation, and need no sequential steps.
  if (shift == 0.0 && em <= CHOL_NEAR_ID) {
    for (int i = wave; i < n; i += 4)
      for (int j = lane; j < n; j += 64) {
        const double sv = Sl[i * CB_LD + j];
        const double u = j > i ? sv : (j == i ? 0.5 * (sv - 1.0) : 0.0);
        const double id = (i == j) ? 1.0 : 0.0;
        R[(int64_t)i * ld + j] = id + u;
        Ri[(int64_t)i * ld + j] = id - u;
      }
    if (threadIdx.x == 0 && J.status != 0) *(int32_t*)J.status = 0;
    return;
  }
  int sa[16], sb[16];
  tnx_d4 acc[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    chol_slot(wave, s, sa[s], sb[s]);
    acc[s] = tnx_d4{0.0, 0.0, 0.0, 0.0};
    if (s <= 8 && sb[s] < nb) {
      acc[s] = chol_ld(Sl + 16 * sa[s] * CB_LD + 16 * sb[s], CB_LD, lane);
      if (sa[s] == sb[s]) {   // the shift on the diagonal (not on the padding)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * sa[s] + (lane >> 4) + 4 * r;
          if ((lane >> 4) + 4 * r == (lane & 15) && i < n) acc[s][r] += shift;
        }
      }
    }
  }
  int bad = 0;
  for (int p = 0; p < nb; ++p) {
    const double* Pp = Sl + 16 * p * CB_LD;       // row panel p
    if (p > 0) {   // trailing update of panel p-1, diagonal blocks first
      const double* Pq = Pp - 16 * CB_LD;
#pragma unroll
      for (int s = 0; s < 2; ++s)
        if (sa[s] > p - 1 && sa[s] < nb)
          acc[s] = chol_mma(acc[s], Pq + 16 * sa[s], 1, CB_LD, Pq + 16 * sb[s], CB_LD, lane, -1.0);
    }
    if (min(p, 7 - p) == wave) {   // factor the diagonal block of panel p
      tnx_d4 dd = (p == wave) ? acc[0] : acc[1], ff;
      chol_diag(dd, ff, lane, bad);
      chol_st(Sl + 16 * p * CB_LD + 16 * p, CB_LD, lane, dd);
      chol_st(Xd + 16 * p * CB_XLD, CB_XLD, lane, ff);
    }
    if (p > 0) {
      const double* Pq = Pp - 16 * CB_LD;
      const double* Fq = Xd + 16 * (p - 1) * CB_XLD;
#pragma unroll
      for (int s = 2; s < 9; ++s)
        if (sa[s] > p - 1 && sb[s] < nb)
          acc[s] = chol_mma(acc[s], Pq + 16 * sa[s], 1, CB_LD, Pq + 16 * sb[s], CB_LD, lane, -1.0);
#pragma unroll
      for (int s = 9; s < 16; ++s)
        if (sa[s] > p - 1 && sa[s] < nb && sb[s] <= p - 1) {
          const bool diag = sb[s] == p - 1;
          acc[s] = chol_mma(acc[s], Pq + 16 * sa[s], 1, CB_LD, diag ? Fq : Pq + 16 * sb[s],
                            diag ? CB_XLD : CB_LD, lane, -1.0);
        }
    }
    // row panel p of S (above the diagonal) and E (below it) as updated
#pragma unroll
    for (int s = 2; s < 16; ++s)
      if (sa[s] == p && sb[s] < nb) chol_st(Sl + 16 * p * CB_LD + 16 * sb[s], CB_LD, lane, acc[s]);
    __syncthreads();
    // R_pb = F_pp S_pb, F_pe = F_pp E_pe
    const double* Fp = Xd + 16 * p * CB_XLD;
#pragma unroll
    for (int s = 2; s < 16; ++s)
      if (sa[s] == p && sb[s] < nb) {
        const tnx_d4 c = chol_mma(tnx_d4{0.0, 0.0, 0.0, 0.0}, Fp, CB_XLD, 1, Pp + 16 * sb[s], CB_LD,
                                  lane, 1.0);
        chol_st(Sl + 16 * p * CB_LD + 16 * sb[s], CB_LD, lane, c);
      }
    __syncthreads();
  }
  if (bad && lane == 0) bad_s = 1;
  __syncthreads();
  // R (upper) and R^{-1} = F^T: F_pp from its buffer, the blocks below the
  // diagonal of F from the lower part of Sl
  for (int i = wave; i < n; i += 4)
    for (int j = lane; j < n; j += 64) {
      double rv = 0.0, xv = 0.0;
      if (j >= i) {
        rv = Sl[i * CB_LD + j];
        xv = ((i >> 4) == (j >> 4)) ? Xd[16 * (j >> 4) * CB_XLD + (j & 15) * CB_XLD + (i & 15)]
                                    : Sl[j * CB_LD + i];
      }
      R[(int64_t)i * ld + j] = rv;
      Ri[(int64_t)i * ld + j] = xv;
    }
  if (threadIdx.x == 0 && J.status != 0) *(int32_t*)J.status = bad_s;
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
