/*
 * tnalg_hip.h - C ABI of the MI355X (gfx950) finite-size DMRG kernels.
 *
 * The reference (ranshiju/T-Nalg) is pure Python: it has no FFI.  Each entry
 * point below replaces a numpy/scipy routine on the path of
 * dmrg_finite_size (algorithms/DMRG_anyH.py:18-101); the comment over it cites
 * the reference function (file:line) whose arithmetic it takes over.  A
 * Python ctypes binding of exactly these symbols lives in
 * t-nalg_amd/tnalg_amd/_lib.py.
 *
 * Conventions
 *  - All pointers are DEVICE pointers unless a parameter says "host".  The
 *    library never allocates persistent memory: every buffer (operands,
 *    descriptors, scratch) belongs to the caller.
 *  - Matrices are row-major.  MPS tensors are (Dl, d, Dr) row-major; an
 *    environment matrix is indexed [bra, ket].
 *  - Every call is ordered on the HIP stream it is given (NULL = the default
 *    stream) and returns 0 on success or a negative tnx_status; the message of
 *    the last failure on the calling thread is tnx_last_error().
 *  - Descriptor arrays are plain structs (layouts below are part of the ABI;
 *    sizes are checked by tnx_abi_sizes()).
 */
#ifndef TNALG_HIP_H
#define TNALG_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* tnx_stream_t; /* a hipStream_t */

enum tnx_status {
  TNX_OK = 0,
  TNX_EINVAL = -1,   /* invalid argument (shape, dtype, count) */
  TNX_EHIP = -2,     /* a HIP runtime call failed */
  TNX_ENOCONV = -3,  /* eigensolver hit its restart limit (result still written) */
};

/* element types of operands; TNX_DTYPE_PER_SOLVE as the x_dtype / w_dtype
 * argument of a Lanczos call takes each solve's own x_dtype / w_dtype field */
enum tnx_dtype { TNX_F64 = 0, TNX_F32 = 1, TNX_BF16 = 2, TNX_DTYPE_PER_SOLVE = -1 };

/* GEMM precision modes: (operand type -> accumulate -> output type)       */
enum tnx_gemm_mode {
  TNX_GEMM_F64 = 0,       /* f64 x f64 -> MFMA f64 -> f64                   */
  TNX_GEMM_F32 = 1,       /* f32 x f32 -> MFMA f32 -> f32                   */
  TNX_GEMM_BF16_F32 = 2,  /* bf16 x bf16 -> MFMA bf16 (f32 acc) -> f32      */
  TNX_GEMM_BF16_BF16 = 3, /* bf16 x bf16 -> MFMA bf16 (f32 acc) -> bf16     */
};

/* One K-segment of a grouped GEMM problem: C += alpha * op(A) * op(B).
 * op(A) is M x K: A[i*lda+k] (transA=0) or A[k*lda+i] (transA=1).
 * op(B) is K x N: B[k*ldb+j] (transB=0) or B[j*ldb+k] (transB=1).           */
typedef struct tnx_gemm_seg {
  uint64_t a, b;
  int32_t lda, ldb;
  int32_t transA, transB;
  int32_t K, reserved;
  double alpha;
} tnx_gemm_seg; /* 48 bytes */

/* One output matrix: C = sum_seg alpha*op(A)op(B) + beta*C0 (C0 may be 0). */
typedef struct tnx_gemm_problem {
  uint64_t c, c0;
  int32_t M, N;
  int32_t ldc, ldc0;
  int32_t seg_begin, seg_count;
  int32_t tile_begin;  /* prefix sum of 64x64 output tiles over problems */
  int32_t active;      /* index into the active-flag array, or -1 */
  double beta;
} tnx_gemm_problem; /* 56 bytes */

/* One output tensor of the elementwise physical-index mixer:
 *   dst[a,s,b] (=|+=) sum_terms coef * sum_t O[s,t] src[a,t,b]
 * over an (Dl, d, Dr) tensor; a D x D matrix is (D, 1, D); op < 0 means the
 * identity.                                                                  */
typedef struct tnx_mix_job {
  uint64_t dst;
  int32_t Dl, d, Dr;   /* d <= 4 */
  int32_t dst_dtype;   /* tnx_dtype */
  int32_t term_begin, term_count;
  int32_t accumulate;  /* 1: add into dst */
  int32_t active;      /* index into the active-flag array, or -1 */
  int32_t block_begin; /* prefix sum over jobs of ceil(Dl * Dr / P) workgroups,
                          P = 1024 for d <= 2, 512 for d > 2 (a workgroup
                          covers P (a, b) pairs, every s) */
  int32_t reserved;
} tnx_mix_job; /* 48 bytes */

typedef struct tnx_mix_term {
  uint64_t src;
  double coef;
  int32_t op;          /* index into the d x d f64 operator table, or -1 */
  int32_t src_dtype;   /* tnx_dtype */
} tnx_mix_term; /* 24 bytes */

/* The effective-Hamiltonian matvec of one local step of a batch of solves,
 * expressed as   T_k = F_k . x   (GEMM 1, grouped),
 *                U   = mix(T, x) (elementwise),
 *                y   = U_0 + sum U_r . R_r^T  (GEMM 2, K-concatenated).
 * The arrays are device descriptor arrays built by the host planner.       */
typedef struct tnx_matvec_plan {
  const tnx_gemm_problem* g1_prob; const tnx_gemm_seg* g1_seg;
  int32_t g1_nprob, g1_ntiles;
  const tnx_mix_job* mix_job; const tnx_mix_term* mix_term;
  int32_t mix_njob, mix_nblocks;
  const tnx_gemm_problem* g2_prob; const tnx_gemm_seg* g2_seg;
  int32_t g2_nprob, g2_ntiles;
  const double* ops;   /* operator table for the mixer */
  int32_t mode;        /* tnx_gemm_mode of GEMM 1 and GEMM 2 (see below)   */
  int32_t mode2;
  double work[3];      /* algorithmic flops of GEMM 1, bytes of the mixer and
                          flops of GEMM 2 with every solve active; read only
                          by the sampled kernel timing (0: not known)       */
} tnx_matvec_plan;

/* One eigenproblem of a batched Lanczos solve.                              */
typedef struct tnx_lanczos_solve {
  uint64_t V;          /* Krylov basis (v_dtype), ncv + 1 rows of stride ldv */
  uint64_t w;          /* matvec output (f64 or f32, see w_dtype)           */
  uint64_t x;          /* matvec input copy (tnx_dtype x_dtype)             */
  uint64_t init;       /* f64 start vector (read at start)                  */
  uint64_t out;        /* f64 result vector (written at the end)            */
  int32_t n, ldv;      /* dimension, basis stride (elements)                */
  int32_t m;           /* reserved: not read by the kernels (the call's ncv
                          bounds the basis, and a solve whose basis spans all
                          n dimensions stops as converged)                   */
  int32_t chunk_begin; /* prefix sum of 1024-element chunks over solves     */
  double tol_scale;    /* converged when |r| <= tol_scale*|1 - tau*theta|   */
  double tau;          /*   (ARPACK's test on I - tau*H, MPSClass.py:886)   */
  double floor_rel;    /*   or |r| <= floor_rel*|theta| (precision floor)   */
  int32_t x_dtype;     /* tnx_dtype of x and of w, read when the call passes */
  int32_t w_dtype;     /*   TNX_DTYPE_PER_SOLVE                             */
  int32_t v_dtype;     /* storage of the basis: TNX_F64, or TNX_F32 for a
                          solve whose matvec runs in f32 or bf16 (half the
                          basis traffic; dots, norms and the tridiagonal
                          problem stay f64, and the second projection is
                          left out below 1e-6 instead of 1e-12 of the new
                          coupling, the basis's own rounding level)        */
  int32_t reserved;
} tnx_lanczos_solve; /* 96 bytes */

/* Per-solve results of tnx_lanczos_run (host-visible after the call).      */
typedef struct tnx_lanczos_result {
  double theta;        /* Ritz value (lowest eigenvalue of H)               */
  double resid;        /* |beta_j y_j| at exit                              */
  int32_t iters;       /* matvecs spent                                     */
  int32_t converged;
} tnx_lanczos_result; /* 24 bytes */

/* ------------------------------------------------------------------------ */
/* version / errors                                                          */
int tnx_version(void);
const char* tnx_last_error(void);
/* writes sizeof of each ABI struct in declaration order into sizes[0..n)    */
int tnx_abi_sizes(int32_t* sizes, int n);

/* Grouped GEMM over descriptor arrays (device).  Replaces the numpy
 * contractions of library/TensorBasicModule.py:489-613 (absorb_matrix2tensor,
 * absorb_matrices2tensor[_full_fast]) and :632-721 (bound_vec_operator_left2right
 * / right2left) as used by MPSClass.py:837-858 and MPSClass.py:555-611.
 * ntiles = total 64x64 tiles; active = device int flags or NULL.
 * mode = tnx_gemm_mode | (variant << 8): a mode may have several kernel
 * variants (LDS staging depth and occupancy), all giving the same bits;
 * variant 0 is the default.  The matvec plan's mode / mode2 fields take the
 * same encoding.                                                           */
int tnx_gemm_grouped(int mode, const tnx_gemm_problem* prob, int nprob,
                     const tnx_gemm_seg* seg, int ntiles, const int32_t* active,
                     tnx_stream_t stream);

/* Number of kernel variants of a GEMM mode (TNX_EINVAL for an unknown mode).
 * No reference counterpart: the engine's tuner times the variants of a mode
 * on the first launches of each shape class and keeps the fastest.         */
int tnx_gemm_variants(int mode);

/* Elementwise physical-index mixer / linear combination / dtype cast.
 * Replaces absorb_matrix2tensor(tensor, op.T, 1) (TensorBasicModule.py:489-526
 * with bond=1) and the opt_env accumulation of classify_and_update_env
 * (MPSClass.py:766-815).                                                     */
int tnx_mix(const tnx_mix_job* job, int njob, const tnx_mix_term* term, int nblocks,
            const double* ops, const int32_t* active, tnx_stream_t stream);

/* One effective-Hamiltonian matvec (GEMM 1, mix, GEMM 2) for a batch.
 * Replaces update_tensor_eigs_f_handle_optimized (MPSClass.py:837-858) without
 * the "psi - tau*" shift.                                                    */
int tnx_heff_matvec(const tnx_matvec_plan* plan /* host struct */, const int32_t* active,
                    tnx_stream_t stream);

/* Batched Lanczos eigensolver for the lowest eigenpair of each solve's Heff,
 * with f64 basis, two-pass classical Gram-Schmidt reorthogonalisation,
 * on-device tridiagonal eigensolve and restart with the Ritz vector; the
 * matvec is `plan` (its x / w buffers are the solves' x / w).  Replaces
 * scipy eigsh(LinearOperator(I - tau*Heff), k=1, which='LM', v0, tol) at
 * MPSClass.py:883-889 (semantics of Eigs_Module_sjr.eigs_fh, EIG:18-58).
 * solves: device array; state: device scratch of tnx_lanczos_state_bytes();
 * results: host array of nsolve entries.                                    */
size_t tnx_lanczos_state_bytes(int nsolve, int nchunks, int ncv);
int tnx_lanczos_run(const tnx_matvec_plan* plan, const tnx_lanczos_solve* solves, int nsolve,
                    int nchunks, int ncv, int max_restarts, int x_dtype, int w_dtype,
                    void* state, tnx_lanczos_result* results, tnx_stream_t stream);
/* As tnx_lanczos_run; the convergence test is applied only from Krylov step
 * min_iters on (min_iters = ncv reproduces ARPACK's habit of completing its
 * first ncv-step factorisation before testing).  The basis V of every solve
 * needs ncv + 1 rows (row ncv is scratch).                                  */
int tnx_lanczos_run_ex(const tnx_matvec_plan* plan, const tnx_lanczos_solve* solves, int nsolve,
                       int nchunks, int ncv, int max_restarts, int min_iters, int x_dtype,
                       int w_dtype, void* state, tnx_lanczos_result* results,
                       tnx_stream_t stream);
/* As tnx_lanczos_run_ex for solves whose matvecs run at different
 * precisions: each Krylov step applies the nplan matvec plans (host array)
 * in turn, and the solves' own x_dtype / w_dtype fields give the dtypes.
 * The `active` field of a plan's GEMM problems and mixer jobs is the
 * solve's index in `solves`.  The same replacement of MPSClass.py:883-889
 * as tnx_lanczos_run, for a batch that mixes the precisions of the sweep
 * schedule (one Lanczos loop instead of one per precision).               */
int tnx_lanczos_run_multi(const tnx_matvec_plan* plans, int nplan,
                          const tnx_lanczos_solve* solves, int nsolve, int nchunks, int ncv,
                          int max_restarts, int min_iters, void* state,
                          tnx_lanczos_result* results, tnx_stream_t stream);

/* No counterpart in the reference: tnx_lanczos_run_multi for callers that
 * cannot allocate device or host memory (code under HIP graph capture of
 * its neighbours, tight serving loops, harnesses that forbid allocation).
 * The state lives in a library-owned device array of
 * TNX_LANCZOS_STATIC_BYTES per GPU instead of a caller workspace, and no
 * host staging buffer is used; results are copied into the caller's host
 * array.  A call whose state (tnx_lanczos_state_bytes(nsolve, nchunks, ncv)
 * plus 24 bytes per solve, each rounded up to 256) exceeds that size fails
 * with TNX_EINVAL.  Every Krylov step of a cycle is queued without polling
 * (converged solves are masked on the device) and the host reads the active
 * count once per cycle; the results are bit-identical to
 * tnx_lanczos_run_multi.  All calls on one GPU share its array: calls that
 * overlap in time on one GPU race (serialise them).                        */
#define TNX_LANCZOS_STATIC_BYTES (4u << 20)
int tnx_lanczos_run_multi_static(const tnx_matvec_plan* plans, int nplan,
                                 const tnx_lanczos_solve* solves, int nsolve, int nchunks,
                                 int ncv, int max_restarts, int min_iters,
                                 tnx_lanczos_result* results, tnx_stream_t stream);

/* Batched f64 Householder QR gauge move.  For each matrix i, the logical
 * matrix M (m x n) is src[r*ld+c] (trans=0) or src[c*ld+r] (trans=1); on
 * return q holds the k=min(m,n) orthonormal columns (row-major m x k when
 * q_trans=0, i.e. q[r*k+c]; column-major k x m rows when q_trans=1, i.e.
 * q[c*m+r]) and r holds R (k x n row-major).  work: m*n f64 per matrix at
 * work + i*work_stride.  Replaces left2right_decompose_tensor /
 * right2left_decompose_tensor with way='qr' (TensorBasicModule.py:416-486).  */
typedef struct tnx_qr_job {
  uint64_t src, q, r, work;
  int32_t m, n, ld, trans;
  int32_t q_trans, reserved;
} tnx_qr_job; /* 56 bytes */
int tnx_qr_f64(const tnx_qr_job* jobs, int njob, tnx_stream_t stream);

/* Batched f64 Cholesky factor and inverse of small SPD matrices, the core of
 * the shifted Cholesky-QR (CholeskyQR3) gauge move on tall panels:
 *   S = G + shift*I, shift = shift_scale * trace(G);  S = R^T R (R upper);
 *   r <- R, rinv <- R^{-1}  (all n x n row-major, n <= 128).
 * Three passes (G = M^T M with shift, then Q^T Q twice) with the Gram and
 * Q = M R^{-1} products on the MFMA GEMM give M = Q R with Q orthonormal to
 * working accuracy for any conditioning of M.  Replaces np.linalg.qr of
 * left2right_decompose_tensor / right2left_decompose_tensor for tall panels
 * (TensorBasicModule.py:416-486).  status (device int32 per job, may be 0)
 * receives 1 if a pivot was not positive (replaced by a tiny value).       */
typedef struct tnx_chol_job {
  uint64_t g, r, rinv, status;
  uint64_t shift_ref;  /* device f64*: if set, shift = shift_scale * *shift_ref
                          (a blocked factorisation passes the full trace)    */
  int32_t n, ld;       /* order (<= 128) and leading dimension of g, r, rinv */
  double shift_scale;
} tnx_chol_job; /* 56 bytes */
int tnx_chol_inv_f64(const tnx_chol_job* jobs, int njob, tnx_stream_t stream);

/* Batched f64 one-sided Jacobi (Hestenes) SVD of M (m x n, m >= n given by
 * the same addressing as tnx_qr_job): u (m x n, layout as q above) holds the
 * left singular vectors, s the singular values in descending order, vt
 * (n x n row-major) the right ones scaled by s (i.e. diag(s) V^T, the matrix
 * the reference absorbs).  Replaces np.linalg.svd in
 * left2right_decompose_tensor / right2left_decompose_tensor with way='svd'
 * (TensorBasicModule.py:438-444,474-480) as used by
 * calculate_entanglement_spectrum (MPSClass.py:965-985).                    */
typedef struct tnx_svd_job {
  uint64_t src, u, s, svt, work;
  int32_t m, n, ld, trans;
  int32_t u_trans, reserved;
} tnx_svd_job; /* 64 bytes */
int tnx_svd_jacobi_f64(const tnx_svd_job* jobs, int njob, int max_sweeps, tnx_stream_t stream);

/* Batched f64 SVD of the n x n triangular factor R of an SVD gauge move
 * (n <= 1024), by block one-sided Jacobi over many workgroups, for
 * calculate_entanglement_spectrum (MPSClass.py:965-985) at large bond
 * dimension: with A = Q R from CholeskyQR3, np.linalg.svd(A) of
 * left2right_decompose_tensor / right2left_decompose_tensor
 * (TensorBasicModule.py:438-444,474-480) is A = (Q u^T) (v), lm = s.
 * Outputs (row-major n x n): row k of u is the k-th left singular vector of
 * R, row k of v is s_k times the k-th right singular vector, so R = u^T v;
 * s holds the singular values in descending order.  The columns of R^T are
 * rotated in blocks of bw (a power of two, 4 <= bw <= 32, bw * n <= 4096);
 * one launch per round of the round-robin over block pairs, sweeps until
 * none rotates (flag: one device int32 shared by the call).  Returns
 * TNX_ENOCONV after max_sweeps sweeps that still rotated.
 * A job with u == 0 computes the singular values only (work: n^2 + n + 1
 * doubles; bw * n <= 8192; r may equal work with ld_r == n), the
 * entanglement spectrum of a bond whose gauge move was a QR.              */
typedef struct tnx_svdj_job {
  uint64_t r;          /* R, row-major n x n, leading dimension ld_r (read)   */
  uint64_t work;       /* scratch: 2 n^2 + n + 1 doubles                      */
  uint64_t u, v, s;    /* outputs                                             */
  int32_t n, ld_r;
  int32_t bw;          /* columns per block                                   */
  int32_t nblk;        /* ceil(n / bw), rounded up to even                    */
  int32_t pair_begin;  /* prefix sum of nblk / 2 over the jobs                */
  int32_t reserved;
} tnx_svdj_job; /* 64 bytes */
int tnx_svd_block_jacobi(const tnx_svdj_job* jobs, int njob, int npairs, int max_nblk, int max_n,
                         int32_t* flag, int max_sweeps, tnx_stream_t stream);

/* Sampled kernel timing (no counterpart in the reference: it measures the
 * running engine for its roofline).  tnx_profile_begin(every, max_launches)
 * arms it: in every `every`-th Lanczos call made afterwards, on any host
 * thread, the launches of the first Krylov step (every solve active) are
 * each timed by their own blocks, which stamp their start and end with the
 * device's constant wall clock (a launch's time is first start to last end,
 * the duration a profiler's kernel trace shows), up to max_launches
 * launches; it allocates the stamp slots (1 KiB per launch) on the current
 * device.  tnx_profile_end waits for the device, fills out[0..cap) with
 * one entry per kernel class (TNX_PROF_NCLASS of them) and returns that
 * count.  The work of a timed launch is its algorithmic flops (GEMMs, from
 * the plan's `work`) or bytes (mixer; Lanczos passes: basis rows read and
 * written over nchunks * 1024 elements, only for calls with one plan; the
 * rows count 8 bytes for an f64 matvec plan and 4 otherwise, the f32 basis
 * of a low-precision matvec).                                             */
enum { TNX_PROF_NCLASS = 10 };
typedef struct tnx_prof_stat {
  char name[48];       /* kernel class                                      */
  int64_t launches;    /* launches of the class while armed (timed or not)  */
  int64_t samples;     /* launches timed                                    */
  double seconds;      /* their summed duration                             */
  double work;         /* their summed algorithmic work                     */
  double work_all;     /* the algorithmic work of every launch counted      */
  int32_t unit;        /* 0: flops, 1: bytes                                */
  int32_t reserved;
} tnx_prof_stat; /* 96 bytes */
int tnx_profile_begin(int every, int max_launches);
int tnx_profile_end(tnx_prof_stat* out, int cap);

/* Batched traces / Frobenius inner products: out[i] = sum_ab X_i[a,b] Y_i[a,b]
 * (Y = NULL: trace of X).  Replaces the np.trace closings of observation_s1
 * and observation_s1_s2 (MPSClass.py:1105-1157).                            */
typedef struct tnx_dot_job {
  uint64_t x, y;
  int32_t rows, cols;
  int32_t ldx, ldy;
} tnx_dot_job; /* 32 bytes */
int tnx_trace_dot(const tnx_dot_job* jobs, int njob, double* out, tnx_stream_t stream);

/* Batched in-place normalisation: x_i /= ||x_i||_2 over n_i f64 elements
 * (a zero vector is left as it is); one workgroup per job.  Replaces the
 * normalize=True of orthogonalize_mps / norm_mps on the centre tensor in the
 * time-evolution drivers (MPSClass.py:1421-1447, TEBDalgo.py:51-52).        */
typedef struct tnx_norm_job {
  uint64_t x;
  int64_t n;
} tnx_norm_job; /* 16 bytes */
int tnx_normalize_f64(const tnx_norm_job* jobs, int njob, tnx_stream_t stream);

/* Batched standard-normal fill: dst[i] = N(0, 1) number offset + i of
 * the stream of `seed` (i < n), f64, one workgroup per job.  The stream is
 * counter-based (Philox4x32-10 of the pair index (offset + i) / 2 under
 * the 64-bit key `seed`; Box-Muller of its two 53-bit uniforms gives the
 * pair), so a job's values do not depend on how a stream is cut into jobs.
 * Replaces the host np.random.randn draws of the initial random MPS
 * (TensorBasicModule.py:179-211) when the draw is made on the device.     */
typedef struct tnx_fill_job {
  uint64_t dst;
  int64_t n;
  uint64_t seed;
  int64_t offset;
} tnx_fill_job; /* 32 bytes */
int tnx_fill_normal_f64(const tnx_fill_job* jobs, int njob, tnx_stream_t stream);

/* Batched Kronecker sums: dst[(i,k),(j,l)] = sum_t coef_t A_t[i,j] B_t[k,l]
 * over the job's terms (row-major (p q) x (p q) dst with leading dimension
 * ldd; A_t p x p with lda, B_t q x q with ldb; a null factor is the
 * identity), KRON = 256 dst elements per workgroup.  Builds the blocks of
 * the infinite-size driver's effective Hamiltonian with one cell absorbed
 * into each environment: env_L (x) T (the cont of env, T, T, env in
 * effective_hamiltonian, MPSClass.py:1831-1845) and bath (x) 1 +
 * sum j eff (x) op (the np.kron sums of effective_hamilt_from_op,
 * MPSClass.py:1709-1727).                                                   */
typedef struct tnx_kron_term {
  uint64_t a, b;       /* f64 factors (0: identity)                          */
  int32_t lda, ldb;
  double coef;
} tnx_kron_term; /* 32 bytes */
typedef struct tnx_kron_job {
  uint64_t dst;
  int32_t p, q, ldd;
  int32_t term_begin, term_count;
  int32_t block_begin; /* prefix sum of ceil((p q)^2 / 256) over jobs        */
} tnx_kron_job; /* 32 bytes */
int tnx_kron_f64(const tnx_kron_job* jobs, int njob, const tnx_kron_term* terms, int nblocks,
                 tnx_stream_t stream);
/* Batched in-place symmetrisation X <- (X + X^T) / 2 of `count` n x n
 * blocks (leading dimension ld, block m at x + m * stride elements), 256
 * elements of the blocks per workgroup.  Replaces the symmetrisations of
 * the environments and the bath of the infinite-size driver
 * (MPSClass.py:1652-1664,1763).                                             */
typedef struct tnx_sym_job {
  uint64_t x;
  int32_t n, ld, count, stride;
  int32_t block_begin; /* prefix sum of ceil(count n^2 / 256) over jobs      */
  int32_t reserved;
} tnx_sym_job; /* 32 bytes */
int tnx_symmetrize_f64(const tnx_sym_job* jobs, int njob, int nblocks, tnx_stream_t stream);

/* Recorded launch sequence (no counterpart in the reference: the engine
 * records the launches of a step's gauge move and environment update once
 * per resident descriptor set, the calls above that dmrg_finite_size's
 * local step makes through MPSClass.py:860-891, and replays them with one
 * call per step).  The ops run in order on `stream`, each as the entry
 * point its kind names with the op's fields as arguments (null active
 * flags).  Returns the first failing op's status (the rest are not
 * issued); a block Jacobi SVD that reaches its sweep limit does not stop
 * the sequence, and the call then returns TNX_ENOCONV.                     */
enum tnx_op_kind {
  TNX_OP_GEMM = 1,   /* tnx_gemm_grouped(mode, a0, n0, a1, n1, 0)          */
  TNX_OP_MIX = 2,    /* tnx_mix(a0, n0, a1, n1, a2 (operator table), 0)     */
  TNX_OP_QR = 3,     /* tnx_qr_f64(a0, n0)                                  */
  TNX_OP_SVD = 4,    /* tnx_svd_jacobi_f64(a0, n0, mode = max sweeps)       */
  TNX_OP_CHOL = 5,   /* tnx_chol_inv_f64(a0, n0)                            */
  TNX_OP_SVDJ = 6,   /* tnx_svd_block_jacobi(a0, n0, n1, n2, n3, a1 (flag),
                        mode = max sweeps)                                  */
  TNX_OP_NORM = 7,   /* tnx_normalize_f64(a0, n0)                           */
};
typedef struct tnx_op {
  int32_t kind, mode;
  int32_t n0, n1, n2, n3;
  uint64_t a0, a1, a2;
} tnx_op; /* 48 bytes */
int tnx_run_ops(const tnx_op* ops /* host */, int nops, tnx_stream_t stream);
/* No counterpart in the reference (its eigensolve returns before the next
 * step's tensor moves start, MPSClass.py:883-889 then TBM:416-526): as
 * tnx_lanczos_run_multi, with the recorded launches `next` (host array,
 * nnext entries, may be 0) queued on the stream behind the Lanczos kernels
 * and result copies.  The call returns once the results are on the host;
 * the queued launches may still be running (later work on the stream
 * follows them).  *next_rc (may be null) receives tnx_run_ops's code:
 * TNX_ENOCONV there is a block Jacobi SVD of `next` at its sweep limit; a
 * launch failure is also the call's return code.  A host that books one
 * step's eigensolves while the device already runs the next step's gauge
 * move and environment update uses this.                                   */
int tnx_lanczos_run_multi_then(const tnx_matvec_plan* plans, int nplan,
                               const tnx_lanczos_solve* solves, int nsolve, int nchunks, int ncv,
                               int max_restarts, int min_iters, void* state,
                               tnx_lanczos_result* results, const tnx_op* next, int nnext,
                               int32_t* next_rc, tnx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TNALG_HIP_H */
