// Recorded launch sequences: a list of the library's own launches (grouped
// GEMMs, mixer, QR, SVD, Cholesky, block Jacobi SVD, normalisation) issued
// in order by one call, so that a host which replays the same step many
// times crosses into the library once per step instead of once per launch.
#include "tnx_common.h"

extern "C" int tnx_run_ops(const tnx_op* ops, int nops, tnx_stream_t stream) {
  if (nops == 0) return TNX_OK;
  if (ops == nullptr || nops < 0) {
    tnx_set_error("tnx_run_ops: bad arguments (nops=%d)", nops);
    return TNX_EINVAL;
  }
  int result = TNX_OK;
  for (int i = 0; i < nops; ++i) {
    const tnx_op& o = ops[i];
    int rc;
    switch (o.kind) {
      case TNX_OP_GEMM:
        rc = tnx_gemm_grouped(o.mode, (const tnx_gemm_problem*)o.a0, o.n0,
                              (const tnx_gemm_seg*)o.a1, o.n1, nullptr, stream);
        break;
      case TNX_OP_MIX:
        rc = tnx_mix((const tnx_mix_job*)o.a0, o.n0, (const tnx_mix_term*)o.a1, o.n1,
                     (const double*)o.a2, nullptr, stream);
        break;
      case TNX_OP_QR:
        rc = tnx_qr_f64((const tnx_qr_job*)o.a0, o.n0, stream);
        break;
      case TNX_OP_SVD:
        rc = tnx_svd_jacobi_f64((const tnx_svd_job*)o.a0, o.n0, o.mode, stream);
        break;
      case TNX_OP_CHOL:
        rc = tnx_chol_inv_f64((const tnx_chol_job*)o.a0, o.n0, stream);
        break;
      case TNX_OP_SVDJ:
        rc = tnx_svd_block_jacobi((const tnx_svdj_job*)o.a0, o.n0, o.n1, o.n2, o.n3,
                                  (int32_t*)o.a1, o.mode, stream);
        if (rc == TNX_ENOCONV) {   // recorded; the sequence goes on
          result = TNX_ENOCONV;
          rc = TNX_OK;
        }
        break;
      case TNX_OP_NORM:
        rc = tnx_normalize_f64((const tnx_norm_job*)o.a0, o.n0, stream);
        break;
      default:
        tnx_set_error("tnx_run_ops: unknown kind %d at op %d", o.kind, i);
        return TNX_EINVAL;
    }
    if (rc != TNX_OK) return rc;   // the failing call set the message
  }
  return result;
}
