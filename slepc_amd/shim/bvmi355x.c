/*
 * BV type "mi355x": registers libslepc_amd.so as a SLEPc basis-vectors
 * implementation (BVRegister, src/sys/classes/bv/interface/bvfunc.c:646-652),
 * modelled on BVCreate_Mat (src/sys/classes/bv/impls/mat/bvmat.c:385-524).
 * Compiles against an installed SLEPc/PETSc built with HIP; it is not built in
 * this repository (neither library is available here).  The call sequences it
 * issues are exercised by tests/harness/bv_sequence.c against the C ABI.
 *
 * Columns: SLEPc column j (counted after the nc constraint columns) is
 * library column nc + j; SLEPc's active range [l, k) is [nc + l, nc + k)
 * (l = -nc while Gram-Schmidt runs against the constraints too).
 *
 * Ops table (include/slepc/private/bvimpl.h:25-61):
 *   mult         -> sa_bv_mult             (bvmat.c:17-36)
 *   multvec      -> sa_bv_multvec          (bvmat.c:38-53)
 *   multinplace  -> sa_bv_multinplace      (bvmat.c:55-71)
 *   dot          -> sa_bv_dot              (bvmat.c:89-108)
 *   dotvec       -> sa_bv_dotvec           (bvmat.c:110-130)
 *   scale        -> sa_bv_scale_column     (bvmat.c:151-163)
 *   norm         -> sa_bv_norm_column / sa_bv_norm (bvmat.c:165-176; NORM_2, NORM_FROBENIUS)
 *   matmult      -> sa_bv_matmult          (bvmat.c:204-232; A of type aijmi355x)
 *   copy         -> sa_bv_copy             (bvmat.c:234-248)
 *   copycolumn   -> sa_bv_copy_column      (bvmat.c:250-260)
 *   resize       -> sa_bv_resize           (bvmat.c:262-284)
 *   getcolumn / restorecolumn -> VECHIP views placed on the device column (bvmat.c:286-310)
 *   getarray* / restorearray* -> host copies of the device array (bvmat.c:312-346)
 *   view         -> a MATSEQDENSE over the host copy (bvmat.c:348-371)
 *   gramschmidt  -> sa_bv_gs_pass: one CGS1 pass on the device per call of
 *                   BVOrthogonalizeGS's DGKS loop (bvorthog.c:134,145-217)
 *   destroy
 * The coefficient-buffer convention of q == NULL (bvmat.c:47,124): the
 * coefficients start at the beginning of bv->buffer.
 */
#include <slepc/private/bvimpl.h>
#include <petsc/private/vecimpl.h>

#include "../../include/slepc_amd.h"

typedef struct {
  sa_bv        V;
  PetscScalar *host;       /* host copy handed out by getarray*, or NULL */
  PetscBool    host_write; /* restorearray uploads it */
} BV_MI355X;

#define SACall(call) do { int _r = (call); PetscCheck(!_r, PETSC_COMM_SELF, PETSC_ERR_LIB, "slepc_amd: %s", sa_last_error()); } while (0)

/* the library handle of an aijmi355x matrix (matmi355x.c) */
PETSC_EXTERN PetscErrorCode MatMI355XGetHandle(Mat, sa_mat *);

static PetscErrorCode BVMI355XActive(BV bv)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  SACall(sa_bv_set_active_columns(ctx->V, (int)(bv->nc + bv->l), (int)(bv->nc + bv->k)));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* Q (SLEPc indices, lda ldq) re-indexed by library column (nc rows and
   columns of zeros in front); returns q itself when nc == 0 */
static PetscErrorCode BVMI355XShiftQ(PetscInt nc, const PetscScalar *q, PetscInt ldq, PetscInt ncols, PetscScalar **qs, PetscInt *ldqs)
{
  PetscFunctionBegin;
  if (!nc) {
    *qs   = (PetscScalar *)q;
    *ldqs = ldq;
    PetscFunctionReturn(PETSC_SUCCESS);
  }
  *ldqs = ldq + nc;
  PetscCall(PetscCalloc1((*ldqs) * (ncols + nc), qs));
  for (PetscInt j = 0; j < ncols; j++)
    for (PetscInt i = 0; i < ldq; i++) (*qs)[(nc + i) + (nc + j) * (*ldqs)] = q[i + j * ldq];
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMult_MI355X(BV Y, PetscScalar alpha, PetscScalar beta, BV X, Mat Q)
{
  BV_MI355X         *x = (BV_MI355X *)X->data, *y = (BV_MI355X *)Y->data;
  const PetscScalar *q;
  PetscScalar       *qs;
  PetscInt           ldq, ldqs, nq;

  PetscFunctionBegin;
  PetscCall(BVMI355XActive(X));
  PetscCall(BVMI355XActive(Y));
  if (!Q) {
    SACall(sa_bv_mult(y->V, alpha, beta, x->V, NULL, 0));
    PetscFunctionReturn(PETSC_SUCCESS);
  }
  PetscCheck(X->nc == Y->nc, PETSC_COMM_SELF, PETSC_ERR_SUP, "BVMult with different numbers of constraints");
  PetscCall(MatDenseGetLDA(Q, &ldq));
  PetscCall(MatGetSize(Q, NULL, &nq));
  PetscCall(MatDenseGetArrayRead(Q, &q));
  PetscCall(BVMI355XShiftQ(X->nc, q, ldq, nq, &qs, &ldqs));
  SACall(sa_bv_mult(y->V, alpha, beta, x->V, qs, (int)ldqs));
  if (qs != q) PetscCall(PetscFree(qs));
  PetscCall(MatDenseRestoreArrayRead(Q, &q));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMultVec_MI355X(BV X, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar *q)
{
  BV_MI355X   *ctx = (BV_MI355X *)X->data;
  PetscScalar *py, *b;

  PetscFunctionBegin;
  PetscCall(BVMI355XActive(X));
  PetscCall(VecHIPGetArray(y, &py));
  if (!q) {   /* coefficients at the start of the buffer (bvmat.c:47) */
    PetscCall(VecGetArray(X->buffer, &b));
    SACall(sa_bv_multvec(ctx->V, alpha, beta, py, b));
    PetscCall(VecRestoreArray(X->buffer, &b));
  } else SACall(sa_bv_multvec(ctx->V, alpha, beta, py, q));
  PetscCall(VecHIPRestoreArray(y, &py));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMultInPlace_MI355X(BV V, Mat Q, PetscInt s, PetscInt e)
{
  BV_MI355X         *ctx = (BV_MI355X *)V->data;
  const PetscScalar *q;
  PetscScalar       *qs;
  PetscInt           ldq, ldqs, nq;

  PetscFunctionBegin;
  if (s >= e || !V->n) PetscFunctionReturn(PETSC_SUCCESS);
  PetscCall(MatDenseGetLDA(Q, &ldq));
  PetscCall(MatGetSize(Q, NULL, &nq));
  PetscCall(MatDenseGetArrayRead(Q, &q));
  PetscCall(BVMI355XShiftQ(V->nc, q, ldq, nq, &qs, &ldqs));
  PetscCall(BVMI355XActive(V));
  SACall(sa_bv_multinplace(ctx->V, qs, (int)ldqs, (int)(V->nc + s), (int)(V->nc + e)));
  if (qs != q) PetscCall(PetscFree(qs));
  PetscCall(MatDenseRestoreArrayRead(Q, &q));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVDot_MI355X(BV X, BV Y, Mat M)
{
  BV_MI355X   *x = (BV_MI355X *)X->data, *y = (BV_MI355X *)Y->data;
  PetscScalar *m, *tmp;
  PetscInt     ldm, ldt = Y->nc + Y->k;

  PetscFunctionBegin;
  PetscCall(BVMI355XActive(X));
  PetscCall(BVMI355XActive(Y));
  PetscCall(MatDenseGetLDA(M, &ldm));
  PetscCall(MatDenseGetArray(M, &m));
  /* the library writes M(Y.l+i, X.l+j) in its own column indices */
  PetscCall(PetscCalloc1(ldt * (X->nc + X->k), &tmp));
  SACall(sa_bv_dot(x->V, y->V, tmp, (int)ldt));
  for (PetscInt j = X->l; j < X->k; j++)
    for (PetscInt i = Y->l; i < Y->k; i++) m[i + j * ldm] = tmp[(Y->nc + i) + (X->nc + j) * ldt];
  PetscCall(PetscFree(tmp));
  PetscCall(MatDenseRestoreArray(M, &m));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVDotVec_MI355X(BV X, Vec y, PetscScalar *q)
{
  BV_MI355X         *ctx = (BV_MI355X *)X->data;
  const PetscScalar *py;
  PetscScalar       *b;

  PetscFunctionBegin;
  PetscCheck(!X->matrix, PETSC_COMM_SELF, PETSC_ERR_SUP, "BV type mi355x has no non-standard inner product");
  PetscCall(BVMI355XActive(X));
  PetscCall(VecHIPGetArrayRead(y, &py));
  if (!q) {   /* results at the start of the buffer (bvmat.c:124) */
    PetscCall(VecGetArray(X->buffer, &b));
    SACall(sa_bv_dotvec(ctx->V, py, b));
    PetscCall(VecRestoreArray(X->buffer, &b));
  } else SACall(sa_bv_dotvec(ctx->V, py, q));
  PetscCall(VecHIPRestoreArrayRead(y, &py));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVScale_MI355X(BV bv, PetscInt j, PetscScalar alpha)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  PetscCall(BVMI355XActive(bv));
  SACall(sa_bv_scale_column(ctx->V, (int)(j < 0 ? -1 : bv->nc + j), alpha));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVNorm_MI355X(BV bv, PetscInt j, NormType type, PetscReal *nrm)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;
  double     v;

  PetscFunctionBegin;
  PetscCheck(type == NORM_2 || type == NORM_FROBENIUS, PETSC_COMM_SELF, PETSC_ERR_SUP, "BV type mi355x computes 2-norms and Frobenius norms");
  PetscCheck(j >= 0 || type == NORM_FROBENIUS, PETSC_COMM_SELF, PETSC_ERR_SUP, "The 2-norm of a block of columns is not available");
  PetscCall(BVMI355XActive(bv));
  if (j >= 0) SACall(sa_bv_norm_column(ctx->V, (int)(bv->nc + j), &v));
  else SACall(sa_bv_norm(ctx->V, &v));
  *nrm = (PetscReal)v;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMatMult_MI355X(BV V, Mat A, BV W)
{
  BV_MI355X *v = (BV_MI355X *)V->data, *w = (BV_MI355X *)W->data;
  sa_mat     M;

  PetscFunctionBegin;
  PetscCall(MatMI355XGetHandle(A, &M));
  PetscCall(BVMI355XActive(V));
  PetscCall(BVMI355XActive(W));
  SACall(sa_bv_matmult(v->V, M, w->V));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVCopy_MI355X(BV V, BV W)
{
  BV_MI355X *v = (BV_MI355X *)V->data, *w = (BV_MI355X *)W->data;

  PetscFunctionBegin;
  PetscCall(BVMI355XActive(V));
  PetscCall(BVMI355XActive(W));
  SACall(sa_bv_copy(v->V, w->V));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVCopyColumn_MI355X(BV V, PetscInt j, PetscInt i)
{
  BV_MI355X *ctx = (BV_MI355X *)V->data;

  PetscFunctionBegin;
  SACall(sa_bv_copy_column(ctx->V, (int)(V->nc + j), (int)(V->nc + i)));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVResize_MI355X(BV bv, PetscInt m, PetscBool copy)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  SACall(sa_bv_resize(ctx->V, (int)(bv->nc + m), copy ? 1 : 0));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVGetColumn_MI355X(BV bv, PetscInt j, Vec *v)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;
  void      *base;
  PetscInt   l;

  PetscFunctionBegin;
  l = BVAvailableVec;
  SACall(sa_bv_get_array(ctx->V, &base));
  PetscCall(VecHIPPlaceArray(bv->cv[l], (PetscScalar *)base + (bv->nc + j) * bv->ld));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVRestoreColumn_MI355X(BV bv, PetscInt j, Vec *v)
{
  PetscInt l;

  PetscFunctionBegin;
  l = (j == bv->ci[0]) ? 0 : 1;
  PetscCall(VecHIPResetArray(bv->cv[l]));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* Host copies of the whole array (ld x (nc+m)) for the callers of
   BVGetArray/BVGetArrayRead, which address it on the host */
static PetscErrorCode BVMI355XDownload(BV bv, PetscBool write)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;
  void      *base;

  PetscFunctionBegin;
  PetscCheck(!ctx->host, PETSC_COMM_SELF, PETSC_ERR_ORDER, "BVGetArray called twice without BVRestoreArray");
  PetscCall(PetscMalloc1(bv->ld * (bv->nc + bv->m), &ctx->host));
  SACall(sa_bv_get_array(ctx->V, &base));
  SACall(sa_memcpy_d2h(ctx->host, base, (int64_t)sizeof(PetscScalar) * bv->ld * (bv->nc + bv->m)));
  ctx->host_write = write;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMI355XRelease(BV bv)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;
  void      *base;

  PetscFunctionBegin;
  if (ctx->host_write) {
    SACall(sa_bv_get_array(ctx->V, &base));
    SACall(sa_memcpy_h2d(base, ctx->host, (int64_t)sizeof(PetscScalar) * bv->ld * (bv->nc + bv->m)));
  }
  PetscCall(PetscFree(ctx->host));
  ctx->host_write = PETSC_FALSE;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVGetArray_MI355X(BV bv, PetscScalar **a)
{
  PetscFunctionBegin;
  PetscCall(BVMI355XDownload(bv, PETSC_TRUE));
  *a = ((BV_MI355X *)bv->data)->host;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVRestoreArray_MI355X(BV bv, PetscScalar **a)
{
  PetscFunctionBegin;
  PetscCall(BVMI355XRelease(bv));
  if (a) *a = NULL;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVGetArrayRead_MI355X(BV bv, const PetscScalar **a)
{
  PetscFunctionBegin;
  PetscCall(BVMI355XDownload(bv, PETSC_FALSE));
  *a = ((BV_MI355X *)bv->data)->host;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVRestoreArrayRead_MI355X(BV bv, const PetscScalar **a)
{
  PetscFunctionBegin;
  PetscCall(BVMI355XRelease(bv));
  if (a) *a = NULL;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVView_MI355X(BV bv, PetscViewer viewer)
{
  const PetscScalar *a;
  Mat                M;
  PetscBool          isascii;
  PetscViewerFormat  format;

  PetscFunctionBegin;
  PetscCall(PetscObjectTypeCompare((PetscObject)viewer, PETSCVIEWERASCII, &isascii));
  if (isascii) {
    PetscCall(PetscViewerGetFormat(viewer, &format));
    if (format == PETSC_VIEWER_ASCII_INFO || format == PETSC_VIEWER_ASCII_INFO_DETAIL) PetscFunctionReturn(PETSC_SUCCESS);
  }
  PetscCall(BVGetArrayRead_MI355X(bv, &a));
  PetscCall(MatCreateSeqDense(PETSC_COMM_SELF, bv->n, bv->nc + bv->m, (PetscScalar *)a, &M));
  PetscCall(MatDenseSetLDA(M, bv->ld));
  PetscCall(PetscObjectSetName((PetscObject)M, ((PetscObject)bv)->name));
  PetscCall(MatView(M, viewer));
  PetscCall(MatDestroy(&M));
  PetscCall(BVRestoreArrayRead_MI355X(bv, &a));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* One CGS1 pass (bvorthog.c:91-132) for BVOrthogonalizeGS's loop.  A column
   (v == NULL) runs on the device with the buffer convention of
   BV_CleanCoefficients/BV_AddCoefficients (h == NULL: column j of the buffer,
   c: its start); an external vector uses dotvec/multvec of the active
   columns. */
static PetscErrorCode BVGramSchmidt_MI355X(BV bv, PetscInt j, Vec v, PetscBool *which, PetscScalar *h, PetscScalar *c, PetscReal *onorm, PetscReal *norm)
{
  BV_MI355X   *ctx = (BV_MI355X *)bv->data;
  PetscScalar *buf = NULL, *hh = h, *cc = c, *pv;
  double       on = 0.0, nr = 0.0;

  PetscFunctionBegin;
  PetscCheck(!bv->indef && !bv->matrix, PETSC_COMM_SELF, PETSC_ERR_SUP, "BV type mi355x orthogonalizes with the standard inner product");
  PetscCheck(bv->orthog_type == BV_ORTHOG_CGS, PETSC_COMM_SELF, PETSC_ERR_SUP, "The gramschmidt slot of BV type mi355x is classical Gram-Schmidt");
  if (!v) {
    bv->k = j;
    PetscCall(BVMI355XActive(bv));
    if (!h) {
      PetscCall(VecGetArray(bv->buffer, &buf));
      hh = buf + j * (bv->nc + bv->m);
      cc = buf;
    }
    SACall(sa_bv_gs_pass(ctx->V, (int)(bv->nc + j), hh, cc, &on, &nr));
    if (buf) PetscCall(VecRestoreArray(bv->buffer, &buf));
  } else {
    PetscReal vn;
    PetscScalar s = 0.0;
    PetscCall(BVMI355XActive(bv));
    PetscCall(VecNorm(v, NORM_2, &vn));
    PetscCall(VecHIPGetArray(v, &pv));
    SACall(sa_bv_dotvec(ctx->V, pv, c));
    SACall(sa_bv_multvec(ctx->V, -1.0, 1.0, pv, c));
    PetscCall(VecHIPRestoreArray(v, &pv));
    for (PetscInt i = 0; i < bv->nc + bv->k; i++) {
      s += c[i] * c[i];
      h[i] += c[i];
    }
    on = vn;
    nr = vn * vn - PetscRealPart(s);
    if (nr <= 0.0) {
      PetscCall(VecNorm(v, NORM_2, &vn));
      nr = vn;
    } else nr = PetscSqrtReal(nr);
  }
  if (onorm) *onorm = (PetscReal)on;
  if (norm) *norm = (PetscReal)nr;
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVDestroy_MI355X(BV bv)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  if (ctx->V) SACall(sa_bv_destroy(ctx->V));
  PetscCall(PetscFree(ctx->host));
  PetscCall(VecDestroy(&bv->cv[0]));
  PetscCall(VecDestroy(&bv->cv[1]));
  PetscCall(PetscFree(bv->data));
  PetscFunctionReturn(PETSC_SUCCESS);
}

SLEPC_EXTERN PetscErrorCode BVCreate_MI355X(BV bv)
{
  BV_MI355X *ctx;
  PetscInt   nloc;
  int64_t    ld;

  PetscFunctionBegin;
  PetscCall(PetscNew(&ctx));
  bv->data = (void *)ctx;
  PetscCall(PetscLayoutGetLocalSize(bv->map, &nloc));
  PetscCheck(nloc == bv->N, PETSC_COMM_SELF, PETSC_ERR_SUP, "BV type mi355x runs on one process");
  SACall(sa_init(0));
  SACall(sa_bv_create(nloc, (int)(bv->m + bv->nc), sizeof(PetscScalar) == 8 ? SA_FP64 : SA_FP32, &ctx->V));
  /* the leading dimension is the library's (a multiple of 64 rows, for
     512-byte column segments); BV_SetDefaultLD (bvimpl.h:471-484) would
     otherwise pick one */
  SACall(sa_bv_get_info(ctx->V, NULL, NULL, &ld, NULL));
  PetscCheck(!bv->ld || bv->ld == (PetscInt)ld, PETSC_COMM_SELF, PETSC_ERR_SUP, "BV type mi355x uses leading dimension %" PetscInt_FMT, (PetscInt)ld);
  bv->ld = (PetscInt)ld;
  PetscCall(VecCreateSeqHIPWithArray(PETSC_COMM_SELF, 1, nloc, NULL, &bv->cv[0]));
  PetscCall(VecCreateSeqHIPWithArray(PETSC_COMM_SELF, 1, nloc, NULL, &bv->cv[1]));
  bv->ops->mult             = BVMult_MI355X;
  bv->ops->multvec          = BVMultVec_MI355X;
  bv->ops->multinplace      = BVMultInPlace_MI355X;
  bv->ops->dot              = BVDot_MI355X;
  bv->ops->dotvec           = BVDotVec_MI355X;
  bv->ops->scale            = BVScale_MI355X;
  bv->ops->norm             = BVNorm_MI355X;
  bv->ops->matmult          = BVMatMult_MI355X;
  bv->ops->copy             = BVCopy_MI355X;
  bv->ops->copycolumn       = BVCopyColumn_MI355X;
  bv->ops->resize           = BVResize_MI355X;
  bv->ops->getcolumn        = BVGetColumn_MI355X;
  bv->ops->restorecolumn    = BVRestoreColumn_MI355X;
  bv->ops->getarray         = BVGetArray_MI355X;
  bv->ops->restorearray     = BVRestoreArray_MI355X;
  bv->ops->getarrayread     = BVGetArrayRead_MI355X;
  bv->ops->restorearrayread = BVRestoreArrayRead_MI355X;
  bv->ops->view             = BVView_MI355X;
  bv->ops->gramschmidt      = BVGramSchmidt_MI355X;
  bv->ops->destroy          = BVDestroy_MI355X;
  PetscFunctionReturn(PETSC_SUCCESS);
}
