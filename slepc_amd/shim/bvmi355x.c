/*
 * BV type "mi355x": registers libslepc_amd.so as a SLEPc basis-vectors
 * implementation (BVRegister, src/sys/classes/bv/interface/bvfunc.c:646-652).
 * Compiles against an installed SLEPc/PETSc built with HIP; it is not built in
 * this repository (neither library is available here).
 *
 * Mapping of the ops table (include/slepc/private/bvimpl.h:25-61):
 *   multvec      -> sa_bv_multvec        (bvmat.c:38-53)
 *   dotvec       -> sa_bv_dotvec         (bvmat.c:110-130)
 *   multinplace  -> sa_bv_multinplace    (bvmat.c:55-71)
 *   scale        -> sa_bv_scale_column   (bvmat.c:151-163)
 *   norm         -> sa_bv_norm_column / sa_bv_norm
 *   copycolumn   -> sa_bv_copy_column    (bvmat.c:250-260)
 *   matmult col  -> sa_bv_mat_mult_column when the Mat is of type aijmi355x
 *   getcolumn / restorecolumn -> VECHIP views placed on the device column
 */
#include <slepc/private/bvimpl.h>
#include <petsc/private/vecimpl.h>

#include "../../include/slepc_amd.h"

typedef struct {
  sa_bv V;
  Vec   view[2];
} BV_MI355X;

#define SACall(call) do { int _r = (call); PetscCheck(!_r, PETSC_COMM_SELF, PETSC_ERR_LIB, "slepc_amd: %s", sa_last_error()); } while (0)

static PetscErrorCode BVMultVec_MI355X(BV X, PetscScalar alpha, PetscScalar beta, Vec y, PetscScalar *q)
{
  BV_MI355X   *ctx = (BV_MI355X *)X->data;
  PetscScalar *py;

  PetscFunctionBegin;
  SACall(sa_bv_set_active_columns(ctx->V, X->l, X->k));
  PetscCall(VecHIPGetArray(y, &py));
  if (!q) {   /* coefficients of the internal buffer (column 0): copy them down */
    const PetscScalar *b;
    PetscCall(VecGetArrayRead(X->buffer, &b));
    SACall(sa_bv_multvec(ctx->V, alpha, beta, py, b + X->nc + X->l));
    PetscCall(VecRestoreArrayRead(X->buffer, &b));
  } else SACall(sa_bv_multvec(ctx->V, alpha, beta, py, q));
  PetscCall(VecHIPRestoreArray(y, &py));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVDotVec_MI355X(BV X, Vec y, PetscScalar *q)
{
  BV_MI355X         *ctx = (BV_MI355X *)X->data;
  const PetscScalar *py;
  PetscScalar       *b;

  PetscFunctionBegin;
  SACall(sa_bv_set_active_columns(ctx->V, X->l, X->k));
  PetscCall(VecHIPGetArrayRead(y, &py));
  if (!q) {
    PetscCall(VecGetArray(X->buffer, &b));
    SACall(sa_bv_dotvec(ctx->V, py, b + X->nc + X->l));
    PetscCall(VecRestoreArray(X->buffer, &b));
  } else SACall(sa_bv_dotvec(ctx->V, py, q));
  PetscCall(VecHIPRestoreArrayRead(y, &py));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVMultInPlace_MI355X(BV V, Mat Q, PetscInt s, PetscInt e)
{
  BV_MI355X         *ctx = (BV_MI355X *)V->data;
  const PetscScalar *q;
  PetscInt          ldq;

  PetscFunctionBegin;
  PetscCall(MatDenseGetLDA(Q, &ldq));
  PetscCall(MatDenseGetArrayRead(Q, &q));
  SACall(sa_bv_set_active_columns(ctx->V, V->l, V->k));
  SACall(sa_bv_multinplace(ctx->V, q, (int)ldq, (int)s, (int)e));
  PetscCall(MatDenseRestoreArrayRead(Q, &q));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVScale_MI355X(BV bv, PetscInt j, PetscScalar alpha)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  SACall(sa_bv_set_active_columns(ctx->V, bv->l, bv->k));
  SACall(sa_bv_scale_column(ctx->V, (int)(j < 0 ? -1 : bv->nc + j), alpha));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVNorm_MI355X(BV bv, PetscInt j, NormType type, PetscReal *nrm)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  PetscCheck(type == NORM_2 || type == NORM_FROBENIUS, PETSC_COMM_SELF, PETSC_ERR_SUP, "Only 2-norm and Frobenius norm");
  SACall(sa_bv_set_active_columns(ctx->V, bv->l, bv->k));
  if (j >= 0) SACall(sa_bv_norm_column(ctx->V, (int)(bv->nc + j), nrm));
  else SACall(sa_bv_norm(ctx->V, nrm));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVCopyColumn_MI355X(BV V, PetscInt j, PetscInt i)
{
  BV_MI355X *ctx = (BV_MI355X *)V->data;

  PetscFunctionBegin;
  SACall(sa_bv_copy_column(ctx->V, (int)(V->nc + j), (int)(V->nc + i)));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode BVGetColumn_MI355X(BV bv, PetscInt j, Vec *v)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;
  void      *base;
  int64_t    ld;
  PetscInt   l;

  PetscFunctionBegin;
  l = BVAvailableVec;
  SACall(sa_bv_get_array(ctx->V, &base));
  SACall(sa_bv_get_info(ctx->V, NULL, NULL, &ld, NULL));
  PetscCall(VecHIPPlaceArray(bv->cv[l], (PetscScalar *)base + (bv->nc + j) * ld));
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

static PetscErrorCode BVDestroy_MI355X(BV bv)
{
  BV_MI355X *ctx = (BV_MI355X *)bv->data;

  PetscFunctionBegin;
  if (ctx->V) SACall(sa_bv_destroy(ctx->V));
  PetscCall(VecDestroy(&bv->cv[0]));
  PetscCall(VecDestroy(&bv->cv[1]));
  PetscCall(PetscFree(bv->data));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* BVCreate_Mat model: bvmat.c:385-524 */
SLEPC_EXTERN PetscErrorCode BVCreate_MI355X(BV bv)
{
  BV_MI355X *ctx;

  PetscFunctionBegin;
  PetscCall(PetscNew(&ctx));
  bv->data = (void *)ctx;
  SACall(sa_init(0));
  SACall(sa_bv_create(bv->n, (int)(bv->m + bv->nc), sizeof(PetscScalar) == 8 ? SA_FP64 : SA_FP32, &ctx->V));
  PetscCall(VecCreateSeqHIPWithArray(PETSC_COMM_SELF, 1, bv->n, NULL, &bv->cv[0]));
  PetscCall(VecCreateSeqHIPWithArray(PETSC_COMM_SELF, 1, bv->n, NULL, &bv->cv[1]));
  bv->ops->multvec       = BVMultVec_MI355X;
  bv->ops->dotvec        = BVDotVec_MI355X;
  bv->ops->multinplace   = BVMultInPlace_MI355X;
  bv->ops->scale         = BVScale_MI355X;
  bv->ops->norm          = BVNorm_MI355X;
  bv->ops->copycolumn    = BVCopyColumn_MI355X;
  bv->ops->getcolumn     = BVGetColumn_MI355X;
  bv->ops->restorecolumn = BVRestoreColumn_MI355X;
  bv->ops->destroy       = BVDestroy_MI355X;
  PetscFunctionReturn(PETSC_SUCCESS);
}
