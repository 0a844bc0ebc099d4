/*
 * Mat type "aijmi355x": a MATSEQAIJ whose MatMult runs the HIP SpMV of
 * libslepc_amd.so (MatRegister, PETSc).  After MatAssemblyEnd the CSR arrays
 * (ascending columns, as AIJ keeps them) are uploaded once; MatMult then
 * takes device vectors (VECHIP), which is what the BV type "mi355x" hands out
 * as column views, so BVMatMultColumn (bvops.c:862-885) and
 * MatMult_STOperator (stsolve.c:244-259) run without host copies.
 * Compiles against an installed PETSc with HIP; not built in this repository.
 */
#include <petsc/private/matimpl.h>
#include <../src/mat/impls/aij/seq/aij.h>

#include "../../include/slepc_amd.h"

typedef struct {
  sa_mat          A;
  PetscErrorCode (*assemblyend_aij)(Mat, MatAssemblyType);
  PetscErrorCode (*destroy_aij)(Mat);
} Mat_MI355X;

#define SACall(call) do { int _r = (call); PetscCheck(!_r, PETSC_COMM_SELF, PETSC_ERR_LIB, "slepc_amd: %s", sa_last_error()); } while (0)

static PetscErrorCode MatAssemblyEnd_AIJMI355X(Mat M, MatAssemblyType t)
{
  Mat_MI355X *ctx = (Mat_MI355X *)M->spptr;
  Mat_SeqAIJ *a   = (Mat_SeqAIJ *)M->data;

  PetscFunctionBegin;
  PetscCall((*ctx->assemblyend_aij)(M, t));
  if (t == MAT_FLUSH_ASSEMBLY) PetscFunctionReturn(PETSC_SUCCESS);
  if (ctx->A) SACall(sa_mat_destroy(ctx->A));
  SACall(sa_mat_create_csr((int)M->rmap->n, (int64_t)a->nz, (const int32_t *)a->i, (const int32_t *)a->j, a->a,
                           SA_FP64, &ctx->A));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode MatMult_AIJMI355X(Mat M, Vec x, Vec y)
{
  Mat_MI355X        *ctx = (Mat_MI355X *)M->spptr;
  const PetscScalar *px;
  PetscScalar       *py;

  PetscFunctionBegin;
  PetscCall(VecHIPGetArrayRead(x, &px));
  PetscCall(VecHIPGetArrayWrite(y, &py));
  SACall(sa_mat_mult(ctx->A, px, py));
  PetscCall(VecHIPRestoreArrayRead(x, &px));
  PetscCall(VecHIPRestoreArrayWrite(y, &py));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode MatDestroy_AIJMI355X(Mat M)
{
  Mat_MI355X *ctx = (Mat_MI355X *)M->spptr;

  PetscFunctionBegin;
  if (ctx->A) SACall(sa_mat_destroy(ctx->A));
  M->ops->destroy = ctx->destroy_aij;
  PetscCall(PetscFree(M->spptr));
  PetscCall((*M->ops->destroy)(M));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The library handle behind an aijmi355x matrix (used by BV type mi355x for
   BVMatMult); an error for any other type. */
PETSC_EXTERN PetscErrorCode MatMI355XGetHandle(Mat M, sa_mat *A)
{
  PetscBool match;

  PetscFunctionBegin;
  PetscCall(PetscObjectTypeCompare((PetscObject)M, "aijmi355x", &match));
  PetscCheck(match, PETSC_COMM_SELF, PETSC_ERR_SUP, "BVMatMult of BV type mi355x needs a matrix of type aijmi355x");
  *A = ((Mat_MI355X *)M->spptr)->A;
  PetscCheck(*A, PETSC_COMM_SELF, PETSC_ERR_ORDER, "The aijmi355x matrix is not assembled");
  PetscFunctionReturn(PETSC_SUCCESS);
}

PETSC_EXTERN PetscErrorCode MatCreate_AIJMI355X(Mat M)
{
  Mat_MI355X *ctx;

  PetscFunctionBegin;
  PetscCheck(sizeof(PetscInt) == 4, PETSC_COMM_SELF, PETSC_ERR_SUP, "aijmi355x needs 32-bit PetscInt (int32 CSR indices)");
  PetscCall(MatSetType(M, MATSEQAIJ));
  PetscCall(PetscNew(&ctx));
  M->spptr             = ctx;
  SACall(sa_init(0));
  ctx->assemblyend_aij = M->ops->assemblyend;
  ctx->destroy_aij     = M->ops->destroy;
  M->ops->assemblyend  = MatAssemblyEnd_AIJMI355X;
  M->ops->mult         = MatMult_AIJMI355X;
  M->ops->destroy      = MatDestroy_AIJMI355X;
  PetscCall(PetscFree(M->defaultvectype));
  PetscCall(PetscStrallocpy(VECHIP, &M->defaultvectype));
  PetscCall(PetscObjectChangeTypeName((PetscObject)M, "aijmi355x"));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* Call once after SlepcInitialize: makes -mat_type aijmi355x and -bv_type mi355x available. */
SLEPC_EXTERN PetscErrorCode BVCreate_MI355X(BV);
PETSC_EXTERN PetscErrorCode SlepcAmdRegisterAll(void)
{
  PetscFunctionBegin;
  PetscCall(MatRegister("aijmi355x", MatCreate_AIJMI355X));
  PetscCall(BVRegister("mi355x", BVCreate_MI355X));
  PetscFunctionReturn(PETSC_SUCCESS);
}
