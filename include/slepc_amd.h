/*
 * slepc_amd.h -- C ABI of libslepc_amd.so, an MI355X (gfx950) Krylov-Schur hot path.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * nonzero code on failure; sa_last_error() returns the message (the analogue
 * of PetscErrorCode + PetscCheck).  "_dev" pointers are device (HBM) addresses,
 * "_host" pointers are host memory.  Each host thread has its own HIP stream
 * (sa_stream() returns the calling thread's); a call orders its work on that
 * stream and returns once the results it reports are ready.  Objects may be
 * read from several threads at once (a matrix shared by concurrent solves)
 * but must not be modified from two threads at once.
 *
 * The functions map onto the SLEPc/PETSc interfaces the reference's Krylov-Schur
 * calls (paths relative to the SLEPc 3.22.2 source tree); the registration
 * shims that bind them as a BV type and a Mat type are in
 * slepc_amd/shim/ (see the integration notes).
 */
#ifndef SLEPC_AMD_H
#define SLEPC_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sa_mat_s *sa_mat;   /* CSR operator resident in HBM            */
typedef struct sa_bv_s  *sa_bv;    /* n x m column-major basis in HBM         */
typedef struct sa_eps_s *sa_eps;   /* Krylov-Schur eigensolver context        */
typedef struct sa_mlp_s *sa_mlp;   /* learned warm-start network (BF16 MFMA)  */

enum { SA_FP64 = 0, SA_FP32 = 1 };
enum { SA_LARGEST_MAGNITUDE = 1, SA_SMALLEST_MAGNITUDE = 2, SA_LARGEST_REAL = 3, SA_SMALLEST_REAL = 4,
       SA_LARGEST_IMAGINARY = 5, SA_SMALLEST_IMAGINARY = 6, SA_TARGET_MAGNITUDE = 7, SA_TARGET_REAL = 8,
       SA_WHICH_USER = 11 };
enum { SA_ERROR_ABSOLUTE = 0, SA_ERROR_RELATIVE = 1, SA_ERROR_BACKWARD = 2 };

/* ---- library ----------------------------------------------------------- */
/* SlepcInitialize analogue: selects the device and creates the stream. */
int         sa_init(int device);
const char *sa_last_error(void);
void       *sa_stream(void);                    /* hipStream_t */
int         sa_synchronize(void);
/* Device memory helpers for hosts without their own allocator. */
int sa_malloc(void **ptr_dev, int64_t bytes);
int sa_free(void *ptr_dev);
int sa_memcpy_h2d(void *dst_dev, const void *src_host, int64_t bytes);
int sa_memcpy_d2h(void *dst_host, const void *src_dev, int64_t bytes);
/* HIP events on the library stream around a timed region (milliseconds). */
int sa_timer_start(void);
int sa_timer_stop(double *ms);

/* ---- Mat: CSR (MATAIJ) ----------------------------------------------------
 * Replaces MatCreateSeqAIJWithArrays/MatAssembly + MatMult of PETSc's AIJ
 * (called at src/sys/classes/bv/interface/bvops.c:879 and
 * src/sys/classes/st/interface/stsolve.c:22).  Arrays are host CSR with
 * ascending column indices per row; prec selects the stored value type. */
int sa_mat_create_csr(int n, int64_t nnz, const int32_t *rowptr_host, const int32_t *col_host,
                      const double *val_host, int prec, sa_mat *A);
/* Storage layout threshold: operators whose values and column indices take
 * more than `bytes` are created for the stream SpMV and SpMM (row blocks of
 * at most 2,048 nonzeros staged in LDS); smaller ones keep the
 * vector-CSR kernel.  Default 64 MiB; returns the previous value; operators
 * already created keep their layout.  No counterpart in the reference (a
 * choice of storage format, as a MatType is); tests use 0 to run the stream
 * kernels on small operators. */
int64_t sa_set_spmv_stream_bytes(int64_t bytes);
/* Partition hint: the number of solves the application keeps in flight at
 * once on the device (one host thread each).  With 2 or more, bases that fit
 * the last-level cache (<= 256 MiB) created afterwards give each
 * Gram-Schmidt workgroup two 64-row chunks instead of one, so the slot
 * kernels of concurrent solves share the CUs.  Default 1; returns the
 * previous value.  No counterpart in the reference (PETSc runs one solve per
 * process); results differ from a solve alone only in the rounding of the
 * reductions. */
int sa_set_concurrent_solves(int nsolves);
int sa_mat_destroy(sa_mat A);
int sa_mat_get_info(sa_mat A, int *n, int64_t *nnz, int *prec);
/* y = A x with x, y device vectors of the matrix precision (MatMult). */
int sa_mat_mult(sa_mat A, const void *x_dev, void *y_dev);

/* Residual norms of approximate eigenpairs, EPSComputeResidualNorm_Private
 * (src/eps/interface/epssolve.c:666-718) with trans = FALSE and B = I:
 * res[i] = ||A x_i - lambda_i x_i|| for the k device columns of X (the
 * matrix precision, leading dimension ldx >= n), lambda_i = lr[i] + i li[i]
 * (li may be NULL).  When li[i] != 0 (and not below |lr[i]| eps) columns i
 * and i+1 hold the real and imaginary parts of the vector of lr[i] + i li[i]
 * and res[i] = res[i+1] = sqrt(||.||_r^2 + ||.||_i^2); |lambda| <= eps drops
 * the lambda terms.  One fused SpMV + axpy + norm kernel per 16 vectors; the
 * partial sums live in device memory the library allocates per call. */
int sa_mat_residual_norms(sa_mat A, int k, const void *X_dev, int64_t ldx, const double *lr_host,
                          const double *li_host, double *res_host);
/* The same computation without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * partial sums live in a static device array of the library (131,200 bytes
 * per GPU, held until the process ends), so at most 16 vectors (one launch)
 * per call; more is an error.  Calls that overlap on one GPU share that
 * array and race; nothing in it needs clearing between calls. */
int sa_mat_residual_norms_noalloc(sa_mat A, int k, const void *X_dev, int64_t ldx, const double *lr_host,
                                  const double *li_host, double *res_host);

/* ---- MatSetValues-style assembly and the stencil families ----------------
 * COO triplets -> CSR with ascending columns and INSERT_VALUES semantics
 * (the last value given for a duplicate (row,col) wins), as PETSc's
 * MatSetValues + MatAssemblyEnd produce.  Returns nnz; fills the outputs when
 * non-NULL. */
int64_t sa_csr_assemble(int n, int64_t ntrip, const int32_t *rows, const int32_t *cols,
                        const double *vals, int32_t *rowptr, int32_t *col, double *val);
/* Generators used by the workloads; each returns nnz and fills when non-NULL. */
int64_t sa_gen_lap2d(int nx, int ny, double pot_scale, uint64_t seed,
                     int32_t *rowptr, int32_t *col, double *val);
int64_t sa_gen_convdiff3d(int nx, int ny, int nz, int32_t *rowptr, int32_t *col, double *val);
int64_t sa_gen_stencil27(int nx, int ny, int nz, uint64_t seed,
                         int32_t *rowptr, int32_t *col, double *val);
int64_t sa_gen_fv2d(int nx, int ny, const double *theta8, double *sigma,
                    int32_t *rowptr, int32_t *col, double *val);
int64_t sa_gen_markov(int m, int32_t *rowptr, int32_t *col, double *val);
void    sa_fv2d_theta(uint64_t seed, int sample, double *theta8);
void    sa_rand_uniform(uint64_t seed, int64_t n, double *out_host); /* U[-1,1) */
double  sa_u01(uint64_t seed, uint64_t i);
uint64_t sa_sub_seed(uint64_t seed, uint64_t p);

/* ---- file formats ---------------------------------------------------------
 * MatLoad of a PETSc binary AIJ file (as src/eps/tutorials/ex4.c loads the
 * .petsc files of share/slepc/datafiles/matrices): big-endian int32 header
 * {1211216, M, N, nz}, M row lengths, nz column indices, nz float64 values.
 * Called with rowptr = col = val = NULL it returns the sizes only. */
int sa_mat_load_petsc_binary(const char *path, int *nrows, int *ncols, int64_t *nnz,
                             int32_t *rowptr, int32_t *col, double *val);
/* VecView / VecLoad with a binary viewer: {1211214, n} then n float64 values;
   append != 0 writes after the vectors already in the file (EPSVectorsView). */
int sa_vec_write_petsc_binary(const char *path, int64_t n, const double *x, int append);
int sa_vec_load_petsc_binary(const char *path, int64_t index, int64_t *n, double *x);
/* Text of EPSValuesView (src/eps/interface/epsview.c:602-640): format 0 is the
   default ASCII viewer ("%.5f" per value), 1 is PETSC_VIEWER_ASCII_MATLAB
   ("Lambda_<name> = [" with "%18.16e").  Returns the text length and writes at
   most len-1 characters plus a zero into buf when buf != NULL; -1 on error. */
int64_t sa_format_eigenvalues(int n, const double *re, const double *im, int format, const char *name,
                              char *buf, int64_t len);

/* ---- BV: basis vectors ------------------------------------------------------
 * Replaces BVCreate_Mat storage (src/sys/classes/bv/impls/mat/bvmat.c:385-524)
 * and the ops table slots of struct _BVOps (include/slepc/private/bvimpl.h:25-61).
 * Storage: m columns of n rows, leading dimension *ld (n rounded up to 64). */
int sa_bv_create(int64_t n, int m, int prec, sa_bv *V);
int sa_bv_destroy(sa_bv V);
int sa_bv_get_info(sa_bv V, int64_t *n, int *m, int64_t *ld, int *prec);
int sa_bv_get_array(sa_bv V, void **data_dev);                 /* BVGetArray   */
int sa_bv_set_active_columns(sa_bv V, int l, int k);           /* BVSetActiveColumns bvbasic.c */
int sa_bv_get_active_columns(sa_bv V, int *l, int *k);
int sa_bv_set_column(sa_bv V, int j, const double *x_host);    /* BVInsertVec without orth. */
int sa_bv_get_column(sa_bv V, int j, double *x_host);          /* BVCopyVec to host        */
/* ops->dotvec (bvmat.c:110-130): q = V(:,l:k)^T x; q_host==NULL stores into
   column 0 of the coefficient buffer (the q==NULL convention, bvimpl.h:289-415). */
int sa_bv_dotvec(sa_bv V, const void *x_dev, double *q_host);
/* ops->multvec (bvmat.c:38-53): y = beta*y + alpha*V(:,l:k) q */
int sa_bv_multvec(sa_bv V, double alpha, double beta, void *y_dev, const double *q_host);
/* ops->multinplace (bvmat.c:55-71 -> bvblas.c:74-106): V(:,s:e) = V(:,l:k) Q(l:k,s:e),
   Q host column-major with leading dimension ldq (MATSEQDENSE); MFMA GEMM. */
int sa_bv_multinplace(sa_bv V, const double *Q_host, int ldq, int s, int e);
/* ops->scale (bvmat.c:151-163): column j (j<0: all active columns) times alpha */
int sa_bv_scale_column(sa_bv V, int j, double alpha);
/* BVNormColumn NORM_2 (bvglobal.c:662-685) and BVNorm NORM_FROBENIUS over active columns */
int sa_bv_norm_column(sa_bv V, int j, double *nrm);
int sa_bv_norm(sa_bv V, double *nrm_fro);
/* ops->copycolumn (bvmat.c:250-260) */
int sa_bv_copy_column(sa_bv V, int src, int dst);
/* BVMatMultColumn (bvops.c:862-885): V(:,j+1) = A V(:,j) */
int sa_bv_mat_mult_column(sa_bv V, sa_mat A, int j);
/* BVOrthonormalizeColumn (bvorthog.c:380-427) with classical Gram-Schmidt and
   DGKS refinement "if needed" (bvorthog.c:91-217), eta as in BVSetOrthogonalization.
   Coefficients land in column j of the coefficient buffer (norm at row j). */
/* BVSetOrthogonalization's refinement (src/sys/classes/bv/interface/bvbasic.c;
 * BVOrthogonalizeGS bvorthog.c:172-199): SA_REFINE_IFNEEDED (default: DGKS,
 * repeat while the estimated norm falls below eta times the norm before the
 * pass, at most three passes), SA_REFINE_NEVER (one pass, then the true norm;
 * lindep only for a zero norm), SA_REFINE_ALWAYS (two passes).  Applies to
 * sa_bv_orthonormalize_column, sa_bv_mat_arnoldi and
 * sa_bv_orthogonalize_some_column of this basis. */
enum { SA_REFINE_IFNEEDED = 0, SA_REFINE_NEVER = 1, SA_REFINE_ALWAYS = 2 };
int sa_bv_set_orthogonalization(sa_bv V, int refine);
/* BVOrthogonalizeSomeColumn (src/sys/classes/bv/interface/bvorthog.c:317-378):
 * modified Gram-Schmidt of column j against the columns i < j with which[i]
 * != 0 (which: j host ints), with the basis's refinement; h (optional, host,
 * j+1 entries) receives the coefficients with the norm at index j. */
int sa_bv_orthogonalize_some_column(sa_bv V, int j, const int *which, double eta, double *h, double *norm,
                                    int *lindep);
int sa_bv_orthonormalize_column(sa_bv V, int j, double eta, double *norm, int *lindep);
/* Copy columns [c0,c1) of the m x m coefficient buffer to the host (ld m). */
int sa_bv_get_coefficients(sa_bv V, int c0, int c1, double *buf_host);
/* The same SpMV + CGS/DGKS step fused for an Arnoldi/Lanczos sweep
   (BVMatArnoldi/BVMatLanczos bvkrylov.c:56-113,165-226): columns k+1..m are
   produced from V(:,k); *m is reduced on breakdown. */
int sa_bv_mat_arnoldi(sa_bv V, sa_mat A, int k, int *m, double eta, double *beta, int *breakdown);
/* Block operations (BVDot bvglobal.c:86-120, BVMatMult bvops.c:692-722,
   BVOrthogonalize bvorthog.c:741-767).
   sa_bv_dot: M(Y.l+i, X.l+j) = Y(:,Y.l+i)^T X(:,X.l+j) over the active columns,
   M host column-major with ldm >= Y.k; other entries untouched.
   sa_bv_matmult: Y(:, Y.l+j) = A X(:, X.l+j); X and Y need equally many active
   columns.
   sa_bv_orthogonalize: QR of the active columns V(:, l:k), the leading ones
   assumed orthonormal: method SA_ORTHOG_BLOCK_GS (column by column, CGS+DGKS),
   _CHOL (block GS against the leading columns, R = chol(V2^T V2), V2 = V2 R^-1)
   or _SVQB (V2 = V2 D U Lambda^-1/2 from the eigendecomposition of the scaled
   Gram matrix; R is then not triangular).  R (host, ldr >= k, or NULL)
   receives columns l..k-1 of the factor: rows 0..l-1 the block-GS
   coefficients, rows l..k-1 the factor of V2 (zeros below the diagonal for GS
   and CHOL).  A dependent column stops GS with an error, as in the reference.
   _TSQR (bvorthog.c:615-631, bvlapack.c:347-480): Householder QR over a tree
   of 256-row tiles, Q formed explicitly; _TSQRCHOL (bvorthog.c:638-659): the
   same tree for R only, then V2 = V2 R^-1.  Both return the Q and R that
   geqrf/orgqr give on one process (same signs).  Above 64 active columns
   the block is factored in panels of 64 (each orthogonalised twice against
   the previous panels, then by the tree), for both methods. */
#define SA_ORTHOG_BLOCK_GS 0
#define SA_ORTHOG_BLOCK_CHOL 1
#define SA_ORTHOG_BLOCK_SVQB 2
#define SA_ORTHOG_BLOCK_TSQR 3
#define SA_ORTHOG_BLOCK_TSQRCHOL 4
int sa_bv_dot(sa_bv X, sa_bv Y, double *M_host, int ldm);
int sa_bv_matmult(sa_bv X, sa_mat A, sa_bv Y);
int sa_bv_orthogonalize(sa_bv V, int method, double *R_host, int ldr);
/* BVMult (src/sys/classes/bv/interface/bvops.c:46-108; bvmat.c:17-36):
   Y(:, Y.l:Y.k) = beta Y(:, Y.l:Y.k) + alpha X(:, X.l:X.k) Q(X.l:X.k, Y.l:Y.k)
   with Q host column-major (ldq >= X.k); Q == NULL gives BVAXPY,
   Y = beta Y + alpha X (equal numbers of active columns).  X != Y. */
int sa_bv_mult(sa_bv Y, double alpha, double beta, sa_bv X, const double *Q_host, int ldq);
/* BVCopy (bvmat.c:234-248): Y(:, Y.l + j) = X(:, X.l + j), j < X.k - X.l. */
int sa_bv_copy(sa_bv X, sa_bv Y);
/* BVResize (bvbasic.c:338-395; bvmat.c:262-284): m columns (< 128); copy != 0
   keeps the first min(m, old m) columns; the active columns are clipped. */
int sa_bv_resize(sa_bv V, int m, int copy);
/* One pass of classical Gram-Schmidt of column j against the active columns
   V(:, l:j) with one global reduction, BVOrthogonalizeCGS1
   (src/sys/classes/bv/interface/bvorthog.c:91-132): the pass behind the
   optional gramschmidt slot of struct _BVOps (bvimpl.h:53, called at
   bvorthog.c:134 up to three times by the DGKS "if needed" loop).  c (host,
   j-l entries) receives the pass's coefficients, h (host, or NULL) adds them
   (BV_AddCoefficients), onorm = ||v|| before the pass, norm = the estimate
   sqrt(onorm^2 - sum c^2), or the true norm when that is <= 0. */
int sa_bv_gs_pass(sa_bv V, int j, double *h_host, double *c_host, double *onorm, double *norm);
/* Gram-Schmidt passes (CGS1 + DGKS refinements) run on V so far. */
int sa_bv_get_pass_count(sa_bv V, int64_t *passes);

/* ---- DS: projected dense problem (host FP64, LAPACK) ---------------------------
 * DSSolve + DSSort of a compact DSHEP problem (src/sys/classes/ds/impls/hep/dshep.c:
 * 221-347,383-426): diagonal d[n], off-diagonal/arrow e[n-1], l locked rows,
 * arrowhead up to row k.  wr[n] receives the sorted eigenvalues and Q
 * ((n+1) x (n+1), column-major) the eigenvectors. */
int sa_ds_hep_compact(int n, int l, int k, const double *d, const double *e, int which,
                      double *wr, double *Q);
/* Path of the LAPACK library resolved at run time (0 when one was found). */
int sa_lapack_info(char *buf, int len);

/* ---- EPS: Krylov-Schur --------------------------------------------------------
 * Mirrors EPSCreate/EPSSetOperators/EPSSetProblemType/EPSSetWhichEigenpairs/
 * EPSSetDimensions/EPSSetTolerances/EPSSetInitialSpace/EPSSolve/EPSGetConverged/
 * EPSGetEigenpair/EPSComputeError (the files of src/eps/interface) for the default
 * Krylov-Schur variant (src/eps/impls/krylov/krylovschur/krylovschur.c:93-348). */
int sa_eps_create(sa_eps *eps);
int sa_eps_destroy(sa_eps eps);
int sa_eps_set_operators(sa_eps eps, sa_mat A);
/* Optional second operator in FP32 used by the mixed-precision bulk phase. */
int sa_eps_set_operators_fp32(sa_eps eps, sa_mat A32);
int sa_eps_set_problem_type(sa_eps eps, int hermitian);
/* EPSSetType: SA_EPS_KRYLOVSCHUR (default, krylovschur.c), SA_EPS_ARNOLDI
   (explicitly restarted Arnoldi, arnoldi.c:52-131: always a non-symmetric
   projected problem, ncv <= nev + mpd) or SA_EPS_LANCZOS (explicitly
   restarted Lanczos with full reorthogonalization, lanczos.c:518-697 with
   -eps_lanczos_reorthog full; symmetric problems, ncv <= nev + mpd). */
#define SA_EPS_KRYLOVSCHUR 0
#define SA_EPS_ARNOLDI 1
#define SA_EPS_LANCZOS 2
int sa_eps_set_type(sa_eps eps, int type);
/* EPSSetWhichEigenpairs (the EPSWhich values of slepceps.h; the target ones
   compare distances to the target, slepcsc.c:240-267) and EPSSetTarget. */
int sa_eps_set_which(sa_eps eps, int which);
int sa_eps_set_target(sa_eps eps, double target);
/* EPSSetExtraction: SA_EPS_RITZ (default) or SA_EPS_HARMONIC (harmonic Ritz
   values and vectors with respect to the target: the projected matrix is
   translated by DSTranslateHarmonic before each solve, krylovschur.c:270-271,
   311-320, arnoldi.c:96-97; Krylov-Schur and Arnoldi). */
#define SA_EPS_RITZ 0
#define SA_EPS_HARMONIC 1
int sa_eps_set_extraction(sa_eps eps, int extraction);
/* STSetType + STSetShift: SA_ST_SHIFT (sigma = 0, the default) or
   SA_ST_SINVERT (shift-and-invert, src/sys/classes/st/impls/sinvert: the
   operator is (A - sigma I)^-1, eigenvalues mapped back by sigma + 1/theta;
   needs which SA_TARGET_MAGNITUDE or SA_TARGET_REAL).  Without shift_given
   sigma is the target (STSetDefaultShift).  The solves run on the device,
   by default as the reference's KSPPREONLY + PCLU (stsles.c:53-54): a banded
   LU with partial pivoting of A - sigma I factored once per solve, or by
   BiCGStab (sa_eps_set_st_ksp). */
#define SA_ST_SHIFT 0
#define SA_ST_SINVERT 1
int sa_eps_set_st(sa_eps eps, int st_type, int shift_given, double shift);
/* KSP of the spectral transformation: SA_PC_LU (default) is KSPPREONLY +
   PCLU, a banded LU on the device (dgbtf2/dgbtrs steps, one workgroup; the
   band is taken from the CSR pattern and limited to 1 GiB of storage and
   2e9 multiply-adds, beyond which the call errors); SA_PC_JACOBI or
   SA_PC_NONE select BiCGStab with that preconditioner, rtol relative to
   ||b|| (default 1e-12) and max_it (10000) — Jacobi needs the diagonal of
   A - sigma I bounded away from zero, and a solve that breaks down or does
   not converge ends the solve with an error (indefinite non-symmetric
   A - sigma I can stagnate it; the LU does not). */
#define SA_PC_NONE 0
#define SA_PC_JACOBI 1
#define SA_PC_LU 2
int sa_eps_set_st_ksp(sa_eps eps, int pc_type, double rtol, int max_it);
/* KSP type of the iterative solves (with SA_PC_JACOBI or SA_PC_NONE):
   SA_KSP_BCGS (BiCGStab, the default here) or SA_KSP_GMRES (GMRES(restart),
   right preconditioned, classical Gram-Schmidt without refinement, the
   convergence test on the recurrence residual: KSPGMRES with
   KSPGMRESSetRestart, PETSc's default KSP type with restart 30;
   src/ksp/ksp/impls/gmres in PETSc, selected by STGetKSP + KSPSetType at
   the reference's st/interface/stsles.c).  restart in [1, 126]. */
#define SA_KSP_BCGS 0
#define SA_KSP_GMRES 1
int sa_eps_set_st_ksp_type(sa_eps eps, int ksp_type, int restart);
/* EPSBuildBalance_Krylov (src/eps/interface/epsdefault.c:443-504) for an
 * FP64 operator: the diagonal D (n host doubles) of D A D^-1 after its
 * iterations, twoside with A^T; the +-1 vectors come from the library's
 * generator (U[0,1) seed 0xBA1A + j, sign by < 0.5) in place of PetscRandom.
 * cutoff <= 0 takes 1e-8.  Exposes what sa_eps_set_balance applies. */
int sa_mat_build_balance(sa_mat A, int twoside, int its, double cutoff, double *D_host);
/* EPSSetBalance (epsopts.c:1160-1215; EPSBuildBalance_Krylov,
   epsdefault.c:443-504): balancing of non-Hermitian problems by a diagonal D
   built from its iterations (its <= 0: 5) of random +-1 vectors, the update
   cut off below cutoff ||A||_inf (cutoff <= 0: 1e-8).  The solve runs on
   D A D^-1 and the eigenvectors returned are D^-1 y, normalised; ignored for
   Hermitian problems, as in the reference.  FP64, one-sided solvers, no
   spectral transformation, Ritz extraction. */
#define SA_BALANCE_NONE 0
#define SA_BALANCE_ONESIDE 1
#define SA_BALANCE_TWOSIDE 2
int sa_eps_set_balance(sa_eps eps, int balance, int its, double cutoff);
/* EPSKrylovSchurSetRestart / GetRestart (krylovschur.c:350-410): the
   fraction keep in [0.1, 0.9] of the non-converged basis kept at restart
   (default 0.5).  EPSKrylovSchurSetLocking / GetLocking (krylovschur.c:
   427-490): with lock 0 converged pairs stay in the active basis (the
   restart keeps them among the l vectors, krylovschur.c:295); default 1. */
int sa_eps_krylovschur_set_restart(sa_eps eps, double keep);
int sa_eps_krylovschur_get_restart(sa_eps eps, double *keep);
int sa_eps_krylovschur_set_locking(sa_eps eps, int lock);
int sa_eps_krylovschur_get_locking(sa_eps eps, int *lock);
/* EPSArnoldiSetDelayed / EPSArnoldiGetDelayed (arnoldi.c:148-220): delayed
   reorthogonalization in SA_EPS_ARNOLDI (EPSDelayedArnoldi,
   epskrylov.c:24-114, writing the projected Hessenberg matrix; the reference
   itself warns that it may be unstable, test9.c). */
int sa_eps_arnoldi_set_delayed(sa_eps eps, int delayed);
int sa_eps_arnoldi_get_delayed(sa_eps eps, int *delayed);
/* EPSLanczosSetReorthog / EPSLanczosGetReorthog (lanczos.c:716-775; enum
   EPSLanczosReorthogType, include/slepceps.h:445): reorthogonalization of
   SA_EPS_LANCZOS.  FULL (the default here; the reference requires an explicit
   choice, lanczos.c:54): classical Gram-Schmidt against every column.  LOCAL
   (lanczos.c:89-118, with the restart steps of lanczos.c:570-575,633-660,
   678-681): modified Gram-Schmidt against the locked columns and the two
   previous Lanczos vectors only, purging repeated and spurious Ritz values at
   each restart.  PERIODIC and PARTIAL (EPSPartialLanczos, lanczos.c:338-441):
   local steps plus a second MGS against the columns of the current sweep
   when the omega recurrence estimates a loss of orthogonality above
   sqrt(eps/ncv) (periodic: all of them, twice in a row; partial: the
   intervals around the largest estimates).  SELECTIVE (EPSSelectiveLanczos,
   lanczos.c:181-258): local steps plus orthogonalization against the Ritz
   vectors of the current sweep whose bound is below sqrt(eps) ||A||.
   DELAYED (EPSDelayedArnoldi, epskrylov.c:24-114): one classical
   Gram-Schmidt pass per step with the normalisation delayed by one step and
   a second pass by two.  As in the reference it keeps orthogonality only
   while the restart vectors are far from converged: restarted from nearly
   converged Ritz vectors it loses it (the reference tests it at tol 1e-8 on
   a 30 x 30 tridiagonal).  Other values are rejected with an error. */
#define SA_LANCZOS_REORTHOG_LOCAL 0
#define SA_LANCZOS_REORTHOG_FULL 1
#define SA_LANCZOS_REORTHOG_SELECTIVE 2
#define SA_LANCZOS_REORTHOG_PERIODIC 3
#define SA_LANCZOS_REORTHOG_PARTIAL 4
#define SA_LANCZOS_REORTHOG_DELAYED 5
int sa_eps_lanczos_set_reorthog(sa_eps eps, int reorthog);
int sa_eps_lanczos_get_reorthog(sa_eps eps, int *reorthog);
/* EPSSetTwoSided: two-sided Krylov-Schur (ks-twosided.c:126-241) that also
   computes left eigenvectors (y^T A = lambda y^T), with a left basis swept by
   A^T (built from the operator's CSR at the first solve).  Non-symmetric
   problems, FP64, Ritz extraction. */
int sa_eps_set_two_sided(sa_eps eps, int twosided);
/* EPSGetLeftEigenvector: left eigenvector i in the order of sa_eps_get_eigenpair. */
int sa_eps_get_left_eigenvector(sa_eps eps, int i, double *yr_host, double *yi_host);
/* EPSSetTrueResidual / EPSGetTrueResidual (src/eps/interface/epsopts.c:
   1333-1370): the convergence test of Krylov-Schur (also two-sided, for the
   right residual) and Arnoldi takes the explicit residual
   ||A x - lambda x|| / |lambda| of each candidate Ritz vector x = V X(:,k)
   (src/eps/impls/krylov/epskrylov.c:225-265), lambda back-transformed under
   shift-and-invert, instead of the Ritz estimate beta |y_last|.  Formed on
   the device (restart GEMM kernel + fused SpMV residual).  Not with
   EPSLANCZOS (the reference ignores it there) or balancing. */
int sa_eps_set_true_residual(sa_eps eps, int trueres);
int sa_eps_get_true_residual(sa_eps eps, int *trueres);
/* EPSSetDimensions (epsopts.c:401-425): ncv or mpd <= 0 take the defaults of
   EPSSetDimensions_Default at solve time.  ncv is at most 126 (the fused
   Gram-Schmidt kernels handle 127 basis columns; an FP64 tile wider than 108
   columns takes more than 64 KB of LDS, which gfx950 allows a workgroup). */
int sa_eps_set_dimensions(sa_eps eps, int nev, int ncv, int mpd);
int sa_eps_set_tolerances(sa_eps eps, double tol, int max_it);
/* EPSSetInitialSpace (epssetup.c:613-626): nvec host vectors of length n,
   stored consecutively.  As in the reference, Krylov-Schur starts from the
   first one after orthonormalisation. */
int sa_eps_set_initial_space(sa_eps eps, int nvec, const double *vecs_host);
/* EPSSetDeflationSpace (epssetup.c:544-583): nvec host vectors of length n,
   stored consecutively.  The solve seeks eigenpairs in their orthogonal
   complement: they are orthonormalized in order (a linearly dependent one is
   dropped, BVInsertVecs) and every Krylov vector is kept orthogonal to them.
   As in the reference the space is removed at the end of each solve, so it is
   set before every solve.  Krylov-Schur and Arnoldi, FP64, one-sided; up to
   128 vectors. */
int sa_eps_set_deflation_space(sa_eps eps, int nvec, const double *vecs_host);
/* Precision: SA_FP64 plain, or mixed (bulk FP32 to tol32, then FP64 polish
   seeded with the FP32 Ritz vectors). */
int sa_eps_set_mixed(sa_eps eps, int mixed, double tol32);
int sa_eps_solve(sa_eps eps);
int sa_eps_get_converged(sa_eps eps, int *nconv);
int sa_eps_get_iteration_number(sa_eps eps, int *its);
int sa_eps_get_converged_reason(sa_eps eps, int *reason);
/* Converged reasons (EPSConvergedReason, include/slepceps.h:179-185). */
enum { SA_EPS_CONVERGED_TOL = 1, SA_EPS_CONVERGED_USER = 2, SA_EPS_DIVERGED_ITS = -1,
       SA_EPS_DIVERGED_BREAKDOWN = -2, SA_EPS_CONVERGED_ITERATING = 0 };
/* EPSMonitorSet (src/eps/interface/epsmon.c:96-122): adds a function called
 * at the end of every outer iteration (EPSMonitor, epsmon.c:39-46, called at
 * krylovschur.c:344 and the matching places of the other solvers) with the
 * iteration number, the number of converged pairs, the Ritz values of the
 * projected problem (eigr, eigi; with shift-and-invert these are theta, and
 * lambda = sigma + 1/theta, as the reference's monitors back-transform them),
 * the error estimates and their count nest.  At most 5 monitors; they run in
 * the order set.  A nonzero return ends the solve with an error.  With mixed
 * precision the monitors are called in both phases (the iteration count
 * restarts at 1 in the FP64 phase).  EPSMonitorCancel: removes them all. */
typedef int (*sa_eps_monitor_fn)(sa_eps eps, int its, int nconv, const double *eigr, const double *eigi,
                                 const double *errest, int nest, void *ctx);
int sa_eps_monitor_set(sa_eps eps, sa_eps_monitor_fn monitor, void *ctx);
int sa_eps_monitor_cancel(sa_eps eps);
/* EPSGetBV + BVSetOrthogonalization for the solver's basis (-bv_orthog_refine,
 * -bv_orthog_eta): refinement SA_REFINE_* and eta (<= 0 keeps 0.7071).  With
 * SA_REFINE_NEVER the delayed variants (EPSArnoldiSetDelayed,
 * -eps_lanczos_reorthog delayed) run EPSDelayedArnoldi1 (epskrylov.c:120-157)
 * as arnoldi.c:84-87 and lanczos.c:500 select it. */
int sa_eps_set_orthogonalization(sa_eps eps, int refine, double eta);
/* EPSSetStoppingTestFunction (src/eps/interface/epsopts.c:863-885): replaces
 * EPSStoppingBasic (src/eps/interface/epsdefault.c:290-302).  Called once per
 * outer iteration after the convergence test (krylovschur.c:284) with the
 * iteration number, max_it, the number of converged pairs and nev; it sets
 * *reason (SA_EPS_CONVERGED_ITERATING to continue, any other value ends the
 * iteration and becomes the converged reason).  A nonzero return ends the
 * solve with an error.  NULL restores the basic test. */
typedef int (*sa_eps_stopping_fn)(sa_eps eps, int its, int max_it, int nconv, int nev, int *reason, void *ctx);
int sa_eps_set_stopping_test_function(sa_eps eps, sa_eps_stopping_fn stopping, void *ctx);
/* EPSSetConvergenceTest (epsopts.c:796-810): the error estimate of a pair
   with eigenvalue lambda and residual norm r is r (SA_CONV_ABS), r / |lambda|
   (SA_CONV_REL, the default) or r / (||A||_inf + |lambda| ||B||_inf)
   (SA_CONV_NORM; ||A|| is computed at the solve, and ||B|| stays 0 for a
   standard problem unless a backward error was computed before, as in the
   reference, epssetup.c:351-359 and epssolve.c:794-809).  Under
   shift-and-invert SA_CONV_NORM takes the back-transformed lambda
   (epskrylov.c:245). */
enum { SA_CONV_ABS = 0, SA_CONV_REL = 1, SA_CONV_NORM = 2, SA_CONV_USER = 3 };
int sa_eps_set_convergence_test(sa_eps eps, int conv);
/* EPSSetConvergenceTestFunction (epsopts.c:751-770): *errest from the
   eigenvalue (eigr, eigi) and the residual norm res; selects SA_CONV_USER.
   A nonzero return ends the solve with an error.  NULL restores SA_CONV_REL. */
typedef int (*sa_eps_converged_fn)(sa_eps eps, double eigr, double eigi, double res, double *errest, void *ctx);
int sa_eps_set_convergence_test_function(sa_eps eps, sa_eps_converged_fn fn, void *ctx);
/* The matrix norms of SA_CONV_NORM and of backward errors, ||A||_inf and
   ||B||_inf (EPSView's "computed matrix norms"); 0 until computed. */
int sa_eps_get_matrix_norms(sa_eps eps, double *nrma, double *nrmb);
/* EPSSetEigenvalueComparison (epsopts.c:675-684): the order of the eigenvalues is the
   user's: cmp(ar, ai, br, bi, &res, ctx) sets res < 0 when a goes before b,
   > 0 when after, 0 when equal (under shift-and-invert the values are the
   back-transformed ones, SlepcSCCompare, slepcsc.c:40-63); selects
   SA_WHICH_USER.  A nonzero return ends the solve with an error.  NULL
   restores SA_LARGEST_MAGNITUDE. */
typedef int (*sa_eps_compare_fn)(double ar, double ai, double br, double bi, int *res, void *ctx);
int sa_eps_set_eigenvalue_comparison(sa_eps eps, sa_eps_compare_fn cmp, void *ctx);
/* Eigenpair i in the order EPSGetEigenpair returns them; vectors may be NULL. */
int sa_eps_get_eigenpair(sa_eps eps, int i, double *re, double *im, double *xr_host, double *xi_host);
int sa_eps_get_error_estimate(sa_eps eps, int i, double *errest);
/* EPSComputeError (epssolve.c:742-815) in FP64 on the device. */
int sa_eps_compute_error(sa_eps eps, int i, int type, double *error);
/* Counters of the last solve: matvecs, Gram-Schmidt passes, restarts of each
   phase, seconds of each phase, algorithmic HBM bytes moved by the SpMV/BV
   kernels (per the byte model of the kernels). */
typedef struct {
  int64_t matvecs, gs_passes, matvecs32, gs_passes32;
  int     its, its32, nconv, reason;
  double  seconds, seconds32, seconds_ds;
  double  bytes_spmv, bytes_bv, bytes_restart;
  int64_t true_residuals;   /* explicit residuals taken by the convergence test */
  int64_t ksp_its;          /* iterations of the shift-and-invert linear solves (KSPGetTotalIterations) */
} sa_eps_stats;
int sa_eps_get_stats(sa_eps eps, sa_eps_stats *st);

/* ---- batched Krylov-Schur: many independent symmetric problems -----------------
 * Each problem runs the default Krylov-Schur of krylovschur.c:227-348 (Lanczos
 * + DSHEP, locking, keep 0.5) on its own; all problems share one block-diagonal
 * operator and one basis in HBM, and one restart phase of every problem is one
 * kernel launch (a workgroup per problem).  No counterpart type in the
 * reference: it is the drop-in for a loop of EPSSolve calls over the problems.
 * Input: P problems, n[p] rows each; rowptr holds the P local row-pointer
 * arrays (n[p]+1 entries each) back to back, col/val the P local CSR arrays. */
typedef struct sa_batch_s *sa_batch;
int sa_batch_create(int P, const int32_t *n, const int32_t *rowptr, const int32_t *col,
                    const double *val, sa_batch *batch);
int sa_batch_destroy(sa_batch batch);
/* mixed != 0: FP32 bulk to tol32, then FP64 polish seeded with the sum of the
   FP32 Ritz vectors.  v0 optional (concatenated start vectors); NULL draws
   U[-1,1) with seed sa_sub_seed(seed, p).  ncv > 0: every problem uses it
   (nev < ncv <= min(64, n)); ncv == 0: the per-problem choice of
   sa_batch_autotune_ncv, else EPS's default max(2 nev, nev + 15). */
int sa_batch_solve(sa_batch batch, int which, int nev, int ncv, double tol, int max_it,
                   int mixed, double tol32, const double *v0, uint64_t seed);
int sa_batch_get_converged(sa_batch batch, int p, int *nconv, int *its, int *reason);
/* EPSSetTrueResidual (epsopts.c:1333-1370) for every problem: the convergence
   test takes the explicit residual of each candidate Ritz vector
   (epskrylov.c:225-265), formed and evaluated for all problems at once. */
int sa_batch_set_true_residual(sa_batch batch, int trueres);
int sa_batch_get_eigenpair(sa_batch batch, int p, int i, double *re, double *x_host);
/* err[p*nev + i] = ||A x - lambda x|| / |lambda| in FP64 on the device. */
int sa_batch_compute_errors(sa_batch batch, int nev, double *err);
/* counters6: matvecs, passes, restart rounds (FP64), then the same for FP32;
   seconds3: total, FP32 phase, host projected problems. */
int sa_batch_get_stats(sa_batch batch, int64_t *counters6, double *seconds3);
/* Sweep kernel of the batch: 0 autotune (default; both are timed on the first
   restart round of the first solve phase of each precision and basis width,
   and the faster is kept for later solves of the batch), 1 one workgroup
   per problem runs the whole sweep (ncv < 64), 2 lock-step: each Gram-Schmidt
   slot of every active problem is one launch over all their 64-row chunks.
   The getter returns the mode used by the last FP64 phase and the autotuner's
   timings of one sweep (milliseconds, -1 when not timed). */
int sa_batch_set_sweep(sa_batch batch, int mode);
int sa_batch_get_sweep(sa_batch batch, int *mode, double *ms_persistent, double *ms_wide);
/* The schedule autotuner (mode 0) times, on the first restart round, the
   one-workgroup sweep (when ncv < 64) and three lock-step tile schedules:
   the default (non-temporal tile loads for a basis above 256 MiB, SpMV lanes
   per row from the mean row length), the other tile-load kind, and twice the
   lanes; it reruns the fastest so the round's results are its own.  The
   getter returns the last solve's schedule (kind 1 one workgroup per
   problem, 2 lock-step; nt; lanes) and the four timings in milliseconds (-1:
   not timed).  No counterpart in the reference. */
int sa_batch_get_schedule(sa_batch batch, int *kind, int *nt, int *lanes, double *ms4);
/* Lock-step restart sweeps replayed from HIP graphs (default on): each sweep
   (state reset, 5 launches per step, export) is captured once per launch
   signature (grids, LDS sizes, step count, schedule) and replayed; the
   problem lists, k0, nv and chunk lists it reads are device buffers refreshed
   before each replay.  At most 64 graphs per precision (least recently used
   evicted).  The stats count captures, replays and cached graphs.  No
   counterpart in the reference. */
int sa_batch_set_graphs(sa_batch batch, int on);
int sa_batch_get_graph_stats(sa_batch batch, int64_t *captures, int64_t *replays, int *cached);
/* Basis-size autotuner: the problems, sorted by n, form nclass classes of
   (nearly) equal count, and nsample problems spread over each class form one
   stratified sample, solved as a sub-batch with every candidate ncv (nev <
   cand <= 64; FP64, lock-step sweeps, this batch's true-residual setting,
   start vectors U[-1,1) seed sa_sub_seed(seed, p)).  Each candidate's wall
   time is split over the classes in proportion to their problems'
   algorithmic bytes, and the candidate with the least time per converged
   pair becomes the ncv of every problem of its class (ncv_out[p], may be
   NULL), used by later sa_batch_solve calls with ncv == 0.  The tuning
   getter returns the class bounds (largest n of each class) and the seconds
   per pair of every (class, candidate), 1e300 where a candidate was not run
   (ncv above the smallest sampled n) or left pairs unconverged.
   sa_batch_get_ncv: the ncv of each problem in the last solve.  No
   counterpart in the reference (EPSSetDimensions is set by hand there). */
int sa_batch_autotune_ncv(sa_batch batch, int which, int nev, double tol, const int *cand, int ncand,
                          int nclass, int nsample, uint64_t seed, int *ncv_out);
int sa_batch_get_ncv_tuning(sa_batch batch, int *nclass, int *ncand, int *class_hi, double *scores);
int sa_batch_get_ncv(sa_batch batch, int *ncv);

/* ---- learned warm start: MLP in BF16 on MFMA ---------------------------------
 * theta (din <= 32) -> hidden -> hidden -> dout, ReLU, weights and activations
 * BF16 on MFMA with FP32 accumulation, FP32 master weights, Adam.  No
 * counterpart in the reference: it produces the vector a caller passes to
 * EPSSetInitialSpace (epssetup.c:613-626).  W1 is stored hidden x 32 (input
 * columns beyond din are zero), W2 hidden x hidden, W3 dout x hidden, row-major. */
int sa_mlp_create(int din, int hidden, int dout, uint64_t seed, sa_mlp *net);
int sa_mlp_destroy(sa_mlp net);
/* Forward for a batch: X_host [batch x din] row-major -> Y_host [batch x dout] (FP32). */
int sa_mlp_forward(sa_mlp net, int batch, const float *X_host, float *Y_host);
/* One Adam step on a batch with mean-squared loss against T_host; returns the loss. */
int sa_mlp_train_step(sa_mlp net, int batch, const float *X_host, const float *T_host,
                      float lr, float *loss);
/* Gradients of the mean-squared loss (no update) as FP32 host arrays. */
int sa_mlp_gradients(sa_mlp net, int batch, const float *X_host, const float *T_host,
                     float *gW1, float *gb1, float *gW2, float *gb2, float *gW3, float *gb3, float *loss);
/* FP32 master parameters. */
int sa_mlp_get_params(sa_mlp net, float *W1, float *b1, float *W2, float *b2, float *W3, float *b3);
int sa_mlp_set_params(sa_mlp net, const float *W1, const float *b1, const float *W2,
                      const float *b2, const float *W3, const float *b3);
/* Bilinear prolongation of nmodes g x g modes (cell-centred) to nx x ny cell grids. */
int sa_prolong_bilinear(int nmodes, int g, int nx, int ny, const float *modes_host, double *out_host);

#ifdef __cplusplus
}
#endif
#endif
