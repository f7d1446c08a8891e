/*
 * blast_hip.h - C ABI of the MI355X (gfx950) batched dense-kernel layer.
 *
 * Every entry point is the batched, device-side replacement of one BLAST
 * function template (paths relative to the mkatliar/blast checkout).  The
 * reference has no FFI: its boundary is C++20 templates instantiated in the
 * caller's translation unit.  This header is what those templates' batched
 * overloads (include/blast/hip/Blast.hpp) call; any other language binds it
 * directly (plain C types only: ctypes, cgo, JNI and the like).
 *
 * Conventions
 *  - All matrix data pointers are DEVICE pointers (hipMalloc'd or torch CUDA
 *    tensors).  Descriptors, scalars and sizes are host values.
 *  - A batch operand is `batch` matrices, `batch_stride` elements apart, each
 *    stored column-major in BLAST's dense format (element (i,j) at
 *    i + spacing*j; dense/DynamicMatrix.hpp:149) or panel format (element
 *    (i,j) at (i/PS)*spacing + i%PS + j*PS, PS = 4 for double and 8 for float;
 *    panel/DynamicPanelMatrix.hpp:184-188).  `rows` x `cols` is the LOGICAL
 *    size; trans = 1 views the storage transposed (BLAST trans(); a row-major
 *    matrix is the trans = 1 view of its transpose's column-major storage).
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    stream-ordered, do not synchronise and do not allocate.
 *  - Return value: 0 on success; -k if argument k (1-based, `stream` = 1) is
 *    invalid; BLAST_HIP_ERR_RUNTIME if the HIP runtime reported an error
 *    (blast_hip_last_error() describes it).  Per-instance numerical status
 *    goes to the optional device array `info` (int32 per instance): 0 = ok,
 *    k > 0 = column k (1-based) had a pivot that is not positive (potrf:
 *    <= 0 or NaN) or exactly zero (getrf).  Instances are independent: one bad instance never
 *    affects another.
 *  - Aliasing allowed as in BLAST: C == D (gemm, syrk), X == B (trsm),
 *    L == A (potrf).
 */
#ifndef BLAST_HIP_H
#define BLAST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLAST_HIP_ERR_RUNTIME (-1000)

typedef struct blast_matrix_t {
    void* data;           /* device pointer to instance 0 */
    int64_t rows;         /* logical rows */
    int64_t cols;         /* logical columns */
    int64_t spacing;      /* dense: leading dimension; panel: elements between panels */
    int64_t batch_stride; /* elements between consecutive instances */
    int32_t panel;        /* 1 = panel storage, 0 = dense column-major storage */
    int32_t trans;        /* 1 = logical matrix is the transpose of the storage */
} blast_matrix_t;

/* Algorithm selection for all entry points of the calling process.
 *  BLAST_HIP_ALGO_AUTO      fastest kernel for the shape (MFMA tiles, fused
 *                           kernels); results agree with BLAST to backward
 *                           error c*n*eps.
 *  BLAST_HIP_ALGO_REFERENCE column-sweep kernels that perform BLAST's
 *                           per-element operation sequence (same FMA order,
 *                           IEEE division and square root): bit-identical
 *                           to BLAST's CPU results. */
enum { BLAST_HIP_ALGO_AUTO = 0, BLAST_HIP_ALGO_REFERENCE = 1 };
int blast_hip_set_algorithm(int algo);
int blast_hip_get_algorithm(void);

const char* blast_hip_last_error(void);
const char* blast_hip_version(void);

/* D := alpha*A*B + beta*C.  Replaces blast::gemm(alpha, A, B, beta, C, D)
 * (include/blast/math/algorithm/Gemm.hpp:80-97) and gemm(A, B, C, D)
 * (:113-118, alpha = beta = 1); the NT form is B with trans = 1
 * (bench/blast/math/panel/StaticGemm.cpp:37). */
int blast_hip_dgemm_batched(void* stream, double alpha, const blast_matrix_t* A, const blast_matrix_t* B,
                            double beta, const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch);
int blast_hip_sgemm_batched(void* stream, float alpha, const blast_matrix_t* A, const blast_matrix_t* B,
                            float beta, const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch);

/* D := alpha*A*A^T + beta*C, lower triangle only (i >= j written).
 * Replaces blast::syrkLower (include/blast/math/dense/Syrk.hpp:15-145). */
int blast_hip_dsyrk_lower_batched(void* stream, double alpha, const blast_matrix_t* A, double beta,
                                  const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch);
int blast_hip_ssyrk_lower_batched(void* stream, float alpha, const blast_matrix_t* A, float beta,
                                  const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch);

/* C := alpha*A*B, A upper triangular (side_right = 0), or C := alpha*B*A,
 * A lower triangular (side_right = 1).  Replaces blast::trmm
 * (include/blast/math/algorithm/Trmm.hpp:260-310; the other side/triangle
 * combinations throw std::logic_error there and return -3 here). */
int blast_hip_dtrmm_batched(void* stream, int side_right, int upper, double alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* C, int64_t batch);
int blast_hip_strmm_batched(void* stream, int side_right, int upper, float alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* C, int64_t batch);

/* Left (side_right = 0): A*X = alpha*B; right: X*A = alpha*B; A lower
 * (upper = 0) or upper triangular, unit or non-unit diagonal.  X may alias B.
 * Replaces blast::trsm<UpLo, UNIT>(A, B, X) (include/blast/math/dense/Trsm.hpp:17-66,
 * alpha = 1) and the reference forms trsm(A, uplo, diag, X, alpha, B) /
 * trsm(X, A, uplo, diag, alpha, B) (include/blast/math/reference/Trsm.hpp:164,206). */
int blast_hip_dtrsm_batched(void* stream, int side_right, int upper, int unit, double alpha,
                            const blast_matrix_t* A, const blast_matrix_t* B, const blast_matrix_t* X,
                            int64_t batch);
int blast_hip_strsm_batched(void* stream, int side_right, int upper, int unit, float alpha,
                            const blast_matrix_t* A, const blast_matrix_t* B, const blast_matrix_t* X,
                            int64_t batch);

/* trsvLeftLower / trsvLeftUpper (dense/Trsv.hpp:15-68): A x = b with A
 * n x n lower (upper = 0) or upper (upper = 1) triangular, non-unit, one
 * right-hand side per instance: b and x are n x 1 descriptors (any storage,
 * e.g. a row or column of a panel matrix).  x may alias b.  "Invalid argument
 * size" (-3) when A is not n x n. */
int blast_hip_dtrsv_batched(void* stream, int upper, const blast_matrix_t* A, const blast_matrix_t* b,
                            const blast_matrix_t* x, int64_t batch);
int blast_hip_strsv_batched(void* stream, int upper, const blast_matrix_t* A, const blast_matrix_t* b,
                            const blast_matrix_t* x, int64_t batch);

/* L := chol(A), lower, M >= N (tall allowed).  Replaces blast::potrf
 * (include/blast/math/dense/Potrf.hpp:66-106 and panel/Potrf.hpp:59-106).
 * Dense L: only i >= j written.  Panel L: the strict upper part of each 4x4
 * diagonal block is written as 0.  Non-positive pivots produce NaN as in
 * BLAST and are reported in info. */
int blast_hip_dpotrf_batched(void* stream, const blast_matrix_t* A, const blast_matrix_t* L, int64_t batch,
                             int32_t* info);
/* FP32 storage: on ALGO_AUTO, square n <= 256 is factored with FP64
 * arithmetic inside (L rounded to FP32 on store). */
int blast_hip_spotrf_batched(void* stream, const blast_matrix_t* A, const blast_matrix_t* L, int64_t batch,
                             int32_t* info);

/* In-place P*A = L*U with partial pivoting (column-major storage) or without
 * pivoting (trans = 1, i.e. row-major storage).  ipiv: device int64 array,
 * min(M,N) entries per instance, ipiv_stride apart, 0-based, LAPACK
 * sequential-interchange meaning.  Replaces blast::getrf
 * (include/blast/math/dense/Getrf.hpp:52-119 and :141-162). */
int blast_hip_dgetrf_batched(void* stream, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride,
                             int64_t batch, int32_t* info);
/* FP32 storage: on ALGO_AUTO a square column-major A with n <= 256 is
 * factored with FP64 arithmetic inside (factors rounded to FP32); other
 * shapes and ALGO_REFERENCE follow BLAST's FP32 operation order. */
int blast_hip_sgetrf_batched(void* stream, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride,
                             int64_t batch, int32_t* info);

/* Cholesky solve, the sequence potrf(A, L); trsm<Lower>(L, B, Y);
 * trsm<Upper>(trans(L), Y, X) fused into one launch (LAPACK dposv contract:
 * L and X are written).  X may alias B, L may alias A.  A is read from its
 * lower triangle only. */
int blast_hip_dposv_batched(void* stream, const blast_matrix_t* A, const blast_matrix_t* L,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info);
/* The same on FP32 storage (BLAST's float instantiation of the sequence).
 * ALGO_AUTO, square n <= 256, nrhs <= 8: the arithmetic is FP64 inside
 * (loads widened, L and X rounded to FP32 on store); other shapes and
 * ALGO_REFERENCE: BLAST's FP32 operation order. */
int blast_hip_sposv_batched(void* stream, const blast_matrix_t* A, const blast_matrix_t* L,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info);

/* Mixed-precision Cholesky solve with FP64 iterative refinement (the LAPACK
 * dsposv contract).  Not in BLAST; it solves the same systems as
 * blast_hip_dposv_batched and is checked against the FP64 oracle by
 * backward error.  The factorisation runs in FP32 (precision = 0) or with
 * BF16 products and FP32 accumulation (precision = 1); the residual
 * r = b - A x, the update x += d and the convergence test
 * max|r| <= max|x| * ||A||_inf * eps * sqrt(n) (per right-hand side) are
 * FP64; tol > 0 replaces eps*sqrt(n) in that test (tol = 0: LAPACK's test).
 * itmax bounds the refinement steps (LAPACK uses 30).
 * iter (device int32 per instance, required): refinement steps taken, or
 * -1 (FP32 factorisation broke down), -2 (no convergence), -3 (shape not
 * covered by the mixed kernel: n > 128 or nrhs > 4).  Instances with
 * iter < 0 are solved by the FP64 path, which writes their Cholesky factor
 * into A (lower triangle) as LAPACK dsposv does; otherwise A is unchanged.
 * info as for blast_hip_dposv_batched.  X may alias B. */
int blast_hip_dsposv_batched(void* stream, int precision, int itmax, double tol, const blast_matrix_t* A,
                             const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info,
                             int32_t* iter);

/* ---------------------------------------------------------------- adjoints
 * Reverse-mode derivatives, forward/backward pairs of the entry points above.
 * Not in BLAST; closed forms from Giles 2008 and Murray 2016, checked
 * against central finite differences and an FP64 CPU reference.  Symmetric
 * inputs (potrf, posv) are read from their lower triangle, so their adjoint
 * is with respect to the stored lower entries (Abar_ij = G_ij + G_ji below
 * the diagonal for a full gradient G; strict upper of Abar written as 0).
 * Output buffers may be uninitialised; they also serve as workspace.  */

/* D = alpha A B + beta C:  Abar = alpha Dbar B^T, Bbar = alpha A^T Dbar,
 * Cbar = beta Dbar.  Abar, Bbar, Cbar may be NULL (not computed). */
int blast_hip_dgemm_backward_batched(void* stream, double alpha, const blast_matrix_t* A, const blast_matrix_t* B,
                                     double beta, const blast_matrix_t* Dbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, const blast_matrix_t* Cbar, int64_t batch);
int blast_hip_sgemm_backward_batched(void* stream, float alpha, const blast_matrix_t* A, const blast_matrix_t* B,
                                     float beta, const blast_matrix_t* Dbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, const blast_matrix_t* Cbar, int64_t batch);

/* A X = B (left, non-unit, A lower or upper):  Bbar = A^{-T} Xbar,
 * Abar = -Bbar X^T on A's triangle (other triangle 0; Abar may be NULL). */
int blast_hip_dtrsm_backward_batched(void* stream, int upper, const blast_matrix_t* A, const blast_matrix_t* X,
                                     const blast_matrix_t* Xbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, int64_t batch);
int blast_hip_strsm_backward_batched(void* stream, int upper, const blast_matrix_t* A, const blast_matrix_t* X,
                                     const blast_matrix_t* Xbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, int64_t batch);

/* L = chol(A):  Abar = fold(L^{-T} Phi(L^T tril(Lbar)) L^{-1}),
 * Phi = lower triangle with halved diagonal. */
int blast_hip_dpotrf_backward_batched(void* stream, const blast_matrix_t* L, const blast_matrix_t* Lbar,
                                      const blast_matrix_t* Abar, int64_t batch);
int blast_hip_spotrf_backward_batched(void* stream, const blast_matrix_t* L, const blast_matrix_t* Lbar,
                                      const blast_matrix_t* Abar, int64_t batch);

/* X = A^{-1} B (SPD):  Bbar = A^{-1} Xbar, Abar = -fold(Bbar X^T).
 * recompute = 0: F is the factor L kept by the forward pass;
 * recompute = 1: F is A and the factor is recomputed into Abar (for callers
 * that do not keep L: no memory beyond the outputs).  info as for potrf. */
int blast_hip_dposv_backward_batched(void* stream, int recompute, const blast_matrix_t* F, const blast_matrix_t* X,
                                     const blast_matrix_t* Xbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, int64_t batch, int32_t* info);
int blast_hip_sposv_backward_batched(void* stream, int recompute, const blast_matrix_t* F, const blast_matrix_t* X,
                                     const blast_matrix_t* Xbar, const blast_matrix_t* Abar,
                                     const blast_matrix_t* Bbar, int64_t batch, int32_t* info);

/* X = A^{-1} B from getrf's factors (LU in place, ipiv as written by
 * blast_hip_?getrf_batched): X = U^{-1} L^{-1} P B (LAPACK getrs). */
int blast_hip_dgetrs_batched(void* stream, const blast_matrix_t* LU, const int64_t* ipiv, int64_t ipiv_stride,
                             const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch);
int blast_hip_sgetrs_batched(void* stream, const blast_matrix_t* LU, const int64_t* ipiv, int64_t ipiv_stride,
                             const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch);

/* getrs adjoint:  Bbar = A^{-T} Xbar = P^T L^{-T} U^{-T} Xbar,
 * Abar = -Bbar X^T (Abar may be NULL). */
int blast_hip_dgetrs_backward_batched(void* stream, const blast_matrix_t* LU, const int64_t* ipiv,
                                      int64_t ipiv_stride, const blast_matrix_t* X, const blast_matrix_t* Xbar,
                                      const blast_matrix_t* Abar, const blast_matrix_t* Bbar, int64_t batch);
int blast_hip_sgetrs_backward_batched(void* stream, const blast_matrix_t* LU, const int64_t* ipiv,
                                      int64_t ipiv_stride, const blast_matrix_t* X, const blast_matrix_t* Xbar,
                                      const blast_matrix_t* Abar, const blast_matrix_t* Bbar, int64_t batch);

/* D := S element by element between any two storage layouts of the same
 * logical size (dense or panel, either transposed view): the batched device
 * form of assigning a panel matrix to a dense one and back
 * (include/blast/math/expressions/PanelMatrix.hpp:81-167 and :170-239).
 * Only the logical rows x cols elements of D are written; S and D must not
 * overlap unless they are the same descriptor. */
int blast_hip_dcopy_batched(void* stream, const blast_matrix_t* S, const blast_matrix_t* D, int64_t batch);
int blast_hip_scopy_batched(void* stream, const blast_matrix_t* S, const blast_matrix_t* D, int64_t batch);

/* ---- Solves for 128 < n <= 256 (and any n <= 256), nrhs <= 8 ----
 * BLAST has no solve driver; these compose its getrf (dense/Getrf.hpp:52-119)
 * or potrf (dense/Potrf.hpp:66-106) with trsm (dense/Trsm.hpp:17-66) and
 * follow the LAPACK driver contracts named below.  The AUTO algorithm also
 * routes blast_hip_dpotrf/dposv/dgetrf_batched with square n <= 256 to the
 * same kernels. */

/* FP32 factor buffer needed per instance by the two mixed-precision calls
 * below (floats; the buffer holds batch * this many, 0 if n is not covered).
 * LAPACK's SWORK. */
int64_t blast_hip_mixed_work_elems(int64_t n);

/* Size, in floats, of the device scratch behind the allocation-free entry
 * point blast_hip_dsgesv_noalloc_batched (64 MiB). */
#define BLAST_HIP_NOALLOC_SCRATCH_FLOATS (16ll << 20)

/* As blast_hip_dsposv_batched, for n <= 256 and nrhs <= 8, with the FP32
 * factor in `work` (device, batch * blast_hip_mixed_work_elems(n) floats;
 * unused, may be NULL, when n <= 128 and nrhs <= 4).  LAPACK dsposv. */
int blast_hip_dsposv_ws_batched(void* stream, int precision, int itmax, double tol, const blast_matrix_t* A,
                                const blast_matrix_t* B, const blast_matrix_t* X, float* work, int64_t batch,
                                int32_t* info, int32_t* iter);

/* LU with partial pivoting in place of A (column-major storage), pivots to
 * ipiv as blast_hip_dgetrf_batched, then X = A^{-1} B.  LAPACK dgesv; info
 * as getrf (the solve still runs; a zero pivot gives inf/NaN in X). */
int blast_hip_dgesv_batched(void* stream, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info);

/* Mixed precision LU (precision 0: FP32; 1: BF16 products with FP32
 * accumulation in the trailing updates) with FP64 iterative refinement, the
 * LAPACK dsgesv contract: A is read and left unchanged unless the instance
 * falls back, in which case A holds its FP64 LU; ipiv holds the pivots of
 * the factorisation used; iter as dsposv (-1: the low-precision factor hit an
 * exact zero pivot; -2: no convergence in itmax refinements; -3: shape not
 * covered, FP64 path used).  work: device, batch *
 * blast_hip_mixed_work_elems(n) floats. */
int blast_hip_dsgesv_batched(void* stream, int precision, int itmax, double tol, const blast_matrix_t* A,
                             int64_t* ipiv, int64_t ipiv_stride, const blast_matrix_t* B, const blast_matrix_t* X,
                             float* work, int64_t batch, int32_t* info, int32_t* iter);

/* Allocation-free form of blast_hip_dsgesv_batched.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The FP32 factor lives in a device array of the library,
 * BLAST_HIP_NOALLOC_SCRATCH_FLOATS floats (64 MiB) per GPU, held until the
 * process ends; a call whose batch * blast_hip_mixed_work_elems(n) exceeds it
 * returns -10 (batch).  Every call on one GPU uses the same array: calls that
 * overlap on one GPU (other streams, other host threads) race.  Same
 * arguments as blast_hip_dsgesv_batched without `work` (so batch is argument
 * 10, info 11, iter 12); the results equal that entry point's. */
int blast_hip_dsgesv_noalloc_batched(void* stream, int precision, int itmax, double tol, const blast_matrix_t* A,
                                     int64_t* ipiv, int64_t ipiv_stride, const blast_matrix_t* B,
                                     const blast_matrix_t* X, int64_t batch, int32_t* info, int32_t* iter);

/* ---- Uneven batched Cholesky solve (FP64) ----
 * Instances of different sizes in ONE persistent launch: potrf
 * (dense/Potrf.hpp:66-106, panel/Potrf.hpp:59-106), and when b/x are given
 * the two triangular solves (dense/Trsm.hpp:17-66), per instance as
 * blast_hip_dposv_batched.  Each instance brings its own descriptors (any
 * storage; batch_stride is not used).  b.data = x.data = NULL: potrf only.
 * L may alias A and X may alias B. */
typedef struct blast_solve_instance_t {
    blast_matrix_t a, l, b, x;
} blast_solve_instance_t;

/* Largest n the uneven kernels take (192). */
int blast_hip_uneven_max_n(void);

/* Host-side queue order: instance indices sorted by 16-row block count
 * ceil(n/16) descending (largest first, so that the small instances fill
 * the tail of the launch), stable within a block count. */
int blast_hip_uneven_order(int64_t count, const int32_t* n, int32_t* order);

/* inst: DEVICE array of count descriptors.  queue: DEVICE int32 array of
 * count + 1 entries, entries 1..count a permutation of 0..count-1 (e.g.
 * blast_hip_uneven_order's); entry 0 is the work counter, reset by a kernel
 * at the start of every call (no memset: capturable in a hipGraph).  info
 * (optional DEVICE int32[count]): 0, k > 0 the first pivot column that is
 * not positive, -1 an instance whose descriptors are inconsistent or whose
 * n exceeds blast_hip_uneven_max_n() (it is skipped).  Calls that share
 * `queue` must not overlap. */
int blast_hip_dposv_uneven_batched(void* stream, const blast_solve_instance_t* inst, int64_t count, int32_t* queue,
                                   int32_t* info);

/* Uneven batched adjoint of the SPD solve (FP64), from the factor kept by
 * the forward pass: per instance Bbar = A^{-1} Xbar = L^{-T} L^{-1} Xbar and
 * Abar = -fold(Bbar X^T) (lower entries, 0 above), as
 * blast_hip_dposv_backward_batched with recompute = 0, for instances of
 * different n (<= blast_hip_uneven_max_n()) and 1..8 columns of X in ONE
 * persistent launch.  No counterpart in the reference (BLAST has no
 * adjoints); it serves callers that differentiate through uneven batches of
 * blast_hip_dposv_uneven_batched.  Descriptors per instance: l (n x n
 * lower), x and xb (n x r), ab (n x n, written whole), bb (n x r); any
 * storage.  queue and info as for blast_hip_dposv_uneven_batched (info 0,
 * or -1 for an instance whose descriptors are inconsistent). */
typedef struct blast_adjoint_instance_t {
    blast_matrix_t l, x, xb, ab, bb;
} blast_adjoint_instance_t;
int blast_hip_dposv_backward_uneven_batched(void* stream, const blast_adjoint_instance_t* inst, int64_t count,
                                            int32_t* queue, int32_t* info);

/* ---- Uneven batched Riccati recursion (LQR / MPC backward pass), FP64 ----
 * No BLAST function is replaced: this composite strings together the entry
 * points above (gemm -> syrkLower -> potrf -> trsm per step) the way the
 * reference's syrk+potrf benchmarks and BLASFEO wrapper do
 * (bench/blas/SyrkPotrf.cpp:15-57, include/blasfeo/blasfeo_api/SyrkPotrf.hpp:14-39).
 *
 * Instance i has nx states and nu inputs.  Its inputs A (nx x nx), B (nx x nu),
 * Q (nx x nx, SPD) and R (nu x nu, SPD) are dense column-major with leading
 * dimension = rows, at element offsets a, b, q, r of the input buffer `in`.
 * Starting from P_N = Q, for k = N-1 .. 0:
 *   K_k = -(R + B^T P B)^{-1} B^T P A,   P <- Q + A^T P A + A^T P B K_k,
 * computed in square-root form on chol(P).  Outputs in `out`: K_k (nu x nx,
 * column-major, k = 0..N-1 consecutive) at offset k and L_0 = chol(P_0)
 * (nx x nx lower, upper triangle zeroed) at offset l.  `work` offsets
 * index the FP64 scratch buffer (2*nx*(nx+nu) + (nx+nu)^2 elements each).
 * Only Q's and R's lower triangles are read. */
typedef struct blast_riccati_instance_t {
    int32_t nx, nu;
    int64_t a, b, q, r; /* element offsets into `in` */
    int64_t k, l;       /* element offsets into `out` */
    int64_t work;       /* element offset into `work` */
} blast_riccati_instance_t;

/* Host-side plan: fills inst[count] with consecutive offsets, each rounded
 * up to 8 elements (inputs A, B, Q, R per instance in that order; outputs K
 * then L), the queue order (order[],
 * instances sorted by nx bucket nx/16 descending, then by (nx+nu)^3
 * descending, so the tail of a launch holds the small instances), and the
 * total element counts of `in`, `out` and `work`.  1 <= nx, nu <= 4096. */
int blast_hip_riccati_plan(int64_t count, const int32_t* nx, const int32_t* nu, int horizon,
                           blast_riccati_instance_t* inst, int32_t* order, int64_t* in_elems, int64_t* out_elems,
                           int64_t* work_elems);

/* One persistent launch over all `count` instances.  inst: DEVICE array of
 * count descriptors.  queue: DEVICE int32 array of count + 1 entries whose
 * entries 1..count hold the queue order (a permutation of 0..count-1, e.g.
 * the plan's order); entry 0 is the work counter, reset by a kernel at the
 * start of every call (no memset, so the call can be captured in a hipGraph
 * and replayed).  info (optional DEVICE int32[count]): 0; c when chol(Q)
 * meets a pivot that is not positive at column c (1-based); 65536 (k + 1) + c
 * when the first such pivot is in the factorisation of
 * Lam = [[S, .], [., .]] at step k, column c (c <= nu: S = R + B^T P B is not
 * positive definite).  1 <= horizon <= 32767, count <= 2^31 - 4097.  Calls
 * that share `queue` or `work` must not overlap. */
int blast_hip_driccati_batched(void* stream, int horizon, const blast_riccati_instance_t* inst, int64_t count,
                               const double* in, double* out, double* work, int32_t* queue, int32_t* info);

/* Strip-width tuning table (MFMA 16x16 tiles per wavefront in the trailing
 * updates, 1, 2 or 4; 4 runs as 2, the kernel's strips holding at most two
 * tiles) indexed by nx bucket min(nx/16, 15); process-wide. */
int blast_hip_riccati_set_tuning(const int8_t* strip_width16);
int blast_hip_riccati_get_tuning(int8_t* strip_width16);

/* Autotuner: for every nx bucket of the queue order (order_host: the queue's
 * entries 1..count, on the host; inst_host: the descriptors, on the host)
 * times one launch over that bucket's instances with each strip width,
 * `reps` times after a warm-up, and keeps the fastest in the table.
 * Synchronises the stream.  ms_out (optional, host float[48]) receives the
 * average time per bucket and width (index bucket*3 + {0: 1, 1: 2, 2: 4}). */
int blast_hip_riccati_autotune(void* stream, int horizon, const blast_riccati_instance_t* inst,
                               const blast_riccati_instance_t* inst_host, int64_t count, const double* in,
                               double* out, double* work, int32_t* queue, const int32_t* order_host, int reps,
                               float* ms_out);

/* Persist / restore the tuning table (text file, first line names the
 * format and target architecture; a file for another target is rejected). */
int blast_hip_save_tuning(const char* path);
int blast_hip_load_tuning(const char* path);

/* ---- BLASFEO-format entry points ------------------------------------------
 * Batched forms of the BLASFEO routines that BLAST wraps in
 * include/blasfeo/blasfeo_api/ (Gemm.hpp, Syrk.hpp, Potrf.hpp, SyrkPotrf.hpp,
 * Trmm.hpp), on matrices in BLASFEO's panel-major format: element (i, j) of
 * an instance at pA[(i/bs)*bs*cn + i%bs + j*bs].  That is BLAST's panel
 * storage with spacing bs*cn, so each operand becomes a panel view and the
 * same kernels run.  Names and argument order follow BLASFEO's C API
 * (blasfeo_dgemm_nt(m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci,
 * cj, sD, di, dj), ...), with a stream first and the batch count last.
 * Differences from BLASFEO: the panel height bs must be BLAST's (4 for
 * double, 8 for float) and row offsets (ai, bi, ...) multiples of bs;
 * potrf writes 0 in the strict upper part of each bs x bs diagonal block of
 * D (BLAST's panel contract) where BLASFEO leaves it untouched.  Outputs
 * must not overlap inputs, except C == D (same matrix and offsets).
 * Return 0, -k for invalid argument k (1-based, stream = 1), or
 * BLAST_HIP_ERR_RUNTIME. */
typedef struct blast_blasfeo_mat_t {
    void* pA;             /* device pointer to instance 0 (blasfeo_dmat / blasfeo_smat .pA) */
    int64_t m;            /* rows (.m) */
    int64_t n;            /* columns (.n) */
    int64_t cn;           /* padded columns (.cn >= n): panels are bs*cn elements apart */
    int64_t bs;           /* panel height of the BLASFEO build (D_PS / S_PS) */
    int64_t batch_stride; /* elements between consecutive instances */
} blast_blasfeo_mat_t;

/* D := beta*C + alpha*op(A)*op(B), D and C m x n; nn: A m x k, B k x n;
 * nt: B stored n x k; tn: A stored k x m; tt: both.  BLAST wraps these as
 * blasfeo::gemm_nn / gemm_nt / gemm_tn (include/blasfeo/blasfeo_api/Gemm.hpp). */
int blast_hip_blasfeo_dgemm_nn(void* stream, int m, int n, int k, double alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_dgemm_nt(void* stream, int m, int n, int k, double alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_dgemm_tn(void* stream, int m, int n, int k, double alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_dgemm_tt(void* stream, int m, int n, int k, double alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_sgemm_nn(void* stream, int m, int n, int k, float alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_sgemm_nt(void* stream, int m, int n, int k, float alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_sgemm_tn(void* stream, int m, int n, int k, float alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_sgemm_tt(void* stream, int m, int n, int k, float alpha, const blast_blasfeo_mat_t* sA,
                               int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);

/* D := beta*C + alpha*A*B^T, lower triangle of the m x m D written; A and B
 * stored m x k.  blasfeo::syrk_ln (include/blasfeo/blasfeo_api/Syrk.hpp). */
int blast_hip_blasfeo_dsyrk_ln(void* stream, int m, int k, double alpha, const blast_blasfeo_mat_t* sA, int ai,
                               int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);
int blast_hip_blasfeo_ssyrk_ln(void* stream, int m, int k, float alpha, const blast_blasfeo_mat_t* sA, int ai,
                               int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch);

/* D := chol(C), lower, m x m (C read from its lower triangle; D may be C).
 * info as blast_hip_dpotrf_batched.  blasfeo::potrf
 * (include/blasfeo/blasfeo_api/Potrf.hpp, blasfeo_dpotrf_l). */
int blast_hip_blasfeo_dpotrf_l(void* stream, int m, const blast_blasfeo_mat_t* sC, int ci, int cj,
                               const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch, int32_t* info);
int blast_hip_blasfeo_spotrf_l(void* stream, int m, const blast_blasfeo_mat_t* sC, int ci, int cj,
                               const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch, int32_t* info);

/* D := chol(C + A*B^T), lower, m x m; A and B stored m x k.
 * blasfeo::syrk_potrf (include/blasfeo/blasfeo_api/SyrkPotrf.hpp,
 * blasfeo_dsyrk_dpotrf_ln): the lower syrk into D, then potrf in place. */
int blast_hip_blasfeo_dsyrk_dpotrf_ln(void* stream, int m, int k, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                      const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sC,
                                      int ci, int cj, const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch,
                                      int32_t* info);
int blast_hip_blasfeo_ssyrk_spotrf_ln(void* stream, int m, int k, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                      const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sC,
                                      int ci, int cj, const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch,
                                      int32_t* info);

/* rlnn: D := alpha*B*A, A lower n x n (its strict upper part not read);
 * rutn: D := alpha*B*A^T, A upper n x n (strict lower part not read); B and
 * D m x n.  blasfeo::trmm_rlnn / trmm_rutn (include/blasfeo/blasfeo_api/Trmm.hpp). */
int blast_hip_blasfeo_dtrmm_rlnn(void* stream, int m, int n, double alpha, const blast_blasfeo_mat_t* sA, int ai,
                                 int aj, const blast_blasfeo_mat_t* sB, int bi, int bj,
                                 const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch);
int blast_hip_blasfeo_dtrmm_rutn(void* stream, int m, int n, double alpha, const blast_blasfeo_mat_t* sA, int ai,
                                 int aj, const blast_blasfeo_mat_t* sB, int bi, int bj,
                                 const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch);
int blast_hip_blasfeo_strmm_rlnn(void* stream, int m, int n, float alpha, const blast_blasfeo_mat_t* sA, int ai,
                                 int aj, const blast_blasfeo_mat_t* sB, int bi, int bj,
                                 const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch);
int blast_hip_blasfeo_strmm_rutn(void* stream, int m, int n, float alpha, const blast_blasfeo_mat_t* sA, int ai,
                                 int aj, const blast_blasfeo_mat_t* sB, int bi, int bj,
                                 const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch);

#ifdef __cplusplus
}
#endif

#endif /* BLAST_HIP_H */
