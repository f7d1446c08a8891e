// Reverse-mode (adjoint) kernels for the batched factorisations and solves,
// and the getrf-based solve they differentiate.  These operations are not
// in BLAST; the closed forms are the standard matrix-derivative results
// (Giles 2008, "An extended collection of matrix derivative results";
// Murray 2016, "Differentiation of the Cholesky decomposition").
//
// Symmetric inputs (potrf, posv) are read from their lower triangle, so
// their adjoint is taken with respect to the stored lower entries: a change
// of stored a_ij (i > j) changes A_ij and A_ji.  For a full-matrix gradient
// G this gives Abar_ij = G_ij + G_ji (i > j), Abar_ii = G_ii; the strict
// upper part of Abar is written as 0.
//
// Each backward pass is composed from the batched kernels of this library
// plus the small elementwise kernels below; all stream-ordered, no
// allocation (the caller's output buffers double as workspace).

#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {

constexpr int kEw = 256;

// M := Phi(M): lower triangle with the diagonal halved, strict upper zeroed.
template <typename T>
__global__ __launch_bounds__(kEw) void phi_lower_kernel(Mat<T> Mm) {
    Ref<T> m(Mm, blockIdx.x);
    int const n = Mm.rows;
    for (int e = threadIdx.x; e < n * n; e += kEw) {
        int const i = e % n, j = e / n;
        T const v = m(i, j);
        m(i, j) = i > j ? v : (i == j ? T(0.5) * v : T(0));
    }
}

// Fold a full gradient G onto stored lower entries, in place:
// G_ij + G_ji (i > j), G_ii, 0 above.  Each thread owns one pair (i, j), i >= j.
template <typename T>
__global__ __launch_bounds__(kEw) void fold_lower_kernel(Mat<T> Gm) {
    Ref<T> g(Gm, blockIdx.x);
    int const n = Gm.rows;
    for (int e = threadIdx.x; e < n * n; e += kEw) {
        int const i = e % n, j = e / n;
        if (i > j) {
            T const s = g(i, j) + g(j, i);
            g(i, j) = s;
            g(j, i) = T(0);
        }
    }
}

// Abar := -(Bbar X^T + X Bbar^T) folded onto lower entries, i.e.
// -(sum_c bb_ic x_jc + bb_jc x_ic) below the diagonal, -sum_c bb_ic x_ic on it.
// Workgroup (instance, y) writes columns 8y..8y+7, one row per thread.
constexpr int kFoldCols = 8;
template <typename T>
__global__ __launch_bounds__(kEw) void sym_outer_lower_kernel(Mat<T> Bb, Mat<T> Xm, Mat<T> Ab) {
    int64_t const b = blockIdx.x;
    Ref<T> bb(Bb, b), x(Xm, b), a(Ab, b);
    int const n = Ab.rows, r = Xm.cols, j0 = kFoldCols * int(blockIdx.y);
    for (int i = int(threadIdx.x); i < n; i += kEw)
        for (int j = j0; j < j0 + kFoldCols && j < n; ++j) {
            T v = T(0);
            if (i > j)
                for (int c = 0; c < r; ++c) v = v - (bb(i, c) * x(j, c) + bb(j, c) * x(i, c));
            else if (i == j)
                for (int c = 0; c < r; ++c) v = v - bb(i, c) * x(i, c);
            a(i, j) = v;
        }
}

// Zero the strict upper (upper = 0) or strict lower (upper = 1) triangle.
template <typename T>
__global__ __launch_bounds__(kEw) void zero_triangle_kernel(Mat<T> Mm, int keep_upper) {
    Ref<T> m(Mm, blockIdx.x);
    int const rows = Mm.rows, cols = Mm.cols;
    for (int e = threadIdx.x; e < rows * cols; e += kEw) {
        int const i = e % rows, j = e / rows;
        if (keep_upper ? i > j : i < j) m(i, j) = T(0);
    }
}

// X := alpha * X
template <typename T>
__global__ __launch_bounds__(kEw) void scale_kernel(Mat<T> Mm, T alpha) {
    Ref<T> m(Mm, blockIdx.x);
    int const rows = Mm.rows, cols = Mm.cols;
    for (int e = threadIdx.x; e < rows * cols; e += kEw) {
        int const i = e % rows, j = e / rows;
        m(i, j) = alpha * m(i, j);
    }
}

// Row interchanges of getrf's ipiv on B (laswp, dense/Laswp.hpp:37-43):
// forward applies k = 0..K-1 (P B), backward k = K-1..0 (P^T B).
template <typename T>
__global__ __launch_bounds__(kEw) void laswp_kernel(Mat<T> Bm, const int64_t* ipiv, int64_t ipiv_stride, int K,
                                                    int backward) {
    int64_t const b = blockIdx.x;
    Ref<T> m(Bm, b);
    int64_t const* piv = ipiv + b * ipiv_stride;
    for (int j = threadIdx.x; j < Bm.cols; j += kEw)
        for (int s = 0; s < K; ++s) {
            int const k = backward ? K - 1 - s : s;
            int const p = int(piv[k]);
            if (p != k) {
                T const t = m(k, j);
                m(k, j) = m(p, j);
                m(p, j) = t;
            }
        }
}

// B := copy of A (same logical shape)
template <typename T>
__global__ __launch_bounds__(kEw) void copy_kernel(Mat<T> Am, Mat<T> Bm) {
    Ref<T> a(Am, blockIdx.x), bb(Bm, blockIdx.x);
    int const rows = Am.rows, cols = Am.cols;
    for (int e = threadIdx.x; e < rows * cols; e += kEw) {
        int const i = e % rows, j = e / rows;
        bb(i, j) = a(i, j);
    }
}

#define LAUNCH_EW(kernel, batch, ...)                                                         \
    do {                                                                                      \
        if ((batch) > 0) {                                                                    \
            hipLaunchKernelGGL(kernel, dim3(unsigned(batch)), dim3(kEw), 0, s, __VA_ARGS__); \
            int const r_ = check_launch(#kernel);                                             \
            if (r_) return r_;                                                                \
        }                                                                                     \
    } while (0)

#define TRY(expr)             \
    do {                      \
        int const r_ = (expr); \
        if (r_) return r_;    \
    } while (0)

// gemm: D = alpha A B + beta C.  Abar = alpha Dbar B^T, Bbar = alpha A^T Dbar,
// Cbar = beta Dbar.  Any output may be null (descriptor data == nullptr).
template <typename T>
int gemm_backward(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> Db, Mat<T> Ab, Mat<T> Bb, Mat<T> Cb,
                  int64_t batch) {
    if (Ab.data) TRY(launch_gemm_overwrite<T>(s, alpha, Db, transposed(B), Ab, batch));
    if (Bb.data) TRY(launch_gemm_overwrite<T>(s, alpha, transposed(A), Db, Bb, batch));
    if (Cb.data) {
        LAUNCH_EW((copy_kernel<T>), batch, Db, Cb);
        LAUNCH_EW((scale_kernel<T>), batch, Cb, beta);
    }
    return 0;
}

// trsm (left, non-unit): A X = B with A lower (upper = 0) or upper.
// Bbar = A^{-T} Xbar, Abar = -Bbar X^T restricted to A's triangle.
template <typename T>
int trsm_backward(hipStream_t s, int upper, Mat<T> A, Mat<T> X, Mat<T> Xb, Mat<T> Ab, Mat<T> Bb, int64_t batch) {
    TRY(dispatch_trsm<T>(s, 0, !upper, 0, T(1), transposed(A), Xb, Bb, batch));
    if (Ab.data) {
        TRY(launch_gemm_overwrite<T>(s, T(-1), Bb, transposed(X), Ab, batch));
        LAUNCH_EW((zero_triangle_kernel<T>), batch, Ab, upper);
    }
    return 0;
}

// potrf: A = L L^T.  Abar (lower entries) from Lbar (lower):
// G = L^{-T} Phi(L^T Lbar) L^{-1}, folded onto the lower entries.
template <typename T>
int potrf_backward(hipStream_t s, Mat<T> L, Mat<T> Lb, Mat<T> Ab, int64_t batch) {
    // Ab := L^T tril(Lbar), Phi, then the two solves in place.
    TRY(launch_upper_times_lower<T>(s, T(1), transposed(L), Lb, Ab, batch));
    LAUNCH_EW((phi_lower_kernel<T>), batch, Ab);
    TRY(dispatch_trsm<T>(s, 1, 0, 0, T(1), L, Ab, Ab, batch));              // Y L = P  (right, lower)
    TRY(dispatch_trsm<T>(s, 0, 1, 0, T(1), transposed(L), Ab, Ab, batch));  // L^T G = Y (left, upper)
    LAUNCH_EW((fold_lower_kernel<T>), batch, Ab);
    return 0;
}

// Cholesky solve x = A^{-1} b: Bbar = A^{-1} Xbar, Abar = -sym(Bbar X^T)
// on lower entries.  recompute = 0: F holds the factor L saved by the
// forward pass; recompute = 1: F holds A, the factor is recomputed into
// Abar (which then receives the gradient) - no extra memory.
template <typename T>
int posv_backward(hipStream_t s, int recompute, Mat<T> F, Mat<T> X, Mat<T> Xb, Mat<T> Ab, Mat<T> Bb, int64_t batch,
                  int32_t* info) {
    if (recompute) {
        TRY(dispatch_posv<T>(s, F, Ab, Xb, Bb, batch, info));  // L -> Abar, Bbar = A^{-1} Xbar
    } else {
        TRY(dispatch_trsm<T>(s, 0, 0, 0, T(1), F, Xb, Bb, batch));
        TRY(dispatch_trsm<T>(s, 0, 1, 0, T(1), transposed(F), Bb, Bb, batch));
    }
    if (batch > 0 && Ab.rows > 0) {
        unsigned const groups = unsigned((Ab.rows + kFoldCols - 1) / kFoldCols);
        hipLaunchKernelGGL((sym_outer_lower_kernel<T>), dim3(unsigned(batch), groups), dim3(kEw), 0, s, Bb, X, Ab);
        int const r_ = check_launch("sym_outer_lower_kernel");
        if (r_) return r_;
    }
    return 0;
}

// getrs: solve A X = B from getrf's in-place LU and pivots:
// X = U^{-1} L^{-1} P B.
template <typename T>
int getrs(hipStream_t s, Mat<T> LU, const int64_t* ipiv, int64_t ipiv_stride, Mat<T> B, Mat<T> X, int64_t batch) {
    LAUNCH_EW((copy_kernel<T>), batch, B, X);
    LAUNCH_EW((laswp_kernel<T>), batch, X, ipiv, ipiv_stride, LU.rows, 0);
    TRY(dispatch_trsm<T>(s, 0, 0, 1, T(1), LU, X, X, batch));  // unit lower L
    TRY(dispatch_trsm<T>(s, 0, 1, 0, T(1), LU, X, X, batch));  // upper U
    return 0;
}

// getrs adjoint: Bbar = A^{-T} Xbar = P^T L^{-T} U^{-T} Xbar,
// Abar = -Bbar X^T (general matrix).
template <typename T>
int getrs_backward(hipStream_t s, Mat<T> LU, const int64_t* ipiv, int64_t ipiv_stride, Mat<T> X, Mat<T> Xb,
                   Mat<T> Ab, Mat<T> Bb, int64_t batch) {
    TRY(dispatch_trsm<T>(s, 0, 0, 0, T(1), transposed(LU), Xb, Bb, batch));  // U^T (lower, non-unit)
    TRY(dispatch_trsm<T>(s, 0, 1, 1, T(1), transposed(LU), Bb, Bb, batch));  // L^T (upper, unit)
    LAUNCH_EW((laswp_kernel<T>), batch, Bb, ipiv, ipiv_stride, LU.rows, 1);
    if (Ab.data) TRY(launch_gemm_overwrite<T>(s, T(-1), Bb, transposed(X), Ab, batch));
    return 0;
}

#define INSTANTIATE(T)                                                                                          \
    template int gemm_backward<T>(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, Mat<T>, Mat<T>, int64_t); \
    template int trsm_backward<T>(hipStream_t, int, Mat<T>, Mat<T>, Mat<T>, Mat<T>, Mat<T>, int64_t);          \
    template int potrf_backward<T>(hipStream_t, Mat<T>, Mat<T>, Mat<T>, int64_t);                              \
    template int posv_backward<T>(hipStream_t, int, Mat<T>, Mat<T>, Mat<T>, Mat<T>, Mat<T>, int64_t, int32_t*); \
    template int getrs<T>(hipStream_t, Mat<T>, const int64_t*, int64_t, Mat<T>, Mat<T>, int64_t);              \
    template int getrs_backward<T>(hipStream_t, Mat<T>, const int64_t*, int64_t, Mat<T>, Mat<T>, Mat<T>, Mat<T>, \
                                   int64_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace blast_hip
