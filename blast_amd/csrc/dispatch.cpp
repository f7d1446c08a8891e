// Chooses the kernel for each call.  BLAST_HIP_ALGO_REFERENCE always takes the
// reference-order kernels; AUTO takes the MFMA / fused kernels where the shape
// and layout qualify.
#include "dispatch.hpp"

#include <type_traits>

namespace blast_hip {

// Fused FP64 Cholesky (+ solve): the one-wavefront kernel for full-tile
// panel shapes, the workgroup kernels otherwise.  Returns 1 when no fused
// kernel covers the shape.
int launch_chol_fused(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if) {
    int r = launch_posv_tiles(s, A, L, B, X, batch, info, only_if);
    if (r != 1) return r;
    r = launch_chol_team(s, A, L, B, X, batch, info, only_if);
    if (r != 1) return r;
    return launch_chol_owner(s, A, L, B, X, batch, info, only_if);
}

template <typename T>
int dispatch_gemm(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO) return launch_gemm_tiles<T>(s, alpha, A, B, beta, C, D, batch);
    return launch_gemm_ref<T>(s, alpha, A, B, beta, C, D, batch);
}

template <typename T>
int dispatch_syrk(hipStream_t s, T alpha, Mat<T> A, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO) return launch_syrk_tiles<T>(s, alpha, A, beta, C, D, batch);
    return launch_syrk_ref<T>(s, alpha, A, beta, C, D, batch);
}

template <typename T>
int dispatch_trmm(hipStream_t s, int side_right, T alpha, Mat<T> A, Mat<T> B, Mat<T> C, int64_t batch) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO) return launch_trmm_tiles<T>(s, side_right, alpha, A, B, C, batch);
    return launch_trmm_ref<T>(s, side_right, alpha, A, B, C, batch);
}

template <typename T>
int dispatch_trsm(hipStream_t s, int side_right, int upper, int unit, T alpha, Mat<T> A, Mat<T> B, Mat<T> X,
                  int64_t batch) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO) {
        // X A = alpha B  <=>  A^T X^T = alpha B^T (A^T has the other triangle)
        int const r = side_right ? launch_trsm_tiles<T>(s, !upper, unit, alpha, transposed(A), transposed(B),
                                                        transposed(X), batch)
                                 : launch_trsm_tiles<T>(s, upper, unit, alpha, A, B, X, batch);
        if (r != 1) return r;
    }
    return launch_trsm_ref<T>(s, side_right, upper, unit, alpha, A, B, X, batch);
}

template <typename T>
int dispatch_potrf(hipStream_t s, Mat<T> A, Mat<T> L, int64_t batch, int32_t* info) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO) {
        Mat<T> none{};
        int r;
        if constexpr (std::is_same_v<T, double>)
            r = launch_chol_fused(s, A, L, none, none, batch, info);
        else
            r = launch_chol_owner_f32(s, A, L, none, none, batch, info);
        if (r != 1) return r;
    }
    return launch_potrf_ref<T>(s, A, L, batch, info);
}

template <typename T>
int dispatch_getrf(hipStream_t s, Mat<T> A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    // column-major storage (partial pivoting); the row-major form has no pivoting
    if (algorithm() == BLAST_HIP_ALGO_AUTO && !A.trans && A.rows == A.cols && wg_covers(A.rows, 0)) {
        if constexpr (std::is_same_v<T, double>) {
            Mat<double> none{};
            if (lu_wg_fits(A, none, 0)) return launch_wg_gesv(s, A, ipiv, ipiv_stride, none, none, 0, batch, info);
        } else {
            if (lu_wg_fits_f32(A)) return launch_wg_getrf_f32(s, A, ipiv, ipiv_stride, batch, info);
        }
    }
    return launch_getrf_ref<T>(s, A, ipiv, ipiv_stride, batch, info);
}

template <typename T>
int dispatch_posv(hipStream_t s, Mat<T> A, Mat<T> L, Mat<T> B, Mat<T> X, int64_t batch, int32_t* info) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO && X.cols > 0) {
        int r;
        if constexpr (std::is_same_v<T, double>)
            r = launch_chol_fused(s, A, L, B, X, batch, info);
        else
            r = launch_chol_owner_f32(s, A, L, B, X, batch, info);
        if (r != 1) return r;
    }
    int r = launch_potrf_ref<T>(s, A, L, batch, info);
    if (r) return r;
    r = launch_trsm_ref<T>(s, 0, 0, 0, T(1), L, B, X, batch);
    if (r) return r;
    return launch_trsm_ref<T>(s, 0, 1, 0, T(1), transposed(L), X, X, batch);
}

// Mixed precision + FP64 refinement; instances that do not converge (or
// shapes the mixed kernel does not cover) are solved by the FP64 path with
// the factor written into A, as LAPACK dsposv does.
int dispatch_dsposv(hipStream_t s, int bf16, int itmax, double tol, Mat<double> A, Mat<double> B, Mat<double> X,
                    int64_t batch, int32_t* info, int32_t* iter) {
    int r = launch_sposv_ir(s, bf16, A, B, X, itmax, tol, batch, info, iter);
    if (r == 0) return launch_chol_fused(s, A, A, B, X, batch, info, iter);
    if (r != 1) return r;
    r = launch_fill_i32(s, iter, -3, batch);
    if (r) return r;
    return dispatch_posv<double>(s, A, A, B, X, batch, info);
}

// As dispatch_dsposv, for n <= 256 and nrhs <= 8 with a caller-provided FP32
// factor buffer (LAPACK dsposv's SWORK).
int dispatch_dsposv_ws(hipStream_t s, int bf16, int itmax, double tol, Mat<double> A, Mat<double> B, Mat<double> X,
                       float* work, int64_t batch, int32_t* info, int32_t* iter) {
    if (A.rows <= 128 && X.cols <= 4) return dispatch_dsposv(s, bf16, itmax, tol, A, B, X, batch, info, iter);
    if (wg_covers(A.rows, X.cols) && work) {
        int const r = launch_wg_mixed(s, 0, bf16, itmax, tol, A, nullptr, 0, B, X, X.cols, work, batch, info, iter);
        if (r) return r;
        return launch_chol_fused(s, A, A, B, X, batch, info, iter);  // covers every shape wg_covers accepts
    }
    int const r = launch_fill_i32(s, iter, -3, batch);
    if (r) return r;
    return dispatch_posv<double>(s, A, A, B, X, batch, info);
}

// LU with partial pivoting in place of A and the solve (LAPACK dgesv).
int dispatch_dgesv(hipStream_t s, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride, Mat<double> B, Mat<double> X,
                   int64_t batch, int32_t* info) {
    if (algorithm() == BLAST_HIP_ALGO_AUTO && wg_covers(A.rows, X.cols) && lu_wg_fits(A, X, X.cols))
        return launch_wg_gesv(s, A, ipiv, ipiv_stride, B, X, X.cols, batch, info);
    int const r = launch_getrf_ref<double>(s, A, ipiv, ipiv_stride, batch, info);
    if (r) return r;
    return getrs<double>(s, A, ipiv, ipiv_stride, B, X, batch);
}

// Mixed-precision LU + FP64 refinement (LAPACK dsgesv): instances that do
// not converge are solved by the FP64 LU in place of A.
int dispatch_dsgesv(hipStream_t s, int bf16, int itmax, double tol, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride,
                    Mat<double> B, Mat<double> X, float* work, int64_t batch, int32_t* info, int32_t* iter) {
    if (wg_covers(A.rows, X.cols) && work && lu_wg_fits(A, X, X.cols)) {
        int const r =
            launch_wg_mixed(s, 1, bf16, itmax, tol, A, ipiv, ipiv_stride, B, X, X.cols, work, batch, info, iter);
        if (r) return r;
        return launch_wg_gesv(s, A, ipiv, ipiv_stride, B, X, X.cols, batch, info, iter);
    }
    int const r = launch_fill_i32(s, iter, -3, batch);
    if (r) return r;
    return dispatch_dgesv(s, A, ipiv, ipiv_stride, B, X, batch, info);
}

#define INSTANTIATE(T)                                                                              \
    template int dispatch_gemm<T>(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);     \
    template int dispatch_syrk<T>(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);             \
    template int dispatch_trmm<T>(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);           \
    template int dispatch_trsm<T>(hipStream_t, int, int, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t); \
    template int dispatch_potrf<T>(hipStream_t, Mat<T>, Mat<T>, int64_t, int32_t*);                \
    template int dispatch_getrf<T>(hipStream_t, Mat<T>, int64_t*, int64_t, int64_t, int32_t*);     \
    template int dispatch_posv<T>(hipStream_t, Mat<T>, Mat<T>, Mat<T>, Mat<T>, int64_t, int32_t*);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace blast_hip
