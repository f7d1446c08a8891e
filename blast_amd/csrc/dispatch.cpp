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

// ---- Blocked Cholesky path for n > 256 (the MFMA kernels above cover
// n <= 256): 256-row blocks, each the fused n <= 256 kernels, trsm_tiles and
// the syrk / gemm tile kernels on sub-blocks.  Stream-ordered launches, no
// allocation.  (BLAST's own blocked algorithms: dense/Potrf.hpp:66-106 and
// dense/Trsm.hpp:17-66.)
constexpr int kBlockedNb = 256;

// Rows i0.., columns j0.. (logical) of every instance: the same storage
// with the base moved; in panel storage the first storage row must lie on a
// panel boundary (i0, j0 are multiples of 256 here).
template <typename T>
Mat<T> subview(Mat<T> m, int i0, int j0, int rows, int cols) {
    constexpr int PS = PanelSize<T>::value;
    int const si = m.trans ? j0 : i0, sj = m.trans ? i0 : j0;
    int64_t const off = m.panel ? int64_t(si / PS) * m.spacing + si % PS + int64_t(sj) * PS
                                : int64_t(si) + int64_t(m.spacing) * sj;
    Mat<T> r = m;
    r.data = m.data + off;
    r.rows = rows;
    r.cols = cols;
    return r;
}

// X = A^{-1} B for A lower (upper = 0) or upper triangular, n > 256, by
// blocks: the diagonal solve on trsm_tiles, the rest of the right-hand sides
// updated by gemm.  X may alias B.
template <typename T>
int trsm_left_blocked(hipStream_t s, int upper, int unit, T alpha, Mat<T> A, Mat<T> B, Mat<T> X, int64_t batch) {
    int const n = A.rows, nrhs = X.cols, NB = kBlockedNb;
    if (X.data != B.data || alpha != T(1)) {
        int const r = launch_copy_scaled<T>(s, B, X, alpha, batch);
        if (r) return r;
    }
    int const nblk = (n + NB - 1) / NB;
    for (int t = 0; t < nblk; ++t) {
        int const kb = upper ? nblk - 1 - t : t, k0 = kb * NB, nb = n - k0 < NB ? n - k0 : NB;
        Mat<T> const Akk = subview(A, k0, k0, nb, nb), Xk = subview(X, k0, 0, nb, nrhs);
        int r = launch_trsm_tiles<T>(s, upper, unit, T(1), Akk, Xk, Xk, batch);
        if (r) return r;
        if (!upper && k0 + nb < n) {  // rows below: X2 -= A21 X1
            int const rest = n - k0 - nb;
            Mat<T> const X2 = subview(X, k0 + nb, 0, rest, nrhs);
            r = launch_gemm_tiles<T>(s, T(-1), subview(A, k0 + nb, k0, rest, nb), Xk, T(1), X2, X2, batch);
        } else if (upper && k0 > 0) {  // rows above: X0 -= A01 X1
            Mat<T> const X0 = subview(X, 0, 0, k0, nrhs);
            r = launch_gemm_tiles<T>(s, T(-1), subview(A, 0, k0, k0, nb), Xk, T(1), X0, X0, batch);
        }
        if (r) return r;
    }
    return 0;
}

// L = chol(A), n > 256, right-looking by 256-column blocks: the diagonal
// block on the fused n <= 256 kernel, the panel by a right-side trsm
// (L21 L11^T = A21), the trailing matrix by syrk.  The lower triangle of A
// is copied into L first (unless L is A); the rest of L is left as the
// kernels leave it for n <= 256 (zeros in the 4x4 diagonal panel blocks).
// info: from the factor's diagonal (a pivot that is not positive leaves a
// NaN or non-positive entry there and NaN in the later ones).
template <typename T>
int potrf_blocked(hipStream_t s, Mat<T> A, Mat<T> L, int64_t batch, int32_t* info) {
    int const n = A.rows, NB = kBlockedNb;
    int r = 0;
    if (A.data != L.data) r = launch_copy_lower<T>(s, A, L, batch);
    if (r) return r;
    Mat<T> const none{};
    for (int k0 = 0; k0 < n; k0 += NB) {
        int const nb = n - k0 < NB ? n - k0 : NB, rest = n - k0 - nb;
        Mat<T> const L11 = subview(L, k0, k0, nb, nb);
        if constexpr (std::is_same_v<T, double>)
            r = launch_chol_fused(s, L11, L11, none, none, batch, nullptr);
        else
            r = launch_chol_owner_f32(s, L11, L11, none, none, batch, nullptr);
        if (r) return r < 0 ? r : BLAST_HIP_ERR_RUNTIME;
        if (rest > 0) {
            Mat<T> const L21 = subview(L, k0 + nb, k0, rest, nb), L22 = subview(L, k0 + nb, k0 + nb, rest, rest);
            // X L11^T = L21  <=>  L11 X^T = L21^T
            r = launch_trsm_tiles<T>(s, 0, 0, T(1), L11, transposed(L21), transposed(L21), batch);
            if (r) return r;
            r = launch_syrk_tiles<T>(s, T(-1), L21, T(1), L22, L22, batch);
            if (r) return r;
        }
    }
    return launch_diag_info<T>(s, L, batch, info);
}

// LU with partial pivoting, square 256 < n <= 1024, right-looking by
// column blocks of the panel kernel's width (24 FP64, 32 FP32): the panel
// kernel (factor, interchanges on the columns to its right, U12), then
// A22 -= L21 U12 on the gemm tile kernel; the interchanges of the columns
// left of each panel in one pass at the end.  info is cleared first; an
// instance whose panel meets a zero pivot records it and its later panels do
// nothing (BLAST stops there, dense/Getrf.hpp:52-119).
template <typename T>
int getrf_blocked(hipStream_t s, Mat<T> A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    int const n = A.rows, kLuPanel = getrf_panel_width<T>();
    int r = info ? launch_fill_i32(s, info, 0, batch) : 0;
    for (int k0 = 0; k0 < n && r == 0; k0 += kLuPanel) {
        int const nb = n - k0 < kLuPanel ? n - k0 : kLuPanel, rest = n - k0 - nb;
        r = launch_getrf_panel<T>(s, A, k0, nb, ipiv, ipiv_stride, batch, info);
        if (r == 0 && rest > 0) {
            Mat<T> const A22 = subview(A, k0 + nb, k0 + nb, rest, rest);
            r = launch_gemm_tiles<T>(s, T(-1), subview(A, k0 + nb, k0, rest, nb), subview(A, k0, k0 + nb, nb, rest),
                                     T(1), A22, A22, batch);
        }
    }
    if (r == 0) r = launch_getrf_left_swaps<T>(s, A, ipiv, ipiv_stride, batch, info);
    return r;
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
        if (A.rows == A.cols && A.rows > 16 * 16)  // n > 256: by 256-row blocks
            return side_right ? trsm_left_blocked<T>(s, !upper, unit, alpha, transposed(A), transposed(B),
                                                     transposed(X), batch)
                              : trsm_left_blocked<T>(s, upper, unit, alpha, A, B, X, batch);
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
        if (A.rows == A.cols && A.rows > kBlockedNb) return potrf_blocked<T>(s, A, L, batch, info);
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
    if (algorithm() == BLAST_HIP_ALGO_AUTO && getrf_panel_covers(A.rows, A.cols, A.trans))
        return getrf_blocked<T>(s, A, ipiv, ipiv_stride, batch, info);
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
        if (A.rows == A.cols && A.rows > kBlockedNb) {
            r = potrf_blocked<T>(s, A, L, batch, info);
            if (r) return r;
            r = trsm_left_blocked<T>(s, 0, 0, T(1), L, B, X, batch);
            if (r) return r;
            return trsm_left_blocked<T>(s, 1, 0, T(1), transposed(L), X, X, batch);
        }
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
    int const r = dispatch_getrf<double>(s, A, ipiv, ipiv_stride, batch, info);
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
