// Kernel launchers and the per-operation dispatch between the
// reference-order kernels and the MFMA / fused kernels.
#pragma once

#include "common.hpp"

namespace blast_hip {

int algorithm();

// reference_order.hip
template <typename T> int launch_gemm_ref(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_syrk_ref(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_trmm_ref(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_trsm_ref(hipStream_t, int, int, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_potrf_ref(hipStream_t, Mat<T>, Mat<T>, int64_t, int32_t*);
template <typename T> int launch_getrf_ref(hipStream_t, Mat<T>, int64_t*, int64_t, int64_t, int32_t*);

// gemm_tiles.hip (MFMA tiles, any size and layout)
template <typename T> int launch_gemm_tiles(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_syrk_tiles(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_syrk2_tiles(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_trmm_tiles(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_upper_times_lower(hipStream_t, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int launch_gemm_overwrite(hipStream_t, T, Mat<T>, Mat<T>, Mat<T>, int64_t);

// trsm_tiles.hip: left solve, square A, n <= 256; returns 1 when not covered
template <typename T>
int launch_trsm_tiles(hipStream_t, int upper, int unit, T alpha, Mat<T> A, Mat<T> B, Mat<T> X, int64_t batch);

// adjoint.hip
template <typename T>
int gemm_backward(hipStream_t, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> Db, Mat<T> Ab, Mat<T> Bb, Mat<T> Cb,
                  int64_t batch);
template <typename T>
int trsm_backward(hipStream_t, int upper, Mat<T> A, Mat<T> X, Mat<T> Xb, Mat<T> Ab, Mat<T> Bb, int64_t batch);
template <typename T> int potrf_backward(hipStream_t, Mat<T> L, Mat<T> Lb, Mat<T> Ab, int64_t batch);
template <typename T>
int posv_backward(hipStream_t, int recompute, Mat<T> F, Mat<T> X, Mat<T> Xb, Mat<T> Ab, Mat<T> Bb, int64_t batch,
                  int32_t* info);
template <typename T>
int getrs(hipStream_t, Mat<T> LU, const int64_t* ipiv, int64_t ipiv_stride, Mat<T> B, Mat<T> X, int64_t batch);
template <typename T>
int getrs_backward(hipStream_t, Mat<T> LU, const int64_t* ipiv, int64_t ipiv_stride, Mat<T> X, Mat<T> Xb, Mat<T> Ab,
                   Mat<T> Bb, int64_t batch);

// posv_ll.hip: one wavefront per instance (two per SIMD), n = 16 NT <= 128, A and L in
// panel storage, nrhs <= 4; returns 1 when the shape is not covered.  only_if
// as launch_chol_team.
int launch_posv_tiles(hipStream_t, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if = nullptr);
// dispatch.cpp: the first of posv_tiles, chol_team, chol_owner that covers
// the shape; 1 when none does.
int launch_chol_fused(hipStream_t, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if = nullptr);
// chol_owner.hip: fused FP64 Cholesky (+ solve) for n <= 256, nrhs <= 8;
// returns 1 when the shape is not covered.  only_if as launch_chol_team.
int launch_chol_owner(hipStream_t, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if = nullptr);
// chol_owner.hip: the same for FP32 storage with FP64 arithmetic (loads
// widened, stores rounded); n <= 256, nrhs <= 8, 1 when not covered.
int launch_chol_owner_f32(hipStream_t, Mat<float> A, Mat<float> L, Mat<float> B, Mat<float> X, int64_t batch,
                          int32_t* info);
// chol_team.hip: fused FP64 Cholesky (+ solve) for n <= 128, nrhs <= 4, one
// workgroup per instance; returns 1 when the shape is not covered.  only_if:
// when non-null, only instances b with only_if[b] < 0 are processed.
int launch_chol_team(hipStream_t, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                     int32_t* info, const int32_t* only_if = nullptr);
// chol_owner.hip: uneven batch, one persistent launch (queue: count + 1
// int32, entries 1..count the order); n <= uneven_max_n().
int launch_posv_uneven(hipStream_t, const blast_solve_instance_t* inst, int64_t count, int32_t* queue,
                       int32_t* info, int grid);
int uneven_max_n();
// adjoint_uneven.hip: the posv adjoint over an uneven batch, one launch.
int launch_posv_backward_uneven(hipStream_t, const blast_adjoint_instance_t* inst, int64_t count, int32_t* queue,
                                int32_t* info, int grid);
// sposv_ir.hip: returns 1 when the shape is not covered
int launch_sposv_ir(hipStream_t, int bf16, Mat<double> A, Mat<double> B, Mat<double> X, int itmax, double tol,
                    int64_t batch, int32_t* info, int32_t* iter);
int launch_fill_i32(hipStream_t, int32_t* p, int32_t v, int64_t count);
int dispatch_dsposv(hipStream_t, int bf16, int itmax, double tol, Mat<double> A, Mat<double> B, Mat<double> X,
                    int64_t batch, int32_t* info, int32_t* iter);

// convert.hip: D := S between any two storage layouts (same logical size)
template <typename T> int launch_convert(hipStream_t, Mat<T> S, Mat<T> D, int64_t batch);

// wg_solve.hip (mixed precision) and lu_wg.hip (FP64 LU): one workgroup per
// instance, square n <= 256, nrhs <= 8.
bool wg_covers(int n, int nrhs);
int64_t wg_work_elems(int n);  // floats per instance of the mixed-precision factor
int launch_wg_gesv(hipStream_t, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride, Mat<double> B, Mat<double> X,
                   int nrhs, int64_t batch, int32_t* info, const int32_t* only_if = nullptr);
int launch_wg_getrf_f32(hipStream_t, Mat<float> A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                        int32_t* info);  // FP32 storage, FP64 arithmetic
bool lu_wg_fits(Mat<double> A, Mat<double> X, int nrhs);  // the LU kernel's 32-bit offsets cover A and X
bool lu_wg_fits_f32(Mat<float> A);
int launch_wg_mixed(hipStream_t, int lu, int bf16, int itmax, double tol, Mat<double> A, int64_t* ipiv,
                    int64_t ipiv_stride, Mat<double> B, Mat<double> X, int nrhs, float* work, int64_t batch,
                    int32_t* info, int32_t* iter);
int dispatch_dsposv_ws(hipStream_t, int bf16, int itmax, double tol, Mat<double> A, Mat<double> B, Mat<double> X,
                       float* work, int64_t batch, int32_t* info, int32_t* iter);
int dispatch_dgesv(hipStream_t, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride, Mat<double> B, Mat<double> X,
                   int64_t batch, int32_t* info);
int dispatch_dsgesv(hipStream_t, int bf16, int itmax, double tol, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride,
                    Mat<double> B, Mat<double> X, float* work, int64_t batch, int32_t* info, int32_t* iter);

// riccati.hip: trailing-update strip width (MFMA tiles per wavefront) per
// nx bucket (nx / 16), passed by value.
struct RiccatiTuning {
    int8_t tn[16];
};
int launch_riccati(hipStream_t, int horizon, const void* inst, int64_t start, int64_t count, const double* in,
                   double* out, double* work, int32_t* queue, RiccatiTuning const& tuning, int32_t* info, int grid);

// blocked.hip: helpers of the blocked Cholesky path for n > 256
template <typename T> int launch_copy_lower(hipStream_t, Mat<T> A, Mat<T> L, int64_t batch);
template <typename T> int launch_copy_scaled(hipStream_t, Mat<T> B, Mat<T> X, T alpha, int64_t batch);
template <typename T> int launch_diag_info(hipStream_t, Mat<T> L, int64_t batch, int32_t* info);

// getrf_blocked.hip: the panel step of the blocked LU (square 256 < n <=
// 1024, not transposed): factor columns k0..k0+nb-1 (nb <= the panel width) with partial
// pivoting, apply the interchanges to the columns right of the panel and
// form U12.
bool getrf_panel_covers(int m, int n, int trans);
template <typename T> int getrf_panel_width();  // columns per panel (24 FP64, 32 FP32)
template <typename T>
int launch_getrf_panel(hipStream_t, Mat<T> A, int k0, int nb, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                       int32_t* info);
// after the last panel: the later panels' interchanges on the columns left
// of each panel (skipped for instances with info != 0)
template <typename T>
int launch_getrf_left_swaps(hipStream_t, Mat<T> A, const int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                            const int32_t* info);

// dispatch.cpp
template <typename T> int dispatch_gemm(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int dispatch_syrk(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);
template <typename T> int dispatch_trmm(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int dispatch_trsm(hipStream_t, int, int, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
template <typename T> int dispatch_potrf(hipStream_t, Mat<T>, Mat<T>, int64_t, int32_t*);
template <typename T> int dispatch_getrf(hipStream_t, Mat<T>, int64_t*, int64_t, int64_t, int32_t*);
template <typename T> int dispatch_posv(hipStream_t, Mat<T>, Mat<T>, Mat<T>, Mat<T>, int64_t, int32_t*);

template <typename T>
inline Mat<T> transposed(Mat<T> m) {
    Mat<T> r = m;
    r.rows = m.cols;
    r.cols = m.rows;
    r.trans = !m.trans;
    return r;
}

}  // namespace blast_hip
