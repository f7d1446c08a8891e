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

// posv_tiles.hip: returns 1 when the shape is not covered
int launch_posv_tiles(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int64_t, int32_t*);

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
