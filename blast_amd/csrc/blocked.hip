// Helpers of the blocked Cholesky path for n > 256 (dispatch.cpp): the lower
// triangle of A copied into L, the solves' right-hand sides copied and
// scaled, and potrf's info read off the factor's diagonal.

#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {
namespace {

constexpr int kBl = 256;
constexpr int kCopyCols = 8;  // columns per workgroup of copy_lower_kernel

// L(i, j) = A(i, j) for i >= j (the rest of L untouched); workgroup
// (instance, y) copies 8 panels (panel storage) or 8 columns (otherwise),
// consecutive addresses per thread
template <typename T>
__global__ __launch_bounds__(kBl) void copy_lower_kernel(Mat<T> Am, Mat<T> Lm) {
    Ref<T> a(Am, blockIdx.x), l(Lm, blockIdx.x);
    int const n = Am.rows;
    if (Am.panel && Lm.panel && !Am.trans && !Lm.trans) {
        // panel storage: workgroup y copies panels 8y..8y+7 (PS rows each,
        // stored contiguously column after column) in storage order
        constexpr int PS = PanelSize<T>::value, LOG = PanelSize<T>::log2;
        int const p0 = kCopyCols * int(blockIdx.y), np = (n + PS - 1) / PS;
        for (int p = p0; p < p0 + kCopyCols && p < np; ++p) {
            int const jmax = PS * p + PS < n ? PS * p + PS : n;  // columns 0..jmax-1 meet the panel's lower part
            for (int e = int(threadIdx.x); e < PS * jmax; e += kBl) {
                int const i = PS * p + (e & (PS - 1)), j = e >> LOG;
                if (i < n && i >= j) l(i, j) = a(i, j);
            }
        }
        return;
    }
    int const j0 = kCopyCols * int(blockIdx.y);  // columns 8y..8y+7, rows down each column
    for (int j = j0; j < j0 + kCopyCols && j < n; ++j)
        for (int i = j + int(threadIdx.x); i < n; i += kBl) l(i, j) = a(i, j);
}

// X = alpha B (same logical shape)
template <typename T>
__global__ __launch_bounds__(kBl) void copy_scaled_kernel(Mat<T> Bm, Mat<T> Xm, T alpha) {
    Ref<T> bb(Bm, blockIdx.x), x(Xm, blockIdx.x);
    int const rows = Bm.rows, cols = Bm.cols;
    for (int e = threadIdx.x; e < rows * cols; e += kBl) {
        int const i = e % rows, j = e / rows;
        x(i, j) = alpha * bb(i, j);
    }
}

// info = 1 + the first i with L(i, i) not positive (a pivot <= 0 or NaN
// leaves NaN or a non-positive entry there, and NaN in every later
// diagonal entry), 0 if none
template <typename T>
__global__ __launch_bounds__(kBl) void diag_info_kernel(Mat<T> Lm, int32_t* info) {
    __shared__ int sFirst;
    Ref<T> l(Lm, blockIdx.x);
    int const n = Lm.rows;
    if (threadIdx.x == 0) sFirst = 0x7fffffff;
    __syncthreads();
    int first = 0x7fffffff;
    for (int i = threadIdx.x; i < n; i += kBl)
        if (!(double(l(i, i)) > 0.0)) first = min(first, i);
    if (first != 0x7fffffff) atomicMin(&sFirst, first);
    __syncthreads();
    if (threadIdx.x == 0) info[blockIdx.x] = sFirst == 0x7fffffff ? 0 : sFirst + 1;
}

}  // namespace

template <typename T>
int launch_copy_lower(hipStream_t s, Mat<T> A, Mat<T> L, int64_t batch) {
    if (batch == 0) return 0;
    unsigned const groups = unsigned((A.rows + kCopyCols - 1) / kCopyCols);
    if (groups == 0) return 0;
    hipLaunchKernelGGL((copy_lower_kernel<T>), dim3(unsigned(batch), groups), dim3(kBl), 0, s, A, L);
    return check_launch("copy_lower_kernel");
}
template <typename T>
int launch_copy_scaled(hipStream_t s, Mat<T> B, Mat<T> X, T alpha, int64_t batch) {
    if (batch == 0) return 0;
    hipLaunchKernelGGL((copy_scaled_kernel<T>), dim3(unsigned(batch)), dim3(kBl), 0, s, B, X, alpha);
    return check_launch("copy_scaled_kernel");
}
template <typename T>
int launch_diag_info(hipStream_t s, Mat<T> L, int64_t batch, int32_t* info) {
    if (batch == 0 || !info) return 0;
    hipLaunchKernelGGL((diag_info_kernel<T>), dim3(unsigned(batch)), dim3(kBl), 0, s, L, info);
    return check_launch("diag_info_kernel");
}

template int launch_copy_lower<double>(hipStream_t, Mat<double>, Mat<double>, int64_t);
template int launch_copy_lower<float>(hipStream_t, Mat<float>, Mat<float>, int64_t);
template int launch_copy_scaled<double>(hipStream_t, Mat<double>, Mat<double>, double, int64_t);
template int launch_copy_scaled<float>(hipStream_t, Mat<float>, Mat<float>, float, int64_t);
template int launch_diag_info<double>(hipStream_t, Mat<double>, int64_t, int32_t*);
template int launch_diag_info<float>(hipStream_t, Mat<float>, int64_t, int32_t*);

}  // namespace blast_hip
