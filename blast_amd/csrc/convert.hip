// Batched storage conversion D := S between any two of BLAST's layouts
// (dense column-major, panel, and their transposed views), element by element
// on the logical matrix: the device counterpart of assigning a PanelMatrix to
// a dense matrix and back (expressions/PanelMatrix.hpp:81-167, :170-239).
//
// HBM-bound.  One 256-thread workgroup moves one 32 x 32 logical tile of one
// instance through LDS: it reads the tile in the order in which S stores it
// (so consecutive lanes read consecutive addresses) and writes it in the order
// in which D stores it, so a transposing conversion is coalesced on both sides.

#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {
namespace {

constexpr int kCvT = 32;  // tile edge

// Tile-local element e (0..1023) -> logical (i, j), enumerated in storage
// order of a matrix with the given layout (PS = panel size).
__device__ __forceinline__ void storage_order(int e, bool panel, bool trans, int ps, int& i, int& j) {
    int r, c;  // storage row / column within the tile
    if (panel) {
        r = ps * (e / (ps * kCvT)) + e % ps;
        c = (e / ps) % kCvT;
    } else {
        r = e % kCvT;
        c = e / kCvT;
    }
    i = trans ? c : r;
    j = trans ? r : c;
}

template <typename T>
__global__ __launch_bounds__(256) void convert_kernel(Mat<T> S, Mat<T> D, int tiles_m, int tiles_n) {
    __shared__ T sT[kCvT][kCvT + 1];
    int64_t const tiles = int64_t(tiles_m) * tiles_n;
    int64_t const b = blockIdx.x / tiles;
    int const t = int(blockIdx.x % tiles);
    int const i0 = kCvT * (t / tiles_n), j0 = kCvT * (t % tiles_n);
    int const m = D.rows, n = D.cols;
    constexpr int PS = PanelSize<T>::value;
    Ref<T> s(S, b), d(D, b);
#pragma unroll
    for (int k = 0; k < kCvT * kCvT / 256; ++k) {
        int i, j;
        storage_order(threadIdx.x + 256 * k, S.panel != 0, S.trans != 0, PS, i, j);
        if (i0 + i < m && j0 + j < n) sT[i][j] = s(i0 + i, j0 + j);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCvT * kCvT / 256; ++k) {
        int i, j;
        storage_order(threadIdx.x + 256 * k, D.panel != 0, D.trans != 0, PS, i, j);
        if (i0 + i < m && j0 + j < n) d(i0 + i, j0 + j) = sT[i][j];
    }
}

}  // namespace

template <typename T>
int launch_convert(hipStream_t s, Mat<T> S, Mat<T> D, int64_t batch) {
    if (batch == 0 || D.rows == 0 || D.cols == 0) return 0;
    int const tm = (D.rows + kCvT - 1) / kCvT, tn = (D.cols + kCvT - 1) / kCvT;
    int64_t const blocks = batch * tm * tn;
    if (blocks > 0x7fffffffLL) {
        set_error("convert: batch * tiles exceeds the grid");
        return -4;
    }
    hipLaunchKernelGGL((convert_kernel<T>), dim3(unsigned(blocks)), dim3(256), 0, s, S, D, tm, tn);
    return check_launch("convert_kernel");
}

template int launch_convert<double>(hipStream_t, Mat<double>, Mat<double>, int64_t);
template int launch_convert<float>(hipStream_t, Mat<float>, Mat<float>, int64_t);

}  // namespace blast_hip
