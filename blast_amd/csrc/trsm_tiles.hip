// Batched triangular solve X = alpha * op(A)^{-1} B on MFMA tiles, n <= 256,
// FP64 and FP32, left side (the right side X A = alpha B is the left solve
// with the transposed views), lower or upper, unit or non-unit diagonal.
//
// One 256-thread workgroup per instance and 16-column chunk of B.  The
// right-hand-side chunk lives in LDS.  The 16x16 diagonal tiles of A are
// inverted first, all at once (one tile per wavefront at a time, tri_inverse),
// so the block substitution is MFMA work only:
//   lower: for k = 0..NT-1:  y_k = W_k r_k;  r_i -= A_ik y_k  (i > k)
//   upper: for k = NT-1..0:  y_k = W_k r_k;  r_i -= A_ik y_k  (i < k)
// with W_k = A_kk^{-1}; the off-diagonal tiles of A are read from HBM once,
// four per wavefront per step with their loads in flight together.
// Algorithm reference: dense/Trsm.hpp:17-66 and reference/Trsm.hpp:44-66,
// :105-127 (the same solve; the block order reorders BLAST's sums, agreement
// to backward error, not bits).

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kTrT = 256;
constexpr int kTrWaves = kTrT / 64;
constexpr int kTrMaxNT = 16;
constexpr int kTrCols = 16;  // right-hand-side columns per workgroup

template <typename T>
__global__ __launch_bounds__(kTrT) void trsm_tiles_kernel(T alpha, Mat<T> A, Mat<T> B, Mat<T> X, int upper, int unit,
                                                          int nchunks) {
    using Mf = Mfma<T>;
    using acc_t = typename Mf::acc_t;
    __shared__ T sW[kTrMaxNT][kTileElems];   // W_k = A_kk^{-1}
    __shared__ T sR[kTrMaxNT * 16 * kTrCols];  // right-hand sides -> solution, [row][kTrCols]

    int64_t const b = blockIdx.x / nchunks;
    int const c0 = kTrCols * int(blockIdx.x % nchunks);
    int const tid = threadIdx.x, lane = tid & 63;
    int const w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const n = A.rows, nr = X.cols, NT = (n + 15) >> 4;
    Ref<T> a(A, b), bb(B, b), x(X, b);
    // A outside n x n reads as the identity
    auto aget = [&](int i, int j) -> T { return (i < n && j < n) ? a(i, j) : T(i == j ? 1 : 0); };

    for (int e = tid; e < NT * 16 * kTrCols; e += kTrT) {
        int const r = e / kTrCols, c = e % kTrCols;
        sR[e] = (r < n && c0 + c < nr) ? alpha * bb(r, c0 + c) : T(0);
    }
    for (int k = w; k < NT; k += kTrWaves) {  // diagonal inverses
        for (int e = lane; e < 256; e += 64) {
            int const m = e & 15, c = e >> 4;
            sW[k][c * kTileStride + m] = aget(16 * k + m, 16 * k + c);
        }
        tri_inverse<T>(sW[k], sW[k], lane, upper != 0, unit != 0);
    }
    __syncthreads();

    auto rb = [&](int blk, int k) -> T { return sR[(16 * blk + k) * kTrCols + (lane & 15)]; };
    for (int step = 0; step < NT; ++step) {
        int const k = upper ? NT - 1 - step : step;
        if (w == 0) {
            acc_t y{};
#pragma unroll
            for (int q = 0; q < 4; ++q) y = Mf::mma(op_read(sW[k], q, lane), rb(k, 4 * q + (lane >> 4)), y);
#pragma unroll
            for (int r = 0; r < 4; ++r) sR[(16 * k + Mf::row(lane, r)) * kTrCols + acc_col(lane)] = y[r];
        }
        __syncthreads();
        // remaining blocks: i > k (lower) or i < k (upper), four per wavefront
        int const first = upper ? 0 : k + 1, count = upper ? k : NT - 1 - k;
        T av[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const t = w + kTrWaves * u, i = first + t;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                av[u][q] = t < count ? aget(16 * i + (lane & 15), 16 * k + 4 * q + (lane >> 4)) : T(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const t = w + kTrWaves * u, i = first + t;
            if (t >= count) continue;
            acc_t v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = sR[(16 * i + Mf::row(lane, r)) * kTrCols + acc_col(lane)];
#pragma unroll
            for (int q = 0; q < 4; ++q) v = Mf::mma(-av[u][q], rb(k, 4 * q + (lane >> 4)), v);
#pragma unroll
            for (int r = 0; r < 4; ++r) sR[(16 * i + Mf::row(lane, r)) * kTrCols + acc_col(lane)] = v[r];
        }
        __syncthreads();
    }
    for (int e = tid; e < n * kTrCols; e += kTrT) {
        int const r = e / kTrCols, c = e % kTrCols;
        if (c0 + c < nr) x(r, c0 + c) = sR[e];
    }
}

}  // namespace

// Left solve A X = alpha B for square A, n <= 256.  X may alias B (each
// workgroup reads its chunk of B before writing the same chunk of X).
// Returns 1 when not covered.
template <typename T>
int launch_trsm_tiles(hipStream_t s, int upper, int unit, T alpha, Mat<T> A, Mat<T> B, Mat<T> X, int64_t batch) {
    int const n = A.rows;
    if (A.rows != A.cols || n > 16 * kTrMaxNT) return 1;
    if (batch == 0 || n == 0 || X.cols == 0) return 0;
    int const nchunks = (X.cols + kTrCols - 1) / kTrCols;
    int64_t const blocks = batch * nchunks;
    if (blocks > 0x7fffffffLL) return 1;
    hipLaunchKernelGGL((trsm_tiles_kernel<T>), dim3(unsigned(blocks)), dim3(kTrT), 0, s, alpha, A, B, X, upper, unit,
                       nchunks);
    return check_launch("trsm_tiles_kernel");
}

template int launch_trsm_tiles<double>(hipStream_t, int, int, double, Mat<double>, Mat<double>, Mat<double>, int64_t);
template int launch_trsm_tiles<float>(hipStream_t, int, int, float, Mat<float>, Mat<float>, Mat<float>, int64_t);

}  // namespace blast_hip
