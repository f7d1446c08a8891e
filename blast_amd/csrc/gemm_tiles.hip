// Batched D = alpha*op(A)*op(B) + beta*C on MFMA tiles (FP64: v_mfma_f64_16x16x4f64,
// FP32: v_mfma_f32_16x16x4f32), 256-thread workgroups, grid y = the 16 x 64
// blocks of D in groups of four (one block per wavefront).
//
// Each wavefront owns 16-row strips of D (strip R = wavefront + 4 rs, ...) and
// computes up to four 16x16 tiles of a strip at once: per K-chunk of 4 it
// loads one A-operand element per lane (reused across the four tiles) and one
// B-operand element per tile.  The epilogue is BLAST's: d = fma(beta, c,
// alpha*acc) (register_matrix/Gemm.hpp:89-91).  The same kernel serves
//   syrkLower  D = alpha*A*A^T + beta*C, lower triangle stored
//              (dense/Syrk.hpp:15-145; B is the transposed view of A);
//   trmm       C = alpha*A*B with A upper (left form) or alpha*B*A with A
//              lower (right form) (algorithm/Trmm.hpp:93-242): the
//              triangular operand is masked on load, beta = 0.
// Summation order differs from BLAST's per-element FMA chain (agreement to
// backward error; bit-exact on exactly representable data).

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {

enum GemmMode : int { kGemm = 0, kSyrkLower = 1, kTrmmLeftUpper = 2, kTrmmRightLower = 3, kUpperTimesLower = 4,
                     kGemmOverwrite = 5 };

template <typename T, int MODE>
__global__ __launch_bounds__(256) void gemm_tile_kernel(T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D,
                                                        int K, unsigned ts_count) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x & 63;
    int const w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    Ref<T> a(A, b), bb(B, b), c(C, b), d(D, b);
    int const Mr = D.rows, Nc = D.cols;
    int const MT = (Mr + 15) / 16, NTt = (Nc + 15) / 16, KQ = (K + 3) / 4;
    int const m = lane & 15, g = lane >> 4;

    // Large outputs spread over the workgroups of grid y = RS x TS: strips
    // R = w + 4 rs, + 4 RS, ..., column groups T0 = 4 ts, + 4 TS, ... (one
    // 16 x 64 block per wavefront when the grid covers the output; the many
    // short wavefronts hide each other's memory latency)
    int const TS = int(ts_count), RS = int(gridDim.y) / TS;
    int const rs = int(blockIdx.y) / TS, ts = int(blockIdx.y) % TS;
    for (int R = w + 4 * rs; R < MT; R += 4 * RS) {
        int const row = 16 * R + m;  // A-operand row of this lane
        for (int T0 = 4 * ts; T0 < NTt; T0 += 4 * TS) {
            if (MODE == kSyrkLower && 16 * T0 > 16 * R + 15) break;  // strictly upper tiles
            acc_t acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = acc_t{};
            // C tiles (modes that read C) loaded with the first operands, in
            // flight during the MFMAs instead of after them
            constexpr bool kReadC = MODE == kGemm || MODE == kSyrkLower;
            T cv[4][4];
            if constexpr (kReadC) {
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        int const i = 16 * R + M::row(lane, r), j = 16 * (T0 + t) + acc_col(lane);
                        cv[t][r] = (i < Mr && j < Nc && (MODE != kSyrkLower || i >= j)) ? c(i, j) : T(0);
                    }
            }
            constexpr int U = 4;  // K-chunks of loads in flight (8 costs occupancy)
            for (int q0 = 0; q0 < KQ; q0 += U) {
                T av[U], bv[U][4];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int const k = 4 * (q0 + u) + g;
                    bool const ina =
                        row < Mr && k < K && ((MODE != kTrmmLeftUpper && MODE != kUpperTimesLower) || k >= row);
                    av[u] = ina ? a(row, k) : T(0);
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        int const col = 16 * (T0 + t) + m;  // B-operand column of this lane
                        bool const inb = col < Nc && k < K &&
                                         ((MODE != kTrmmRightLower && MODE != kUpperTimesLower) || k >= col);
                        bv[u][t] = inb ? bb(k, col) : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
#pragma unroll
                    for (int t = 0; t < 4; ++t) acc[t] = M::mma(av[u], bv[u][t], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const i = 16 * R + M::row(lane, r), j = 16 * (T0 + t) + acc_col(lane);
                    if (i < Mr && j < Nc && (MODE != kSyrkLower || i >= j)) {
                        T const ab = acc[t][r] * alpha;
                        if constexpr (kReadC)
                            d(i, j) = fma_t(beta, cv[t][r], ab);
                        else
                            d(i, j) = ab;
                    }
                }
        }
    }
}

// Plain gemm for M, N <= 64 and K <= 64: one 256-thread workgroup per
// instance, wavefront w owning rows 16w..16w+15 of D.  Each operand is read
// once: a quarter of B per wavefront, staged in LDS for all four (the tile
// kernel reads all of B in every wavefront, in four rounds); after the
// barrier the wavefront's A strip (already in A-operand layout) and C strip
// (accumulator layout), whose loads overlap the MFMAs (issued this late, they
// reuse the staging registers: four wavefronts per SIMD).  Every element is
// touched once, so the accesses are nontemporal.  Zero padding outside K x N
// keeps the 16 K-chunks branch-free; the MFMA sequence and the epilogue are
// the tile kernel's, so the results are identical.
constexpr int kSmallLdb = 64 + 4;  // LDS row stride of staged B (offsets the four row groups' banks)

template <typename T>
__global__ __launch_bounds__(256) void gemm_small_kernel(T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D) {
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    __shared__ T sB[64 * kSmallLdb];  // B[k][n] at k * kSmallLdb + n
    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x & 63;
    int const w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    Ref<T> a(A, b), bb(B, b), c(C, b), d(D, b);
    int const Mr = D.rows, Nc = D.cols, K = A.cols;
    int const m = lane & 15, g = lane >> 4;

    // B rows 16w..16w+15: lane l holds row 16w + 4(j % 4) + l % 4, column
    // 16(j / 4) + l / 4 (panel storage: 64 consecutive elements per load)
    T bst[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int const k = 16 * w + 4 * (j & 3) + (lane & 3), n = 16 * (j >> 2) + (lane >> 2);
        bst[j] = (k < K && n < Nc) ? __builtin_nontemporal_load(&bb(k, n)) : T(0);
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        int const k = 16 * w + 4 * (j & 3) + (lane & 3), n = 16 * (j >> 2) + (lane >> 2);
        sB[k * kSmallLdb + n] = bst[j];
    }
    __syncthreads();
    int const row = 16 * w + m;  // A-operand row of this lane
    T av[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        int const k = 4 * q + g;
        av[q] = (row < Mr && k < K) ? __builtin_nontemporal_load(&a(row, k)) : T(0);
    }
    // C strip: issued after the staging (its registers are the B staging
    // registers'), in flight while the MFMAs run
    T cv[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const i = 16 * w + M::row(lane, r), jj = 16 * t + acc_col(lane);
            cv[t][r] = (i < Mr && jj < Nc) ? __builtin_nontemporal_load(&c(i, jj)) : T(0);
        }
    acc_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = acc_t{};
#pragma unroll
    for (int q = 0; q < 16; ++q)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = M::mma(av[q], sB[(4 * q + g) * kSmallLdb + 16 * t + m], acc[t]);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const i = 16 * w + M::row(lane, r), jj = 16 * t + acc_col(lane);
            if (i < Mr && jj < Nc) __builtin_nontemporal_store(fma_t(beta, cv[t][r], acc[t][r] * alpha), &d(i, jj));
        }
}

template <typename T, int MODE>
static int launch_mode(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D, int K, int64_t batch) {
    if (batch == 0 || D.rows == 0 || D.cols == 0) return 0;
    // one 16 x 64 block per wavefront: RS x TS workgroups per instance
    // (grid y at most 65535; beyond that the wavefronts loop)
    unsigned const RS = unsigned((((D.rows + 15) / 16) + 3) / 4), TSn = unsigned((((D.cols + 15) / 16) + 3) / 4);
    unsigned const TS = TSn < 65535u / RS ? TSn : (65535u / RS > 0 ? 65535u / RS : 1u);
    unsigned const RSc = RS < 65535u ? RS : 65535u;
    hipLaunchKernelGGL((gemm_tile_kernel<T, MODE>), dim3(unsigned(batch), RSc * TS), dim3(256), 0, s, alpha, A, B,
                       beta, C, D, K, TS);
    return check_launch("gemm_tile_kernel");
}

template <typename T>
int launch_gemm_tiles(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    // 64 x 64 blocks with K in (32, 64]: the single-pass kernel (shorter K
    // would spend most of its 16 fixed K-chunks on padding)
    if (D.rows <= 64 && D.cols <= 64 && A.cols > 32 && A.cols <= 64) {
        if (batch == 0 || D.rows == 0 || D.cols == 0) return 0;
        hipLaunchKernelGGL((gemm_small_kernel<T>), dim3(unsigned(batch)), dim3(256), 0, s, alpha, A, B, beta, C, D);
        return check_launch("gemm_small_kernel");
    }
    return launch_mode<T, kGemm>(s, alpha, A, B, beta, C, D, A.cols, batch);
}

template <typename T>
int launch_syrk_tiles(hipStream_t s, T alpha, Mat<T> A, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    return launch_mode<T, kSyrkLower>(s, alpha, A, transposed(A), beta, C, D, A.cols, batch);
}

// D = alpha*A*B^T + beta*C, lower triangle stored: the two-operand form of
// syrkLower (BLASFEO's syrk_ln; B == A gives syrkLower).
template <typename T>
int launch_syrk2_tiles(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    return launch_mode<T, kSyrkLower>(s, alpha, A, transposed(B), beta, C, D, A.cols, batch);
}

// side_right = 0: C = alpha*A*B (A upper); 1: C = alpha*B*A (A lower).
template <typename T>
int launch_trmm_tiles(hipStream_t s, int side_right, T alpha, Mat<T> A, Mat<T> B, Mat<T> C, int64_t batch) {
    if (!side_right) return launch_mode<T, kTrmmLeftUpper>(s, alpha, A, B, T(0), C, C, A.cols, batch);
    return launch_mode<T, kTrmmRightLower>(s, alpha, B, A, T(0), C, C, B.cols, batch);
}

// D = alpha * A * B, D not read (used by the adjoints on uninitialised outputs).
template <typename T>
int launch_gemm_overwrite(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, Mat<T> D, int64_t batch) {
    return launch_mode<T, kGemmOverwrite>(s, alpha, A, B, T(0), D, D, A.cols, batch);
}

// C = alpha * triu(A) * tril(B) (used by the potrf adjoint: L^T Lbar).
template <typename T>
int launch_upper_times_lower(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, Mat<T> C, int64_t batch) {
    return launch_mode<T, kUpperTimesLower>(s, alpha, A, B, T(0), C, C, A.cols, batch);
}

#define INSTANTIATE(T)                                                                              \
    template int launch_gemm_tiles<T>(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t); \
    template int launch_syrk_tiles<T>(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);         \
    template int launch_syrk2_tiles<T>(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t); \
    template int launch_trmm_tiles<T>(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);              \
    template int launch_upper_times_lower<T>(hipStream_t, T, Mat<T>, Mat<T>, Mat<T>, int64_t);           \
    template int launch_gemm_overwrite<T>(hipStream_t, T, Mat<T>, Mat<T>, Mat<T>, int64_t);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace blast_hip
