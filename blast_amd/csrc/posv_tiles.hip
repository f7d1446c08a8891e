// Fused batched Cholesky factor + solve for n <= 128, FP64, on MFMA tiles.
//
// One wavefront per instance.  The matrix is cut into 16x16 tiles; the 36
// lower-triangle tiles (n = 128) live in registers as v_mfma_f64_16x16x4
// accumulators.  Right-looking blocked Cholesky, block column kb = 0..NT-1:
//   stage tile (kb,kb) and the panel tiles (i,kb) in LDS;
//   factor the 16x16 diagonal tile and build W = L_kk^{-1} in registers
//   (DPP row broadcasts, Newton-refined rsqrt);
//   L_ik = A_ik W^T on MFMA, written to HBM and back to the LDS panel;
//   forward substitution y_kb = W b_kb, b_i -= L_ik y_kb on the RHS;
//   trailing update A_ij -= L_ik L_jk^T on MFMA for kb < j <= i.
// Then the backward substitution x = L^{-T} y, column block by column block,
// with L read straight from the accumulators.
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// factorisation reorders BLAST's sums (agreement to backward error, not bits).
// HBM traffic per instance: the lower triangle of A is read, the lower
// triangle of L written, B read and X written.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {



// One wavefront per instance.  acc holds the lower tiles (i, j), j <= i, as
// MFMA accumulators (AGPR/VGPR resident); LDS stages the current block
// column, keeps W_kb = L_kb,kb^{-1} for the backward solve and holds the
// right-hand sides.  A single wavefront needs no barriers: its LDS accesses
// complete in program order.
template <int NT, bool PANEL>
__global__ __launch_bounds__(64) void posv_wave_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                       int n, int nrhs, int32_t* info, const int32_t* only_if) {
    __shared__ double sP[NT][kTileElems];          // block column kb: tile i at sP[i] (i >= kb)
    __shared__ double sW[NT][kTileElems];          // W_kb
    __shared__ double sR[NT * 16 * kMaxRhs + 64];  // RHS -> y -> x as [row][rhs]; tail: partial sums

    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;  // fallback mode: only the flagged instances
    int const lane = threadIdx.x;
    Ref<double> a(A, b), l(L, b);
    bool const lpanel = L.panel != 0;
    constexpr int NTILES = NT * (NT + 1) / 2;

    d4 acc[NTILES];
    static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            acc[tri(i, j)] = load_tile<PANEL, double>(a, n, i, j, lane);
        });
    });
    if (nrhs > 0) {
        Ref<double> bb(B, b);
        for (int e = lane; e < NT * 16 * kMaxRhs; e += 64) {
            int const row = e >> 2, c = e & 3;
            sR[e] = (row < n && c < nrhs) ? bb(row, c) : 0.0;
        }
    }
    int inf = 0;

    static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
        constexpr int kb = decltype(KB)::value;
        // stage block column kb
        static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                if constexpr (j == kb) acc_to_lds(sP[i], acc[tri(i, j)], lane);
            });
        });
        // diagonal tile: L_kk in sP[kb], W_kb in sW[kb]
        int const bad = diag_factor(sP[kb], sW[kb], lane);
        if (bad && inf == 0 && 16 * kb + bad <= n) inf = 16 * kb + bad;
        store_tile<PANEL>(l, lpanel, n, kb, kb, lds_to_acc<double>(sP[kb], lane), lane);
        // panel: L_ik = A_ik W_kb^T
        static_for<1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            static_for<0, i>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                if constexpr (j == kb) {
                    d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) z = Mfma<double>::mma(op_read(sP[i], q, lane), op_read(sW[kb], q, lane), z);
                    acc[tri(i, j)] = z;
                    store_tile<PANEL>(l, lpanel, n, i, j, z, lane);
                    acc_to_lds(sP[i], z, lane);
                }
            });
        });
        // forward substitution: y_kb = W_kb b_kb; b_i -= L_ik y_kb
        if (nrhs > 0) {
            int const m = lane >> 2, c = lane & 3;
            double rv[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) rv[k] = sR[(16 * kb + k) * kMaxRhs + c];
            double y = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) y = __builtin_fma(sW[kb][k * kTileStride + m], rv[k], y);
            sR[(16 * kb + m) * kMaxRhs + c] = y;
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                double v = sR[(16 * i + m) * kMaxRhs + c];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    v = __builtin_fma(-sP[i][k * kTileStride + m], sR[(16 * kb + k) * kMaxRhs + c], v);
                sR[(16 * i + m) * kMaxRhs + c] = v;
                __builtin_amdgcn_sched_barrier(0);
            });
        }
        // trailing update: A_ij -= L_ik L_jk^T for kb < j <= i, operands
        // loaded once per block column
        double f[NT][4];
        static_for<1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            if constexpr (i > kb)
#pragma unroll
                for (int q = 0; q < 4; ++q) f[i][q] = op_read(sP[i], q, lane);
        });
#pragma unroll
        for (int q = 0; q < 4; ++q)
            static_for<1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                static_for<1, i + 1>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value;
                    if constexpr (j > kb) acc[tri(i, j)] = Mfma<double>::mma(-f[i][q], f[j][q], acc[tri(i, j)]);
                });
            });
    });

    if (nrhs > 0) {
        // backward substitution x = L^{-T} y with L_ik from the accumulators
        double* sSum = sR + NT * 16 * kMaxRhs;
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            double p[kMaxRhs] = {0.0, 0.0, 0.0, 0.0};
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const row = 16 * i + Mfma<double>::row(lane, r);
#pragma unroll
                    for (int c = 0; c < kMaxRhs; ++c)
                        p[c] = __builtin_fma(acc[tri(i, kb)][r], sR[row * kMaxRhs + c], p[c]);
                }
                __builtin_amdgcn_sched_barrier(0);
            });
#pragma unroll
            for (int c = 0; c < kMaxRhs; ++c) {
                p[c] += __shfl_xor(p[c], 16);
                p[c] += __shfl_xor(p[c], 32);
            }
            if (lane < 16)
#pragma unroll
                for (int c = 0; c < kMaxRhs; ++c) sSum[lane * kMaxRhs + c] = p[c];
            int const m = lane >> 2, c = lane & 3;
            double tv[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) tv[k] = sR[(16 * kb + k) * kMaxRhs + c] - sSum[k * kMaxRhs + c];
            double x = 0.0;
#pragma unroll
            for (int k = 0; k < 16; ++k) x = __builtin_fma(sW[kb][m * kTileStride + k], tv[k], x);
            sR[(16 * kb + m) * kMaxRhs + c] = x;
        });
        Ref<double> xx(X, b);
        for (int e = lane; e < n * nrhs; e += 64) {
            int const row = e % n, c = e / n;
            xx(row, c) = sR[row * kMaxRhs + c];
        }
    }
    if (info && lane == 0) info[b] = inf;
}

template <int NT>
static int launch_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                     int64_t batch, int32_t* info, const int32_t* only_if) {
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    if (panel)
        hipLaunchKernelGGL((posv_wave_kernel<NT, true>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n, nrhs,
                           info, only_if);
    else
        hipLaunchKernelGGL((posv_wave_kernel<NT, false>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n,
                           nrhs, info, only_if);
    return check_launch("posv_wave_kernel");
}

// Fused potrf (+ forward/backward solve when nrhs > 0) for n <= 128,
// nrhs <= 4.  Returns 1 when the shape is not covered.
int launch_posv_tiles(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    if (L.rows != L.cols || n > 128 || nrhs > kMaxRhs) return 1;
    if (batch == 0) return 0;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_nt<1>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 2: return launch_nt<2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 3: return launch_nt<3>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 4: return launch_nt<4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 5: return launch_nt<5>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 6: return launch_nt<6>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 7: return launch_nt<7>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        default: return launch_nt<8>(s, A, L, B, X, n, nrhs, batch, info, only_if);
    }
}

}  // namespace blast_hip
