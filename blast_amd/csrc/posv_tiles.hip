// Fused batched Cholesky factor + solve for n <= 128, FP64, on MFMA tiles.
//
// One wavefront per instance.  The matrix is cut into 16x16 tiles; the 36
// lower-triangle tiles (n = 128) live in registers as v_mfma_f64_16x16x4
// accumulators.  Right-looking blocked Cholesky, block column kb = 0..NT-1:
//   stage tile (kb,kb) and the panel tiles (i,kb) in LDS;
//   factor the 16x16 diagonal tile and build W = L_kk^{-1} in registers
//   (DPP row broadcasts, Newton-refined rsqrt);
//   L_ik = A_ik W^T on MFMA, written to HBM and back to the LDS panel;
//   forward substitution y_kb = W b_kb, b_i -= L_ik y_kb on MFMA;
//   trailing update A_ij -= L_ik L_jk^T on MFMA for kb < j <= i.
// Then the backward substitution x = L^{-T} y on MFMA, block by block, with
// L read straight from the accumulators.
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
//
// The solves run on MFMA with the right-hand sides as the B operand (columns
// 0..kMaxRhs-1 of the 16; the others are zero):
//   forward, at block column kb:  y_kb = W_kb r_kb,  r_i -= L_ik y_kb
//   backward, kb = NT-1..0:       x_kb = W_kb^T (y_kb - sum_{i>kb} L_ik^T x_i)
// Register r of an accumulator tile M (row l/16 + 4r, column l%16 of lane l)
// is the B operand of M for K-chunk r and the A operand of M^T for the same
// chunk, so y, x and the factor tiles feed the MFMAs without a transpose.
// Block columns whose factor tiles are not kept in registers after the panel
// step (the backward solve reads them back from L), so that nothing is
// spilled: at kb = 0 the 28 trailing tiles alone fill the accumulator
// registers.
constexpr int kReloadCols = 2;
// Solve tiles processed together, their MFMA chains interleaved.
constexpr int kSolveGroup = 2;

// FULL: n == 16 NT and panel storage, tiles loaded without bounds tests
// (load_tile_full).
template <int NT, bool PANEL, bool FULL>
__global__ __launch_bounds__(64) void posv_wave_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                       int n, int nrhs, int32_t* info, const int32_t* only_if) {
    using M = Mfma<double>;
    constexpr int NR = NT * 16 * kMaxRhs;  // right-hand-side elements, [row][rhs]
    constexpr int RPL = NR / 64;           // per lane
    __shared__ double sP[NT][kTileElems];  // block column kb: tile i at sP[i] (i >= kb)
    __shared__ double sW[NT][kTileElems];  // W_kb
    __shared__ double sR[NR];              // b -> y -> x

    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;  // fallback mode: only the flagged instances
    int const lane = threadIdx.x;
    Ref<double> a(A, b), l(L, b);
    bool const lpanel = L.panel != 0;
    constexpr int NTILES = NT * (NT + 1) / 2;

    // right-hand sides first, then A: every load is issued before the first wait
    double rv[RPL];
    if (nrhs > 0) {
        Ref<double> bb(B, b);
#pragma unroll
        for (int e = 0; e < RPL; ++e) {
            int const idx = lane + 64 * e, row = idx / kMaxRhs, c = idx % kMaxRhs;
            rv[e] = (row < n && c < nrhs) ? bb(row, c) : 0.0;
        }
    }
    d4 acc[NTILES];
    static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            // (per-element addressing here: with the scalar-addressed L
            // stores below, the scalar form exceeds the register budget)
            acc[tri(i, j)] = FULL ? load_tile_full<PANEL, double, false>(a, i, j, lane) : load_tile<PANEL, double>(a, n, i, j, lane);
        });
    });
    if (nrhs > 0)
#pragma unroll
        for (int e = 0; e < RPL; ++e) sR[lane + 64 * e] = rv[e];

    // right-hand-side block blk in accumulator layout (zero in columns >= kMaxRhs)
    int const cn = lane & 15;
    bool const rl = cn < kMaxRhs;
    auto rload = [&](int blk) -> d4 {
        d4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double const t = sR[(16 * blk + M::row(lane, r)) * kMaxRhs + (cn & (kMaxRhs - 1))];
            v[r] = rl ? t : 0.0;
        }
        return v;
    };
    auto rstore = [&](int blk, d4 v) {
        if (rl)
#pragma unroll
            for (int r = 0; r < 4; ++r) sR[(16 * blk + M::row(lane, r)) * kMaxRhs + cn] = v[r];
    };
    int inf = 0;

    static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
        constexpr int kb = decltype(KB)::value;
        // stage block column kb
        static_for<kb, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            acc_to_lds(sP[i], acc[tri(i, kb)], lane);
        });
        // diagonal tile: L_kk in sP[kb], W_kb in sW[kb]
        int const bad = diag_factor(sP[kb], sW[kb], lane);
        if (bad && inf == 0 && 16 * kb + bad <= n) inf = 16 * kb + bad;
        store_tile<PANEL>(l, lpanel, n, kb, kb, lds_to_acc<double>(sP[kb], lane), lane);
        // panel: L_ik = A_ik W_kb^T, two tiles' MFMA chains interleaved
        // (more would raise the peak register count at kb = 0)
        static_for<0, (NT - kb) / 2>([&](auto P) __attribute__((always_inline)) {
            constexpr int i0 = kb + 1 + 2 * decltype(P)::value, i1 = i0 + 1 < NT ? i0 + 1 : i0;
            d4 z0 = {0.0, 0.0, 0.0, 0.0}, z1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double const w = op_read(sW[kb], q, lane);
                z0 = M::mma(op_read(sP[i0], q, lane), w, z0);
                if constexpr (i1 != i0) z1 = M::mma(op_read(sP[i1], q, lane), w, z1);
            }
            if constexpr (FULL) store_tile_panel_full(l, i0, kb, z0, lane);
            else store_tile<PANEL>(l, lpanel, n, i0, kb, z0, lane);
            acc_to_lds(sP[i0], z0, lane);
            if constexpr (kb >= kReloadCols) acc[tri(i0, kb)] = z0;
            if constexpr (i1 != i0) {
                if constexpr (FULL) store_tile_panel_full(l, i1, kb, z1, lane);
                else store_tile<PANEL>(l, lpanel, n, i1, kb, z1, lane);
                acc_to_lds(sP[i1], z1, lane);
                if constexpr (kb >= kReloadCols) acc[tri(i1, kb)] = z1;
            }
        });
        // operands L_ik (A-operand layout), shared by the forward solve and
        // the trailing update
        double f[NT][4];
        static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
#pragma unroll
            for (int q = 0; q < 4; ++q) f[i][q] = op_read(sP[i], q, lane);
        });
        // forward substitution: y_kb = W_kb r_kb; r_i -= L_ik y_kb
        if (nrhs > 0) {
            d4 const rk = rload(kb);
            d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) y = M::mma(op_read(sW[kb], q, lane), rk[q], y);
            rstore(kb, y);
            // kSolveGroup tiles at a time, their MFMA chains interleaved
            static_for<0, (NT - 1 - kb + kSolveGroup - 1) / kSolveGroup>([&](auto GI) __attribute__((always_inline)) {
                constexpr int i0 = kb + 1 + kSolveGroup * decltype(GI)::value;
                d4 v[kSolveGroup];
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (i0 + t < NT) v[t] = rload(i0 + t);
                });
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                        constexpr int t = decltype(T)::value;
                        if constexpr (i0 + t < NT) v[t] = M::mma(-f[i0 + t][q], y[q], v[t]);
                    });
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (i0 + t < NT) rstore(i0 + t, v[t]);
                });
            });
        }
        // trailing update: A_ij -= L_ik L_jk^T for kb < j <= i
#pragma unroll
        for (int q = 0; q < 4; ++q)
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                static_for<kb + 1, i + 1>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value;
                    acc[tri(i, j)] = M::mma(-f[i][q], f[j][q], acc[tri(i, j)]);
                });
            });
    });

    if (nrhs > 0) {
        // backward substitution with L_ik^T straight from the accumulators;
        // the first kReloadCols block columns are read back from L (this
        // lane's own stores, so program order makes them visible), issued
        // here so they arrive while the later columns are solved
        static_for<0, kReloadCols>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            static_for<j + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                acc[tri(i, j)] = FULL ? load_tile_full<PANEL, double, false>(l, i, j, lane) : load_tile<PANEL, double>(l, n, i, j, lane);
            });
        });
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            // sum_{i>kb} L_ik^T x_i in kSolveGroup partial sums (independent
            // MFMA chains), added at the end
            d4 tp[kSolveGroup];
#pragma unroll
            for (int g = 0; g < kSolveGroup; ++g) tp[g] = d4{0.0, 0.0, 0.0, 0.0};
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                d4 const xi = rload(i);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    tp[(i - kb - 1) % kSolveGroup] = M::mma(acc[tri(i, kb)][r], xi[r], tp[(i - kb - 1) % kSolveGroup]);
            });
            d4 t = tp[0];
#pragma unroll
            for (int g = 1; g < kSolveGroup; ++g) t += tp[g];
            d4 const u = rload(kb) - t;
            d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q)  // A operand W^T: W[4q + l/16][l%16]
                x = M::mma(sW[kb][cn * kTileStride + 4 * q + (lane >> 4)], u[q], x);
            rstore(kb, x);
        });
        Ref<double> xx(X, b);
        for (int e = lane; e < n * nrhs; e += 64) {
            int const row = e % n, c = e / n;
            xx(row, c) = sR[row * kMaxRhs + c];
        }
    }
    if (info && lane == 0) info[b] = inf;
}

template <int NT, bool PANEL>
static void launch_nt_full(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                           int64_t batch, int32_t* info, const int32_t* only_if) {
    if (PANEL && n == 16 * NT)  // (dense storage: the generic addressing would spill)
        hipLaunchKernelGGL((posv_wave_kernel<NT, PANEL, PANEL>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n,
                           nrhs, info, only_if);
    else
        hipLaunchKernelGGL((posv_wave_kernel<NT, PANEL, false>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n,
                           nrhs, info, only_if);
}

template <int NT>
static int launch_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                     int64_t batch, int32_t* info, const int32_t* only_if) {
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    if (panel)
        launch_nt_full<NT, true>(s, A, L, B, X, n, nrhs, batch, info, only_if);
    else
        launch_nt_full<NT, false>(s, A, L, B, X, n, nrhs, batch, info, only_if);
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
