// Fused batched Cholesky factor + solve, FP64, n = 16 NT <= 128 with A and L
// in panel storage (every tile inside the matrix), on MFMA tiles.
//
// One wavefront per instance.  The matrix is cut into 16x16 tiles; the 36
// lower-triangle tiles (n = 128) live in registers as v_mfma_f64_16x16x4
// accumulators.  Right-looking blocked Cholesky with a one-block lookahead,
// block column kb = 0..NT-1:
//   panel:    L_ik = A_ik W_kb^T on MFMA (W_kb = L_kb,kb^{-1}, formed with
//             the diagonal factor), written to HBM and kept as operands;
//   forward:  y_kb = W_kb b_kb on MFMA, the right-hand sides held in
//             registers as accumulator tiles;
//   lookahead: tile (kb+1, kb+1) and b_kb+1 are updated first and the tile
//             is factored at once (16x16 Cholesky and W on the VALU, DPP row
//             broadcasts); the updates b_i -= L_ik y_kb and the trailing
//             update A_ij -= L_ik L_jk^T of the other tiles are spread
//             between the 32 steps of that factorisation.
// Then the backward substitution x = L^{-T} y on MFMA, right-looking, block
// by block, with L read straight from the accumulators and the right-hand
// sides kept in registers.  The kernel makes no calls and
// tests no bounds; shapes with partial tiles or dense storage go to the
// workgroup kernel (chol_team.hpp).
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// factorisation reorders BLAST's sums (agreement to backward error, not bits).
// HBM traffic per instance: the lower triangle of A is read, the lower
// triangle of L written, B read and X written.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

// Block columns whose factor tiles are not kept in registers after the panel
// step (the backward solve reads them back from L), so that nothing is
// spilled: at kb = 0 the 28 trailing tiles alone fill the accumulator
// registers.
constexpr int kReloadCols = 2;

// The trailing update of step kb minus the lookahead tile: tiles (i, j),
// kb + 1 <= j <= i < NT, (i, j) != (kb + 1, kb + 1), each four MFMAs (K-chunks
// q = 0..3).  Block column by block column (at kb = 0 the later columns may
// still be arriving from HBM), and within a column q outer, so that
// consecutive MFMAs accumulate into different tiles.
template <int NT, int KB>
struct Trailing {
    static constexpr int tiles = (NT - KB - 1) * (NT - KB) / 2 - 1;
    static constexpr int ops = 4 * (tiles > 0 ? tiles : 0);
    static constexpr int encode(int o) {  // (q << 16) | (i << 8) | j
        int c = 0;
        for (int j = KB + 1; j < NT; ++j) {
            int const i0 = j == KB + 1 ? j + 1 : j, cnt = NT - i0;  // tiles of column j
            if (o < c + 4 * cnt) {
                int const q = (o - c) / cnt, i = i0 + (o - c) % cnt;
                return (q << 16) | (i << 8) | j;
            }
            c += 4 * cnt;
        }
        return -1;
    }
};

template <int NT>
__global__ __launch_bounds__(64) void posv_wave_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                       int nrhs, int32_t* info, const int32_t* only_if) {
    using M = Mfma<double>;
    constexpr int n = 16 * NT;
    constexpr int NR = NT * 16 * kMaxRhs;  // right-hand-side elements, [row][rhs]
    constexpr int RPL = NR / 64;           // per lane
    __shared__ double sP[NT][kTileElems];  // block column kb: tile i at sP[i]; sP[kb + 1]: the next diagonal tile
    __shared__ double sW[NT][kTileElems];  // W_kb
    __shared__ double sR[NR];              // b -> y -> x

    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;  // fallback mode: only the flagged instances
    int const lane = threadIdx.x;
    Ref<double> a(A, b), l(L, b);
    constexpr int NTILES = NT * (NT + 1) / 2;

    // right-hand sides first, then A: every load is issued before the first wait
    double rv[RPL];
    if (nrhs > 0) {
        Ref<double> bb(B, b);
#pragma unroll
        for (int e = 0; e < RPL; ++e) {
            int const idx = lane + 64 * e, row = idx / kMaxRhs, c = idx % kMaxRhs;
            rv[e] = c < nrhs ? bb(row, c) : 0.0;
        }
    }
    // block column by block column, so that the first steps wait only for
    // their own tiles
    d4 acc[NTILES];
    static_for<0, NT>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        static_for<j, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            // (per-element addressing here: with the scalar-addressed L
            // stores below, the scalar form exceeds the register budget)
            acc[tri(i, j)] = load_tile_full<true, double, false>(a, i, j, lane);
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
    // right-hand sides in registers: r[i] is block i in accumulator layout
    // (row 4q + l/16, column l%16 of lane l); it holds b until the forward
    // step of block i, whose y = L^{-1} b waits in LDS for the backward steps
    d4 r[NT];
    static_for<0, NT>([&](auto I) __attribute__((always_inline)) { r[I] = rload(I); });
    int inf = 0;

    // diagonal tile 0
    acc_to_lds(sP[0], acc[tri(0, 0)], lane);
    {
        int const bad = diag_factor_inl<double>(sP[0], sW[0], lane);
        if (bad) inf = bad;
        store_tile<true>(l, true, n, 0, 0, lds_to_acc<double>(sP[0], lane), lane);
    }

    static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
        constexpr int kb = decltype(KB)::value;
        // stage the panel tiles (updated through block column kb - 1)
        static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            acc_to_lds(sP[i], acc[tri(i, kb)], lane);
        });
        // panel: L_ik = A_ik W_kb^T, two tiles' MFMA chains interleaved
        static_for<0, (NT - kb) / 2>([&](auto P) __attribute__((always_inline)) {
            constexpr int i0 = kb + 1 + 2 * decltype(P)::value, i1 = i0 + 1 < NT ? i0 + 1 : i0;
            d4 z0 = {0.0, 0.0, 0.0, 0.0}, z1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double const w = op_read(sW[kb], q, lane);
                z0 = M::mma(op_read(sP[i0], q, lane), w, z0);
                if constexpr (i1 != i0) z1 = M::mma(op_read(sP[i1], q, lane), w, z1);
            }
            store_tile_panel_full(l, i0, kb, z0, lane);
            acc_to_lds(sP[i0], z0, lane);
            if constexpr (kb >= kReloadCols) acc[tri(i0, kb)] = z0;
            if constexpr (i1 != i0) {
                store_tile_panel_full(l, i1, kb, z1, lane);
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
        // forward substitution, y_kb = W_kb r_kb (two partial chains); the
        // updates r_i -= L_ik y_kb follow in the lookahead below
        d4 y = {0.0, 0.0, 0.0, 0.0};
        if (nrhs > 0) {
            d4 y1 = {0.0, 0.0, 0.0, 0.0};
            y = M::mma(op_read(sW[kb], 0, lane), r[kb][0], y);
            y1 = M::mma(op_read(sW[kb], 1, lane), r[kb][1], y1);
            y = M::mma(op_read(sW[kb], 2, lane), r[kb][2], y);
            y1 = M::mma(op_read(sW[kb], 3, lane), r[kb][3], y1);
            y += y1;
            rstore(kb, y);  // (kept in LDS until the backward step)
        }
        if constexpr (kb + 1 < NT) {
            constexpr int kd = kb + 1;
            // lookahead: the next diagonal tile first (its panel tile is
            // no longer needed in LDS: the operands are in f)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[tri(kd, kd)] = M::mma(-f[kd][q], f[kd][q], acc[tri(kd, kd)]);
            acc_to_lds(sP[kd], acc[tri(kd, kd)], lane);
            // its factorisation, with the rest of the trailing update spread
            // over the 32 steps (about ops / 32 MFMAs after each)
            using Tr = Trailing<NT, kb>;
            // the next block's right-hand side first (its y is formed right
            // after the factorisation)
            if (nrhs > 0)
#pragma unroll
                for (int q = 0; q < 4; ++q) r[kd] = M::mma(-f[kd][q], y[q], r[kd]);
            int const bad = diag_factor_hooked<double>(sP[kd], sW[kd], lane, [&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                // forward updates of the other blocks in the first steps
                if constexpr (s < NT - kd - 1) {
                    constexpr int i = kd + 1 + s;
                    if (nrhs > 0)
#pragma unroll
                        for (int q = 0; q < 4; ++q) r[i] = M::mma(-f[i][q], y[q], r[i]);
                }
                static_for<Tr::ops * s / 32, Tr::ops * (s + 1) / 32>([&](auto O) __attribute__((always_inline)) {
                    constexpr int e = Tr::encode(decltype(O)::value);
                    constexpr int q = e >> 16, i = (e >> 8) & 255, j = e & 255;
                    acc[tri(i, j)] = M::mma(-f[i][q], f[j][q], acc[tri(i, j)]);
                });
            });
            if (bad && inf == 0) inf = 16 * kd + bad;
            store_tile<true>(l, true, n, kd, kd, lds_to_acc<double>(sP[kd], lane), lane);
        }
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
                acc[tri(i, j)] = load_tile_full<true, double, false>(l, i, j, lane);
            });
        });
        // right-looking: x_kb = W_kb^T r_kb (r_kb final), then r_j -= L_kb,j^T
        // x_kb for j < kb, the next block's update first.  Accumulator
        // register q of x_kb is the B operand of its rows 4q..4q+3 and
        // register q of L_kb,j the A operand of L_kb,j^T, so nothing passes
        // through LDS.
        static_for<0, NT>([&](auto I) __attribute__((always_inline)) { r[I] = rload(I); });
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            d4 x = {0.0, 0.0, 0.0, 0.0}, x1 = {0.0, 0.0, 0.0, 0.0};
            // A operand W^T: W[4q + l/16][l%16]
            x = M::mma(sW[kb][cn * kTileStride + 0 + (lane >> 4)], r[kb][0], x);
            x1 = M::mma(sW[kb][cn * kTileStride + 4 + (lane >> 4)], r[kb][1], x1);
            x = M::mma(sW[kb][cn * kTileStride + 8 + (lane >> 4)], r[kb][2], x);
            x1 = M::mma(sW[kb][cn * kTileStride + 12 + (lane >> 4)], r[kb][3], x1);
            x += x1;
            r[kb] = x;
            static_for<0, kb>([&](auto JJ) __attribute__((always_inline)) {
                constexpr int j = kb - 1 - decltype(JJ)::value;
#pragma unroll
                for (int q = 0; q < 4; ++q) r[j] = M::mma(acc[tri(kb, j)][q], -x[q], r[j]);
            });
        });
        static_for<0, NT>([&](auto I) __attribute__((always_inline)) { rstore(I, r[I]); });
        Ref<double> xx(X, b);
        for (int e = lane; e < n * nrhs; e += 64) {
            int const row = e % n, c = e / n;
            xx(row, c) = sR[row * kMaxRhs + c];
        }
    }
    if (info && lane == 0) info[b] = inf;
}

template <int NT>
int launch_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int nrhs, int64_t batch,
              int32_t* info, const int32_t* only_if) {
    hipLaunchKernelGGL((posv_wave_kernel<NT>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, nrhs, info, only_if);
    return check_launch("posv_wave_kernel");
}

}  // namespace

// Fused potrf (+ forward/backward solve when nrhs > 0) for n = 16 NT <= 128
// with A and L in (untransposed) panel storage, nrhs <= 4.  Returns 1 when
// the shape is not covered.
int launch_posv_tiles(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    if (L.rows != L.cols || n > 128 || n == 0 || n % 16 != 0 || !panel || nrhs > kMaxRhs) return 1;
    if (batch == 0) return 0;
    switch (n / 16) {
        case 1: return launch_nt<1>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 2: return launch_nt<2>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 3: return launch_nt<3>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 4: return launch_nt<4>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 5: return launch_nt<5>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 6: return launch_nt<6>(s, A, L, B, X, nrhs, batch, info, only_if);
        case 7: return launch_nt<7>(s, A, L, B, X, nrhs, batch, info, only_if);
        default: return launch_nt<8>(s, A, L, B, X, nrhs, batch, info, only_if);
    }
}

}  // namespace blast_hip
