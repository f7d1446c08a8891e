// Fused batched Cholesky factor + solve, FP64, n <= 256, nrhs <= 8, on MFMA
// tiles: one workgroup of NW wavefronts per instance.
//
// The NT(NT+1)/2 lower 16x16 tiles are spread over the wavefronts (table
// below) and stay in their owner's registers as v_mfma_f64_16x16x4f64
// accumulators for the whole factorisation.  Right-looking blocked Cholesky
// with a one-block lookahead on the diagonal; per block column kb, two
// barriers:
//   U(kb-1)  trailing update A_ij -= L_i,kb-1 L_j,kb-1^T (kb <= j <= i).  The
//            owner of tile (kb,kb) updates it first and factors it at once
//            (16x16 Cholesky and W = L_kk^{-1} on one wavefront, inlined:
//            the kernel makes no calls), writes L_kk, keeps W in the tile's
//            registers and forms y_kb = W r_kb, while the other wavefronts
//            run their share of the update;
//   P(kb)    the owners of the panel tiles (i,kb) form L_ik = A_ik W^T, write
//            it to L and to LDS, and update r_i -= L_ik y_kb.
// The backward solve x = L^{-T} y reads W_kk and L_kj from the registers:
// accumulator register r of a tile M is the MFMA A operand of M^T for
// K-chunk r.  HBM traffic is the algorithmic one: A's lower triangle and B
// in, L's lower triangle and X out.
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// algorithm reorders BLAST's sums (agreement to backward error, not bits).

#pragma once

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace team {

// Returns v unchanged, opaque to the optimiser: a tile's HBM address is then
// computed where the tile is stored instead of being hoisted out of the
// block loop (and held in registers) for every tile of the wavefront.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// Tile ownership.  Diagonal tile (j,j) belongs to wavefront j % NW (so
// consecutive diagonal factorisations alternate between wavefronts); the
// off-diagonal tiles are dealt round-robin column by column, top to
// bottom.  slot[w][s] = (i << 8 | j) of the s-th tile of wavefront w, or -1.
template <int NT, int NW>
struct TeamMap {
    static constexpr int kTiles = NT * (NT + 1) / 2;
    static constexpr int kSlots = [] {
        int cnt[NW] = {};
        int next = 0;
        for (int j = 0; j < NT; ++j)
            for (int i = j; i < NT; ++i) {
                int const w = i == j ? j % NW : (next++) % NW;
                ++cnt[w];
            }
        int m = 0;
        for (int w = 0; w < NW; ++w) m = cnt[w] > m ? cnt[w] : m;
        return m;
    }();
    short slot[NW][kSlots];
    signed char diag_slot[NT];
    constexpr TeamMap() : slot{}, diag_slot{} {
        int cnt[NW] = {};
        int next = 0;
        for (int w = 0; w < NW; ++w)
            for (int s = 0; s < kSlots; ++s) slot[w][s] = -1;
        for (int j = 0; j < NT; ++j)
            for (int i = j; i < NT; ++i) {
                int const w = i == j ? j % NW : (next++) % NW;
                if (i == j) diag_slot[j] = static_cast<signed char>(cnt[w]);
                slot[w][cnt[w]++] = static_cast<short>(i << 8 | j);
            }
    }
};

template <int NT, int NW>
inline constexpr TeamMap<NT, NW> kTeamMap{};
// (i << 8 | j) of slot s of wavefront W, or -1
template <int NT, int NW, int W, int s>
inline constexpr int kTeamTile = kTeamMap<NT, NW>.slot[W][s];

// Shared memory of one instance.
template <int NT, int RHS>
struct TeamLds {
    double sP[NT][kTileElems];  // panel tiles L_ik of the current block column, T[m][k]
    double sD[kTileElems];      // diagonal tile being factored -> L_kk
    double sW[kTileElems];      // W = L_kk^{-1} of the current block
    double sR[NT * 16 * RHS];   // right-hand sides -> y -> x, [row][RHS]
    int bad;
};

// The work of wavefront W of the instance: every tile coordinate is a
// compile-time constant, so the tiles stay in fixed registers and every LDS
// operand address is a constant offset from one per-lane base.  All
// wavefronts pass the same sequence of barriers.
template <int NT, int NW, int W, int RHS, bool PANEL, bool FULL>
__device__ __forceinline__ void team_wave(TeamLds<NT, RHS>& sm, Mat<double> const& A, Mat<double> const& L,
                                          Mat<double> const& B, Mat<double> const& X, int n, int nrhs, int64_t b,
                                          int32_t* info) {
    using Mf = Mfma<double>;
    constexpr int S = TeamMap<NT, NW>::kSlots;
    int const lane = threadIdx.x & 63;
    bool const lpanel = L.panel != 0;

    d4 acc[S];
    {
        Ref<double> const a(A, b);
        static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
            constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
            if constexpr (e >= 0)
                acc[SI] = FULL ? load_tile_full<PANEL, double, true>(a, e >> 8, e & 255, lane)
                               : load_tile<PANEL, double>(a, n, e >> 8, e & 255, lane);
            else
                acc[SI] = d4{0.0, 0.0, 0.0, 0.0};
        });
    }
    if (threadIdx.x == 0) sm.bad = 0;
    if (nrhs > 0) {
        Ref<double> const bb(B, b);
        for (int e = threadIdx.x; e < NT * 16 * RHS; e += 64 * NW) {
            int const row = e / RHS, c = e % RHS;
            sm.sR[e] = (row < n && c < nrhs) ? bb(row, c) : 0.0;
        }
    }
    __syncthreads();

    int const cn = lane & 15;
    // right-hand-side block blk: MFMA B operand row k, column lane % 16
    // (the loads are unconditional, in bounds for every lane, and the value
    // is selected afterwards: a conditional load would become a branch)
    auto rhs_b = [&](int blk, int k) -> double {
        double const t = sm.sR[(16 * blk + k) * RHS + (cn & (RHS - 1))];
        return cn < RHS ? t : 0.0;
    };
    auto rhs_load = [&](int blk) {
        d4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double const t = sm.sR[(16 * blk + Mf::row(lane, r)) * RHS + (cn & (RHS - 1))];
            v[r] = cn < RHS ? t : 0.0;
        }
        return v;
    };
    auto rhs_store = [&](int blk, d4 v) {
        if (cn < RHS)
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.sR[(16 * blk + Mf::row(lane, r)) * RHS + cn] = v[r];
    };

#pragma unroll 1
    for (int kb = 0;; ++kb) {
        // U(kb-1) and the lookahead factorisation of tile (kb, kb)
        if (kb % NW == W) {
            d4 dt = {0.0, 0.0, 0.0, 0.0};
            static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
                constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
                if constexpr (e >= 0 && (e >> 8) == (e & 255))
                    if (kb == (e & 255)) dt = acc[SI];
            });
            if (kb > 0)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    dt = Mf::mma(-op_read(sm.sP[kb], q, lane), op_read(sm.sP[kb], q, lane), dt);
            acc_to_lds(sm.sD, dt, lane);
            int const bk = diag_factor_inl<double>(sm.sD, sm.sW, lane);
            if (lane == 0 && bk && sm.bad == 0 && 16 * kb + bk <= n) sm.bad = 16 * kb + bk;
            {
                Ref<double> const l(L, b);
                store_tile<PANEL>(l, lpanel, n, kb, kb, lds_to_acc<double>(sm.sD, lane), lane);
            }
            d4 const wv = lds_to_acc<double>(sm.sW, lane);
            static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
                constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
                if constexpr (e >= 0 && (e >> 8) == (e & 255))
                    if (kb == (e & 255)) acc[SI] = wv;  // W_kk, kept for the backward solve
            });
            if (nrhs > 0) {
                d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    y = Mf::mma(op_read(sm.sW, q, lane), rhs_b(kb, 4 * q + (lane >> 4)), y);
                rhs_store(kb, y);
            }
        }
        if (kb > 0)
            static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
                constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
                if constexpr (e >= 0) {
                    constexpr int i = e >> 8, j = e & 255;
                    if (j >= kb && !(i == j && j == kb))
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[SI] = Mf::mma(-op_read(sm.sP[i], q, lane), op_read(sm.sP[j], q, lane), acc[SI]);
                }
            });
        if (kb == NT - 1) break;
        __syncthreads();
        // P(kb): panel tiles of block column kb
        static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
            constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
            if constexpr (e >= 0 && (e >> 8) > (e & 255)) {
                constexpr int i = e >> 8, j = e & 255;
                if (j == kb) {
                    acc_to_lds(sm.sP[i], acc[SI], lane);
                    d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) z = Mf::mma(op_read(sm.sP[i], q, lane), op_read(sm.sW, q, lane), z);
                    acc[SI] = z;
                    {
                        Ref<double> const l(L, b);
                        if constexpr (FULL)
                            store_tile_panel_full(l, opaque(i), opaque(j), z, lane);
                        else
                            store_tile<PANEL>(l, lpanel, n, opaque(i), opaque(j), z, lane);
                    }
                    acc_to_lds(sm.sP[i], z, lane);
                    if (nrhs > 0) {
                        d4 r = rhs_load(i);
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            r = Mf::mma(-op_read(sm.sP[i], q, lane), rhs_b(j, 4 * q + (lane >> 4)), r);
                        rhs_store(i, r);
                    }
                }
            }
        });
        __syncthreads();
    }

    if (nrhs > 0) {  // x = L^{-T} y: W_kk and L_kj from the registers (A operand of the transpose)
#pragma unroll 1
        for (int bk = NT - 1; bk >= 0; --bk) {
            __syncthreads();
            if (bk % NW == W) {
                d4 wv = {0.0, 0.0, 0.0, 0.0};
                static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
                    constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
                    if constexpr (e >= 0 && (e >> 8) == (e & 255))
                        if (bk == (e & 255)) wv = acc[SI];
                });
                d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) x = Mf::mma(wv[q], rhs_b(bk, Mf::row(lane, q)), x);
                rhs_store(bk, x);
            }
            __syncthreads();
            static_for<0, S>([&](auto SI) __attribute__((always_inline)) {
                constexpr int e = kTeamTile<NT, NW, W, decltype(SI)::value>;
                if constexpr (e >= 0 && (e >> 8) > (e & 255)) {
                    constexpr int i = e >> 8, j = e & 255;
                    if (i == bk) {
                        d4 y = rhs_load(j);
#pragma unroll
                        for (int q = 0; q < 4; ++q) y = Mf::mma(acc[SI][q], -rhs_b(i, Mf::row(lane, q)), y);
                        rhs_store(j, y);
                    }
                }
            });
        }
        __syncthreads();
        Ref<double> const xx(X, b);
        for (int e = threadIdx.x; e < n * nrhs; e += 64 * NW) {
            int const row = e % n, c = e / n;
            xx(row, c) = sm.sR[row * RHS + c];
        }
    } else {
        __syncthreads();
    }
    if (info && threadIdx.x == 0) info[b] = sm.bad;
}

template <int NT, int NW, int WPE, int RHS, bool PANEL, bool FULL>
__global__ __launch_bounds__(64 * NW, WPE) void chol_team_kernel(Mat<double> A, Mat<double> L, Mat<double> B,
                                                                 Mat<double> X, int n, int nrhs, int32_t* info,
                                                                 const int32_t* only_if) {
    __shared__ TeamLds<NT, RHS> sm;
    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;
    int const w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    static_for<0, NW>([&](auto WI) __attribute__((always_inline)) {
        if (w == decltype(WI)::value)
            team_wave<NT, NW, decltype(WI)::value, RHS, PANEL, FULL>(sm, A, L, B, X, n, nrhs, b, info);
    });
}

template <int NT, int NW, int WPE, int RHS>
int launch_team_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                   int64_t batch, int32_t* info, const int32_t* only_if) {
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    dim3 const grid = dim3(unsigned(batch)), block = dim3(64 * NW);
    if (panel && n == 16 * NT)
        hipLaunchKernelGGL((chol_team_kernel<NT, NW, WPE, RHS, true, true>), grid, block, 0, s, A, L, B, X, n, nrhs,
                           info, only_if);
    else if (panel)
        hipLaunchKernelGGL((chol_team_kernel<NT, NW, WPE, RHS, true, false>), grid, block, 0, s, A, L, B, X, n, nrhs,
                           info, only_if);
    else
        hipLaunchKernelGGL((chol_team_kernel<NT, NW, WPE, RHS, false, false>), grid, block, 0, s, A, L, B, X, n, nrhs,
                           info, only_if);
    return check_launch("chol_team_kernel");
}

}  // namespace team

// Launches the team kernel for NT 16-column blocks (n <= 16 NT) with NW
// wavefronts per instance and WPE wavefronts per SIMD; the larger NT are
// instantiated in translation units of their own (chol_team_nt*.hip) so
// that they compile in parallel.
template <int NT, int NW, int WPE>
int launch_team(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                int64_t batch, int32_t* info, const int32_t* only_if) {
    return team::launch_team_nt<NT, NW, WPE, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
}

}  // namespace blast_hip
