// Fused batched Cholesky factor + solve, FP64, n = 16 NT <= 128 with A and L
// in panel storage (every tile inside the matrix), on MFMA tiles, two
// wavefronts per SIMD.
//
// One wavefront per instance, left-looking by 16-column block columns, so
// that only the factor tiles still needed by later columns are held in
// registers (at most 24 of the 36 lower tiles at n = 128) and the kernel fits
// 256 registers and 20 KB of LDS: two instances share each SIMD, one
// instance's MFMA chains filling the other's dependency stalls (diagonal
// factorisation, loads).
//
// Register tiles are held transposed, in the layout in which the MFMA
// consumes them: register q of lane l holds M[l % 16][4q + l / 16], which is
// both the A operand of M and the B operand of M^T for K-chunk q, and is what
// v_mfma_f64_16x16x4f64 produces when it computes M^T.  So every product of
// the factorisation is formed transposed and nothing passes through LDS:
//   update   A_ij^T -= L_jk L_ik^T            (k < j <= i)
//   factor   the diagonal tile in the same layout, each column of L_jj
//            through LDS (chol16_cols)
//   inverse  W_j = L_jj^{-1} (tri_inv16_regs), in the same layout
//   panel    L_ij^T = W_j A_ij^T              (i > j)
//   forward  r_j = b_j - sum_k L_jk y_k, y_j = W_j r_j
// Column j+1 is loaded while column j is factored.  The backward solve
// x = L^{-T} y needs the other layout: it reads the factor tiles back from L
// (written by this wavefront; they are still in the cache hierarchy).
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// factorisation reorders BLAST's sums (agreement to backward error, not bits).
// HBM traffic per instance: the lower triangle of A is read, the lower
// triangle of L written, B read and X written (plus the re-read of L's
// off-diagonal tiles by the backward solve, served by L2 / MALL).

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

// LDS: one tile per block column, element (m, k) at k*17 + m (conflict-free
// whether the lanes of a 16-lane row vary m or k; every access is a per-lane
// base plus an immediate offset).  Tile j holds W_j = L_jj^{-1} in its lower
// triangle and the right-hand-side block j (b_j, then y_j: 16 rows x 4
// columns) in its strict upper triangle: rows 0..11 in columns 12..15, rows
// 12..15 in columns 8..11, i.e. row rr, column c at ys(rr, c).  While tile j
// is factored, the whole tile holds the multipliers -l_ij (zeros above the
// diagonal).
constexpr int kLlTile = 16 * kTileStride;
__device__ __forceinline__ int ys(int rr, int c) {
    return rr < 12 ? (12 + c) * kTileStride + rr : (8 + c) * kTileStride + rr - 12;
}

template <int NT>
struct LlLds {
    double sW[NT][kLlTile];
    double sR[16];  // r_k = 1/sqrt(pivot k) of the diagonal tile being factored
};

// Panel storage (PS = 4) offsets of the transposed register layout:
// element (16 ti + l % 16, 16 tj + 4q + l / 16) of tile (ti, tj) is at
// (4 ti) spacing + 64 tj + 16 q  +  ((l % 16) / 4) spacing + l % 4 + 4 (l / 16).
__device__ __forceinline__ int t_lane_off(int lane, int spacing) {
    return ((lane & 15) >> 2) * spacing + (lane & 3) + 4 * (lane >> 4);
}

// Off-diagonal tile (ti, tj) of A, transposed layout.
__device__ __forceinline__ d4 load_t(const double* p, int spacing, int lo, int ti, int tj) {
    const double* const q = p + int64_t(opaque_s(4 * ti * spacing)) + 64 * tj + lo;
    return d4{q[0], q[16], q[32], q[48]};
}

// Diagonal tile (tj, tj) of A, transposed layout, from the lower triangle:
// element (m, k) with k > m is read from its mirror (k, m).
__device__ __forceinline__ d4 load_t_diag(const double* p, int spacing, int tj, int lane_) {
    int const lane = opaque_v(lane_), m = lane & 15, g = lane >> 4;
    const double* const base = p + int64_t(opaque_s(4 * tj * spacing)) + 64 * tj;
    d4 v;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int const k = 4 * q + g, r = m > k ? m : k, c = m > k ? k : m;
        v[q] = base[(r >> 2) * spacing + (r & 3) + 4 * c];
    }
    return v;
}

// Accumulator layout of tile (ti, tj) (register r of lane l = M[l/16 + 4r][l%16]).
__device__ __forceinline__ d4 load_acc(const double* p, int spacing, int ti, int tj, int lane) {
    const double* const q = p + int64_t(opaque_s(4 * ti * spacing)) + 64 * tj + (lane >> 4) + 4 * (lane & 15);
    return d4{q[0], q[spacing], q[2 * spacing], q[3 * spacing]};
}

// The operations of step j that do not wait for its diagonal factorisation,
// issued between the 16 steps of that factorisation (their MFMAs fill the
// dependency stalls of its VALU chain), in this order:
//   kind 0: the deferred panel of block column j-1, L_i,j-1^T = W_j-1 A_i,j-1^T
//           for i = j+1..NT-1 (tile (j, j-1) was formed at the end of step
//           j-1), four K-chunks per tile;
//   kind 1: y_j-1 = W_j-1 r_j-1, four K-chunks;
//   kind 3: the next diagonal tile's update A_j+1,j+1 -= L_j+1,k L_j+1,k^T
//           for k < j (its k = j term opens step j+1), k outer, q inner;
//   kind 2: for k < j (outer), K-chunk q, then row i = j+1..NT-1 (the update
//           A_ij^T -= L_jk L_ik^T) and i = NT (the forward update
//           r_j -= L_jk y_k), so that consecutive MFMAs accumulate into
//           different tiles.
// The k = j-1 terms of kinds 3 and 2 come after the panel tiles and y_j-1
// they read.  Encoded as (kind << 24) | (k << 16) | (q << 8) | i.
template <int NT, int J>
struct StepOps {
    static constexpr int panel = J >= 1 ? 4 * (NT - 1 - J) : 0;
    static constexpr int yops = J >= 1 ? 4 : 0;
    static constexpr int nextdiag = J + 1 < NT ? 4 * J : 0;
    static constexpr int rows = NT - J;  // NT - 1 - J tiles and the right-hand sides
    static constexpr int ops = panel + yops + nextdiag + J * 4 * rows;
    static constexpr int encode(int o) {
        if (o < panel) return (0 << 24) | ((o % 4) << 8) | (J + 1 + o / 4);
        o -= panel;
        if (o < yops) return (1 << 24) | (o << 8);
        o -= yops;
        if (o < nextdiag) return (3 << 24) | ((o / 4) << 16) | ((o % 4) << 8) | (J + 1);
        o -= nextdiag;
        int const k = o / (4 * rows), q = (o / rows) % 4, i = J + 1 + o % rows;
        return (2 << 24) | (k << 16) | (q << 8) | i;
    }
};

template <int NT>
__global__ __launch_bounds__(64, 2) void posv_ll_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                        int nrhs, int32_t* info, const int32_t* only_if) {
    using M = Mfma<double>;
    constexpr int NTILES = NT * (NT + 1) / 2;
    constexpr int NR = NT * 16 * kMaxRhs;
    __shared__ LlLds<NT> sm;

    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;  // fallback mode: only the flagged instances
    int const lane = threadIdx.x;
    int const m = lane & 15, g = lane >> 4;
    const double* const pa = A.inst(b);
    double* const pl = L.inst(b);
    int const sa = A.spacing, sl = L.spacing;
    int const loa = t_lane_off(lane, sa), lol = t_lane_off(lane, sl);

    // t[tri(i, j)]: tile (i, j) of A while column j is updated, then L_ij;
    // transposed layout.
    d4 t[NTILES];
    // column j's tiles below the diagonal and the next diagonal tile (j+1,
    // j+1), whose update starts during step j (column j's own diagonal tile
    // came with column j-1)
    auto load_col = [&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        if constexpr (j + 1 < NT) t[tri(j + 1, j + 1)] = load_t_diag(pa, sa, j + 1, lane);
        static_for<j + 1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            t[tri(i, j)] = load_t(pa, sa, loa, i, j);
        });
    };
    t[tri(0, 0)] = load_t_diag(pa, sa, 0, lane);
    load_col(std::integral_constant<int, 0>{});
    if (nrhs > 0) {  // every load issued before the first LDS store
        Ref<double> bb(B, b);
        double v[NR / 64];
#pragma unroll
        for (int e = 0; e < NR / 64; ++e) {
            int const idx = lane + 64 * e, row = idx / kMaxRhs, c = idx % kMaxRhs;
            v[e] = c < nrhs ? bb(row, c) : 0.0;
        }
#pragma unroll
        for (int e = 0; e < NR / 64; ++e) {
            int const idx = lane + 64 * e, row = idx / kMaxRhs, c = idx % kMaxRhs;
            sm.sW[row >> 4][ys(row & 15, c)] = v[e];
        }
    }

    // Right-hand-side blocks (16 rows x kMaxRhs columns) live in LDS and take
    // part in v_mfma_f64_4x4x4f64 products: four 4x4x4 blocks b = 0..3, A
    // operand A_b[i][k] in lane 16k + 4b + i, B operand B_b[k][c] in lane
    // 16k + 4b + c, result D_b[i][c] in lane 16i + 4b + c.  With block b =
    // rows 4b..4b+3 of a 16-row product, register q of the transposed layout
    // of M is the A operand of M (K-chunk q), and the accumulator layout's
    // that of M^T.
    // rhs_b(k, q): block k as the B operand of K-chunk q, rows 4q + l/16,
    // column l%4 (the same for the four 4x4 blocks).
    int const ybase = kTileStride * (m & (kMaxRhs - 1)) + g;  // ys(4q + g, m%4) - ys(4q, 0)
    auto rhs_b = [&](int k, int q) -> double { return sm.sW[k][ybase + ys(4 * q, 0)]; };
    // the result layout: row 4 (l%16 / 4) + l/16, column l%4
    int const yd = ys(4 * (m >> 2) + g, m & (kMaxRhs - 1));
    auto mma4 = [](double a, double bv, double c) { return __builtin_amdgcn_mfma_f64_4x4x4f64(a, bv, c, 0, 0, 0); };
    int inf = 0;
    // factor tiles L_kb,j in the accumulator layout, read back for the
    // backward solve (row kb, j < kb)
    d4 lt[NTILES];
    auto load_row = [&](auto KB) __attribute__((always_inline)) {
        constexpr int kb = decltype(KB)::value;
        static_for<0, kb>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            lt[tri(kb, j)] = load_acc(pl, sl, kb, j, lane);
        });
    };

    auto store_l = [&](int i, int jj, d4 v) __attribute__((always_inline)) {  // tile (i, jj), transposed layout
        double* const q0 = pl + int64_t(opaque_s(4 * i * sl)) + 64 * jj + lol;
#pragma unroll
        for (int q = 0; q < 4; ++q) q0[16 * q] = v[q];
    };
    double wp[4] = {0.0, 0.0, 0.0, 0.0};  // W of the previous block column
    d4 pz;                                // deferred panel tile being formed
    double ypz[2];

    static_for<0, NT>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        using Ops = StepOps<NT, j>;
        double* const sw = sm.sW[j];
        if constexpr (j == NT - 1 && NT > 1) {
            // every off-diagonal factor tile is stored: the backward solve's
            // read-back of the last row starts here, after a barrier (a
            // workgroup-scope release and acquire; the workgroup is this one
            // wavefront) that makes the stores visible to the other lanes
            if (nrhs > 0) {
                __syncthreads();
                load_row(std::integral_constant<int, NT - 1>{});
            }
        }
        // diagonal tile: A_jj -= L_j,j-1 L_j,j-1^T (whole tile, symmetric; the
        // terms k < j-1 were applied during step j-1)
        if constexpr (j >= 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                t[tri(j, j)] = M::mma(-t[tri(j, j - 1)][q], t[tri(j, j - 1)][q], t[tri(j, j)]);
        }
        // forward: r_j = b_j - sum_k L_jk y_k (result layout)
        double r = 0.0;
        if (nrhs > 0) r = sw[yd];
        // the diagonal factorisation in the transposed layout, each column
        // of L_jj through the LDS tile (b_j has been read: its slots are
        // free until r is stored back), with the column's other updates and
        // the forward update spread over its 16 steps
        double z[4];  // the inverse's running rows, W_j = -z r (chol16_cols_inv)
        chol16_cols_inv(t[tri(j, j)], sw, sm.sR, lane, z, [&](auto S) __attribute__((always_inline)) {
            constexpr int s = decltype(S)::value;
            static_for<Ops::ops * s / 16, Ops::ops * (s + 1) / 16>([&](auto O) __attribute__((always_inline)) {
                constexpr int e = Ops::encode(decltype(O)::value);
                constexpr int kind = e >> 24, k = (e >> 16) & 255, q = (e >> 8) & 255, i = e & 255;
                if constexpr (kind == 0) {  // deferred panel tile (i, j-1)
                    pz = M::mma(wp[q], t[tri(i, j - 1)][q], q == 0 ? d4{0.0, 0.0, 0.0, 0.0} : pz);
                    if constexpr (q == 3) {
                        t[tri(i, j - 1)] = pz;
                        store_l(i, j - 1, pz);
                    }
                } else if constexpr (kind == 3) {  // next diagonal tile, term k
                    t[tri(j + 1, j + 1)] = M::mma(-t[tri(j + 1, k)][q], t[tri(j + 1, k)][q], t[tri(j + 1, j + 1)]);
                } else if constexpr (kind == 1) {  // y_j-1 = W_j-1 r_j-1 (r_j-1 in LDS)
                    if constexpr (q < 2) ypz[q] = mma4(wp[q], rhs_b(j - 1, q), 0.0);
                    else ypz[q - 2] = mma4(wp[q], rhs_b(j - 1, q), ypz[q - 2]);
                    if constexpr (q == 3) sm.sW[j - 1][yd] = ypz[0] + ypz[1];
                } else if constexpr (i < NT) {
                    t[tri(i, j)] = M::mma(-t[tri(j, k)][q], t[tri(i, k)][q], t[tri(i, j)]);
                } else {  // (without a branch on nrhs: r is not used when nrhs = 0)
                    r = mma4(-t[tri(j, k)][q], rhs_b(k, q), r);
                }
            });
        });
        double dt[4];  // L_jj in the transposed layout (L[m][4u + g])
#pragma unroll
        for (int u = 0; u < 4; ++u) dt[u] = sw[(4 * u + g) * kTileStride + m];
        // the next column's tiles, while this one is inverted and applied
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (j + 1 < NT) load_col(std::integral_constant<int, j + 1>{});
        if constexpr (j == NT - 1 && NT > 2)
            if (nrhs > 0) load_row(std::integral_constant<int, NT - 2>{});
        // L_jj: lower part; zeros in the strict upper part of each 4x4
        // diagonal block (panel storage)
        {
            double* const q0 = pl + int64_t(opaque_s(4 * j * sl)) + 64 * j + lol;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                int const c = 4 * u + g;
                if (c <= m || (c >> 2) == (m >> 2)) q0[16 * u] = c <= m ? dt[u] : 0.0;
            }
        }
        // W_j = L_jj^{-1} = -z r (formed with the factorisation), l_ii from
        // the diagonal of the LDS factor for the pivot check
        double const lii = sw[m * kTileStride + m];
        int const bad = first_bad_pivot(lii, lane);
        // (settled now: left to the end, the eight steps' ballots stay live
        // in SGPR pairs and spill)
        inf = opaque_s(bad && inf == 0 ? 16 * j + bad : inf);
        double const rm = sm.sR[m];
        double w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) w[u] = (4 * u + g <= m) ? -z[u] * rm : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)  // lower triangle, for the backward solve
            if (4 * u + g <= m) sw[(4 * u + g) * kTileStride + m] = w[u];
        // r_j final: kept in LDS (B-operand layout) for y_j = W_j r_j, which
        // the next step forms among its hooked operations
        sw[yd] = r;
        // the panel tile the next diagonal tile waits for; the others are
        // formed during the next diagonal factorisation
        if constexpr (j + 1 < NT) {
            d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) z = M::mma(w[q], t[tri(j + 1, j)][q], z);
            t[tri(j + 1, j)] = z;
            store_l(j + 1, j, z);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) wp[q] = w[q];
    });
    {  // y of the last block column
        double ya = mma4(wp[0], rhs_b(NT - 1, 0), 0.0), yb = mma4(wp[1], rhs_b(NT - 1, 1), 0.0);
        ya = mma4(wp[2], rhs_b(NT - 1, 2), ya);
        yb = mma4(wp[3], rhs_b(NT - 1, 3), yb);
        sm.sW[NT - 1][yd] = ya + yb;
    }

    if (nrhs > 0) {
        // right-looking: x_kb = W_kb^T r_kb (r_kb final), then
        // r_j -= L_kb,j^T x_kb for j < kb; each result passes through LDS
        // into the B-operand layout.
        double r[NT];
        static_for<0, NT>([&](auto I) __attribute__((always_inline)) { r[I] = sm.sW[I][yd]; });
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            if constexpr (kb >= 3) load_row(std::integral_constant<int, kb - 2>{});
            double* const sw = sm.sW[kb];
            // A operand W^T: W[4q + l/16][l%16], zero above the diagonal
            // (where the tile holds y)
            auto wt = [&](int q) {
                double const v = sw[m * kTileStride + 4 * q + g];
                return 4 * q + g >= m ? v : 0.0;
            };
            sw[yd] = r[kb];
            double x = mma4(wt(0), rhs_b(kb, 0), 0.0);
            double x1 = mma4(wt(1), rhs_b(kb, 1), 0.0);
            x = mma4(wt(2), rhs_b(kb, 2), x);
            x1 = mma4(wt(3), rhs_b(kb, 3), x1);
            r[kb] = x + x1;
            if constexpr (kb > 0) {
                sw[yd] = r[kb];
                double xb[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) xb[q] = -rhs_b(kb, q);
                static_for<0, kb>([&](auto JJ) __attribute__((always_inline)) {
                    constexpr int j = kb - 1 - decltype(JJ)::value;
#pragma unroll
                    for (int q = 0; q < 4; ++q) r[j] = mma4(lt[tri(kb, j)][q], xb[q], r[j]);
                });
            }
        });
        Ref<double> xx(X, b);
        if ((m & (kMaxRhs - 1)) < nrhs)
            static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                xx(16 * i + 4 * (m >> 2) + g, m & (kMaxRhs - 1)) = r[i];
            });
    }
    if (info && lane == 0) info[b] = inf;
}

template <int NT>
int launch_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int nrhs, int64_t batch,
              int32_t* info, const int32_t* only_if) {
    hipLaunchKernelGGL((posv_ll_kernel<NT>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, nrhs, info, only_if);
    return check_launch("posv_ll_kernel");
}

}  // namespace

// Fused potrf (+ forward/backward solve when nrhs > 0) for n = 16 NT <= 128
// with A and L in (untransposed) panel storage, nrhs <= 4.  L may alias A and
// X may alias B.  Returns 1 when the shape is not covered.
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
