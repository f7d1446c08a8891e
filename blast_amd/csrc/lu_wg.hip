// Batched FP64 LU with partial pivoting and solve (LAPACK dgetrf + dgetrs,
// BLAST getrf), square n <= 256, nrhs <= 8: A is factored in place (unit
// lower L strictly below the diagonal, U on and above), the pivots go to
// ipiv, X = A^{-1} B.
//
// One 256-thread workgroup (4 wavefronts) per instance.  The factor lives in
// global memory (512 KB at n = 256 does not fit one CU) and is updated
// right-looking by 32-column block columns kb (rows r0 = 32 kb..), so that
// each trailing tile makes one round trip through HBM per 32 columns
// (2.3 MB per instance at n = 256, against 5 MB for 16-column blocks) and
// the per-block phases run n / 32 times:
//   panel: the block column's rows r0.. are held in registers, one row of
//     32 values per thread, and factored there column by column (getf2).
//     Per column each wavefront reduces its (|value|, position) pairs with
//     DPP and lane swaps and posts its winning row to LDS; after one
//     barrier every thread combines the four candidates itself.  Rows never
//     move: a swap only exchanges the positions two threads record, and
//     each row is written to its final position after the panel.  The
//     multipliers and the rank-1 update stay in registers: one barrier per
//     column.
//   swaps: applied to the columns right of the panel from the panel's
//     position table (which row each position received), 32 lanes per
//     column: rows r0..r0+31 land in LDS as the block row A_kj and the
//     displaced rows are written back; the columns left of the panel receive
//     every later block's swaps at the end, in one pass per block, from the
//     composite row permutation (each thread traces its row back through
//     the swaps).
//   U_kj = L_kk^{-1} A_kj and the trailing update A_ij -= L_ik U_kj on MFMA
//     (K = 32: two 16x16 tiles per operand), both computed transposed
//     (U_kj^T = A_kj^T L_kk^{-T}, A_ij^T -= U_kj^T L_ik^T) so that each
//     lane's accumulator elements run down a column of A: coalesced
//     column-major loads and stores.  The trailing tiles stream through
//     registers eight per wavefront, the next group's loads issued before
//     the current group's MFMAs.
// Solve: B's rows permuted once (the composite permutation of all swaps),
// the 16x16 diagonal blocks' inverses (unit-lower and upper) formed in
// parallel, then block substitution on MFMA.
// Algorithm references: dense/Getrf.hpp:52-119 (blocking, ipiv += k, laswp
// on both sides, trsm, gemm), dense/Getf2.hpp:52-77 (pivot, swap, division
// by the pivot, ger), dense/Iamax.hpp:82-95 (first maximum),
// dense/Laswp.hpp:37-43.  Within the panel the per-element operations are
// getf2's (a division by the pivot, one fma per update, in column order);
// the trailing sums are reordered on MFMA tiles, so pivots agree with the
// sequential algorithm except at near-ties.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kLuThreads = 256;
constexpr int kLuWaves = kLuThreads / 64;
constexpr int kLuMaxNT = 16;  // n <= 256: one panel row per thread
constexpr int kLuBs = 32;     // block column width (two 16x16 tiles)
constexpr int kLuRhs = 8;     // right-hand-side columns per instance
constexpr int kLuGroup = 5;   // trailing tiles per wavefront in flight

// One step of the wavefront argmax: combine with the lane that DPP control
// CTRL names (quad_perm, row_half_mirror, row_mirror: all within a row).
template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& v, int& i) {
    unsigned long long const u = __builtin_bit_cast(unsigned long long, v);
    unsigned const lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u)), CTRL, 0xF, 0xF, false));
    unsigned const hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u >> 32)), CTRL, 0xF, 0xF, false));
    int const i2 = __builtin_amdgcn_mov_dpp(i, CTRL, 0xF, 0xF, false);
    argmax_combine(v, i, __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo), i2);
}
// Across 16-lane rows: v_permlane16_swap (odd/even rows) or
// v_permlane32_swap (halves) hands every lane both rows' pairs.
template <bool ROW16>
__device__ __forceinline__ void argmax_swap(double& v, int& i) {
    unsigned long long const u = __builtin_bit_cast(unsigned long long, v);
    unsigned l0, l1, h0, h1, i0, i1;
    lane_swap<ROW16>(unsigned(u), l0, l1);
    lane_swap<ROW16>(unsigned(u >> 32), h0, h1);
    lane_swap<ROW16>(unsigned(i), i0, i1);
    v = __builtin_bit_cast(double, (static_cast<unsigned long long>(h0) << 32) | l0);
    i = int(i0);
    argmax_combine(v, i, __builtin_bit_cast(double, (static_cast<unsigned long long>(h1) << 32) | l1), int(i1));
}
// The wavefront's (key, row) argmax under argmax_combine (an order-free
// combination), in every lane, without LDS.
__device__ __forceinline__ void wave_argmax(double& v, int& i) {
    argmax_dpp<0xB1>(v, i);   // quad_perm [1,0,3,2]
    argmax_dpp<0x4E>(v, i);   // quad_perm [2,3,0,1]
    argmax_dpp<0x141>(v, i);  // row_half_mirror
    argmax_dpp<0x140>(v, i);  // row_mirror
    argmax_swap<true>(v, i);
    argmax_swap<false>(v, i);
}

// T: the storage type (the arithmetic is FP64 for FP32 storage too: values
// are widened on load and rounded on store).
template <typename T>
struct LuArgs {
    Mat<T> A, B, X;
    int n, nrhs;
    int32_t* info;
    const int32_t* only_if;
    int64_t* ipiv;
    int64_t ipiv_stride;
};

// LDS of one instance (80.4 KB: two instances per CU); the factor phase's
// tiles and the solve phase's overlay the same storage.  Tiles are 16x16
// with element (row m, col c) at c*17 + m.
struct LuFactorLds {
    double sP[kLuMaxNT][2][kTileElems];  // panel: tile (row tile t, column tile s) of rows r0 + 16 t..
    double sW[3][kTileElems];            // L_kk^{-1} (unit lower, 32 x 32): W00, W10, W11
    double sRow[2][kLuWaves][kLuBs];     // each wavefront's candidate pivot row
    double sCandV[2][kLuWaves];
    int sCandI[2][kLuWaves];
    int sNonFinite[kLuWaves];
};
struct LuSolveLds {
    double sInv[kLuMaxNT][kTileElems];  // U_kk^{-1}
    double sX[kLuMaxNT * 16 * kLuRhs];  // right-hand sides, [row][rhs]
};

// Element offsets of one instance's storage in 32 bits (the launcher checks
// that every offset of the matrix fits), as row part + column part:
// off(i, j) = rows(i) + cols(j), the panel or dense map of Mat::off.  Each
// part is (x >> s) * a + (x & msk) * c with per-matrix constants: the panel
// index (x / PS) spacing + x % PS, the dense row x, or the column x * PS /
// x * spacing, with no branch on the storage kind.
struct AxisMap {
    int s, a, msk, c;
    __device__ __forceinline__ int operator()(int x) const { return (x >> s) * a + (x & msk) * c; }
};
template <typename T>
struct Sep {
    T* p;
    AxisMap rows, cols;
    // a 32-bit byte offset from the instance base: the uniform base stays in
    // SGPRs and every address costs one VGPR
    __device__ __forceinline__ T& at(int off) const {
        return *reinterpret_cast<T*>(reinterpret_cast<char*>(p) + unsigned(off) * unsigned(sizeof(T)));
    }
    static __device__ __forceinline__ Sep make(Mat<T> const& m, T* p) {
        constexpr int PS = PanelSize<T>::value, LOG = PanelSize<T>::log2;
        AxisMap const P = m.panel ? AxisMap{LOG, m.spacing, PS - 1, 1} : AxisMap{31, 0, 0x7fffffff, 1};
        AxisMap const Q{31, 0, 0x7fffffff, m.panel ? PS : m.spacing};
        return Sep{p, m.trans ? Q : P, m.trans ? P : Q};
    }
};

// The value of a load, opaque to the compiler: a select on it cannot be turned
// into a branch around the load (which would issue the loads one at a time).
__device__ __forceinline__ double pin(double x) {
    asm volatile("" : "+v"(x));
    return x;
}

// B operand of a 16x16 LDS tile M (element (m, c) at c*17 + m) for K-chunk q:
// M[4q + l/16][l%16].
__device__ __forceinline__ double op_read_t(const double* tile, int q, int lane) {
    return tile[(lane & 15) * kTileStride + 4 * q + (lane >> 4)];
}

template <typename T>
__global__ __launch_bounds__(kLuThreads, 2) void lu_wg_kernel(LuArgs<T> g) {
    using M = Mfma<double>;
    __shared__ union {
        LuFactorLds f;
        LuSolveLds s;
    } sm;
    __shared__ int sPiv[kLuMaxNT * 16];  // pivot row of every row (absolute)
    __shared__ int sSrc[kLuMaxNT * 16];  // panel position -> its row before the block's swaps
    __shared__ int sFlag;
    auto& sP = sm.f.sP;
    auto& sW = sm.f.sW;
    auto& sRow = sm.f.sRow;
    auto& sCandV = sm.f.sCandV;
    auto& sCandI = sm.f.sCandI;

    int64_t const b = blockIdx.x;
    if (g.only_if && g.only_if[b] >= 0) return;
    int const tid = threadIdx.x, lane = tid & 63;
    int const w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const m = lane & 15, gq = lane >> 4;
    int const n = g.n, nrhs = g.nrhs, NB = (n + kLuBs - 1) / kLuBs, npad = kLuBs * NB, NT = 2 * NB;
    bool const rhs = nrhs > 0;
    Sep<T> const sa = Sep<T>::make(g.A, g.A.inst(b));
    // A outside n x n reads as the identity (the factorisation of blkdiag(A,
    // I)); loads are clamped into the matrix and their value replaced
    auto fget = [&](int i, int j) -> double {
        double const v = double(sa.at(sa.rows(min(i, n - 1)) + sa.cols(min(j, n - 1))));
        return (i < n && j < n) ? v : (i == j ? 1.0 : 0.0);
    };
    // the right-hand sides ride along as column tile NT of [A | X]: the row
    // swaps and the elimination turn X = B into y = L^{-1} P B
    Sep<T> const sx = Sep<T>::make(g.X, rhs ? g.X.inst(b) : nullptr);
    auto xget = [&](int i, int c) -> double {
        double const v = double(sx.at(sx.rows(min(i, n - 1)) + sx.cols(min(c, nrhs - 1))));
        return (i < n && c < nrhs) ? v : 0.0;
    };
    if (rhs && g.X.data != g.B.data) {  // X = B (X may alias B)
        Ref<T> bb(g.B, b);
        for (int e = tid; e < n * nrhs; e += kLuThreads) {
            int const r = e % n, c = e / n;
            sx.at(sx.rows(r) + sx.cols(c)) = bb(r, c);
        }
    }
    if (tid == 0) sFlag = 0;

    for (int kb = 0; kb < NB; ++kb) {
        int const r0 = kLuBs * kb, P = npad - r0, T0 = 2 * kb + 2, mt = NT - T0;
        // ---- panel: row r0 + tid in registers.  Rows do not move: each
        // thread tracks the position its row holds under the swaps so far
        // (pos < c: a finished row of U), and the pivot search breaks ties by
        // position, as the sequential scan does.
        bool const live = tid < P;
        int pos = tid;
        double v[kLuBs];
        // full: every column of the block lies inside the matrix (uniform),
        // so that per-column conditions are needed only in the last block
        bool const full = r0 + kLuBs <= n;
        {
            // (the identity fill of the padding is formed here, not hoisted
            // out of the block loop as 32 loop-invariant registers)
            int const me = opaque_v(tid);
            int const rowoff = sa.rows(min(r0 + me, n - 1));
            bool const rowok = r0 + me < n;
#pragma unroll
            for (int c = 0; c < kLuBs; ++c) v[c] = double(sa.at(rowoff + sa.cols(min(r0 + c, n - 1))));
            if (full) {
#pragma unroll
                for (int c = 0; c < kLuBs; ++c) v[c] = rowok ? pin(v[c]) : (me == c && live ? 1.0 : 0.0);
            } else {
#pragma unroll
                for (int c = 0; c < kLuBs; ++c)
                    v[c] = (rowok && r0 + c < n) ? pin(v[c]) : (me == c && live ? 1.0 : 0.0);
            }
        }
        // The update of a column runs on every row, finished ones with a zero
        // multiplier (no per-lane branch, no selects); that leaves them exact
        // unless the pivot row holds an infinity or a NaN, so a panel with a
        // non-finite entry takes the masked update instead.
        bool nf = false;
#pragma unroll
        for (int c = 0; c < kLuBs; ++c) nf |= !(__builtin_fabs(v[c]) <= __DBL_MAX__);
        if (lane == 0) sm.f.sNonFinite[w] = __ballot(nf) != 0;
        __syncthreads();
        bool const masked = sm.f.sNonFinite[0] | sm.f.sNonFinite[1] | sm.f.sNonFinite[2] | sm.f.sNonFinite[3];
        auto panel = [&](auto MASKED) __attribute__((always_inline)) {
            constexpr bool kMasked = decltype(MASKED)::value;
            static_for<0, kLuBs>([&](auto C) __attribute__((always_inline)) {
                constexpr int c = decltype(C)::value, buf = c & 1;
                // The pivot is the candidate (position >= c) with the largest
                // key, the lowest position among equal keys: the first maximum
                // of BLAST's scan from position c with a strict '>'.  Keys:
                // |v|; a NaN at position c (the scan's first element, never
                // replaced) +inf; a NaN elsewhere -1 (never chosen); no
                // candidate -2.  So the keys hold no NaN, and the search is a
                // wavefront maximum plus a ballot for the lowest position
                // holding it (the full (key, position) reduction only on a
                // tie).
                bool const cand = live && pos >= c;
                double const av = __builtin_fabs(v[c]);
                double const key = !cand ? -2.0 : (av != av ? (pos == c ? __builtin_inf() : -1.0) : av);
                double const wm = wave_max(key);
                int wpos = 0x7fffffff;
                if (wm > -2.0) {
                    unsigned long long const tie = __ballot(key == wm);
                    if (__popcll(tie) == 1) {
                        wpos = __builtin_amdgcn_readlane(pos, int(__builtin_ctzll(tie)));
                    } else {
                        double kv = key;
                        int ki = cand ? pos : 0x7fffffff;
                        wave_argmax(kv, ki);
                        wpos = ki;
                    }
                }
                if (pos == wpos)
#pragma unroll
                    for (int cc = c; cc < kLuBs; ++cc) sRow[buf][w][cc] = v[cc];
                if (lane == 0) sCandV[buf][w] = wm, sCandI[buf][w] = wpos;
                __syncthreads();
                double const k0 = sCandV[buf][0], k1 = sCandV[buf][1], k2 = sCandV[buf][2], k3 = sCandV[buf][3];
                int const p0 = sCandI[buf][0], p1 = sCandI[buf][1], p2 = sCandI[buf][2], p3 = sCandI[buf][3];
                double const km = __builtin_fmax(__builtin_fmax(k0, k1), __builtin_fmax(k2, k3));
                int const pp = min(min(k0 == km ? p0 : 0x7fffffff, k1 == km ? p1 : 0x7fffffff),
                                   min(k2 == km ? p2 : 0x7fffffff, k3 == km ? p3 : 0x7fffffff));
                int const wp = p1 == pp ? 1 : (p2 == pp ? 2 : (p3 == pp ? 3 : 0));
                // position c always holds a candidate, so pp is a real position
                pos = pos == pp ? c : (pos == c ? pp : pos);
                if (tid == 0) sPiv[r0 + c] = r0 + pp;
                double const* const prow = sRow[buf][wp];
                double const piv = prow[c];
                if (piv == 0.0 && tid == 0 && sFlag == 0 && r0 + c < n) sFlag = r0 + c + 1;
                bool const upd = live && pos > c && piv != 0.0;
                double const l = v[c] / piv;
                if constexpr (kMasked) {
                    if (upd) {
                        v[c] = l;
#pragma unroll
                        for (int cc = c + 1; cc < kLuBs; ++cc) v[cc] = __builtin_fma(-l, prow[cc], v[cc]);
                    }
                } else {
                    double const lm = upd ? l : 0.0;
                    v[c] = upd ? l : v[c];
#pragma unroll
                    for (int cc = c + 1; cc < kLuBs; ++cc) v[cc] = __builtin_fma(-lm, prow[cc], v[cc]);
                }
                // the next column's search starts after this update (no
                // overlap: keeps the pivot row's 31 values and the next
                // search's registers from being live together)
                __builtin_amdgcn_sched_barrier(0);
            });
        };
        if (masked)
            panel(std::true_type{});
        else
            panel(std::false_type{});
        if (live) {
            sSrc[pos] = tid;
#pragma unroll
            for (int cc = 0; cc < kLuBs; ++cc) sP[pos >> 4][cc >> 4][(cc & 15) * kTileStride + (pos & 15)] = v[cc];
            if (r0 + pos < n) {
                int const rowoff = sa.rows(r0 + pos);
                if (full) {
#pragma unroll
                    for (int cc = 0; cc < kLuBs; ++cc) sa.at(rowoff + sa.cols(r0 + cc)) = T(v[cc]);
                } else {
#pragma unroll
                    for (int cc = 0; cc < kLuBs; ++cc)
                        if (r0 + cc < n) sa.at(rowoff + sa.cols(r0 + cc)) = T(v[cc]);
                }
            }
        }
        __syncthreads();
        if (tid < kLuBs && r0 + tid < n) g.ipiv[b * g.ipiv_stride + r0 + tid] = sPiv[r0 + tid];

        // ---- W = L_kk^{-1} (32 x 32 unit lower): the diagonal 16x16 blocks'
        // inverses on two wavefronts (tri_inv16_regs: lane (g, i) forms row
        // i's columns 4u + g), then W10 = -W11 (L10 W00) on MFMA
        if (w < 2) {
            double const* const lt = sP[w][w];
            double* const wt = sW[2 * w];
            int const i = lane & 15;
            double wv[4];
            tri_inv16_regs(wv, lane, 1.0, [&](int j) { return j < i ? -lt[j * kTileStride + i] : 0.0; }, NoHook{});
#pragma unroll
            for (int u = 0; u < 4; ++u) wt[(4 * u + gq) * kTileStride + i] = wv[u];
        }
        __syncthreads();
        if (w == 0) {
            d4 t = {0.0, 0.0, 0.0, 0.0}, u = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) t = M::mma(op_read(sP[1][0], q, lane), op_read_t(sW[0], q, lane), t);
#pragma unroll
            for (int q = 0; q < 4; ++q) u = M::mma(-op_read(sW[2], q, lane), t[q], u);
            acc_to_lds(sW[1], u, lane);
        }
        __syncthreads();

        // ---- the columns right of the panel, and the right-hand sides: each
        // wavefront owns whole column tiles (tile T0 + w, T0 + w + 4, ...; tile
        // NT is X).  Per tile every load comes first, through the block's row
        // permutation (row R now holds row r0 + sSrc[R - r0]): the block row
        // A_kj as a B operand, the trailing tiles in the accumulator layout;
        // U_kj = W A_kj leaves the MFMA in the layout of the trailing
        // update's A operand (U_kj^T), A_ij^T -= U_kj^T L_ik^T (K = 32, L_ik
        // from the panel tiles); then U_kj and the trailing tiles are stored.
        // No row is stored before every load of the tile has returned (the
        // stores depend on all of them), so the swaps need no separate pass.
        // Then the tiles left of the panel (columns < r0, the L of earlier
        // blocks) take the block's swaps: rows r0.. of the block receive rows
        // r0 + sSrc[i], and each row below that a swap displaced (position p
        // = ipiv - r0 >= 32) receives row r0 + sSrc[p], a row of the block.
        int const ntile = mt + (rhs ? 1 : 0), nleft = 2 * kb;
        for (int jt = w; jt < ntile + nleft; jt += kLuWaves) {
            if (jt >= ntile) {
                int const col = 16 * (jt - ntile) + m, cpc = sa.cols(col);
                double bv[2][4], dv[2][4];
                int drow[2][4];
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        int const i = 16 * t + 4 * q + gq, pi = sPiv[r0 + i];
                        int const src = r0 + sSrc[i];
                        drow[t][q] = pi >= r0 + kLuBs ? pi : n;
                        int const dsrc = pi >= r0 + kLuBs ? r0 + sSrc[pi - r0] : src;
                        bv[t][q] = double(sa.at(sa.rows(min(src, n - 1)) + cpc));
                        dv[t][q] = double(sa.at(sa.rows(min(dsrc, n - 1)) + cpc));
                    }
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): every row read before any is written
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        int const R = r0 + 16 * t + 4 * q + gq;
                        if (R < n) sa.at(sa.rows(R) + cpc) = T(bv[t][q]);
                        if (drow[t][q] < n) sa.at(sa.rows(drow[t][q]) + cpc) = T(dv[t][q]);
                    }
                continue;
            }
            int const tj = T0 + jt;
            bool const isx = tj == NT;
            Sep<T> const S = isx ? sx : sa;
            int const c0 = isx ? 0 : 16 * tj, cmax = isx ? nrhs : n;
            // every element outside n x n (and every right-hand-side column
            // >= nrhs) reads as 0 here and is not stored: the padding takes no
            // part in the real rows' and columns' results
            int const cg = c0 + m, cpg = S.cols(min(cg, cmax - 1));
            bool const vcg = cg < cmax;
            double kv[2][4];
            int kr[2][4];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    kr[t][q] = r0 + sSrc[16 * t + 4 * q + gq];
                    kv[t][q] = double(S.at(S.rows(min(kr[t][q], n - 1)) + cpg));
                }
            int cpr[4];
            bool vcr[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const c = c0 + gq + 4 * r;
                vcr[r] = c < cmax;
                cpr[r] = S.cols(min(c, cmax - 1));
            }
            // trailing tiles, kLuGroup at a time (every load of a group in
            // flight together; the other wavefronts of the CU, two instances'
            // worth, run meanwhile); the padding outside n x n is replaced by
            // 0 once the group has arrived.  A row
            // below the block holds either its own content or a row of the
            // block (a swap at step c exchanges position c with a later one),
            // so no trailing store clobbers another trailing tile's source; the
            // block's rows are overwritten by U_kj only after every load of
            // the column tile.
            auto load_group = [&](int i0, d4 (&dst)[kLuGroup], int (&rows)[kLuGroup]) __attribute__((always_inline)) {
#pragma unroll
                for (int u = 0; u < kLuGroup; ++u) {
                    int const ti = T0 + i0 + u;
                    int const R = ti < NT ? r0 + sSrc[16 * ti + m - r0] : n;
                    int const rpR = S.rows(min(R, n - 1));
                    rows[u] = R;
#pragma unroll
                    for (int r = 0; r < 4; ++r) dst[u][r] = double(S.at(rpR + cpr[r]));
                }
            };
            d4 cur[kLuGroup];
            int curR[kLuGroup];
            if (mt > 0) load_group(0, cur, curR);
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int q = 0; q < 4; ++q) kv[t][q] = (vcg && kr[t][q] < n) ? pin(kv[t][q]) : 0.0;
            d4 u0 = {0.0, 0.0, 0.0, 0.0}, u1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                u0 = M::mma(op_read(sW[0], q, lane), kv[0][q], u0);
                u1 = M::mma(op_read(sW[1], q, lane), kv[0][q], u1);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) u1 = M::mma(op_read(sW[2], q, lane), kv[1][q], u1);
            for (int i0 = 0; i0 < mt; i0 += kLuGroup) {
                if (i0 > 0) load_group(i0, cur, curR);
#pragma unroll
                for (int u = 0; u < kLuGroup; ++u)
#pragma unroll
                    for (int r = 0; r < 4; ++r) cur[u][r] = (curR[u] < n && vcr[r]) ? pin(cur[u][r]) : 0.0;
#pragma unroll
                for (int u = 0; u < kLuGroup; ++u) {
                    int const ti = T0 + i0 + u;
                    if (ti >= NT) break;
                    double const* const l0 = sP[ti - 2 * kb][0];
                    double const* const l1 = sP[ti - 2 * kb][1];
#pragma unroll
                    for (int q = 0; q < 4; ++q) cur[u] = M::mma(-u0[q], op_read(l0, q, lane), cur[u]);
#pragma unroll
                    for (int q = 0; q < 4; ++q) cur[u] = M::mma(-u1[q], op_read(l1, q, lane), cur[u]);
                    int const R = 16 * ti + m;
                    if (R < n) {
                        int const rpR = S.rows(R);
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (vcr[r]) S.at(rpR + cpr[r]) = T(cur[u][r]);
                    }
                }
            }
            // every load of this column tile has returned before the block's
            // rows are overwritten
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            if (vcg)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const R0 = r0 + gq + 4 * r, R1 = R0 + 16;
                    if (R0 < n) S.at(S.rows(R0) + cpg) = T(u0[r]);
                    if (R1 < n) S.at(S.rows(R1) + cpg) = T(u1[r]);
                }
        }
        __syncthreads();
    }

    __syncthreads();  // the solve's LDS overlays the factor's

    int const bad = sFlag;
    if (tid == 0 && g.info) g.info[b] = bad;
    if (!rhs) return;

    // ---- U x = y (X holds y): the diagonal blocks' inverses U_kk^{-1}
    // (tri_inverse, one wavefront per block), then block back substitution on
    // MFMA.  A zero pivot (info > 0) gives inf / NaN in X.
    auto& sInv = sm.s.sInv;
    auto& sX = sm.s.sX;
    if (tid < npad)
#pragma unroll
        for (int c = 0; c < kLuRhs; ++c) sX[tid * kLuRhs + c] = xget(tid, c);
    for (int k = w; k < NT; k += kLuWaves) {
        // U_kk^T is lower triangular: lane (g, i) forms row i of (U_kk^T)^{-1},
        // i.e. column i of U_kk^{-1}, from column i of U_kk; every read of the
        // tile precedes the wavefront's writes
        double* const t = sInv[k];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            int const c = gq + 4 * e;
            t[c * kTileStride + m] = fget(16 * k + m, 16 * k + c);
        }
        double const myr = 1.0 / t[m * kTileStride + m];
        double wv[4];
        tri_inv16_regs(wv, lane, myr, [&](int j) { return j < m ? -t[m * kTileStride + j] : 0.0; }, NoHook{});
#pragma unroll
        for (int u = 0; u < 4; ++u) t[m * kTileStride + 4 * u + gq] = wv[u];
    }
    __syncthreads();
    // right-hand-side block as MFMA B operand (rows 16 blk + 4 q + l / 16,
    // column l % 16) and in accumulator layout
    int const rc = lane & 15;
    auto rhs_op = [&](int blk, int q) -> double {
        return rc < kLuRhs ? sX[(16 * blk + 4 * q + (lane >> 4)) * kLuRhs + rc] : 0.0;
    };
    auto rhs_load = [&](int blk) {
        d4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = rc < kLuRhs ? sX[(16 * blk + M::row(lane, r)) * kLuRhs + rc] : 0.0;
        return v;
    };
    auto rhs_store = [&](int blk, d4 v) {
        if (rc < kLuRhs)
#pragma unroll
            for (int r = 0; r < 4; ++r) sX[(16 * blk + M::row(lane, r)) * kLuRhs + rc] = v[r];
    };
    // U x = y by 16-row blocks from the bottom; the U tiles of the next
    // block are loaded while this one is solved
    double fv[4][4], fn[4][4];
    auto load_u = [&](int bk, double (&dst)[4][4]) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const i = w + kLuWaves * u;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                dst[u][q] = i < bk ? fget(16 * i + (lane & 15), 16 * bk + 4 * q + (lane >> 4)) : 0.0;
        }
    };
    load_u(NT - 1, fv);
    for (int bk = NT - 1; bk >= 0; --bk) {
        if (bk > 0) load_u(bk - 1, fn);
        if (w == 0) {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = M::mma(op_read(sInv[bk], q, lane), rhs_op(bk, q), acc);
            rhs_store(bk, acc);
        }
        __syncthreads();
        {  // y_i -= U_ib x_b, i < b
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                int const i = w + kLuWaves * u;
                if (i >= bk) continue;
                d4 acc = rhs_load(i);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = M::mma(-fv[u][q], rhs_op(bk, q), acc);
                rhs_store(i, acc);
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) fv[u][q] = fn[u][q];
    }
    Ref<T> x(g.X, b);
    for (int e = tid; e < n * nrhs; e += kLuThreads) {
        int const r = e % n, c = e / n;
        x(r, c) = T(sX[r * kLuRhs + c]);
    }
}

template <typename T>
bool offsets_fit(Mat<T> const& m, int rows, int cols) {
    if (rows <= 0 || cols <= 0) return true;
    constexpr int PS = PanelSize<T>::value, LOG = PanelSize<T>::log2;
    int64_t const i = rows - 1, j = cols - 1, r = m.trans ? j : i, c = m.trans ? i : j;
    int64_t const off = m.panel ? (r >> LOG) * m.spacing + (r & (PS - 1)) + c * PS : r + int64_t(m.spacing) * c;
    return off >= 0 && off * int64_t(sizeof(T)) < (int64_t(1) << 31);
}

}  // namespace

// The kernel addresses an instance with 32-bit byte offsets: true when every
// element of A (and of X's first nrhs columns) is within 2 GiB of its base.
bool lu_wg_fits(Mat<double> A, Mat<double> X, int nrhs) {
    return offsets_fit(A, A.rows, A.cols) && (nrhs == 0 || offsets_fit(X, X.rows, nrhs));
}
bool lu_wg_fits_f32(Mat<float> A) { return offsets_fit(A, A.rows, A.cols); }

// FP64 LU with partial pivoting in place of A, pivots to ipiv, and (nrhs > 0)
// the solve into X; only_if: solve only the instances with only_if[b] < 0.
int launch_wg_gesv(hipStream_t s, Mat<double> A, int64_t* ipiv, int64_t ipiv_stride, Mat<double> B, Mat<double> X,
                   int nrhs, int64_t batch, int32_t* info, const int32_t* only_if) {
    if (batch == 0) return 0;
    LuArgs<double> g;
    g.A = A, g.B = B, g.X = X;
    g.n = A.rows, g.nrhs = nrhs;
    g.info = info, g.only_if = only_if;
    g.ipiv = ipiv, g.ipiv_stride = ipiv_stride;
    hipLaunchKernelGGL(lu_wg_kernel<double>, dim3(unsigned(batch)), dim3(kLuThreads), 0, s, g);
    return check_launch("lu_wg_kernel");
}

// The LU alone on FP32 storage (FP64 arithmetic), pivots to ipiv.
int launch_wg_getrf_f32(hipStream_t s, Mat<float> A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                        int32_t* info) {
    if (batch == 0) return 0;
    LuArgs<float> g;
    g.A = A, g.B = Mat<float>{}, g.X = Mat<float>{};
    g.n = A.rows, g.nrhs = 0;
    g.info = info, g.only_if = nullptr;
    g.ipiv = ipiv, g.ipiv_stride = ipiv_stride;
    hipLaunchKernelGGL(lu_wg_kernel<float>, dim3(unsigned(batch)), dim3(kLuThreads), 0, s, g);
    return check_launch("lu_wg_kernel<float>");
}

}  // namespace blast_hip
