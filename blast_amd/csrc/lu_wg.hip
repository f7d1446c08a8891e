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
constexpr int kLuGroup = 8;   // trailing tiles per wavefront in flight

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

// LDS of one instance: the factor phase's tiles and the solve phase's
// overlay the same storage.  Tiles are 16x16 with element (row m, col c) at
// c*17 + m unless noted.
struct LuFactorLds {
    double sP[kLuMaxNT][2][kTileElems];      // panel: tile (row tile t, column tile s) of rows r0 + 16 t..
    double sR[kLuMaxNT - 2][2][kTileElems];  // block row, column tile t, row tile s: A_kj (element (k, m)
                                             // at k*17 + m), then U_kj^T (accumulator layout)
    double sW[2][2][kTileElems];             // L_kk^{-1} (unit lower, 32 x 32)
    double sRow[2][kLuWaves][kLuBs];         // each wavefront's candidate pivot row
    double sCandV[2][kLuWaves];
    int sCandI[2][kLuWaves];
};
struct LuSolveLds {
    double sT[3][kLuMaxNT][kTileElems];    // unit-lower inverses, upper inverses, diagonal blocks
    double sX[kLuMaxNT * 16 * kLuRhs];     // right-hand sides, [row][rhs]
};

template <typename T>
__global__ __launch_bounds__(kLuThreads) void lu_wg_kernel(LuArgs<T> g) {
    using M = Mfma<double>;
    __shared__ union {
        LuFactorLds f;
        LuSolveLds s;
    } sm;
    __shared__ int sPiv[kLuMaxNT * 16];  // pivot row of every row (absolute)
    __shared__ int sSrc[kLuMaxNT * 16];  // panel position -> its row before the block's swaps
    __shared__ int sFlag;
    auto& sP = sm.f.sP;
    auto& sR = sm.f.sR;
    auto& sW = sm.f.sW;
    auto& sRow = sm.f.sRow;
    auto& sCandV = sm.f.sCandV;
    auto& sCandI = sm.f.sCandI;

    int64_t const b = blockIdx.x;
    if (g.only_if && g.only_if[b] >= 0) return;
    int const tid = threadIdx.x, lane = tid & 63;
    int const w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const n = g.n, nrhs = g.nrhs, NB = (n + kLuBs - 1) / kLuBs, npad = kLuBs * NB, NT = 2 * NB;
    Ref<T> f(g.A, b);
    // A outside n x n reads as the identity (the factorisation of blkdiag(A, I))
    auto fget = [&](int i, int j) -> double { return (i < n && j < n) ? double(f(i, j)) : (i == j ? 1.0 : 0.0); };
    auto fset = [&](int i, int j, double v) {
        if (i < n && j < n) f(i, j) = T(v);
    };
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
#pragma unroll
        for (int c = 0; c < kLuBs; ++c) v[c] = live ? fget(r0 + tid, r0 + c) : 0.0;
#pragma unroll
        for (int c = 0; c < kLuBs; ++c) {
            int const buf = c & 1;
            bool const cand = live && pos >= c;
            double kv = cand ? iamax_key(v[c], pos - c) : -1.0;
            int ki = cand ? pos : 0x7fffffff;
            wave_argmax(kv, ki);
            if (pos == ki)
#pragma unroll
                for (int cc = c; cc < kLuBs; ++cc) sRow[buf][w][cc] = v[cc];
            if (lane == 0) sCandV[buf][w] = kv, sCandI[buf][w] = ki;
            __syncthreads();
            double pv = sCandV[buf][0];
            int pp = sCandI[buf][0], wp = 0;
#pragma unroll
            for (int q = 1; q < kLuWaves; ++q) {
                int const before = pp;
                argmax_combine(pv, pp, sCandV[buf][q], sCandI[buf][q]);
                wp = pp != before ? q : wp;
            }
            // position c always holds a candidate, so pp is a real position
            pos = pos == pp ? c : (pos == c ? pp : pos);
            if (tid == 0) sPiv[r0 + c] = r0 + pp;
            double const piv = sRow[buf][wp][c];
            if (piv == 0.0) {
                if (tid == 0 && sFlag == 0 && r0 + c < n) sFlag = r0 + c + 1;
            } else if (live && pos > c) {
                double const l = v[c] / piv;
                v[c] = l;
#pragma unroll
                for (int cc = c + 1; cc < kLuBs; ++cc) v[cc] = __builtin_fma(-l, sRow[buf][wp][cc], v[cc]);
            }
        }
        if (live) {
            sSrc[pos] = tid;
#pragma unroll
            for (int cc = 0; cc < kLuBs; ++cc) {
                fset(r0 + pos, r0 + cc, v[cc]);
                sP[pos >> 4][cc >> 4][(cc & 15) * kTileStride + (pos & 15)] = v[cc];
            }
        }
        __syncthreads();
        if (tid < kLuBs && r0 + tid < n) g.ipiv[b * g.ipiv_stride + r0 + tid] = sPiv[r0 + tid];

        if (tid < kLuBs) {
            // column tid of L_kk^{-1} (32 x 32 unit lower), forward substitution
            double x[kLuBs];
#pragma unroll
            for (int i = 0; i < kLuBs; ++i) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < i; ++k) s = __builtin_fma(-sP[i >> 4][k >> 4][(k & 15) * kTileStride + (i & 15)], x[k], s);
                x[i] = (i == tid) ? 1.0 : (i < tid ? 0.0 : s);
            }
#pragma unroll
            for (int i = 0; i < kLuBs; ++i) sW[i >> 4][tid >> 4][(tid & 15) * kTileStride + (i & 15)] = x[i];
        }
        {
            // the block's swaps on the columns right of the panel, 32 lanes
            // (one row each) per column: block row i becomes old row
            // r0 + sSrc[i] (A_kj, to LDS) and a displaced row p_i >= r0 + 32
            // becomes old row r0 + sSrc[p_i - r0] (back to A).  In each
            // column every load precedes every store.
            int const i = lane & 31, grp = tid >> 5;
            int const pi = sPiv[r0 + i];
            bool const disp = pi >= r0 + kLuBs;
            int const srcb = r0 + sSrc[i], srcd = disp ? r0 + sSrc[pi - r0] : 0;
            int const ncol = npad - r0 - kLuBs;
            constexpr int kU = 4, kGroups = kLuThreads / kLuBs;
            for (int j0 = grp; j0 < ncol; j0 += kGroups * kU) {
                double xb[kU], xd[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    int const j = r0 + kLuBs + j0 + kGroups * u;
                    bool const in = j0 + kGroups * u < ncol;
                    xb[u] = in ? fget(srcb, j) : 0.0;
                    xd[u] = in && disp ? fget(srcd, j) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    int const jr = j0 + kGroups * u;  // column relative to the trailing block
                    if (jr >= ncol) continue;
                    sR[jr >> 4][i >> 4][(i & 15) * kTileStride + (jr & 15)] = xb[u];
                    if (disp) fset(pi, r0 + kLuBs + jr, xd[u]);
                }
            }
        }
        __syncthreads();

        // ---- U_kj^T = A_kj^T L_kk^{-T}: to A (rows r0..) and to LDS as an
        // A operand.  U^T's first 16 columns take the block row's first row
        // tile only (L_kk^{-1} is lower triangular).
        for (int jt = w; jt < mt; jt += kLuWaves) {
            d4 u0 = {0.0, 0.0, 0.0, 0.0}, u1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double const a0 = op_read(sR[jt][0], q, lane);
                u0 = M::mma(a0, op_read(sW[0][0], q, lane), u0);
                u1 = M::mma(a0, op_read(sW[1][0], q, lane), u1);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) u1 = M::mma(op_read(sR[jt][1], q, lane), op_read(sW[1][1], q, lane), u1);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                fset(r0 + (lane & 15), 16 * (T0 + jt) + M::row(lane, r), u0[r]);
                fset(r0 + 16 + (lane & 15), 16 * (T0 + jt) + M::row(lane, r), u1[r]);
            }
            acc_to_lds(sR[jt][0], u0, lane);
            acc_to_lds(sR[jt][1], u1, lane);
        }
        __syncthreads();

        // ---- trailing A_ij^T -= U_kj^T L_ik^T, i, j >= T0 (16-tiles), K = 32;
        // wavefront w takes tiles w, w + 4, ..., kLuGroup at a time, the next
        // group's loads issued before the current group's MFMAs
        int const cnt = mt * mt;
        auto tile_of = [&](int t, int& ti, int& tj) {
            ti = t < cnt ? T0 + t % mt : -1;
            tj = T0 + (t < cnt ? t / mt : 0);
        };
        auto load_group = [&](int t0, d4 (&acc)[kLuGroup]) {
#pragma unroll
            for (int u = 0; u < kLuGroup; ++u) {
                int ti, tj;
                tile_of(t0 + kLuWaves * u, ti, tj);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[u][r] = ti >= 0 ? fget(16 * ti + (lane & 15), 16 * tj + M::row(lane, r)) : 0.0;
            }
        };
        constexpr int kStep = kLuWaves * kLuGroup;
        d4 cur[kLuGroup], nxt[kLuGroup];
        if (w < cnt) load_group(w, cur);
        for (int t0 = w; t0 < cnt; t0 += kStep) {
            if (t0 + kStep < cnt) load_group(t0 + kStep, nxt);
#pragma unroll
            for (int u = 0; u < kLuGroup; ++u) {
                int ti, tj;
                tile_of(t0 + kLuWaves * u, ti, tj);
                if (ti < 0) continue;
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) {
                    double const* const ut = sR[tj - T0][s2];
                    double const* const lt = sP[ti - 2 * kb][s2];
#pragma unroll
                    for (int q = 0; q < 4; ++q) cur[u] = M::mma(-op_read(ut, q, lane), op_read(lt, q, lane), cur[u]);
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) fset(16 * ti + (lane & 15), 16 * tj + M::row(lane, r), cur[u][r]);
            }
            if (t0 + kStep < cnt)
#pragma unroll
                for (int u = 0; u < kLuGroup; ++u) cur[u] = nxt[u];
        }
        __syncthreads();
    }

    // ---- the columns left of each panel get the later blocks' swaps: thread
    // r traces its row back through them (last swap first) to the row whose
    // content ends up in row r; sig accumulates from the last block down.
    // Blocks are handled kChunk at a time: their gathers, one barrier, their
    // stores (different blocks touch different columns).
    int sig = tid;
    constexpr int kChunk = 2;
    for (int kb0 = NB - 2; kb0 >= 0; kb0 -= kChunk) {
        double t[kChunk][kLuBs];
        bool mv[kChunk];
#pragma unroll
        for (int u = 0; u < kChunk; ++u) {
            int const kb = kb0 - u;
            if (kb >= 0) {
                int const s0 = kLuBs * (kb + 1);
#pragma unroll
                for (int s = kLuBs - 1; s >= 0; --s) {
                    int const rr = s0 + s, p = sPiv[rr];
                    sig = sig == rr ? p : (sig == p ? rr : sig);
                }
            }
            mv[u] = kb >= 0 && tid < npad && sig != tid;
#pragma unroll
            for (int c = 0; c < kLuBs; ++c) t[u][c] = mv[u] ? fget(sig, kLuBs * kb + c) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kChunk; ++u)
            if (mv[u])
#pragma unroll
                for (int c = 0; c < kLuBs; ++c) fset(tid, kLuBs * (kb0 - u) + c, t[u][c]);
    }
#pragma unroll
    for (int s = kLuBs - 1; s >= 0; --s) {  // block 0's swaps: sig = the source row of row tid for B
        int const p = sPiv[s];
        sig = sig == s ? p : (sig == p ? s : sig);
    }
    __syncthreads();  // the solve's LDS overlays the factor's

    int const bad = sFlag;
    if (tid == 0 && g.info) g.info[b] = bad;
    if (bad || nrhs == 0) return;

    // ---- solve
    auto& sT = sm.s.sT;
    auto& sX = sm.s.sX;
    Ref<T> bb(g.B, b);
    if (tid < npad)
#pragma unroll
        for (int c = 0; c < kLuRhs; ++c) sX[tid * kLuRhs + c] = (tid < n && c < nrhs) ? double(bb(sig, c)) : 0.0;
    for (int e = tid; e < NT * 256; e += kLuThreads) {  // the diagonal blocks
        int const k = e >> 8, c = (e >> 4) & 15, m = e & 15;
        sT[2][k][c * kTileStride + m] = fget(16 * k + m, 16 * k + c);
    }
    __syncthreads();
    for (int task = tid; task < 2 * NT * 16; task += kLuThreads) {
        int const k = (task >> 4) % NT, cc = task & 15;
        double const* const d = sT[2][k];
        double x[16];
        if (task < NT * 16) {  // column cc of the unit-lower inverse
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                double s = 0.0;
#pragma unroll
                for (int kk = 0; kk < i; ++kk) s = __builtin_fma(-d[kk * kTileStride + i], x[kk], s);
                x[i] = (i == cc) ? 1.0 : (i < cc ? 0.0 : s);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) sT[0][k][cc * kTileStride + i] = x[i];
        } else {  // column cc of the upper inverse
#pragma unroll
            for (int i = 15; i >= 0; --i) {
                double s = 0.0;
#pragma unroll
                for (int kk = i + 1; kk < 16; ++kk) s = __builtin_fma(-d[kk * kTileStride + i], x[kk], s);
                x[i] = (i > cc) ? 0.0 : ((i == cc ? 1.0 : s) / d[i * kTileStride + i]);
            }
#pragma unroll
            for (int i = 0; i < 16; ++i) sT[1][k][cc * kTileStride + i] = x[i];
        }
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
    for (int bk = 0; bk < NT; ++bk) {  // L y = P b
        // the update's L tiles first: their loads are in flight during the
        // diagonal step and the barrier
        double fv[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const i = bk + 1 + w + kLuWaves * u;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                fv[u][q] = i < NT ? fget(16 * i + (lane & 15), 16 * bk + 4 * q + (lane >> 4)) : 0.0;
        }
        if (w == 0) {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = M::mma(op_read(sT[0][bk], q, lane), rhs_op(bk, q), acc);
            rhs_store(bk, acc);
        }
        __syncthreads();
        {  // y_i -= L_ib y_b, up to four tiles per wavefront
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                int const i = bk + 1 + w + kLuWaves * u;
                if (i >= NT) continue;
                d4 acc = rhs_load(i);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = M::mma(-fv[u][q], rhs_op(bk, q), acc);
                rhs_store(i, acc);
            }
        }
        __syncthreads();
    }
    for (int bk = NT - 1; bk >= 0; --bk) {  // U x = y
        double fv[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const i = w + kLuWaves * u;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                fv[u][q] = i < bk ? fget(16 * i + (lane & 15), 16 * bk + 4 * q + (lane >> 4)) : 0.0;
        }
        if (w == 0) {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = M::mma(op_read(sT[1][bk], q, lane), rhs_op(bk, q), acc);
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
    }
    Ref<T> x(g.X, b);
    for (int e = tid; e < n * nrhs; e += kLuThreads) {
        int const r = e % n, c = e / n;
        x(r, c) = T(sX[r * kLuRhs + c]);
    }
}

}  // namespace

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
