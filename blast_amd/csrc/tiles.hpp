// MFMA tile helpers shared by the fused Cholesky kernels (FP64 and FP32).
//
// Tiles are 16x16.  Register tiles use the accumulator layout of
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32; LDS tiles store T[m][k] at
// k*17 + m (the 17 avoids bank conflicts for both the accumulator writes and
// the operand reads).
#pragma once

#include <type_traits>

#include "common.hpp"

namespace blast_hip {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kTileStride = 17;
constexpr int kTileElems = 16 * kTileStride;
constexpr int kMaxRhs = 4;

constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Returns v unchanged, opaque to the optimiser: lane-dependent values
// (addresses, select masks) are recomputed where they are used instead of
// being hoisted and held in registers (or spilled) across a whole kernel.
__device__ __forceinline__ int opaque_v(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
// The same for wavefront-uniform values (tile offsets in scalar registers).
__device__ __forceinline__ int opaque_s(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// Operand layouts (both precisions, lane l): A operand A[l % 16][l / 16],
// B operand B[l / 16][l % 16] for the K-chunk.  Accumulator register r of
// lane l holds D[row][l % 16]:
//   f64 16x16x4: row = l / 16 + 4 r
//   f32 16x16x4: row = 4 (l / 16) + r
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    using acc_t = d4;
    static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    using acc_t = f4;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ int acc_col(int lane) { return lane & 15; }

// First index of max |x_r| with BLAST's strict '>' scan (dense/Iamax.hpp:82-95),
// as a reduction over (value, index) pairs.  The scan starts from element 0
// unconditionally (a NaN there is never replaced) and skips any later NaN
// (NaN > v is false), so the key of element r is |x_r|, except that a NaN
// after element 0 becomes -1, below every real |x|.  Pairs combine as "the
// higher index wins only if strictly greater"; the identity is (-1, INT_MAX).
template <typename T>
__device__ __forceinline__ T iamax_key(T x, int r) {
    T const v = x < T(0) ? -x : x;
    return (r != 0 && v != v) ? T(-1) : v;
}
template <typename T>
__device__ __forceinline__ void argmax_combine(T& v, int& i, T v2, int i2) {
    bool const swap = i2 < i;
    T const va = swap ? v2 : v, vb = swap ? v : v2;
    int const ia = swap ? i2 : i, ib = swap ? i : i2;
    if (vb > va) { v = vb; i = ib; } else { v = va; i = ia; }
}

typedef short s4 __attribute__((ext_vector_type(4)));

// Four consecutive-k operand elements T[l % 16][4 (l / 16) + e], e = 0..3,
// of an FP32 LDS tile rounded to BF16 (round to nearest even), the operand
// layout of v_mfma_f32_16x16x16bf16_1k (A and B alike).
__device__ __forceinline__ s4 op_read_bf16(const float* tile, int lane) {
    s4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        unsigned const u = __builtin_bit_cast(unsigned, tile[(4 * (lane >> 4) + e) * kTileStride + (lane & 15)]);
        v[e] = short((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    return v;
}

// Operand element for K-chunk q: T[l % 16][4q + l / 16] of an LDS tile.
template <typename T>
__device__ __forceinline__ T op_read(const T* tile, int q, int lane) {
    return tile[(4 * q + (lane >> 4)) * kTileStride + (lane & 15)];
}

template <typename T>
__device__ __forceinline__ void acc_to_lds(T* tile, typename Mfma<T>::acc_t v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[acc_col(lane) * kTileStride + Mfma<T>::row(lane, r)] = v[r];
}

template <typename T>
__device__ __forceinline__ typename Mfma<T>::acc_t lds_to_acc(const T* tile, int lane) {
    typename Mfma<T>::acc_t v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = tile[acc_col(lane) * kTileStride + Mfma<T>::row(lane, r)];
    return v;
}

// Element access; PANEL = column-major panel storage, not transposed (the
// layout of BLAST's StaticPanelMatrix), which needs no runtime dispatch.
template <bool PANEL>
__device__ __forceinline__ double& at(Ref<double> const& a, int row, int col) {
    if constexpr (PANEL)
        return a.p[int64_t(row >> 2) * a.m.spacing + (row & 3) + 4 * col];
    else
        return a(row, col);
}

// Tile (ti, tj) of the FP64 matrix A in accumulator layout of precision T.
// Outside n x n the matrix is padded with the identity; the strict upper
// part of diagonal tiles is not read (BLAST reads A's lower triangle only).
template <bool PANEL, typename T>
__device__ __forceinline__ typename Mfma<T>::acc_t load_tile(Ref<double> const& a, int n, int ti, int tj, int lane) {
    typename Mfma<T>::acc_t v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + Mfma<T>::row(lane, r), col = 16 * tj + acc_col(lane);
        v[r] = T((row < n && col < n) ? (row >= col ? at<PANEL>(a, row, col) : 0.0) : (row == col ? 1.0 : 0.0));
    }
    return v;
}

// Panel storage (PS = 4) address of element (16 ti + row(l, r), 16 tj + l%16)
// of an accumulator-layout tile, split into a per-lane part (the same for
// every tile) and a part that is uniform over the wavefront (scalar
// arithmetic):  p + (4 ti + row/4) * spacing + row % 4 + 4 (16 tj + l % 16)
//   FP64 (row = l/16 + 4r):   lane (l/16) + 4 (l%16);          uniform (4 ti + r) spacing + 64 tj
//   FP32 (row = 4 (l/16) + r): lane (l/16) spacing + 4 (l%16);  uniform 4 ti spacing + r + 64 tj
template <typename T>
__device__ __forceinline__ int panel_lane_off(int lane, int spacing) {
    if constexpr (std::is_same_v<T, double>)
        return (lane >> 4) + 4 * (lane & 15);
    else
        return (lane >> 4) * spacing + 4 * (lane & 15);
}
template <typename T>
__device__ __forceinline__ int64_t panel_tile_off(int ti, int tj, int r, int spacing) {
    if constexpr (std::is_same_v<T, double>)
        return int64_t(4 * ti + r) * spacing + 64 * tj;
    else
        return int64_t(4 * ti) * spacing + r + 64 * tj;
}

// Tile (ti, tj) of A when n is a multiple of 16 (every tile inside the
// matrix): no bounds to test, so the loads carry no branches and all of them
// issue before the first use.  Every address is a lower-triangle element
// that load_tile also reads: the strict upper part of a diagonal tile is
// read from its mirror (a lower element) and then zeroed.
// SCALAR_ADDR: panel tiles addressed as scalar base + lane offset (fewer
// vector instructions; the per-element form can need fewer registers).
template <bool PANEL, typename T, bool SCALAR_ADDR = true>
__device__ __forceinline__ typename Mfma<T>::acc_t load_tile_full(Ref<double> const& a, int ti, int tj, int lane) {
    typename Mfma<T>::acc_t v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + Mfma<T>::row(lane, r), col = 16 * tj + acc_col(lane);
        if (PANEL && SCALAR_ADDR && ti != tj) {
            v[r] = T((a.p + panel_tile_off<T>(ti, tj, r, a.m.spacing))[panel_lane_off<T>(lane, a.m.spacing)]);
        } else if (ti == tj) {
            double const x = at<PANEL>(a, row >= col ? row : col, row >= col ? col : row);
            v[r] = T(row >= col ? x : 0.0);
        } else {
            v[r] = T(at<PANEL>(a, row, col));
        }
    }
    return v;
}

// Store FP64 off-diagonal tile (ti, tj), ti > tj, of L in panel storage when
// the tile lies inside the matrix: no tests, scalar tile addressing.
__device__ __forceinline__ void store_tile_panel_full(Ref<double> const& l, int ti, int tj, d4 v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r)
        (l.p + panel_tile_off<double>(ti, tj, r, l.m.spacing))[panel_lane_off<double>(lane, l.m.spacing)] = v[r];
}

// Store FP64 tile (ti, tj) of L.  Dense storage: i >= j only.  Panel storage:
// also zeros in the strict upper part of each 4x4 diagonal block (what
// BLAST's panel potrf leaves there is rounding residue of 0).
template <bool PANEL>
__device__ __forceinline__ void store_tile(Ref<double> const& l, bool panel, int n, int ti, int tj, d4 v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + Mfma<double>::row(lane, r), col = 16 * tj + acc_col(lane);
        if (row < n && col < n && (row >= col || (panel && (row >> 2) == (col >> 2))))
            at<PANEL>(l, row, col) = row >= col ? v[r] : 0.0;
    }
}

// Broadcast lane C of every 16-lane row to the whole row (DPP row_newbcast).
// FP64: one v_mov_b64_dpp (row_newbcast is a legal 64-bit DPP control).
template <int C>
__device__ __forceinline__ double row_bcast(double v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + C, 0xF, 0xF, false);
}
template <int C>
__device__ __forceinline__ float row_bcast(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + C, 0xF, 0xF, false));
}

// 1/sqrt(d): the hardware estimate r refined by one third-order step (about
// 1 ulp).  With eps = (1 - d r^2) / 2, 1/sqrt(d) = r (1 - 2 eps)^(-1/2)
// = r (1 + eps + 1.5 eps^2 + O(eps^3)); four dependent operations after the
// estimate instead of the six of two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double const r = __builtin_amdgcn_rsq(d);
    double const eps = fma_t(-0.5 * d * r, r, 0.5);
    return fma_t(r * eps, fma_t(1.5, eps, 1.0), r);
}
__device__ __forceinline__ float rsqrt_nr(float d) {
    float r = __builtin_amdgcn_rsqf(d);
    float const h = 0.5f * d;
    return __builtin_fmaf(r, __builtin_fmaf(-h * r, r, 0.5f), r);
}

// Sum over the four 16-lane rows of the wavefront (lane i of every row gets
// the sum of lane i of all four rows), without LDS: v_permlane16_swap
// exchanges odd and even rows, v_permlane32_swap the two halves.  Every lane
// adds the same operands in the same order, so all four copies are equal.
// The swap operands and results pass through empty asm statements (not
// volatile, so that they do not bound scheduling regions): given
// one value for both operands, the optimiser folds the two results into one
// register (and then adds x + x).
__device__ __forceinline__ unsigned distinct_copy(unsigned x) {
    asm("" : "+v"(x));
    return x;
}
template <bool ROW16>
__device__ __forceinline__ void lane_swap(unsigned u, unsigned& mine, unsigned& other) {
    auto const p = ROW16 ? __builtin_amdgcn_permlane16_swap(u, distinct_copy(u), false, false)
                         : __builtin_amdgcn_permlane32_swap(u, distinct_copy(u), false, false);
    mine = distinct_copy(p[0]);
    other = distinct_copy(p[1]);
}
template <bool ROW16>
__device__ __forceinline__ float swap_add(float x) {
    unsigned a, b;
    lane_swap<ROW16>(__builtin_bit_cast(unsigned, x), a, b);
    return __builtin_bit_cast(float, a) + __builtin_bit_cast(float, b);
}
template <bool ROW16>
__device__ __forceinline__ double swap_add(double x) {
    unsigned long long const u = __builtin_bit_cast(unsigned long long, x);
    unsigned al, bl, ah, bh;
    lane_swap<ROW16>(unsigned(u), al, bl);
    lane_swap<ROW16>(unsigned(u >> 32), ah, bh);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(ah) << 32) | al) +
           __builtin_bit_cast(double, (static_cast<unsigned long long>(bh) << 32) | bl);
}
template <typename T>
__device__ __forceinline__ T rows_sum(T x) {
    return swap_add<false>(swap_add<true>(x));
}

// Factorisation of a 16x16 diagonal tile and its inverse by one wavefront,
// in two phases that work on registers.  Lane (g, i) = (lane / 16, lane % 16)
// holds tile row i in the columns c = 4u + g (u = 0..3): the four 16-lane rows
// split the columns, and every broadcast of a tile row is a DPP row_newbcast.
// This is the accumulator layout of the transposed tile (register u of lane l
// is T[l % 16][4u + l / 16]), i.e. the MFMA A-operand layout of T itself.
//
// chol16_regs, Cholesky, left-looking: for column k, each row-group forms its
// share of a_ik - <l_i, l_k> over its own columns j < k (l_kj broadcast from
// lane k), the shares are summed across the groups (rows_sum), then
// r = 1/sqrt(d), l_kk = d*r, l_ik = (.)*r, kept by the owner of column k.  The
// pivot arithmetic (reciprocal square root, scaling) is FP64 for FP32 tiles
// too; l_ik is rounded to FP32 afterwards.  On exit t[u] holds L in the
// columns c <= i (the other entries are never used), and every lane of row i
// holds lii = l_ii and rii = 1/sqrt(d_i).  Rows i < k hold values that are
// never used.  hook(integral_constant<int, s>) runs after step s = 0..15.
template <typename T, typename Hook>
__device__ __forceinline__ void chol16_regs(T (&t)[4], int lane, T& lii, double& rii, Hook&& hook) {
    int const i = lane & 15, g = lane >> 4;
    lii = T(0);
    rii = 0.0;
    static_for<0, 16>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value, gk = k & 3, uk = k >> 2;
        T p0 = (g == gk) ? t[uk] : T(0), p1 = T(0);
        static_for<0, 4>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value;
            if constexpr (4 * u < k) {
                T& p = u % 2 == 0 ? p0 : p1;
                T const pr = fma_t(-t[u], row_bcast<k>(t[u]), p);
                // column 4u + g is finished (j < k) in every group when
                // 4u + 3 < k; otherwise only where 4u + g < k, the others
                // holding the diagonal and upper part of rows i and k.  The
                // product is dropped by a select, not multiplied by 0: a NaN
                // in the upper part (a full-tile update mirrors the lower one
                // there) must not reach an earlier pivot.
                p = (4 * u + 3 < k || 4 * u + g < k) ? pr : p;
            }
        });
        T const v = rows_sum(k > 4 ? p0 + p1 : p0);
        T const d = row_bcast<k>(v);  // equals v in lane k
        double const r = rsqrt_nr(double(d));
        T const lkk = T(double(v) * r);
        t[uk] = (g == gk) ? lkk : t[uk];
        lii = (i == k) ? lkk : lii;
        rii = (i == k) ? r : rii;
        hook(K);
    });
}

// chol16_rows: the same factorisation with one tile row per lane: lane
// (g, m) holds row m in all 16 columns (the four 16-lane rows hold copies).
// Right-looking: at step k the pivot a_kk is broadcast from lane k,
// r = 1/sqrt(a_kk), column k is scaled (l_mk = a_mk r, l_kk = a_kk r) and the
// columns c > k are updated, a_mc -= l_mk l_ck with l_ck broadcast from lane
// c.  A step's dependent chain is only the pivot broadcast, the reciprocal
// square root and the update of a_{k+1,k+1}; the other updates are
// independent of it.  On exit row[c] = l_mc for c <= m; the entries above
// the diagonal hold values that are never used (and never reach a pivot).
// hook(integral_constant<int, k>) runs after column k is scaled.
template <typename Hook>
__device__ __forceinline__ void chol16_rows(double (&row)[16], Hook&& hook) {
    static_for<0, 16>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value;
        double const r = rsqrt_nr(row_bcast<k>(row[k]));
        row[k] *= r;
        hook(K);
        static_for<k + 1, 16>([&](auto C) __attribute__((always_inline)) {
            constexpr int c = decltype(C)::value;
            row[c] = fma_t(-row[k], row_bcast<c>(row[k]), row[c]);
        });
    });
}

// chol16_cols: the same factorisation on the tile in the transposed register
// layout t (lane (g, m) holds row m in the columns 4u + g, t[u] = A[m][4u + g],
// the layout the MFMA updates leave it in), each column passing through LDS:
// column c of the 16x16 LDS tile sw (element (m, c) at c*17 + m) receives
// column c of L.  At step k the pivot a_kk is broadcast from lane k of each
// 16-lane row (DPP; the owners of column k, group k % 4, receive a_kk),
// r = 1/sqrt(a_kk), the owners of column k write
// l_mk = a_mk r to LDS, and every lane reads l_mk and the l_ck of its four
// columns c = 4u + g and updates a_mc -= l_mk l_ck: four FMAs per step
// instead of a broadcast and an FMA per trailing column.  The per-element
// operations are chol16_rows' (same FMAs in the same order).  Entries of
// finished columns, and rows above the pivot, receive values that are never
// read again: the factor is the one in LDS (its strict upper triangle holds
// garbage and is overwritten by the caller; rows below the diagonal of each
// column are exact).  hook(integral_constant<int, k>) runs after column k is
// written.
template <typename Hook>
__device__ __forceinline__ void chol16_cols(d4& t, double* sw, int lane, Hook&& hook) {
    int const m = lane & 15, g = lane >> 4;
    static_for<0, 16>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value, kq = k >> 2, kg = k & 3;
        double const r = rsqrt_nr(row_bcast<k>(t[kq]));  // a_kk in the lanes of group k % 4
        double const lk = t[kq] * r;
        if (g == kg) sw[k * kTileStride + m] = lk;
        // the other lanes read what group k % 4 wrote: without this fence the
        // compiler, reasoning per lane, lets a lane that did not store load
        // before the store (LDS itself keeps a wavefront's accesses in order)
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        hook(K);
        double const lmk = sw[k * kTileStride + m];
        static_for<0, 4>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value;
            if constexpr (4 * u + 3 > k) t[u] = fma_t(-lmk, sw[k * kTileStride + 4 * u + g], t[u]);
        });
    });
}

// chol16_cols_inv: chol16_cols fused with the inverse W = L^{-1}, whose step
// k needs only column k of L: the two dependent chains interleave.  The
// owners of column k also write r_k = 1/sqrt(a_kk) (= 1/l_kk to an ulp or
// two) to sr[k].  Inverse (as tri_inv16_regs): lane (g, i) keeps z_i[c] for
// c = 4u + g, z_i[c] = -delta_ic at first; at step k row k is final and rows
// i > k apply it, z_i += (-l_ik r_k) z_k with z_k broadcast from lane k of the
// lane's 16-lane row.  On exit W[i][c] = -z_i[c] r_i (the caller scales).
template <typename Hook>
__device__ __forceinline__ void chol16_cols_inv(d4& t, double* sw, double* sr, int lane, double (&z)[4],
                                                Hook&& hook) {
    int const m = lane & 15, g = lane >> 4;
#pragma unroll
    for (int u = 0; u < 4; ++u) z[u] = (4 * u + g == m) ? -1.0 : 0.0;
    static_for<0, 16>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value, kq = k >> 2, kg = k & 3;
        double const r = rsqrt_nr(row_bcast<k>(t[kq]));  // a_kk in the lanes of group k % 4
        double const lk = t[kq] * r;
        if (g == kg) {
            sw[k * kTileStride + m] = lk;
            sr[k] = r;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // as in chol16_cols
        hook(K);
        double const lmk = sw[k * kTileStride + m];
        // rows below the pivot only: a sign mask, not a per-step compare (16
        // of those kept in SGPR pairs spill)
        unsigned long long const below = static_cast<unsigned long long>(static_cast<long long>((k - m) >> 31));
        double const f = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, -lmk * sr[k]) & below);
        static_for<0, 4>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value;
            if constexpr (4 * u <= k) z[u] = fma_t(f, row_bcast<k>(z[u]), z[u]);
            if constexpr (4 * u + 3 > k) t[u] = fma_t(-lmk, sw[k * kTileStride + 4 * u + g], t[u]);
        });
    });
}

// tri_inv16_regs: W = L^{-1} of the 16x16 lower-triangular tile whose
// multipliers are read by mult(j) = -l_ij (row i = lane % 16; 0 for j >= i),
// with myr = 1/l_ii of the lane's row.  Lane (g, i) computes W[i][c] for its
// columns c = 4u + g, kept as z_i[c] = acc_i[c] - delta_ic, where acc_i[c] is
// the sum of l_im W[m][c] over the rows m < i applied so far; once all are
// in, W[i][c] = -z_i[c] r_i.  At step j row j is final and rows i > j apply
// it: z_i += l_ij W[j] = (-l_ij r_j) z_j, one FMA per column with z_j
// broadcast from lane j.  z_i[c] = 0 for c > i throughout.  On exit w[u] =
// W[i][4u + g] (0 above the diagonal).  hook(integral_constant<int, 16 + j>)
// runs after step j = 0..15.
template <typename T, typename Mult, typename Hook>
__device__ __forceinline__ void tri_inv16_regs(T (&w)[4], int lane, T myr, Mult&& mult, Hook&& hook) {
    int const i = lane & 15, g = lane >> 4;
    T z[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) z[u] = (4 * u + g == i) ? T(-1) : T(0);
    static_for<0, 16>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        T const f = mult(j) * row_bcast<j>(myr);  // -l_ij r_j, 0 for i <= j
        static_for<0, 4>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value;
            if constexpr (4 * u <= j) z[u] = fma_t(f, row_bcast<j>(z[u]), z[u]);
        });
        hook(std::integral_constant<int, 16 + j>{});
    });
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = (4 * u + g <= i) ? -z[u] * myr : T(0);
}

// 1-based tile column of the first pivot that is not positive (d <= 0 or NaN
// makes l_ii NaN), 0 if none; the same value in every lane.
template <typename T>
__device__ __forceinline__ int first_bad_pivot(T lii, int lane) {
    unsigned long long const badl = __ballot(lane < 16 && !(double(lii) > 0.0));
    return badl ? __builtin_ctzll(badl) + 1 : 0;
}

// The two phases on a tile in LDS: sD holds the tile (T[m][k] at k*17 + m);
// on exit sD holds L_kk (zeros above the diagonal) and sW holds W, same
// layout; with sW == sD the tile holds W.  The multipliers -l_ij are parked in
// sW between the phases.  Returns first_bad_pivot.  Branch-free: every lane
// runs the same instructions; all 64 lanes must be active.  hook(s) runs
// after step s = 0..31 (s < 16: column s of L_kk; s >= 16: row s - 16 of W),
// so that a caller can interleave independent work (MFMA updates) with the
// chain; the body alone needs about 30 VGPRs and 24 SGPRs inlined.
struct NoHook {
    template <typename S>
    __device__ __forceinline__ void operator()(S) const {}
};

template <typename T, typename Hook>
__device__ __forceinline__ int diag_factor_hooked(T* sD_generic, T* sW_generic, int lane, Hook&& hook) {
    // LDS-typed pointers: one generic-to-LDS conversion here rather than one
    // per access when the caller's LDS is not visible
    auto* const sD = (__attribute__((address_space(3))) T*)sD_generic;
    auto* const sW = (__attribute__((address_space(3))) T*)sW_generic;
    int const i = lane & 15, g = lane >> 4;
    T t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = sD[(4 * u + g) * kTileStride + i];
    T lii_unused;
    double rii_unused;
    chol16_regs(t, lane, lii_unused, rii_unused, hook);
    // (sW may be sD: L_kk is written and its diagonal read back before the
    // multipliers overwrite it)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int const c = 4 * u + g;
        sD[c * kTileStride + i] = (c <= i) ? t[u] : T(0);
    }
    T const lii = sD[i * kTileStride + i];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int const c = 4 * u + g;
        sW[c * kTileStride + i] = (c < i) ? -t[u] : T(0);
    }
    T w[4];
    tri_inv16_regs(w, lane, T(1.0 / double(lii)), [&](int j) { return sW[j * kTileStride + i]; }, hook);
#pragma unroll
    for (int u = 0; u < 4; ++u) sW[(4 * u + g) * kTileStride + i] = w[u];
    return first_bad_pivot(lii, lane);
}

template <typename T>
__device__ __forceinline__ int diag_factor_inl(T* sD, T* sW, int lane) {
    return diag_factor_hooked<T>(sD, sW, lane, NoHook{});
}

// Called form, for kernels whose register budget cannot hold the inlined body.
template <typename T>
__device__ __attribute__((noinline)) int diag_factor(T* sD, T* sW, int lane) {
    return diag_factor_inl<T>(sD, sW, lane);
}

// Inverse of the 16x16 triangular tile in sA (T[m][k] at k*17 + m; lower,
// or upper when `upper`; unit diagonal when `unit`: the stored diagonal is
// not read) into sW, same layout, same triangle.  One wavefront; may run in
// place (sW == sA).  Row-by-row substitution as in diag_factor: lane i holds
// row i of the lower-triangular form, rows finalised at step j are
// broadcast by DPP from lane j.
template <typename T>
__device__ __forceinline__ void tri_inverse_inl(const T* sA, T* sW, int lane, bool upper, bool unit) {
    int const i = lane & 15;
    T t[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) t[c] = upper ? sA[i * kTileStride + c] : sA[c * kTileStride + i];
    T di = T(1);
#pragma unroll
    for (int c = 0; c < 16; ++c) di = (c == i) ? t[c] : di;
    T const myr = unit ? T(1) : T(1) / di;
    T w[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) w[c] = T(0);
    static_for<0, 16>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        static_for<0, j>([&](auto C) __attribute__((always_inline)) {
            constexpr int c = decltype(C)::value;
            w[c] = (i == j) ? -w[c] * myr : w[c];
        });
        w[j] = (i == j) ? myr : w[j];
        T const f = (i > j) ? t[j] : T(0);
        static_for<0, j + 1>([&](auto C) __attribute__((always_inline)) {
            constexpr int c = decltype(C)::value;
            w[c] = fma_t(f, row_bcast<j>(w[c]), w[c]);
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            T const v = (c <= i) ? w[c] : T(0);  // W_lower[i][c]
            if (upper)
                sW[i * kTileStride + c] = v;  // W = W_lower^T
            else
                sW[c * kTileStride + i] = v;
        }
    }
}
template <typename T>
__device__ void tri_inverse(const T* sA, T* sW, int lane, bool upper, bool unit) {
    tri_inverse_inl(sA, sW, lane, upper, unit);
}

// Wavefront maximum of an FP64 value without NaNs, in every lane: four DPP
// steps within each 16-lane row, then the two row swaps.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    unsigned long long const u = __builtin_bit_cast(unsigned long long, v);
    unsigned const lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u)), CTRL, 0xF, 0xF, false));
    unsigned const hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <bool ROW16>
__device__ __forceinline__ double swap_max(double v) {
    unsigned long long const u = __builtin_bit_cast(unsigned long long, v);
    unsigned l0, l1, h0, h1;
    lane_swap<ROW16>(unsigned(u), l0, l1);
    lane_swap<ROW16>(unsigned(u >> 32), h0, h1);
    return __builtin_fmax(__builtin_bit_cast(double, (static_cast<unsigned long long>(h0) << 32) | l0),
                          __builtin_bit_cast(double, (static_cast<unsigned long long>(h1) << 32) | l1));
}
__device__ __forceinline__ double wave_max(double v) {
    v = __builtin_fmax(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
    v = __builtin_fmax(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
    v = __builtin_fmax(v, dpp_f64<0x141>(v));  // row_half_mirror
    v = __builtin_fmax(v, dpp_f64<0x140>(v));  // row_mirror
    v = swap_max<true>(v);
    return swap_max<false>(v);
}


}  // namespace blast_hip
