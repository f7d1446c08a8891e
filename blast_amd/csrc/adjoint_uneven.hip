// Uneven batched adjoint of the SPD solve X = A^{-1} B (FP64), from the
// factor L kept by the forward pass, in one persistent launch over instances
// of different n (n <= 192, up to 8 right-hand sides):
//   Bbar = A^{-1} Xbar = L^{-T} L^{-1} Xbar,
//   Abar = -fold(Bbar X^T): -(sum_c bb_ic x_jc + bb_jc x_ic) below the
//          diagonal, -sum_c bb_ic x_ic on it, 0 above,
// the same quantities as the uniform posv backward with recompute = 0
// (adjoint.hip: two trsm and sym_outer_lower_kernel), per instance.
//
// One 256-thread workgroup per instance, instances pulled from a queue
// (atomic counter in queue[0], the order in queue[1..count], largest block
// count first) as in the uneven forward launch.  The right-hand sides live
// in LDS ([row][8]); the triangular solves go by 16-row blocks: the block
// column of L below and on the diagonal (forward) or the block row left of
// and on it (backward) is staged into LDS, wavefront 0 solves the 16x16
// diagonal system (one lane per row and right-hand side, the finished value
// broadcast within the 16-lane row by DPP; the diagonal's reciprocals formed
// first, so the dependent chain multiplies), and every thread then updates
// one row of the remaining right-hand sides from the staged block (loops
// specialised to 1, 2, 4 or 8 right-hand sides).  Block loads are issued two
// steps ahead (two register sets).
// Algorithm references: the potrf/trsm adjoints of BLAST's dense kernels are
// not in BLAST; the formulas are the standard ones (Giles 2008, Murray 2016)
// the uniform entry points and oracle/adjoint.py use.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kAdjT = 256;
constexpr int kAdjNT = 12;               // n <= 192
constexpr int kAdjMaxN = 16 * kAdjNT;
constexpr int kAdjRhs = 8;               // right-hand sides (columns of X)
constexpr int kAdjLs = 17;               // LDS row stride of a staged 16-column block (doubles)
constexpr int kAdjStage = kAdjMaxN * kAdjLs;  // staged block: up to 192 rows x 16 (or 16 x 192)
constexpr int kAdjPer = (kAdjMaxN * 16 + kAdjT - 1) / kAdjT;  // staged elements per thread (12)

struct AdjLds {
    double v[kAdjMaxN * kAdjRhs];  // Xbar -> y -> Bbar, [row][kAdjRhs]
    double x[kAdjMaxN * kAdjRhs];  // X, [row][kAdjRhs]
    double l[kAdjStage];           // the staged block of L
};

__device__ __forceinline__ bool adj_shape_ok(blast_adjoint_instance_t const& d) {
    int64_t const n = d.l.rows, r = d.x.cols;
    return n >= 0 && n <= kAdjMaxN && d.l.cols == n && r >= 1 && r <= kAdjRhs && d.x.rows == n &&
           d.xb.rows == n && d.xb.cols == r && d.bb.rows == n && d.bb.cols == r && d.ab.rows == n && d.ab.cols == n &&
           d.l.data && d.x.data && d.xb.data && d.bb.data && d.ab.data;
}

// Block kb of L staged for the solve: forward (column): element (row r0 + i,
// column r0 + k), i = 0..n - r0 - 1, k < 16, at [i * 17 + k]; backward (row):
// element (row r0 + k, column j), j = 0..r0 + 15, at [j * 17 + k].  Entries
// outside the lower triangle and beyond n read 0 (1 on the padded diagonal).
template <bool FWD>
__device__ __forceinline__ void adj_load_block(Ref<double> const& L, int n, int kb, double (&reg)[kAdjPer]) {
    int const r0 = 16 * kb, span = FWD ? kAdjMaxN - r0 : r0 + 16;
#pragma unroll
    for (int e = 0; e < kAdjPer; ++e) {
        int const idx = threadIdx.x + kAdjT * e, a = idx >> 4, k = idx & 15;  // a: the long index, k: within block
        int const row = FWD ? r0 + a : r0 + k, col = FWD ? r0 + k : a;
        bool const in = a < span && row < n && col < n && col <= row;
        reg[e] = in ? L(row, col) : (a < span && row == col ? 1.0 : 0.0);
    }
}
__device__ __forceinline__ void adj_store_block(double* s, double const (&reg)[kAdjPer]) {
#pragma unroll
    for (int e = 0; e < kAdjPer; ++e) {
        int const idx = threadIdx.x + kAdjT * e;
        if (idx < kAdjStage / kAdjLs * 16) s[(idx >> 4) * kAdjLs + (idx & 15)] = reg[e];
    }
}

// One instance: the two block solves on the right-hand sides in LDS and the
// outputs, for R >= r right-hand-side columns (R = 1, 2, 4, 8: the row
// updates and the diagonal solves cover R columns, not all 8).
template <int R>
__device__ __forceinline__ void adj_instance(AdjLds& sm, Ref<double> const& L, Ref<double> const& X,
                                             Ref<double> const& XB, Ref<double> const& AB, Ref<double> const& BB,
                                             int n, int r) {
    int const tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int const li = lane & 15, lc = lane >> 4;  // wavefront 0's diagonal solve: row li, right-hand side lc (+4)
    int const NT = (n + 15) / 16, npad = 16 * NT;
    // block loads of step st (forward block column st, then backward block
    // row 2NT-1-st); nothing beyond the last step
    auto load_step = [&](int st, double (&reg)[kAdjPer]) __attribute__((always_inline)) {
        if (st < NT)
            adj_load_block<true>(L, n, st, reg);
        else if (st < 2 * NT)
            adj_load_block<false>(L, n, 2 * NT - 1 - st, reg);
    };
    double rA[kAdjPer], rB[kAdjPer];
    load_step(0, rA);
    load_step(1, rB);
    for (int e = tid; e < npad * kAdjRhs; e += kAdjT) {
        int const row = e / kAdjRhs, c = e % kAdjRhs;
        bool const in = row < n && c < r;
        sm.v[e] = in ? XB(row, c) : 0.0;
        sm.x[e] = in ? X(row, c) : 0.0;
    }
    constexpr int H = (R + 3) / 4;  // diagonal-solve passes (4 columns each)
    // ---- the block steps: s = 0..NT-1 the forward solve y = L^{-1} Xbar by
    // 16-row blocks (block column s of L staged), s = NT..2NT-1 the
    // backward solve Bbar = L^{-T} y from the bottom (block row 2NT-1-s).
    // Two register sets alternate (the step pair is unrolled), so each
    // block's loads are issued two steps before it is staged.
    auto step = [&](int st, double (&reg)[kAdjPer]) __attribute__((always_inline)) {
        __syncthreads();  // the previous step's reads of sm.l are done
        adj_store_block(sm.l, reg);
        __syncthreads();
        load_step(st + 2, reg);
        if (st < NT) {
            int const kb = st, r0 = 16 * kb;
            if (w == 0) {
                // 1 / l_jj formed for all 16 rows at once, off the chain
                double const dinv = 1.0 / sm.l[li * kAdjLs + li];
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    int const c = lc + 4 * h;
                    if (c < R) {
                        double y = sm.v[(r0 + li) * kAdjRhs + c];
                        static_for<0, 16>([&](auto J) __attribute__((always_inline)) {
                            constexpr int j = decltype(J)::value;
                            if (li == j) y = y * dinv;
                            double const yj = row_bcast<j>(y);
                            if (li > j) y = __builtin_fma(-sm.l[li * kAdjLs + j], yj, y);
                        });
                        sm.v[(r0 + li) * kAdjRhs + c] = y;
                    }
                }
            }
            __syncthreads();
            // rows below the block: v_i -= L_i,kb y_kb
            for (int i = r0 + 16 + tid; i < npad; i += kAdjT) {
                double const* const lrow = sm.l + (i - r0) * kAdjLs;
                double acc[R];
#pragma unroll
                for (int c = 0; c < R; ++c) acc[c] = sm.v[i * kAdjRhs + c];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    double const lk = lrow[k];
#pragma unroll
                    for (int c = 0; c < R; ++c) acc[c] = __builtin_fma(-lk, sm.v[(r0 + k) * kAdjRhs + c], acc[c]);
                }
#pragma unroll
                for (int c = 0; c < R; ++c) sm.v[i * kAdjRhs + c] = acc[c];
            }
        } else {
            int const kb = 2 * NT - 1 - st, r0 = 16 * kb;
            if (w == 0) {
                // L_kk^T z = v (upper): z_j final from the bottom, rows i < j apply it
                double const dinv = 1.0 / sm.l[(r0 + li) * kAdjLs + li];
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    int const c = lc + 4 * h;
                    if (c < R) {
                        double z = sm.v[(r0 + li) * kAdjRhs + c];
                        static_for<0, 16>([&](auto JJ) __attribute__((always_inline)) {
                            constexpr int j = 15 - decltype(JJ)::value;
                            if (li == j) z = z * dinv;
                            double const zj = row_bcast<j>(z);
                            if (li < j) z = __builtin_fma(-sm.l[(r0 + li) * kAdjLs + j], zj, z);
                        });
                        sm.v[(r0 + li) * kAdjRhs + c] = z;
                    }
                }
            }
            __syncthreads();
            // rows above the block: v_i -= sum_k L(r0 + k, i) z_k
            for (int i = tid; i < r0; i += kAdjT) {
                double const* const lcol = sm.l + i * kAdjLs;
                double acc[R];
#pragma unroll
                for (int c = 0; c < R; ++c) acc[c] = sm.v[i * kAdjRhs + c];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    double const lk = lcol[k];
#pragma unroll
                    for (int c = 0; c < R; ++c) acc[c] = __builtin_fma(-lk, sm.v[(r0 + k) * kAdjRhs + c], acc[c]);
                }
#pragma unroll
                for (int c = 0; c < R; ++c) sm.v[i * kAdjRhs + c] = acc[c];
            }
        }
    };
    for (int st = 0; st < 2 * NT; st += 2) {
        step(st, rA);
        step(st + 1, rB);
    }
    __syncthreads();
    // ---- outputs: Bbar, and Abar = -fold(Bbar X^T) over the whole n x n
    for (int e = tid; e < n * r; e += kAdjT) {
        int const row = e % n, c = e / n;
        BB(row, c) = sm.v[row * kAdjRhs + c];
    }
    // element e = tid + 256 t at (i, j) = (e mod n, e div n), stepped
    // without a division per element
    int const si = kAdjT % (n > 0 ? n : 1), sj = kAdjT / (n > 0 ? n : 1);
    int i = tid % (n > 0 ? n : 1), j = tid / (n > 0 ? n : 1);
    for (int e = tid; e < n * n; e += kAdjT) {
        double v = 0.0;
        if (i > j) {
#pragma unroll
            for (int c = 0; c < R; ++c)
                if (c < r)
                    v = v - (sm.v[i * kAdjRhs + c] * sm.x[j * kAdjRhs + c] + sm.v[j * kAdjRhs + c] * sm.x[i * kAdjRhs + c]);
        } else if (i == j) {
#pragma unroll
            for (int c = 0; c < R; ++c)
                if (c < r) v = v - sm.v[i * kAdjRhs + c] * sm.x[i * kAdjRhs + c];
        }
        AB(i, j) = v;
        i += si;
        j += sj;
        if (i >= n) {
            i -= n;
            ++j;
        }
    }
}

__global__ __launch_bounds__(kAdjT, 3) void posv_backward_uneven_kernel(const blast_adjoint_instance_t* inst,
                                                                        int64_t count, int32_t* queue, int32_t* info) {
    __shared__ AdjLds sm;
    __shared__ int sIdx;
    int const tid = threadIdx.x;
    for (;;) {
        if (tid == 0) sIdx = atomicAdd(&queue[0], 1);
        __syncthreads();
        int const q = sIdx;
        __syncthreads();
        if (q >= count) break;  // every workgroup reaches this exit
        int const idx = queue[1 + q];
        if (idx < 0 || idx >= count) continue;
        blast_adjoint_instance_t const d = inst[idx];
        if (!adj_shape_ok(d)) {
            if (info && tid == 0) info[idx] = -1;
            continue;
        }
        Mat<double> const Lm = Mat<double>::from(d.l), Xm = Mat<double>::from(d.x), XBm = Mat<double>::from(d.xb);
        Mat<double> const ABm = Mat<double>::from(d.ab), BBm = Mat<double>::from(d.bb);
        Ref<double> const L(Lm, 0), X(Xm, 0), XB(XBm, 0), AB(ABm, 0), BB(BBm, 0);
        int const n = int(d.l.rows), r = int(d.x.cols);
        // (uniform: one descriptor per workgroup)
        if (r <= 1)
            adj_instance<1>(sm, L, X, XB, AB, BB, n, r);
        else if (r <= 2)
            adj_instance<2>(sm, L, X, XB, AB, BB, n, r);
        else if (r <= 4)
            adj_instance<4>(sm, L, X, XB, AB, BB, n, r);
        else
            adj_instance<8>(sm, L, X, XB, AB, BB, n, r);
        if (info && tid == 0) info[idx] = 0;
    }
}

__global__ void zero_adj_counter_kernel(int32_t* p) { *p = 0; }

}  // namespace

int launch_posv_backward_uneven(hipStream_t s, const blast_adjoint_instance_t* inst, int64_t count, int32_t* queue,
                                int32_t* info, int grid) {
    if (count == 0) return 0;
    hipLaunchKernelGGL(zero_adj_counter_kernel, dim3(1), dim3(1), 0, s, queue);
    int r = check_launch("zero_adj_counter_kernel");
    if (r) return r;
    int64_t const g = grid > 0 ? (grid < count ? grid : count) : (count < 2048 ? count : 2048);
    hipLaunchKernelGGL(posv_backward_uneven_kernel, dim3(unsigned(g)), dim3(kAdjT), 0, s, inst, count, queue, info);
    return check_launch("posv_backward_uneven_kernel");
}

}  // namespace blast_hip
