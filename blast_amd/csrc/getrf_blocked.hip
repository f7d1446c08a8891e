// Panel step of the blocked LU with partial pivoting for 256 < n <= 1024
// (column-major or panel storage, not transposed).  dispatch.cpp drives it:
// per block of W columns (24 FP64, 32 FP32) starting at k0, this kernel, then
// A22 -= L21 U12 on the MFMA gemm tile kernel.  BLAST's getrf is the same
// right-looking scheme with a narrower block (dense/Getrf.hpp:52-119: getf2
// on the panel, laswp left and right, U12 = L11^{-1} A12, gemm update).
//
// One 512-thread workgroup per instance.  Thread t holds rows k0 + t and
// k0 + 512 + t of the panel in registers, rotated one place per column so
// that the current column is always register position 0 (a runtime column
// loop: a fully unrolled one runs out of instruction cache).  Per column c:
//   * pivot search: the first row of largest |a| (dense/Iamax.hpp:82-95):
//     a DPP wavefront max of the key (|a|; NaN on the diagonal -> +inf, NaN
//     elsewhere -> -1, never chosen), the lowest lane holding it by ballot,
//     then the eight wavefront results through LDS (one barrier);
//   * swap: the two owners put their rows into LDS and take each other's
//     (second barrier); every thread reads the pivot row from there;
//   * an exact zero pivot ends the factorisation (dense/Getf2.hpp:68-69):
//     info = k0 + c + 1, later panels of the instance do nothing;
//   * rows below divide by the pivot and take the rank-1 update
//     v = fma(-l, u, v) (dense/Getf2.hpp:71-72, Ger.hpp).
// Two LDS buffers alternate between columns, so two barriers per column
// suffice.  Thread 0 replays the interchanges on a row map (rowat[r]: the
// row whose content ends in row k0 + r) as they happen.  After the panel,
// every other column j takes its interchanges as one gather and scatter
// through that map (laswp, dense/Laswp.hpp:37-43), and the columns right of
// the panel are solved with the unit lower L11 in registers before the
// store (U12, increasing order as BLAST's trsm).  ipiv is 0-based and
// absolute, LAPACK's sequential-interchange meaning.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kPt = 512;              // threads per workgroup
constexpr int kPr = 2;                // panel rows per thread
constexpr int kPnb = 32;              // widest panel (FP32); FP64 uses 24 columns
// panel width: 24 FP64 columns keep the register panel (96 VGPRs) and the
// column loop's working set within 256 VGPRs; 32 FP32 columns
template <typename T>
constexpr int kPw = sizeof(T) == 8 ? 24 : 32;
constexpr int kPmax = kPt * kPr;      // rows of the panel: n - k0 <= 1024
static_assert(kPmax == 1024 && kPt == 512, "row r of the panel is thread r % 512, register set r / 512");

template <typename T>
struct PanelLds {
    double key[2][kPt / 64];  // per wavefront: best key and its row
    int row[2][kPt / 64];
    T prow[2][kPnb];                      // the pivot row
    T trow[2][kPnb];                      // row k0 + c before the swap
    T l11[kPnb][kPnb + 1];                // unit lower L11 (strict part)
    int rowat[kPmax];
    int piv[kPnb];
};

// Pivot-search key of a candidate (as a double, exact for both storage
// types): |x|; NaN on the diagonal -> +inf (kept, as BLAST's first-max
// search keeps it: nothing compares greater), NaN elsewhere -> -1 (never
// chosen); rows that are not candidates carry -2.
template <typename T>
__device__ __forceinline__ double pivot_key(T x, bool diag) {
    double const m = __builtin_fabs(double(x));
    return m == m ? m : (diag ? __builtin_huge_val() : -1.0);
}

// Element offsets of a non-transposed matrix as a row part plus a column
// part (32-bit within one instance): a uniform part goes into the scalar
// base pointer and each address costs one VGPR.
template <typename T>
struct Split {
    int spacing, panel;
    __device__ __forceinline__ int r(int i) const {
        constexpr int PS = PanelSize<T>::value, LOG = PanelSize<T>::log2;
        return panel ? (i >> LOG) * spacing + (i & (PS - 1)) : i;
    }
    __device__ __forceinline__ int c(int j) const { return panel ? j * PanelSize<T>::value : spacing * j; }
};
template <typename T>
__device__ __forceinline__ T& at(T* base, int off) {
    return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + unsigned(off) * unsigned(sizeof(T)));
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
// c ? x : y on values the compiler cannot trace back to their loads (a
// select between two elements of the register panel would otherwise become a
// load through a selected address, and the panel would go to scratch memory)
template <typename T>
__device__ __forceinline__ T pick(bool c, T x, T y) {
    asm volatile("" : "+v"(x), "+v"(y));
    return c ? x : y;
}

template <typename T>
__global__ __launch_bounds__(kPt) void getrf_panel_kernel(Mat<T> Am, int k0, int nb, int64_t* ipiv,
                                                          int64_t ipiv_stride, int32_t* info) {
    constexpr int W = kPw<T>;
    __shared__ PanelLds<T> sm;
    int64_t const b = blockIdx.x;
    if (info && info[b] != 0) return;  // an earlier panel met a zero pivot
    T* const base = Am.inst(b);
    Split<T> const ax{Am.spacing, Am.panel};
    int const N = Am.cols, rows = Am.rows - k0;
    int const tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // this thread's panel rows: offsets of (k0 + tid + 512 i, k0), stepped
    // one column at a time (opaque steps: 32 column offsets precomputed into
    // scalar registers would spill)
    int const cstep = ax.c(1);
    bool okr[kPr];  // the row exists
    int roff[kPr];
#pragma unroll
    for (int i = 0; i < kPr; ++i) {
        okr[i] = tid + kPt * i < rows;
        roff[i] = ax.r(okr[i] ? k0 + tid + kPt * i : k0) + ax.c(k0);
    }

    // Rows beyond the matrix read row k0 and columns beyond the panel column
    // k0 + nb - 1: such padding never takes part in a pivot choice or a
    // store, and unconditional loads keep no per-element exec masks live.
    T v[kPr][W];
    {
        int o[kPr] = {roff[0], roff[1]};
#pragma unroll
        for (int c = 0; c < W; ++c) {
            int const step = c + 1 < nb ? cstep : 0;
#pragma unroll
            for (int i = 0; i < kPr; ++i) {
                v[i][c] = at(base, o[i]);
                o[i] = opaque_v(o[i] + step);
            }
        }
    }
    for (int r = tid; r < rows; r += kPt) sm.rowat[r] = k0 + r;

    // Runtime column loop (a fully unrolled one is ~20K instructions, run
    // once per panel: instruction-fetch bound).  The register row is rotated
    // one place per column, so the current column is always position 0:
    // at step c, position q holds column (q + c) mod W.
    int c = 0, fail = 0;  // fail: 1 + the column of a zero pivot (uniform)
    for (; c < nb; ++c) {
        int const P = c & 1;
        // ---- pivot: largest key, smallest row among equals
        double best = -2.0;
        int brow = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < kPr; ++i) {
            int const r = tid + kPt * i;
            double const k = (r >= c && r < rows) ? pivot_key(v[i][0], r == c) : -2.0;
            if (k > best) {  // rows ascend with i: ties keep the smaller
                best = k;
                brow = r;
            }
        }
        double const wk = wave_max(best);
        // the smallest row holding wk: a register-set-0 row (rows < 512) of
        // the lowest such lane, else the lowest lane's register-set-1 row
        unsigned long long const b0 = __ballot(best == wk && brow < kPt), b1 = __ballot(best == wk);
        int const wr = b0 ? 64 * w + __builtin_ctzll(b0) : kPt + 64 * w + __builtin_ctzll(b1);
        if (lane == 0) {
            sm.key[P][w] = wk;
            sm.row[P][w] = wr;
        }
        __syncthreads();
        double gk = sm.key[P][0];
        int pv = sm.row[P][0];
#pragma unroll
        for (int q = 1; q < kPt / 64; ++q) {
            double const k = sm.key[P][q];
            int const rr = sm.row[P][q];
            if (k > gk || (k == gk && rr < pv)) {
                gk = k;
                pv = rr;
            }
        }
        int const p = uni(pv);
        // ---- swap rows c and p of the panel
        int const owner = p & (kPt - 1), pi = p / kPt;
        if (tid == owner) {
#pragma unroll
            for (int q = 0; q < W; ++q) sm.prow[P][q] = pick(pi == 0, v[0][q], v[1][q]);
        }
        if (tid == c) {
#pragma unroll
            for (int q = 0; q < W; ++q) sm.trow[P][q] = v[0][q];
        }
        if (tid == 0) {
            sm.piv[c] = k0 + p;
            int const x = sm.rowat[c], y = sm.rowat[p];
            sm.rowat[c] = y;
            sm.rowat[p] = x;
        }
        __syncthreads();
        {  // branch-free per element (with p == c both rows are the same one)
            bool const top = tid == c, own0 = tid == owner && pi == 0, own1 = tid == owner && pi != 0;
#pragma unroll
            for (int q = 0; q < W; ++q) {
                T const u = sm.prow[P][q], t = sm.trow[P][q];
                v[0][q] = pick(top, u, pick(own0, t, v[0][q]));
                v[1][q] = pick(own1, t, v[1][q]);
            }
        }
        T const piv = sm.prow[P][0];
        if (uni(piv == T(0))) {
            fail = c + 1;
            break;
        }
        // ---- rows below: l = a / pivot, then the rank-1 update of the
        // columns still to come (positions 1 .. nb - 1 - c)
        int const lim = nb - c;
#pragma unroll
        for (int i = 0; i < kPr; ++i) {
            int const r = tid + kPt * i;
            if (r > c && r < rows) {
                T const l = v[i][0] / piv;
                v[i][0] = l;
#pragma unroll
                for (int q = 1; q < W; ++q)
                    if (q < lim) v[i][q] = fma_t(-l, sm.prow[P][q], v[i][q]);
            }
        }
#pragma unroll
        for (int i = 0; i < kPr; ++i) {
            T const t0 = v[i][0];
#pragma unroll
            for (int q = 0; q + 1 < W; ++q) v[i][q] = v[i][q + 1];
            v[i][W - 1] = t0;
        }
    }
    int const ce = c;  // position q holds column (q + ce) mod W

    // ---- the panel, its pivots and (on a zero pivot) info
#pragma unroll
    for (int q = 0; q < W; ++q) {
        int const col = q + ce < W ? q + ce : q + ce - W;
        if (uni(col < nb)) {  // uniform: only the two row masks stay live
#pragma unroll
            for (int i = 0; i < kPr; ++i)
                if (okr[i]) at(base, roff[i] + col * cstep) = v[i][q];
        }
    }
    int const done = fail ? fail : nb;  // columns whose pivot was chosen
    if (tid < done) ipiv[b * ipiv_stride + k0 + tid] = sm.piv[tid];
    if (fail) {
        // the rest of the panel's pivots: no interchange (keeps ipiv a valid
        // sequence when there is no info array to stop later panels)
        if (tid >= done && tid < nb) ipiv[b * ipiv_stride + k0 + tid] = k0 + tid;
        if (info && tid == 0) info[b] = k0 + fail;
        return;
    }
    if (tid < nb) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            int const col = q + ce < W ? q + ce : q + ce - W;
            if (col < tid) sm.l11[tid][col] = v[0][q];
        }
    }
    __syncthreads();

    // ---- the columns right of the panel: interchanges (gather, then
    // scatter) and U12.  The columns left of it take the interchanges of all
    // later panels in one pass at the end (getrf_left_swaps_kernel): no
    // later step reads them.
    for (int q = tid; q < N - k0 - nb; q += kPt) {
        int const j = k0 + nb + q, coff = ax.c(j);
        // LDS reads indexed through an opaque zero: read again per column,
        // not hoisted out of the loop into (too many) registers
        int const z0 = opaque_s(0);
        int const* const rowat = sm.rowat + z0;
        int const* const pivs = sm.piv + z0;
        T const* const l11 = &sm.l11[0][0] + z0;
        // slots c >= nb (last panel only) repeat slot nb - 1: no per-slot
        // guards (their masks would be kept in scalar registers)
        int const nbq = opaque_s(nb);  // the clamped slots formed here, not hoisted
        T x[W], z[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
            int const cs = c < nbq ? c : nbq - 1;
            x[c] = at(base, ax.r(rowat[cs]) + coff);
            z[c] = at(base, ax.r(rowat[pivs[cs] - k0]) + coff);
        }
        {
#pragma unroll
            for (int c = 1; c < W; ++c) {
#pragma unroll
                for (int cc = 0; cc < c; ++cc) x[c] = fma_t(-l11[c * (kPnb + 1) + cc], x[cc], x[c]);
            }
        }
        // rows piv[c] first: where one is also a row of the block, the
        // block's store (the solved value) comes last; the block's stores in
        // decreasing c, so slot nb - 1's own value lands after its repeats
#pragma unroll
        for (int c = 0; c < W; ++c) at(base, ax.r(pivs[c < nbq ? c : nbq - 1]) + coff) = z[c];
        int const kx = opaque_s(k0);
#pragma unroll
        for (int c = W - 1; c >= 0; --c) at(base, ax.r(kx + (c < nbq ? c : nbq - 1)) + coff) = x[c];
    }
}

// The row interchanges of panels 1.. applied to the columns left of them,
// after the last panel: for panel P (columns k0_P.., width W) every column
// left of it needs the interchanges of panels P, P+1, ... in order.  From
// the last panel down: sig = panel P's interchanges replayed on the identity
// (thread 0, W swaps), g = sig o g (g: the row each row takes its content
// from under panels P..last), then panel P - 1's columns take rows k0_P..n
// through g, all loads before all stores.  Rows before k0_P never move.
// Instances with info != 0 (a zero pivot stopped them) are left as they are.
template <typename T>
__global__ __launch_bounds__(kPt) void getrf_left_swaps_kernel(Mat<T> Am, int W, const int64_t* ipiv,
                                                               int64_t ipiv_stride, const int32_t* info) {
    constexpr int kCols = kPnb;  // columns of one panel in registers (W <= 32)
    __shared__ int g[kPmax], sig[kPmax];
    int64_t const b = blockIdx.x;
    if (info && info[b] != 0) return;
    T* const base = Am.inst(b);
    Split<T> const ax{Am.spacing, Am.panel};
    int const n = Am.rows, tid = threadIdx.x;
    const int64_t* const piv = ipiv + b * ipiv_stride;
    for (int r = tid; r < n; r += kPt) g[r] = r;
    int const npan = (n + W - 1) / W;
    for (int P = npan - 1; P >= 1; --P) {
        int const kP = P * W, nbP = n - kP < W ? n - kP : W;
        for (int r = kP + tid; r < n; r += kPt) sig[r] = r;
        __syncthreads();
        if (tid == 0) {
            for (int c = 0; c < nbP; ++c) {
                int const row = kP + c;
                int64_t p = piv[row];
                if (p < row || p >= n) p = row;  // (a malformed entry swaps nothing)
                int const x = sig[row], y = sig[p];
                sig[row] = y;
                sig[p] = x;
            }
        }
        __syncthreads();
        for (int r = kP + tid; r < n; r += kPt) g[r] = sig[g[r]];
        __syncthreads();
        // panel P - 1's columns, rows kP..n-1 (at most 2 per thread)
        // (offsets stepped one column at a time through an opaque add, and
        // columns beyond W repeat the last one: no per-column scalar values
        // or masks held live)
        int const kQ = kP - W, cstep = ax.c(1), cq = ax.c(kQ);
        T v[2][kCols];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            int const r = kP + tid + kPt * i;
            int o = ax.r(r < n ? g[r] : kP) + cq;
#pragma unroll
            for (int c = 0; c < kCols; ++c) {
                v[i][c] = at(base, o);
                o = opaque_v(o + (c + 1 < W ? cstep : 0));
            }
        }
        __syncthreads();  // every load of these columns before any store
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            int const r = kP + tid + kPt * i;
            if (r < n) {
                // stores in decreasing column order from column W - 1 (the
                // repeats of it hold its own value)
                int o = ax.r(r) + cq + (W - 1) * cstep;
#pragma unroll
                for (int c = kCols - 1; c >= 0; --c) {
                    at(base, o) = v[i][c];
                    o = opaque_v(o - (c <= W - 1 ? cstep : 0));
                }
            }
        }
        __syncthreads();  // (sig is rebuilt next)
    }
}

}  // namespace

bool getrf_panel_covers(int m, int n, int trans) { return !trans && m == n && n > 256 && n <= kPmax; }
template <typename T>
int getrf_panel_width() {
    return kPw<T>;
}

template <typename T>
int launch_getrf_panel(hipStream_t s, Mat<T> A, int k0, int nb, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                       int32_t* info) {
    if (batch == 0) return 0;
    if (A.trans || A.rows - k0 > kPmax || nb < 1 || nb > kPw<T> || k0 < 0 || k0 + nb > A.cols) {
        set_error("getrf_panel: shape not covered (rows %d, k0 %d, nb %d)", A.rows, k0, nb);
        return BLAST_HIP_ERR_RUNTIME;
    }
    hipLaunchKernelGGL((getrf_panel_kernel<T>), dim3(unsigned(batch)), dim3(kPt), 0, s, A, k0, nb, ipiv, ipiv_stride,
                       info);
    return check_launch("getrf_panel_kernel");
}

template <typename T>
int launch_getrf_left_swaps(hipStream_t s, Mat<T> A, const int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                            const int32_t* info) {
    if (batch == 0 || A.rows <= kPw<T>) return 0;
    if (A.trans || A.rows > kPmax || A.rows != A.cols) {
        set_error("getrf_left_swaps: shape not covered (%d x %d)", A.rows, A.cols);
        return BLAST_HIP_ERR_RUNTIME;
    }
    hipLaunchKernelGGL((getrf_left_swaps_kernel<T>), dim3(unsigned(batch)), dim3(kPt), 0, s, A, kPw<T>, ipiv,
                       ipiv_stride, info);
    return check_launch("getrf_left_swaps_kernel");
}

template int launch_getrf_left_swaps<double>(hipStream_t, Mat<double>, const int64_t*, int64_t, int64_t,
                                             const int32_t*);
template int launch_getrf_left_swaps<float>(hipStream_t, Mat<float>, const int64_t*, int64_t, int64_t,
                                            const int32_t*);
template int getrf_panel_width<double>();
template int getrf_panel_width<float>();
template int launch_getrf_panel<double>(hipStream_t, Mat<double>, int, int, int64_t*, int64_t, int64_t, int32_t*);
template int launch_getrf_panel<float>(hipStream_t, Mat<float>, int, int, int64_t*, int64_t, int64_t, int32_t*);

}  // namespace blast_hip
