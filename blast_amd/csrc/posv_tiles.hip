// Fused batched Cholesky factor + solve for n <= 128, FP64, on MFMA tiles.
//
// One 256-thread workgroup (4 wavefronts) per instance.  The matrix is cut
// into 16x16 tiles; the lower-triangle tiles live in VGPRs as
// v_mfma_f64_16x16x4 accumulators, owned by wavefront owner(i) for tile row i
// (rows {w, 7-w} per wavefront: 9 tiles each at n = 128).  Right-looking
// blocked Cholesky, block column kb = 0..NT-1:
//   A  owners stage tile (kb,kb) and the panel tiles (i,kb) in LDS;
//   B  the owner of row kb factors the 16x16 diagonal tile and builds
//      W = L_kk^{-1} in registers (DPP row broadcasts, Newton-refined rsqrt);
//   C  L_ik = A_ik W^T on MFMA, written to HBM and to the LDS panel;
//      forward substitution y_kb = W b_kb on the RHS;
//   D  trailing update A_ij -= L_ik L_jk^T on MFMA for every owned tile with
//      kb < j <= i, and b_i -= L_ik y_kb.
// Then the backward substitution x = L^{-T} y, column block by column block,
// with L read straight from the accumulators.
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// factorisation reorders BLAST's sums (agreement to backward error, not bits).
// HBM traffic per instance: the lower triangle of A is read, the lower
// triangle of L written, B read and X written.

#include <type_traits>

#include "common.hpp"
#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kTileStride = 17;               // LDS tile: T[m][k] at k*17 + m (bank-conflict padding)
constexpr int kTileElems = 16 * kTileStride;  // doubles per LDS tile
constexpr int kMaxRhs = 4;

// v_mfma_f64_16x16x4f64 operand and accumulator layouts (lane l, register r).
// A operand: A[l % 16][l / 16]; B operand: B[l / 16][l % 16];
// accumulator register r of lane l holds D[row][l % 16] with row given here.
__device__ __forceinline__ int acc_row(int lane, int r) { return (lane >> 4) + 4 * r; }
__device__ __forceinline__ int acc_col(int lane) { return lane & 15; }

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Operand element for K-chunk q: T[l % 16][4q + l / 16] of an LDS tile.
__device__ __forceinline__ double op_read(const double* tile, int q, int lane) {
    return tile[(4 * q + (lane >> 4)) * kTileStride + (lane & 15)];
}

__device__ __forceinline__ void acc_to_lds(double* tile, d4 v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) tile[acc_col(lane) * kTileStride + acc_row(lane, r)] = v[r];
}

__device__ __forceinline__ d4 lds_to_acc(const double* tile, int lane) {
    d4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = tile[acc_col(lane) * kTileStride + acc_row(lane, r)];
    return v;
}

// Load tile (ti, tj) of A into accumulator layout.  Outside n x n the matrix
// is padded with the identity; the strict upper part of diagonal tiles is
// not read (BLAST reads A's lower triangle only).
// Element access; PANEL = column-major panel storage, not transposed (the
// layout BLAST's StaticPanelMatrix uses), which needs no runtime dispatch.
template <bool PANEL>
__device__ __forceinline__ double& at(Ref<double> const& a, int row, int col) {
    if constexpr (PANEL)
        return a.p[int64_t(row >> 2) * a.m.spacing + (row & 3) + 4 * col];
    else
        return a(row, col);
}

template <bool PANEL>
__device__ __forceinline__ d4 load_tile(Ref<double> const& a, int n, int ti, int tj, int lane) {
    d4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + acc_row(lane, r), col = 16 * tj + acc_col(lane);
        v[r] = (row < n && col < n) ? (row >= col ? at<PANEL>(a, row, col) : 0.0) : (row == col ? 1.0 : 0.0);
    }
    return v;
}

// Store tile (ti, tj) of L.  Dense storage: i >= j only.  Panel storage: also
// zeros in the strict upper part of each 4x4 diagonal block (what BLAST's
// panel potrf leaves there is rounding residue of 0).
template <bool PANEL>
__device__ __forceinline__ void store_tile(Ref<double> const& l, bool panel, int n, int ti, int tj, d4 v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + acc_row(lane, r), col = 16 * tj + acc_col(lane);
        if (row < n && col < n && (row >= col || (panel && (row >> 2) == (col >> 2))))
            at<PANEL>(l, row, col) = row >= col ? v[r] : 0.0;
    }
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Broadcast lane C of every 16-lane row to the whole row (DPP row_newbcast).
template <int C>
__device__ __forceinline__ double row_bcast(double v) {
    long long const x = __builtin_bit_cast(long long, v);
    int const lo = __builtin_amdgcn_mov_dpp(int(x), 0x150 + C, 0xF, 0xF, false);
    int const hi = __builtin_amdgcn_mov_dpp(int(x >> 32), 0x150 + C, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// 1/sqrt(d): v_rsq_f64 refined by two Newton steps (about 1 ulp).
__device__ __forceinline__ double rsqrt_nr(double d) {
    double r = __builtin_amdgcn_rsq(d);
    double const h = 0.5 * d;
    r = __builtin_fma(r, __builtin_fma(-h * r, r, 0.5), r);
    r = __builtin_fma(r, __builtin_fma(-h * r, r, 0.5), r);
    return r;
}

// Factorisation of the 16x16 diagonal tile in sD, and its inverse.
// Each 16-lane row of the wavefront holds a copy of the tile, lane i of the
// row holding tile row i; every broadcast is a DPP row_newbcast.
// 1. Cholesky, left-looking: for column k, lane i forms a_ik - <l_i, l_k>
//    (row k broadcast from lane k), then r = 1/sqrt(d), l_kk = d*r,
//    l_ik = (.)*r.
// 2. W = L_kk^{-1}, lane i computing row i: row j is finalised at step j
//    and broadcast from lane j to the lanes below it.
// On exit sD holds L_kk (zeros above the diagonal) and sW holds W, both as
// T[m][k] at k*17 + m.  Returns 0 or the 1-based tile column of the first
// non-positive pivot.  Branch-free: every lane runs the same instructions.
__device__ int diag_factor(double* sD, double* sW, int lane) {
    int const i = lane & 15;
    double t[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) t[c] = sD[c * kTileStride + i];
    int bad = 0;
    double myr = 0.0;
    static_for<0, 16>([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value;
        double acc0 = t[k], acc1 = 0.0;
        static_for<0, k>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            if constexpr (j % 2 == 0)
                acc0 = __builtin_fma(-t[j], row_bcast<k>(t[j]), acc0);
            else
                acc1 = __builtin_fma(-t[j], row_bcast<k>(t[j]), acc1);
        });
        double const v = acc0 + acc1;
        double const d = row_bcast<k>(v);
        bad = (bad == 0 && !(d > 0.0)) ? k + 1 : bad;
        double const r = rsqrt_nr(d);
        myr = (i == k) ? r : myr;
        double lik = ((i == k) ? d : v) * r;
        lik = (i < k) ? 0.0 : lik;
        t[k] = lik;
        __builtin_amdgcn_sched_barrier(0);
    });
    // W row by row: at step j lane j finalises W[j][:] = (e_j - acc) / l_jj;
    // lanes i > j then accumulate acc[c] += l_ij W[j][c], c <= j.
    double w[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) w[c] = 0.0;
    static_for<0, 16>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        static_for<0, j>([&](auto C) __attribute__((always_inline)) {
            constexpr int c = decltype(C)::value;
            w[c] = (i == j) ? -w[c] * myr : w[c];
        });
        w[j] = (i == j) ? myr : w[j];
        double const f = (i > j) ? t[j] : 0.0;
        static_for<0, j + 1>([&](auto C) __attribute__((always_inline)) {
            constexpr int c = decltype(C)::value;
            w[c] = __builtin_fma(f, row_bcast<j>(w[c]), w[c]);
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            sD[c * kTileStride + i] = (c <= i) ? t[c] : 0.0;  // L[i][c]
            sW[c * kTileStride + i] = (c <= i) ? w[c] : 0.0;  // W[i][c]
        }
    }
    return bad;
}

constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }

// One wavefront per instance.  acc holds the lower tiles (i, j), j <= i, as
// MFMA accumulators (AGPR/VGPR resident); LDS stages the current block
// column, keeps W_kb = L_kb,kb^{-1} for the backward solve and holds the
// right-hand sides.  A single wavefront needs no barriers: its LDS accesses
// complete in program order.
template <int NT, bool PANEL>
__global__ __launch_bounds__(64) void posv_wave_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                       int n, int nrhs, int32_t* info) {
    __shared__ double sP[NT][kTileElems];          // block column kb: tile i at sP[i] (i >= kb)
    __shared__ double sW[NT][kTileElems];          // W_kb
    __shared__ double sR[NT * 16 * kMaxRhs + 64];  // RHS -> y -> x as [row][rhs]; tail: partial sums

    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x;
    Ref<double> a(A, b), l(L, b);
    bool const lpanel = L.panel != 0;
    constexpr int NTILES = NT * (NT + 1) / 2;

    d4 acc[NTILES];
    static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            acc[tri(i, j)] = load_tile<PANEL>(a, n, i, j, lane);
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
        store_tile<PANEL>(l, lpanel, n, kb, kb, lds_to_acc(sP[kb], lane), lane);
        // panel: L_ik = A_ik W_kb^T
        static_for<1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            static_for<0, i>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                if constexpr (j == kb) {
                    d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) z = mfma(op_read(sP[i], q, lane), op_read(sW[kb], q, lane), z);
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
                    if constexpr (j > kb) acc[tri(i, j)] = mfma(-f[i][q], f[j][q], acc[tri(i, j)]);
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
                    int const row = 16 * i + acc_row(lane, r);
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
                     int64_t batch, int32_t* info) {
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    if (panel)
        hipLaunchKernelGGL((posv_wave_kernel<NT, true>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n, nrhs,
                           info);
    else
        hipLaunchKernelGGL((posv_wave_kernel<NT, false>), dim3(unsigned(batch)), dim3(64), 0, s, A, L, B, X, n,
                           nrhs, info);
    return check_launch("posv_wave_kernel");
}

// Fused potrf (+ forward/backward solve when nrhs > 0) for n <= 128,
// nrhs <= 4.  Returns 1 when the shape is not covered.
int launch_posv_tiles(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    if (L.rows != L.cols || n > 128 || nrhs > kMaxRhs) return 1;
    if (batch == 0) return 0;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_nt<1>(s, A, L, B, X, n, nrhs, batch, info);
        case 2: return launch_nt<2>(s, A, L, B, X, n, nrhs, batch, info);
        case 3: return launch_nt<3>(s, A, L, B, X, n, nrhs, batch, info);
        case 4: return launch_nt<4>(s, A, L, B, X, n, nrhs, batch, info);
        case 5: return launch_nt<5>(s, A, L, B, X, n, nrhs, batch, info);
        case 6: return launch_nt<6>(s, A, L, B, X, n, nrhs, batch, info);
        case 7: return launch_nt<7>(s, A, L, B, X, n, nrhs, batch, info);
        default: return launch_nt<8>(s, A, L, B, X, n, nrhs, batch, info);
    }
}

}  // namespace blast_hip
