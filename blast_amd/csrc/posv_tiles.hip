// Fused batched Cholesky factor + solve for n <= 128, FP64, on MFMA tiles.
//
// One 256-thread workgroup (4 wavefronts) per instance.  The matrix is cut
// into 16x16 tiles; the lower-triangle tiles live in VGPRs as
// v_mfma_f64_16x16x4 accumulators, owned by wavefront owner(i) for tile row i
// (rows {w, 7-w} per wavefront: 9 tiles each at n = 128).  Right-looking
// blocked Cholesky, block column kb = 0..NT-1:
//   A  owners stage tile (kb,kb) and the panel tiles (i,kb) in LDS;
//   B  the owner of row kb factors the 16x16 diagonal tile with a
//      lane-per-row sweep (readlane broadcasts, IEEE sqrt/divide) that also
//      builds W = L_kk^{-1};
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

__device__ __forceinline__ double readlane_d(double v, int l) {
    long long const x = __builtin_bit_cast(long long, v);
    int const lo = __builtin_amdgcn_readlane(int(x), l);
    int const hi = __builtin_amdgcn_readlane(int(x >> 32), l);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Tile row -> owning wavefront: rows {w, 7-w} go to wavefront w.
__device__ __forceinline__ int owner(int i) { return i < 4 ? i : 7 - i; }

// Load tile (ti, tj) of A into accumulator layout.  Outside n x n the matrix
// is padded with the identity; the strict upper part of diagonal tiles is
// not read (BLAST reads A's lower triangle only).
__device__ __forceinline__ d4 load_tile(Ref<double> const& a, int n, int ti, int tj, int lane) {
    d4 v;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + acc_row(lane, r), col = 16 * tj + acc_col(lane);
        v[r] = (row < n && col < n) ? (row >= col ? a(row, col) : 0.0) : (row == col ? 1.0 : 0.0);
    }
    return v;
}

// Store tile (ti, tj) of L.  Dense storage: i >= j only.  Panel storage: also
// zeros in the strict upper part of each 4x4 diagonal block (what BLAST's
// panel potrf leaves there is rounding residue of 0).
__device__ __forceinline__ void store_tile(Ref<double> const& l, bool panel, int n, int ti, int tj, d4 v, int lane) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int const row = 16 * ti + acc_row(lane, r), col = 16 * tj + acc_col(lane);
        if (row < n && col < n && (row >= col || (panel && (row >> 2) == (col >> 2)))) l(row, col) = row >= col ? v[r] : 0.0;
    }
}

// Lane-per-row factorisation of the 16x16 tile in sD (lanes 0..15 hold rows;
// the other lanes shadow them).  On exit sD holds L_kk (zeros above the
// diagonal) and sW holds W = L_kk^{-1}.  Returns 0 or the 1-based tile column
// of the first non-positive pivot.
__device__ int diag_factor(double* sD, double* sW, int lane) {
    int const i = lane & 15;
    double t[16], wv[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        t[c] = sD[c * kTileStride + i];
        wv[c] = (c == i) ? 1.0 : 0.0;
    }
    int bad = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        double const d = readlane_d(t[k], k);
        if (!(d > 0.0) && bad == 0) bad = k + 1;
        double const s = __builtin_sqrt(d);
        double const lkk = d / s;
        double const lik = (i > k) ? t[k] / s : (i == k ? lkk : 0.0);
        t[k] = lik;
        double const sc = (i == k) ? 1.0 / lkk : 1.0;
#pragma unroll
        for (int c = 0; c <= k; ++c) wv[c] *= sc;
#pragma unroll
        for (int c = k + 1; c < 16; ++c) t[c] = __builtin_fma(-lik, readlane_d(lik, c), t[c]);
        double const f = (i > k) ? lik : 0.0;
#pragma unroll
        for (int c = 0; c <= k; ++c) wv[c] = __builtin_fma(-f, readlane_d(wv[c], k), wv[c]);
    }
    if (lane < 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            sD[c * kTileStride + i] = (c <= i) ? t[c] : 0.0;
            sW[c * kTileStride + i] = (c <= i) ? wv[c] : 0.0;
        }
    }
    return bad;
}

template <int NT>
__global__ __launch_bounds__(256) void posv_tile_kernel(Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X,
                                                        int n, int nrhs, int32_t* info) {
    __shared__ double sP[2][NT][kTileElems];
    __shared__ double sD[kTileElems];
    __shared__ double sW[NT][kTileElems];
    __shared__ double sR[NT * 16 * kMaxRhs];
    __shared__ double sS[4][16 * kMaxRhs];
    __shared__ int sInfo;

    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x & 63;
    int const w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));  // wavefront-uniform
    Ref<double> a(A, b), l(L, b);
    bool const lpanel = L.panel != 0;

    // Tile slots of this wavefront: row r0 = w (tiles (r0, 0..r0)), row r1 = 7-w.
    int const r0 = w, r1 = 7 - w;
    constexpr int S = 9;
    d4 acc[S];
    int ti[S], tj[S];
    bool ok[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
        bool const first = t <= r0;
        ti[t] = first ? r0 : r1;
        tj[t] = first ? t : t - r0 - 1;
        ok[t] = ti[t] < NT && tj[t] <= ti[t] && (first || r1 >= 4);
        acc[t] = ok[t] ? load_tile(a, n, ti[t], tj[t], lane) : d4{0.0, 0.0, 0.0, 0.0};
    }
    if (nrhs > 0) {
        Ref<double> bb(B, b);
        for (int e = threadIdx.x; e < NT * 16 * kMaxRhs; e += 256) {
            int const row = e >> 2, c = e & 3;
            sR[e] = (row < n && c < nrhs) ? bb(row, c) : 0.0;
        }
    }
    if (threadIdx.x == 0) sInfo = 0;
    __syncthreads();

    for (int kb = 0; kb < NT; ++kb) {
        double(*P)[kTileElems] = sP[kb & 1];
        // A: stage column kb
#pragma unroll
        for (int t = 0; t < S; ++t)
            if (ok[t] && tj[t] == kb) acc_to_lds(ti[t] == kb ? sD : P[ti[t]], acc[t], lane);
        __syncthreads();
        // B: diagonal tile
        if (w == owner(kb)) {
            int const bad = diag_factor(sD, sW[kb], lane);
            if (bad && lane == 0 && sInfo == 0 && 16 * kb + bad <= n) sInfo = 16 * kb + bad;
        }
        __syncthreads();
        // C: panel solve L_ik = A_ik W^T, forward substitution on the RHS
#pragma unroll
        for (int t = 0; t < S; ++t)
            if (ok[t] && tj[t] == kb && ti[t] > kb) {
                d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) z = mfma(op_read(P[ti[t]], q, lane), op_read(sW[kb], q, lane), z);
                acc[t] = z;
                store_tile(l, lpanel, n, ti[t], kb, z, lane);
                acc_to_lds(P[ti[t]], z, lane);
            }
        if (w == owner(kb)) {
            store_tile(l, lpanel, n, kb, kb, lds_to_acc(sD, lane), lane);
            if (nrhs > 0) {
                int const m = lane >> 2, c = lane & 3;
                double rv[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) rv[k] = sR[(16 * kb + k) * kMaxRhs + c];
                double y = 0.0;
#pragma unroll
                for (int k = 0; k < 16; ++k) y = __builtin_fma(sW[kb][k * kTileStride + m], rv[k], y);
                sR[(16 * kb + m) * kMaxRhs + c] = y;
            }
        }
        __syncthreads();
        // D: trailing update and RHS update
#pragma unroll
        for (int t = 0; t < S; ++t)
            if (ok[t] && tj[t] > kb) {
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[t] = mfma(-op_read(P[ti[t]], q, lane), op_read(P[tj[t]], q, lane), acc[t]);
            }
            __builtin_amdgcn_sched_barrier(0);  // keep one slot's operands live at a time
        if (nrhs > 0) {
            int const m = lane >> 2, c = lane & 3;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                int const i = s == 0 ? r0 : r1;
                if (i > kb && i < NT && (s == 0 || r1 >= 4)) {
                    double v = sR[(16 * i + m) * kMaxRhs + c];
#pragma unroll
                    for (int k = 0; k < 16; ++k)
                        v = __builtin_fma(-P[i][k * kTileStride + m], sR[(16 * kb + k) * kMaxRhs + c], v);
                    sR[(16 * i + m) * kMaxRhs + c] = v;
                }
            }
        }
    }

    if (nrhs > 0) {
        __syncthreads();
        // Backward substitution x = L^{-T} y, block column kb = NT-1..0.
        for (int kb = NT - 1; kb >= 0; --kb) {
            double p[kMaxRhs] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int t = 0; t < S; ++t)
                if (ok[t] && tj[t] == kb && ti[t] > kb) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        int const row = 16 * ti[t] + acc_row(lane, r);
#pragma unroll
                        for (int c = 0; c < kMaxRhs; ++c) p[c] = __builtin_fma(acc[t][r], sR[row * kMaxRhs + c], p[c]);
                    }
                }
#pragma unroll
            for (int c = 0; c < kMaxRhs; ++c) {
                p[c] += __shfl_xor(p[c], 16);
                p[c] += __shfl_xor(p[c], 32);
            }
            if (lane < 16)
#pragma unroll
                for (int c = 0; c < kMaxRhs; ++c) sS[w][lane * kMaxRhs + c] = p[c];
            __syncthreads();
            if (w == owner(kb)) {
                int const m = lane >> 2, c = lane & 3;
                double tv[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    int const e = k * kMaxRhs + c;
                    tv[k] = sR[(16 * kb + k) * kMaxRhs + c] - ((sS[0][e] + sS[1][e]) + (sS[2][e] + sS[3][e]));
                }
                double x = 0.0;
#pragma unroll
                for (int k = 0; k < 16; ++k) x = __builtin_fma(sW[kb][m * kTileStride + k], tv[k], x);
                sR[(16 * kb + m) * kMaxRhs + c] = x;
            }
            __syncthreads();
        }
        Ref<double> xx(X, b);
        for (int e = threadIdx.x; e < n * nrhs; e += 256) {
            int const row = e % n, c = e / n;
            xx(row, c) = sR[row * kMaxRhs + c];
        }
    }
    if (info && threadIdx.x == 0) info[b] = sInfo;
}

template <int NT>
static int launch_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                     int64_t batch, int32_t* info) {
    hipLaunchKernelGGL(posv_tile_kernel<NT>, dim3(unsigned(batch)), dim3(256), 0, s, A, L, B, X, n, nrhs, info);
    return check_launch("posv_tile_kernel");
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
