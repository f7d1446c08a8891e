// Batched mixed-precision factor + solve for n <= 256, nrhs <= 8: Cholesky
// (potrf + two trsm) or LU with partial pivoting in FP32 (optionally BF16
// products in the trailing updates) with FP64 iterative refinement (LAPACK
// dsposv / dsgesv contracts).  The FP64 factorisations have their own
// kernels (posv_tiles.hip and chol_owner.hip for Cholesky, lu_wg.hip for LU).
//
// One 256-thread workgroup (4 wavefronts) per instance.  The factor F lives
// in global memory (the caller's L / A buffer in FP64, a dense work buffer in
// FP32) and is updated in place, right-looking by 16-column blocks:
//   Cholesky block kb: the block column is staged in LDS; wavefront 0
//     factors the diagonal tile and forms W = L_kk^{-1} (diag_factor); the
//     panel is L_ik = A_ik W^T on MFMA; the trailing tiles get
//     A_ij -= L_ik L_jk^T on MFMA with both operands from LDS.
//   LU block kb: the block column is staged in LDS and factored there
//     (getf2: first-maximum pivot search over the column, row swap, scale,
//     rank-1 update); the row swaps are applied to the other columns in
//     global memory; U_kj = L_kk^{-1} A_kj on MFMA (kept in LDS); the
//     trailing tiles get A_ij -= L_ik U_kj.
// The diagonal-tile inverses stay in LDS for the triangular solves, which run
// block by block on MFMA (16 x 16 tiles times the 16 x 8 block of right-hand
// sides).  Refinement: r = b - A x in FP64 (MFMA f64, A re-read from HBM),
// d = F^{-1} r in the factor's precision, x += d in FP64, until
// max|r| <= max|x| * ||A||_inf * tol per right-hand side (tol = eps*sqrt(n)
// when 0, LAPACK's test) or itmax refinements.
// Algorithm references: dense/Potrf.hpp:66-106, dense/Trsm.hpp:17-66,
// dense/Getrf.hpp:52-119, dense/Getf2.hpp:52-77, dense/Iamax.hpp:82-95,
// dense/Laswp.hpp:37-43 (same pivots; sums reordered on MFMA tiles).

#include <type_traits>

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kWgThreads = 256;
constexpr int kWgWaves = kWgThreads / 64;
constexpr int kWgMaxNT = 16;  // n <= 256 (the solves assume at most 4 tiles per wavefront per step)
constexpr int kWgRhs = 8;     // right-hand-side columns held per instance

struct WgArgs {
    Mat<double> A, B, X;
    int n, nrhs, itmax;
    double tol;
    int32_t* info;
    int32_t* iter;
    const int32_t* only_if;
    int64_t* ipiv;
    int64_t ipiv_stride;
};

__device__ __forceinline__ unsigned long long abs_bits(double v) {
    return __builtin_bit_cast(unsigned long long, v < 0.0 ? -v : v);
}

template <typename T, bool LU, bool BF16>
__global__ __launch_bounds__(kWgThreads) void wg_solve_kernel(WgArgs g, Mat<T> F) {
    using Mf = Mfma<T>;
    using acc_t = typename Mf::acc_t;
    using Md = Mfma<double>;
    constexpr bool kMixed = std::is_same_v<T, float>;
    constexpr int kFactorBytes = (LU ? 2 : 1) * kWgMaxNT * kTileElems * int(sizeof(T));
    constexpr int kSolveBytes = kWgMaxNT * 16 * kWgRhs * int(sizeof(double) + sizeof(T));
    constexpr int kUnionBytes = kFactorBytes > kSolveBytes ? kFactorBytes : kSolveBytes;

    __shared__ T sW1[kWgMaxNT][kTileElems];            // Cholesky: L_kk^{-1}; LU: unit L_kk^{-1}
    __shared__ T sW2[LU ? kWgMaxNT : 1][kTileElems];   // LU: U_kk^{-1}
    __shared__ __attribute__((aligned(16))) unsigned char sBuf[kUnionBytes];
    __shared__ int sPiv[LU ? kWgMaxNT * 16 : 1];
    __shared__ unsigned long long sMax[2 * kWgRhs];
    __shared__ T sRedV[kWgWaves];
    __shared__ int sRedI[kWgWaves];
    __shared__ unsigned long long sAnrmBits;  // ||A||_inf as the bits of a non-negative double
    __shared__ int sFlag;

    // factor phase: block column (and LU: block row) tiles, T[m][k] at k*17+m
    T(*sP)[kTileElems] = reinterpret_cast<T(*)[kTileElems]>(sBuf);
    T(*sU)[kTileElems] = sP + kWgMaxNT;  // LU: U_kj stored as element (k, n) at k*17 + n
    // solve phase: x (FP64) and the working right-hand side (T), [row][kWgRhs]
    double* sX = reinterpret_cast<double*>(sBuf);
    T* sR = reinterpret_cast<T*>(sBuf + kWgMaxNT * 16 * kWgRhs * sizeof(double));

    int64_t const b = blockIdx.x;
    if (g.only_if && g.only_if[b] >= 0) return;
    int const tid = threadIdx.x, lane = tid & 63;
    int const w = __builtin_amdgcn_readfirstlane(tid >> 6);
    int const n = g.n, nrhs = g.nrhs, NT = (n + 15) >> 4, npad = 16 * NT;
    Ref<double> a(g.A, b);
    Ref<T> f(F, b);
    // F outside n x n reads as the identity (the factorisation of blkdiag(A, I))
    auto fget = [&](int i, int j) -> T { return (i < n && j < n) ? f(i, j) : T(i == j ? 1 : 0); };
    auto fset = [&](int i, int j, T v) {
        if (i < n && j < n) f(i, j) = v;
    };
    auto asym = [&](int i, int j) -> double { return LU ? a(i, j) : (i >= j ? a(i, j) : a(j, i)); };

    if (tid == 0) sFlag = 0;
    if (kMixed && tid == 0 && g.info) g.info[b] = 0;  // an FP64 fallback overwrites it
    if (tid == 0) sAnrmBits = 0ull;
    // ---- F = A (converted; Cholesky: lower triangle), eight loads in flight per thread
    if (static_cast<const void*>(F.data) != static_cast<const void*>(g.A.data) || F.stride != g.A.stride) {
        constexpr int U = 8;
        int const nn = n * n;
        for (int e0 = tid; e0 < nn; e0 += kWgThreads * U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int const e = e0 + u * kWgThreads, r = e % n, c = e / n;
                v[u] = (e < nn && (LU || r >= c)) ? a(r, c) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int const e = e0 + u * kWgThreads, r = e % n, c = e / n;
                if (e < nn && (LU || r >= c)) f(r, c) = T(v[u]);
            }
        }
    }
    __syncthreads();

    // ---- factorisation
    for (int kb = 0; kb < NT; ++kb) {
        int const r0 = 16 * kb;
        {
            int const ne = (NT - kb) * 256;
            for (int e0 = tid; e0 < ne; e0 += 4 * kWgThreads) {
                T v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const e = e0 + u * kWgThreads, t = kb + (e >> 8), c = (e >> 4) & 15, m = e & 15;
                    v[u] = (e >= ne || (!LU && t == kb && m < c)) ? T(0) : fget(16 * t + m, r0 + c);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const e = e0 + u * kWgThreads, t = kb + (e >> 8), c = (e >> 4) & 15, m = e & 15;
                    if (e < ne) sP[t][c * kTileStride + m] = v[u];
                }
            }
        }
        __syncthreads();
        if constexpr (!LU) {
            if (w == 0) {
                int const bk = diag_factor_inl<T>(sP[kb], sW1[kb], lane);
                if (lane == 0 && bk && sFlag == 0 && r0 + bk <= n) sFlag = r0 + bk;
            }
            __syncthreads();
            for (int i = kb + 1 + w; i < NT; i += kWgWaves) {  // L_ik = A_ik W^T
                acc_t acc{};
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = Mf::mma(op_read(sP[i], q, lane), op_read(sW1[kb], q, lane), acc);
                acc_to_lds(sP[i], acc, lane);
#pragma unroll
                for (int r = 0; r < 4; ++r) fset(16 * i + Mf::row(lane, r), r0 + acc_col(lane), acc[r]);
            }
            for (int e = tid; e < 256; e += kWgThreads) {
                int const c = e >> 4, m = e & 15;
                if (m >= c) fset(r0 + m, r0 + c, sP[kb][c * kTileStride + m]);
            }
            __syncthreads();
            int const mt = NT - kb - 1, ntr = mt * (mt + 1) / 2;
            constexpr int U = 4;  // tiles per wavefront per pass: their loads overlap
            for (int t0 = w; t0 < ntr; t0 += kWgWaves * U) {  // A_ij -= L_ik L_jk^T, kb < j <= i
                acc_t acc[U];
                int ti[U], tj[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int const t = t0 + u * kWgWaves;
                    int ii = int((__builtin_sqrtf(8.f * float(t) + 1.f) - 1.f) * 0.5f);
                    while ((ii + 1) * (ii + 2) / 2 <= t) ++ii;
                    while (ii * (ii + 1) / 2 > t) --ii;
                    ti[u] = t < ntr ? kb + 1 + ii : -1;
                    tj[u] = kb + 1 + (t - ii * (ii + 1) / 2);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[u][r] = ti[u] >= 0 ? fget(16 * ti[u] + Mf::row(lane, r), 16 * tj[u] + acc_col(lane)) : T(0);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ti[u] < 0) continue;
                    if constexpr (BF16) {
                        acc_t const p = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            op_read_bf16(sP[ti[u]], lane), op_read_bf16(sP[tj[u]], lane), acc_t{}, 0, 0, 0);
                        acc[u] = acc[u] - p;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[u] = Mf::mma(-op_read(sP[ti[u]], q, lane), op_read(sP[tj[u]], q, lane), acc[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ti[u] < 0) continue;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        int const row = 16 * ti[u] + Mf::row(lane, r), col = 16 * tj[u] + acc_col(lane);
                        if (ti[u] != tj[u] || row >= col) fset(row, col, acc[u][r]);
                    }
                }
            }
            __syncthreads();
        } else {
            int const prow = npad - r0;
            T(*const pp)[kTileElems] = sP + kb;  // panel rows relative to r0
            for (int c = 0; c < 16; ++c) {  // getf2 on the LDS panel
                T v = T(-1);
                int idx = 0x7fffffff;
                for (int rr = c + tid; rr < prow; rr += kWgThreads)
                    argmax_combine(v, idx, iamax_key(pp[rr >> 4][c * kTileStride + (rr & 15)], rr - c), rr);
                for (int o = 32; o > 0; o >>= 1) {
                    T const v2 = __shfl_xor(v, o);
                    int const i2 = __shfl_xor(idx, o);
                    argmax_combine(v, idx, v2, i2);
                }
                if (lane == 0) sRedV[w] = v, sRedI[w] = idx;
                __syncthreads();
                v = sRedV[0], idx = sRedI[0];
                for (int q = 1; q < kWgWaves; ++q) argmax_combine(v, idx, sRedV[q], sRedI[q]);
                int const p = idx;
                if (tid == 0) sPiv[r0 + c] = r0 + p;
                if (p != c && tid < 16) {
                    T* const x = &pp[c >> 4][tid * kTileStride + (c & 15)];
                    T* const y = &pp[p >> 4][tid * kTileStride + (p & 15)];
                    T const t = *x;
                    *x = *y, *y = t;
                }
                __syncthreads();
                T const piv = pp[c >> 4][c * kTileStride + (c & 15)];
                if (piv == T(0)) {
                    if (tid == 0 && sFlag == 0 && r0 + c < n) sFlag = r0 + c + 1;
                } else {
                    for (int rr = c + 1 + tid; rr < prow; rr += kWgThreads) {
                        T* const row = &pp[rr >> 4][rr & 15];
                        T const l = row[c * kTileStride] / piv;
                        row[c * kTileStride] = l;
                        for (int cc = c + 1; cc < 16; ++cc)
                            row[cc * kTileStride] =
                                fma_t(-l, pp[c >> 4][cc * kTileStride + (c & 15)], row[cc * kTileStride]);
                    }
                }
                __syncthreads();
            }
            for (int e = tid; e < (NT - kb) * 256; e += kWgThreads) {
                int const t = kb + (e >> 8), c = (e >> 4) & 15, m = e & 15;
                fset(16 * t + m, r0 + c, sP[t][c * kTileStride + m]);
            }
            if (tid < 16 && r0 + tid < n) g.ipiv[b * g.ipiv_stride + r0 + tid] = sPiv[r0 + tid];
            for (int j = tid; j < n; j += kWgThreads) {  // laswp on the other columns, in registers
                if (j >= r0 && j < r0 + 16) continue;
                // rows r0..r0+15 and the pivot rows below the block: load all, swap, store
                T v[16], o[16];
                int pr[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    pr[c] = sPiv[r0 + c];
                    v[c] = r0 + c < n ? f(r0 + c, j) : T(0);
                    o[c] = pr[c] >= r0 + 16 && pr[c] < n ? f(pr[c], j) : T(0);
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    int const p = pr[c];
                    if (p == r0 + c) continue;
                    if (p < r0 + 16) {  // both rows inside the block
#pragma unroll
                        for (int d = c + 1; d < 16; ++d)
                            if (p == r0 + d) {
                                T const t = v[c];
                                v[c] = v[d], v[d] = t;
                            }
                    } else {  // the current content of row p: the latest earlier swap with it, else memory
                        T cur = o[c];
#pragma unroll
                        for (int d = 0; d < c; ++d)
                            if (pr[d] == p) cur = o[d];
                        o[c] = v[c];
                        v[c] = cur;
                    }
                }
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    if (r0 + c < n) f(r0 + c, j) = v[c];
                    if (pr[c] >= r0 + 16 && pr[c] < n) f(pr[c], j) = o[c];  // in c order: the last swap wins
                }
            }
            if (tid < 32) {  // inverses of the unit-lower and upper diagonal blocks, one column per lane
                int const cc = tid & 15;
                T const* const d = sP[kb];
                T x[16];
                if (tid < 16) {
#pragma unroll
                    for (int i = 0; i < 16; ++i) {
                        T s = T(0);
#pragma unroll
                        for (int k = 0; k < i; ++k) s = fma_t(-d[k * kTileStride + i], x[k], s);
                        x[i] = (i == cc) ? T(1) : (i < cc ? T(0) : s);
                    }
#pragma unroll
                    for (int i = 0; i < 16; ++i) sW1[kb][cc * kTileStride + i] = x[i];
                } else {
#pragma unroll
                    for (int i = 15; i >= 0; --i) {
                        T s = T(0);
#pragma unroll
                        for (int k = i + 1; k < 16; ++k) s = fma_t(-d[k * kTileStride + i], x[k], s);
                        T const u = d[i * kTileStride + i];
                        x[i] = (i > cc) ? T(0) : ((i == cc ? T(1) : s) / u);
                    }
#pragma unroll
                    for (int i = 0; i < 16; ++i) sW2[kb][cc * kTileStride + i] = x[i];
                }
            }
            __syncthreads();
            for (int j = kb + 1 + w; j < NT; j += kWgWaves) {  // U_kj = L_kk^{-1} A_kj
                acc_t acc{};
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc = Mf::mma(op_read(sW1[kb], q, lane), fget(r0 + 4 * q + (lane >> 4), 16 * j + (lane & 15)), acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const row = Mf::row(lane, r), col = acc_col(lane);
                    fset(r0 + row, 16 * j + col, acc[r]);
                    sU[j][row * kTileStride + col] = acc[r];
                }
            }
            __syncthreads();
            int const mt = NT - kb - 1;
            constexpr int U = 4;
            for (int t0 = w; t0 < mt * mt; t0 += kWgWaves * U) {  // A_ij -= L_ik U_kj
                acc_t acc[U];
                int ti[U], tj[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int const t = t0 + u * kWgWaves;
                    ti[u] = t < mt * mt ? kb + 1 + t / mt : -1;
                    tj[u] = kb + 1 + t % mt;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[u][r] = ti[u] >= 0 ? fget(16 * ti[u] + Mf::row(lane, r), 16 * tj[u] + acc_col(lane)) : T(0);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ti[u] < 0) continue;
                    if constexpr (BF16) {
                        acc_t const p = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                            op_read_bf16(sP[ti[u]], lane), op_read_bf16(sU[tj[u]], lane), acc_t{}, 0, 0, 0);
                        acc[u] = acc[u] - p;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[u] = Mf::mma(-op_read(sP[ti[u]], q, lane), op_read(sU[tj[u]], q, lane), acc[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ti[u] < 0) continue;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        fset(16 * ti[u] + Mf::row(lane, r), 16 * tj[u] + acc_col(lane), acc[u][r]);
                }
            }
            __syncthreads();
        }
    }

    int const bad = sFlag;
    if constexpr (!kMixed) {
        if (!LU && F.panel && !F.trans) {  // panel L: zeros above the diagonal inside 4x4 diagonal blocks
            for (int e = tid; e < n * 4; e += kWgThreads) {
                int const j = e >> 2, i = (j & ~3) + (e & 3);
                if (i < j) f(i, j) = T(0);
            }
        }
        if (tid == 0 && g.info) g.info[b] = bad;
        if (bad || nrhs == 0) return;
    } else {
        if (bad) {
            if (tid == 0) g.iter[b] = -1;  // factorisation broke down: FP64 fallback
            return;
        }
    }

    // ---- solves
    Ref<double> bb(g.B, b);
    for (int e = tid; e < npad * kWgRhs; e += kWgThreads) {
        int const r = e / kWgRhs, c = e % kWgRhs;
        sX[e] = 0.0;
        sR[e] = (r < n && c < nrhs) ? T(bb(r, c)) : T(0);
    }
    __syncthreads();
    // right-hand-side block as MFMA B operand: rows 16 blk + 4 q + l / 16, column l % 16
    auto rhs_op = [&](int blk, int q) -> T {
        int const c = lane & 15;
        return c < kWgRhs ? sR[(16 * blk + 4 * q + (lane >> 4)) * kWgRhs + c] : T(0);
    };
    auto rhs_load = [&](int blk) {
        acc_t v;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const c = acc_col(lane);
            v[r] = c < kWgRhs ? sR[(16 * blk + Mf::row(lane, r)) * kWgRhs + c] : T(0);
        }
        return v;
    };
    auto rhs_store = [&](int blk, acc_t v) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const c = acc_col(lane);
            if (c < kWgRhs) sR[(16 * blk + Mf::row(lane, r)) * kWgRhs + c] = v[r];
        }
    };
    auto solve = [&]() {
        if constexpr (LU) {
            if (tid < kWgRhs)
                for (int k = 0; k < n; ++k) {
                    int const p = sPiv[k];
                    if (p != k) {
                        T const t = sR[k * kWgRhs + tid];
                        sR[k * kWgRhs + tid] = sR[p * kWgRhs + tid];
                        sR[p * kWgRhs + tid] = t;
                    }
                }
            __syncthreads();
        }
        for (int bk = 0; bk < NT; ++bk) {  // L y = r
            if (w == 0) {
                acc_t acc{};
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = Mf::mma(op_read(sW1[bk], q, lane), rhs_op(bk, q), acc);
                rhs_store(bk, acc);
            }
            __syncthreads();
            {  // r_i -= L_ib y_b, up to four tiles per wavefront, their loads issued together
                T fv[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const i = bk + 1 + w + kWgWaves * u;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        fv[u][q] = i < NT ? fget(16 * i + (lane & 15), 16 * bk + 4 * q + (lane >> 4)) : T(0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const i = bk + 1 + w + kWgWaves * u;
                    if (i >= NT) continue;
                    acc_t acc = rhs_load(i);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc = Mf::mma(-fv[u][q], rhs_op(bk, q), acc);
                    rhs_store(i, acc);
                }
            }
            __syncthreads();
        }
        for (int bk = NT - 1; bk >= 0; --bk) {  // U x = y (Cholesky: U = L^T)
            if (w == 0) {
                acc_t acc{};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    T const av = LU ? op_read(sW2[bk], q, lane)
                                    : sW1[bk][(lane & 15) * kTileStride + 4 * q + (lane >> 4)];
                    acc = Mf::mma(av, rhs_op(bk, q), acc);
                }
                rhs_store(bk, acc);
            }
            __syncthreads();
            {  // y_i -= U_ib x_b, i < b
                T fv[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const i = w + kWgWaves * u;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        int const m = 16 * i + (lane & 15), k = 16 * bk + 4 * q + (lane >> 4);
                        fv[u][q] = i < bk ? (LU ? fget(m, k) : fget(k, m)) : T(0);
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const i = w + kWgWaves * u;
                    if (i >= bk) continue;
                    acc_t acc = rhs_load(i);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc = Mf::mma(-fv[u][q], rhs_op(bk, q), acc);
                    rhs_store(i, acc);
                }
            }
            __syncthreads();
        }
    };
    auto accumulate = [&]() {  // x += d
        for (int e = tid; e < npad * kWgRhs; e += kWgThreads) sX[e] += double(sR[e]);
        __syncthreads();
    };
    // r = b - A x in FP64 into sR; returns the stopping test.  The first call
    // also forms ||A||_inf from the same loads (row sums of |A|).
    auto residual = [&](bool first) -> bool {
        if (tid < 2 * kWgRhs) sMax[tid] = 0ull;
        __syncthreads();
        for (int i = w; i < NT; i += kWgWaves) {
            Md::acc_t acc{};
            int const m = 16 * i + (lane & 15);
            double rs = 0.0;
            for (int j0 = 0; j0 < NT; j0 += 4) {
                double av[4][4], xv[4][4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        int const k = 16 * (j0 + u) + 4 * q + (lane >> 4);
                        bool const in = j0 + u < NT;
                        av[u][q] = (in && m < n && k < n) ? asym(m, k) : 0.0;
                        xv[u][q] = (in && (lane & 15) < kWgRhs) ? sX[k * kWgRhs + (lane & 15)] : 0.0;
                    }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        acc = Md::mma(av[u][q], xv[u][q], acc);
                        if (first) rs += __builtin_fabs(av[u][q]);
                    }
            }
            if (first) {
                rs += __shfl_xor(rs, 16);
                rs += __shfl_xor(rs, 32);
                if (lane < 16) atomicMax(&sAnrmBits, abs_bits(rs));
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const row = 16 * i + Md::row(lane, r), c = acc_col(lane);
                if (c < kWgRhs) {
                    bool const live = row < n && c < nrhs;
                    double const rv = live ? bb(row, c) - acc[r] : 0.0;
                    sR[row * kWgRhs + c] = T(rv);
                    if (live) {
                        atomicMax(&sMax[c], abs_bits(rv));
                        atomicMax(&sMax[kWgRhs + c], abs_bits(sX[row * kWgRhs + c]));
                    }
                }
            }
        }
        __syncthreads();
        double const cte = __builtin_bit_cast(double, sAnrmBits) *
                           (g.tol > 0.0 ? g.tol : 1.1102230246251565e-16 * __builtin_sqrt(double(n)));
        bool ok = true;
        for (int c = 0; c < nrhs; ++c) {
            double const rn = __builtin_bit_cast(double, sMax[c]), xn = __builtin_bit_cast(double, sMax[kWgRhs + c]);
            ok = ok && (rn <= xn * cte);
        }
        return ok;
    };

    solve();
    accumulate();
    if constexpr (kMixed) {
        int it = 0, status;
        for (;;) {
            if (residual(it == 0)) {
                status = it;
                break;
            }
            if (it == g.itmax) {
                status = -2;
                break;
            }
            ++it;
            solve();
            accumulate();
        }
        if (tid == 0) g.iter[b] = status;
        if (status < 0) return;
    }
    Ref<double> x(g.X, b);
    for (int e = tid; e < n * nrhs; e += kWgThreads) {
        int const r = e % n, c = e / n;
        x(r, c) = sX[r * kWgRhs + c];
    }
}

template <typename T, bool LU, bool BF16>
int launch_wg(hipStream_t s, WgArgs const& g, Mat<T> F, int64_t batch) {
    if (batch == 0) return 0;
    hipLaunchKernelGGL((wg_solve_kernel<T, LU, BF16>), dim3(unsigned(batch)), dim3(kWgThreads), 0, s, g, F);
    return check_launch("wg_solve_kernel");
}

WgArgs make_args(Mat<double> A, Mat<double> B, Mat<double> X, int nrhs, int itmax, double tol, int32_t* info,
                 int32_t* iter, const int32_t* only_if, int64_t* ipiv, int64_t ipiv_stride) {
    WgArgs g;
    g.A = A, g.B = B, g.X = X;
    g.n = A.rows, g.nrhs = nrhs, g.itmax = itmax, g.tol = tol;
    g.info = info, g.iter = iter, g.only_if = only_if;
    g.ipiv = ipiv, g.ipiv_stride = ipiv_stride;
    return g;
}

}  // namespace

int64_t wg_work_elems(int n) {
    int64_t const npad = 16 * ((int64_t(n) + 15) / 16);
    return (npad * npad + 63) & ~int64_t(63);
}

bool wg_covers(int n, int nrhs) { return n >= 1 && n <= 16 * kWgMaxNT && nrhs >= 0 && nrhs <= kWgRhs; }

// Mixed precision: factor in FP32 (bf16: BF16 products in the trailing
// updates) into `work` (wg_work_elems(n) floats per instance), refine in FP64.
int launch_wg_mixed(hipStream_t s, int lu, int bf16, int itmax, double tol, Mat<double> A, int64_t* ipiv,
                    int64_t ipiv_stride, Mat<double> B, Mat<double> X, int nrhs, float* work, int64_t batch,
                    int32_t* info, int32_t* iter) {
    Mat<float> F;
    F.data = work;
    F.stride = wg_work_elems(A.rows);
    F.rows = A.rows, F.cols = A.cols;
    F.spacing = 16 * ((A.rows + 15) / 16);
    F.panel = 0, F.trans = 0;
    WgArgs const g = make_args(A, B, X, nrhs, itmax, tol, info, iter, nullptr, ipiv, ipiv_stride);
    if (lu) return bf16 ? launch_wg<float, true, true>(s, g, F, batch) : launch_wg<float, true, false>(s, g, F, batch);
    return bf16 ? launch_wg<float, false, true>(s, g, F, batch) : launch_wg<float, false, false>(s, g, F, batch);
}

}  // namespace blast_hip
