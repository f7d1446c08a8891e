// Mixed-precision batched Cholesky solve with FP64 iterative refinement,
// n <= 128, nrhs <= 4 (the LAPACK dsposv contract: A is read, X written;
// A is overwritten by its FP64 factor only for instances that fall back).
//
// One wavefront per instance, two per SIMD (the FP32 factor lives in
// registers, LDS holds one block column and the vectors, about 17 KB):
//   1. factor A (rounded to FP32) with the right-looking tile algorithm of
//      posv_tiles.hip on v_mfma_f32_16x16x4f32 (or, with BF16, trailing
//      updates on v_mfma_f32_16x16x16bf16_1k: BF16 products, FP32
//      accumulation); the off-diagonal factor tiles stay in the accumulator
//      registers, the diagonal ones are replaced by W = L_kk^{-1} in LDS;
//   2. x = 0, r = b; repeat: d = L^{-T} L^{-1} r in FP32 on MFMA, x += d in FP64,
//      r = b - A x in FP64 with A read again from HBM (L2/MALL resident),
//      until max|r| <= max|x| * ||A||_inf * eps * sqrt(n) per right-hand side
//      (LAPACK dsposv's test, BWDMAX = 1; or tol instead of eps*sqrt(n)
//      when tol > 0) or itmax refinements.
// iter[b] = number of refinement steps (0 = the first solve converged),
// -1 = the FP32 factorisation broke down, -2 = no convergence; negative
// instances are then solved by the FP64 kernel (posv_tiles.hip) in place.
// This mixed scheme is not in BLAST; its result is checked against the FP64
// oracle by backward error.

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {

// max that keeps a NaN operand (fmax drops it): a NaN residual, solution or
// norm must fail the convergence test, not pass it
__device__ __forceinline__ double max_keep_nan(double a, double b) { return (b > a || b != b) ? b : a; }

template <bool PANEL>
__device__ __forceinline__ double a_sym(Ref<double> const& a, int row, int col) {
    return row >= col ? at<PANEL>(a, row, col) : at<PANEL>(a, col, row);
}

// FULL: n == 16 NT and panel storage, tiles loaded without bounds tests
// (load_tile_full).
constexpr int kSolveGroup = 4;

template <int NT, bool PANEL, bool BF16, bool FULL>
__global__ __launch_bounds__(64, 2) void sposv_ir_kernel(Mat<double> A, Mat<double> B, Mat<double> X, int n,
                                                         int nrhs, int itmax, double tol, int32_t* info,
                                                         int32_t* iter) {
    constexpr int NTILES = NT * (NT + 1) / 2;
    constexpr int NV = NT * 16 * kMaxRhs;
    __shared__ __align__(16) float sP[NT][kTileElems];  // block column kb (operand layout); sP[kb]: the diagonal tile
    __shared__ float sW[kSolveGroup][kTileElems];  // factor tiles staged for the forward solve
    __shared__ double sX[NV];             // x, [row][rhs]
    __shared__ float sV[NV];              // FP32 solve vector, [row][rhs]

    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x;
    Ref<double> a(A, b), bb(B, b);
    using M = Mfma<float>;
    using acc_t = M::acc_t;

    // ---------------------------------------------------------- factor
    // acc holds the FP32 factor tiles L_ij below the diagonal for the whole
    // kernel; sP[k] ends up holding W_k = L_kk^{-1}.
    acc_t acc[NTILES];
    int bad = 0;
    static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
        constexpr int i = decltype(I)::value;
        static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            acc[tri(i, j)] = FULL ? load_tile_full<PANEL, float>(a, i, j, lane) : load_tile<PANEL, float>(a, n, i, j, lane);
        });
    });
    static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
        constexpr int kb = decltype(KB)::value;
        static_for<kb, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            acc_to_lds<float>(sP[i], acc[tri(i, kb)], lane);
        });
        // W_kb = L_kk^{-1} replaces the diagonal tile in sP[kb] (the factor
        // L_kk itself is not needed); sP[kb] is not written again
        int const bk = diag_factor<float>(sP[kb], sP[kb], lane);
        if (bk && bad == 0 && 16 * kb + bk <= n) bad = 16 * kb + bk;
        static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            acc_t z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q) z = M::mma(op_read(sP[i], q, lane), op_read(sP[kb], q, lane), z);
            acc[tri(i, kb)] = z;
            acc_to_lds<float>(sP[i], z, lane);
        });
        if constexpr (BF16) {
            s4 f[NT];
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                f[i] = op_read_bf16(sP[i], lane);
            });
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                static_for<kb + 1, i + 1>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value;
                    s4 neg = f[i];
#pragma unroll
                    for (int e = 0; e < 4; ++e) neg[e] = short(neg[e] ^ short(0x8000));
                    acc[tri(i, j)] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(neg, f[j], acc[tri(i, j)], 0, 0, 0);
                });
            });
        } else {
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                static_for<kb + 1, i + 1>([&](auto J) __attribute__((always_inline)) {
                    constexpr int j = decltype(J)::value;
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        acc[tri(i, j)] = M::mma(-op_read(sP[i], q, lane), op_read(sP[j], q, lane), acc[tri(i, j)]);
                });
            });
        }
    });
    if (bad) {
        if (lane == 0) iter[b] = -1;
        return;
    }

    // v <- L^{-T} L^{-1} v in FP32 on MFMA, the right-hand sides as the B
    // operand (columns 0..kMaxRhs-1 of the 16).  Forward: y_k = W_k v_k,
    // v_i -= L_ik y_k (tiles staged through sW into the A-operand layout).
    // Backward: x_k = W_k^T v_k, v_j -= L_kj^T x_k, straight from the
    // registers: register r of an accumulator tile M is the A operand of M^T
    // for K-chunk r with k = M::row(lane, r).
    int const cn = lane & 15;
    auto vb = [&](int blk, int k) -> float { return cn < kMaxRhs ? sV[(16 * blk + k) * kMaxRhs + cn] : 0.f; };
    auto vload = [&](int blk) {
        acc_t v;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = cn < kMaxRhs ? sV[(16 * blk + M::row(lane, r)) * kMaxRhs + cn] : 0.f;
        return v;
    };
    auto vstore = [&](int blk, acc_t v) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (cn < kMaxRhs) sV[(16 * blk + M::row(lane, r)) * kMaxRhs + cn] = v[r];
    };
    // Tiles are processed kSolveGroup at a time, their MFMA chains
    // interleaved (one chain alone waits on each MFMA's result).
    auto solve = [&]() {
        static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
            constexpr int kb = decltype(KB)::value;
            acc_t y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q) y = M::mma(op_read(sP[kb], q, lane), vb(kb, 4 * q + (lane >> 4)), y);
            vstore(kb, y);
            static_for<0, (NT - 1 - kb + kSolveGroup - 1) / kSolveGroup>([&](auto GI) __attribute__((always_inline)) {
                constexpr int i0 = kb + 1 + kSolveGroup * decltype(GI)::value;
                acc_t v[kSolveGroup];
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (i0 + t < NT) {
                        acc_to_lds<float>(sW[t], acc[tri(i0 + t, kb)], lane);
                        v[t] = vload(i0 + t);
                    }
                });
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float const yq = vb(kb, 4 * q + (lane >> 4));
                    static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                        constexpr int t = decltype(T)::value;
                        if constexpr (i0 + t < NT) v[t] = M::mma(-op_read(sW[t], q, lane), yq, v[t]);
                    });
                }
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (i0 + t < NT) vstore(i0 + t, v[t]);
                });
            });
        });
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            acc_t x = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q)  // A operand W^T: W[row(l, q)][l % 16]
                x = M::mma(sP[kb][cn * kTileStride + M::row(lane, q)], vb(kb, M::row(lane, q)), x);
            vstore(kb, x);
            static_for<0, (kb + kSolveGroup - 1) / kSolveGroup>([&](auto GI) __attribute__((always_inline)) {
                constexpr int j0 = kSolveGroup * decltype(GI)::value;
                acc_t v[kSolveGroup];
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (j0 + t < kb) v[t] = vload(j0 + t);
                });
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float const xq = vb(kb, M::row(lane, q));
                    static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                        constexpr int t = decltype(T)::value;
                        if constexpr (j0 + t < kb) v[t] = M::mma(-acc[tri(kb, j0 + t)][q], xq, v[t]);
                    });
                }
                static_for<0, kSolveGroup>([&](auto T) __attribute__((always_inline)) {
                    constexpr int t = decltype(T)::value;
                    if constexpr (j0 + t < kb) vstore(j0 + t, v[t]);
                });
            });
        });
    };

    // ---------------------------------------------------------- refine
    // Residual lanes: one row each, 64 rows per pass.
    for (int e = lane; e < NV; e += 64) {
        int const row = e >> 2, cc = e & 3;
        sX[e] = 0.0;
        sV[e] = (row < n && cc < nrhs) ? float(bb(row, cc)) : 0.f;
    }
    double anrm = 0.0;  // ||A||_inf, accumulated during the first residual
    int it = 0;
    bool converged = false;
    for (;; ++it) {
        solve();  // d = L^{-T} L^{-1} v
        for (int e = lane; e < NV; e += 64) sX[e] += double(sV[e]);
        double rmax[kMaxRhs] = {0.0, 0.0, 0.0, 0.0}, xmax[kMaxRhs] = {0.0, 0.0, 0.0, 0.0};
        if constexpr (FULL) {
            // r = b - A x in FP64, lane = row, A symmetric from its lower
            // triangle in panel storage: element (row, col) at
            // (row/4)*spacing + row%4 + 4 col for col <= row, its mirror
            // (col/4)*spacing + col%4 + 4 row above.  The column is the same
            // in every lane, so the mirror's panel offset is scalar
            // arithmetic; consecutive columns reuse each lane's cache lines.
            // x is broadcast from LDS.
            for (int h = 0; h < NT * 16; h += 64) {
                int const row = h + lane;
                bool const act = row < NT * 16;
                int const rr = act ? row : NT * 16 - 1;  // idle lanes read a valid row
                int64_t const lowbase = int64_t(rr >> 2) * A.spacing + (rr & 3);
                double r[kMaxRhs];
#pragma unroll
                for (int cc = 0; cc < kMaxRhs; ++cc) r[cc] = cc < nrhs ? bb(rr, cc) : 0.0;
                double asum = 0.0;
                // one 16-column batch at a time (unrolling the batches of the
                // smaller sizes holds several batches' loads and spills)
#pragma unroll 1
                for (int c0 = 0; c0 < NT * 16; c0 += 16) {
                    double av[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        int const col = c0 + k;
                        int64_t const up = int64_t(col >> 2) * A.spacing + (col & 3);  // scalar
                        av[k] = a.p[rr >= col ? lowbase + 4 * col : up + 4 * rr];
                    }
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        asum += __builtin_fabs(av[k]);
#pragma unroll
                        for (int cc = 0; cc < kMaxRhs; ++cc)
                            r[cc] = fma_t(-av[k], sX[(c0 + k) * kMaxRhs + cc], r[cc]);
                    }
                }
                if (act) {
                    anrm = max_keep_nan(anrm, asum);
#pragma unroll
                    for (int cc = 0; cc < kMaxRhs; ++cc) {
                        sV[row * kMaxRhs + cc] = float(r[cc]);
                        rmax[cc] = max_keep_nan(rmax[cc], __builtin_fabs(r[cc]));
                        xmax[cc] = max_keep_nan(xmax[cc], __builtin_fabs(sX[row * kMaxRhs + cc]));
                    }
                }
            }
        } else {
            // r = b - A x in FP64, A (symmetric, lower triangle stored) read in
            // 16-column batches so the loads overlap; v = float(r)
            for (int h = 0; h < NT * 16; h += 64) {
                int const row = h + lane;
                double r[kMaxRhs];
#pragma unroll
                for (int cc = 0; cc < kMaxRhs; ++cc) r[cc] = (row < n && cc < nrhs) ? bb(row, cc) : 0.0;
                double asum = 0.0;
                // one 16-column batch at a time (unrolling the batches of the
                // smaller sizes holds several batches' loads and spills)
#pragma unroll 1
                for (int c0 = 0; c0 < NT * 16; c0 += 16) {
                    double av[16];
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        int const col = c0 + k;
                        bool const in = row < n && col < n;
                        int const lo = row >= col ? row : col, hi = row >= col ? col : row;
                        int64_t const off = in ? (PANEL ? int64_t(lo >> 2) * A.spacing + (lo & 3) + 4 * hi : A.off(lo, hi)) : 0;
                        double const v = a.p[off];
                        av[k] = in ? v : 0.0;
                    }
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        asum += __builtin_fabs(av[k]);
#pragma unroll
                        for (int cc = 0; cc < kMaxRhs; ++cc)
                            r[cc] = __builtin_fma(-av[k], sX[(c0 + k) * kMaxRhs + cc], r[cc]);
                    }
                }
                anrm = max_keep_nan(anrm, asum);
                if (row < NT * 16)
#pragma unroll
                    for (int cc = 0; cc < kMaxRhs; ++cc) {
                        sV[row * kMaxRhs + cc] = float(r[cc]);
                        rmax[cc] = max_keep_nan(rmax[cc], __builtin_fabs(r[cc]));
                        xmax[cc] = max_keep_nan(xmax[cc], __builtin_fabs(sX[row * kMaxRhs + cc]));
                    }
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) anrm = max_keep_nan(anrm, __shfl_xor(anrm, o));
        // LAPACK dsposv: ||r|| <= ||x|| ||A|| eps sqrt(n); with tol > 0:
        // ||r|| <= tol ||x|| ||A|| (infinity norms, per right-hand side).
        double const cte = anrm * (tol > 0.0 ? tol : 1.1102230246251565e-16 * __builtin_sqrt(double(n)));
        bool ok = true;
#pragma unroll
        for (int cc = 0; cc < kMaxRhs; ++cc) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                rmax[cc] = max_keep_nan(rmax[cc], __shfl_xor(rmax[cc], o));
                xmax[cc] = max_keep_nan(xmax[cc], __shfl_xor(xmax[cc], o));
            }
            ok = ok && (rmax[cc] <= xmax[cc] * cte);
        }
        if (ok) {
            converged = true;
            break;
        }
        if (it >= itmax) break;
    }
    if (!converged) {
        if (lane == 0) iter[b] = -2;
        return;
    }
    Ref<double> xx(X, b);
    for (int e = lane; e < n * nrhs; e += 64) {
        int const row = e % n, cc = e / n;
        xx(row, cc) = sX[row * kMaxRhs + cc];
    }
    if (lane == 0) {
        iter[b] = it;
        if (info) info[b] = 0;
    }
}

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t count) {
    int64_t const i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

int launch_fill_i32(hipStream_t s, int32_t* p, int32_t v, int64_t count) {
    if (count == 0) return 0;
    hipLaunchKernelGGL(fill_i32_kernel, dim3(unsigned((count + 255) / 256)), dim3(256), 0, s, p, v, count);
    return check_launch("fill_i32_kernel");
}

template <int NT, bool PANEL, bool BF16>
static void launch_ir_full(hipStream_t s, Mat<double> A, Mat<double> B, Mat<double> X, int n, int nrhs, int itmax,
                           double tol, int64_t batch, int32_t* info, int32_t* iter) {
    if (PANEL && n == 16 * NT)  // (dense storage: the generic addressing would spill)
        hipLaunchKernelGGL((sposv_ir_kernel<NT, PANEL, BF16, PANEL>), dim3(unsigned(batch)), dim3(64), 0, s, A, B, X, n,
                           nrhs, itmax, tol, info, iter);
    else
        hipLaunchKernelGGL((sposv_ir_kernel<NT, PANEL, BF16, false>), dim3(unsigned(batch)), dim3(64), 0, s, A, B, X,
                           n, nrhs, itmax, tol, info, iter);
}

template <int NT, bool BF16>
static int launch_ir(hipStream_t s, Mat<double> A, Mat<double> B, Mat<double> X, int n, int nrhs, int itmax,
                     double tol, int64_t batch, int32_t* info, int32_t* iter) {
    if (A.panel && !A.trans)
        launch_ir_full<NT, true, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
    else
        launch_ir_full<NT, false, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
    return check_launch("sposv_ir_kernel");
}

template <bool BF16>
static int launch_ir_n(hipStream_t s, Mat<double> A, Mat<double> B, Mat<double> X, int itmax, double tol,
                       int64_t batch, int32_t* info, int32_t* iter) {
    int const n = A.rows, nrhs = X.cols;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_ir<1, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 2: return launch_ir<2, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 3: return launch_ir<3, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 4: return launch_ir<4, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 5: return launch_ir<5, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 6: return launch_ir<6, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        case 7: return launch_ir<7, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
        default: return launch_ir<8, BF16>(s, A, B, X, n, nrhs, itmax, tol, batch, info, iter);
    }
}

int launch_sposv_ir(hipStream_t s, int bf16, Mat<double> A, Mat<double> B, Mat<double> X, int itmax, double tol,
                    int64_t batch, int32_t* info, int32_t* iter) {
    if (A.rows != A.cols || A.rows > 128 || X.cols > kMaxRhs || A.rows == 0 || X.cols == 0) return 1;
    if (batch == 0) return 0;
    return bf16 ? launch_ir_n<true>(s, A, B, X, itmax, tol, batch, info, iter)
                : launch_ir_n<false>(s, A, B, X, itmax, tol, batch, info, iter);
}

}  // namespace blast_hip
