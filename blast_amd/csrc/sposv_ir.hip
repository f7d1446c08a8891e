// Mixed-precision batched Cholesky solve with FP64 iterative refinement,
// n <= 128, nrhs <= 4 (the LAPACK dsposv contract: A is read, X written;
// A is overwritten by its FP64 factor only for instances that fall back).
//
// One wavefront per instance:
//   1. factor A (rounded to FP32) with the right-looking tile algorithm of
//      posv_tiles.hip on v_mfma_f32_16x16x4f32 (or, with BF16, trailing
//      updates on v_mfma_f32_16x16x16bf16_1k: BF16 products, FP32
//      accumulation); the finished FP32 tiles stay in LDS, each diagonal
//      tile replaced by W = L_kk^{-1};
//   2. x = 0, r = b; repeat: d = L^{-T} L^{-1} r in FP32, x += d in FP64,
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

template <bool PANEL>
__device__ __forceinline__ double a_sym(Ref<double> const& a, int row, int col) {
    return row >= col ? at<PANEL>(a, row, col) : at<PANEL>(a, col, row);
}

template <int NT, bool PANEL, bool BF16>
__global__ __launch_bounds__(64) void sposv_ir_kernel(Mat<double> A, Mat<double> B, Mat<double> X, int n, int nrhs,
                                                      int itmax, double tol, int32_t* info, int32_t* iter) {
    constexpr int NTILES = NT * (NT + 1) / 2;
    constexpr int NV = NT * 16 * kMaxRhs;
    __shared__ float sL[NTILES][kTileElems];  // FP32 factor; tile (k,k) holds W_k = L_kk^{-1}
    __shared__ double sX[NV];                 // x, [row][rhs]
    __shared__ float sV[NV];                  // FP32 solve vector, [row][rhs]
    __shared__ float sT[16 * kMaxRhs];        // one block of the substitution

    int64_t const b = blockIdx.x;
    int const lane = threadIdx.x;
    Ref<double> a(A, b), bb(B, b);
    using M = Mfma<float>;
    using acc_t = M::acc_t;

    // ---------------------------------------------------------- factor
    int bad = 0;
    {
        acc_t acc[NTILES];
        static_for<0, NT>([&](auto I) __attribute__((always_inline)) {
            constexpr int i = decltype(I)::value;
            static_for<0, i + 1>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                acc[tri(i, j)] = load_tile<PANEL, float>(a, n, i, j, lane);
            });
        });
        static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
            constexpr int kb = decltype(KB)::value;
            static_for<kb, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                acc_to_lds<float>(sL[tri(i, kb)], acc[tri(i, kb)], lane);
            });
            float* const W = sL[tri(kb, kb)];
            int const bk = diag_factor<float>(W, W, lane);  // second write (W) wins
            if (bk && bad == 0 && 16 * kb + bk <= n) bad = 16 * kb + bk;
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                acc_t z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int q = 0; q < 4; ++q) z = M::mma(op_read(sL[tri(i, kb)], q, lane), op_read(W, q, lane), z);
                acc_to_lds<float>(sL[tri(i, kb)], z, lane);
            });
            if constexpr (BF16) {
                s4 f[NT];
                static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                    constexpr int i = decltype(I)::value;
                    s4 v = op_read_bf16(sL[tri(i, kb)], lane);
                    f[i] = v;
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
                float f[NT][4];
                static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                    constexpr int i = decltype(I)::value;
#pragma unroll
                    for (int q = 0; q < 4; ++q) f[i][q] = op_read(sL[tri(i, kb)], q, lane);
                });
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                        constexpr int i = decltype(I)::value;
                        static_for<kb + 1, i + 1>([&](auto J) __attribute__((always_inline)) {
                            constexpr int j = decltype(J)::value;
                            acc[tri(i, j)] = M::mma(-f[i][q], f[j][q], acc[tri(i, j)]);
                        });
                    });
            }
        });
    }
    if (bad) {
        if (lane == 0) iter[b] = -1;
        return;
    }

    // ---------------------------------------------------------- refine
    // Solve lanes: (m, c) = (row within a 16-row block, right-hand side).
    // Residual lanes: one row each, 64 rows per pass.
    int const m = lane >> 2, c = lane & 3;
    for (int e = lane; e < NV; e += 64) {
        int const row = e >> 2, cc = e & 3;
        sX[e] = 0.0;
        sV[e] = (row < n && cc < nrhs) ? float(bb(row, cc)) : 0.f;
    }
    double anrm = 0.0;  // ||A||_inf, accumulated during the first residual
    int it = 0;
    bool converged = false;
    for (;; ++it) {
        // d = L^{-T} L^{-1} v in FP32, by 16-row blocks
        static_for<0, NT>([&](auto KB) __attribute__((always_inline)) {
            constexpr int kb = decltype(KB)::value;
            float t = sV[(16 * kb + m) * kMaxRhs + c];
            static_for<0, kb>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                const float* Lt = sL[tri(kb, j)];
#pragma unroll
                for (int k = 0; k < 16; ++k) t = __builtin_fmaf(-Lt[k * kTileStride + m], sV[(16 * j + k) * kMaxRhs + c], t);
            });
            sT[m * kMaxRhs + c] = t;
            const float* W = sL[tri(kb, kb)];
            float y = 0.f;
#pragma unroll
            for (int k = 0; k < 16; ++k) y = __builtin_fmaf(W[k * kTileStride + m], sT[k * kMaxRhs + c], y);
            sV[(16 * kb + m) * kMaxRhs + c] = y;
        });
        static_for<0, NT>([&](auto KK) __attribute__((always_inline)) {
            constexpr int kb = NT - 1 - decltype(KK)::value;
            float t = sV[(16 * kb + m) * kMaxRhs + c];
            static_for<kb + 1, NT>([&](auto I) __attribute__((always_inline)) {
                constexpr int i = decltype(I)::value;
                const float* Lt = sL[tri(i, kb)];
#pragma unroll
                for (int k = 0; k < 16; ++k) t = __builtin_fmaf(-Lt[m * kTileStride + k], sV[(16 * i + k) * kMaxRhs + c], t);
            });
            sT[m * kMaxRhs + c] = t;
            const float* W = sL[tri(kb, kb)];
            float x = 0.f;
#pragma unroll
            for (int k = 0; k < 16; ++k) x = __builtin_fmaf(W[m * kTileStride + k], sT[k * kMaxRhs + c], x);
            sV[(16 * kb + m) * kMaxRhs + c] = x;
        });
        for (int e = lane; e < NV; e += 64) sX[e] += double(sV[e]);
        // r = b - A x in FP64, A (symmetric, lower triangle stored) read in
        // 16-column batches so the loads overlap; v = float(r)
        double rmax[kMaxRhs] = {0.0, 0.0, 0.0, 0.0}, xmax[kMaxRhs] = {0.0, 0.0, 0.0, 0.0};
        for (int h = 0; h < NT * 16; h += 64) {
            int const row = h + lane;
            double r[kMaxRhs];
#pragma unroll
            for (int cc = 0; cc < kMaxRhs; ++cc) r[cc] = (row < n && cc < nrhs) ? bb(row, cc) : 0.0;
            double asum = 0.0;
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
            anrm = __builtin_fmax(anrm, asum);
            if (row < NT * 16)
#pragma unroll
                for (int cc = 0; cc < kMaxRhs; ++cc) {
                    sV[row * kMaxRhs + cc] = float(r[cc]);
                    rmax[cc] = __builtin_fmax(rmax[cc], __builtin_fabs(r[cc]));
                    xmax[cc] = __builtin_fmax(xmax[cc], __builtin_fabs(sX[row * kMaxRhs + cc]));
                }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) anrm = __builtin_fmax(anrm, __shfl_xor(anrm, o));
        // LAPACK dsposv: ||r|| <= ||x|| ||A|| eps sqrt(n); with tol > 0:
        // ||r|| <= tol ||x|| ||A|| (infinity norms, per right-hand side).
        double const cte = anrm * (tol > 0.0 ? tol : 1.1102230246251565e-16 * __builtin_sqrt(double(n)));
        bool ok = true;
#pragma unroll
        for (int cc = 0; cc < kMaxRhs; ++cc) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                rmax[cc] = __builtin_fmax(rmax[cc], __shfl_xor(rmax[cc], o));
                xmax[cc] = __builtin_fmax(xmax[cc], __shfl_xor(xmax[cc], o));
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

template <int NT, bool BF16>
static int launch_ir(hipStream_t s, Mat<double> A, Mat<double> B, Mat<double> X, int n, int nrhs, int itmax,
                     double tol, int64_t batch, int32_t* info, int32_t* iter) {
    if (A.panel && !A.trans)
        hipLaunchKernelGGL((sposv_ir_kernel<NT, true, BF16>), dim3(unsigned(batch)), dim3(64), 0, s, A, B, X, n, nrhs,
                           itmax, tol, info, iter);
    else
        hipLaunchKernelGGL((sposv_ir_kernel<NT, false, BF16>), dim3(unsigned(batch)), dim3(64), 0, s, A, B, X, n,
                           nrhs, itmax, tol, info, iter);
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
