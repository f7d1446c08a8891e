// Reference-order kernels: one workgroup per instance, column sweeps in
// global memory (L1/L2 resident), every element computed with the same
// operation sequence as BLAST's CPU kernels:
//   gemm   acc = fma(a_ik, b_kj, acc), k increasing; acc *= alpha; fma(beta, c, acc)
//          (register_matrix/Gemm.hpp:71-92)
//   syrk   acc = beta*c; acc = fma(alpha*a_ik, a_jk, acc)   (dense/Syrk.hpp:15-145)
//   trmm   acc = fma(alpha*a_ik, b_kj, acc) over the triangle (algorithm/Trmm.hpp)
//   trsm   v = alpha*b; v = fma(-a_ik, x_k, v) in sweep order; v / a_ii
//          (reference/Trsm.hpp:44-127, dense/Trsm.hpp:17-66)
//   potrf  right-looking sweep; each element sees its subtractions in the
//          same increasing-column order as BLAST's left-looking tiles, with
//          BLAST's sqrt / divide pattern (dense/Potrf.hpp:20-106,
//          RegisterMatrix.hpp:709-739)
//   getrf  BLAST's NB = TileSize blocked algorithm step by step
//          (dense/Getrf.hpp:52-162, Getf2.hpp:52-77, Iamax.hpp:82-95)
// Division and square root are IEEE correctly rounded (HIP's default double
// and float '/' and sqrt), so results are bit-identical to BLAST's.
// These kernels are the generic path for every size, layout and transpose
// combination; the MFMA kernels cover the hot shapes.

#include "common.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {

constexpr int kThreads = 256;

template <typename T>
__global__ __launch_bounds__(kThreads) void gemm_ref_kernel(T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C,
                                                             Mat<T> D) {
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b), bb(B, b), c(C, b), d(D, b);
    int const M = D.rows, N = D.cols, K = A.cols;
    for (int e = threadIdx.x; e < M * N; e += kThreads) {
        int const i = e % M, j = e / M;
        T acc = T(0);
        for (int k = 0; k < K; ++k) acc = fma_t(a(i, k), bb(k, j), acc);
        acc = acc * alpha;
        d(i, j) = fma_t(beta, c(i, j), acc);
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void syrk_ref_kernel(T alpha, Mat<T> A, T beta, Mat<T> C, Mat<T> D) {
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b), c(C, b), d(D, b);
    int const M = D.rows, K = A.cols;
    for (int e = threadIdx.x; e < M * M; e += kThreads) {
        int const i = e % M, j = e / M;
        if (i < j) continue;
        T acc = beta * c(i, j);
        for (int k = 0; k < K; ++k) acc = fma_t(alpha * a(i, k), a(j, k), acc);
        d(i, j) = acc;
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void trmm_ref_kernel(int side_right, T alpha, Mat<T> A, Mat<T> B,
                                                             Mat<T> C) {
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b), bb(B, b), c(C, b);
    int const M = C.rows, N = C.cols;
    for (int e = threadIdx.x; e < M * N; e += kThreads) {
        int const i = e % M, j = e / M;
        T acc = T(0);
        if (!side_right)
            for (int k = i; k < M; ++k) acc = fma_t(alpha * a(i, k), bb(k, j), acc);
        else
            for (int k = j; k < N; ++k) acc = fma_t(alpha * bb(i, k), a(k, j), acc);
        c(i, j) = acc;
    }
}

// Left: A X = alpha B.  Right: X A = alpha B.  X may alias B (B is read once,
// before X is written).  BLAST's element recurrence (reference/Trsm.hpp:44-127)
// subtracts the known terms in increasing k.  For left/lower and right/upper
// those terms become known in increasing k, so a column sweep (parallel over
// the other index) keeps the order.  For left/upper and right/lower they
// become known in decreasing k, so each right-hand side (left) or row (right)
// runs the recurrence as a sequential dot product in one thread.
template <typename T>
__global__ __launch_bounds__(kThreads) void trsm_ref_kernel(int side_right, int upper, int unit, T alpha,
                                                             Mat<T> A, Mat<T> B, Mat<T> X) {
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b), bb(B, b), x(X, b);
    int const M = X.rows, N = X.cols;
    bool const sweep = side_right ? upper : !upper;
    if (!sweep) {
        if (!side_right) {
            for (int j = threadIdx.x; j < N; j += kThreads)
                for (int i = M - 1; i >= 0; --i) {
                    T v = alpha * bb(i, j);
                    for (int k = i + 1; k < M; ++k) v = fma_t(-a(i, k), x(k, j), v);
                    x(i, j) = unit ? v : v / a(i, i);
                }
        } else {
            for (int i = threadIdx.x; i < M; i += kThreads)
                for (int j = N - 1; j >= 0; --j) {
                    T v = alpha * bb(i, j);
                    for (int k = j + 1; k < N; ++k) v = fma_t(-x(i, k), a(k, j), v);
                    x(i, j) = unit ? v : v / a(j, j);
                }
        }
        return;
    }
    for (int e = threadIdx.x; e < M * N; e += kThreads) {
        int const i = e % M, j = e / M;
        x(i, j) = alpha * bb(i, j);
    }
    __syncthreads();
    if (!side_right) {
        for (int k = 0; k < M; ++k) {
            if (!unit) {
                T const akk = a(k, k);
                for (int j = threadIdx.x; j < N; j += kThreads) x(k, j) = x(k, j) / akk;
                __syncthreads();
            }
            int const nr = M - k - 1;
            for (int e = threadIdx.x; e < nr * N; e += kThreads) {
                int const i = k + 1 + e % nr, j = e / nr;
                x(i, j) = fma_t(-a(i, k), x(k, j), x(i, j));
            }
            __syncthreads();
        }
    } else {
        for (int k = 0; k < N; ++k) {
            if (!unit) {
                T const akk = a(k, k);
                for (int i = threadIdx.x; i < M; i += kThreads) x(i, k) = x(i, k) / akk;
                __syncthreads();
            }
            int const nc = N - k - 1;
            for (int e = threadIdx.x; e < M * nc; e += kThreads) {
                int const i = e % M, j = k + 1 + e / M;
                x(i, j) = fma_t(-x(i, k), a(k, j), x(i, j));
            }
            __syncthreads();
        }
    }
}

// Cholesky, right-looking sweep in place in L.  For column k of the 4-column
// block k0: s = sqrt(acc_kk), l_kk = acc_kk / s; rows inside BLAST's diagonal
// register tile [k0, k0+KM) divide by s, rows below divide by l_kk.
template <typename T>
__global__ __launch_bounds__(kThreads) void potrf_ref_kernel(Mat<T> A, Mat<T> L, int32_t* info) {
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b), l(L, b);
    int const M = L.rows, N = L.cols;
    constexpr int TS = tile_size<T>();
    bool const alias = (a.p == l.p) && A.panel == L.panel && A.trans == L.trans && A.spacing == L.spacing;
    if (!alias)
        for (int e = threadIdx.x; e < M * N; e += kThreads) {
            int const i = e % M, j = e / M;
            if (i >= j) l(i, j) = a(i, j);
        }
    __syncthreads();
    int32_t inf = 0;
    for (int k0 = 0; k0 < N; k0 += 4) {
        int const KM = (k0 + 2 * TS < M) ? 3 * TS : (k0 + TS < M) ? 2 * TS : TS;
        int const k1 = min(k0 + 4, N);
        for (int k = k0; k < k1; ++k) {
            T const d = l(k, k);
            if (!(d > T(0)) && inf == 0) inf = k + 1;
            T const s = sqrt(d);
            T const lkk = d / s;
            for (int i = k + 1 + threadIdx.x; i < M; i += kThreads) l(i, k) = l(i, k) / (i < k0 + KM ? s : lkk);
            __syncthreads();  // every thread has read l(k,k) before it is replaced
            if (threadIdx.x == 0) l(k, k) = lkk;
            int const nr = M - k - 1, nc = N - k - 1;
            for (int e = threadIdx.x; e < nr * nc; e += kThreads) {
                int const i = k + 1 + e % nr, j = k + 1 + e / nr;
                if (i >= j) l(i, j) = fma_t(-l(i, k), l(j, k), l(i, j));
            }
            __syncthreads();
        }
        if (L.panel)
            for (int e = threadIdx.x; e < 16; e += kThreads) {
                int const i = k0 + (e & 3), k = k0 + (e >> 2);
                if (i < k && k < k1) l(i, k) = T(0);
            }
    }
    if (info && threadIdx.x == 0) info[b] = inf;
}

// ---------------------------------------------------------------- getrf
template <typename T>
__device__ int block_iamax(Ref<T> const& a, int r0, int c, int n, T* sv, int* si) {
    T v = T(-1);
    int idx = 0x7fffffff;
    for (int r = threadIdx.x; r < n; r += kThreads) {
        argmax_combine(v, idx, iamax_key(a(r0 + r, c), r), r);
    }
    for (int o = 32; o > 0; o >>= 1) {
        T const v2 = __shfl_xor(v, o);
        int const i2 = __shfl_xor(idx, o);
        argmax_combine(v, idx, v2, i2);
    }
    int const w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = idx; }
    __syncthreads();
    v = sv[0];
    idx = si[0];
    for (int q = 1; q < kThreads / 64; ++q) argmax_combine(v, idx, sv[q], si[q]);
    __syncthreads();
    return idx;
}

// Unblocked LU (Getf2.hpp:52-77) of the m x n block at (r0, c0); pivots
// relative to r0 written to piv.  Returns 1-based relative column of an
// exact-zero pivot, 0 otherwise.
template <typename T>
__device__ int getf2_block(Ref<T> const& a, int r0, int c0, int m, int n, int64_t* piv, T* sv, int* si) {
    for (int k = 0; k < m && k < n; ++k) {
        int const ip = block_iamax(a, r0 + k, c0 + k, m - k, sv, si) + k;
        if (threadIdx.x == 0) piv[k] = ip;
        if (ip != k)
            for (int j = threadIdx.x; j < n; j += kThreads) {
                T const t = a(r0 + k, c0 + j);
                a(r0 + k, c0 + j) = a(r0 + ip, c0 + j);
                a(r0 + ip, c0 + j) = t;
            }
        __syncthreads();
        T const p = a(r0 + k, c0 + k);
        if (p == T(0)) return k + 1;
        for (int i = k + 1 + threadIdx.x; i < m; i += kThreads) a(r0 + i, c0 + k) = a(r0 + i, c0 + k) / p;
        __syncthreads();
        int const nr = m - k - 1, nc = n - k - 1;
        for (int e = threadIdx.x; e < nr * nc; e += kThreads) {
            int const i = k + 1 + e % nr, j = k + 1 + e / nr;
            a(r0 + i, c0 + j) = fma_t(-a(r0 + i, c0 + k), a(r0 + k, c0 + j), a(r0 + i, c0 + j));
        }
        __syncthreads();
    }
    return 0;
}

template <typename T>
__device__ void laswp_cols(Ref<T> const& a, int c0, int c1, int k0, int k1, int64_t const* piv) {
    for (int j = c0 + threadIdx.x; j < c1; j += kThreads)
        for (int k = k0; k < k1; ++k) {
            int const p = int(piv[k]);
            if (p != k) {
                T const t = a(k, j);
                a(k, j) = a(p, j);
                a(p, j) = t;
            }
        }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void getrf_ref_kernel(Mat<T> A, int64_t* ipiv, int64_t ipiv_stride,
                                                              int32_t* info) {
    __shared__ T sv[kThreads / 64];
    __shared__ int si[kThreads / 64];
    int64_t const b = blockIdx.x;
    Ref<T> a(A, b);
    int64_t* piv = ipiv + b * ipiv_stride;
    int const M = A.rows, N = A.cols;
    constexpr int NB = tile_size<T>();
    int32_t inf = 0;
    if (A.trans) {
        // row-major storage: no pivoting (dense/Getrf.hpp:141-162)
        for (int k = 0; k < M && k < N; ++k) {
            T const p = a(k, k);
            for (int i = k + 1 + threadIdx.x; i < M; i += kThreads) a(i, k) = a(i, k) / p;
            __syncthreads();
            int const nr = M - k - 1, nc = N - k - 1;
            for (int e = threadIdx.x; e < nr * nc; e += kThreads) {
                int const i = k + 1 + e % nr, j = k + 1 + e / nr;
                a(i, j) = fma_t(-a(i, k), a(k, j), a(i, j));
            }
            if (threadIdx.x == 0) piv[k] = k;
            __syncthreads();
        }
        if (info && threadIdx.x == 0) info[b] = 0;
        return;
    }
    int k = 0;
    for (; k + NB < M && k + NB < N; k += NB) {
        int const r = getf2_block(a, k, k, M - k, NB, piv + k, sv, si);
        if (r) { inf = k + r; break; }
        if (threadIdx.x < NB) piv[k + threadIdx.x] += k;
        __syncthreads();
        laswp_cols(a, 0, k, k, k + NB, piv);
        laswp_cols(a, k + NB, N, k, k + NB, piv);
        __syncthreads();
        // U12 := L11^{-1} A12, unit lower, one column per thread
        for (int j = k + NB + threadIdx.x; j < N; j += kThreads)
            for (int i = 0; i < NB; ++i) {
                T v = a(k + i, j);
                for (int kk = 0; kk < i; ++kk) v = fma_t(-a(k + i, k + kk), a(k + kk, j), v);
                a(k + i, j) = v;
            }
        __syncthreads();
        // A22 := -1*A21*A12 + 1*A22 with gemm's per-element sequence
        int const nr = M - k - NB, nc = N - k - NB;
        for (int e = threadIdx.x; e < nr * nc; e += kThreads) {
            int const i = k + NB + e % nr, j = k + NB + e / nr;
            T acc = T(0);
            for (int kk = 0; kk < NB; ++kk) acc = fma_t(a(i, k + kk), a(k + kk, j), acc);
            acc = acc * T(-1);
            a(i, j) = fma_t(T(1), a(i, j), acc);
        }
        __syncthreads();
    }
    if (inf == 0) {
        if (k < M && k < N) {
            int const r = getf2_block(a, k, k, M - k, N - k, piv + k, sv, si);
            if (r) inf = k + r;
        }
        if (inf == 0) {
            int const kn = min(M, N);
            if (threadIdx.x < kn - k) piv[k + threadIdx.x] += k;
            __syncthreads();
            laswp_cols(a, 0, k, k, kn, piv);
        }
    }
    if (info && threadIdx.x == 0) info[b] = inf;
}

// ---------------------------------------------------------------- launchers
template <typename T>
int launch_gemm_ref(hipStream_t s, T alpha, Mat<T> A, Mat<T> B, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    if (batch == 0 || D.rows == 0 || D.cols == 0) return 0;
    hipLaunchKernelGGL(gemm_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, alpha, A, B, beta, C, D);
    return check_launch("gemm_ref_kernel");
}
template <typename T>
int launch_syrk_ref(hipStream_t s, T alpha, Mat<T> A, T beta, Mat<T> C, Mat<T> D, int64_t batch) {
    if (batch == 0 || D.rows == 0) return 0;
    hipLaunchKernelGGL(syrk_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, alpha, A, beta, C, D);
    return check_launch("syrk_ref_kernel");
}
template <typename T>
int launch_trmm_ref(hipStream_t s, int side_right, T alpha, Mat<T> A, Mat<T> B, Mat<T> C, int64_t batch) {
    if (batch == 0 || C.rows == 0 || C.cols == 0) return 0;
    hipLaunchKernelGGL(trmm_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, side_right, alpha, A, B, C);
    return check_launch("trmm_ref_kernel");
}
template <typename T>
int launch_trsm_ref(hipStream_t s, int side_right, int upper, int unit, T alpha, Mat<T> A, Mat<T> B, Mat<T> X,
                    int64_t batch) {
    if (batch == 0 || X.rows == 0 || X.cols == 0) return 0;
    hipLaunchKernelGGL(trsm_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, side_right, upper, unit,
                       alpha, A, B, X);
    return check_launch("trsm_ref_kernel");
}
template <typename T>
int launch_potrf_ref(hipStream_t s, Mat<T> A, Mat<T> L, int64_t batch, int32_t* info) {
    if (batch == 0) return 0;
    hipLaunchKernelGGL(potrf_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, A, L, info);
    return check_launch("potrf_ref_kernel");
}
template <typename T>
int launch_getrf_ref(hipStream_t s, Mat<T> A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    if (batch == 0) return 0;
    hipLaunchKernelGGL(getrf_ref_kernel<T>, dim3(unsigned(batch)), dim3(kThreads), 0, s, A, ipiv, ipiv_stride, info);
    return check_launch("getrf_ref_kernel");
}

#define INSTANTIATE(T)                                                                                       \
    template int launch_gemm_ref<T>(hipStream_t, T, Mat<T>, Mat<T>, T, Mat<T>, Mat<T>, int64_t);            \
    template int launch_syrk_ref<T>(hipStream_t, T, Mat<T>, T, Mat<T>, Mat<T>, int64_t);                    \
    template int launch_trmm_ref<T>(hipStream_t, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);                  \
    template int launch_trsm_ref<T>(hipStream_t, int, int, int, T, Mat<T>, Mat<T>, Mat<T>, int64_t);        \
    template int launch_potrf_ref<T>(hipStream_t, Mat<T>, Mat<T>, int64_t, int32_t*);                       \
    template int launch_getrf_ref<T>(hipStream_t, Mat<T>, int64_t*, int64_t, int64_t, int32_t*);
INSTANTIATE(double)
INSTANTIATE(float)

}  // namespace blast_hip
