// CPU oracle for the batched dense-kernel layer.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain scalar C++, the
// floating-point algorithm of BLAST's CPU kernels so that the HIP kernels can
// be checked against it.  Only the tests, the smoke check and the CPU-baseline
// leg of the benchmark load it; the product library never links or calls it.
//
// Every routine names the reference file:line it follows (paths relative to
// the mkatliar/blast checkout).  The restatement reproduces BLAST's per-element
// operation order:
//   * every multiply-add that BLAST performs through RegisterMatrix
//     (fmadd / fnmadd, register_matrix/RegisterMatrix.hpp:605-706) is an
//     explicit std::fma here;
//   * scalar loops of the form `acc -= a * b` (dense/Trsm.hpp:57-60,
//     reference/Trsm.hpp:50-52) are restated as fma(-a, b, acc), i.e. the
//     contraction clang performs at its default -ffp-contract=on;
//   * this file is compiled with -ffp-contract=off so that no other
//     contraction happens.
//
// Storage: a matrix operand is described by OMat (same field order as the
// product's blast_matrix_t).  Element (i, j) of the logical matrix lives at
//   trans == 0: storage(i, j);  trans == 1: storage(j, i)
// where storage is column-major, either dense (idx = i + spacing*j,
// dense/DynamicMatrix.hpp:149) or panel (idx = (i/PS)*spacing + i%PS + j*PS,
// panel/DynamicPanelMatrix.hpp:187; PS = 4 for double, 8 for float,
// panel/PanelSize.hpp:25-26).  A row-major panel or dense matrix is the
// trans == 1 view of the column-major storage of its transpose.
//
// Parity status of each routine is recorded in the test files under tests/.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>
#include <sched.h>

extern "C" {
typedef struct {
    void* data;
    int64_t rows;
    int64_t cols;
    int64_t spacing;
    int64_t batch_stride;
    int32_t panel;
    int32_t trans;
} OMat;
}

namespace {

template <typename T> struct PanelSize;
template <> struct PanelSize<double> { static constexpr int64_t value = 4; };
template <> struct PanelSize<float> { static constexpr int64_t value = 8; };

// TileSize_v (system/Tile.hpp:19-34): 4 for double, 8 for float.
template <typename T> constexpr int64_t tile_size() { return PanelSize<T>::value; }

template <typename T>
struct View {
    T* p;
    int64_t rows, cols, spacing;
    bool panel, trans;

    View(OMat const& m, int64_t b)
        : p(static_cast<T*>(m.data) + b * m.batch_stride), rows(m.rows), cols(m.cols),
          spacing(m.spacing), panel(m.panel != 0), trans(m.trans != 0) {}

    int64_t idx(int64_t i, int64_t j) const {
        if (trans) std::swap(i, j);
        constexpr int64_t PS = PanelSize<T>::value;
        return panel ? (i / PS) * spacing + (i % PS) + j * PS : i + spacing * j;
    }
    T& operator()(int64_t i, int64_t j) const { return p[idx(i, j)]; }
};

// ---------------------------------------------------------------- gemm
// algorithm/Gemm.hpp:41-62 tiles D and runs register_matrix/Gemm.hpp:71-92 on
// every tile: ker.reset(); acc = fma(a_ik, b_kj, acc) for k = 0..K-1;
// ker *= alpha; ker.axpy(beta, C) = fma(beta, c, acc); store.  The tile
// shape does not change the per-element operation sequence.
template <typename T>
int gemm(T alpha, OMat const& A, OMat const& B, T beta, OMat const& C, OMat const& D, int64_t batch) {
    int64_t const M = A.rows, K = A.cols, N = B.cols;
    if (B.rows != K || C.rows != M || C.cols != N || D.rows != M || D.cols != N) return -1;
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), bb(B, b), c(C, b), d(D, b);
        for (int64_t j = 0; j < N; ++j)
            for (int64_t i = 0; i < M; ++i) {
                T acc = T(0);
                for (int64_t k = 0; k < K; ++k) acc = std::fma(a(i, k), bb(k, j), acc);
                acc = acc * alpha;
                d(i, j) = std::fma(beta, c(i, j), acc);
            }
    }
    return 0;
}

// ---------------------------------------------------------------- syrkLower
// dense/Syrk.hpp:15-145: ker.load(beta, C) (v = beta*c),
// gemm(ker, K, alpha, A(i,:), trans(A(j,:))) = ger(alpha, ...) with
// ax = alpha*a_ik, v = fma(ax, a_jk, v) (RegisterMatrix.hpp:610-627),
// storeLower on diagonal tiles.  Only i >= j is written.
template <typename T>
int syrk_lower(T alpha, OMat const& A, T beta, OMat const& C, OMat const& D, int64_t batch) {
    int64_t const M = A.rows, K = A.cols;
    if (C.rows != M || C.cols != M || D.rows != M || D.cols != M) return -1;
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), c(C, b), d(D, b);
        for (int64_t j = 0; j < M; ++j)
            for (int64_t i = j; i < M; ++i) {
                T acc = beta * c(i, j);
                for (int64_t k = 0; k < K; ++k) acc = std::fma(alpha * a(i, k), a(j, k), acc);
                d(i, j) = acc;
            }
    }
    return 0;
}

// ---------------------------------------------------------------- trmm
// Left/upper: C = alpha*A*B, A upper triangular M x M, B M x N
// (algorithm/Trmm.hpp:93-126 with RegisterMatrix::trmm :742-790 and the
// gemm remainder): acc = 0; acc = fma(alpha*a_ik, b_kj, acc), k = i..M-1.
// The strict lower part of A is not referenced.
template <typename T>
int trmm_left_upper(T alpha, OMat const& A, OMat const& B, OMat const& C, int64_t batch) {
    int64_t const M = B.rows, N = B.cols;
    if (A.rows != M || A.cols != M || C.rows != M || C.cols != N) return -1;
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), bb(B, b), c(C, b);
        for (int64_t j = 0; j < N; ++j)
            for (int64_t i = 0; i < M; ++i) {
                T acc = T(0);
                for (int64_t k = i; k < M; ++k) acc = std::fma(alpha * a(i, k), bb(k, j), acc);
                c(i, j) = acc;
            }
    }
    return 0;
}

// Right/lower: C = alpha*B*A, A lower triangular N x N, B M x N
// (algorithm/Trmm.hpp:146-242, RegisterMatrix::trmm :793-880):
// acc = 0; acc = fma(alpha*b_ik, a_kj, acc), k = j..N-1.
template <typename T>
int trmm_right_lower(T alpha, OMat const& B, OMat const& A, OMat const& C, int64_t batch) {
    int64_t const M = B.rows, N = B.cols;
    if (A.rows != N || A.cols != N || C.rows != M || C.cols != N) return -1;
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), bb(B, b), c(C, b);
        for (int64_t j = 0; j < N; ++j)
            for (int64_t i = 0; i < M; ++i) {
                T acc = T(0);
                for (int64_t k = j; k < N; ++k) acc = std::fma(alpha * bb(i, k), a(k, j), acc);
                c(i, j) = acc;
            }
    }
    return 0;
}

// ---------------------------------------------------------------- trsm
// Left: A*X = alpha*B (reference/Trsm.hpp:44-66; dense/Trsm.hpp:17-66 is the
// alpha = 1 case).  v = alpha*b_ij; v = fma(-a_ik, x_kj, v) over k in
// increasing (lower) or decreasing (upper) order; x_ij = unit ? v : v/a_ii.
// X may alias B: column j of B is read only before it is overwritten.
template <typename T>
int trsm_left(int lower, int unit, T alpha, OMat const& A, OMat const& B, OMat const& X, int64_t batch) {
    int64_t const M = B.rows, N = B.cols;
    if (A.rows != M || A.cols != M || X.rows != M || X.cols != N) return -1;
    std::vector<T> col(M);
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), bb(B, b), x(X, b);
        for (int64_t j = 0; j < N; ++j) {
            if (lower) {
                for (int64_t i = 0; i < M; ++i) {
                    T v = alpha * bb(i, j);
                    for (int64_t k = 0; k < i; ++k) v = std::fma(-a(i, k), x(k, j), v);
                    x(i, j) = unit ? v : v / a(i, i);
                }
            } else {
                for (int64_t i = M; i-- > 0;) {
                    T v = alpha * bb(i, j);
                    for (int64_t k = i + 1; k < M; ++k) v = std::fma(-a(i, k), x(k, j), v);
                    x(i, j) = unit ? v : v / a(i, i);
                }
            }
        }
    }
    return 0;
}

// Right: X*A = alpha*B (reference/Trsm.hpp:105-127).
template <typename T>
int trsm_right(int lower, int unit, T alpha, OMat const& A, OMat const& B, OMat const& X, int64_t batch) {
    int64_t const M = B.rows, N = B.cols;
    if (A.rows != N || A.cols != N || X.rows != M || X.cols != N) return -1;
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), bb(B, b), x(X, b);
        for (int64_t i = 0; i < M; ++i) {
            if (!lower) {
                for (int64_t j = 0; j < N; ++j) {
                    T v = alpha * bb(i, j);
                    for (int64_t k = 0; k < j; ++k) v = std::fma(-x(i, k), a(k, j), v);
                    x(i, j) = unit ? v : v / a(j, j);
                }
            } else {
                for (int64_t j = N; j-- > 0;) {
                    T v = alpha * bb(i, j);
                    for (int64_t k = j + 1; k < N; ++k) v = std::fma(-x(i, k), a(k, j), v);
                    x(i, j) = unit ? v : v / a(j, j);
                }
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------- potrf
// dense/Potrf.hpp:66-106 (driver) + :20-63 (backend); panel/Potrf.hpp:59-106.
// Left-looking by 4-column blocks k0; rows tiled 3*TS / 2*TS / TS.  For every
// element the arithmetic is
//   acc = a_ij;  acc = fma(-l_il, l_jl, acc) for l = 0..j-1 (increasing)
// (ger(-1, ...) for l < k0, RegisterMatrix::potrf / ::trsm fnmadd inside the
// block, RegisterMatrix.hpp:709-739 and :566-591).  For column k of block k0:
//   s = sqrt(acc_kk);  l_kk = acc_kk / s   (the diagonal row is divided too,
//   RegisterMatrix.hpp:728-737);
//   rows inside the diagonal register tile [k0, k0+KM) divide by s;
//   rows of the off-diagonal tiles divide by the stored l_kk (trsm, :584-589).
// KM is the height of the first row tile: 3*TS if k0+2*TS < M, else 2*TS if
// k0+TS < M, else TS (Potrf.hpp:97-104).
// Dense storage: only i >= j is written (storeLower).  Panel storage: BLAST
// stores whole tiles, leaving rounding residue (~1e-16) in the strict upper
// part of each 4x4 diagonal block; here that residue is written as 0.
// No SPD check: a non-positive pivot yields NaN which propagates, as in BLAST.
// info[b] reports the first column (1-based) whose pivot was not > 0.
template <typename T>
int potrf(OMat const& A, OMat const& L, int64_t batch, int32_t* info) {
    int64_t const M = A.rows, N = A.cols;
    if (N > M || L.rows != M || L.cols != N) return -1;
    constexpr int64_t TS = tile_size<T>();
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b), l(L, b);
        int32_t inf = 0;
        for (int64_t j = 0; j < N; ++j)
            for (int64_t i = j; i < M; ++i) l(i, j) = a(i, j);
        for (int64_t k0 = 0; k0 < N; k0 += 4) {
            int64_t const KM = (k0 + 2 * TS < M) ? 3 * TS : (k0 + TS < M) ? 2 * TS : TS;
            int64_t const k1 = std::min(k0 + 4, N);
            for (int64_t k = k0; k < k1; ++k) {
                T const d = l(k, k);
                if (!(d > T(0)) && inf == 0) inf = int32_t(k + 1);
                T const s = std::sqrt(d);
                T const lkk = d / s;
                l(k, k) = lkk;
                for (int64_t i = k + 1; i < M; ++i) l(i, k) = l(i, k) / (i < k0 + KM ? s : lkk);
                for (int64_t j = k + 1; j < N; ++j) {
                    T const ljk = l(j, k);
                    for (int64_t i = j; i < M; ++i) l(i, j) = std::fma(-l(i, k), ljk, l(i, j));
                }
            }
            if (l.panel)
                for (int64_t k = k0; k < k1; ++k)
                    for (int64_t i = k0; i < k; ++i) l(i, k) = T(0);
        }
        if (info) info[b] = inf;
    }
    return 0;
}

// ---------------------------------------------------------------- iamax
// dense/Iamax.hpp:66-95 (active scalar branch): first index of max |x_i|
// with strict '>' so ties keep the lowest index.
template <typename T>
int64_t iamax(int64_t n, T const* x, int64_t incx) {
    if (n <= 0) return -1;
    T value = std::abs(x[0]);
    int64_t index = 0;
    for (int64_t i = 1; i < n; ++i) {
        T const v = std::abs(x[i * incx]);
        if (v > value) { value = v; index = i; }
    }
    return index;
}

template <typename T>
void swap_rows(View<T> const& a, int64_t r0, int64_t r1, int64_t c0, int64_t c1) {
    for (int64_t j = c0; j < c1; ++j) std::swap(a(r0, j), a(r1, j));
}

// ---------------------------------------------------------------- getf2
// dense/Getf2.hpp:52-77 on the m x n block of `a` with top-left (r0, c0):
// pivot = first max |.| (Iamax), swap rows across the block's n columns
// (Swap.hpp:40-44), exact-zero pivot => singular (:68-69), divide the
// sub-column by the pivot (:71-72), rank-1 update ger(-1) (Ger.hpp:40-135:
// v = fma(-x_i, y_j, v)).  Pivots are relative to r0.  Returns 0 or the
// 1-based relative column of the zero pivot.
template <typename T>
int64_t getf2_block(View<T> const& a, int64_t r0, int64_t c0, int64_t m, int64_t n, int64_t* ipiv) {
    for (int64_t k = 0; k < m && k < n; ++k) {
        T value = std::abs(a(r0 + k, c0 + k));
        int64_t ip = k;
        for (int64_t i = k + 1; i < m; ++i) {
            T const v = std::abs(a(r0 + i, c0 + k));
            if (v > value) { value = v; ip = i; }
        }
        ipiv[k] = ip;
        if (ip != k) swap_rows(a, r0 + k, r0 + ip, c0, c0 + n);
        T const piv = a(r0 + k, c0 + k);
        if (piv == T(0)) return k + 1;
        for (int64_t i = k + 1; i < m; ++i) a(r0 + i, c0 + k) = a(r0 + i, c0 + k) / piv;
        if (k + 1 < m && k + 1 < n)
            for (int64_t j = k + 1; j < n; ++j) {
                T const y = a(r0 + k, c0 + j);
                for (int64_t i = k + 1; i < m; ++i)
                    a(r0 + i, c0 + j) = std::fma(-a(r0 + i, c0 + k), y, a(r0 + i, c0 + j));
            }
    }
    return 0;
}

// laswp (dense/Laswp.hpp:37-43) on columns [c0, c1): for k in [k0, k1) swap
// rows k and ipiv[k] when they differ.
template <typename T>
void laswp_cols(View<T> const& a, int64_t c0, int64_t c1, int64_t k0, int64_t k1, int64_t const* ipiv) {
    for (int64_t k = k0; k < k1; ++k)
        if (k != ipiv[k]) swap_rows(a, k, ipiv[k], c0, c1);
}

// ---------------------------------------------------------------- getrf
// Column-major: dense/Getrf.hpp:52-119, blocked right-looking with
// NB = TileSize (4 double, 8 float).  Per block: getf2 on the (M-k) x NB
// panel, ipiv += k, laswp left and right, U12 = trsm<Lower,unit>(A11, A12)
// (v = fma(-a, x, v) in increasing order), then
// gemm(-1, A21, A12, 1, A22, A22): acc = sum fma chain from 0, acc *= -1,
// d = fma(1, c, acc).  Tail: getf2 on the remainder, ipiv += k, laswp left.
// Row-major storage (trans == 1): dense/Getrf.hpp:141-162, no pivoting,
// ipiv[k] = k.
// ipiv is 0-based with LAPACK's sequential-interchange meaning.
template <typename T>
int getrf(OMat const& A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    int64_t const M = A.rows, N = A.cols;
    constexpr int64_t NB = tile_size<T>();
    for (int64_t b = 0; b < batch; ++b) {
        View<T> a(A, b);
        int64_t* piv = ipiv + b * ipiv_stride;
        int32_t inf = 0;
        if (a.trans) {
            for (int64_t k = 0; k < M && k < N; ++k) {
                T const p = a(k, k);
                for (int64_t i = k + 1; i < M; ++i) a(i, k) = a(i, k) / p;
                for (int64_t j = k + 1; j < N; ++j)
                    for (int64_t i = k + 1; i < M; ++i) a(i, j) = std::fma(-a(i, k), a(k, j), a(i, j));
                piv[k] = k;
            }
            if (info) info[b] = 0;
            continue;
        }
        int64_t k = 0;
        for (; k + NB < M && k + NB < N; k += NB) {
            int64_t const r = getf2_block(a, k, k, M - k, NB, piv + k);
            if (r) { inf = int32_t(k + r); break; }
            for (int64_t i = k; i < k + NB; ++i) piv[i] += k;
            laswp_cols(a, 0, k, k, k + NB, piv);
            laswp_cols(a, k + NB, N, k, k + NB, piv);
            // U12 := L11^{-1} A12 (unit lower, in place)
            for (int64_t j = k + NB; j < N; ++j)
                for (int64_t i = 0; i < NB; ++i) {
                    T v = a(k + i, j);
                    for (int64_t kk = 0; kk < i; ++kk) v = std::fma(-a(k + i, k + kk), a(k + kk, j), v);
                    a(k + i, j) = v;
                }
            // A22 := -1 * A21 * A12 + 1 * A22
            for (int64_t j = k + NB; j < N; ++j)
                for (int64_t i = k + NB; i < M; ++i) {
                    T acc = T(0);
                    for (int64_t kk = 0; kk < NB; ++kk) acc = std::fma(a(i, k + kk), a(k + kk, j), acc);
                    acc = acc * T(-1);
                    a(i, j) = std::fma(T(1), a(i, j), acc);
                }
        }
        if (inf == 0) {
            if (k < M && k < N) {
                int64_t const r = getf2_block(a, k, k, M - k, N - k, piv + k);
                if (r) inf = int32_t(k + r);
            }
            if (inf == 0) {
                for (int64_t i = k; i < M && i < N; ++i) piv[i] += k;
                laswp_cols(a, 0, k, k, std::min(M, N), piv);
            }
        }
        if (info) info[b] = inf;
    }
    return 0;
}

// ---------------------------------------------------------------- ger
// dense/Ger.hpp:40-135: B = A + alpha*x*y^T with v = fma(alpha*x_i, y_j, a_ij).
template <typename T>
int ger(T alpha, T const* x, T const* y, OMat const& A, OMat const& B) {
    if (A.rows != B.rows || A.cols != B.cols) return -1;
    View<T> a(A, 0), bb(B, 0);
    for (int64_t j = 0; j < A.cols; ++j)
        for (int64_t i = 0; i < A.rows; ++i) bb(i, j) = std::fma(alpha * x[i], y[j], a(i, j));
    return 0;
}

// ---------------------------------------------------------------- posv
// The Cholesky solve composed by the caller from BLAST calls:
// potrf(A, L); trsm<Lower,false>(L, B, Y); trsm<Upper,false>(trans(L), Y, X).
template <typename T>
int posv(OMat const& A, OMat const& L, OMat const& B, OMat const& X, int64_t batch, int32_t* info) {
    int r = potrf<T>(A, L, batch, info);
    if (r) return r;
    r = trsm_left<T>(1, 0, T(1), L, B, X, batch);
    if (r) return r;
    OMat Lt = L;
    Lt.trans = !L.trans;
    return trsm_left<T>(0, 0, T(1), Lt, X, X, batch);
}

// One single-threaded instance stream per worker, each worker pinned to one
// CPU of the process affinity mask, static round-robin over the batch.
template <typename F>
int run_parallel(int64_t batch, int threads, F&& f) {
    cpu_set_t mask;
    std::vector<int> cpus;
    if (sched_getaffinity(0, sizeof(mask), &mask) == 0)
        for (int c = 0; c < CPU_SETSIZE; ++c)
            if (CPU_ISSET(c, &mask)) cpus.push_back(c);
    if (threads <= 0) threads = int(cpus.size());
    if (threads <= 0) threads = 1;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t]() {
            if (!cpus.empty()) {
                cpu_set_t one;
                CPU_ZERO(&one);
                CPU_SET(cpus[size_t(t) % cpus.size()], &one);
                sched_setaffinity(0, sizeof(one), &one);
            }
            for (int64_t b = t; b < batch; b += threads) f(b);
        });
    for (auto& th : pool) th.join();
    return threads;
}

OMat shift(OMat m, int64_t b, size_t elem) {
    m.data = static_cast<char*>(m.data) + b * m.batch_stride * int64_t(elem);
    return m;
}

}  // namespace

extern "C" {

int oracle_dgemm(double alpha, OMat const* A, OMat const* B, double beta, OMat const* C, OMat const* D, int64_t batch) {
    return gemm<double>(alpha, *A, *B, beta, *C, *D, batch);
}
int oracle_sgemm(float alpha, OMat const* A, OMat const* B, float beta, OMat const* C, OMat const* D, int64_t batch) {
    return gemm<float>(alpha, *A, *B, beta, *C, *D, batch);
}
int oracle_dsyrk_lower(double alpha, OMat const* A, double beta, OMat const* C, OMat const* D, int64_t batch) {
    return syrk_lower<double>(alpha, *A, beta, *C, *D, batch);
}
int oracle_ssyrk_lower(float alpha, OMat const* A, float beta, OMat const* C, OMat const* D, int64_t batch) {
    return syrk_lower<float>(alpha, *A, beta, *C, *D, batch);
}
int oracle_dtrmm_left_upper(double alpha, OMat const* A, OMat const* B, OMat const* C, int64_t batch) {
    return trmm_left_upper<double>(alpha, *A, *B, *C, batch);
}
int oracle_strmm_left_upper(float alpha, OMat const* A, OMat const* B, OMat const* C, int64_t batch) {
    return trmm_left_upper<float>(alpha, *A, *B, *C, batch);
}
int oracle_dtrmm_right_lower(double alpha, OMat const* B, OMat const* A, OMat const* C, int64_t batch) {
    return trmm_right_lower<double>(alpha, *B, *A, *C, batch);
}
int oracle_strmm_right_lower(float alpha, OMat const* B, OMat const* A, OMat const* C, int64_t batch) {
    return trmm_right_lower<float>(alpha, *B, *A, *C, batch);
}
int oracle_dtrsm_left(int lower, int unit, double alpha, OMat const* A, OMat const* B, OMat const* X, int64_t batch) {
    return trsm_left<double>(lower, unit, alpha, *A, *B, *X, batch);
}
int oracle_strsm_left(int lower, int unit, float alpha, OMat const* A, OMat const* B, OMat const* X, int64_t batch) {
    return trsm_left<float>(lower, unit, alpha, *A, *B, *X, batch);
}
int oracle_dtrsm_right(int lower, int unit, double alpha, OMat const* A, OMat const* B, OMat const* X, int64_t batch) {
    return trsm_right<double>(lower, unit, alpha, *A, *B, *X, batch);
}
int oracle_strsm_right(int lower, int unit, float alpha, OMat const* A, OMat const* B, OMat const* X, int64_t batch) {
    return trsm_right<float>(lower, unit, alpha, *A, *B, *X, batch);
}
int oracle_dpotrf(OMat const* A, OMat const* L, int64_t batch, int32_t* info) {
    return potrf<double>(*A, *L, batch, info);
}
int oracle_spotrf(OMat const* A, OMat const* L, int64_t batch, int32_t* info) {
    return potrf<float>(*A, *L, batch, info);
}
int oracle_dgetrf(OMat const* A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    return getrf<double>(*A, ipiv, ipiv_stride, batch, info);
}
int oracle_sgetrf(OMat const* A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch, int32_t* info) {
    return getrf<float>(*A, ipiv, ipiv_stride, batch, info);
}
// getf2 on the whole matrix (dense/Getf2.hpp:103-107).  Returns the
// 1-based column of an exact-zero pivot, 0 otherwise.
int oracle_dgetf2(OMat const* A, int64_t* ipiv) {
    View<double> a(*A, 0);
    return int(getf2_block(a, 0, 0, A->rows, A->cols, ipiv));
}
int oracle_sgetf2(OMat const* A, int64_t* ipiv) {
    View<float> a(*A, 0);
    return int(getf2_block(a, 0, 0, A->rows, A->cols, ipiv));
}
int64_t oracle_diamax(int64_t n, double const* x, int64_t incx) { return iamax<double>(n, x, incx); }
int64_t oracle_siamax(int64_t n, float const* x, int64_t incx) { return iamax<float>(n, x, incx); }
void oracle_dlaswp(OMat const* A, int64_t k0, int64_t k1, int64_t const* ipiv) {
    View<double> a(*A, 0);
    laswp_cols(a, 0, A->cols, k0, k1, ipiv);
}
int oracle_dger(double alpha, double const* x, double const* y, OMat const* A, OMat const* B) {
    return ger<double>(alpha, x, y, *A, *B);
}
int oracle_dposv(OMat const* A, OMat const* L, OMat const* B, OMat const* X, int64_t batch, int32_t* info) {
    return posv<double>(*A, *L, *B, *X, batch, info);
}
int oracle_sposv(OMat const* A, OMat const* L, OMat const* B, OMat const* X, int64_t batch, int32_t* info) {
    return posv<float>(*A, *L, *B, *X, batch, info);
}

// Multithreaded CPU baseline: one single-threaded posv per instance, one
// worker per CPU of the affinity mask (threads <= 0) or `threads` workers.
// Returns the number of workers used.
int oracle_dposv_mt(OMat const* A, OMat const* L, OMat const* B, OMat const* X, int64_t batch,
                    int32_t* info, int threads) {
    return run_parallel(batch, threads, [&](int64_t b) {
        OMat a = shift(*A, b, 8), l = shift(*L, b, 8), bb = shift(*B, b, 8), x = shift(*X, b, 8);
        posv<double>(a, l, bb, x, 1, info ? info + b : nullptr);
    });
}
int oracle_dgemm_mt(double alpha, OMat const* A, OMat const* B, double beta, OMat const* C, OMat const* D,
                    int64_t batch, int threads) {
    return run_parallel(batch, threads, [&](int64_t b) {
        OMat a = shift(*A, b, 8), bb = shift(*B, b, 8), c = shift(*C, b, 8), d = shift(*D, b, 8);
        gemm<double>(alpha, a, bb, beta, c, d, 1);
    });
}

}  // extern "C"
