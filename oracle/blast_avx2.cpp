// BLAST-restated CPU kernels (AVX2 + FMA) for the benchmark's CPU baseline.
//
// TEST INFRASTRUCTURE ONLY: the benchmark times this beside the GPU path, and
// tests/test_oracle_avx2.py checks it bit for bit against the scalar oracle
// (blast_oracle.cpp).  The product library never links or calls it.
//
// What it restates: BLAST's panel Cholesky (panel/Potrf.hpp:59-106, backend
// :23-56) as BLAST's AVX2 build runs it, on RegisterMatrix tiles
// (register_matrix/RegisterMatrix.hpp): the matrix is factored by 4-column
// blocks k0 (KN = min((RC - (MAX_RM + 1)) / MAX_RM, SS) = 4 with 16 AVX2
// registers, MAX_RM = 3, SS = 4), each block column cut into row tiles of
// 12, 8 and 4 rows (12 while i + 8 < M, then 8 while i + 4 < M, then 4:
// Potrf.hpp:97-104).  A tile is (KM/4) x 4 registers of 4 doubles (one panel
// column each); it is loaded from A, receives ger(-1, L(i.., l), L(k0.., l)^T)
// for l < k0 (RegisterMatrix::ger, fnmadd, :605-706), then the in-register
// Cholesky of the diagonal tile (RegisterMatrix::potrf, :709-739: sqrt of the
// pivot, the pivot row divided too) or the right-upper solve with L(k0, k0)^T
// (RegisterMatrix::trsm, :563-602: division by the stored diagonal), and is
// stored whole.  The two triangular solves of the Cholesky solve follow
// dense/Trsm.hpp:17-66, a scalar loop in BLAST, restated as such.
//
// Per element the arithmetic is exactly the scalar oracle's (the same fma
// sequence in the same order), so the results agree bit for bit; only the
// instruction mix differs (4-wide FMAs, register tiles).  Panel storage with
// n a multiple of 4, as BLAST's StaticPanelMatrix<double, 128, 128> of the
// benchmark.  Built -O3 -march=x86-64-v3 (AVX2 + FMA) -ffp-contract=off.

#include <immintrin.h>
#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

extern "C" {
typedef struct {
    void* data;
    int64_t rows;
    int64_t cols;
    int64_t spacing;
    int64_t batch_stride;
    int32_t panel;
    int32_t trans;
} OMatAvx;
}

namespace {

constexpr int64_t PS = 4;  // panel size of double on AVX2 (panel/PanelSize.hpp:25-26)

// Column j of panel p (rows 4p..4p+3) of a column-major panel matrix.
inline double* panel_col(double* base, int64_t spacing, int64_t p, int64_t j) { return base + p * spacing + j * PS; }

// One row tile of KM rows (KM / 4 registers per column) of block column k0.
template <int R>
void potrf_tile(double const* a, double* l, int64_t sp, int64_t M, int64_t N, int64_t k0, int64_t i, int32_t* inf) {
    (void)M;
    __m256d acc[R][4];
    int64_t const kn = std::min<int64_t>(4, N - k0);
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < 4; ++c)
            acc[r][c] = c < kn ? _mm256_loadu_pd(panel_col(const_cast<double*>(a), sp, i / PS + r, k0 + c))
                               : _mm256_setzero_pd();
    // ger(-1, L(i.., l), L(k0.., l)^T) for l < k0, l increasing
    for (int64_t lc = 0; lc < k0; ++lc) {
        __m256d b[4];
        double const* bcol = panel_col(l, sp, k0 / PS, lc);  // rows k0..k0+3 of column l
        for (int c = 0; c < 4; ++c) b[c] = _mm256_broadcast_sd(bcol + c);
        for (int r = 0; r < R; ++r) {
            __m256d const av = _mm256_loadu_pd(panel_col(l, sp, i / PS + r, lc));
            for (int c = 0; c < 4; ++c) acc[r][c] = _mm256_fnmadd_pd(av, b[c], acc[r][c]);
        }
    }
    alignas(32) double lane[4];
    if (i == k0) {
        // RegisterMatrix::potrf: column k: s = sqrt(d), every row of the tile
        // divided by s (the pivot row gives l_kk = d / s), then the later
        // columns of the block updated with fnmadd
        for (int k = 0; k < kn; ++k) {
            _mm256_store_pd(lane, acc[0][k]);
            double const d = lane[k];
            if (!(d > 0.0) && *inf == 0) *inf = int32_t(k0 + k + 1);
            __m256d const s = _mm256_set1_pd(std::sqrt(d));
            for (int r = 0; r < R; ++r) acc[r][k] = _mm256_div_pd(acc[r][k], s);
            _mm256_store_pd(lane, acc[0][k]);
            for (int j = k + 1; j < kn; ++j) {
                __m256d const ljk = _mm256_set1_pd(lane[j]);
                for (int r = 0; r < R; ++r) acc[r][j] = _mm256_fnmadd_pd(acc[r][k], ljk, acc[r][j]);
            }
        }
        // the strict upper part of the 4x4 diagonal block is written as 0
        // (BLAST leaves rounding residue there; the scalar oracle writes 0)
        for (int k = 0; k < kn; ++k) {
            _mm256_store_pd(lane, acc[0][k]);
            for (int rr = 0; rr < k; ++rr) lane[rr] = 0.0;
            acc[0][k] = _mm256_load_pd(lane);
        }
    } else {
        // RegisterMatrix::trsm(Right, Upper, L(k0, k0)^T): column k divided by
        // the stored l_kk, then the later columns updated with fnmadd
        double const* dk = panel_col(l, sp, k0 / PS, k0);
        for (int k = 0; k < kn; ++k) {
            __m256d const lkk = _mm256_set1_pd(dk[k * PS + k]);
            for (int r = 0; r < R; ++r) acc[r][k] = _mm256_div_pd(acc[r][k], lkk);
            for (int j = k + 1; j < kn; ++j) {
                __m256d const ljk = _mm256_set1_pd(dk[k * PS + j]);  // L(k0 + j, k0 + k)
                for (int r = 0; r < R; ++r) acc[r][j] = _mm256_fnmadd_pd(acc[r][k], ljk, acc[r][j]);
            }
        }
    }
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < kn; ++c) _mm256_storeu_pd(panel_col(l, sp, i / PS + r, k0 + c), acc[r][c]);
}

// panel/Potrf.hpp:59-106 on one n x n instance (n % 4 == 0).
int32_t potrf_panel(double const* a, double* l, int64_t sp, int64_t n) {
    int32_t inf = 0;
    for (int64_t k0 = 0; k0 < n; k0 += 4) {
        int64_t i = k0;
        for (; i + 8 < n; i += 12) potrf_tile<3>(a, l, sp, n, n, k0, i, &inf);
        for (; i + 4 < n; i += 8) potrf_tile<2>(a, l, sp, n, n, k0, i, &inf);
        for (; i < n; i += 4) potrf_tile<1>(a, l, sp, n, n, k0, i, &inf);
    }
    return inf;
}

inline int64_t idx(OMatAvx const& m, int64_t i, int64_t j) {
    int64_t const r = m.trans ? j : i, c = m.trans ? i : j;
    return m.panel ? (r / PS) * m.spacing + r % PS + c * PS : r + m.spacing * c;
}

// dense/Trsm.hpp:17-66: x_i = (b_i - sum_k a_ik x_k) / a_ii, the sum as
// fma(-a_ik, x_k, .) in increasing k (lower) or decreasing k (upper, on
// trans(L)), one right-hand-side column after the other.
void solve(double const* l, int64_t sp, int64_t n, OMatAvx const& B, double const* b, OMatAvx const& X, double* x) {
    auto L = [&](int64_t i, int64_t j) { return l[(i / PS) * sp + i % PS + j * PS]; };
    for (int64_t c = 0; c < B.cols; ++c) {
        for (int64_t i = 0; i < n; ++i) {
            double v = b[idx(B, i, c)];
            for (int64_t k = 0; k < i; ++k) v = std::fma(-L(i, k), x[idx(X, k, c)], v);
            x[idx(X, i, c)] = v / L(i, i);
        }
        for (int64_t i = n; i-- > 0;) {
            double v = x[idx(X, i, c)];
            for (int64_t k = i + 1; k < n; ++k) v = std::fma(-L(k, i), x[idx(X, k, c)], v);
            x[idx(X, i, c)] = v / L(i, i);
        }
    }
}

}  // namespace

extern "C" {

// Cholesky solve (potrf + trsm<Lower> + trsm<Upper>(trans(L))) of `batch`
// instances, A and L n x n column-major panel storage (n % 4 == 0), B and X
// any storage.  One single-threaded instance per worker, each worker pinned
// to one CPU of the affinity mask (threads <= 0: one per CPU), static
// round-robin over the batch.  Returns the number of workers, or -1 for an
// unsupported shape.
int oracle_dposv_avx2_mt(OMatAvx const* A, OMatAvx const* L, OMatAvx const* B, OMatAvx const* X, int64_t batch,
                         int32_t* info, int threads) {
    int64_t const n = A->rows;
    if (A->cols != n || n % PS || !A->panel || !L->panel || A->trans || L->trans || L->rows != n ||
        L->cols != n || L->spacing != A->spacing || B->rows != n || X->rows != n || X->cols != B->cols)
        return -1;
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
            for (int64_t b = t; b < batch; b += threads) {
                double const* a = static_cast<double const*>(A->data) + b * A->batch_stride;
                double* l = static_cast<double*>(L->data) + b * L->batch_stride;
                int32_t const inf = potrf_panel(a, l, A->spacing, n);
                if (info) info[b] = inf;
                solve(l, L->spacing, n, *B, static_cast<double const*>(B->data) + b * B->batch_stride, *X,
                      static_cast<double*>(X->data) + b * X->batch_stride);
            }
        });
    for (auto& th : pool) th.join();
    return threads;
}

}  // extern "C"
