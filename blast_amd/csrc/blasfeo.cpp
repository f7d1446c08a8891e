// BLASFEO-format entry points: the BLASFEO routines that BLAST wraps in
// include/blasfeo/blasfeo_api/ (Gemm.hpp, Syrk.hpp, Potrf.hpp,
// SyrkPotrf.hpp, Trmm.hpp), batched, on BLASFEO's panel-major storage.
//
// A BLASFEO matrix keeps element (i, j) at pA[(i/bs)*bs*cn + i%bs + j*bs]:
// BLAST's panel storage with spacing bs*cn.  A sub-block at (ai, aj) whose
// row offset is a multiple of bs is again such a matrix, starting at
// pA + (ai/bs)*bs*cn + aj*bs, so every operand becomes a panel Mat view and
// the AUTO kernels (MFMA tiles, fused Cholesky) run unchanged.  This file is
// host code only: argument checks and views.

#include <type_traits>

#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {
namespace {

// Largest BLASFEO matrix dimension accepted (the C ABI's limit).
constexpr int64_t kMaxBlasfeoDim = 1 << 20;

// Panel view of the srows x scols block of s at (i, j); trans: the logical
// operand is that block's transpose.  Returns false (error text set) when
// the descriptor or the block is unusable.
template <typename T>
bool block_view(const blast_blasfeo_mat_t* s, int64_t i, int64_t j, int64_t srows, int64_t scols, bool trans,
                const char* name, Mat<T>& out) {
    constexpr int64_t ps = PanelSize<T>::value;
    if (!s) { set_error("%s: null descriptor", name); return false; }
    if (s->bs != ps) {
        set_error("%s: panel height bs = %lld; this library's panels are %lld rows (%s)", name, (long long)s->bs,
                  (long long)ps, std::is_same_v<T, double> ? "double" : "float");
        return false;
    }
    if (s->m < 0 || s->n < 0 || s->m > kMaxBlasfeoDim || s->n > kMaxBlasfeoDim || s->cn < s->n ||
        s->cn > kMaxBlasfeoDim) {
        set_error("%s: bad size m = %lld, n = %lld, cn = %lld", name, (long long)s->m, (long long)s->n,
                  (long long)s->cn);
        return false;
    }
    if (s->batch_stride < 0) { set_error("%s: negative batch_stride", name); return false; }
    if (srows < 0 || scols < 0) { set_error("%s: negative block size", name); return false; }
    if (i < 0 || j < 0 || i + srows > s->m || j + scols > s->n) {
        set_error("%s: block (%lld, %lld) of %lld x %lld exceeds the %lld x %lld matrix", name, (long long)i,
                  (long long)j, (long long)srows, (long long)scols, (long long)s->m, (long long)s->n);
        return false;
    }
    if (i % ps != 0) {
        set_error("%s: row offset %lld is not a multiple of the panel height %lld", name, (long long)i,
                  (long long)ps);
        return false;
    }
    if (srows > 0 && scols > 0 && !s->pA) { set_error("%s: null pA", name); return false; }
    Mat<T> m{};
    m.data = s->pA ? static_cast<T*>(s->pA) + (i / ps) * ps * s->cn + j * ps : nullptr;
    m.stride = s->batch_stride;
    m.rows = int32_t(trans ? scols : srows);
    m.cols = int32_t(trans ? srows : scols);
    m.spacing = int32_t(ps * s->cn);
    m.panel = 1;
    m.trans = trans ? 1 : 0;
    out = m;
    return true;
}

bool batch_ok(int64_t batch) {
    if (batch < 0 || batch > 0x7fffffffLL) { set_error("batch %lld out of range", (long long)batch); return false; }
    return true;
}

bool sizes_ok(std::initializer_list<int> sizes, int first_arg) {
    int k = first_arg;
    for (int v : sizes) {
        if (v < 0) { set_error("argument %d: negative size %d", k, v); return false; }
        ++k;
    }
    return true;
}

// C and D name the same storage block: the in-place form BLASFEO allows.
bool same_block(const blast_blasfeo_mat_t* a, int ai, int aj, const blast_blasfeo_mat_t* b, int bi, int bj) {
    return a && b && a->pA == b->pA && a->batch_stride == b->batch_stride && a->cn == b->cn && ai == bi && aj == bj;
}

// Arguments: stream 1, m 2, n 3, k 4, alpha 5, sA 6, ai 7, aj 8, sB 9, bi 10,
// bj 11, beta 12, sC 13, ci 14, cj 15, sD 16, di 17, dj 18, batch 19.
template <typename T, bool TA, bool TB>
int gemm_fe(void* stream, int m, int n, int k, T alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
            const blast_blasfeo_mat_t* sB, int bi, int bj, T beta, const blast_blasfeo_mat_t* sC, int ci, int cj,
            const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch) {
    if (!sizes_ok({m, n, k}, 2)) return m < 0 ? -2 : (n < 0 ? -3 : -4);
    Mat<T> A, B, C, D;
    if (!block_view<T>(sA, ai, aj, TA ? k : m, TA ? m : k, TA, "sA", A)) return -6;
    if (!block_view<T>(sB, bi, bj, TB ? n : k, TB ? k : n, TB, "sB", B)) return -9;
    if (!block_view<T>(sC, ci, cj, m, n, false, "sC", C)) return -13;
    if (!block_view<T>(sD, di, dj, m, n, false, "sD", D)) return -16;
    if (!batch_ok(batch)) return -19;
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    if (m == 0 || n == 0 || batch == 0) return 0;
    return dispatch_gemm<T>(hipStream_t(stream), alpha, A, B, beta, C, D, batch);
}

// Arguments: stream 1, m 2, k 3, alpha 4, sA 5, ai 6, aj 7, sB 8, bi 9, bj 10,
// beta 11, sC 12, ci 13, cj 14, sD 15, di 16, dj 17, batch 18.
template <typename T>
int syrk_ln_fe(void* stream, int m, int k, T alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
               const blast_blasfeo_mat_t* sB, int bi, int bj, T beta, const blast_blasfeo_mat_t* sC, int ci, int cj,
               const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch) {
    if (!sizes_ok({m, k}, 2)) return m < 0 ? -2 : -3;
    Mat<T> A, B, C, D;
    if (!block_view<T>(sA, ai, aj, m, k, false, "sA", A)) return -5;
    if (!block_view<T>(sB, bi, bj, m, k, false, "sB", B)) return -8;
    if (!block_view<T>(sC, ci, cj, m, m, false, "sC", C)) return -12;
    if (!block_view<T>(sD, di, dj, m, m, false, "sD", D)) return -15;
    if (!batch_ok(batch)) return -18;
    (void)hipGetLastError();
    if (m == 0 || batch == 0) return 0;
    return launch_syrk2_tiles<T>(hipStream_t(stream), alpha, A, B, beta, C, D, batch);
}

// Arguments: stream 1, m 2, sC 3, ci 4, cj 5, sD 6, di 7, dj 8, batch 9, info 10.
template <typename T>
int potrf_l_fe(void* stream, int m, const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
               int di, int dj, int64_t batch, int32_t* info) {
    if (m < 0) { set_error("argument 2: negative size %d", m); return -2; }
    Mat<T> C, D;
    if (!block_view<T>(sC, ci, cj, m, m, false, "sC", C)) return -3;
    if (!block_view<T>(sD, di, dj, m, m, false, "sD", D)) return -6;
    if (!batch_ok(batch)) return -9;
    (void)hipGetLastError();
    if (batch == 0) return 0;
    return dispatch_potrf<T>(hipStream_t(stream), C, D, batch, info);
}

// Arguments: stream 1, m 2, k 3, sA 4, ai 5, aj 6, sB 7, bi 8, bj 9, sC 10,
// ci 11, cj 12, sD 13, di 14, dj 15, batch 16, info 17.
template <typename T>
int syrk_potrf_ln_fe(void* stream, int m, int k, const blast_blasfeo_mat_t* sA, int ai, int aj,
                     const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sC, int ci, int cj,
                     const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch, int32_t* info) {
    if (!sizes_ok({m, k}, 2)) return m < 0 ? -2 : -3;
    Mat<T> A, B, C, D;
    if (!block_view<T>(sA, ai, aj, m, k, false, "sA", A)) return -4;
    if (!block_view<T>(sB, bi, bj, m, k, false, "sB", B)) return -7;
    if (!block_view<T>(sC, ci, cj, m, m, false, "sC", C)) return -10;
    if (!block_view<T>(sD, di, dj, m, m, false, "sD", D)) return -13;
    if (!batch_ok(batch)) return -16;
    (void)hipGetLastError();
    if (batch == 0) return 0;
    hipStream_t const s = hipStream_t(stream);
    if (m > 0) {
        int const r = launch_syrk2_tiles<T>(s, T(1), A, B, T(1), C, D, batch);
        if (r) return r;
    }
    return dispatch_potrf<T>(s, D, D, batch, info);
}

// Arguments: stream 1, m 2, n 3, alpha 4, sA 5, ai 6, aj 7, sB 8, bi 9, bj 10,
// sD 11, di 12, dj 13, batch 14.  UPPER_T: rutn (A upper, used transposed).
template <typename T, bool UPPER_T>
int trmm_r_fe(void* stream, int m, int n, T alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
              const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sD, int di, int dj,
              int64_t batch) {
    if (!sizes_ok({m, n}, 2)) return m < 0 ? -2 : -3;
    Mat<T> A, B, D;
    // the kernel reads the lower triangle of the logical A: A itself (rlnn)
    // or the transpose of the stored upper A (rutn)
    if (!block_view<T>(sA, ai, aj, n, n, UPPER_T, "sA", A)) return -5;
    if (!block_view<T>(sB, bi, bj, m, n, false, "sB", B)) return -8;
    if (!block_view<T>(sD, di, dj, m, n, false, "sD", D)) return -11;
    if (!batch_ok(batch)) return -14;
    if (same_block(sB, bi, bj, sD, di, dj) && m > 0 && n > 0) {
        set_error("trmm: D must not overlap B");
        return -11;
    }
    (void)hipGetLastError();
    if (m == 0 || n == 0 || batch == 0) return 0;
    return dispatch_trmm<T>(hipStream_t(stream), 1, alpha, A, B, D, batch);
}

}  // namespace
}  // namespace blast_hip

using namespace blast_hip;

extern "C" {

#define BLASFEO_GEMM(P, T, SUF, TA, TB)                                                                             \
    int blast_hip_blasfeo_##P##gemm_##SUF(void* s, int m, int n, int k, T alpha, const blast_blasfeo_mat_t* sA,    \
                                          int ai, int aj, const blast_blasfeo_mat_t* sB, int bi, int bj, T beta,  \
                                          const blast_blasfeo_mat_t* sC, int ci, int cj,                          \
                                          const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch) {         \
        return gemm_fe<T, TA, TB>(s, m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, batch); \
    }
BLASFEO_GEMM(d, double, nn, false, false)
BLASFEO_GEMM(d, double, nt, false, true)
BLASFEO_GEMM(d, double, tn, true, false)
BLASFEO_GEMM(d, double, tt, true, true)
BLASFEO_GEMM(s, float, nn, false, false)
BLASFEO_GEMM(s, float, nt, false, true)
BLASFEO_GEMM(s, float, tn, true, false)
BLASFEO_GEMM(s, float, tt, true, true)
#undef BLASFEO_GEMM

int blast_hip_blasfeo_dsyrk_ln(void* s, int m, int k, double alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                               const blast_blasfeo_mat_t* sB, int bi, int bj, double beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch) {
    return syrk_ln_fe<double>(s, m, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, batch);
}
int blast_hip_blasfeo_ssyrk_ln(void* s, int m, int k, float alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                               const blast_blasfeo_mat_t* sB, int bi, int bj, float beta,
                               const blast_blasfeo_mat_t* sC, int ci, int cj, const blast_blasfeo_mat_t* sD,
                               int di, int dj, int64_t batch) {
    return syrk_ln_fe<float>(s, m, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj, batch);
}

int blast_hip_blasfeo_dpotrf_l(void* s, int m, const blast_blasfeo_mat_t* sC, int ci, int cj,
                               const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch, int32_t* info) {
    return potrf_l_fe<double>(s, m, sC, ci, cj, sD, di, dj, batch, info);
}
int blast_hip_blasfeo_spotrf_l(void* s, int m, const blast_blasfeo_mat_t* sC, int ci, int cj,
                               const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch, int32_t* info) {
    return potrf_l_fe<float>(s, m, sC, ci, cj, sD, di, dj, batch, info);
}

int blast_hip_blasfeo_dsyrk_dpotrf_ln(void* s, int m, int k, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                      const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sC,
                                      int ci, int cj, const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch,
                                      int32_t* info) {
    return syrk_potrf_ln_fe<double>(s, m, k, sA, ai, aj, sB, bi, bj, sC, ci, cj, sD, di, dj, batch, info);
}
int blast_hip_blasfeo_ssyrk_spotrf_ln(void* s, int m, int k, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                      const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sC,
                                      int ci, int cj, const blast_blasfeo_mat_t* sD, int di, int dj, int64_t batch,
                                      int32_t* info) {
    return syrk_potrf_ln_fe<float>(s, m, k, sA, ai, aj, sB, bi, bj, sC, ci, cj, sD, di, dj, batch, info);
}

int blast_hip_blasfeo_dtrmm_rlnn(void* s, int m, int n, double alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                 const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sD,
                                 int di, int dj, int64_t batch) {
    return trmm_r_fe<double, false>(s, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, batch);
}
int blast_hip_blasfeo_dtrmm_rutn(void* s, int m, int n, double alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                 const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sD,
                                 int di, int dj, int64_t batch) {
    return trmm_r_fe<double, true>(s, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, batch);
}
int blast_hip_blasfeo_strmm_rlnn(void* s, int m, int n, float alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                 const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sD,
                                 int di, int dj, int64_t batch) {
    return trmm_r_fe<float, false>(s, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, batch);
}
int blast_hip_blasfeo_strmm_rutn(void* s, int m, int n, float alpha, const blast_blasfeo_mat_t* sA, int ai, int aj,
                                 const blast_blasfeo_mat_t* sB, int bi, int bj, const blast_blasfeo_mat_t* sD,
                                 int di, int dj, int64_t batch) {
    return trmm_r_fe<float, true>(s, m, n, alpha, sA, ai, aj, sB, bi, bj, sD, di, dj, batch);
}

}  // extern "C"
