// C ABI entry points: argument validation (BLAST's size checks), algorithm
// dispatch, error reporting.  No allocation, no synchronisation.

#include <cstdarg>
#include <cstdio>

#include "common.hpp"
#include "dispatch.hpp"
#include "launch.hpp"

namespace blast_hip {

namespace {
thread_local char g_error[512] = "";
int g_algo = BLAST_HIP_ALGO_AUTO;
}  // namespace

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return BLAST_HIP_ERR_RUNTIME;
    }
    return 0;
}

int algorithm() { return g_algo; }

namespace {

constexpr int64_t kMaxDim = 1 << 20;

// Validates one operand descriptor; returns false (and records why) when the
// storage it describes is inconsistent.
template <typename T>
bool valid(const blast_matrix_t* m, const char* name) {
    if (!m) { set_error("%s: null descriptor", name); return false; }
    if (m->rows < 0 || m->cols < 0 || m->rows > kMaxDim || m->cols > kMaxDim) {
        set_error("%s: bad size %lld x %lld", name, (long long)m->rows, (long long)m->cols);
        return false;
    }
    if (m->batch_stride < 0) { set_error("%s: negative batch_stride", name); return false; }
    int64_t const sr = m->trans ? m->cols : m->rows, sc = m->trans ? m->rows : m->cols;
    if (sr == 0 || sc == 0) return true;
    if (!m->data) { set_error("%s: null data", name); return false; }
    int64_t const need = m->panel ? int64_t(PanelSize<T>::value) * sc : sr;
    if (m->spacing < need || m->spacing > (int64_t(1) << 30)) {
        set_error("%s: spacing %lld < required %lld", name, (long long)m->spacing, (long long)need);
        return false;
    }
    return true;
}

bool valid_batch(int64_t batch) {
    if (batch < 0 || batch > 0x7fffffffLL) { set_error("batch %lld out of range", (long long)batch); return false; }
    return true;
}

template <typename T>
int gemm_entry(void* stream, T alpha, const blast_matrix_t* A, const blast_matrix_t* B, T beta,
               const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch) {
    if (!valid<T>(A, "A")) return -3;
    if (!valid<T>(B, "B")) return -4;
    if (!valid<T>(C, "C")) return -6;
    if (!valid<T>(D, "D")) return -7;
    if (!valid_batch(batch)) return -8;
    // algorithm/Gemm.hpp:87-94: "Matrix sizes do not match"
    if (B->rows != A->cols) { set_error("gemm: Matrix sizes do not match (B)"); return -4; }
    if (C->rows != A->rows || C->cols != B->cols) { set_error("gemm: Matrix sizes do not match (C)"); return -6; }
    if (D->rows != A->rows || D->cols != B->cols) { set_error("gemm: Matrix sizes do not match (D)"); return -7; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_gemm<T>(hipStream_t(stream), alpha, Mat<T>::from(*A), Mat<T>::from(*B), beta,
                            Mat<T>::from(*C), Mat<T>::from(*D), batch);
}

template <typename T>
int syrk_entry(void* stream, T alpha, const blast_matrix_t* A, T beta, const blast_matrix_t* C,
               const blast_matrix_t* D, int64_t batch) {
    if (!valid<T>(A, "A")) return -3;
    if (!valid<T>(C, "C")) return -5;
    if (!valid<T>(D, "D")) return -6;
    if (!valid_batch(batch)) return -7;
    if (C->rows != A->rows || C->cols != A->rows) { set_error("syrk: Matrix sizes do not match (C)"); return -5; }
    if (D->rows != A->rows || D->cols != A->rows) { set_error("syrk: Matrix sizes do not match (D)"); return -6; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_syrk<T>(hipStream_t(stream), alpha, Mat<T>::from(*A), beta, Mat<T>::from(*C),
                            Mat<T>::from(*D), batch);
}

template <typename T>
int trmm_entry(void* stream, int side_right, int upper, T alpha, const blast_matrix_t* A,
               const blast_matrix_t* B, const blast_matrix_t* C, int64_t batch) {
    // algorithm/Trmm.hpp: only left-upper and right-lower are implemented.
    if (side_right ? upper : !upper) {
        set_error("trmm: only left/upper and right/lower are implemented");
        return -3;
    }
    if (!valid<T>(A, "A")) return -5;
    if (!valid<T>(B, "B")) return -6;
    if (!valid<T>(C, "C")) return -7;
    if (!valid_batch(batch)) return -8;
    int64_t const na = side_right ? B->cols : B->rows;
    if (A->rows != na || A->cols != na) { set_error("trmm: Matrix sizes do not match (A)"); return -5; }
    if (C->rows != B->rows || C->cols != B->cols) { set_error("trmm: Matrix sizes do not match (C)"); return -7; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_trmm<T>(hipStream_t(stream), side_right, alpha, Mat<T>::from(*A), Mat<T>::from(*B),
                            Mat<T>::from(*C), batch);
}

template <typename T>
int trsm_entry(void* stream, int side_right, int upper, int unit, T alpha, const blast_matrix_t* A,
               const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch) {
    if (!valid<T>(A, "A")) return -6;
    if (!valid<T>(B, "B")) return -7;
    if (!valid<T>(X, "X")) return -8;
    if (!valid_batch(batch)) return -9;
    int64_t const na = side_right ? B->cols : B->rows;
    // dense/Trsm.hpp:244-251: "Invalid argument size"
    if (A->rows != na || A->cols != na) { set_error("trsm: Invalid argument size (A)"); return -6; }
    if (X->rows != B->rows || X->cols != B->cols) { set_error("trsm: Invalid argument size (X)"); return -8; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_trsm<T>(hipStream_t(stream), side_right, upper, unit, alpha, Mat<T>::from(*A),
                            Mat<T>::from(*B), Mat<T>::from(*X), batch);
}

template <typename T>
int potrf_entry(void* stream, const blast_matrix_t* A, const blast_matrix_t* L, int64_t batch, int32_t* info) {
    if (!valid<T>(A, "A")) return -2;
    if (!valid<T>(L, "L")) return -3;
    if (!valid_batch(batch)) return -4;
    // dense/Potrf.hpp:78-85: "Invalid matrix size"
    if (A->cols > A->rows) { set_error("potrf: Invalid matrix size (columns > rows)"); return -2; }
    if (L->rows != A->rows || L->cols != A->cols) { set_error("potrf: Invalid matrix size (L)"); return -3; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_potrf<T>(hipStream_t(stream), Mat<T>::from(*A), Mat<T>::from(*L), batch, info);
}

template <typename T>
int getrf_entry(void* stream, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                int32_t* info) {
    if (!valid<T>(A, "A")) return -2;
    int64_t const k = A->rows < A->cols ? A->rows : A->cols;
    if (k > 0 && !ipiv) { set_error("getrf: null ipiv"); return -3; }
    if (ipiv_stride < k) { set_error("getrf: ipiv_stride < min(M, N)"); return -4; }
    if (!valid_batch(batch)) return -5;
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_getrf<T>(hipStream_t(stream), Mat<T>::from(*A), ipiv, ipiv_stride, batch, info);
}

template <typename T>
int posv_entry(void* stream, const blast_matrix_t* A, const blast_matrix_t* L, const blast_matrix_t* B,
               const blast_matrix_t* X, int64_t batch, int32_t* info) {
    if (!valid<T>(A, "A")) return -2;
    if (!valid<T>(L, "L")) return -3;
    if (!valid<T>(B, "B")) return -4;
    if (!valid<T>(X, "X")) return -5;
    if (!valid_batch(batch)) return -6;
    if (A->rows != A->cols) { set_error("posv: A must be square"); return -2; }
    if (L->rows != A->rows || L->cols != A->cols) { set_error("posv: Invalid matrix size (L)"); return -3; }
    if (B->rows != A->rows) { set_error("posv: Invalid argument size (B)"); return -4; }
    if (X->rows != B->rows || X->cols != B->cols) { set_error("posv: Invalid argument size (X)"); return -5; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_posv<T>(hipStream_t(stream), Mat<T>::from(*A), Mat<T>::from(*L), Mat<T>::from(*B),
                            Mat<T>::from(*X), batch, info);
}

int sposv_entry(void* stream, int precision, int itmax, double tol, const blast_matrix_t* A,
                const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info, int32_t* iter) {
    if (precision != 0 && precision != 1) { set_error("dsposv: precision must be 0 (FP32) or 1 (BF16)"); return -2; }
    if (itmax < 0 || itmax > 1000) { set_error("dsposv: itmax out of range"); return -3; }
    if (!(tol >= 0.0 && tol < 1.0)) { set_error("dsposv: tol must be in [0, 1)"); return -4; }
    if (!valid<double>(A, "A")) return -5;
    if (!valid<double>(B, "B")) return -6;
    if (!valid<double>(X, "X")) return -7;
    if (!valid_batch(batch)) return -8;
    if (batch > 0 && !iter) { set_error("dsposv: iter is required"); return -10; }
    if (A->rows != A->cols) { set_error("dsposv: A must be square"); return -5; }
    if (B->rows != A->rows) { set_error("dsposv: Invalid argument size (B)"); return -6; }
    if (X->rows != B->rows || X->cols != B->cols) { set_error("dsposv: Invalid argument size (X)"); return -7; }
    (void)hipGetLastError();  // drop a stale error left by an earlier, unrelated HIP call
    return dispatch_dsposv(hipStream_t(stream), precision, itmax, tol, Mat<double>::from(*A), Mat<double>::from(*B),
                           Mat<double>::from(*X), batch, info, iter);
}

}  // namespace
}  // namespace blast_hip

using namespace blast_hip;

extern "C" {

int blast_hip_set_algorithm(int algo) {
    if (algo != BLAST_HIP_ALGO_AUTO && algo != BLAST_HIP_ALGO_REFERENCE) return -1;
    g_algo = algo;
    return 0;
}
int blast_hip_get_algorithm(void) { return g_algo; }
const char* blast_hip_last_error(void) { return g_error; }
const char* blast_hip_version(void) { return "blast-hip 0.1 (gfx950)"; }

int blast_hip_dgemm_batched(void* s, double alpha, const blast_matrix_t* A, const blast_matrix_t* B, double beta,
                            const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch) {
    return gemm_entry<double>(s, alpha, A, B, beta, C, D, batch);
}
int blast_hip_sgemm_batched(void* s, float alpha, const blast_matrix_t* A, const blast_matrix_t* B, float beta,
                            const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch) {
    return gemm_entry<float>(s, alpha, A, B, beta, C, D, batch);
}
int blast_hip_dsyrk_lower_batched(void* s, double alpha, const blast_matrix_t* A, double beta,
                                  const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch) {
    return syrk_entry<double>(s, alpha, A, beta, C, D, batch);
}
int blast_hip_ssyrk_lower_batched(void* s, float alpha, const blast_matrix_t* A, float beta,
                                  const blast_matrix_t* C, const blast_matrix_t* D, int64_t batch) {
    return syrk_entry<float>(s, alpha, A, beta, C, D, batch);
}
int blast_hip_dtrmm_batched(void* s, int side_right, int upper, double alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* C, int64_t batch) {
    return trmm_entry<double>(s, side_right, upper, alpha, A, B, C, batch);
}
int blast_hip_strmm_batched(void* s, int side_right, int upper, float alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* C, int64_t batch) {
    return trmm_entry<float>(s, side_right, upper, alpha, A, B, C, batch);
}
int blast_hip_dtrsm_batched(void* s, int side_right, int upper, int unit, double alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch) {
    return trsm_entry<double>(s, side_right, upper, unit, alpha, A, B, X, batch);
}
int blast_hip_strsm_batched(void* s, int side_right, int upper, int unit, float alpha, const blast_matrix_t* A,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch) {
    return trsm_entry<float>(s, side_right, upper, unit, alpha, A, B, X, batch);
}
int blast_hip_dpotrf_batched(void* s, const blast_matrix_t* A, const blast_matrix_t* L, int64_t batch,
                             int32_t* info) {
    return potrf_entry<double>(s, A, L, batch, info);
}
int blast_hip_spotrf_batched(void* s, const blast_matrix_t* A, const blast_matrix_t* L, int64_t batch,
                             int32_t* info) {
    return potrf_entry<float>(s, A, L, batch, info);
}
int blast_hip_dgetrf_batched(void* s, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                             int32_t* info) {
    return getrf_entry<double>(s, A, ipiv, ipiv_stride, batch, info);
}
int blast_hip_sgetrf_batched(void* s, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride, int64_t batch,
                             int32_t* info) {
    return getrf_entry<float>(s, A, ipiv, ipiv_stride, batch, info);
}
int blast_hip_dposv_batched(void* s, const blast_matrix_t* A, const blast_matrix_t* L, const blast_matrix_t* B,
                            const blast_matrix_t* X, int64_t batch, int32_t* info) {
    return posv_entry<double>(s, A, L, B, X, batch, info);
}

int blast_hip_dsposv_batched(void* s, int precision, int itmax, double tol, const blast_matrix_t* A,
                             const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info,
                             int32_t* iter) {
    return sposv_entry(s, precision, itmax, tol, A, B, X, batch, info, iter);
}

}  // extern "C"
