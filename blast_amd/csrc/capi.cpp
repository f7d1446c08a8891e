// C ABI entry points: argument validation (BLAST's size checks), algorithm
// dispatch, error reporting.  No allocation, no synchronisation.

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <numeric>
#include <vector>

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

// trsvLeftLower / trsvLeftUpper (dense/Trsv.hpp:15-68): one right-hand side
// per instance (b, x: n x 1), non-unit diagonal, on the trsm kernels.
template <typename T>
int trsv_entry(void* stream, int upper, const blast_matrix_t* A, const blast_matrix_t* b, const blast_matrix_t* x,
               int64_t batch) {
    if (!valid<T>(A, "A")) return -3;
    if (!valid<T>(b, "b")) return -4;
    if (!valid<T>(x, "x")) return -5;
    if (!valid_batch(batch)) return -6;
    if (b->cols != 1) { set_error("trsv: b must be a single column"); return -4; }
    if (A->rows != b->rows || A->cols != b->rows) { set_error("trsv: Invalid argument size"); return -3; }
    if (x->rows != b->rows || x->cols != 1) { set_error("trsv: Invalid argument size (x)"); return -5; }
    (void)hipGetLastError();
    return dispatch_trsm<T>(hipStream_t(stream), 0, upper != 0, 0, T(1), Mat<T>::from(*A), Mat<T>::from(*b),
                            Mat<T>::from(*x), batch);
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

template <typename T>
Mat<T> opt(const blast_matrix_t* m) {
    if (m) return Mat<T>::from(*m);
    Mat<T> r{};
    return r;
}

#define CHECK_SHAPE(cond, code, msg) \
    if (!(cond)) {                   \
        set_error("%s", msg);        \
        return code;                 \
    }

template <typename T>
int gemm_backward_entry(void* s, T alpha, const blast_matrix_t* A, const blast_matrix_t* B, T beta,
                        const blast_matrix_t* Db, const blast_matrix_t* Ab, const blast_matrix_t* Bb,
                        const blast_matrix_t* Cb, int64_t batch) {
    if (!valid<T>(A, "A")) return -3;
    if (!valid<T>(B, "B")) return -4;
    if (!valid<T>(Db, "Dbar")) return -6;
    if (Ab && !valid<T>(Ab, "Abar")) return -7;
    if (Bb && !valid<T>(Bb, "Bbar")) return -8;
    if (Cb && !valid<T>(Cb, "Cbar")) return -9;
    if (!valid_batch(batch)) return -10;
    CHECK_SHAPE(B->rows == A->cols, -4, "gemm_backward: Matrix sizes do not match (B)");
    CHECK_SHAPE(Db->rows == A->rows && Db->cols == B->cols, -6, "gemm_backward: Matrix sizes do not match (Dbar)");
    CHECK_SHAPE(!Ab || (Ab->rows == A->rows && Ab->cols == A->cols), -7, "gemm_backward: Abar shape");
    CHECK_SHAPE(!Bb || (Bb->rows == B->rows && Bb->cols == B->cols), -8, "gemm_backward: Bbar shape");
    CHECK_SHAPE(!Cb || (Cb->rows == Db->rows && Cb->cols == Db->cols), -9, "gemm_backward: Cbar shape");
    (void)hipGetLastError();
    return gemm_backward<T>(hipStream_t(s), alpha, Mat<T>::from(*A), Mat<T>::from(*B), beta, Mat<T>::from(*Db),
                            opt<T>(Ab), opt<T>(Bb), opt<T>(Cb), batch);
}

template <typename T>
int trsm_backward_entry(void* s, int upper, const blast_matrix_t* A, const blast_matrix_t* X,
                        const blast_matrix_t* Xb, const blast_matrix_t* Ab, const blast_matrix_t* Bb, int64_t batch) {
    if (!valid<T>(A, "A")) return -3;
    if (!valid<T>(X, "X")) return -4;
    if (!valid<T>(Xb, "Xbar")) return -5;
    if (Ab && !valid<T>(Ab, "Abar")) return -6;
    if (!valid<T>(Bb, "Bbar")) return -7;
    if (!valid_batch(batch)) return -8;
    CHECK_SHAPE(A->rows == A->cols && A->rows == X->rows, -3, "trsm_backward: Invalid argument size (A)");
    CHECK_SHAPE(Xb->rows == X->rows && Xb->cols == X->cols, -5, "trsm_backward: Xbar shape");
    CHECK_SHAPE(!Ab || (Ab->rows == A->rows && Ab->cols == A->cols), -6, "trsm_backward: Abar shape");
    CHECK_SHAPE(Bb->rows == X->rows && Bb->cols == X->cols, -7, "trsm_backward: Bbar shape");
    (void)hipGetLastError();
    return trsm_backward<T>(hipStream_t(s), upper, Mat<T>::from(*A), Mat<T>::from(*X), Mat<T>::from(*Xb), opt<T>(Ab),
                            Mat<T>::from(*Bb), batch);
}

template <typename T>
int potrf_backward_entry(void* s, const blast_matrix_t* L, const blast_matrix_t* Lb, const blast_matrix_t* Ab,
                         int64_t batch) {
    if (!valid<T>(L, "L")) return -2;
    if (!valid<T>(Lb, "Lbar")) return -3;
    if (!valid<T>(Ab, "Abar")) return -4;
    if (!valid_batch(batch)) return -5;
    CHECK_SHAPE(L->rows == L->cols, -2, "potrf_backward: L must be square");
    CHECK_SHAPE(Lb->rows == L->rows && Lb->cols == L->cols, -3, "potrf_backward: Lbar shape");
    CHECK_SHAPE(Ab->rows == L->rows && Ab->cols == L->cols, -4, "potrf_backward: Abar shape");
    (void)hipGetLastError();
    return potrf_backward<T>(hipStream_t(s), Mat<T>::from(*L), Mat<T>::from(*Lb), Mat<T>::from(*Ab), batch);
}

template <typename T>
int posv_backward_entry(void* s, int recompute, const blast_matrix_t* F, const blast_matrix_t* X,
                        const blast_matrix_t* Xb, const blast_matrix_t* Ab, const blast_matrix_t* Bb, int64_t batch,
                        int32_t* info) {
    if (!valid<T>(F, "F")) return -3;
    if (!valid<T>(X, "X")) return -4;
    if (!valid<T>(Xb, "Xbar")) return -5;
    if (!valid<T>(Ab, "Abar")) return -6;
    if (!valid<T>(Bb, "Bbar")) return -7;
    if (!valid_batch(batch)) return -8;
    CHECK_SHAPE(F->rows == F->cols && F->rows == X->rows, -3, "posv_backward: Invalid argument size (F)");
    CHECK_SHAPE(Xb->rows == X->rows && Xb->cols == X->cols, -5, "posv_backward: Xbar shape");
    CHECK_SHAPE(Ab->rows == F->rows && Ab->cols == F->cols, -6, "posv_backward: Abar shape");
    CHECK_SHAPE(Bb->rows == X->rows && Bb->cols == X->cols, -7, "posv_backward: Bbar shape");
    (void)hipGetLastError();
    return posv_backward<T>(hipStream_t(s), recompute, Mat<T>::from(*F), Mat<T>::from(*X), Mat<T>::from(*Xb),
                            Mat<T>::from(*Ab), Mat<T>::from(*Bb), batch, info);
}

template <typename T>
int getrs_entry(void* s, const blast_matrix_t* LU, const int64_t* ipiv, int64_t ipiv_stride, const blast_matrix_t* B,
                const blast_matrix_t* X, int64_t batch) {
    if (!valid<T>(LU, "LU")) return -2;
    if (LU->rows > 0 && !ipiv) { set_error("getrs: null ipiv"); return -3; }
    if (ipiv_stride < LU->rows) { set_error("getrs: ipiv_stride < n"); return -4; }
    if (!valid<T>(B, "B")) return -5;
    if (!valid<T>(X, "X")) return -6;
    if (!valid_batch(batch)) return -7;
    CHECK_SHAPE(LU->rows == LU->cols && B->rows == LU->rows, -5, "getrs: Invalid argument size (B)");
    CHECK_SHAPE(X->rows == B->rows && X->cols == B->cols, -6, "getrs: Invalid argument size (X)");
    (void)hipGetLastError();
    return getrs<T>(hipStream_t(s), Mat<T>::from(*LU), ipiv, ipiv_stride, Mat<T>::from(*B), Mat<T>::from(*X), batch);
}

template <typename T>
int getrs_backward_entry(void* s, const blast_matrix_t* LU, const int64_t* ipiv, int64_t ipiv_stride,
                         const blast_matrix_t* X, const blast_matrix_t* Xb, const blast_matrix_t* Ab,
                         const blast_matrix_t* Bb, int64_t batch) {
    if (!valid<T>(LU, "LU")) return -2;
    if (LU->rows > 0 && !ipiv) { set_error("getrs_backward: null ipiv"); return -3; }
    if (ipiv_stride < LU->rows) { set_error("getrs_backward: ipiv_stride < n"); return -4; }
    if (!valid<T>(X, "X")) return -5;
    if (!valid<T>(Xb, "Xbar")) return -6;
    if (Ab && !valid<T>(Ab, "Abar")) return -7;
    if (!valid<T>(Bb, "Bbar")) return -8;
    if (!valid_batch(batch)) return -9;
    CHECK_SHAPE(LU->rows == LU->cols && X->rows == LU->rows, -5, "getrs_backward: Invalid argument size (X)");
    CHECK_SHAPE(Xb->rows == X->rows && Xb->cols == X->cols, -6, "getrs_backward: Xbar shape");
    CHECK_SHAPE(!Ab || (Ab->rows == LU->rows && Ab->cols == LU->cols), -7, "getrs_backward: Abar shape");
    CHECK_SHAPE(Bb->rows == X->rows && Bb->cols == X->cols, -8, "getrs_backward: Bbar shape");
    (void)hipGetLastError();
    return getrs_backward<T>(hipStream_t(s), Mat<T>::from(*LU), ipiv, ipiv_stride, Mat<T>::from(*X),
                             Mat<T>::from(*Xb), opt<T>(Ab), Mat<T>::from(*Bb), batch);
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

template <typename T>
int convert_entry(void* stream, const blast_matrix_t* S, const blast_matrix_t* D, int64_t batch) {
    if (!valid<T>(S, "S")) return -2;
    if (!valid<T>(D, "D")) return -3;
    if (!valid_batch(batch)) return -4;
    if (S->rows != D->rows || S->cols != D->cols) { set_error("copy: Matrix sizes do not match"); return -3; }
    (void)hipGetLastError();
    return launch_convert<T>(hipStream_t(stream), Mat<T>::from(*S), Mat<T>::from(*D), batch);
}

// Shared checks of the mixed-precision and LU solve entry points.
int solve_shapes(const char* who, const blast_matrix_t* A, const blast_matrix_t* B, const blast_matrix_t* X, int ia,
                 int ib, int ix) {
    if (!valid<double>(A, "A")) return -ia;
    if (!valid<double>(B, "B")) return -ib;
    if (!valid<double>(X, "X")) return -ix;
    if (A->rows != A->cols) { set_error("%s: A must be square", who); return -ia; }
    if (B->rows != A->rows) { set_error("%s: Invalid argument size (B)", who); return -ib; }
    if (X->rows != B->rows || X->cols != B->cols) { set_error("%s: Invalid argument size (X)", who); return -ix; }
    return 0;
}

int mixed_params(const char* who, int precision, int itmax, double tol) {
    if (precision != 0 && precision != 1) { set_error("%s: precision must be 0 (FP32) or 1 (BF16)", who); return -2; }
    if (itmax < 0 || itmax > 1000) { set_error("%s: itmax out of range", who); return -3; }
    if (!(tol >= 0.0 && tol < 1.0)) { set_error("%s: tol must be in [0, 1)", who); return -4; }
    return 0;
}

}  // namespace
}  // namespace blast_hip

using namespace blast_hip;

namespace {

// Device-resident scratch of the allocation-free entry point (the FP32 factor
// of blast_hip_dsgesv_noalloc_batched): a file-scope device array, present on
// every GPU the library runs on until the process ends.  Host code uses only
// the per-device address from hipGetSymbolAddress, never the array's name.
__device__ __align__(256) float scratch[BLAST_HIP_NOALLOC_SCRATCH_FLOATS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's copy of `scratch`, resolved on the first
// call on that device; null (with the error recorded) on failure.
float* device_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("noalloc: no current HIP device (or device index >= %d)", kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

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
int blast_hip_dtrsv_batched(void* s, int upper, const blast_matrix_t* A, const blast_matrix_t* b,
                            const blast_matrix_t* x, int64_t batch) {
    return trsv_entry<double>(s, upper, A, b, x, batch);
}
int blast_hip_strsv_batched(void* s, int upper, const blast_matrix_t* A, const blast_matrix_t* b,
                            const blast_matrix_t* x, int64_t batch) {
    return trsv_entry<float>(s, upper, A, b, x, batch);
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
int blast_hip_sposv_batched(void* s, const blast_matrix_t* A, const blast_matrix_t* L, const blast_matrix_t* B,
                            const blast_matrix_t* X, int64_t batch, int32_t* info) {
    return posv_entry<float>(s, A, L, B, X, batch, info);
}

int blast_hip_dsposv_batched(void* s, int precision, int itmax, double tol, const blast_matrix_t* A,
                             const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info,
                             int32_t* iter) {
    return sposv_entry(s, precision, itmax, tol, A, B, X, batch, info, iter);
}

int blast_hip_dcopy_batched(void* s, const blast_matrix_t* S, const blast_matrix_t* D, int64_t batch) {
    return convert_entry<double>(s, S, D, batch);
}
int blast_hip_scopy_batched(void* s, const blast_matrix_t* S, const blast_matrix_t* D, int64_t batch) {
    return convert_entry<float>(s, S, D, batch);
}

int64_t blast_hip_mixed_work_elems(int64_t n) { return (n >= 1 && n <= 256) ? wg_work_elems(int(n)) : 0; }

int blast_hip_dsposv_ws_batched(void* s, int precision, int itmax, double tol, const blast_matrix_t* A,
                                const blast_matrix_t* B, const blast_matrix_t* X, float* work, int64_t batch,
                                int32_t* info, int32_t* iter) {
    if (int r = mixed_params("dsposv", precision, itmax, tol)) return r;
    if (int r = solve_shapes("dsposv", A, B, X, 5, 6, 7)) return r;
    if (!valid_batch(batch)) return -9;
    if (batch > 0 && !iter) { set_error("dsposv: iter is required"); return -11; }
    if (batch > 0 && !work && wg_covers(int(A->rows), int(X->cols)) && !(A->rows <= 128 && X->cols <= 4)) {
        set_error("dsposv: work is required for n > 128 or nrhs > 4");
        return -8;
    }
    (void)hipGetLastError();
    return dispatch_dsposv_ws(hipStream_t(s), precision, itmax, tol, Mat<double>::from(*A), Mat<double>::from(*B),
                              Mat<double>::from(*X), work, batch, info, iter);
}

int blast_hip_dgesv_batched(void* s, const blast_matrix_t* A, int64_t* ipiv, int64_t ipiv_stride,
                            const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch, int32_t* info) {
    if (int r = solve_shapes("dgesv", A, B, X, 2, 5, 6)) return r;
    if (A->trans) { set_error("dgesv: A must be column-major (trans = 0)"); return -2; }
    if (A->rows > 0 && !ipiv) { set_error("dgesv: null ipiv"); return -3; }
    if (ipiv_stride < A->rows) { set_error("dgesv: ipiv_stride < n"); return -4; }
    if (!valid_batch(batch)) return -7;
    (void)hipGetLastError();
    return dispatch_dgesv(hipStream_t(s), Mat<double>::from(*A), ipiv, ipiv_stride, Mat<double>::from(*B),
                          Mat<double>::from(*X), batch, info);
}

int blast_hip_uneven_max_n(void) { return uneven_max_n(); }

int blast_hip_uneven_order(int64_t count, const int32_t* n, int32_t* order) {
    if (count < 0 || count > 0x7fffffff - 4096) { set_error("uneven_order: count out of range"); return -1; }
    if (count > 0 && (!n || !order)) { set_error("uneven_order: null array"); return -2; }
    std::vector<int32_t> idx(static_cast<size_t>(count));
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return (n[x] + 15) / 16 > (n[y] + 15) / 16; });
    std::copy(idx.begin(), idx.end(), order);
    return 0;
}

int blast_hip_dposv_uneven_batched(void* s, const blast_solve_instance_t* inst, int64_t count, int32_t* queue,
                                   int32_t* info) {
    // the work counter bumps once past count per workgroup (at most 2048)
    if (count < 0 || count > 0x7fffffff - 4096) { set_error("posv_uneven: count out of range"); return -3; }
    if (count > 0 && !inst) { set_error("posv_uneven: null instance array"); return -2; }
    if (count > 0 && !queue) { set_error("posv_uneven: null queue"); return -4; }
    (void)hipGetLastError();
    return launch_posv_uneven(hipStream_t(s), inst, count, queue, info, 0);
}

int blast_hip_dposv_backward_uneven_batched(void* s, const blast_adjoint_instance_t* inst, int64_t count,
                                            int32_t* queue, int32_t* info) {
    if (count < 0 || count > 0x7fffffff - 4096) { set_error("posv_backward_uneven: count out of range"); return -3; }
    if (count > 0 && !inst) { set_error("posv_backward_uneven: null instance array"); return -2; }
    if (count > 0 && !queue) { set_error("posv_backward_uneven: null queue"); return -4; }
    (void)hipGetLastError();
    return launch_posv_backward_uneven(hipStream_t(s), inst, count, queue, info, 0);
}

int blast_hip_dsgesv_noalloc_batched(void* s, int precision, int itmax, double tol, const blast_matrix_t* A,
                                     int64_t* ipiv, int64_t ipiv_stride, const blast_matrix_t* B,
                                     const blast_matrix_t* X, int64_t batch, int32_t* info, int32_t* iter) {
    // the checks of blast_hip_dsgesv_batched (argument indices of this entry)
    if (int r = mixed_params("dsgesv_noalloc", precision, itmax, tol)) return r;
    if (int r = solve_shapes("dsgesv_noalloc", A, B, X, 5, 8, 9)) return r;
    if (A->trans) { set_error("dsgesv_noalloc: A must be column-major (trans = 0)"); return -5; }
    if (A->rows > 0 && !ipiv) { set_error("dsgesv_noalloc: null ipiv"); return -6; }
    if (ipiv_stride < A->rows) { set_error("dsgesv_noalloc: ipiv_stride < n"); return -7; }
    if (!valid_batch(batch)) return -10;
    if (batch > 0 && !iter) { set_error("dsgesv_noalloc: iter is required"); return -12; }
    float* work = nullptr;
    if (batch > 0 && wg_covers(int(A->rows), int(X->cols))) {
        int64_t const need = batch * wg_work_elems(int(A->rows));
        if (need > int64_t(BLAST_HIP_NOALLOC_SCRATCH_FLOATS)) {
            set_error("dsgesv_noalloc: batch %lld of n = %lld needs %lld scratch floats, more than the %lld available",
                      (long long)batch, (long long)A->rows, (long long)need,
                      (long long)BLAST_HIP_NOALLOC_SCRATCH_FLOATS);
            return -10;
        }
        work = device_scratch();
        if (!work) return BLAST_HIP_ERR_RUNTIME;
    }
    (void)hipGetLastError();
    return dispatch_dsgesv(hipStream_t(s), precision, itmax, tol, Mat<double>::from(*A), ipiv, ipiv_stride,
                           Mat<double>::from(*B), Mat<double>::from(*X), work, batch, info, iter);
}

int blast_hip_dsgesv_batched(void* s, int precision, int itmax, double tol, const blast_matrix_t* A, int64_t* ipiv,
                             int64_t ipiv_stride, const blast_matrix_t* B, const blast_matrix_t* X, float* work,
                             int64_t batch, int32_t* info, int32_t* iter) {
    if (int r = mixed_params("dsgesv", precision, itmax, tol)) return r;
    if (int r = solve_shapes("dsgesv", A, B, X, 5, 8, 9)) return r;
    if (A->trans) { set_error("dsgesv: A must be column-major (trans = 0)"); return -5; }
    if (A->rows > 0 && !ipiv) { set_error("dsgesv: null ipiv"); return -6; }
    if (ipiv_stride < A->rows) { set_error("dsgesv: ipiv_stride < n"); return -7; }
    if (!valid_batch(batch)) return -11;
    if (batch > 0 && !iter) { set_error("dsgesv: iter is required"); return -13; }
    if (batch > 0 && !work && wg_covers(int(A->rows), int(X->cols))) {
        set_error("dsgesv: work is required");
        return -10;
    }
    (void)hipGetLastError();
    return dispatch_dsgesv(hipStream_t(s), precision, itmax, tol, Mat<double>::from(*A), ipiv, ipiv_stride,
                           Mat<double>::from(*B), Mat<double>::from(*X), work, batch, info, iter);
}

#define BOTH(P, T)                                                                                                \
    int blast_hip_##P##gemm_backward_batched(void* s, T alpha, const blast_matrix_t* A, const blast_matrix_t* B,    \
                                             T beta, const blast_matrix_t* Db, const blast_matrix_t* Ab,            \
                                             const blast_matrix_t* Bb, const blast_matrix_t* Cb, int64_t batch) {   \
        return gemm_backward_entry<T>(s, alpha, A, B, beta, Db, Ab, Bb, Cb, batch);                                 \
    }                                                                                                               \
    int blast_hip_##P##trsm_backward_batched(void* s, int upper, const blast_matrix_t* A, const blast_matrix_t* X,  \
                                             const blast_matrix_t* Xb, const blast_matrix_t* Ab,                    \
                                             const blast_matrix_t* Bb, int64_t batch) {                             \
        return trsm_backward_entry<T>(s, upper, A, X, Xb, Ab, Bb, batch);                                           \
    }                                                                                                               \
    int blast_hip_##P##potrf_backward_batched(void* s, const blast_matrix_t* L, const blast_matrix_t* Lb,           \
                                              const blast_matrix_t* Ab, int64_t batch) {                            \
        return potrf_backward_entry<T>(s, L, Lb, Ab, batch);                                                        \
    }                                                                                                               \
    int blast_hip_##P##posv_backward_batched(void* s, int recompute, const blast_matrix_t* F,                       \
                                             const blast_matrix_t* X, const blast_matrix_t* Xb,                     \
                                             const blast_matrix_t* Ab, const blast_matrix_t* Bb, int64_t batch,     \
                                             int32_t* info) {                                                       \
        return posv_backward_entry<T>(s, recompute, F, X, Xb, Ab, Bb, batch, info);                                 \
    }                                                                                                               \
    int blast_hip_##P##getrs_batched(void* s, const blast_matrix_t* LU, const int64_t* ipiv, int64_t ipiv_stride,   \
                                     const blast_matrix_t* B, const blast_matrix_t* X, int64_t batch) {             \
        return getrs_entry<T>(s, LU, ipiv, ipiv_stride, B, X, batch);                                               \
    }                                                                                                               \
    int blast_hip_##P##getrs_backward_batched(void* s, const blast_matrix_t* LU, const int64_t* ipiv,               \
                                              int64_t ipiv_stride, const blast_matrix_t* X,                         \
                                              const blast_matrix_t* Xb, const blast_matrix_t* Ab,                   \
                                              const blast_matrix_t* Bb, int64_t batch) {                            \
        return getrs_backward_entry<T>(s, LU, ipiv, ipiv_stride, X, Xb, Ab, Bb, batch);                             \
    }
BOTH(d, double)
BOTH(s, float)

}  // extern "C"
