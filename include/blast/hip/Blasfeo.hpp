// C++ counterpart of BLAST's BLASFEO wrapper (include/blasfeo/blasfeo_api/
// Gemm.hpp, Syrk.hpp, Potrf.hpp, SyrkPotrf.hpp, Trmm.hpp in the reference):
// the same function names and argument order, on batches of BLASFEO
// panel-major matrices in device memory, over the blast_hip_blasfeo_* C ABI.
//
//   blasfeo::gemm_nt(m, n, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj)
// becomes
//   blast::hip::blasfeo::gemm_nt(m, n, k, alpha, A, ai, aj, B, bi, bj, beta, C, ci, cj, D, di, dj)
// with A..D BlasfeoBatch views; the batch count is D's.  Errors throw as in
// blast/hip/Blast.hpp (std::invalid_argument for bad arguments).
#pragma once

#include <blast/hip/Blast.hpp>

#include <type_traits>

namespace blast::hip::blasfeo {

// Non-owning view of `count` BLASFEO matrices (m x n, padded column count
// cn, panel height PanelSize_v<T>) stored `stride` elements apart from
// `pA` in device memory: the fields of blasfeo_dmat / blasfeo_smat.
template <typename T>
struct BlasfeoBatch {
    T* pA = nullptr;
    std::int64_t m = 0, n = 0, cn = 0, stride = 0, count = 0;

    BlasfeoBatch() = default;
    BlasfeoBatch(T* p, std::int64_t rows, std::int64_t cols, std::int64_t padded_cols, std::int64_t batch_stride,
                 std::int64_t batch)
        : pA(p), m(rows), n(cols), cn(padded_cols), stride(batch_stride), count(batch) {}

    // Elements of one instance (panel rows rounded up), a 64-byte multiple.
    static std::int64_t storage(std::int64_t rows, std::int64_t padded_cols) {
        std::int64_t const ps = std::int64_t(PanelSize_v<std::remove_const_t<T>>);
        std::int64_t const elems = (rows + ps - 1) / ps * ps * padded_cols;
        std::int64_t const align = std::int64_t(64 / sizeof(T));
        return (elems + align - 1) / align * align;
    }
    // Offset of element (i, j) inside one instance.
    std::int64_t index(std::int64_t i, std::int64_t j) const {
        std::int64_t const ps = std::int64_t(PanelSize_v<std::remove_const_t<T>>);
        return i / ps * ps * cn + i % ps + j * ps;
    }
    blast_blasfeo_mat_t desc() const {
        return blast_blasfeo_mat_t{const_cast<void*>(static_cast<void const*>(pA)), m, n, cn,
                                   std::int64_t(PanelSize_v<std::remove_const_t<T>>), stride};
    }
    operator BlasfeoBatch<T const>() const { return BlasfeoBatch<T const>(pA, m, n, cn, stride, count); }
};

namespace detail {
// input operand: a non-deduced context, so T comes from alpha / the output
template <typename T>
using In = std::type_identity_t<BlasfeoBatch<T const>> const&;
}

#define BLAST_HIP_BLASFEO_GEMM(NAME)                                                                                \
    template <typename T>                                                                                          \
    void NAME(int m, int n, int k, T alpha, detail::In<T> sA, int ai, int aj, detail::In<T> sB, int bi, int bj,    \
              T beta, detail::In<T> sC, int ci, int cj, BlasfeoBatch<T> const& sD, int di, int dj,                 \
              hipStream_t s = nullptr) {                                                                           \
        static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>);                                     \
        auto const a = sA.desc(), b = sB.desc(), c = sC.desc(), d = sD.desc();                                     \
        if constexpr (std::is_same_v<T, double>)                                                                   \
            check(blast_hip_blasfeo_d##NAME(s, m, n, k, alpha, &a, ai, aj, &b, bi, bj, beta, &c, ci, cj, &d, di,  \
                                            dj, sD.count));                                                        \
        else                                                                                                       \
            check(blast_hip_blasfeo_s##NAME(s, m, n, k, alpha, &a, ai, aj, &b, bi, bj, beta, &c, ci, cj, &d, di,  \
                                            dj, sD.count));                                                        \
    }
// blasfeo::gemm_nn / gemm_nt / gemm_tn (blasfeo_api/Gemm.hpp); gemm_tt as BLASFEO's blasfeo_dgemm_tt
BLAST_HIP_BLASFEO_GEMM(gemm_nn)
BLAST_HIP_BLASFEO_GEMM(gemm_nt)
BLAST_HIP_BLASFEO_GEMM(gemm_tn)
BLAST_HIP_BLASFEO_GEMM(gemm_tt)
#undef BLAST_HIP_BLASFEO_GEMM

// blasfeo::syrk_ln (blasfeo_api/Syrk.hpp): lower triangle of D = beta*C + alpha*A*B^T.
template <typename T>
void syrk_ln(int m, int k, T alpha, detail::In<T> sA, int ai, int aj, detail::In<T> sB, int bi, int bj, T beta,
             detail::In<T> sC, int ci, int cj, BlasfeoBatch<T> const& sD, int di, int dj, hipStream_t s = nullptr) {
    auto const a = sA.desc(), b = sB.desc(), c = sC.desc(), d = sD.desc();
    if constexpr (std::is_same_v<T, double>)
        check(blast_hip_blasfeo_dsyrk_ln(s, m, k, alpha, &a, ai, aj, &b, bi, bj, beta, &c, ci, cj, &d, di, dj,
                                         sD.count));
    else
        check(blast_hip_blasfeo_ssyrk_ln(s, m, k, alpha, &a, ai, aj, &b, bi, bj, beta, &c, ci, cj, &d, di, dj,
                                         sD.count));
}

// blasfeo::potrf (blasfeo_api/Potrf.hpp, blasfeo_dpotrf_l): D = chol(C).
// info (device, optional): per instance 0 or the first non-positive pivot.
template <typename T>
void potrf(int m, detail::In<T> sC, int ci, int cj, BlasfeoBatch<T> const& sD, int di, int dj,
           std::int32_t* info = nullptr, hipStream_t s = nullptr) {
    auto const c = sC.desc(), d = sD.desc();
    if constexpr (std::is_same_v<T, double>)
        check(blast_hip_blasfeo_dpotrf_l(s, m, &c, ci, cj, &d, di, dj, sD.count, info));
    else
        check(blast_hip_blasfeo_spotrf_l(s, m, &c, ci, cj, &d, di, dj, sD.count, info));
}

// blasfeo::syrk_potrf (blasfeo_api/SyrkPotrf.hpp, blasfeo_dsyrk_dpotrf_ln): D = chol(C + A*B^T).
template <typename T>
void syrk_potrf(int m, int k, detail::In<T> sA, int ai, int aj, detail::In<T> sB, int bi, int bj, detail::In<T> sC,
                int ci, int cj, BlasfeoBatch<T> const& sD, int di, int dj, std::int32_t* info = nullptr,
                hipStream_t s = nullptr) {
    auto const a = sA.desc(), b = sB.desc(), c = sC.desc(), d = sD.desc();
    if constexpr (std::is_same_v<T, double>)
        check(blast_hip_blasfeo_dsyrk_dpotrf_ln(s, m, k, &a, ai, aj, &b, bi, bj, &c, ci, cj, &d, di, dj, sD.count,
                                                info));
    else
        check(blast_hip_blasfeo_ssyrk_spotrf_ln(s, m, k, &a, ai, aj, &b, bi, bj, &c, ci, cj, &d, di, dj, sD.count,
                                                info));
}

// blasfeo::trmm_rlnn / trmm_rutn (blasfeo_api/Trmm.hpp): D = alpha*B*A (A lower)
// or D = alpha*B*A^T (A upper).
template <typename T>
void trmm_rlnn(int m, int n, T alpha, detail::In<T> sA, int ai, int aj, detail::In<T> sB, int bi, int bj,
               BlasfeoBatch<T> const& sD, int di, int dj, hipStream_t s = nullptr) {
    auto const a = sA.desc(), b = sB.desc(), d = sD.desc();
    if constexpr (std::is_same_v<T, double>)
        check(blast_hip_blasfeo_dtrmm_rlnn(s, m, n, alpha, &a, ai, aj, &b, bi, bj, &d, di, dj, sD.count));
    else
        check(blast_hip_blasfeo_strmm_rlnn(s, m, n, alpha, &a, ai, aj, &b, bi, bj, &d, di, dj, sD.count));
}
template <typename T>
void trmm_rutn(int m, int n, T alpha, detail::In<T> sA, int ai, int aj, detail::In<T> sB, int bi, int bj,
               BlasfeoBatch<T> const& sD, int di, int dj, hipStream_t s = nullptr) {
    auto const a = sA.desc(), b = sB.desc(), d = sD.desc();
    if constexpr (std::is_same_v<T, double>)
        check(blast_hip_blasfeo_dtrmm_rutn(s, m, n, alpha, &a, ai, aj, &b, bi, bj, &d, di, dj, sD.count));
    else
        check(blast_hip_blasfeo_strmm_rutn(s, m, n, alpha, &a, ai, aj, &b, bi, bj, &d, di, dj, sD.count));
}

}  // namespace blast::hip::blasfeo
