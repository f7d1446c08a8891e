// Blast.hpp - C++20 host front end of the MI355X batched dense-kernel layer.
//
// Mirrors the public surface of BLAST (mkatliar/blast) on the path it
// accelerates:
//   * storage types with BLAST's byte layout: DynamicMatrix / StaticMatrix
//     (dense, include/blast/math/dense/DynamicMatrix.hpp:26-151,
//     StaticMatrix.hpp:27-166) and DynamicPanelMatrix / StaticPanelMatrix
//     (panel, include/blast/math/panel/DynamicPanelMatrix.hpp:34-190,
//     StaticPanelMatrix.hpp:35-185), plus trans() views;
//   * the entry points gemm, syrkLower, trmm, trsm<UpLo, UNIT>, potrf and
//     getrf with BLAST's argument order and exceptions
//     (std::invalid_argument for size mismatches and exact-zero getrf pivots,
//     std::logic_error for unimplemented trmm forms);
//   * batched overloads on device memory (hip::Batch views), the drop-in for
//     many independent instances: stream-ordered, no allocation, no sync.
// The single-instance overloads stage the host matrices through device
// memory, run the batched kernel with batch = 1 and copy the results back.
// Every computation goes through the C ABI in blast_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../blast_hip.h"

namespace blast {

// include/blast/math/StorageOrder.hpp, UpLo.hpp, Side.hpp
enum StorageOrder : bool { rowMajor = false, columnMajor = true };
enum class UpLo : bool { Lower = false, Upper = true };
enum class Side : bool { Left = false, Right = true };

// SIMD width of BLAST's AVX2 build = panel size (panel/PanelSize.hpp:25-26).
template <typename T> inline constexpr std::size_t PanelSize_v = 0;
template <> inline constexpr std::size_t PanelSize_v<double> = 4;
template <> inline constexpr std::size_t PanelSize_v<float> = 8;

constexpr std::size_t nextMultiple(std::size_t x, std::size_t m) { return (x + m - 1) / m * m; }

namespace detail {

inline constexpr std::size_t kAlignment = 64;  // system/CacheLine.hpp:16

template <typename T>
struct AlignedStore {
    T* p = nullptr;
    std::size_t size = 0;
    explicit AlignedStore(std::size_t n) : p(new (std::align_val_t{kAlignment}) T[n]{}), size(n) {}
    AlignedStore(AlignedStore const&) = delete;
    AlignedStore& operator=(AlignedStore const&) = delete;
    AlignedStore(AlignedStore&& o) noexcept : p(std::exchange(o.p, nullptr)), size(o.size) {}
    ~AlignedStore() {
        if (p) ::operator delete[](p, std::align_val_t{kAlignment});
    }
};

}  // namespace detail

// Common storage interface: storage is column-major (SO = columnMajor) or the
// transpose of it (rowMajor), dense (idx = i + spacing*j) or panel
// (idx = i/PS*spacing + i%PS + j*PS) in storage coordinates.
template <typename T, bool SO, bool PANEL>
class StorageBase {
public:
    using ElementType = T;
    static constexpr bool storageOrder = SO;
    static constexpr bool isPanel = PANEL;

    std::size_t rows() const noexcept { return m_; }
    std::size_t columns() const noexcept { return n_; }
    std::size_t spacing() const noexcept { return spacing_; }
    std::size_t storageSize() const noexcept { return store_.size; }
    T* data() noexcept { return store_.p; }
    T const* data() const noexcept { return store_.p; }

    std::size_t offset(std::size_t i, std::size_t j) const noexcept {
        std::size_t const r = SO == columnMajor ? i : j, c = SO == columnMajor ? j : i;
        constexpr std::size_t PS = PanelSize_v<T>;
        return PANEL ? r / PS * spacing_ + r % PS + c * PS : r + spacing_ * c;
    }
    // Row of logical element (i, j) in the (column-major) storage.
    std::size_t storageRow(std::size_t i, std::size_t j) const noexcept { return SO == columnMajor ? i : j; }
    // The whole storage block (what the host overloads stage to the device).
    T* storageData() noexcept { return store_.p; }
    T const* storageData() const noexcept { return store_.p; }
    T& operator()(std::size_t i, std::size_t j) noexcept { return store_.p[offset(i, j)]; }
    T const& operator()(std::size_t i, std::size_t j) const noexcept { return store_.p[offset(i, j)]; }

    // C ABI descriptor of this matrix's storage placed at `device` (instance 0).
    blast_matrix_t descriptor(void* device, std::int64_t batch_stride) const noexcept {
        return blast_matrix_t{device, std::int64_t(m_), std::int64_t(n_), std::int64_t(spacing_), batch_stride,
                              PANEL ? 1 : 0, SO == columnMajor ? 0 : 1};
    }

protected:
    StorageBase(std::size_t m, std::size_t n)
        : m_(m), n_(n), spacing_(compute_spacing(m, n)), store_(compute_size(m, n)) {}

    static std::size_t compute_spacing(std::size_t m, std::size_t n) {
        std::size_t const sr = SO == columnMajor ? m : n, sc = SO == columnMajor ? n : m;
        constexpr std::size_t PS = PanelSize_v<T>;
        return PANEL ? PS * sc : nextMultiple(sr, PS);
    }
    static std::size_t compute_size(std::size_t m, std::size_t n) {
        std::size_t const sr = SO == columnMajor ? m : n, sc = SO == columnMajor ? n : m;
        constexpr std::size_t PS = PanelSize_v<T>;
        std::size_t const s = PANEL ? nextMultiple(sr, PS) * sc : nextMultiple(sr, PS) * sc;
        return s > 0 ? s : 1;
    }

    std::size_t m_, n_, spacing_;
    detail::AlignedStore<T> store_;
};

// Dense matrix with runtime size (dense/DynamicMatrix.hpp:26-151).
template <typename T, bool SO = columnMajor>
class DynamicMatrix : public StorageBase<T, SO, false> {
public:
    DynamicMatrix(std::size_t m, std::size_t n) : StorageBase<T, SO, false>(m, n) {}
};

// Dense matrix with compile-time size (dense/StaticMatrix.hpp:27-166).
template <typename T, std::size_t M, std::size_t N, bool SO = columnMajor>
class StaticMatrix : public StorageBase<T, SO, false> {
public:
    StaticMatrix() : StorageBase<T, SO, false>(M, N) {}
};

// Panel matrix with runtime size (panel/DynamicPanelMatrix.hpp:34-190).
template <typename T, bool SO = columnMajor>
class DynamicPanelMatrix : public StorageBase<T, SO, true> {
public:
    DynamicPanelMatrix(std::size_t m, std::size_t n) : StorageBase<T, SO, true>(m, n) {}
};

// Panel matrix with compile-time size (panel/StaticPanelMatrix.hpp:35-185).
template <typename T, std::size_t M, std::size_t N, bool SO = columnMajor>
class StaticPanelMatrix : public StorageBase<T, SO, true> {
public:
    StaticPanelMatrix() : StorageBase<T, SO, true>(M, N) {}
};

template <typename MT>
concept HostMatrix = requires(MT const& m) {
    typename MT::ElementType;
    { m.rows() } -> std::convertible_to<std::size_t>;
    { m.descriptor(nullptr, 0) } -> std::same_as<blast_matrix_t>;
};

// trans() of a matrix: same storage, transposed view (expressions/MatTransExpr.hpp:139-160).
template <HostMatrix MT>
struct Trans {
    MT& m;
    using ElementType = typename MT::ElementType;
    static constexpr bool isPanel = std::remove_const_t<MT>::isPanel;
    std::size_t rows() const noexcept { return m.columns(); }
    std::size_t columns() const noexcept { return m.rows(); }
    std::size_t storageSize() const noexcept { return m.storageSize(); }
    auto* data() const noexcept { return m.data(); }
    auto* storageData() const noexcept { return m.storageData(); }
    std::size_t offset(std::size_t i, std::size_t j) const noexcept { return m.offset(j, i); }
    std::size_t storageRow(std::size_t i, std::size_t j) const noexcept { return m.storageRow(j, i); }
    decltype(auto) operator()(std::size_t i, std::size_t j) const noexcept { return m(j, i); }
    blast_matrix_t descriptor(void* device, std::int64_t stride) const noexcept {
        blast_matrix_t d = m.descriptor(device, stride);
        std::swap(d.rows, d.cols);
        d.trans = !d.trans;
        return d;
    }
};
template <HostMatrix MT>
Trans<MT> trans(MT& m) {
    return Trans<MT>{m};
}
template <HostMatrix MT>
Trans<MT const> trans(MT const& m) {
    return Trans<MT const>{m};
}

// ------------------------------------------------------------------ views
// Every view shares its root's storage: offset(i, j) is the element offset of
// logical element (i, j) from the root storage block, storageRow(i, j) its row
// in that storage, and descriptor(device, stride) describes the view inside a
// device copy of the whole root block whose base is `device`.

// submatrix(M, i, j, m, n): rows i..i+m-1, columns j..j+n-1 of M, a
// non-owning window (views/Submatrix.hpp:22-297, views/submatrix/Panel.hpp:
// 43-282).  Out-of-range blocks throw std::invalid_argument("Invalid
// submatrix specification"); in panel storage the block's first storage row
// must lie on a panel boundary, else std::logic_error, as BLAST's
// PanelSubmatrix (views/submatrix/Panel.hpp:96-102).
template <HostMatrix MT>
struct Submatrix {
    using Root = std::remove_const_t<MT>;
    using ElementType = typename MT::ElementType;
    static constexpr bool isPanel = Root::isPanel;
    using T = std::remove_const_t<ElementType>;

    Submatrix(MT& mm, std::size_t i, std::size_t j, std::size_t rows, std::size_t cols)
        : m(mm), i0(i), j0(j), m_(rows), n_(cols) {
        if (i + rows > mm.rows() || j + cols > mm.columns())
            throw std::invalid_argument("Invalid submatrix specification");
        if (isPanel && rows > 0 && cols > 0 && mm.storageRow(i, j) % PanelSize_v<T> != 0)
            throw std::logic_error("Submatrices of a panel matrix which are not aligned on a panel boundary are "
                                   "not supported");
    }
    std::size_t rows() const noexcept { return m_; }
    std::size_t columns() const noexcept { return n_; }
    std::size_t row() const noexcept { return i0; }
    std::size_t column() const noexcept { return j0; }
    std::size_t storageSize() const noexcept { return m.storageSize(); }
    auto* storageData() const noexcept { return m.storageData(); }
    auto* data() const noexcept { return m.storageData() + offset(0, 0); }
    std::size_t offset(std::size_t i, std::size_t j) const noexcept { return m.offset(i0 + i, j0 + j); }
    std::size_t storageRow(std::size_t i, std::size_t j) const noexcept { return m.storageRow(i0 + i, j0 + j); }
    decltype(auto) operator()(std::size_t i, std::size_t j) const noexcept { return m(i0 + i, j0 + j); }
    blast_matrix_t descriptor(void* device, std::int64_t stride) const noexcept {
        blast_matrix_t d = m.descriptor(device, stride);
        d.data = static_cast<T*>(device) + offset(0, 0);
        d.rows = std::int64_t(m_), d.cols = std::int64_t(n_);
        return d;
    }

    MT& m;
    std::size_t i0, j0, m_, n_;
};

template <HostMatrix MT>
Submatrix<MT> submatrix(MT& m, std::size_t i, std::size_t j, std::size_t rows, std::size_t cols) {
    return Submatrix<MT>(m, i, j, rows, cols);
}
template <HostMatrix MT>
Submatrix<MT const> submatrix(MT const& m, std::size_t i, std::size_t j, std::size_t rows, std::size_t cols) {
    return Submatrix<MT const>(m, i, j, rows, cols);
}

// One row or column of a matrix as a 1 x n (or n x 1) matrix view
// (views/Row.hpp:27-275, math/PanelMatrix.hpp:29-40).  Where the vector runs
// across panels' columns of panel storage at a row inside a panel (a row of
// a column-major panel matrix whose index is not a multiple of the panel
// size), its elements are panel-size apart: it is described to the kernels
// as a dense matrix with that spacing.
template <HostMatrix MT>
struct VectorView {
    using Root = std::remove_const_t<MT>;
    using ElementType = typename MT::ElementType;
    static constexpr bool isPanel = Root::isPanel;
    using T = std::remove_const_t<ElementType>;

    VectorView(MT& mm, std::size_t index, bool is_row) : m(mm), k(index), isRow(is_row) {
        if (index >= (is_row ? mm.rows() : mm.columns())) throw std::invalid_argument("Invalid vector index");
    }
    std::size_t rows() const noexcept { return isRow ? 1 : m.rows(); }
    std::size_t columns() const noexcept { return isRow ? m.columns() : 1; }
    std::size_t storageSize() const noexcept { return m.storageSize(); }
    auto* storageData() const noexcept { return m.storageData(); }
    std::size_t offset(std::size_t i, std::size_t j) const noexcept {
        return isRow ? m.offset(k, j) : m.offset(i, k);
    }
    std::size_t storageRow(std::size_t i, std::size_t j) const noexcept {
        return isRow ? m.storageRow(k, j) : m.storageRow(i, k);
    }
    decltype(auto) operator()(std::size_t i, std::size_t j) const noexcept { return isRow ? m(k, j) : m(i, k); }
    blast_matrix_t descriptor(void* device, std::int64_t stride) const noexcept {
        T* const base = static_cast<T*>(device) + offset(0, 0);
        if (!isPanel || storageRow(0, 0) % PanelSize_v<T> == 0) {  // a window like submatrix()
            blast_matrix_t d = m.descriptor(device, stride);
            d.data = base;
            d.rows = std::int64_t(rows()), d.cols = std::int64_t(columns());
            return d;
        }
        // elements `step` apart: a dense matrix with one storage row (a row
        // view; a column view is its transpose)
        std::size_t const len = isRow ? m.columns() : m.rows();
        std::size_t const step = len > 1 ? (isRow ? offset(0, 1) : offset(1, 0)) - offset(0, 0) : 1;
        return blast_matrix_t{base, std::int64_t(rows()), std::int64_t(columns()), std::int64_t(step), stride, 0,
                              isRow ? 0 : 1};
    }

    MT& m;
    std::size_t k;
    bool isRow;
};

template <HostMatrix MT>
VectorView<MT> row(MT& m, std::size_t i) {
    return VectorView<MT>(m, i, true);
}
template <HostMatrix MT>
VectorView<MT const> row(MT const& m, std::size_t i) {
    return VectorView<MT const>(m, i, true);
}
template <HostMatrix MT>
VectorView<MT> column(MT& m, std::size_t j) {
    return VectorView<MT>(m, j, false);
}
template <HostMatrix MT>
VectorView<MT const> column(MT const& m, std::size_t j) {
    return VectorView<MT const>(m, j, false);
}

// ptr(M, i, j): matrix pointer to element (i, j), the argument of the
// pointer-form gemm (dense/DynamicMatrixPointer.hpp:245-262,
// panel/DynamicPanelMatrixPointer.hpp:274-287).
template <HostMatrix MT>
struct MatrixPointer {
    MT& m;
    std::size_t i, j;
};
template <HostMatrix MT>
MatrixPointer<MT> ptr(MT& m, std::size_t i = 0, std::size_t j = 0) {
    return MatrixPointer<MT>{m, i, j};
}
template <HostMatrix MT>
MatrixPointer<MT const> ptr(MT const& m, std::size_t i = 0, std::size_t j = 0) {
    return MatrixPointer<MT const>{m, i, j};
}

namespace hip {

// Maps a C ABI status to BLAST's exception types.
inline void check(int status) {
    if (status == 0) return;
    std::string const msg = blast_hip_last_error();
    if (status == BLAST_HIP_ERR_RUNTIME) throw std::runtime_error(msg);
    if (msg.find("only left/upper and right/lower") != std::string::npos) throw std::logic_error(msg);
    throw std::invalid_argument(msg);
}

inline void check_hip(hipError_t e) {
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
}

// Non-owning view of a batch of matrices in device memory: `count` instances
// `stride` elements apart, each with the storage of a matrix type MT
// (layout and size taken from a host prototype, or given explicitly).
template <typename T>
struct Batch {
    T* data = nullptr;
    std::int64_t count = 0;
    blast_matrix_t desc{};

    Batch() = default;
    Batch(T* d, std::int64_t n, blast_matrix_t const& shape) : data(d), count(n), desc(shape) {
        desc.data = const_cast<void*>(static_cast<void const*>(d));
    }

    // Batch of instances shaped like host matrix `proto`, packed back to back
    // (stride = proto.storageSize() rounded to 64 bytes).
    template <class MT>
    static Batch like(MT const& proto, T* d, std::int64_t n) {
        std::int64_t const stride = std::int64_t(nextMultiple(proto.storageSize(), detail::kAlignment / sizeof(T)));
        return Batch(d, n, proto.descriptor(d, stride));
    }
    Batch t() const {
        Batch r = *this;
        std::swap(r.desc.rows, r.desc.cols);
        r.desc.trans = !r.desc.trans;
        return r;
    }
    // Sub-block rows i..i+m-1, columns j..j+n-1 of every instance (the
    // batched submatrix(); same alignment rule for panel storage).
    Batch sub(std::int64_t i, std::int64_t j, std::int64_t m, std::int64_t n) const {
        if (i < 0 || j < 0 || m < 0 || n < 0 || i + m > desc.rows || j + n > desc.cols)
            throw std::invalid_argument("Invalid submatrix specification");
        using E = std::remove_const_t<T>;
        std::int64_t const r = desc.trans ? j : i, c = desc.trans ? i : j, ps = std::int64_t(PanelSize_v<E>);
        if (desc.panel && m > 0 && n > 0 && r % ps != 0)
            throw std::logic_error("Submatrices of a panel matrix which are not aligned on a panel boundary are "
                                   "not supported");
        std::int64_t const off = desc.panel ? r / ps * desc.spacing + c * ps : r + desc.spacing * c;
        Batch out = *this;
        out.data = data + off;
        out.desc.data = const_cast<void*>(static_cast<void const*>(out.data));
        out.desc.rows = m, out.desc.cols = n;
        return out;
    }
    operator Batch<T const>() const { return Batch<T const>(data, count, desc); }
    std::int64_t stride() const noexcept { return desc.batch_stride; }
    std::size_t bytes() const noexcept { return std::size_t(count) * std::size_t(desc.batch_stride) * sizeof(T); }
};

// Owning device allocation (host-side convenience; the batched entry points
// themselves never allocate).
template <typename T>
class DeviceArray {
public:
    explicit DeviceArray(std::size_t n) : n_(n) { check_hip(hipMalloc(&p_, (n ? n : 1) * sizeof(T))); }
    DeviceArray(DeviceArray const&) = delete;
    DeviceArray& operator=(DeviceArray const&) = delete;
    ~DeviceArray() {
        if (p_) (void)hipFree(p_);
    }
    T* get() const noexcept { return p_; }
    std::size_t size() const noexcept { return n_; }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

}  // namespace hip

// ------------------------------------------------------------------ batched overloads (device memory)
// Each is the batched replacement of the BLAST function named in its comment.

// blast::gemm(alpha, A, B, beta, C, D)  (algorithm/Gemm.hpp:80-97)
template <typename T>
void gemm(T alpha, hip::Batch<T const> A, hip::Batch<T const> B, T beta, hip::Batch<T const> C, hip::Batch<T> D,
          hipStream_t s = nullptr) {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>);
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dgemm_batched(s, alpha, &A.desc, &B.desc, beta, &C.desc, &D.desc, D.count));
    else
        hip::check(blast_hip_sgemm_batched(s, alpha, &A.desc, &B.desc, beta, &C.desc, &D.desc, D.count));
}

// blast::gemm(A, B, C, D): D = A*B + C  (algorithm/Gemm.hpp:113-118)
template <typename T>
void gemm(hip::Batch<T const> A, hip::Batch<T const> B, hip::Batch<T const> C, hip::Batch<T> D,
          hipStream_t s = nullptr) {
    gemm<T>(T(1), A, B, T(1), C, D, s);
}

// blast::syrkLower(alpha, A, beta, C, D)  (dense/Syrk.hpp:15-19)
template <typename T>
void syrkLower(T alpha, hip::Batch<T const> A, T beta, hip::Batch<T const> C, hip::Batch<T> D,
               hipStream_t s = nullptr) {
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dsyrk_lower_batched(s, alpha, &A.desc, beta, &C.desc, &D.desc, D.count));
    else
        hip::check(blast_hip_ssyrk_lower_batched(s, alpha, &A.desc, beta, &C.desc, &D.desc, D.count));
}

// blast::trmm(alpha, A, uplo, diag, B, C): C = alpha*A*B  (algorithm/Trmm.hpp:260-280)
template <typename T>
void trmm(T alpha, hip::Batch<T const> A, UpLo uplo, bool unit, hip::Batch<T const> B, hip::Batch<T> C,
          hipStream_t s = nullptr) {
    if (unit) throw std::logic_error("Unit-triangular matrices support not implemented in trmm()");
    int const up = uplo == UpLo::Upper;
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dtrmm_batched(s, 0, up, alpha, &A.desc, &B.desc, &C.desc, C.count));
    else
        hip::check(blast_hip_strmm_batched(s, 0, up, alpha, &A.desc, &B.desc, &C.desc, C.count));
}

// blast::trmm(alpha, B, A, uplo, diag, C): C = alpha*B*A  (algorithm/Trmm.hpp:290-310)
template <typename T>
void trmm(T alpha, hip::Batch<T const> B, hip::Batch<T const> A, UpLo uplo, bool unit, hip::Batch<T> C,
          hipStream_t s = nullptr) {
    if (unit) throw std::logic_error("Unit-triangular matrices support not implemented in trmm()");
    int const up = uplo == UpLo::Upper;
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dtrmm_batched(s, 1, up, alpha, &A.desc, &B.desc, &C.desc, C.count));
    else
        hip::check(blast_hip_strmm_batched(s, 1, up, alpha, &A.desc, &B.desc, &C.desc, C.count));
}

// blast::trsm<UPLO, UNIT>(A, B, X): A X = B  (dense/Trsm.hpp:17-66)
template <UpLo UPLO, bool UNIT, typename T>
void trsm(hip::Batch<T const> A, hip::Batch<T const> B, hip::Batch<T> X, hipStream_t s = nullptr) {
    int const up = UPLO == UpLo::Upper;
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dtrsm_batched(s, 0, up, UNIT, 1.0, &A.desc, &B.desc, &X.desc, X.count));
    else
        hip::check(blast_hip_strsm_batched(s, 0, up, UNIT, 1.f, &A.desc, &B.desc, &X.desc, X.count));
}

// reference::trsm(A, uplo, diag, X, alpha, B) / trsm(X, A, uplo, diag, alpha, B)
// (reference/Trsm.hpp:164, 206): left and right solves with alpha.
template <typename T>
void trsm(Side side, UpLo uplo, bool unit, T alpha, hip::Batch<T const> A, hip::Batch<T const> B, hip::Batch<T> X,
          hipStream_t s = nullptr) {
    int const right = side == Side::Right, up = uplo == UpLo::Upper;
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dtrsm_batched(s, right, up, unit, alpha, &A.desc, &B.desc, &X.desc, X.count));
    else
        hip::check(blast_hip_strsm_batched(s, right, up, unit, alpha, &A.desc, &B.desc, &X.desc, X.count));
}

// blast::trsvLeftLower / trsvLeftUpper(A, b, x): A x = b with one right-hand
// side per instance, b and x as n x 1 matrices (dense/Trsv.hpp:15-68).
template <UpLo UPLO, typename T>
void trsv(hip::Batch<T const> A, hip::Batch<T const> b, hip::Batch<T> x, hipStream_t s = nullptr) {
    int const up = UPLO == UpLo::Upper;
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dtrsv_batched(s, up, &A.desc, &b.desc, &x.desc, x.count));
    else
        hip::check(blast_hip_strsv_batched(s, up, &A.desc, &b.desc, &x.desc, x.count));
}

// blast::potrf(A, L)  (dense/Potrf.hpp:66-106, panel/Potrf.hpp:59-106).
// info (device, optional): first non-positive pivot column per instance.
template <typename T>
void potrf(hip::Batch<T const> A, hip::Batch<T> L, std::int32_t* info = nullptr, hipStream_t s = nullptr) {
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dpotrf_batched(s, &A.desc, &L.desc, L.count, info));
    else
        hip::check(blast_hip_spotrf_batched(s, &A.desc, &L.desc, L.count, info));
}

// blast::getrf(A, ipiv)  (dense/Getrf.hpp:52-119; row-major :141-162).
// ipiv (device): min(M, N) entries per instance, ipiv_stride apart.
template <typename T>
void getrf(hip::Batch<T> A, std::int64_t* ipiv, std::int64_t ipiv_stride, std::int32_t* info = nullptr,
           hipStream_t s = nullptr) {
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dgetrf_batched(s, &A.desc, ipiv, ipiv_stride, A.count, info));
    else
        hip::check(blast_hip_sgetrf_batched(s, &A.desc, ipiv, ipiv_stride, A.count, info));
}

// potrf + trsm<Lower> + trsm<Upper>(trans(L)), fused (LAPACK dposv / sposv contract).
inline void posv(hip::Batch<double const> A, hip::Batch<double> L, hip::Batch<double const> B, hip::Batch<double> X,
                 std::int32_t* info = nullptr, hipStream_t s = nullptr) {
    hip::check(blast_hip_dposv_batched(s, &A.desc, &L.desc, &B.desc, &X.desc, X.count, info));
}
inline void posv(hip::Batch<float const> A, hip::Batch<float> L, hip::Batch<float const> B, hip::Batch<float> X,
                 std::int32_t* info = nullptr, hipStream_t s = nullptr) {
    hip::check(blast_hip_sposv_batched(s, &A.desc, &L.desc, &B.desc, &X.desc, X.count, info));
}

// Mixed precision (FP32 or BF16 factor) + FP64 refinement (LAPACK dsposv contract).
inline void sposv(int precision, hip::Batch<double> A, hip::Batch<double const> B, hip::Batch<double> X,
                  std::int32_t* iter, std::int32_t* info = nullptr, int itmax = 30, double tol = 0.0,
                  hipStream_t s = nullptr) {
    hip::check(blast_hip_dsposv_batched(s, precision, itmax, tol, &A.desc, &B.desc, &X.desc, X.count, info, iter));
}

// Storage conversion D := S between any two layouts of the same logical size:
// the batched form of assigning a panel matrix to a dense one and back
// (expressions/PanelMatrix.hpp:81-167, :170-239).
template <typename T>
void assign(hip::Batch<T const> S, hip::Batch<T> D, hipStream_t s = nullptr) {
    if constexpr (std::is_same_v<T, double>)
        hip::check(blast_hip_dcopy_batched(s, &S.desc, &D.desc, D.count));
    else
        hip::check(blast_hip_scopy_batched(s, &S.desc, &D.desc, D.count));
}

// getrf + getrs (LAPACK dgesv): A holds its LU afterwards, ipiv its pivots.
inline void gesv(hip::Batch<double> A, std::int64_t* ipiv, std::int64_t ipiv_stride, hip::Batch<double const> B,
                 hip::Batch<double> X, std::int32_t* info = nullptr, hipStream_t s = nullptr) {
    hip::check(blast_hip_dgesv_batched(s, &A.desc, ipiv, ipiv_stride, &B.desc, &X.desc, X.count, info));
}

// FP32 factor buffer (floats per instance) of the n <= 256 mixed-precision solves.
inline std::int64_t mixedWorkElems(std::int64_t n) { return blast_hip_mixed_work_elems(n); }

// Mixed precision Cholesky solve for n <= 256, nrhs <= 8 with the FP32 factor
// in `work` (LAPACK dsposv with SWORK).
inline void sposv(int precision, hip::Batch<double> A, hip::Batch<double const> B, hip::Batch<double> X, float* work,
                  std::int32_t* iter, std::int32_t* info = nullptr, int itmax = 30, double tol = 0.0,
                  hipStream_t s = nullptr) {
    hip::check(
        blast_hip_dsposv_ws_batched(s, precision, itmax, tol, &A.desc, &B.desc, &X.desc, work, X.count, info, iter));
}

// Mixed precision LU solve with FP64 refinement (LAPACK dsgesv).
inline void sgesv(int precision, hip::Batch<double> A, std::int64_t* ipiv, std::int64_t ipiv_stride,
                  hip::Batch<double const> B, hip::Batch<double> X, float* work, std::int32_t* iter,
                  std::int32_t* info = nullptr, int itmax = 30, double tol = 0.0, hipStream_t s = nullptr) {
    hip::check(blast_hip_dsgesv_batched(s, precision, itmax, tol, &A.desc, ipiv, ipiv_stride, &B.desc, &X.desc, work,
                                        X.count, info, iter));
}

namespace hip {

// Uneven batch of Riccati recursions (LQR backward passes): owns the plan,
// the device buffers and the queue; setInputs() takes per-instance A, B, Q,
// R (dense column-major, ld = rows) in one host array laid out by the plan.
class RiccatiBatch {
public:
    RiccatiBatch(std::vector<std::int32_t> const& nx, std::vector<std::int32_t> const& nu, int horizon)
        : horizon_(horizon), inst_(nx.size()), order_(nx.size()) {
        std::int64_t ni = 0, no = 0, nw = 0;
        check(blast_hip_riccati_plan(std::int64_t(nx.size()), nx.data(), nu.data(), horizon, inst_.data(),
                                     order_.data(), &ni, &no, &nw));
        in_ = std::make_unique<DeviceArray<double>>(std::size_t(ni));
        out_ = std::make_unique<DeviceArray<double>>(std::size_t(no));
        work_ = std::make_unique<DeviceArray<double>>(std::size_t(nw));
        dinst_ = std::make_unique<DeviceArray<blast_riccati_instance_t>>(inst_.size());
        queue_ = std::make_unique<DeviceArray<std::int32_t>>(inst_.size() + 1);
        info_ = std::make_unique<DeviceArray<std::int32_t>>(inst_.size());
        std::vector<std::int32_t> q(inst_.size() + 1, 0);
        std::copy(order_.begin(), order_.end(), q.begin() + 1);
        check_hip(hipMemcpy(dinst_->get(), inst_.data(), inst_.size() * sizeof(blast_riccati_instance_t),
                            hipMemcpyHostToDevice));
        check_hip(hipMemcpy(queue_->get(), q.data(), q.size() * sizeof(std::int32_t), hipMemcpyHostToDevice));
        nin_ = std::size_t(ni), nout_ = std::size_t(no);
    }
    std::vector<blast_riccati_instance_t> const& instances() const noexcept { return inst_; }
    std::size_t inputSize() const noexcept { return nin_; }
    std::size_t outputSize() const noexcept { return nout_; }
    void setInputs(std::vector<double> const& host) {
        if (host.size() != nin_) throw std::invalid_argument("RiccatiBatch: input size does not match the plan");
        check_hip(hipMemcpy(in_->get(), host.data(), nin_ * sizeof(double), hipMemcpyHostToDevice));
    }
    // One persistent launch; no allocation, no memset: capturable in a hipGraph.
    void run(hipStream_t s = nullptr) {
        check(blast_hip_driccati_batched(s, horizon_, dinst_->get(), std::int64_t(inst_.size()), in_->get(),
                                         out_->get(), work_->get(), queue_->get(), info_->get()));
    }
    void autotune(int reps = 3, hipStream_t s = nullptr) {
        check(blast_hip_riccati_autotune(s, horizon_, dinst_->get(), inst_.data(), std::int64_t(inst_.size()),
                                         in_->get(), out_->get(), work_->get(), queue_->get(), order_.data(), reps,
                                         nullptr));
    }
    std::vector<double> outputs() const {
        std::vector<double> host(nout_);
        check_hip(hipMemcpy(host.data(), out_->get(), nout_ * sizeof(double), hipMemcpyDeviceToHost));
        return host;
    }
    std::vector<std::int32_t> info() const {
        std::vector<std::int32_t> host(inst_.size());
        check_hip(hipMemcpy(host.data(), info_->get(), host.size() * sizeof(std::int32_t), hipMemcpyDeviceToHost));
        return host;
    }

private:
    int horizon_;
    std::vector<blast_riccati_instance_t> inst_;
    std::vector<std::int32_t> order_;
    std::size_t nin_ = 0, nout_ = 0;
    std::unique_ptr<DeviceArray<double>> in_, out_, work_;
    std::unique_ptr<DeviceArray<blast_riccati_instance_t>> dinst_;
    std::unique_ptr<DeviceArray<std::int32_t>> queue_, info_;
};

}  // namespace hip

// ------------------------------------------------------------------ single-instance drop-ins (host memory)
namespace detail {

// Device copy of one host matrix (or trans view), copied back on demand.
template <class MT>
class Staged {
public:
    using T = std::remove_const_t<typename std::remove_reference_t<MT>::ElementType>;
    explicit Staged(MT& m) : m_(m), buf_(m.storageSize()) {
        hip::check_hip(hipMemcpy(buf_.get(), m.storageData(), m.storageSize() * sizeof(T), hipMemcpyHostToDevice));
    }
    // (a view's descriptor points inside the staged block: its data pointer is kept)
    hip::Batch<T> batch() const {
        blast_matrix_t const d = m_.descriptor(buf_.get(), 0);
        return hip::Batch<T>(static_cast<T*>(d.data), 1, d);
    }
    operator hip::Batch<T const>() const { return batch(); }
    void copy_back() {
        hip::check_hip(hipMemcpy(const_cast<T*>(m_.storageData()), buf_.get(), m_.storageSize() * sizeof(T),
                                 hipMemcpyDeviceToHost));
    }

private:
    MT& m_;
    hip::DeviceArray<T> buf_;
};

}  // namespace detail

template <typename ST1, class MT1, class MT2, typename ST2, class MT3, class MT4>
    requires HostMatrix<MT1> && HostMatrix<MT2> && HostMatrix<MT3> && HostMatrix<MT4>
void gemm(ST1 alpha, MT1 const& A, MT2 const& B, ST2 beta, MT3 const& C, MT4& D) {
    using T = std::remove_const_t<typename MT4::ElementType>;
    // algorithm/Gemm.hpp:87-94
    if (B.rows() != A.columns() || C.rows() != A.rows() || C.columns() != B.columns() || D.rows() != A.rows() ||
        D.columns() != B.columns())
        throw std::invalid_argument("Matrix sizes do not match");
    detail::Staged<MT1 const> a(A);
    detail::Staged<MT2 const> b(B);
    detail::Staged<MT3 const> c(C);
    detail::Staged<MT4> d(D);
    gemm<T>(T(alpha), a.batch(), b.batch(), T(beta), c.batch(), d.batch());
    d.copy_back();
}

template <class MT1, class MT2, class MT3, class MT4>
    requires HostMatrix<MT1> && HostMatrix<MT2> && HostMatrix<MT3> && HostMatrix<MT4>
void gemm(MT1 const& A, MT2 const& B, MT3 const& C, MT4& D) {
    using T = std::remove_const_t<typename MT4::ElementType>;
    gemm(T(1), A, B, T(1), C, D);
}

// Pointer form: the M x N block of D at ptr D gets alpha * (M x K block at
// ptr A) * (K x N block at ptr B) + beta * (M x N block at ptr C)
// (algorithm/Gemm.hpp:41-62).
template <typename ST1, class MA, class MB, typename ST2, class MC, class MD>
void gemm(std::size_t M, std::size_t N, std::size_t K, ST1 alpha, MatrixPointer<MA> A, MatrixPointer<MB> B, ST2 beta,
          MatrixPointer<MC> C, MatrixPointer<MD> D) {
    auto a = submatrix(A.m, A.i, A.j, M, K);
    auto b = submatrix(B.m, B.i, B.j, K, N);
    auto c = submatrix(C.m, C.i, C.j, M, N);
    auto d = submatrix(D.m, D.i, D.j, M, N);
    gemm(alpha, a, b, beta, c, d);
}

// C = alpha A B, A triangular (algorithm/Trmm.hpp:260-277): only the upper
// form is implemented (std::logic_error otherwise, as BLAST).
template <typename ST, class MT1, class MT2, class MT3>
    requires HostMatrix<MT1> && HostMatrix<MT2> && HostMatrix<MT3>
void trmm(ST alpha, MT1 const& A, UpLo uplo, bool diag, MT2 const& B, MT3& C) {
    using T = std::remove_const_t<typename MT3::ElementType>;
    std::size_t const M = B.rows(), N = B.columns();
    if (A.rows() != M || A.columns() != M) throw std::invalid_argument("Matrix sizes do not match");
    if (C.rows() != M || C.columns() != N) throw std::invalid_argument("Matrix sizes do not match");
    if (uplo != UpLo::Upper) throw std::logic_error("Lower-triangular matrices support not implemented in trmm()");
    detail::Staged<MT1 const> a(A);
    detail::Staged<MT2 const> b(B);
    detail::Staged<MT3> c(C);
    trmm<T>(T(alpha), a.batch(), uplo, diag, b.batch(), c.batch());
    c.copy_back();
}

// C = alpha B A, A triangular (algorithm/Trmm.hpp:280-310): only the lower
// form is implemented (std::logic_error otherwise, as BLAST).
template <typename ST, class MTB, class MTA, class MTC>
    requires HostMatrix<MTB> && HostMatrix<MTA> && HostMatrix<MTC>
void trmm(ST alpha, MTB const& B, MTA const& A, UpLo uplo, bool diag, MTC& C) {
    using T = std::remove_const_t<typename MTC::ElementType>;
    std::size_t const M = B.rows(), N = B.columns();
    if (A.rows() != N || A.columns() != N) throw std::invalid_argument("Matrix sizes do not match");
    if (C.rows() != M || C.columns() != N) throw std::invalid_argument("Matrix sizes do not match");
    if (uplo != UpLo::Lower) throw std::logic_error("Upper-triangular matrices support not implemented in trmm()");
    detail::Staged<MTB const> b(B);
    detail::Staged<MTA const> a(A);
    detail::Staged<MTC> c(C);
    trmm<T>(T(alpha), b.batch(), a.batch(), uplo, diag, c.batch());
    c.copy_back();
}

namespace detail {

// x = A^{-1} b for one instance, b and x host vectors (x is resized to n).
template <UpLo UPLO, class MT, typename T>
void trsv_host(MT const& A, std::vector<T> const& b, std::vector<T>& x) {
    std::size_t const N = b.size();
    if (A.rows() != N || A.columns() != N) throw std::invalid_argument("Invalid argument size");  // dense/Trsv.hpp:25
    x.resize(N);
    if (N == 0) return;
    Staged<MT const> a(A);
    hip::DeviceArray<T> db(N), dx(N);
    hip::check_hip(hipMemcpy(db.get(), b.data(), N * sizeof(T), hipMemcpyHostToDevice));
    blast_matrix_t const vb{db.get(), std::int64_t(N), 1, std::int64_t(N), std::int64_t(N), 0, 0};
    blast_matrix_t const vx{dx.get(), std::int64_t(N), 1, std::int64_t(N), std::int64_t(N), 0, 0};
    trsv<UPLO, T>(a.batch(), hip::Batch<T const>(db.get(), 1, vb), hip::Batch<T>(dx.get(), 1, vx));
    hip::check_hip(hipMemcpy(x.data(), dx.get(), N * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace detail

// Solve A x = b, A lower triangular (dense/Trsv.hpp:15-36).
template <class MT, typename T>
    requires HostMatrix<MT>
void trsvLeftLower(MT const& A, std::vector<T> const& b, std::vector<T>& x) {
    detail::trsv_host<UpLo::Lower>(A, b, x);
}

// Solve A x = b, A upper triangular (dense/Trsv.hpp:41-68).
template <class MT, typename T>
    requires HostMatrix<MT>
void trsvLeftUpper(MT const& A, std::vector<T> const& b, std::vector<T>& x) {
    detail::trsv_host<UpLo::Upper>(A, b, x);
}

template <typename ST1, class MT1, typename ST2, class MT2, class MT3>
    requires HostMatrix<MT1> && HostMatrix<MT2> && HostMatrix<MT3>
void syrkLower(ST1 alpha, MT1 const& A, ST2 beta, MT2 const& C, MT3& D) {
    using T = std::remove_const_t<typename MT3::ElementType>;
    if (C.rows() != A.rows() || C.columns() != A.rows() || D.rows() != A.rows() || D.columns() != A.rows())
        throw std::invalid_argument("Matrix sizes do not match");
    detail::Staged<MT1 const> a(A);
    detail::Staged<MT2 const> c(C);
    detail::Staged<MT3> d(D);
    syrkLower<T>(T(alpha), a.batch(), T(beta), c.batch(), d.batch());
    d.copy_back();
}

template <UpLo UPLO, bool UNIT, class MT0, class MT1, class MT2>
    requires HostMatrix<MT0> && HostMatrix<MT1> && HostMatrix<MT2>
void trsm(MT0 const& A, MT1 const& B, MT2& X) {
    using T = std::remove_const_t<typename MT2::ElementType>;
    // dense/Trsm.hpp:244-251
    if (A.rows() != B.rows() || A.columns() != B.rows() || X.rows() != B.rows() || X.columns() != B.columns())
        throw std::invalid_argument("Invalid argument size");
    detail::Staged<MT0 const> a(A);
    detail::Staged<MT1 const> b(B);
    detail::Staged<MT2> x(X);
    trsm<UPLO, UNIT, T>(a.batch(), b.batch(), x.batch());
    x.copy_back();
}

template <class MT1, class MT2>
    requires HostMatrix<MT1> && HostMatrix<MT2>
void potrf(MT1 const& A, MT2& L) {
    using T = std::remove_const_t<typename MT2::ElementType>;
    // dense/Potrf.hpp:78-85
    if (A.columns() > A.rows() || L.rows() != A.rows() || L.columns() != A.columns())
        throw std::invalid_argument("Invalid matrix size");
    detail::Staged<MT1 const> a(A);
    detail::Staged<MT2> l(L);
    potrf<T>(a.batch(), l.batch());
    l.copy_back();
}

// Throws std::invalid_argument("Matrix is singular") on an exactly zero
// pivot, as BLAST's getf2 does (dense/Getf2.hpp:68-69).
template <class MT>
    requires HostMatrix<MT>
void getrf(MT& A, std::size_t* ipiv) {
    using T = std::remove_const_t<typename MT::ElementType>;
    std::size_t const k = A.rows() < A.columns() ? A.rows() : A.columns();
    detail::Staged<MT> a(A);
    hip::DeviceArray<std::int64_t> piv(k ? k : 1);
    hip::DeviceArray<std::int32_t> info(1);
    getrf<T>(a.batch(), piv.get(), std::int64_t(k ? k : 1), info.get());
    std::int32_t inf = 0;
    hip::check_hip(hipMemcpy(&inf, info.get(), sizeof(inf), hipMemcpyDeviceToHost));
    if (inf > 0) throw std::invalid_argument("Matrix is singular");
    a.copy_back();
    if (k) {
        static_assert(sizeof(std::size_t) == sizeof(std::int64_t));
        hip::check_hip(hipMemcpy(ipiv, piv.get(), k * sizeof(std::int64_t), hipMemcpyDeviceToHost));
    }
}

}  // namespace blast
