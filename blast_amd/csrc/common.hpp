// Shared device-side definitions: storage descriptors and element addressing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blast_hip.h"

namespace blast_hip {

// PanelSize_v (panel/PanelSize.hpp:25-26): SIMD width on AVX2.
template <typename T> struct PanelSize;
template <> struct PanelSize<double> { static constexpr int value = 4, log2 = 2; };
template <> struct PanelSize<float> { static constexpr int value = 8, log2 = 3; };

// TileSize_v (system/Tile.hpp:19-34): BLAST's blocking unit (4 double, 8 float).
template <typename T> constexpr int tile_size() { return PanelSize<T>::value; }

// Device view of one batch operand.  Element (i, j) of the logical matrix of
// instance b lives at data + b*stride + off(i, j).
template <typename T>
struct Mat {
    T* data;
    int64_t stride;
    int32_t rows, cols, spacing;
    int32_t panel, trans;

    __host__ __device__ static Mat from(const blast_matrix_t& m) {
        Mat r;
        r.data = static_cast<T*>(m.data);
        r.stride = m.batch_stride;
        r.rows = int32_t(m.rows);
        r.cols = int32_t(m.cols);
        r.spacing = int32_t(m.spacing);
        r.panel = m.panel;
        r.trans = m.trans;
        return r;
    }

    __device__ __forceinline__ int64_t off(int i, int j) const {
        int const r = trans ? j : i;
        int const c = trans ? i : j;
        constexpr int PS = PanelSize<T>::value, LOG = PanelSize<T>::log2;
        return panel ? int64_t(r >> LOG) * spacing + (r & (PS - 1)) + int64_t(c) * PS
                     : int64_t(r) + int64_t(spacing) * c;
    }
    __device__ __forceinline__ T* inst(int64_t b) const { return data + b * stride; }
};

// Fused multiply-add in the operand type: __builtin_fma is the double
// routine, so float operands would be widened and the result rounded twice.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Per-instance accessor bound to one instance's base pointer.
template <typename T>
struct Ref {
    Mat<T> m;
    T* p;
    __device__ __forceinline__ Ref(Mat<T> const& mm, int64_t b) : m(mm), p(mm.inst(b)) {}
    __device__ __forceinline__ T& operator()(int i, int j) const { return p[m.off(i, j)]; }
};

}  // namespace blast_hip
