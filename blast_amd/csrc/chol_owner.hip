// Fused batched Cholesky factor + solve, n <= 256, nrhs <= 8: FP64, and FP32
// storage with FP64 arithmetic (loads widened, stores rounded).
//
// One workgroup of NW wavefronts per instance; the NT(NT+1)/2 lower 16x16
// tiles are distributed round-robin over the wavefronts and stay in their
// owner's registers as v_mfma_f64_16x16x4f64 accumulators for the whole
// factorisation ("owner computes").  Right-looking blocked Cholesky, block
// column kb = 0..NT-1, two barriers per block:
//   1. the owner of tile (kb,kb) factors it (diag_factor: left-looking inside
//      the tile, Newton-refined rsqrt, DPP row broadcasts), forms
//      W = L_kk^{-1}, writes L_kk, keeps W in the tile's registers for the
//      backward solve and forms y_kb = W r_kb;
//   2. the owners of panel tiles (i,kb) form L_ik = A_ik W^T on MFMA, write
//      it to L and to LDS, and update r_i -= L_ik y_kb;
//   3. every owner of a trailing tile (i,j), kb < j <= i, applies
//      A_ij -= L_ik L_jk^T on MFMA with both operands read from LDS.
// The backward solve x = L^{-T} y reads W_kk and L_kj straight from the
// registers: the f64 accumulator layout of a tile M is the MFMA A-operand
// layout of M^T.  HBM traffic is the algorithmic one: A's lower triangle and
// B in, L's lower triangle and X out.
// Algorithm reference: the potrf -> trsm<Lower> -> trsm<Upper>(trans(L))
// sequence of dense/Potrf.hpp:66-106 and dense/Trsm.hpp:17-66; the tile
// algorithm reorders BLAST's sums (agreement to backward error, not bits).

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {
namespace {

constexpr int kOwnRhs = 8;  // right-hand-side columns held in LDS

// Returns v unchanged, opaque to the optimiser: keeps per-tile address
// arithmetic inside the block loop instead of hoisted (and held in
// registers) for every tile of the wavefront.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// Tile loads and stores in the accumulator layout for storage type T: FP64
// storage uses the panel fast paths of tiles.hpp, FP32 storage (PS = 8
// panels, dense, transposed) the generic element map; the arithmetic is FP64
// either way (FP32 values are widened on load and rounded on store).
template <bool PANEL, typename T>
__device__ __forceinline__ d4 owner_load(Ref<T> const& a, int n, int ti, int tj, int lane) {
    if constexpr (std::is_same_v<T, double>) {
        return load_tile<PANEL, double>(a, n, ti, tj, lane);
    } else {
        d4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const row = 16 * ti + Mfma<double>::row(lane, r), col = 16 * tj + acc_col(lane);
            v[r] = (row < n && col < n) ? (row >= col ? double(a(row, col)) : 0.0) : (row == col ? 1.0 : 0.0);
        }
        return v;
    }
}
template <bool PANEL, typename T>
__device__ __forceinline__ void owner_store(Ref<T> const& l, bool panel, int n, int ti, int tj, d4 v, int lane) {
    if constexpr (std::is_same_v<T, double>) {
        store_tile<PANEL>(l, panel, n, ti, tj, v, lane);
    } else {  // the same elements as store_tile: lower part, zeros above it in 4x4 diagonal blocks (panel)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const row = 16 * ti + Mfma<double>::row(lane, r), col = 16 * tj + acc_col(lane);
            if (row < n && col < n && (row >= col || (panel && (row >> 2) == (col >> 2))))
                l(row, col) = row >= col ? T(v[r]) : T(0);
        }
    }
}

// Shared memory of one instance with up to NT block columns.
template <int NT>
struct OwnerLds {
    double sP[NT][kTileElems];     // panel tiles L_ik of the current block, T[m][k]
    double sD[kTileElems];         // diagonal tile being factored
    double sW[kTileElems];         // W = L_kk^{-1} of the current block
    double sR[NT * 16 * kOwnRhs];  // right-hand sides -> y -> x, [row][kOwnRhs]
    int bad;
};

// One instance (index b of the batch operands) by the NW wavefronts of the
// workgroup; `lds` is at least OwnerLds<NT> (a larger one is reused by the
// uneven kernel, whose instances have different NT).
template <int NT, int NW, bool PANEL, int NTL, typename T = double>
__device__ __forceinline__ void chol_owner_body(Mat<T> const& A, Mat<T> const& L, Mat<T> const& B, Mat<T> const& X,
                                                int n, int nrhs, int64_t b, int32_t* info, OwnerLds<NTL>& lds) {
    static_assert(NT <= NTL);
    using Mf = Mfma<double>;
    constexpr int NTILES = NT * (NT + 1) / 2;
    constexpr int S = (NTILES + NW - 1) / NW;  // register tiles per wavefront
    auto& sP = lds.sP;
    auto& sD = lds.sD;
    auto& sW = lds.sW;
    auto& sR = lds.sR;
    auto& sBad = lds.bad;

    int const lane = threadIdx.x & 63;
    int const w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    Ref<T> a(A, b), l(L, b);
    bool const lpanel = L.panel != 0;

    // tile t = i(i+1)/2 + j lives in wavefront t % NW, register slot t / NW
    int ti[S], tj[S];
    d4 acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int const t = s * NW + w;
        int i = int((__builtin_sqrtf(8.f * float(t) + 1.f) - 1.f) * 0.5f);
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        while (i * (i + 1) / 2 > t) --i;
        ti[s] = t < NTILES ? i : -1;
        tj[s] = t - i * (i + 1) / 2;
        acc[s] = ti[s] >= 0 ? owner_load<PANEL>(a, n, ti[s], tj[s], lane) : d4{0.0, 0.0, 0.0, 0.0};
    }
    if (threadIdx.x == 0) sBad = 0;
    int const nr = nrhs;
    if (nr > 0) {
        Ref<T> bb(B, b);
        for (int e = threadIdx.x; e < NT * 16 * kOwnRhs; e += 64 * NW) {
            int const row = e / kOwnRhs, c = e % kOwnRhs;
            sR[e] = (row < n && c < nr) ? double(bb(row, c)) : 0.0;
        }
    }
    __syncthreads();

    // right-hand-side block blk: MFMA B operand row k, column lane % 16
    auto rhs_b = [&](int blk, int k) -> double {
        int const c = lane & 15;
        return c < kOwnRhs ? sR[(16 * blk + k) * kOwnRhs + c] : 0.0;
    };
    auto rhs_load = [&](int blk) {
        d4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const c = acc_col(lane);
            v[r] = c < kOwnRhs ? sR[(16 * blk + Mf::row(lane, r)) * kOwnRhs + c] : 0.0;
        }
        return v;
    };
    auto rhs_store = [&](int blk, d4 v) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int const c = acc_col(lane);
            if (c < kOwnRhs) sR[(16 * blk + Mf::row(lane, r)) * kOwnRhs + c] = v[r];
        }
    };

    for (int kb = 0; kb < NT; ++kb) {
        {  // 1. diagonal tile, by its owner (one copy of the factor code)
            int const td = kb * (kb + 1) / 2 + kb, sd = td / NW;
            if (w == td % NW) {
                d4 dt = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (s == sd) dt = acc[s];
                acc_to_lds(sD, dt, lane);
                int const bk = diag_factor_inl<double>(sD, sW, lane);
                if (lane == 0 && bk && sBad == 0 && 16 * kb + bk <= n) sBad = 16 * kb + bk;
                // (opaque indices: the store's address terms formed here, not
                // hoisted out of the block loop into registers)
                owner_store<PANEL>(l, lpanel, n, opaque(kb), opaque(kb), lds_to_acc<double>(sD, lane), lane);
                d4 const wv = lds_to_acc<double>(sW, lane);
#pragma unroll
                for (int s = 0; s < S; ++s)
                    if (s == sd) acc[s] = wv;  // W_kk, kept for the backward solve
                if (nr > 0) {
                    d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) y = Mf::mma(op_read(sW, q, lane), rhs_b(kb, 4 * q + (lane >> 4)), y);
                    rhs_store(kb, y);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S; ++s) {  // 2. panel
            if (tj[s] == kb && ti[s] > kb) {
                int const i = opaque(ti[s]);
                acc_to_lds(sP[i], acc[s], lane);
                d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 4; ++q) z = Mf::mma(op_read(sP[i], q, lane), op_read(sW, q, lane), z);
                acc[s] = z;
                owner_store<PANEL>(l, lpanel, n, i, opaque(kb), z, lane);
                acc_to_lds(sP[i], z, lane);
                if (nr > 0) {
                    d4 r = rhs_load(i);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        r = Mf::mma(-op_read(sP[i], q, lane), rhs_b(kb, 4 * q + (lane >> 4)), r);
                    rhs_store(i, r);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < S; ++s) {  // 3. trailing update
            if (tj[s] > kb) {
                int const i = opaque(ti[s]), j = opaque(tj[s]);
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[s] = Mf::mma(-op_read(sP[i], q, lane), op_read(sP[j], q, lane), acc[s]);
            }
        }
    }

    if (nr > 0) {  // x = L^{-T} y: W_kk and L_kj from the registers (A operand of the transpose)
        for (int bk = NT - 1; bk >= 0; --bk) {
            __syncthreads();
            {
                int const td = bk * (bk + 1) / 2 + bk, sd = td / NW;
                if (w == td % NW) {
                    d4 wv = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s = 0; s < S; ++s)
                        if (s == sd) wv = acc[s];
                    d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int q = 0; q < 4; ++q) x = Mf::mma(wv[q], rhs_b(bk, Mf::row(lane, q)), x);
                    rhs_store(bk, x);
                }
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (ti[s] == bk && tj[s] < bk) {
                    int const j = opaque(tj[s]);
                    d4 y = rhs_load(j);
#pragma unroll
                    for (int q = 0; q < 4; ++q) y = Mf::mma(-acc[s][q], rhs_b(bk, Mf::row(lane, q)), y);
                    rhs_store(j, y);
                }
            }
        }
        __syncthreads();
        Ref<T> xx(X, b);
        for (int e = threadIdx.x; e < n * nr; e += 64 * NW) {
            int const row = e % n, c = e / n;
            xx(row, c) = T(sR[row * kOwnRhs + c]);
        }
    } else {
        __syncthreads();
    }
    if (info && threadIdx.x == 0) info[b] = sBad;
}

template <int NT, int NW, int WPE, bool PANEL, typename T = double>
__global__ __launch_bounds__(64 * NW, WPE) void chol_owner_kernel(Mat<T> A, Mat<T> L, Mat<T> B, Mat<T> X, int n,
                                                             int nrhs, int32_t* info, const int32_t* only_if) {
    __shared__ OwnerLds<NT> lds;
    int64_t const b = blockIdx.x;
    if (only_if && only_if[b] >= 0) return;
    chol_owner_body<NT, NW, PANEL, NT, T>(A, L, B, X, n, nrhs, b, info, lds);
}

// Uneven batches: one persistent launch over instances of different n (one
// workgroup of kUnevenNW wavefronts per instance, n <= 16 kUnevenNT).  Each
// workgroup takes the next queue entry (atomic counter in queue[0], the
// order in queue[1..count], largest block count first so that small
// instances fill the tail) until the queue is empty.
constexpr int kUnevenNT = 12, kUnevenNW = 4;

__device__ __forceinline__ bool uneven_shape_ok(blast_solve_instance_t const& d) {
    int64_t const n = d.a.rows;
    bool ok = n >= 0 && n <= 16 * kUnevenNT && d.a.cols == n && d.l.rows == n && d.l.cols == n && d.a.data &&
              d.l.data;
    if (d.b.data || d.x.data)
        ok = ok && d.b.data && d.x.data && d.b.rows == n && d.x.rows == n && d.x.cols == d.b.cols && d.b.cols >= 0 &&
             d.b.cols <= kOwnRhs;
    return ok;
}

__global__ __launch_bounds__(64 * kUnevenNW, 2) void posv_uneven_kernel(const blast_solve_instance_t* inst,
                                                                        int64_t count, int32_t* queue,
                                                                        int32_t* info) {
    __shared__ OwnerLds<kUnevenNT> lds;
    __shared__ int sIdx;
    for (;;) {
        if (threadIdx.x == 0) sIdx = atomicAdd(&queue[0], 1);
        __syncthreads();
        int const q = sIdx;
        __syncthreads();
        if (q >= count) break;  // every workgroup reaches this exit
        int const idx = queue[1 + q];
        if (idx < 0 || idx >= count) continue;
        blast_solve_instance_t const d = inst[idx];
        int32_t* const inf = info ? info + idx : nullptr;
        if (!uneven_shape_ok(d)) {
            if (inf && threadIdx.x == 0) *inf = -1;
            continue;
        }
        Mat<double> const A = Mat<double>::from(d.a), L = Mat<double>::from(d.l);
        Mat<double> const B = Mat<double>::from(d.b), X = Mat<double>::from(d.x);
        int const n = int(d.a.rows), nrhs = d.b.data ? int(d.b.cols) : 0;
        switch ((n + 15) / 16) {
            case 0:
            case 1: chol_owner_body<1, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 2: chol_owner_body<2, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 3: chol_owner_body<3, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 4: chol_owner_body<4, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 5: chol_owner_body<5, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 6: chol_owner_body<6, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 7: chol_owner_body<7, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 8: chol_owner_body<8, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 9: chol_owner_body<9, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 10: chol_owner_body<10, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            case 11: chol_owner_body<11, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
            default: chol_owner_body<12, kUnevenNW, false>(A, L, B, X, n, nrhs, 0, inf, lds); break;
        }
    }
}

__global__ void zero_queue_counter_kernel(int32_t* p) { *p = 0; }

template <int NT, int NW, int WPE = 1>
int launch_owner_nt(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int n, int nrhs,
                    int64_t batch, int32_t* info, const int32_t* only_if) {
    bool const panel = A.panel && !A.trans && L.panel && !L.trans;
    if (panel)
        hipLaunchKernelGGL((chol_owner_kernel<NT, NW, WPE, true>), dim3(unsigned(batch)), dim3(64 * NW), 0, s, A, L, B, X,
                           n, nrhs, info, only_if);
    else
        hipLaunchKernelGGL((chol_owner_kernel<NT, NW, WPE, false>), dim3(unsigned(batch)), dim3(64 * NW), 0, s, A, L, B,
                           X, n, nrhs, info, only_if);
    return check_launch("chol_owner_kernel");
}

}  // namespace

// Fused potrf (+ forward/backward solve when nrhs > 0), square n <= 256,
// nrhs <= 8.  The dispatcher prefers posv_tiles.hip for n <= 128, nrhs <= 4
// (one wavefront per instance with every tile in its registers is faster
// there).  L may alias A, X may alias B.  Returns 1 when the shape is not
// covered.
int launch_chol_owner(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                      int32_t* info, const int32_t* only_if) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    if (L.rows != L.cols || n > 256 || nrhs > kOwnRhs) return 1;
    if (batch == 0) return 0;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_owner_nt<1, 1>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 2: return launch_owner_nt<2, 1>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 3: return launch_owner_nt<3, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 4: return launch_owner_nt<4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 5: return launch_owner_nt<5, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 6: return launch_owner_nt<6, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 7: return launch_owner_nt<7, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 8: return launch_owner_nt<8, 4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 9: return launch_owner_nt<9, 8>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 10: return launch_owner_nt<10, 8>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 11: return launch_owner_nt<11, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 12: return launch_owner_nt<12, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 13: return launch_owner_nt<13, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 14: return launch_owner_nt<14, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 15: return launch_owner_nt<15, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        default: return launch_owner_nt<16, 16>(s, A, L, B, X, n, nrhs, batch, info, only_if);
    }
}

// FP32 storage (any layout), FP64 arithmetic: square n <= 256, nrhs <= 8; L
// may alias A, X may alias B.  Returns 1 when the shape is not covered.
template <int NT, int NW, int WPE = 1>
int launch_owner_nt_f32(hipStream_t s, Mat<float> A, Mat<float> L, Mat<float> B, Mat<float> X, int n, int nrhs,
                        int64_t batch, int32_t* info) {
    hipLaunchKernelGGL((chol_owner_kernel<NT, NW, WPE, false, float>), dim3(unsigned(batch)), dim3(64 * NW), 0, s, A, L,
                       B, X, n, nrhs, info, nullptr);
    return check_launch("chol_owner_kernel<float>");
}
int launch_chol_owner_f32(hipStream_t s, Mat<float> A, Mat<float> L, Mat<float> B, Mat<float> X, int64_t batch,
                          int32_t* info) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    if (L.rows != L.cols || n > 256 || nrhs > kOwnRhs) return 1;
    if (batch == 0) return 0;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_owner_nt_f32<1, 1>(s, A, L, B, X, n, nrhs, batch, info);
        case 2: return launch_owner_nt_f32<2, 1>(s, A, L, B, X, n, nrhs, batch, info);
        case 3: return launch_owner_nt_f32<3, 2>(s, A, L, B, X, n, nrhs, batch, info);
        case 4: return launch_owner_nt_f32<4, 2>(s, A, L, B, X, n, nrhs, batch, info);
        case 5: return launch_owner_nt_f32<5, 4>(s, A, L, B, X, n, nrhs, batch, info);
        case 6: return launch_owner_nt_f32<6, 4>(s, A, L, B, X, n, nrhs, batch, info);
        case 7: return launch_owner_nt_f32<7, 4>(s, A, L, B, X, n, nrhs, batch, info);
        case 8: return launch_owner_nt_f32<8, 4, 2>(s, A, L, B, X, n, nrhs, batch, info);
        case 9: return launch_owner_nt_f32<9, 8>(s, A, L, B, X, n, nrhs, batch, info);
        case 10: return launch_owner_nt_f32<10, 8>(s, A, L, B, X, n, nrhs, batch, info);
        case 11: return launch_owner_nt_f32<11, 16>(s, A, L, B, X, n, nrhs, batch, info);
        case 12: return launch_owner_nt_f32<12, 16>(s, A, L, B, X, n, nrhs, batch, info);
        case 13: return launch_owner_nt_f32<13, 16>(s, A, L, B, X, n, nrhs, batch, info);
        case 14: return launch_owner_nt_f32<14, 16>(s, A, L, B, X, n, nrhs, batch, info);
        case 15: return launch_owner_nt_f32<15, 16>(s, A, L, B, X, n, nrhs, batch, info);
        default: return launch_owner_nt_f32<16, 16>(s, A, L, B, X, n, nrhs, batch, info);
    }
}

// Uneven batched Cholesky (+ solve): queue counter reset by a kernel (no
// memset, capturable), then one persistent launch of at most `grid`
// workgroups (0: min(count, 2048)).
int launch_posv_uneven(hipStream_t s, const blast_solve_instance_t* inst, int64_t count, int32_t* queue,
                       int32_t* info, int grid) {
    if (count == 0) return 0;
    hipLaunchKernelGGL(zero_queue_counter_kernel, dim3(1), dim3(1), 0, s, queue);
    int r = check_launch("zero_queue_counter_kernel");
    if (r) return r;
    int64_t const g = grid > 0 ? (grid < count ? grid : count) : (count < 2048 ? count : 2048);
    hipLaunchKernelGGL(posv_uneven_kernel, dim3(unsigned(g)), dim3(64 * kUnevenNW), 0, s, inst, count, queue, info);
    return check_launch("posv_uneven_kernel");
}

int uneven_max_n() { return 16 * kUnevenNT; }

}  // namespace blast_hip
