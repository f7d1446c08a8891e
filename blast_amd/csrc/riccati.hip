// Uneven batched Riccati recursion (LQR / MPC backward pass), FP64, one
// persistent launch.
//
// Per instance (nx states, nu inputs, horizon N; A nx x nx, B nx x nu,
// Q nx x nx SPD, R nu x nu SPD; dense column-major, leading dimension =
// rows), with P_N = Q and L = chol(P):
//   M   = L^T [B A]                         gemm (triangular operand)
//   Lam = blkdiag(R, Q) + M^T M             syrk, lower
//   chol(Lam) = [Ls 0; Lks Lp]              potrf (blocked, MFMA trailing update)
//   K_k = -Ls^{-T} Lks^T                    trsm
//   L  <- Lp   (P_k = Q + A^T P A - H^T S^{-1} H, S = R + B^T P B, H = B^T P A)
// for k = N-1 .. 0, i.e. the square-root form of
//   K_k = -(R + B^T P B)^{-1} B^T P A,  P <- Q + A^T P A + A^T P B K_k.
// Outputs: K_k (nu x nx, k = 0..N-1) and L_0 = chol(P_0) (lower).
// This composite is not in BLAST; it strings together the entry points BLAST
// provides (gemm, syrkLower, potrf, trsm) the way the reference's
// syrk+potrf benchmarks and BLASFEO wrapper do (bench/blas/SyrkPotrf.cpp,
// include/blasfeo/blasfeo_api/SyrkPotrf.hpp) and is checked against the
// textbook recursion in FP64.
//
// Scheduling: workgroups pull instance indices from queue[1..count] through
// the counter queue[0] (atomic); the host orders the queue by size bucket,
// largest first, so the tail of the launch is made of small instances.
// The trailing-update strip width (1 or 2 MFMA tiles per wavefront; a table
// entry of 4 runs as 2) is taken per nx-bucket from the tuning table.  Three
// workgroups per CU (at most 168 VGPRs).

#include "dispatch.hpp"
#include "launch.hpp"
#include "tiles.hpp"

namespace blast_hip {

using RiccatiInstance = blast_riccati_instance_t;

namespace {

constexpr int kT = 256;

// Dense column-major view: element (i, j) at p[i + ld*j] (or p[j + ld*i] transposed).
struct DV {
    double* p;
    int ld;
    bool t;
    __device__ double& operator()(int i, int j) const { return t ? p[j + int64_t(ld) * i] : p[i + int64_t(ld) * j]; }
    __device__ DV sub(int i, int j) const { return DV{&(*this)(i, j), ld, t}; }
    __device__ DV tr() const { return DV{p, ld, !t}; }
};

enum { kPlain = 0, kSyrkAdd = 1, kUpperA = 2 };

// blkdiag(R, Q) read in place (the C operand of the Lam syrk): no separate
// fill pass over Lam.
struct BlkDiagRQ {
    DV r, q;
    int nu;
    __device__ double operator()(int i, int j) const {
        return i < nu ? (j < nu ? r(i, j) : 0.0) : (j >= nu ? q(i - nu, j - nu) : 0.0);
    }
};

// D = alpha*op(A)*op(B) + beta*C on 16x16 MFMA tiles, whole workgroup,
// TN tiles per wavefront strip.  kSyrkAdd: lower triangle of an M x M
// result only; kUpperA: A masked to its upper triangle.
template <int TN, int MODE, class CV = DV>
__device__ void wg_gemm(double alpha, DV A, DV B, double beta, CV C, DV D, int M, int N, int K) {
    using Mf = Mfma<double>;
    int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int const m = lane & 15, g = lane >> 4;
    int const MT = (M + 15) / 16, NTt = (N + 15) / 16, KQ = (K + 3) / 4;
    for (int R = w; R < MT; R += kT / 64) {
        int const row = 16 * R + m;
        for (int T0 = 0; T0 < NTt; T0 += TN) {
            if (MODE == kSyrkAdd && 16 * T0 > 16 * R + 15) break;
            d4 acc[TN];
#pragma unroll
            for (int t = 0; t < TN; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
            // four K-chunks of loads in flight; an upper-triangular A has no
            // entries left of its strip's diagonal (k < 16 R), so kUpperA starts there
            for (int q0 = MODE == kUpperA ? 4 * R : 0; q0 < KQ; q0 += 4) {
                double av[4], bv[4][TN];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    int const k = 4 * (q0 + u) + g;
                    bool const ina = row < M && k < K && (MODE != kUpperA || k >= row);
                    av[u] = ina ? A(row, k) : 0.0;
#pragma unroll
                    for (int t = 0; t < TN; ++t) {
                        int const col = 16 * (T0 + t) + m;
                        bv[u][t] = (col < N && k < K) ? B(k, col) : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int t = 0; t < TN; ++t) acc[t] = Mf::mma(av[u], bv[u][t], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < TN; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const i = 16 * R + Mf::row(lane, r), j = 16 * (T0 + t) + acc_col(lane);
                    if (i < M && j < N && (MODE != kSyrkAdd || i >= j))
                        D(i, j) = beta == 0.0 ? alpha * acc[t][r] : __builtin_fma(beta, C(i, j), alpha * acc[t][r]);
                }
        }
    }
    __syncthreads();
}

// The same product with both operands staged through LDS: the workgroup
// forms 64 x 16 tn blocks of D (wavefront w: rows 16 w..16 w + 15, tn = 1 or
// 2 column tiles, a runtime choice so that one instantiation serves every
// size bucket), K in chunks of 16 held in two LDS buffers (each element of a chunk
// loaded once from memory and read by all four wavefronts, the next chunk's
// loads in flight during the current chunk's MFMAs).  Loads follow each
// operand's contiguous direction (DV::t).  kSyrkAdd skips the blocks above
// the diagonal; kUpperA starts K at the block's first row.  sg: 4 * 16 * 65
// doubles of LDS.
constexpr int kGemmKc = 16, kGemmLs = 65, kGemmLds = 4 * kGemmKc * kGemmLs;
template <int MODE, class CV = DV>
__device__ void wg_gemm_lds(double alpha, DV A, DV B, double beta, CV C, DV D, int M, int N, int K, int tn,
                            double* sg) {
    using Mf = Mfma<double>;
    // at most 2 tiles per strip: the 32 VGPRs of two more accumulators would
    // cost the third workgroup per CU (14 % on the uneven 20000 x 10 batch)
    constexpr int BM = 64, KC = kGemmKc, LS = kGemmLs, TN = 2;
    tn = tn < TN ? tn : TN;
    int const lbn = tn == 4 ? 6 : (tn == 2 ? 5 : 4), BN = 1 << lbn;
    // (lane-dependent indices formed here: hoisted out of a caller's loop they
    // would stay live, and be saved, across its calls)
    int const tid = opaque_v(int(threadIdx.x)), lane = tid & 63, w = tid >> 6;
    int const m = lane & 15, g = lane >> 4;
    int const MB = (M + BM - 1) / BM, NB = (N + BN - 1) >> lbn;
    double av[4], bv[TN];
    auto load_chunk = [&](int bi, int bj, int kc) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int const r = A.t ? tid / KC + 16 * i : tid % BM, kk = A.t ? tid % KC : tid / BM + 4 * i;
            int const row = bi * BM + r, k = kc + kk;
            bool const in = row < M && k < K && (MODE != kUpperA || k >= row);
            av[i] = in ? A(row, k) : 0.0;
        }
#pragma unroll
        for (int i = 0; i < TN; ++i) {
            if (i >= tn) break;
            int const c = B.t ? tid & (BN - 1) : tid / KC + 16 * i;
            int const kk = B.t ? (tid >> lbn) + (256 >> lbn) * i : tid % KC;
            int const col = bj * BN + c, k = kc + kk;
            bv[i] = (col < N && k < K) ? B(k, col) : 0.0;
        }
    };
    auto store_chunk = [&](int buf) __attribute__((always_inline)) {
        double* const sa = sg + (2 * buf) * KC * LS;
        double* const sb = sa + KC * LS;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int const r = A.t ? tid / KC + 16 * i : tid % BM, kk = A.t ? tid % KC : tid / BM + 4 * i;
            sa[kk * LS + r] = av[i];
        }
#pragma unroll
        for (int i = 0; i < TN; ++i) {
            if (i >= tn) break;
            int const c = B.t ? tid & (BN - 1) : tid / KC + 16 * i;
            int const kk = B.t ? (tid >> lbn) + (256 >> lbn) * i : tid % KC;
            sb[kk * LS + c] = bv[i];
        }
    };
    for (int bi = 0; bi < MB; ++bi) {
        for (int bj = 0; bj < NB; ++bj) {
            if (MODE == kSyrkAdd && bj * BN > bi * BM + BM - 1) break;
            // an upper-triangular A has no entries left of its rows' diagonal
            int const k0 = MODE == kUpperA ? (bi * BM / KC) * KC : 0;
            d4 acc[TN];
#pragma unroll
            for (int t = 0; t < TN; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
            load_chunk(bi, bj, k0);
            int buf = 0;
            for (int kc = k0; kc < K; kc += KC) {
                store_chunk(buf);
                __syncthreads();
                if (kc + KC < K) load_chunk(bi, bj, kc + KC);
                double const* const sa = sg + (2 * buf) * KC * LS;
                double const* const sb = sa + KC * LS;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    double const a = sa[(4 * q + g) * LS + 16 * w + m];
#pragma unroll
                    for (int t = 0; t < TN; ++t)
                        if (t < tn) acc[t] = Mf::mma(a, sb[(4 * q + g) * LS + 16 * t + m], acc[t]);
                }
                buf ^= 1;
            }
#pragma unroll
            for (int t = 0; t < TN; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int const i = bi * BM + 16 * w + Mf::row(lane, r), j = bj * BN + 16 * t + acc_col(lane);
                    if (t < tn && i < M && j < N && (MODE != kSyrkAdd || i >= j))
                        D(i, j) = beta == 0.0 ? alpha * acc[t][r] : __builtin_fma(beta, C(i, j), alpha * acc[t][r]);
                }
            __syncthreads();  // the next block's first chunk reuses buffer 0
        }
    }
}

// In-place lower Cholesky of the n x n matrix a (lower triangle read and
// written), blocked by 16 columns: wavefront 0 factors the 16x16 diagonal
// block in LDS and forms W = L_kk^{-1} (diag_factor, no barriers inside), the
// panel below is L_ik = A_ik W^T on MFMA, the trailing matrix is updated on
// MFMA tiles.  Returns the first non-positive pivot column (1-based) or 0
// (uniform across the workgroup).
__device__ int wg_potrf(DV a, int n, int tn, double* sB, double* sW, int* sBad, double* sG) {
    using Mf = Mfma<double>;
    int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) *sBad = 0;
    for (int k0 = 0; k0 < n; k0 += 16) {
        int const kb = min(16, n - k0);
        // diagonal block -> LDS (T[m][k] at k*17+m), identity beyond kb
        for (int e = threadIdx.x; e < 256; e += kT) {
            int const i = e & 15, j = e >> 4;
            sB[j * kTileStride + i] = (i < kb && j < kb) ? (i >= j ? a(k0 + i, k0 + j) : 0.0) : (i == j ? 1.0 : 0.0);
        }
        __syncthreads();
        if (w == 0) {
            int const bk = diag_factor<double>(sB, sW, lane);
            if (lane == 0 && bk && bk <= kb && *sBad == 0) *sBad = k0 + bk;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < kb * kb; e += kT) {
            int const i = e % kb, j = e / kb;
            if (i >= j) a(k0 + i, k0 + j) = sB[j * kTileStride + i];
        }
        // panel: L_ik = A_ik W^T, 16-row tiles over the wavefronts
        int const r0 = k0 + kb, rest = n - r0;
        for (int R = w; 16 * R < rest; R += kT / 64) {
            int const row = r0 + 16 * R + (lane & 15);
            double av[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const k = 4 * q + (lane >> 4);
                av[q] = (row < n && k < kb) ? a(row, k0 + k) : 0.0;
            }
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = Mf::mma(av[q], op_read(sW, q, lane), acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const i = r0 + 16 * R + Mf::row(lane, r), j = acc_col(lane);
                if (i < n && j < kb) a(i, k0 + j) = acc[r];
            }
        }
        __syncthreads();
        // trailing update, lower triangle, A22 -= L21 L21^T, in passes of
        // four blocks (64 columns): after block gi = 0, 1, 2 of a group only
        // the next block column is updated, with the group's blocks so far
        // (its factorisation needs them); after block 3 the whole trailing
        // matrix takes the group's four blocks at once (K = 64), so it is read
        // and written a quarter as often.  Each block column receives each
        // block's contribution once.  (One call site: two would outgrow the
        // inliner's budget and spill.)
        if (rest > 0) {
            int const gi = (k0 >> 4) & 3;
            DV const l21 = a.sub(r0, k0 - 16 * gi), a22 = a.sub(r0, r0);
            int const nn = gi == 3 ? rest : min(16, rest), kk = 16 * gi + kb;
            wg_gemm_lds<kSyrkAdd>(-1.0, l21, l21.tr(), 1.0, a22, a22, rest, nn, kk, tn, sG);
        }
    }
    __syncthreads();
    return *sBad;
}

// Solve U X = B in place for upper-triangular U (here U = Ls^T as a
// transposed view), n x nrhs, by 16-row blocks from the bottom: the
// diagonal block is inverted in LDS by one wavefront (tri_inverse), applied
// to its 16 rows on MFMA (16-column tiles over the wavefronts, in place:
// each tile's rows are read before any is written), and the rows above are
// updated with one MFMA product.  Four barriers per block instead of two
// per row of a column sweep.
__device__ __attribute__((noinline)) void wg_trsm_upper(DV U, DV Bm, int n, int nrhs, double* sA, double* sW) {
    using Mf = Mfma<double>;
    int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int r0 = ((n - 1) / 16) * 16; r0 >= 0; r0 -= 16) {
        int const kb = min(16, n - r0);
        // diagonal block (upper) -> LDS as T[m][k] at k*17 + m, identity beyond kb
        for (int e = threadIdx.x; e < 256; e += kT) {
            int const i = e & 15, j = e >> 4;
            sA[j * kTileStride + i] = (i < kb && j < kb) ? (i <= j ? U(r0 + i, r0 + j) : 0.0) : (i == j ? 1.0 : 0.0);
        }
        __syncthreads();
        if (w == 0) tri_inverse<double>(sA, sW, lane, true, false);
        __syncthreads();
        // X_b = U_bb^{-1} B_b
        for (int T0 = w; 16 * T0 < nrhs; T0 += kT / 64) {
            int const col = 16 * T0 + (lane & 15);
            d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const k = 4 * q + (lane >> 4);
                double const bv = (k < kb && col < nrhs) ? Bm(r0 + k, col) : 0.0;
                acc = Mf::mma(op_read(sW, q, lane), bv, acc);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int const i = Mf::row(lane, r);
                if (i < kb && col < nrhs) Bm(r0 + i, col) = acc[r];
            }
        }
        __syncthreads();
        // rows above: B_a -= U_ab X_b
        if (r0 > 0) wg_gemm<1, kPlain>(-1.0, U.sub(0, r0), Bm.sub(r0, 0), 1.0, Bm, Bm, r0, nrhs, kb);
    }
}

// info of an instance: 0; c when chol(Q) meets a pivot that is not positive
// at column c; kRiccatiStepInfo (k + 1) + c when the factorisation of Lam at
// step k does (column c <= nu: S = R + B^T P B is not positive definite).
constexpr int kRiccatiStepInfo = 65536;

__device__ void riccati_instance(RiccatiInstance const& ins, int horizon, int tn, const double* in, double* out,
                                 double* work, double* sB, double* sW, int* sBad, double* sG, int32_t* info_out) {
    int const nx = ins.nx, nu = ins.nu, nt = nx + nu;
    DV const A{const_cast<double*>(in) + ins.a, nx, false};
    DV const B{const_cast<double*>(in) + ins.b, nx, false};
    DV const Q{const_cast<double*>(in) + ins.q, nx, false};
    DV const R{const_cast<double*>(in) + ins.r, nu, false};
    double* wp = work + ins.work;
    DV const BA{wp, nx, false};                   // [B A], nx x nt
    DV const Mm{wp + int64_t(nx) * nt, nx, false};  // M = L^T [B A], nx x nt
    DV const Lam{wp + 2 * int64_t(nx) * nt, nt, false};
    DV const Lcur = Lam.sub(nu, nu);              // chol(P) lives in Lam's trailing block
    int bad = 0;
    for (int e = threadIdx.x; e < nx * nt; e += kT) {
        int const i = e % nx, j = e / nx;
        BA(i, j) = j < nu ? B(i, j) : A(i, j - nu);
    }
    for (int e = threadIdx.x; e < nx * nx; e += kT) {
        int const i = e % nx, j = e / nx;
        Lcur(i, j) = i >= j ? Q(i, j) : 0.0;
    }
    __syncthreads();
    int b0 = wg_potrf(Lcur, nx, tn, sB, sW, sBad, sG);
    if (b0) bad = b0;
    for (int k = horizon - 1; k >= 0; --k) {
        // M = L^T [B A]
        wg_gemm_lds<kUpperA>(1.0, Lcur.tr(), BA, 0.0, Mm, Mm, nx, nt, nx, tn, sG);
        // Lam = blkdiag(R, Q) (lower) + M^T M, the block diagonal read in
        // the syrk epilogue
        wg_gemm_lds<kSyrkAdd>(1.0, Mm.tr(), Mm, 1.0, BlkDiagRQ{R, Q, nu}, Lam, nt, nt, nx, tn, sG);
        int const bk = wg_potrf(Lam, nt, tn, sB, sW, sBad, sG);
        if (bk && bad == 0) bad = kRiccatiStepInfo * (k + 1) + bk;  // step k, column bk of Lam
        // K_k = -Ls^{-T} Lks^T
        DV const Kk{out + ins.k + int64_t(k) * nu * nx, nu, false};
        for (int e = threadIdx.x; e < nu * nx; e += kT) {
            int const i = e % nu, j = e / nu;
            Kk(i, j) = -Lam(nu + j, i);
        }
        __syncthreads();
        wg_trsm_upper(Lam.tr(), Kk, nu, nx, sB, sW);
    }
    DV const L0{out + ins.l, nx, false};
    for (int e = threadIdx.x; e < nx * nx; e += kT) {
        int const i = e % nx, j = e / nx;
        L0(i, j) = i >= j ? Lcur(i, j) : 0.0;
    }
    if (threadIdx.x == 0 && info_out) *info_out = bad;
    __syncthreads();
}

__global__ __launch_bounds__(kT, 3) void riccati_persistent_kernel(const RiccatiInstance* inst, int64_t start,
                                                                int64_t count, int horizon, const double* in,
                                                                double* out, double* work, int32_t* queue,
                                                                RiccatiTuning tuning, int32_t* info) {
    __shared__ double sB[kTileElems];
    __shared__ double sW[kTileElems];
    __shared__ double sG[kGemmLds];
    __shared__ int sIdx, sBad;
    for (;;) {
        if (threadIdx.x == 0) sIdx = atomicAdd(&queue[0], 1);
        __syncthreads();
        int const q = sIdx;
        __syncthreads();
        if (q >= count) break;  // every workgroup reaches this exit
        int const idx = queue[1 + start + q];
        RiccatiInstance const ins = inst[idx];
        int const tn = tuning.tn[min(ins.nx / 16, 15)];
        int32_t* inf = info ? info + idx : nullptr;
        riccati_instance(ins, horizon, tn, in, out, work, sB, sW, &sBad, sG, inf);
    }
}

__global__ void zero_counter_kernel(int32_t* p) { *p = 0; }

}  // namespace

int launch_riccati(hipStream_t s, int horizon, const void* inst, int64_t start, int64_t count, const double* in,
                   double* out, double* work, int32_t* queue, RiccatiTuning const& tuning, int32_t* info, int grid) {
    if (count == 0) return 0;
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, s, queue);
    int r = check_launch("zero_counter_kernel");
    if (r) return r;
    int64_t const g = grid > 0 ? (grid < count ? grid : count) : (count < 2048 ? count : 2048);
    hipLaunchKernelGGL(riccati_persistent_kernel, dim3(unsigned(g)), dim3(kT), 0, s,
                       static_cast<const RiccatiInstance*>(inst), start, count, horizon, in, out, work, queue, tuning,
                       info);
    return check_launch("riccati_persistent_kernel");
}

}  // namespace blast_hip
