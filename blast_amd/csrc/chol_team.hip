// Dispatch of the fused FP64 Cholesky team kernels (chol_team.hpp).
#include "chol_team.hpp"

namespace blast_hip {

extern template int launch_team<5, 4, 2>(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int, int,
                                         int64_t, int32_t*, const int32_t*);
extern template int launch_team<6, 4, 2>(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int, int,
                                         int64_t, int32_t*, const int32_t*);
extern template int launch_team<7, 4, 2>(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int, int,
                                         int64_t, int32_t*, const int32_t*);
extern template int launch_team<8, 4, 2>(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int, int,
                                         int64_t, int32_t*, const int32_t*);

// Fused potrf (+ forward/backward solve when nrhs > 0), square n <= 128,
// nrhs <= 4.  L may alias A, X may alias B.  Returns 1 when the shape is not
// covered.
int launch_chol_team(hipStream_t s, Mat<double> A, Mat<double> L, Mat<double> B, Mat<double> X, int64_t batch,
                     int32_t* info, const int32_t* only_if) {
    int const n = L.rows, nrhs = (B.data && X.data) ? X.cols : 0;
    if (L.rows != L.cols || n > 128 || nrhs > 4) return 1;
    if (batch == 0) return 0;
    switch ((n + 15) / 16) {
        case 0:
        case 1: return launch_team<1, 1, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 2: return launch_team<2, 1, 4>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 3: return launch_team<3, 2, 3>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 4: return launch_team<4, 2, 3>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 5: return launch_team<5, 4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 6: return launch_team<6, 4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        case 7: return launch_team<7, 4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
        default: return launch_team<8, 4, 2>(s, A, L, B, X, n, nrhs, batch, info, only_if);
    }
}

}  // namespace blast_hip
