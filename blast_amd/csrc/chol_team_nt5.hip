// Team kernels for n <= 80 (chol_team.hpp), compiled on their own.
#include "chol_team.hpp"

namespace blast_hip {
template int launch_team<5, 4, 2>(hipStream_t, Mat<double>, Mat<double>, Mat<double>, Mat<double>, int, int, int64_t,
                                   int32_t*, const int32_t*);
}  // namespace blast_hip
