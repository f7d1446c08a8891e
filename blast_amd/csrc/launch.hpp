// Error reporting shared by the kernels' launchers and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace blast_hip {

// Records a formatted message for blast_hip_last_error() (thread-local).
void set_error(const char* fmt, ...);

// Checks hipGetLastError() after a launch; returns 0 or BLAST_HIP_ERR_RUNTIME.
int check_launch(const char* what);

}  // namespace blast_hip
