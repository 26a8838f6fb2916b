// The batched SCF kernel built for molecules of up to 64 basis functions: the same source
// (am1g_scf.hip) compiled with 256-thread workgroups and 68 KiB of dynamic LDS, so that two
// molecules share a CU. Each of a CU's four SIMDs then holds one wavefront of each molecule (256
// VGPRs each), so one molecule's dependent chain of barriers, reductions and memory round trips
// overlaps the other's. The register-blocked Householder step of the eigensolver covers 64 rows
// with 256 threads, and every other step of the kernel strides over the workgroup's threads.
//
// The code is compiled inside its own namespace (am1g::wg256), so its kernels and helpers are
// distinct symbols from the 512-thread build's.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>

#include "am1g_common.hpp"
#include "am1g_launch.hpp"
#include "am1g_pair.hpp"

#define AM1G_WG_VARIANT 1
#define AM1G_WG_THREADS 256
#define AM1G_LDS_DOUBLES 8704

namespace am1g {
namespace wg256 {
#include "am1g_scf.hip"
}  // namespace wg256

hipError_t launch_scf_batch_small(const BatchArgs& a, int precision, int grid, hipStream_t stream) {
    return wg256::am1g::launch_scf_batch(a, precision, grid, (size_t)AM1G_LDS_DOUBLES * sizeof(double), stream);
}

}  // namespace am1g
