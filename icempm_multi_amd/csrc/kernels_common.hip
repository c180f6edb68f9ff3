// kernels_common.hip -- small kernels shared by both precision forms.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace icempm {

namespace {

__global__ void k_fill_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}

}  // namespace

// Zero-fills `bytes` bytes with a kernel (graph-safe, unlike a memset node).
// ptr must be 16-byte aligned.
hipError_t launch_fill_zero(void *ptr, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return hipSuccess;
    const size_t n16 = bytes / 16, ntail = bytes % 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (uint4 *)ptr, n16,
                       (unsigned char *)ptr + n16 * 16, ntail);
    return hipGetLastError();
}

}  // namespace icempm
