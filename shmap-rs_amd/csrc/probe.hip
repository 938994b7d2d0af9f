// Calibration kernel for HBM traffic counters: a plain copy of a known number
// of bytes.  The FETCH_SIZE / WRITE_SIZE counters of its one dispatch, set
// against the bytes it must move, give the scale that turns the same counters
// of a mapping kernel into bytes.  Used by the benchmark only.
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ void __launch_bounds__(256) k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                    uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

// Copies `bytes` (a multiple of 16) from a fresh source to a fresh destination
// in one dispatch and waits for it.  Returns 0, or the HIP error code.
extern "C" int shm_probe_copy(uint64_t bytes) {
    void *src = nullptr, *dst = nullptr;
    hipError_t e = hipMalloc(&src, bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, bytes);
    if (e == hipSuccess) e = hipMemset(src, 1, bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) {
        k_probe_copy<<<4096, 256>>>((const uint4 *)src, (uint4 *)dst, bytes / 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (src) (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    return (int)e;
}
