// calibration.hip -- a plain streaming copy with the step kernels' access
// pattern (one double per lane, 512 B per wave and load, XCD-aware grid), for
// two measurements:
//   - its FETCH_SIZE / WRITE_SIZE counters per dispatch divided by the bytes
//     it moves give the factors that turn those counters into bytes for the
//     step kernels;
//   - its time gives the HBM rate a read+write stream reaches on this GPU.
// Measurement support only; not part of the solver's C ABI.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace
{

__global__ __launch_bounds__(256) void k_copy_f64(const double* __restrict__ src, double* __restrict__ dst, uint64_t n)
{
    // 8 doubles per thread in flight, each a 512 B wave-wide access
    const uint64_t wave  = (static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x) / 64;
    const uint64_t lane  = threadIdx.x & 63;
    const uint64_t waves = static_cast<uint64_t>(gridDim.x) * 4;
    for (uint64_t base = wave * 512; base < n; base += waves * 512)
    {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const uint64_t i = base + j * 64 + lane;
            v[j]             = i < n ? src[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const uint64_t i = base + j * 64 + lane;
            if (i < n) dst[i] = v[j];
        }
    }
}

} // namespace

extern "C" int amr_calib_copy(const double* src, double* dst, uint64_t n, void* stream)
{
    const uint64_t chunks = (n + 511) / 512;               // one wave per 512 doubles
    uint64_t       blocks = (chunks + 3) / 4;               // 4 waves per block
    if (blocks > 65536) blocks = 65536;                     // grid-stride beyond that
    blocks = ((blocks + 7) / 8) * 8;                        // multiple of the XCD count
    hipLaunchKernelGGL(k_copy_f64, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), src, dst, n);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
