// Device helpers for gfx950 (wave64).
#pragma once
#include <hip/hip_runtime.h>

namespace aldic {

constexpr int kWave = 64;

// Butterfly all-reduce across the 64 lanes of a wavefront.  Every lane ends with
// the bitwise-identical sum (a + b == b + a at each step), so decisions taken on the
// result are uniform across the wave.
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Orders this wavefront's LDS writes before its later LDS reads from other lanes.
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Keys (1981) cubic convolution weights, a = -1/2, for the taps at
// floor-1, floor, floor+1, floor+2 of a sample with fractional part d.
template <class T>
__device__ __forceinline__ void keys_weights(T d, T &w0, T &w1, T &w2, T &w3)
{
    const T d2 = d * d, d3 = d2 * d;
    w0 = T(0.5) * (-d + T(2) * d2 - d3);
    w1 = T(0.5) * (T(2) - T(5) * d2 + T(3) * d3);
    w2 = T(0.5) * (d + T(4) * d2 - T(3) * d3);
    w3 = T(0.5) * (-d2 + d3);
}

}  // namespace aldic
