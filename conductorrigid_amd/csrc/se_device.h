// se_device.h -- device helpers shared by the spectral-Ewald spreading/contraction kernels.
#pragma once

#include "cr_kernels.h"

namespace cr {
namespace se {

constexpr int kWave = 64;

__device__ __forceinline__ int wrap_idx(int x, int n) { return (x < 0) ? x + n : ((x > n - 1) ? x - n : x); }

// First window node along one axis, with the reference's rule for odd P
// (cppmodule/ConductorRigid.cu:193-210): frac = (r - lo)/L, pf = frac*N, x = int(pf).
__device__ __forceinline__ int window_start(double r, double L, int N, int P)
{
    const double frac = (r - (-0.5 * L)) / L;
    const double pf = frac * (double)N;
    const int x = (int)pf;
    return x - P / 2 + 1 - (P % 2) * ((pf - (double)x) < 0.5);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T gexp(T a);
template <>
__device__ __forceinline__ float gexp<float>(float a) { return __expf(a); }
template <>
__device__ __forceinline__ double gexp<double>(double a) { return exp(a); }

template <typename T>
__device__ __forceinline__ void atomic_add(T* p, T v) { unsafeAtomicAdd(p, v); }

}  // namespace se
}  // namespace cr
