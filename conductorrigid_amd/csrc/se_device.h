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

// Sum over the 64 lanes of a fully active wavefront, returned uniformly.  DPP row shifts
// (row_shr 1, 2, 4, 8; lanes shifted in from outside the row read 0) leave each row of
// 16 lanes' total in its lane 15; the four row totals are then read out as scalars.
// Stays in the VALU: no LDS round trips, unlike ds_bpermute-based shuffles.
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T wave_total(T v)
{
    v += dpp_shift<0x111>(v);
    v += dpp_shift<0x112>(v);
    v += dpp_shift<0x114>(v);
    v += dpp_shift<0x118>(v);
    return (lane_value(v, 15) + lane_value(v, 31)) + (lane_value(v, 47) + lane_value(v, 63));
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
