// k_se.hip -- Gaussian spreading and contraction of the spectral Ewald method.
//
// Reference: spread (cppmodule/ConductorRigid.cu:144-234), contract (:260-360),
// contract_force (:363-466).  The reference launches one P x P x P thread block per
// bead and does one float atomicAdd per window node.  Here one 64-lane wavefront
// owns a bead: lanes cover the (y, z) window plane and walk x, the Gaussian is
// evaluated separably (exp(a+b) = exp(a) exp(b)), and contraction reduces across
// the wavefront with cross-lane shuffles instead of a shared-memory tree.
#include "cr_kernels.h"

#include <cstdint>
#include <cstring>

namespace cr {

namespace {

constexpr int kWave = 64;
constexpr int kBeadsPerBlock = 4;

__device__ __forceinline__ int wrap_idx(int x, int n)
{
    return (x < 0) ? x + n : ((x > n - 1) ? x - n : x);
}

// First window node along one axis, with the reference's rule for odd P
// (cppmodule/ConductorRigid.cu:193-210): frac = (r - lo)/L, pf = frac*N.
__device__ __forceinline__ int window_start(double r, double L, int N, int P)
{
    double frac = (r - (-0.5 * L)) / L;
    double pf = frac * (double)N;
    int x = (int)pf;
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

template <typename T>
__global__ __launch_bounds__(kWave * kBeadsPerBlock) void k_spread_atomic(MeshParams mp, int N,
                                                                         const double4* __restrict__ pos,
                                                                         const double* __restrict__ q,
                                                                         T* __restrict__ grid)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int bead = blockIdx.x * kBeadsPerBlock + (threadIdx.x >> 6);
    if (bead >= N) return;
    const double4 r = pos[bead];
    const int P = mp.P;
    const int sx = window_start(r.x, mp.L[0], mp.Nx, P);
    const int sy = window_start(r.y, mp.L[1], mp.Ny, P);
    const int sz = window_start(r.z, mp.L[2], mp.Nz, P);
    const T qb = (T)q[bead];
    const long plane = (long)mp.Ny * mp.Nz;
    for (int pair = lane; pair < P * P; pair += kWave) {
        const int ty = pair / P, tz = pair - ty * P;
        const int yi = sy + ty, zi = sz + tz;
        const double ry = mp.h[1] * yi - 0.5 * mp.L[1] - r.y;
        const double rz = mp.h[2] * zi - 0.5 * mp.L[2] - r.z;
        const T eyz = gexp<T>(-(T)mp.xiterm * (T)(ry * ry / mp.eta[1] + rz * rz / mp.eta[2]));
        const T c = (T)mp.prefac * eyz * qb;
        const long col = (long)wrap_idx(yi, mp.Ny) * mp.Nz + wrap_idx(zi, mp.Nz);
        for (int tx = 0; tx < P; ++tx) {
            const int xi = sx + tx;
            const double rx = mp.h[0] * xi - 0.5 * mp.L[0] - r.x;
            const T ex = gexp<T>(-(T)mp.xiterm * (T)(rx * rx / mp.eta[0]));
            atomic_add(&grid[(long)wrap_idx(xi, mp.Nx) * plane + col], c * ex);
        }
    }
}

template <typename T, bool FORCE>
__global__ __launch_bounds__(kWave * kBeadsPerBlock) void k_contract(MeshParams mp, int N,
                                                                    const double4* __restrict__ pos,
                                                                    const double* __restrict__ q,
                                                                    const T* __restrict__ grid,
                                                                    T* __restrict__ phi,
                                                                    double* __restrict__ force,
                                                                    const double* __restrict__ fadd)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int bead = blockIdx.x * kBeadsPerBlock + (threadIdx.x >> 6);
    if (bead >= N) return;
    const double4 r = pos[bead];
    const int P = mp.P;
    const int sx = window_start(r.x, mp.L[0], mp.Nx, P);
    const int sy = window_start(r.y, mp.L[1], mp.Ny, P);
    const int sz = window_start(r.z, mp.L[2], mp.Nz, P);
    const long plane = (long)mp.Ny * mp.Nz;
    T acc = 0, fx = 0, fy = 0, fz = 0;
    for (int pair = lane; pair < P * P; pair += kWave) {
        const int ty = pair / P, tz = pair - ty * P;
        const int yi = sy + ty, zi = sz + tz;
        // r = bead - node (cppmodule/ConductorRigid.cu:332,438)
        const double ry = r.y - (mp.h[1] * yi - 0.5 * mp.L[1]);
        const double rz = r.z - (mp.h[2] * zi - 0.5 * mp.L[2]);
        const T eyz = gexp<T>(-(T)mp.xiterm * (T)(ry * ry / mp.eta[1] + rz * rz / mp.eta[2]));
        const long col = (long)wrap_idx(yi, mp.Ny) * mp.Nz + wrap_idx(zi, mp.Nz);
        T sxe = 0, sxr = 0;
        for (int tx = 0; tx < P; ++tx) {
            const int xi = sx + tx;
            const double rx = r.x - (mp.h[0] * xi - 0.5 * mp.L[0]);
            const T ex = gexp<T>(-(T)mp.xiterm * (T)(rx * rx / mp.eta[0]));
            const T g = grid[(long)wrap_idx(xi, mp.Nx) * plane + col];
            sxe += ex * g;
            if (FORCE) sxr += ex * g * (T)(rx / mp.eta[0]);
        }
        acc += eyz * sxe;
        if (FORCE) {
            fx += eyz * sxr;
            fy += eyz * sxe * (T)(ry / mp.eta[1]);
            fz += eyz * sxe * (T)(rz / mp.eta[2]);
        }
    }
    if (!FORCE) {
        acc = wave_sum(acc);
        if (lane == 0) phi[bead] = acc * (T)(mp.quadW * mp.prefac);
    } else {
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        fz = wave_sum(fz);
        if (lane == 0) {
            // F = 2 (h^3 prefac) (2 xi^2) q_i sum phi_g e^{...} r/eta  (cppmodule/ConductorRigid.cu:449)
            const double c = 2.0 * (mp.quadW * mp.prefac) * mp.xiterm * q[bead];
            const long o = 4 * (long)bead;
            // fadd: the real-space force of the same bead, computed beside this pass
            const double a0 = fadd ? fadd[o] : 0.0, a1 = fadd ? fadd[o + 1] : 0.0;
            const double a2 = fadd ? fadd[o + 2] : 0.0;
            force[o + 0] = c * (double)fx + a0;
            force[o + 1] = c * (double)fy + a1;
            force[o + 2] = c * (double)fz + a2;
            force[o + 3] = 0.0;
        }
    }
}

template <typename T>
__global__ void k_fill(T* p, T v, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// 16 bytes per store (the bit pattern of 4 floats or 2 doubles)
__global__ void k_fill16(uint4* p, uint4 v, long n16)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) p[i] = v;
}

template <typename Tin, typename Tout>
__global__ void k_convert(const Tin* in, Tout* out, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (Tout)in[i];
}

inline unsigned grid_1d(long n, int block)
{
    long b = (n + block - 1) / block;
    if (b > 65536L * 16) b = 65536L * 16;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

template <typename T>
hipError_t launch_spread(const MeshParams& mp, const SysView& s, const double* q, T* grid, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    unsigned nb = (unsigned)((s.N + kBeadsPerBlock - 1) / kBeadsPerBlock);
    hipLaunchKernelGGL(k_spread_atomic<T>, dim3(nb), dim3(kWave * kBeadsPerBlock), 0, st, mp, s.N, s.pos, q, grid);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_contract(const MeshParams& mp, const SysView& s, const T* grid, T* phi, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    unsigned nb = (unsigned)((s.N + kBeadsPerBlock - 1) / kBeadsPerBlock);
    hipLaunchKernelGGL((k_contract<T, false>), dim3(nb), dim3(kWave * kBeadsPerBlock), 0, st, mp, s.N, s.pos,
                       (const double*)nullptr, grid, phi, (double*)nullptr, (const double*)nullptr);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_contract_force(const MeshParams& mp, const SysView& s, const double* q, const T* grid,
                                 double* force, hipStream_t st, const double* fadd)
{
    if (s.N <= 0) return hipSuccess;
    unsigned nb = (unsigned)((s.N + kBeadsPerBlock - 1) / kBeadsPerBlock);
    hipLaunchKernelGGL((k_contract<T, true>), dim3(nb), dim3(kWave * kBeadsPerBlock), 0, st, mp, s.N, s.pos, q,
                       grid, (T*)nullptr, force, fadd);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(T* p, T v, long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    constexpr long per16 = 16 / (long)sizeof(T);
    if constexpr (sizeof(T) == 4 || sizeof(T) == 8) {
        if (reinterpret_cast<uintptr_t>(p) % 16 == 0 && n >= 4096) {
            // wide stores for the bulk, element stores for the tail
            unsigned w[4];
            for (int k = 0; k < 4; ++k) std::memcpy(&w[k], reinterpret_cast<const char*>(&v) + (4 * k) % sizeof(T), 4);
            const long n16 = n / per16;
            hipLaunchKernelGGL(k_fill16, dim3(grid_1d(n16, 256)), dim3(256), 0, st, reinterpret_cast<uint4*>(p),
                               make_uint4(w[0], w[1], w[2], w[3]), n16);
            const long done = n16 * per16;
            if (done < n)
                hipLaunchKernelGGL(k_fill<T>, dim3(1), dim3(256), 0, st, p + done, v, n - done);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_1d(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

template <typename Tin, typename Tout>
hipError_t launch_convert(const Tin* in, Tout* out, long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3(grid_1d(n, 256)), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

template hipError_t launch_spread<float>(const MeshParams&, const SysView&, const double*, float*, hipStream_t);
template hipError_t launch_spread<double>(const MeshParams&, const SysView&, const double*, double*, hipStream_t);
template hipError_t launch_contract<float>(const MeshParams&, const SysView&, const float*, float*, hipStream_t);
template hipError_t launch_contract<double>(const MeshParams&, const SysView&, const double*, double*, hipStream_t);
template hipError_t launch_contract_force<float>(const MeshParams&, const SysView&, const double*, const float*,
                                                 double*, hipStream_t, const double*);
template hipError_t launch_contract_force<double>(const MeshParams&, const SysView&, const double*, const double*,
                                                  double*, hipStream_t, const double*);
// k-space factor tables of one axis (SetParams, cppmodule/ConductorRigid.cc:245-274):
// k_i = 2 pi (i < (n+1)/2 ? i : i - n) / L and e_i = exp(-(1 - eta) k_i^2 / (4 xi^2)).
__global__ void k_kspace_axis(int n, double L, double xi, double eta, double* __restrict__ kt, double* __restrict__ et)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double k = (double)((i < (n + 1) / 2) ? i : i - n) * 2.0 * 3.14159265358979323846 / L;
    kt[i] = k;
    et[i] = exp(-(1.0 - eta) * k * k / (4.0 * xi * xi));
}

// Cubic boxes: j0(|k|)^2 / k^2 / G over m2 = |m|^2, k^2 = kappa^2 m2 (0 at m2 = 0), times
// exp(-gauss m2) when gauss >= 0.
template <typename T>
__global__ void k_kspace_radial(long m2max, double kappa, double invG, T* __restrict__ rad, double gauss)
{
    const long m2 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m2 > m2max) return;
    if (m2 == 0) {
        rad[0] = (T)0;
        return;
    }
    const double k2 = kappa * kappa * (double)m2, k = sqrt(k2);
    const double j0 = sin(k) / k;
    const double g = gauss >= 0.0 ? exp(-gauss * (double)m2) : 1.0;
    rad[m2] = (T)(j0 * j0 / k2 * invG * g);
}

hipError_t launch_kspace_axis(int n, double L, double xi, double eta, double* k, double* e, hipStream_t st)
{
    hipLaunchKernelGGL(k_kspace_axis, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, L, xi, eta, k, e);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_kspace_radial(long m2max, double kappa, double invG, T* rad, hipStream_t st, double gauss)
{
    hipLaunchKernelGGL(k_kspace_radial<T>, dim3((unsigned)((m2max + 256) / 256)), dim3(256), 0, st, m2max, kappa,
                       invG, rad, gauss);
    return hipGetLastError();
}
template hipError_t launch_kspace_radial<float>(long, double, double, float*, hipStream_t, double);
template hipError_t launch_kspace_radial<double>(long, double, double, double*, hipStream_t, double);

template hipError_t launch_fill<float>(float*, float, long, hipStream_t);
template hipError_t launch_fill<double>(double*, double, long, hipStream_t);
template hipError_t launch_fill<unsigned>(unsigned*, unsigned, long, hipStream_t);
template hipError_t launch_fill<int>(int*, int, long, hipStream_t);
template hipError_t launch_convert<double, float>(const double*, float*, long, hipStream_t);
template hipError_t launch_convert<float, double>(const float*, double*, long, hipStream_t);
template hipError_t launch_convert<double, double>(const double*, double*, long, hipStream_t);

}  // namespace cr
