// k_fft.hip -- hand-written mixed-radix (2/3/4/5) 3-D FFT for the spectral Ewald mesh.
//
// Replaces the reference's cuFFT C2C plan (cppmodule/ConductorRigid.cc:228,
// cppmodule/ConductorRigid.cu:927,933): unnormalised transforms, forward e^{-i k x},
// inverse e^{+i k x}, z fastest.  Because the mesh data are real and the k-space
// scale S(k) is real and even, the pipeline is real-to-complex along z, complex
// along y and x on the half spectrum [Nx][Ny][Nz/2+1], and complex-to-real back,
// with the scale (cppmodule/ConductorRigid.cu:237-257) fused between the forward
// and inverse x transforms so the spectrum makes one round trip through HBM.
//
// Each workgroup loads a batch of pencils into one LDS buffer (coalesced: along the
// pencil for z; across 16 (FP32) or 8 (FP64) neighbouring z columns -- 128-byte rows --
// for y and x), copies the axis' twiddle table into LDS, and runs the Stockham
// autosort stages (Govindaraju et al., SC'08) in place: each stage reads its
// butterfly inputs into registers, synchronises, and writes the outputs back.
#include "cr_kernels.h"

namespace cr {

namespace {

template <typename T>
struct Cplx;
template <>
struct Cplx<float> { using type = float2; };
template <>
struct Cplx<double> { using type = double2; };

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename T2>
__device__ __forceinline__ T2 twiddle(const T2* tw, int e, int sign)
{
    T2 w = tw[e];
    if (sign > 0) w.y = -w.y;  // table holds e^{-2 pi i e/n}
    return w;
}

constexpr int kThreads = 256;
constexpr int kMaxElems = 4096;  // cols * n per workgroup: <= 16 elements per thread

// One Stockham stage of radix R over `cols` pencils of length n (pitch `pitch`), in place.
template <typename T2, int R, int MAXB>
__device__ __forceinline__ void stage(T2* buf, int n, int pitch, int cols, int Ns, const T2* tw, int sign)
{
    const int nb = n / R;
    const int total = cols * nb;
    const int step = n / (Ns * R);
    T2 v[MAXB][R];
    int cj[MAXB];
#pragma unroll
    for (int k = 0; k < MAXB; ++k) {
        const int b = threadIdx.x + k * kThreads;
        cj[k] = -1;
        if (b < total) {
            const int c = b / nb, j = b - c * nb;
            cj[k] = c * pitch + j;
            const int kk = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                T2 a = buf[c * pitch + j + r * nb];
                if (r > 0) a = cmul(a, twiddle(tw, r * kk * step, sign));
                v[k][r] = a;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MAXB; ++k) {
        if (cj[k] < 0) continue;
        const int b = threadIdx.x + k * kThreads;
        const int c = b / nb, j = b - c * nb;
        const int kk = j % Ns;
        T2* dp = buf + c * pitch + (j / Ns) * Ns * R + kk;
        if (R == 2) {
            T2 o0, o1;
            o0.x = v[k][0].x + v[k][1].x; o0.y = v[k][0].y + v[k][1].y;
            o1.x = v[k][0].x - v[k][1].x; o1.y = v[k][0].y - v[k][1].y;
            dp[0] = o0;
            dp[Ns] = o1;
        } else if (R == 4) {
            T2 a0, a1, a2, a3, b3;
            a0.x = v[k][0].x + v[k][2].x; a0.y = v[k][0].y + v[k][2].y;
            a1.x = v[k][0].x - v[k][2].x; a1.y = v[k][0].y - v[k][2].y;
            a2.x = v[k][1].x + v[k][3].x; a2.y = v[k][1].y + v[k][3].y;
            a3.x = v[k][1].x - v[k][3].x; a3.y = v[k][1].y - v[k][3].y;
            // a3 * (sign * i): W4 = -i forward, +i inverse
            if (sign < 0) { b3.x = a3.y; b3.y = -a3.x; }
            else { b3.x = -a3.y; b3.y = a3.x; }
            T2 o;
            o.x = a0.x + a2.x; o.y = a0.y + a2.y; dp[0] = o;
            o.x = a1.x + b3.x; o.y = a1.y + b3.y; dp[Ns] = o;
            o.x = a0.x - a2.x; o.y = a0.y - a2.y; dp[2 * Ns] = o;
            o.x = a1.x - b3.x; o.y = a1.y - b3.y; dp[3 * Ns] = o;
        } else {
            const int rstep = n / R;
#pragma unroll
            for (int o = 0; o < R; ++o) {
                T2 acc = v[k][0];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const T2 t = cmul(v[k][r], twiddle(tw, ((r * o) % R) * rstep, sign));
                    acc.x += t.x;
                    acc.y += t.y;
                }
                dp[o * Ns] = acc;
            }
        }
    }
    __syncthreads();
}

// All stages of the axis plan, in place in LDS.
template <typename T2>
__device__ void fft_lds(T2* buf, int n, int pitch, int cols, const AxisPlan& plan, const T2* tw, int sign)
{
    int Ns = 1;
    for (int s = 0; s < plan.nstages; ++s) {
        const int R = plan.radix[s];
        if (R == 4) stage<T2, 4, kMaxElems / kThreads / 4>(buf, n, pitch, cols, Ns, tw, sign);
        else if (R == 2) stage<T2, 2, kMaxElems / kThreads / 2>(buf, n, pitch, cols, Ns, tw, sign);
        else if (R == 3) stage<T2, 3, (kMaxElems / kThreads + 2) / 3>(buf, n, pitch, cols, Ns, tw, sign);
        else stage<T2, 5, (kMaxElems / kThreads + 4) / 5>(buf, n, pitch, cols, Ns, tw, sign);
        Ns *= R;
    }
}

// LDS layout: [twiddles n][cols x pitch pencils]
template <typename T2>
__device__ __forceinline__ T2* lds_setup(char* smem_raw, const AxisPlan& plan, T2** tw_out)
{
    T2* tw = reinterpret_cast<T2*>(smem_raw);
    const T2* g = reinterpret_cast<const T2*>(plan.tw);
    for (int e = threadIdx.x; e < plan.n; e += blockDim.x) tw[e] = g[e];
    *tw_out = tw;
    return tw + plan.n;
}

// ---- z: real rows -> half-spectrum rows ----
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_z_r2c(AxisPlan plan, long nrows, int cols, const T* __restrict__ real,
                                                        typename Cplx<T>::type* __restrict__ cplx)
{
    using T2 = typename Cplx<T>::type;
    extern __shared__ char smem_raw[];
    const int n = plan.n, nh = n / 2 + 1, pitch = n + 1;
    T2* tw;
    T2* A = lds_setup<T2>(smem_raw, plan, &tw);
    const long row0 = (long)blockIdx.x * cols;
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, t = idx - c * n;
        const long row = row0 + c;
        T2 v;
        v.x = (row < nrows) ? real[row * n + t] : (T)0;
        v.y = 0;
        A[c * pitch + t] = v;
    }
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, -1);
    for (int idx = threadIdx.x; idx < cols * nh; idx += blockDim.x) {
        const int c = idx / nh, t = idx - c * nh;
        const long row = row0 + c;
        if (row < nrows) cplx[row * nh + t] = A[c * pitch + t];
    }
}

// ---- z: half-spectrum rows -> real rows (Hermitian extension, inverse, real part) ----
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_z_c2r(AxisPlan plan, long nrows, int cols,
                                                        const typename Cplx<T>::type* __restrict__ cplx,
                                                        T* __restrict__ real)
{
    using T2 = typename Cplx<T>::type;
    extern __shared__ char smem_raw[];
    const int n = plan.n, nh = n / 2 + 1, pitch = n + 1;
    T2* tw;
    T2* A = lds_setup<T2>(smem_raw, plan, &tw);
    const long row0 = (long)blockIdx.x * cols;
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, t = idx - c * n;
        const long row = row0 + c;
        T2 v;
        v.x = 0;
        v.y = 0;
        if (row < nrows) {
            if (t < nh) {
                v = cplx[row * nh + t];
            } else {
                v = cplx[row * nh + (n - t)];
                v.y = -v.y;
            }
        }
        A[c * pitch + t] = v;
    }
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, +1);
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, t = idx - c * n;
        const long row = row0 + c;
        if (row < nrows) real[row * n + t] = A[c * pitch + t].x;
    }
}

// ---- z: full complex rows, C2C ----
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_z_c2c(AxisPlan plan, long nrows, int cols,
                                                        typename Cplx<T>::type* __restrict__ data, int sign)
{
    using T2 = typename Cplx<T>::type;
    extern __shared__ char smem_raw[];
    const int n = plan.n, pitch = n + 1;
    T2* tw;
    T2* A = lds_setup<T2>(smem_raw, plan, &tw);
    const long row0 = (long)blockIdx.x * cols;
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, t = idx - c * n;
        const long row = row0 + c;
        T2 v;
        v.x = 0;
        v.y = 0;
        if (row < nrows) v = data[row * n + t];
        A[c * pitch + t] = v;
    }
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, sign);
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, t = idx - c * n;
        const long row = row0 + c;
        if (row < nrows) data[row * n + t] = A[c * pitch + t];
    }
}

// Strided pencils along y (axis 1) or x (axis 0) of [Nx][Ny][Zc]; a workgroup
// takes `cols` neighbouring z columns of one (x or y) line.
struct StridedLayout {
    long base;    // offset of column z0
    long stride;  // between consecutive pencil elements
    int z0;
    int line;
};

__device__ __forceinline__ StridedLayout strided_layout(int axis, int Ny, int Zc, int cols)
{
    StridedLayout s;
    const int zblocks = (Zc + cols - 1) / cols;
    s.line = blockIdx.x / zblocks;
    s.z0 = (blockIdx.x - s.line * zblocks) * cols;
    if (axis == 1) {  // pencil over y at fixed x = line
        s.base = (long)s.line * Ny * Zc + s.z0;
        s.stride = Zc;
    } else {          // pencil over x at fixed y = line
        s.base = (long)s.line * Zc + s.z0;
        s.stride = (long)Ny * Zc;
    }
    return s;
}

template <typename T2>
__device__ __forceinline__ void load_strided(const T2* __restrict__ data, T2* A, int n, int pitch, int cols, int Zc,
                                             const StridedLayout& L)
{
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int t = idx / cols, c = idx - t * cols;
        T2 v;
        v.x = 0;
        v.y = 0;
        if (L.z0 + c < Zc) v = data[L.base + c + t * L.stride];
        A[c * pitch + t] = v;
    }
}

template <typename T2>
__device__ __forceinline__ void store_strided(T2* __restrict__ data, const T2* A, int n, int pitch, int cols, int Zc,
                                              const StridedLayout& L)
{
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int t = idx / cols, c = idx - t * cols;
        if (L.z0 + c < Zc) data[L.base + c + t * L.stride] = A[c * pitch + t];
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_strided(AxisPlan plan, int axis, int Ny, int Zc, int cols,
                                                          typename Cplx<T>::type* __restrict__ data, int sign)
{
    using T2 = typename Cplx<T>::type;
    extern __shared__ char smem_raw[];
    const int n = plan.n, pitch = n + 1;
    T2* tw;
    T2* A = lds_setup<T2>(smem_raw, plan, &tw);
    const StridedLayout L = strided_layout(axis, Ny, Zc, cols);
    load_strided(data, A, n, pitch, cols, Zc, L);
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, sign);
    store_strided(data, A, n, pitch, cols, Zc, L);
}

template <typename T>
__device__ __forceinline__ T sin_t(T x);
template <>
__device__ __forceinline__ float sin_t<float>(float x) { return sinf(x); }
template <>
__device__ __forceinline__ double sin_t<double>(double x) { return sin(x); }

// Forward x FFT, multiply by S(k) = j0(|k|)^2 exp(-sum_d (1-eta_d) k_d^2/4xi^2)/|k|^2/G
// (cppmodule/ConductorRigid.cc:253-268; S(0) = 0), inverse x FFT.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fft_x_scale(AxisPlan plan, int Ny, int Zc, int cols,
                                                          typename Cplx<T>::type* __restrict__ data, ScaleTables sc)
{
    using T2 = typename Cplx<T>::type;
    extern __shared__ char smem_raw[];
    const int n = plan.n, pitch = n + 1;
    T2* tw;
    T2* A = lds_setup<T2>(smem_raw, plan, &tw);
    const StridedLayout L = strided_layout(0, Ny, Zc, cols);
    const int y = L.line;
    load_strided(data, A, n, pitch, cols, Zc, L);
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, -1);
    const double ky = sc.k[1][y], ey = sc.e[1][y];
    for (int idx = threadIdx.x; idx < cols * n; idx += blockDim.x) {
        const int c = idx / n, x = idx - c * n;
        const int z = L.z0 + c;
        if (z >= Zc) continue;
        const T kx = (T)sc.k[0][x], kyt = (T)ky, kz = (T)sc.k[2][z];
        const T k2 = kx * kx + kyt * kyt + kz * kz;
        T S = 0;
        if (k2 > (T)0) {
            const T kk = sqrt(k2);
            const T j0 = sin_t<T>(kk) / kk;
            S = j0 * j0 * (T)(sc.e[0][x] * ey * sc.e[2][z]) / k2 * (T)sc.invG;
        }
        T2 v = A[c * pitch + x];
        v.x *= S;
        v.y *= S;
        A[c * pitch + x] = v;
    }
    __syncthreads();
    fft_lds<T2>(A, n, pitch, cols, plan, tw, +1);
    store_strided(data, A, n, pitch, cols, Zc, L);
}

// Pencils per workgroup: 128-byte rows for the strided passes (16 float2 / 8 double2
// columns), capped so that cols * n <= 4096 elements (16 per thread in registers).
template <typename T>
size_t lds_bytes(int n, int cols)
{
    return ((size_t)(n + 1) * (size_t)cols + (size_t)n) * 2 * sizeof(T);
}

template <typename T>
int pick_cols(int n)
{
    int cols = (sizeof(T) == 4) ? 16 : 8;
    while (cols > 1 && (cols * n > kMaxElems || lds_bytes<T>(n, cols) > 64 * 1024)) cols >>= 1;
    return cols;
}

template <typename T>
bool fits(int n)
{
    return n <= kMaxElems && lds_bytes<T>(n, pick_cols<T>(n)) <= 64 * 1024;
}

}  // namespace

template <typename T>
hipError_t launch_fft_z_r2c(const AxisPlan& pz, int Nx, int Ny, const T* real, T* cplx, int zp, hipStream_t st)
{
    using T2 = typename Cplx<T>::type;
    const long nrows = (long)Nx * Ny;
    if (fftr_radix(pz.n)) return launch_fftr_z_r2c<T>(pz, nrows, zp, real, cplx, st);
    if (!fits<T>(pz.n) || zp != pz.n / 2 + 1) return hipErrorInvalidValue;
    const int cols = pick_cols<T>(pz.n);
    hipLaunchKernelGGL(k_fft_z_r2c<T>, dim3((unsigned)((nrows + cols - 1) / cols)), dim3(kThreads),
                       lds_bytes<T>(pz.n, cols), st, pz, nrows, cols, real, (T2*)cplx);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_z_c2r(const AxisPlan& pz, int Nx, int Ny, const T* cplx, int zp, T* real, hipStream_t st)
{
    using T2 = typename Cplx<T>::type;
    const long nrows = (long)Nx * Ny;
    if (fftr_radix(pz.n)) return launch_fftr_z_c2r<T>(pz, nrows, zp, cplx, real, st);
    if (!fits<T>(pz.n) || zp != pz.n / 2 + 1) return hipErrorInvalidValue;
    const int cols = pick_cols<T>(pz.n);
    hipLaunchKernelGGL(k_fft_z_c2r<T>, dim3((unsigned)((nrows + cols - 1) / cols)), dim3(kThreads),
                       lds_bytes<T>(pz.n, cols), st, pz, nrows, cols, (const T2*)cplx, real);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_z_c2c(const AxisPlan& pz, int Nx, int Ny, T* cplx, int sign, hipStream_t st)
{
    using T2 = typename Cplx<T>::type;
    const long nrows = (long)Nx * Ny;
    if (fftr_radix(pz.n)) return launch_fftr_z_c2c<T>(pz, nrows, cplx, sign, st);
    if (!fits<T>(pz.n)) return hipErrorInvalidValue;
    const int cols = pick_cols<T>(pz.n);
    hipLaunchKernelGGL(k_fft_z_c2c<T>, dim3((unsigned)((nrows + cols - 1) / cols)), dim3(kThreads),
                       lds_bytes<T>(pz.n, cols), st, pz, nrows, cols, (T2*)cplx, sign);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_yx(const AxisPlan& pa, int axis, int Nx, int Ny, int Zc, int Zp, T* cplx, int sign,
                         hipStream_t st)
{
    using T2 = typename Cplx<T>::type;
    if (fftr_radix(pa.n)) return launch_fftr_strided<T>(pa, axis, Nx, Ny, Zc, Zp, cplx, sign, st);
    if (!fits<T>(pa.n) || Zp != Zc) return hipErrorInvalidValue;
    const int cols = pick_cols<T>(pa.n);
    const int lines = (axis == 1) ? Nx : Ny;
    const int zblocks = (Zc + cols - 1) / cols;
    hipLaunchKernelGGL(k_fft_strided<T>, dim3((unsigned)(lines * zblocks)), dim3(kThreads), lds_bytes<T>(pa.n, cols),
                       st, pa, axis, Ny, Zc, cols, (T2*)cplx, sign);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fft_x_scale(const AxisPlan& px, int Nx, int Ny, int Zc, int Zp, T* cplx, const ScaleTables& sc,
                              hipStream_t st)
{
    using T2 = typename Cplx<T>::type;
    (void)Nx;
    if (fftr_radix(px.n)) return launch_fftr_x_scale<T>(px, Ny, Zc, Zp, cplx, sc, st);
    if (!fits<T>(px.n) || Zp != Zc) return hipErrorInvalidValue;
    const int cols = pick_cols<T>(px.n);
    const int zblocks = (Zc + cols - 1) / cols;
    hipLaunchKernelGGL(k_fft_x_scale<T>, dim3((unsigned)(Ny * zblocks)), dim3(kThreads), lds_bytes<T>(px.n, cols), st,
                       px, Ny, Zc, cols, (T2*)cplx, sc);
    return hipGetLastError();
}

#define CR_FFT_INST(T)                                                                                        \
    template hipError_t launch_fft_z_r2c<T>(const AxisPlan&, int, int, const T*, T*, int, hipStream_t);       \
    template hipError_t launch_fft_z_c2r<T>(const AxisPlan&, int, int, const T*, int, T*, hipStream_t);       \
    template hipError_t launch_fft_z_c2c<T>(const AxisPlan&, int, int, T*, int, hipStream_t);                 \
    template hipError_t launch_fft_yx<T>(const AxisPlan&, int, int, int, int, int, T*, int, hipStream_t);     \
    template hipError_t launch_fft_x_scale<T>(const AxisPlan&, int, int, int, int, T*, const ScaleTables&,    \
                                              hipStream_t);
CR_FFT_INST(float)
CR_FFT_INST(double)

}  // namespace cr
