// k_fft_reg.hip -- register-resident FFT passes for pencil lengths n = A * B with
// A = B or A = 2B (powers of two, A <= 32): meshes of 32, 64, 128, 256, 512 or 1024
// nodes per axis.
//
// Four-step decomposition (Bailey): with j = j2 + B j1 and k = k1 + A k2,
//   X[k1 + A k2] = sum_j2 W_B^{j2 k2} [ W_n^{j2 k1} sum_j1 x[j2 + B j1] W_A^{j1 k1} ].
// B threads own one pencil.  Thread j2 loads x[j2 + B j1] (j1 = 0..A-1; for a fixed j1
// the B threads read B consecutive elements), runs an A-point DFT in registers with
// compile-time roots, multiplies by the pass twiddle W_n^{j2 k1}, and hands the values
// over through one padded LDS transpose; thread t then runs the B-point DFTs of
// k1 = t, t + B (A / B of them) and holds X[k1 + A k2].  One barrier per pass instead of
// two per radix stage.  The inverse from that layout (used by the C2R pass and the fused
// x pass) runs the same steps in the other order, ending with thread t holding
// x[t + B j1].
//
// Unnormalised, forward e^{-ikx}, inverse e^{+ikx}, as the generic passes in k_fft.hip
// (which replace the reference's cuFFT plan, cppmodule/ConductorRigid.cc:228).
#include "cr_kernels.h"

namespace cr {

namespace {

template <typename T>
struct Cx;
template <>
struct Cx<float> { using t = float2; };
template <>
struct Cx<double> { using t = double2; };

// cos(2 pi k / 32) for any integer k: cos(pi a / 16), a = 0..8, by symmetry
__device__ __forceinline__ double cos32(int k)
{
    constexpr double c[9] = {1.0,
                             0.98078528040323044913,
                             0.92387953251128675613,
                             0.83146961230254523708,
                             0.70710678118654752440,
                             0.55557023301960222474,
                             0.38268343236508977173,
                             0.19509032201612826785,
                             0.0};
    k &= 31;
    const int a = k <= 16 ? k : 32 - k;
    return a <= 8 ? c[a] : -c[16 - a];
}
__device__ __forceinline__ double sin32(int k) { return cos32(k - 8); }

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <int R>
__device__ __forceinline__ constexpr int bitrev(int i)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

// In-register R-point DFT, natural order in and out, root e^{SIGN 2 pi i / R}
// (radix-2 decimation in time; roots are compile-time constants after unrolling).
template <int R, int SIGN, typename T2>
__device__ __forceinline__ void dft(T2 (&v)[R])
{
    using T = decltype(v[0].x);
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int j = bitrev<R>(i);
        if (j > i) {
            const T2 t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; ++k) {
                const int m = k * (32 / len);
                const T2 a = v[i + k];
                T2 b = v[i + k + len / 2];
                if (m == 8) {  // root = SIGN * i
                    const T bx = b.x;
                    b.x = -(T)SIGN * b.y;
                    b.y = (T)SIGN * bx;
                } else if (m != 0) {
                    T2 w;
                    w.x = (T)cos32(m);
                    w.y = (T)(SIGN * sin32(m));
                    b = cmul(b, w);
                }
                v[i + k].x = a.x + b.x;
                v[i + k].y = a.y + b.y;
                v[i + k + len / 2].x = a.x - b.x;
                v[i + k + len / 2].y = a.y - b.y;
            }
        }
    }
}

// pass twiddle W_n^{e} (table holds e^{-2 pi i e / n}); conjugated for SIGN = +1
template <int SIGN, typename T2>
__device__ __forceinline__ T2 ptw(const T2* __restrict__ tw, int e)
{
    T2 w = tw[e];
    if (SIGN > 0) w.y = -w.y;
    return w;
}

// Pencils (or columns) per workgroup for a pencil of B threads and an A x (B+1)
// transpose buffer: as many as fit ~40 KB of LDS, at most 256 threads.
template <typename T, int A, int B>
constexpr int lanes_per_wg()
{
    constexpr int fit = 40960 / (int)(A * (B + 1) * 2 * sizeof(T));
    constexpr int cap = 256 / B;
    int p = 1;
    while (p * 2 <= fit && p * 2 <= cap) p *= 2;
    return p;
}

// Columns per workgroup of the strided passes (C columns x A*B values in the transpose
// buffer): as many as fit 40 KB (float) or 80 KB (double) of LDS, at most 512 threads, so
// that a column group is 128 or 256 bytes of each row.  On rows padded to whole 128-byte
// lines, 16 float columns (one line, twice the workgroups per CU) measured faster than 32
// at 256^3 (y pass 34.2 -> 32.9 us, x pass 52.2 -> 49.3 us); on the odd pitch Nz/2+1
// (each group split over partial lines) 32 had been faster than 16.
template <typename T, int A, int B>
constexpr int cols_per_wg()
{
    constexpr int fit = (sizeof(T) == 4 ? 40960 : 81920) / (int)(A * B * 2 * sizeof(T));
    constexpr int cap = 512 / B;
    int p = 1;
    while (p * 2 <= fit && p * 2 <= cap) p *= 2;
    return p;
}

// ---- contiguous pencils (z) ---------------------------------------------------
// n = A * B with A = m * B.  Forward: thread j2 (of B) holds x[j2 + B j1], j1 = 0..A-1;
// an A-point DFT over j1 gives k1, the pass twiddle W_n^{j2 k1} is applied, and the
// values go through a padded LDS transpose; thread t then runs m B-point DFTs over j2,
// for k1 = t + B s (s = 0..m-1), and holds v[s B + k2] = X[(t + B s) + A k2].
// buf: this pencil's A x (B+1) transpose buffer.
template <int A, int B, int SIGN, typename T2>
__device__ __forceinline__ void four_step(T2 (&v)[A], T2* buf, const T2* __restrict__ tw, int t)
{
    constexpr int m = A / B;
    dft<A, SIGN>(v);
#pragma unroll
    for (int q = 0; q < A; ++q) buf[q * (B + 1) + t] = q == 0 ? v[q] : cmul(v[q], ptw<SIGN>(tw, t * q));
    __syncthreads();
#pragma unroll
    for (int s = 0; s < m; ++s) {
        T2 u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) u[j] = buf[(t + B * s) * (B + 1) + j];
        dft<B, SIGN>(u);
#pragma unroll
        for (int j = 0; j < B; ++j) v[s * B + j] = u[j];
    }
}

// Inverse from the X layout of four_step (thread t holds v[s B + k2] = X[(t + B s) + A k2])
// to the x layout (thread t holds x[t + B j1] in v[j1]):
//   x[j2 + B j1] = sum_k1 W_A^{j1 k1} W_n^{j2 k1} sum_k2 X[k1 + A k2] W_B^{j2 k2}  (+ roots).
template <int A, int B, typename T2>
__device__ __forceinline__ void four_step_inv_from_X(T2 (&v)[A], T2* buf, const T2* __restrict__ tw, int t)
{
    constexpr int m = A / B;
#pragma unroll
    for (int s = 0; s < m; ++s) {
        T2 u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) u[j] = v[s * B + j];
        dft<B, +1>(u);
        const int k1 = t + B * s;
#pragma unroll
        for (int j = 0; j < B; ++j) buf[k1 * (B + 1) + j] = (j == 0 || k1 == 0) ? u[j] : cmul(u[j], ptw<+1>(tw, j * k1));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < A; ++q) v[q] = buf[q * (B + 1) + t];
    dft<A, +1>(v);
}

template <typename T, int A, int B>
__global__ __launch_bounds__((B * lanes_per_wg<T, A, B>())) void k_fftr_z_r2c(long nrows, int zp,
                                                                            const T* __restrict__ real,
                                                                            typename Cx<T>::t* __restrict__ cplx,
                                                                            const typename Cx<T>::t* __restrict__ tw)
{
    using T2 = typename Cx<T>::t;
    constexpr int n = A * B, nh = n / 2 + 1, m = A / B;
    __shared__ T2 lds[lanes_per_wg<T, A, B>() * A * (B + 1)];
    const int p = threadIdx.x / B, l = threadIdx.x % B;
    const long row = (long)blockIdx.x * lanes_per_wg<T, A, B>() + p;
    const bool ok = row < nrows;
    T2 v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
        v[q].x = ok ? real[row * n + l + B * q] : (T)0;
        v[q].y = 0;
    }
    four_step<A, B, -1>(v, lds + p * A * (B + 1), tw, l);
    if (!ok) return;
#pragma unroll
    for (int s = 0; s < m; ++s)
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = l + B * s + A * j;
            if (k < nh) cplx[row * zp + k] = v[s * B + j];
        }
}

template <typename T, int A, int B>
__global__ __launch_bounds__((B * lanes_per_wg<T, A, B>())) void k_fftr_z_c2r(long nrows, int zp,
                                                                            const typename Cx<T>::t* __restrict__ cplx,
                                                                            T* __restrict__ real,
                                                                            const typename Cx<T>::t* __restrict__ tw)
{
    using T2 = typename Cx<T>::t;
    constexpr int n = A * B, nh = n / 2 + 1, m = A / B;
    __shared__ T2 lds[lanes_per_wg<T, A, B>() * A * (B + 1)];
    const int p = threadIdx.x / B, l = threadIdx.x % B;
    const long row = (long)blockIdx.x * lanes_per_wg<T, A, B>() + p;
    const bool ok = row < nrows;
    T2 v[A];
#pragma unroll
    for (int s = 0; s < m; ++s)
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int k = l + B * s + A * j;
            T2 z;
            z.x = 0;
            z.y = 0;
            if (ok) {
                if (k < nh) {
                    z = cplx[row * zp + k];
                } else {  // Hermitian extension
                    z = cplx[row * zp + (n - k)];
                    z.y = -z.y;
                }
            }
            v[s * B + j] = z;
        }
    four_step_inv_from_X<A, B>(v, lds + p * A * (B + 1), tw, l);
    if (!ok) return;
#pragma unroll
    for (int q = 0; q < A; ++q) real[row * n + l + B * q] = v[q].x;
}

template <typename T, int A, int B, int SIGN>
__global__ __launch_bounds__((B * lanes_per_wg<T, A, B>())) void k_fftr_z_c2c(long nrows,
                                                                            typename Cx<T>::t* __restrict__ data,
                                                                            const typename Cx<T>::t* __restrict__ tw)
{
    using T2 = typename Cx<T>::t;
    constexpr int n = A * B, m = A / B;
    __shared__ T2 lds[lanes_per_wg<T, A, B>() * A * (B + 1)];
    const int p = threadIdx.x / B, l = threadIdx.x % B;
    const long row = (long)blockIdx.x * lanes_per_wg<T, A, B>() + p;
    const bool ok = row < nrows;
    T2 v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
        T2 z;
        z.x = 0;
        z.y = 0;
        if (ok) z = data[row * n + l + B * q];
        v[q] = z;
    }
    four_step<A, B, SIGN>(v, lds + p * A * (B + 1), tw, l);
    if (!ok) return;
#pragma unroll
    for (int s = 0; s < m; ++s)
#pragma unroll
        for (int j = 0; j < B; ++j) data[row * n + l + B * s + A * j] = v[s * B + j];
}

// ---- strided pencils (y: axis 1, x: axis 0) of [Nx][Ny][Zc], rows Zp apart ----------
// A workgroup takes C neighbouring z columns of one line; thread t handles column
// c = t % C (consecutive lanes read consecutive columns) and pencil position t / C.
// Transpose buffer layout [q][l][c] keeps both phases free of bank conflicts.  With the
// row pitch Zp a multiple of 128 bytes, a workgroup's column group is whole cache lines
// of every row it reads (an odd pitch such as 129 splits each 256-byte run over 3 lines).
struct Lines {
    long base, stride;
    int z0, line;
};

template <int C>
__device__ __forceinline__ Lines lines(int axis, int Ny, int Zc, int Zp)
{
    Lines L;
    const int zb = (Zc + C - 1) / C;
    L.line = blockIdx.x / zb;
    L.z0 = (blockIdx.x - L.line * zb) * C;
    if (axis == 1) {
        L.base = (long)L.line * Ny * Zp + L.z0;
        L.stride = Zp;
    } else {
        L.base = (long)L.line * Zp + L.z0;
        L.stride = (long)Ny * Zp;
    }
    return L;
}

// Column forms of four_step / four_step_inv_from_X: thread (c, l), l = 0..B-1, holds
// x[l + B j1] in v[j1] (x layout) or X[(l + B s) + A k2] in v[s B + k2] (X layout).
template <int A, int B, int C, int SIGN, typename T2>
__device__ __forceinline__ void four_step_cols(T2 (&v)[A], T2* lds, const T2* __restrict__ tw, int c, int l)
{
    constexpr int m = A / B;
    dft<A, SIGN>(v);
#pragma unroll
    for (int q = 0; q < A; ++q) lds[(q * B + l) * C + c] = q == 0 ? v[q] : cmul(v[q], ptw<SIGN>(tw, l * q));
    __syncthreads();
#pragma unroll
    for (int s = 0; s < m; ++s) {
        T2 u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) u[j] = lds[((l + B * s) * B + j) * C + c];
        dft<B, SIGN>(u);
#pragma unroll
        for (int j = 0; j < B; ++j) v[s * B + j] = u[j];
    }
}

template <int A, int B, int C, typename T2>
__device__ __forceinline__ void four_step_cols_inv_from_X(T2 (&v)[A], T2* lds, const T2* __restrict__ tw, int c, int l)
{
    constexpr int m = A / B;
#pragma unroll
    for (int s = 0; s < m; ++s) {
        T2 u[B];
#pragma unroll
        for (int j = 0; j < B; ++j) u[j] = v[s * B + j];
        dft<B, +1>(u);
        const int k1 = l + B * s;
#pragma unroll
        for (int j = 0; j < B; ++j)
            lds[(k1 * B + j) * C + c] = (j == 0 || k1 == 0) ? u[j] : cmul(u[j], ptw<+1>(tw, j * k1));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < A; ++q) v[q] = lds[(q * B + l) * C + c];
    dft<A, +1>(v);
}

template <typename T, int A, int B, int SIGN>
__global__ __launch_bounds__((B * cols_per_wg<T, A, B>())) void k_fftr_strided(int axis, int Ny, int Zc, int Zp,
                                                                              typename Cx<T>::t* __restrict__ data,
                                                                              const typename Cx<T>::t* __restrict__ tw)
{
    using T2 = typename Cx<T>::t;
    constexpr int C = cols_per_wg<T, A, B>(), m = A / B;
    __shared__ T2 lds[C * A * B];
    const Lines L = lines<C>(axis, Ny, Zc, Zp);
    const int c = threadIdx.x % C, l = threadIdx.x / C;
    const bool ok = L.z0 + c < Zc;
    T2 v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
        T2 z;
        z.x = 0;
        z.y = 0;
        if (ok) z = data[L.base + c + (long)(l + B * q) * L.stride];
        v[q] = z;
    }
    four_step_cols<A, B, C, SIGN>(v, lds, tw, c, l);
    if (!ok) return;
#pragma unroll
    for (int s = 0; s < m; ++s)
#pragma unroll
        for (int j = 0; j < B; ++j) data[L.base + c + (long)(l + B * s + A * j) * L.stride] = v[s * B + j];
}

template <typename T>
__device__ __forceinline__ T sin_t(T x);
template <>
__device__ __forceinline__ float sin_t<float>(float x) { return sinf(x); }
template <>
__device__ __forceinline__ double sin_t<double>(double x) { return sin(x); }

// Forward x FFT, S(k) = j0(|k|)^2 exp(-sum_d (1-eta_d) k_d^2/4xi^2)/|k|^2/G (S(0) = 0;
// cppmodule/ConductorRigid.cc:253-268), inverse x FFT, on the half spectrum.  RG: the radial
// table holds all of S (sc.radial_gauss), one table load per mode.
template <typename T, int A, int B, bool RG = false>
__device__ __forceinline__ void x_scale_pass(int Ny, int Zc, int Zp, typename Cx<T>::t* __restrict__ data,
                                             const typename Cx<T>::t* __restrict__ tw, const ScaleTables& sc)
{
    using T2 = typename Cx<T>::t;
    constexpr int C = cols_per_wg<T, A, B>(), m = A / B;
    __shared__ T2 lds[C * A * B];
    const Lines L = lines<C>(0, Ny, Zc, Zp);
    const int c = threadIdx.x % C, l = threadIdx.x / C;
    const int z = L.z0 + c;
    const bool ok = z < Zc;
    // RG: the thread's 16 table values are loaded before the mesh values, so that their
    // latency hides behind the loads and the forward transform (x pass 45.8 -> 42.6 us at 256^3)
    T Sv[RG ? A : 1];
    if constexpr (RG) {
        const T* rad = static_cast<const T*>(sc.radial);
        const int my = L.line < (sc.n[1] + 1) / 2 ? L.line : L.line - sc.n[1];
        const int zi = ok ? z : 0;  // padding lanes index the table in range too
        const int mz = zi < (sc.n[2] + 1) / 2 ? zi : zi - sc.n[2];
        const int myz = my * my + mz * mz;
#pragma unroll
        for (int s = 0; s < m; ++s)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int x = l + B * s + A * j;
                const int mx = x < (sc.n[0] + 1) / 2 ? x : x - sc.n[0];
                Sv[s * B + j] = rad[myz + mx * mx];
            }
    }
    T2 v[A];
#pragma unroll
    for (int q = 0; q < A; ++q) {
        T2 u;
        u.x = 0;
        u.y = 0;
        if (ok) u = data[L.base + c + (long)(l + B * q) * L.stride];
        v[q] = u;
    }
    four_step_cols<A, B, C, -1>(v, lds, tw, c, l);
    // thread (c, l) holds X[(l + B s) + A k2] in v[s B + k2]: scale by S(kx, ky, kz)
    if constexpr (RG) {
#pragma unroll
        for (int k = 0; k < A; ++k) {
            const T S = ok ? Sv[k] : (T)0;
            v[k].x *= S;
            v[k].y *= S;
        }
        __syncthreads();  // the transpose buffer is reused
        four_step_cols_inv_from_X<A, B, C>(v, lds, tw, c, l);
        if (!ok) return;
#pragma unroll
        for (int q = 0; q < A; ++q) data[L.base + c + (long)(l + B * q) * L.stride] = v[q];
        return;
    }
    const double ky = sc.k[1][L.line], ey = sc.e[1][L.line];
    const double kz = ok ? sc.k[2][z] : 0.0, ez = ok ? sc.e[2][z] : 0.0;
    if (sc.radial && sc.radial_gauss) {
        const T* rad = static_cast<const T*>(sc.radial);
        const int my = L.line < (sc.n[1] + 1) / 2 ? L.line : L.line - sc.n[1];
        const int zi = ok ? z : 0;
        const int mz = zi < (sc.n[2] + 1) / 2 ? zi : zi - sc.n[2];
        const int myz = my * my + mz * mz;
#pragma unroll
        for (int s = 0; s < m; ++s)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int x = l + B * s + A * j;
                const int mx = x < (sc.n[0] + 1) / 2 ? x : x - sc.n[0];
                const T S = ok ? rad[myz + mx * mx] : (T)0;
                v[s * B + j].x *= S;
                v[s * B + j].y *= S;
            }
    } else if (sc.radial) {
        const T* rad = static_cast<const T*>(sc.radial);
        const int my = L.line < (sc.n[1] + 1) / 2 ? L.line : L.line - sc.n[1];
        const int zi = ok ? z : 0;  // padding lanes index the table in range too
        const int mz = zi < (sc.n[2] + 1) / 2 ? zi : zi - sc.n[2];
        const int myz = my * my + mz * mz;
        const double eyz = ey * ez;
#pragma unroll
        for (int s = 0; s < m; ++s)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int x = l + B * s + A * j;
                const int mx = x < (sc.n[0] + 1) / 2 ? x : x - sc.n[0];
                const T S = ok ? rad[myz + mx * mx] * (T)(sc.e[0][x] * eyz) : (T)0;
                v[s * B + j].x *= S;
                v[s * B + j].y *= S;
            }
    } else {
#pragma unroll
        for (int s = 0; s < m; ++s)
#pragma unroll
            for (int j = 0; j < B; ++j) {
                const int x = l + B * s + A * j;
                const T kx = (T)sc.k[0][x], kyt = (T)ky, kzt = (T)kz;
                const T k2 = kx * kx + kyt * kyt + kzt * kzt;
                T S = 0;
                if (k2 > (T)0) {
                    const T kk = sqrt(k2);
                    const T j0 = sin_t<T>(kk) / kk;
                    S = j0 * j0 * (T)(sc.e[0][x] * ey * ez) / k2 * (T)sc.invG;
                }
                v[s * B + j].x *= S;
                v[s * B + j].y *= S;
            }
    }
    __syncthreads();  // the transpose buffer is reused
    four_step_cols_inv_from_X<A, B, C>(v, lds, tw, c, l);
    if (!ok) return;
#pragma unroll
    for (int q = 0; q < A; ++q) data[L.base + c + (long)(l + B * q) * L.stride] = v[q];
}

template <typename T, int A, int B, bool RG>
__global__ __launch_bounds__((B * cols_per_wg<T, A, B>())) void k_fftr_x_scale(int Ny, int Zc, int Zp,
                                                                              typename Cx<T>::t* __restrict__ data,
                                                                              const typename Cx<T>::t* __restrict__ tw,
                                                                              ScaleTables sc)
{
    x_scale_pass<T, A, B, RG>(Ny, Zc, Zp, data, tw, sc);
}

// Boxes of one mesh stored back to back ([box][Nx][Ny][Zc]): blockIdx.y picks the box
// and its own scale tables.
template <typename T, int A, int B>
__global__ __launch_bounds__((B * cols_per_wg<T, A, B>())) void k_fftr_x_scale_batch(
    int Ny, int Zc, int Zp, long box_stride, typename Cx<T>::t* __restrict__ data,
    const typename Cx<T>::t* __restrict__ tw, const ScaleTables* __restrict__ sct)
{
    x_scale_pass<T, A, B>(Ny, Zc, Zp, data + blockIdx.y * box_stride, tw, sct[blockIdx.y]);
}

inline unsigned ublk(long n, int per) { return (unsigned)((n + per - 1) / per); }

template <typename T, int A, int B>
hipError_t z_r2c(long nrows, int zp, const T* real, T* cplx, const void* tw, hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int P = lanes_per_wg<T, A, B>();
    hipLaunchKernelGGL((k_fftr_z_r2c<T, A, B>), dim3(ublk(nrows, P)), dim3(B * P), 0, st, nrows, zp, real, (T2*)cplx,
                       (const T2*)tw);
    return hipGetLastError();
}

template <typename T, int A, int B>
hipError_t z_c2r(long nrows, int zp, const T* cplx, T* real, const void* tw, hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int P = lanes_per_wg<T, A, B>();
    hipLaunchKernelGGL((k_fftr_z_c2r<T, A, B>), dim3(ublk(nrows, P)), dim3(B * P), 0, st, nrows, zp, (const T2*)cplx,
                       real, (const T2*)tw);
    return hipGetLastError();
}

template <typename T, int A, int B>
hipError_t z_c2c(long nrows, T* data, int sign, const void* tw, hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int P = lanes_per_wg<T, A, B>();
    if (sign < 0)
        hipLaunchKernelGGL((k_fftr_z_c2c<T, A, B, -1>), dim3(ublk(nrows, P)), dim3(B * P), 0, st, nrows, (T2*)data,
                           (const T2*)tw);
    else
        hipLaunchKernelGGL((k_fftr_z_c2c<T, A, B, +1>), dim3(ublk(nrows, P)), dim3(B * P), 0, st, nrows, (T2*)data,
                           (const T2*)tw);
    return hipGetLastError();
}

template <typename T, int A, int B>
hipError_t strided(int axis, int Nx, int Ny, int Zc, int Zp, T* data, int sign, const void* tw, hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int C = cols_per_wg<T, A, B>();
    const int lines_n = axis == 1 ? Nx : Ny;
    const unsigned blocks = (unsigned)(lines_n * ((Zc + C - 1) / C));
    if (sign < 0)
        hipLaunchKernelGGL((k_fftr_strided<T, A, B, -1>), dim3(blocks), dim3(B * C), 0, st, axis, Ny, Zc, Zp,
                           (T2*)data, (const T2*)tw);
    else
        hipLaunchKernelGGL((k_fftr_strided<T, A, B, +1>), dim3(blocks), dim3(B * C), 0, st, axis, Ny, Zc, Zp,
                           (T2*)data, (const T2*)tw);
    return hipGetLastError();
}

template <typename T, int A, int B>
hipError_t x_scale(int Ny, int Zc, int Zp, T* data, const void* tw, const ScaleTables& sc, hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int C = cols_per_wg<T, A, B>();
    const unsigned blocks = (unsigned)(Ny * ((Zc + C - 1) / C));
    if (sc.radial && sc.radial_gauss)
        hipLaunchKernelGGL((k_fftr_x_scale<T, A, B, true>), dim3(blocks), dim3(B * C), 0, st, Ny, Zc, Zp, (T2*)data,
                           (const T2*)tw, sc);
    else
        hipLaunchKernelGGL((k_fftr_x_scale<T, A, B, false>), dim3(blocks), dim3(B * C), 0, st, Ny, Zc, Zp,
                           (T2*)data, (const T2*)tw, sc);
    return hipGetLastError();
}

template <typename T, int A, int B>
hipError_t x_scale_batch(int nbox, int Ny, int Zc, int Zp, T* data, const void* tw, const ScaleTables* sct,
                         hipStream_t st)
{
    using T2 = typename Cx<T>::t;
    constexpr int C = cols_per_wg<T, A, B>();
    const unsigned blocks = (unsigned)(Ny * ((Zc + C - 1) / C));
    const long stride = (long)(A * B) * Ny * Zp;
    hipLaunchKernelGGL((k_fftr_x_scale_batch<T, A, B>), dim3(blocks, (unsigned)nbox), dim3(B * C), 0, st, Ny, Zc,
                       Zp, stride, (T2*)data, (const T2*)tw, sct);
    return hipGetLastError();
}

}  // namespace

// Register path lengths n = A * B (A = B or A = 2B, both powers of two, A <= 32):
// 32 = 8 x 4, 64 = 8 x 8, 128 = 16 x 8, 256 = 16 x 16, 512 = 32 x 16, 1024 = 32 x 32.
int fftr_radix(int n)
{
    for (int v : {32, 64, 128, 256, 512, 1024})
        if (v == n) return n;
    return 0;
}

#define CR_FFTR_DISPATCH(n, CALL)                       \
    switch (fftr_radix(n)) {                            \
        case 32: return CALL(8, 4);                     \
        case 64: return CALL(8, 8);                     \
        case 128: return CALL(16, 8);                   \
        case 256: return CALL(16, 16);                  \
        case 512: return CALL(32, 16);                  \
        case 1024: return CALL(32, 32);                 \
        default: return hipErrorInvalidValue;           \
    }

template <typename T>
hipError_t launch_fftr_z_r2c(const AxisPlan& pz, long nrows, int zp, const T* real, T* cplx, hipStream_t st)
{
#define CALL(A, B) z_r2c<T, A, B>(nrows, zp, real, cplx, pz.tw, st)
    CR_FFTR_DISPATCH(pz.n, CALL)
#undef CALL
}

template <typename T>
hipError_t launch_fftr_z_c2r(const AxisPlan& pz, long nrows, int zp, const T* cplx, T* real, hipStream_t st)
{
#define CALL(A, B) z_c2r<T, A, B>(nrows, zp, cplx, real, pz.tw, st)
    CR_FFTR_DISPATCH(pz.n, CALL)
#undef CALL
}

template <typename T>
hipError_t launch_fftr_z_c2c(const AxisPlan& pz, long nrows, T* data, int sign, hipStream_t st)
{
#define CALL(A, B) z_c2c<T, A, B>(nrows, data, sign, pz.tw, st)
    CR_FFTR_DISPATCH(pz.n, CALL)
#undef CALL
}

template <typename T>
hipError_t launch_fftr_strided(const AxisPlan& pa, int axis, int Nx, int Ny, int Zc, int Zp, T* data, int sign,
                               hipStream_t st)
{
#define CALL(A, B) strided<T, A, B>(axis, Nx, Ny, Zc, Zp, data, sign, pa.tw, st)
    CR_FFTR_DISPATCH(pa.n, CALL)
#undef CALL
}

template <typename T>
hipError_t launch_fftr_x_scale(const AxisPlan& px, int Ny, int Zc, int Zp, T* data, const ScaleTables& sc,
                               hipStream_t st)
{
#define CALL(A, B) x_scale<T, A, B>(Ny, Zc, Zp, data, px.tw, sc, st)
    CR_FFTR_DISPATCH(px.n, CALL)
#undef CALL
}

template <typename T>
hipError_t launch_fftr_x_scale_batch(const AxisPlan& px, int nbox, int Ny, int Zc, int Zp, T* data,
                                     const ScaleTables* sct, hipStream_t st)
{
#define CALL(A, B) x_scale_batch<T, A, B>(nbox, Ny, Zc, Zp, data, px.tw, sct, st)
    CR_FFTR_DISPATCH(px.n, CALL)
#undef CALL
}

#define CR_FFTR_INST(T)                                                                                          \
    template hipError_t launch_fftr_z_r2c<T>(const AxisPlan&, long, int, const T*, T*, hipStream_t);           \
    template hipError_t launch_fftr_z_c2r<T>(const AxisPlan&, long, int, const T*, T*, hipStream_t);           \
    template hipError_t launch_fftr_z_c2c<T>(const AxisPlan&, long, T*, int, hipStream_t);                     \
    template hipError_t launch_fftr_strided<T>(const AxisPlan&, int, int, int, int, int, T*, int, hipStream_t); \
    template hipError_t launch_fftr_x_scale<T>(const AxisPlan&, int, int, int, T*, const ScaleTables&,          \
                                               hipStream_t);                                                   \
    template hipError_t launch_fftr_x_scale_batch<T>(const AxisPlan&, int, int, int, int, T*, const ScaleTables*, \
                                                     hipStream_t);
CR_FFTR_INST(float)
CR_FFTR_INST(double)

}  // namespace cr
