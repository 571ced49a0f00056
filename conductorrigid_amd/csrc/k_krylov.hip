// k_krylov.hip -- FP64 restarted GMRES building blocks, device resident.
//
// The reference runs CUSP's krylov::gmres (cppmodule/ConductorRigid.cu:1134-1136):
// modified Gram-Schmidt Arnoldi on the device with one device->host sync per dot
// product, Hessenberg matrix and Givens rotations on the host in float.  Here the
// basis, Hessenberg matrix, rotations and least-squares right-hand side stay on the
// device in double; orthogonalisation is classical Gram-Schmidt with one
// re-orthogonalisation pass (CGS2), so each pass over the basis is one kernel that
// produces all (j+1) projections at once, reduced deterministically (fixed block
// order, no floating-point atomics).
#include "cr_kernels.h"

namespace cr {

namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 8;       // basis columns per multidot launch row
constexpr int kMaxBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-reduce kChunk values held per thread; thread 0 writes them.
__device__ __forceinline__ void block_reduce_chunk(double (&acc)[kChunk], double* out)
{
    __shared__ double red[kBlock / 64][kChunk];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kChunk) {
        double s = 0;
        for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// partial[(c0 + k) * kMaxBlocks + b] = sum_{i in block b} V[c0+k][i] w[i].  The partials are
// column-major: a reduction over the blocks of one column reads a contiguous run.
__global__ __launch_bounds__(kBlock) void k_multidot(const double* __restrict__ V, long ld, int ncols,
                                                     const double* __restrict__ w, long n,
                                                     double* __restrict__ partial)
{
    const int c0 = blockIdx.y * kChunk;
    double acc[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) acc[k] = 0.0;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        // all loads first (columns past ncols re-read the last one, cached and unused)
        double v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) v[k] = V[(long)min(c0 + k, ncols - 1) * ld + i];
        const double wi = w[i];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (c0 + k < ncols) acc[k] += v[k] * wi;
    }
    __shared__ double res[kChunk];
    block_reduce_chunk(acc, res);
    if (threadIdx.x < kChunk && c0 + (int)threadIdx.x < ncols)
        partial[(long)(c0 + threadIdx.x) * kMaxBlocks + blockIdx.x] = res[threadIdx.x];
}

// out[k] = sum_b partial[k * kMaxBlocks + b]: one workgroup per column; each thread sums a
// strided subset of the blocks (independent loads in flight), then a fixed-order
// tree over the workgroup, so the result is deterministic.
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double* __restrict__ partial, int nblocks,
                                                            double* __restrict__ out)
{
    const int k = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[(long)k * kMaxBlocks + b];
    __shared__ double red[kBlock / 64];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w];
        out[k] = t;
    }
}

// Relative size below which CGS2's second correction is not applied: ||h2|| <= this *
// ||w|| changes w by less than that, far below what the GMRES tolerances resolve.
constexpr double kReorthSkip = 1e-13;

// w -= V[:, 0:ncols] h ; mode 2 also accumulates ||w||^2 partials into partial[b] (column 0).
// Mode 3 is mode 2 for CGS2's second correction with h[ncols] = ||w||^2 before it: when
// ||h[0:ncols]|| <= kReorthSkip ||w|| the update is skipped and the partials carry ||w||^2.
__global__ __launch_bounds__(kBlock) void k_gs_update(const double* __restrict__ V, long ld, int ncols,
                                                      const double* __restrict__ h, double* __restrict__ w, long n,
                                                      int mode, double* __restrict__ partial)
{
    __shared__ double hs[32];
    if (threadIdx.x < 32) hs[threadIdx.x] = (threadIdx.x < (unsigned)ncols) ? h[threadIdx.x] : 0.0;
    __syncthreads();
    if (mode == 3) {
        double h2 = 0.0;
        for (int k = 0; k < ncols; ++k) h2 += hs[k] * hs[k];
        const double wn2 = h[ncols];
        if (h2 <= kReorthSkip * kReorthSkip * wn2) {
            if (threadIdx.x == 0) partial[blockIdx.x] = blockIdx.x == 0 ? wn2 : 0.0;
            return;
        }
    }
    double acc[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) acc[k] = 0.0;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        double wi = w[i];
        for (int k = 0; k < ncols; ++k) wi -= hs[k] * V[(long)k * ld + i];
        w[i] = wi;
        acc[0] += wi * wi;
    }
    if (mode >= 2) {
        __shared__ double res[kChunk];
        block_reduce_chunk(acc, res);
        if (threadIdx.x == 0) partial[blockIdx.x] = res[0];
    }
}

// w -= V[:, 0:ncols] h, fused with the next projection: partial[k * kMaxBlocks + b] = sum over the
// block's elements of V[k][i] * w_new[i] (same layout as k_multidot), so CGS2's first
// update and second projection read the basis once.  with_norm (ncols < 32): column
// ncols also gets the block's sum of w_new[i]^2.
// NC >= ncols + with_norm bounds the columns held per thread, so that a short basis keeps
// few registers live; every NC gives the same sums in the same order.
template <int NC>
__global__ __launch_bounds__(kBlock) void k_gs_update_dot(const double* __restrict__ V, long ld, int ncols,
                                                          const double* __restrict__ h, double* __restrict__ w,
                                                          long n, double* __restrict__ partial, int with_norm)
{
    __shared__ double hs[32];
    if (threadIdx.x < 32) hs[threadIdx.x] = (threadIdx.x < (unsigned)ncols) ? h[threadIdx.x] : 0.0;
    __syncthreads();
    double acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0.0;
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
        // all loads first (columns past ncols re-read the last one, which is cached and
        // never used), so that they are in flight together
        double v[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) v[k] = V[(long)min(k, ncols - 1) * ld + i];
        double wi = w[i];
#pragma unroll
        for (int k = 0; k < NC; ++k)
            if (k < ncols) wi -= hs[k] * v[k];
        w[i] = wi;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            if (k < ncols) acc[k] += v[k] * wi;
            else if (k == ncols && with_norm) acc[k] += wi * wi;
        }
    }
    const int nout = ncols + (with_norm ? 1 : 0);
#pragma unroll
    for (int c0 = 0; c0 < NC; c0 += kChunk) {
        if (c0 >= nout) break;
        double part[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) part[k] = acc[c0 + k];
        __shared__ double res[kChunk];
        block_reduce_chunk(part, res);
        if (threadIdx.x < kChunk && c0 + (int)threadIdx.x < nout)
            partial[(long)(c0 + threadIdx.x) * kMaxBlocks + blockIdx.x] = res[threadIdx.x];
        __syncthreads();
    }
}


__global__ void k_axpby(double a, const double* __restrict__ x, double b, double* __restrict__ y, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

__global__ void k_scale_copy(const double* __restrict__ x, const double* __restrict__ s, double sign,
                             double* __restrict__ y, long n)
{
    const double f = sign * (*s);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = f * x[i];
}

__global__ void k_combine(const double* __restrict__ V, long ld, const double* __restrict__ c,
                          const int* __restrict__ kcount, double* __restrict__ x, long n)
{
    const int K = *kcount;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double xi = x[i];
        for (int k = 0; k < K; ++k) xi += c[k] * V[(long)k * ld + i];
        x[i] = xi;
    }
}

// Givens plane rotation helpers (Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 6.9; the same rotation CUSP's GMRES applies).
__device__ __forceinline__ void apply_rot(double& dx, double& dy, double cs, double sn)
{
    const double t = cs * dx + sn * dy;
    dy = -sn * dx + cs * dy;
    dx = t;
}

__device__ __forceinline__ void gen_rot(double dx, double dy, double& cs, double& sn)
{
    if (dy == 0.0) {
        cs = 1.0;
        sn = 0.0;
    } else if (fabs(dy) > fabs(dx)) {
        const double t = dx / dy;
        sn = 1.0 / sqrt(1.0 + t * t);
        cs = t * sn;
    } else {
        const double t = dy / dx;
        cs = 1.0 / sqrt(1.0 + t * t);
        sn = t * cs;
    }
}

// Column i of H from the two CGS projections and ||w||, then the rotations.
__device__ __forceinline__ void givens_column(GmresState* g, int i, int R)
{
    const int ld = R + 1;
    double* H = g->H;
    for (int k = 0; k <= i; ++k) H[i * ld + k] = g->h1[k] + g->h2[k];
    const double hn = sqrt(g->wnorm2);
    H[i * ld + i + 1] = hn;
    g->inv_norm = (hn != 0.0) ? 1.0 / hn : 0.0;
    for (int k = 0; k < i; ++k) apply_rot(H[i * ld + k], H[i * ld + k + 1], g->cs[k], g->sn[k]);
    gen_rot(H[i * ld + i], H[i * ld + i + 1], g->cs[i], g->sn[i]);
    apply_rot(H[i * ld + i], H[i * ld + i + 1], g->cs[i], g->sn[i]);
    apply_rot(g->s[i], g->s[i + 1], g->cs[i], g->sn[i]);
    g->resid = fabs(g->s[i + 1]);
    g->i = i;
    g->ncols = i + 1;
}

__global__ void k_givens(GmresState* g, int i, int R)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    givens_column(g, i, R);
}

__device__ __forceinline__ void publish_value(double v, HostPub* dst, unsigned long long seq)
{
    *reinterpret_cast<volatile double*>(&dst->value) = v;
    __threadfence_system();
    *reinterpret_cast<volatile unsigned long long*>(&dst->seq) = seq;
}

// NaN-propagating maximum (a NaN charge must reach the FP32-bulk mesh as NaN)
__device__ __forceinline__ float nan_max(float m, float a) { return (a > m || a != a) ? a : m; }

// The end of an Arnoldi step in one launch: ||w||^2 from the partials of the last CGS pass
// (every workgroup sums them in k_reduce_partials' order, so all hold the same value),
// V[:, i+1] = w / ||w|| (as k_scale_copy), and in workgroup 0 column i of H with its
// Givens rotations (k_givens), then the residual published to the host (pub != nullptr,
// as k_publish).  qf != nullptr: also the next operator application's float charges,
// qf = (float) V[0:nq, i+1], and each workgroup's NaN-propagating max |qf| in
// fs->part[blockIdx.x] (the float conversion and maximum of k_charges_fx, so the same
// charges and scale).
__global__ __launch_bounds__(kBlock) void k_step_tail(const double* __restrict__ partial, int nblocks, GmresState* g,
                                                      int i, int R, const double* __restrict__ w,
                                                      double* __restrict__ vnext, long n, HostPub* pub,
                                                      unsigned long long seq, float* __restrict__ qf,
                                                      FxScale* __restrict__ fs, long nq,
                                                      const double* __restrict__ V, long ld, int ncols,
                                                      const double* __restrict__ h2)
{
    __shared__ double red[kBlock / 64];
    __shared__ double tot;
    __shared__ double hs[32];
    double wn2;
    bool apply = false;
    if (h2) {
        // CGS2's second correction folded in: h2[0:ncols] = V^T w, h2[ncols] = ||w||^2; the
        // corrected vector w - V h2 has ||w||^2 - ||h2||^2 as its squared norm (V orthonormal),
        // and the correction is skipped as in k_gs_update's mode 3
        if (threadIdx.x < 32) hs[threadIdx.x] = (int)threadIdx.x < ncols ? h2[threadIdx.x] : 0.0;
        __syncthreads();
        double hsq = 0.0;
        for (int k = 0; k < ncols; ++k) hsq += hs[k] * hs[k];
        const double w2 = h2[ncols];
        apply = hsq > kReorthSkip * kReorthSkip * w2;
        wn2 = apply ? fmax(w2 - hsq, 0.0) : w2;
    } else {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partial[b];
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < kBlock / 64; ++k) t += red[k];
            tot = t;
        }
        __syncthreads();
        wn2 = tot;
    }
    const double hn = sqrt(wn2);
    const double f = (hn != 0.0) ? 1.0 / hn : 0.0;
    // element k of the (corrected) w, subtracting in k_gs_update's order
    auto wk = [&](long k) {
        double v = w[k];
        if (apply) {
            int j = 0;
            for (; j + 4 <= ncols; j += 4) {  // four columns' loads in flight together
                const double a0 = V[(long)j * ld + k], a1 = V[(long)(j + 1) * ld + k];
                const double a2 = V[(long)(j + 2) * ld + k], a3 = V[(long)(j + 3) * ld + k];
                v -= hs[j] * a0;
                v -= hs[j + 1] * a1;
                v -= hs[j + 2] * a2;
                v -= hs[j + 3] * a3;
            }
            for (; j < ncols; ++j) v -= hs[j] * V[(long)j * ld + k];
        }
        return v;
    };
    if (!qf) {
        for (long k = (long)blockIdx.x * kBlock + threadIdx.x; k < n; k += (long)gridDim.x * kBlock)
            vnext[k] = f * wk(k);
    } else {
        float m = 0.0f;
        for (long k = (long)blockIdx.x * kBlock + threadIdx.x; k < n; k += (long)gridDim.x * kBlock) {
            const double v = f * wk(k);
            vnext[k] = v;
            if (k < nq) {
                const float q = (float)v;
                qf[k] = q;
                m = nan_max(m, fabsf(q));
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off, 64));
        __shared__ float wm[kBlock / 64];
        if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            float bm = wm[0];
            for (int k = 1; k < kBlock / 64; ++k) bm = nan_max(bm, wm[k]);
            fs->part[blockIdx.x] = bm;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        g->wnorm2 = wn2;
        givens_column(g, i, R);
        if (pub) publish_value(g->resid, pub, seq);
    }
}

__global__ void k_publish(const double* __restrict__ src, HostPub* dst, unsigned long long seq)
{
    if (threadIdx.x != 0) return;
    publish_value(*src, dst, seq);
}

__global__ void k_cycle_init(GmresState* g, const double* norm2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double beta = sqrt(*norm2);
    g->beta = beta;
    g->inv_beta = (beta != 0.0) ? 1.0 / beta : 0.0;
    for (int k = 0; k < 33; ++k) g->s[k] = 0.0;
    g->s[0] = beta;
    g->resid = beta;
    g->ncols = 0;
}

// Back substitution H(0:i,0:i) y = s(0:i) (R upper triangular after the rotations).  A zero
// pivot (a step taken after an exact breakdown: its column and right-hand side are zero)
// contributes y = 0.
__global__ void k_backsub(GmresState* g, int R)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int ld = R + 1;
    const int m = g->ncols;
    for (int k = 0; k < m; ++k) g->y[k] = g->s[k];
    for (int j = m - 1; j >= 0; --j) {
        const double d = g->H[j * ld + j];
        g->y[j] = d != 0.0 ? g->y[j] / d : 0.0;
        for (int k = j - 1; k >= 0; --k) g->y[k] -= g->H[j * ld + k] * g->y[j];
    }
}

inline unsigned vblocks(long n)
{
    long b = (n + kBlock - 1) / kBlock;
    if (b > kMaxBlocks) b = kMaxBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace

int krylov_blocks(long n) { return (int)vblocks(n); }

hipError_t launch_multidot(const double* V, long ld, int ncols, const double* w, long n, double* partial,
                           hipStream_t st)
{
    if (ncols <= 0 || ncols > 32) return hipErrorInvalidValue;  // the kernel re-reads column ncols - 1
    const unsigned nb = vblocks(n);
    const unsigned nchunk = (unsigned)((ncols + kChunk - 1) / kChunk);
    hipLaunchKernelGGL(k_multidot, dim3(nb, nchunk), dim3(kBlock), 0, st, V, ld, ncols, w, n, partial);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partial, int nblocks, int ncols, double* out, hipStream_t st)
{
    if (ncols <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_partials, dim3(ncols), dim3(kBlock), 0, st, partial, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_step_tail(const double* partial, int nblocks, GmresState* g, int i, int R, const double* w,
                            double* vnext, long n, HostPub* pub, unsigned long long seq, hipStream_t st, float* qf,
                            FxScale* fs, long nq, int* nparts, const double* V, long ld, int ncols,
                            const double* h2)
{
    const unsigned nb = vblocks(n);
    if (qf && (!fs || nb > (unsigned)kFxParts || nq > n)) return hipErrorInvalidValue;
    if (h2 && (ncols < 1 || ncols >= 32 || !V)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_step_tail, dim3(nb), dim3(kBlock), 0, st, partial, nblocks, g, i, R, w, vnext, n, pub, seq,
                       qf, fs, nq, V, ld, ncols, h2);
    if (nparts) *nparts = (int)nb;
    return hipGetLastError();
}

hipError_t launch_publish(const double* src, HostPub* dst, unsigned long long seq, hipStream_t st)
{
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, st, src, dst, seq);
    return hipGetLastError();
}

hipError_t launch_gs_update(const double* V, long ld, int ncols, const double* h, double* w, long n, int mode,
                            double* partial, hipStream_t st)
{
    hipLaunchKernelGGL(k_gs_update, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, ncols, h, w, n, mode, partial);
    return hipGetLastError();
}

hipError_t launch_gs_update_dot(const double* V, long ld, int ncols, const double* h, double* w, long n,
                                double* partial, hipStream_t st, bool with_norm)
{
    const int nout = ncols + (with_norm ? 1 : 0);
    if (ncols <= 0 || ncols > 32 || nout > 32) return hipErrorInvalidValue;  // the kernel re-reads column ncols - 1
    const int wn = with_norm ? 1 : 0;
    if (nout <= 8)
        hipLaunchKernelGGL(k_gs_update_dot<8>, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, ncols, h, w, n, partial, wn);
    else if (nout <= 16)
        hipLaunchKernelGGL(k_gs_update_dot<16>, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, ncols, h, w, n, partial,
                           wn);
    else if (nout <= 24)
        hipLaunchKernelGGL(k_gs_update_dot<24>, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, ncols, h, w, n, partial,
                           wn);
    else
        hipLaunchKernelGGL(k_gs_update_dot<32>, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, ncols, h, w, n, partial,
                           wn);
    return hipGetLastError();
}

hipError_t launch_axpby(double a, const double* x, double b, double* y, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpby, dim3(vblocks(n)), dim3(kBlock), 0, st, a, x, b, y, n);
    return hipGetLastError();
}

hipError_t launch_scale_copy(const double* x, const double* s_ptr, double sign, double* y, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_scale_copy, dim3(vblocks(n)), dim3(kBlock), 0, st, x, s_ptr, sign, y, n);
    return hipGetLastError();
}

hipError_t launch_combine(const double* V, long ld, const double* c, const int* kcount, double* x, long n,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_combine, dim3(vblocks(n)), dim3(kBlock), 0, st, V, ld, c, kcount, x, n);
    return hipGetLastError();
}

hipError_t launch_givens(GmresState* g, int i, int R, hipStream_t st)
{
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, st, g, i, R);
    return hipGetLastError();
}

hipError_t launch_cycle_init(GmresState* g, const double* norm2, hipStream_t st)
{
    hipLaunchKernelGGL(k_cycle_init, dim3(1), dim3(64), 0, st, g, norm2);
    return hipGetLastError();
}

hipError_t launch_backsub(GmresState* g, int R, hipStream_t st)
{
    hipLaunchKernelGGL(k_backsub, dim3(1), dim3(64), 0, st, g, R);
    return hipGetLastError();
}

}  // namespace cr
