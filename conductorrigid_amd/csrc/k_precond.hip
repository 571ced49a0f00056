// k_precond.hip -- near-field block preconditioner for the conductor saddle-point system.
//
// The reference's "PreconditionWrapper" (cppmodule/PreconditionWrapper.cuh:23-75 ->
// precondition, cppmodule/ConductorRigid.cu:816-859) is a diagonal scaling.  At large N
// the bead-bead coupling inside each body (beads touch) dominates the conditioning, and
// diagonally preconditioned GMRES(30) stagnates.  This preconditioner inverts, per body,
// the body's own block of the saddle-point operator
//
//     A_b = [ -G_bb   1 ]       with G_ij the free-space interaction of two unit
//           [  1^T    0 ]       spherical shells: 1/4pi (i = j), (4 - r)/16pi (r < 2),
//                                1/4pi r (r >= 2)
//
// in closed form: u = G^{-1} 1, s = 1^T u,
//     A_b^{-1} = [ -G^{-1} + u u^T / s   u / s ]
//                [  u^T / s              1 / s ].
// G is symmetric positive definite (a Gram matrix of the Coulomb operator), so it is
// inverted by Gauss-Jordan without pivoting, in FP64, one workgroup per body; the
// inverse is stored in FP32 (it only steers convergence, never the converged answer).
#include "cr_kernels.h"

namespace cr {

namespace {

// Entry (i, j) of body t's inverse block in the packed lower triangle (block_inv_stride).
__device__ __forceinline__ int tri_at(int i, int j)
{
    const int a = i > j ? i : j, b = i > j ? j : i;
    return a * (a + 1) / 2 + b;
}
__device__ __forceinline__ long ginv_at(long t, int nmax, int i, int j)
{
    return t * block_inv_stride(nmax) + tri_at(i, j);
}

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double min_image(double d, double L) { return d - L * rint(d / L); }

__device__ __forceinline__ double shell_shell(double r)
{
    if (r == 0.0) return 1.0 / (4.0 * kPi);
    if (r < 2.0) return (4.0 - r) / (16.0 * kPi);
    return 1.0 / (4.0 * kPi * r);
}

// Bead counts per body without atomics (atomics at scattered addresses run at ~27 G/s on
// the memory side: ~40 us at 1M beads): a rigid body holds rigid_size beads, a bead
// outside every rigid body is a body of its own (the layout the fill kernels below use).
__global__ void k_body_count_rigid(SysView s, unsigned* __restrict__ count)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < s.N_rigid && s.rigid_size[r] > 0) count[s.iso[s.rigid_members[r * s.Nmax]]] = s.rigid_size[r];
}

__global__ void k_body_count_single(SysView s, unsigned* __restrict__ count)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.N && s.rigid_membership[i] == NOT_MEMBER) count[s.iso[i]] = 1u;
}

// Members of each body in a fixed order, without atomics: a rigid body's beads in its
// rigid_members order (the rigid body is one isolated body, tag iso[first member]), a
// bead outside every rigid body alone in its own body.
__global__ void k_body_fill_rigid(SysView s, const unsigned* __restrict__ start, unsigned* __restrict__ beads)
{
    const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long r = slot / s.Nmax;
    const int k = (int)(slot - r * s.Nmax);
    if (r >= s.N_rigid || k >= (int)s.rigid_size[r]) return;
    const unsigned* mem = s.rigid_members + r * s.Nmax;
    beads[start[s.iso[mem[0]]] + k] = mem[k];
}

__global__ void k_body_fill_single(SysView s, const unsigned* __restrict__ start, unsigned* __restrict__ beads)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.N && s.rigid_membership[i] == NOT_MEMBER) beads[start[s.iso[i]]] = (unsigned)i;
}

// Sum over the 64 lanes (xor butterfly: every lane ends with the same bits).
__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double lane_bcast(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Bodies of at most NB <= 64 beads: one wavefront per body, lane i holds row i of G in
// registers and the inverse is formed in place by Gauss-Jordan (no pivoting: G is
// symmetric positive definite), the pivot row broadcast element by element with
// readlane -- no LDS, no barriers.  Rows and columns n..NB-1 are padded with the
// identity, which the elimination leaves untouched.  Four bodies per 256-thread
// workgroup.  Body t's inverse, u and 1/s go to entry o of ginv, uvec and inv_s.
//
// cache (single system): G depends on the body's shape only, which a rigid body keeps.  Each
// entry o remembers the body, its bead count, a weighted sum of its bead indices and its
// shape signature (the sum of its beads' squared distances from their centroid, invariant
// under rigid motion); when all match the stored inverse is kept.  The preconditioner only
// steers GMRES, never the converged answer.
struct BlockCache {
    unsigned long long* key;  // t | n << 32 per entry
    double* sig;              // (shape signature, bead-index sum) per entry
    int ok;                   // the entries describe the inverses now stored
};

template <int NB>
__device__ __forceinline__ void block_setup_body(int t, int o, int nmax, const unsigned* __restrict__ start,
                                                 const unsigned* __restrict__ beads,
                                                 const double4* __restrict__ pos, double Lx, double Ly, double Lz,
                                                 float* __restrict__ ginv, float* __restrict__ uvec,
                                                 double* __restrict__ inv_s, int nlo = 0, bool last = true,
                                                 const BlockCache* cache = nullptr)
{
    const int i = threadIdx.x & 63;
    const int b0 = start[t];
    const int n = start[t + 1] - b0;
    if (n <= nlo) return;            // a launch for smaller bodies takes this one
    if (n > NB && !last) return;     // ... and one for larger bodies this one
    if (n > NB) {
        if (i == 0) inv_s[o] = __builtin_nan("");
        return;
    }
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    const unsigned bead = i < n ? beads[b0 + i] : 0u;
    if (i < n) pi = pos[bead];
    double sig = 0.0, hsum = 0.0;
    const unsigned long long key = (unsigned)t | ((unsigned long long)(unsigned)n << 32);
    if (cache) {
        // offsets from bead 0 (minimum image), their centroid, the spread about it
        const double x0 = lane_bcast(pi.x, 0), y0 = lane_bcast(pi.y, 0), z0 = lane_bcast(pi.z, 0);
        const double dx = i < n ? min_image(pi.x - x0, Lx) : 0.0;
        const double dy = i < n ? min_image(pi.y - y0, Ly) : 0.0;
        const double dz = i < n ? min_image(pi.z - z0, Lz) : 0.0;
        const double cx = wave_sum64(dx) / n, cy = wave_sum64(dy) / n, cz = wave_sum64(dz) / n;
        const double ex = dx - cx, ey = dy - cy, ez = dz - cz;
        sig = wave_sum64(i < n ? ex * ex + ey * ey + ez * ez : 0.0);
        hsum = wave_sum64(i < n ? (double)bead * (double)(i + 1) : 0.0);
        if (cache->ok && cache->key[o] == key && cache->sig[2 * o + 1] == hsum &&
            fabs(sig - cache->sig[2 * o]) <= 1e-9 * sig + 1e-300)
            return;  // the same body, beads and shape: its stored inverse stands
    }
    double a[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const double xj = lane_bcast(pi.x, j), yj = lane_bcast(pi.y, j), zj = lane_bcast(pi.z, j);
        double g;
        if (i < n && j < n) {
            if (i == j) {
                g = shell_shell(0.0);
            } else {
                const double dx = min_image(pi.x - xj, Lx), dy = min_image(pi.y - yj, Ly),
                             dz = min_image(pi.z - zj, Lz);
                g = shell_shell(sqrt(dx * dx + dy * dy + dz * dz));
            }
        } else {
            g = (i == j) ? 1.0 : 0.0;
        }
        a[j] = g;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const double inv = 1.0 / lane_bcast(a[k], k);
        const double f = a[k];
        const bool pivot = i == k;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            // lane k's a[j] before this step (each j is read before it is overwritten)
            const double pkj = lane_bcast(a[j], k);
            if (j == k) a[j] = pivot ? inv : -f * inv;
            else a[j] = pivot ? pkj * inv : a[j] - f * inv * pkj;
        }
    }
    double u = 0.0;
    if (i < n) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < n) {
                if (j <= i) ginv[ginv_at(o, nmax, i, j)] = (float)a[j];  // the lower triangle
                u += a[j];
            }
        uvec[(long)o * nmax + i] = (float)u;
    }
    double ssum = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) ssum += (j < n) ? lane_bcast(u, j) : 0.0;
    if (i == 0) {
        inv_s[o] = 1.0 / ssum;
        if (cache) {
            cache->key[o] = key;
            cache->sig[2 * o] = sig;
            cache->sig[2 * o + 1] = hsum;
        }
    }
}

// Slot s takes body order[s] (the block apply's processing order) and stores it at s.
template <int NB>
__global__ __launch_bounds__(256) void k_block_setup_wave(int B, int nmax, const unsigned* __restrict__ order,
                                                          const unsigned* __restrict__ start,
                                                          const unsigned* __restrict__ beads,
                                                          const double4* __restrict__ pos, double Lx, double Ly,
                                                          double Lz, float* __restrict__ ginv,
                                                          float* __restrict__ uvec, double* __restrict__ inv_s,
                                                          BlockCache cache)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= B) return;  // whole wavefront
    block_setup_body<NB>((int)order[s], s, nmax, start, beads, pos, Lx, Ly, Lz, ginv, uvec, inv_s, 0, true,
                         cache.key ? &cache : nullptr);
}

// All bodies of a batch of boxes in one launch (body0 ascending over boxes).
// Bodies with nlo < n <= NB; the last launch (NB >= every body) also marks larger ones.
template <int NB>
__global__ __launch_bounds__(256) void k_block_setup_wave_batch(const BlockBox* __restrict__ boxes, int nbox,
                                                                int total, int nlo, int last)
{
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= total) return;  // whole wavefront
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].body0 <= g) lo = mid;
        else hi = mid - 1;
    }
    const BlockBox& d = boxes[lo];
    block_setup_body<NB>(g - d.body0, g - d.body0, d.nmax, d.start, d.beads, d.pos, d.L[0], d.L[1], d.L[2],
                         const_cast<float*>(d.ginv), const_cast<float*>(d.uvec), const_cast<double*>(d.inv_s), nlo,
                         last != 0);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// w = A_b^{-1} y for body t: bead rows and the body's constraint row (64 threads).
__device__ __forceinline__ void block_apply_body(int t, int N, int nmax, const unsigned* __restrict__ start,
                                                 const unsigned* __restrict__ beads, const float* __restrict__ ginv,
                                                 const float* __restrict__ uvec, const double* __restrict__ inv_s,
                                                 const double* __restrict__ y, double* __restrict__ w)
{
    __shared__ double ys[64];
    const int b0 = start[t];
    const int n = start[t + 1] - b0;
    const int i = threadIdx.x;
    if (n > nmax || n > 64) {
        if (i == 0) w[N + t] = __builtin_nan("");
        return;
    }
    const unsigned bead = (i < n) ? beads[b0 + i] : 0u;
    const double yi = (i < n) ? y[bead] : 0.0;
    ys[i] = yi;
    __syncthreads();
    double gy = 0.0, ui = 0.0;
    if (i < n) {
        for (int j = 0; j < n; ++j) gy += (double)ginv[ginv_at(t, nmax, i, j)] * ys[j];
        ui = (double)uvec[(long)t * nmax + i];
    }
    const double dot = wave_sum(ui * yi);
    const double c = (dot + y[N + t]) * inv_s[t];
    if (i < n) w[bead] = -gy + ui * c;
    if (i == 0) w[N + t] = c;
}

// Processing order of the bodies for k_block_apply: ascending smallest bead index, by a
// counting sort over buckets of 2^kOrderShift bead indices (the bodies of one bucket in any
// order; a body without beads goes to the last bucket).  Each body's result is independent
// of the others, so the order changes no bits.  Bead indices in spatial (cell) order make
// the bodies of one workgroup share the cache lines of y and w that it gathers and scatters.
constexpr int kOrderShift = 5;

__global__ void k_body_order_key(int B, int N, const unsigned* __restrict__ start, const unsigned* __restrict__ beads,
                                 unsigned* __restrict__ bucket_count, uint2* __restrict__ key,
                                 unsigned* __restrict__ mkey)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const unsigned lo = start[b], hi = start[b + 1];
    unsigned m = (unsigned)N + (unsigned)b;  // a body without beads: past every bead, by index
    for (unsigned k = lo; k < hi; ++k) m = min(m, beads[k]);
    const unsigned bk = min(m, (unsigned)N) >> kOrderShift;
    key[b] = make_uint2(bk, atomicAdd(bucket_count + bk, 1u));
    mkey[b] = m;
}

// Body b to a place in its bucket (in the atomics' order), then each bucket sorted by smallest
// bead index (unique per body), so that the slots do not depend on the order of the atomics.
__global__ void k_body_order_place(int B, const uint2* __restrict__ key, const unsigned* __restrict__ bucket_start,
                                   unsigned* __restrict__ order)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint2 kv = key[b];
    order[bucket_start[kv.x] + kv.y] = (unsigned)b;
}

__global__ void k_body_order_sort(long nbk, const unsigned* __restrict__ bucket_start, const unsigned* __restrict__ mkey,
                                  unsigned* __restrict__ order)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nbk) return;
    const unsigned lo = bucket_start[k], hi = bucket_start[k + 1];
    for (unsigned a = lo + 1; a < hi; ++a) {  // insertion sort: buckets hold a few bodies
        const unsigned b = order[a], m = mkey[b];
        unsigned c = a;
        while (c > lo && mkey[order[c - 1]] > m) {
            order[c] = order[c - 1];
            --c;
        }
        order[c] = b;
    }
}

// Slot s: count[s] = its body's bead count, members[s * nmax + k] = the body's k-th bead
// (kNoBead past the count; a body of more than nmax beads keeps nmax).
constexpr unsigned kNoBead = 0xffffffffu;

__global__ void k_body_slot_members(int B, int nmax, const unsigned* __restrict__ order,
                                    const unsigned* __restrict__ start, const unsigned* __restrict__ beads,
                                    unsigned* __restrict__ count, unsigned* __restrict__ members)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)B * nmax) return;
    const long s = q / nmax;
    const int k = (int)(q - s * nmax);
    const unsigned b = order[s], lo = start[b], n = start[b + 1] - lo;
    if (k == 0) count[s] = n;
    members[q] = (unsigned)k < n ? beads[lo + k] : kNoBead;
}

// Workgroup b's place in an XCD-local order: consecutive workgroups go to different XCDs
// (round robin over 8), so XCD x takes the contiguous range of places that follows those
// of XCDs 0 .. x-1.
__device__ __forceinline__ int xcd_local_place(int b, int nwg)
{
    const int x = b & 7, q = nwg >> 3, r = nwg & 7;
    return x * q + min(x, r) + (b >> 3);
}

// Single system: K = 64 / nmax bodies share a wavefront (lane = body slot * nmax + bead),
// four wavefronts per workgroup, so small bodies do not leave most lanes idle.  The bodies
// are taken in slot order (launch_block_order) by XCD-local workgroup places; the inverse
// blocks, u, 1/s, the bead counts and member lists are stored by slot, so that a
// workgroup's bodies are contiguous in each and a lane's bead index is one load away.
template <int R4>
__global__ __launch_bounds__(256) void k_block_apply(int B, int N, int nmax, int K, const unsigned* __restrict__ order,
                                                     const unsigned* __restrict__ count,
                                                     const unsigned* __restrict__ members,
                                                     const float* __restrict__ ginv, const float* __restrict__ uvec,
                                                     const double* __restrict__ inv_s, const double* __restrict__ y,
                                                     double* __restrict__ w, const double* __restrict__ xq)
{
    __shared__ double ys[256], red[256], cs[256], qs[256];
    // R4 > 0 (nmax = 4 R4): the wavefront's K consecutive slots hold their packed triangles
    // back to back; those float4s do not depend on y, so they go out first (overlapping the
    // members -> y chain) and are staged in LDS, where each lane reads its row.
    constexpr int NM = 4 * (R4 > 0 ? R4 : 1), KZ = 64 / NM, PSZ = (int)block_inv_stride(NM);
    constexpr int NQ = KZ * PSZ / 4, NL = (NQ + 63) / 64;
    __shared__ float4 gl[R4 > 0 ? 4 : 1][R4 > 0 ? NQ : 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = lane / nmax, i = lane - k * nmax;
    const int slot0 = (xcd_local_place(blockIdx.x, gridDim.x) * 4 + wave) * K;
    const int slot = slot0 + k;
    const bool live = k < K && slot < B;
    float4 g4[R4 > 0 ? NL : 1];
    if constexpr (R4 > 0) {
        const int nb = min(K, B - slot0);
        const int nq = nb > 0 ? nb * (PSZ / 4) : 0;
        const float4* src = reinterpret_cast<const float4*>(ginv + (long)slot0 * PSZ);
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int q = lane + 64 * m;
            g4[m] = q < nq ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    const unsigned mb = live ? members[(long)slot * nmax + i] : kNoBead;
    const int n = live ? (int)count[slot] : 0;
    const int t = live ? (int)order[slot] : 0;  // the body's own index (its constraint row)
    const bool bad = live && n > nmax;  // body larger than the workspace: poison the solve
    const bool act = live && !bad && i < n;
    const unsigned bead = act ? mb : 0u;
    const double yi = act ? y[bead] : 0.0;
    const double ui = act ? (double)uvec[(long)slot * nmax + i] : 0.0;
    const int base = wave * 64 + k * nmax;  // this body's slots
    ys[threadIdx.x] = yi;
    red[threadIdx.x] = ui * yi;
    if (xq) qs[threadIdx.x] = act ? xq[bead] : 0.0;
    if constexpr (R4 > 0) {
#pragma unroll
        for (int m = 0; m < NL; ++m)
            if (lane + 64 * m < NQ) gl[wave][lane + 64 * m] = g4[m];
    }
    __syncthreads();
    double gy = 0.0;
    if constexpr (R4 > 0) {
        if (act) {
            // entries past n (and slots past B) are zeros or never written: masked
            const float* gk = reinterpret_cast<const float*>(gl[wave]) + k * PSZ;
#pragma unroll
            for (int j = 0; j < NM; ++j)
                if (j < n) gy += (double)gk[tri_at(i, j)] * ys[base + j];
        }
    } else if (act) {
        for (int j = 0; j < n; ++j) gy += (double)ginv[ginv_at(slot, nmax, i, j)] * ys[base + j];
    }
    if (live && i == 0) {
        double dot = 0.0;
        for (int j = 0; j < n; ++j) dot += red[base + j];
        // the body's constraint row: sum_{i in t} q_i (sum_charges / set_charge,
        // cppmodule/ConductorRigid.cu:667-761), formed here from xq or read from y
        double yc = 0.0;
        if (xq)
            for (int j = 0; j < n; ++j) yc += qs[base + j];
        else
            yc = y[N + t];
        cs[base] = bad ? __builtin_nan("") : (dot + yc) * inv_s[slot];
    }
    __syncthreads();
    if (live) {
        const double c = cs[base];
        if (act) w[bead] = -gy + ui * c;
        if (i == 0) w[N + t] = c;
    }
}

// One workgroup per body of any box (body0 ascending over boxes).
__global__ __launch_bounds__(64) void k_block_apply_batch(const BlockBox* __restrict__ boxes, int nbox,
                                                          const double* __restrict__ y, double* __restrict__ w,
                                                          const int* __restrict__ skip)
{
    int lo = 0, hi = nbox - 1;
    const int k = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].body0 <= k) lo = mid;
        else hi = mid - 1;
    }
    const BlockBox& d = boxes[lo];
    if (skip && skip[lo]) return;
    block_apply_body(k - d.body0, d.N, d.nmax, d.start, d.beads, d.ginv, d.uvec, d.inv_s, y + d.xoff, w + d.xoff);
}

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

__global__ void k_zero_u32(unsigned* p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

hipError_t launch_block_setup(const SysView& s, const double L[3], int nmax, unsigned* body_count,
                              unsigned* body_start, unsigned* body_beads,
                              unsigned long long* scan_sums, float* ginv, float* uvec, double* inv_s, hipStream_t st,
                              bool invert, const BlockSlots* slots)
{
    const int B = s.N_isolated;
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks(B + 1, 256)), dim3(256), 0, st, body_count, (long)B + 1);
    if (s.N_rigid > 0)
        hipLaunchKernelGGL(k_body_count_rigid, dim3(blocks(s.N_rigid, 256)), dim3(256), 0, st, s, body_count);
    if (s.N > 0) hipLaunchKernelGGL(k_body_count_single, dim3(blocks(s.N, 256)), dim3(256), 0, st, s, body_count);
    hipError_t e = exclusive_scan_u32(body_count, body_start, (long)B + 1, scan_sums, st);
    if (e != hipSuccess) return e;
    if (s.N_rigid > 0)
        hipLaunchKernelGGL(k_body_fill_rigid, dim3(blocks((long)s.N_rigid * s.Nmax, 256)), dim3(256), 0, st, s,
                           body_start, body_beads);
    if (s.N > 0)
        hipLaunchKernelGGL(k_body_fill_single, dim3(blocks(s.N, 256)), dim3(256), 0, st, s, body_start, body_beads);
    if (!invert) return hipGetLastError();
    if (!slots || nmax > 64) return hipErrorInvalidValue;
    // the apply's slot order first: the inverses are stored by slot
    const long nbk = block_order_buckets(s.N);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks(nbk + 1, 256)), dim3(256), 0, st, slots->bucket_count, nbk + 1);
    hipLaunchKernelGGL(k_body_order_key, dim3(blocks(B, 256)), dim3(256), 0, st, B, s.N, body_start, body_beads,
                       slots->bucket_count, slots->key, slots->mkey);
    // bucket_start has nbk + 1 entries: the scan's total closes the last bucket
    e = exclusive_scan_u32(slots->bucket_count, slots->bucket_start, nbk + 1, slots->scan_sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_body_order_place, dim3(blocks(B, 256)), dim3(256), 0, st, B, slots->key, slots->bucket_start,
                       slots->order);
    hipLaunchKernelGGL(k_body_order_sort, dim3(blocks(nbk, 256)), dim3(256), 0, st, nbk, slots->bucket_start,
                       slots->mkey, slots->order);
    hipLaunchKernelGGL(k_body_slot_members, dim3(blocks((long)B * nmax, 256)), dim3(256), 0, st, B, nmax,
                       slots->order, body_start, body_beads, slots->count, slots->members);
    const unsigned wb = (unsigned)((B + 3) / 4);
    const BlockCache cache{slots->cache_key, slots->cache_sig, slots->cache_ok};
#define CR_BLOCK_WAVE(NB)                                                                                     \
    hipLaunchKernelGGL(k_block_setup_wave<NB>, dim3(wb), dim3(256), 0, st, B, nmax, slots->order, body_start, \
                       body_beads, s.pos, L[0], L[1], L[2], ginv, uvec, inv_s, cache)
    if (nmax <= 8) {
        CR_BLOCK_WAVE(8);
    } else if (nmax <= 16) {
        CR_BLOCK_WAVE(16);
    } else if (nmax <= 20) {
        CR_BLOCK_WAVE(20);  // the common 20-bead body: no padded rows
    } else if (nmax <= 24) {
        CR_BLOCK_WAVE(24);
    } else if (nmax <= 32) {
        CR_BLOCK_WAVE(32);
    } else if (nmax <= 40) {
        CR_BLOCK_WAVE(40);
    } else {
        CR_BLOCK_WAVE(64);
    }
#undef CR_BLOCK_WAVE
    return hipGetLastError();
}

long block_order_buckets(int N) { return ((long)N >> kOrderShift) + 1; }

hipError_t launch_block_apply(const SysView& s, int nmax, const BlockSlots& slots, const float* ginv,
                              const float* uvec, const double* inv_s, const double* y, double* w, hipStream_t st,
                              const double* xq)
{
    if (s.N_isolated <= 0) return hipSuccess;
    if (nmax > 64) return hipErrorInvalidValue;
    const int K = 64 / nmax;
    const unsigned nb = (unsigned)((s.N_isolated + 4 * K - 1) / (4 * K));
#define CR_BLOCK_APPLY(R4)                                                                                       \
    hipLaunchKernelGGL(k_block_apply<R4>, dim3(nb), dim3(256), 0, st, s.N_isolated, s.N, nmax, K, slots.order,   \
                       slots.count, slots.members, ginv, uvec, inv_s, y, w, xq)
    switch (nmax) {
        case 8: CR_BLOCK_APPLY(2); break;
        case 12: CR_BLOCK_APPLY(3); break;
        case 16: CR_BLOCK_APPLY(4); break;
        case 20: CR_BLOCK_APPLY(5); break;
        case 24: CR_BLOCK_APPLY(6); break;
        case 28: CR_BLOCK_APPLY(7); break;
        case 32: CR_BLOCK_APPLY(8); break;
        default: CR_BLOCK_APPLY(0); break;
    }
#undef CR_BLOCK_APPLY
    return hipGetLastError();
}

hipError_t launch_block_setup_batch(const BlockBox* boxes, int nbox, int total_bodies, int nmax_all, hipStream_t st)
{
    if (total_bodies <= 0) return hipSuccess;
    if (nmax_all > 64) return hipErrorInvalidValue;
    // one launch per body-size class up to the largest body: each body is inverted by the
    // smallest instantiation that holds it (boxes mix body sizes)
    const unsigned wb = (unsigned)((total_bodies + 3) / 4);
    const int ladder[] = {8, 16, 20, 24, 32, 40, 64};
    int nlo = 0;
    for (int NB : ladder) {
        const int last = NB >= nmax_all ? 1 : 0;
#define CR_BLOCK_WAVE_B(NBV)                                                                                    \
    hipLaunchKernelGGL(k_block_setup_wave_batch<NBV>, dim3(wb), dim3(256), 0, st, boxes, nbox, total_bodies, nlo, \
                       last)
        switch (NB) {
            case 8: CR_BLOCK_WAVE_B(8); break;
            case 16: CR_BLOCK_WAVE_B(16); break;
            case 20: CR_BLOCK_WAVE_B(20); break;
            case 24: CR_BLOCK_WAVE_B(24); break;
            case 32: CR_BLOCK_WAVE_B(32); break;
            case 40: CR_BLOCK_WAVE_B(40); break;
            default: CR_BLOCK_WAVE_B(64); break;
        }
#undef CR_BLOCK_WAVE_B
        if (NB >= nmax_all) break;
        nlo = NB;
    }
    return hipGetLastError();
}

hipError_t launch_block_apply_batch(const BlockBox* boxes, int nbox, int total_bodies, const double* y, double* w,
                                    const int* skip, hipStream_t st)
{
    if (total_bodies <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_apply_batch, dim3((unsigned)total_bodies), dim3(64), 0, st, boxes, nbox, y, w, skip);
    return hipGetLastError();
}

}  // namespace cr
