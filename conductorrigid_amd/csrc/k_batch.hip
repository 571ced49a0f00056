// k_batch.hip -- operator rows and FP64 Krylov steps for a batch of independent boxes
// solved in lockstep.
//
// The reference solves one box per computeForces call (cppmodule/ConductorRigid.cu:1051-1151)
// with CUSP's GMRES, one device->host sync per dot product.  For many small boxes that
// leaves the GPU idle between tiny launches, so here all boxes advance together: their
// unknown vectors are concatenated, every launch below covers every box (each box owns a
// contiguous run of workgroups, so no dot product crosses a box), and one Arnoldi step of
// the whole batch reads its column index from a device counter.  The step is therefore
// the same sequence of launches every time and can be captured once as a HIP graph.
// The per-box arithmetic is the single-box solver's (k_krylov.hip, k_real.hip): same
// CGS2 projections, Givens rotations, back substitution and row formulas.
#include "cr_kernels.h"

#include <algorithm>

namespace cr {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kB = kBatchBlock;
constexpr int kCols = 8;            // basis columns per workgroup row of launch_bdot
constexpr int kMaxSegBlocks = 64;   // workgroups per box in vector launches

// Box owning global workgroup k: the last box whose first workgroup is <= k (boxes that
// own no workgroups share their successor's first index and are never selected).
template <typename S, int S::*F>
__device__ __forceinline__ int owner(const S* __restrict__ a, int nbox, int k)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a[mid].*F <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int K>
__device__ __forceinline__ void block_sum(double (&acc)[K], double* out)
{
    __shared__ double red[kB / 64][K];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
        for (int w = 0; w < kB / 64; ++w) s += red[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

// The element range of workgroup k inside its box's segment.
__device__ __forceinline__ void seg_range(const BatchSeg& s, int k, long& lo, long& hi)
{
    const long per = ((long)s.n + s.nblk - 1) / s.nblk;
    lo = s.xoff + (long)(k - s.blk0) * per;
    hi = min(lo + per, s.xoff + (long)s.n);
}

// ---- operator rows (PotentialChargeMultiply's real-space and constraint rows,
// cppmodule/ConductorRigid.cu:469-567, 667-761; the single-box forms are in k_real.hip)

// xt: the batch vector in the row precision T (the gathered operand; x itself when T = double)
template <typename T>
__global__ __launch_bounds__(kB) void k_brows(const RowBox* __restrict__ boxes, int nbox, const double* __restrict__ x,
                                              const T* __restrict__ xt, double* __restrict__ y, int precond,
                                              const int* __restrict__ skip)
{
    const RowBox& d = boxes[owner<RowBox, &RowBox::blk0>(boxes, nbox, blockIdx.x)];
    const int b = (int)(&d - boxes);
    if (skip && skip[b]) return;
    const int i = (blockIdx.x - d.blk0) * kB + threadIdx.x;
    if (i >= d.N) return;
    const double* q = x + d.xoff;
    const double* Phi = q + d.N;
    const T* qt = xt + d.xoff;
    const T* phi = static_cast<const T*>(d.phi_recip);
    const T* val = static_cast<const T*>(d.val);
    T acc = phi[i] + (T)d.selfterm * qt[i];
    const unsigned len = d.row_len[i];
    for (unsigned k = 0; k < len; ++k) acc += val[k * d.ldn + i] * qt[d.col[k * d.ldn + i]];
    const unsigned tag = d.iso[i];
    const double out = Phi[tag] - (double)acc;
    y[d.xoff + i] = precond ? -4.0 * kPi * out : out;
    // an isolated bead carries its own body's charge row (set_charge)
    if (d.rigid_membership[i] == NOT_MEMBER) y[d.xoff + d.N + tag] = precond ? q[i] / (4.0 * kPi) : q[i];
}

__global__ __launch_bounds__(kB) void k_brigid_rows(const RowBox* __restrict__ boxes, int nbox,
                                                    const double* __restrict__ x, double* __restrict__ y, int precond,
                                                    const int* __restrict__ skip)
{
    const RowBox& d = boxes[owner<RowBox, &RowBox::rblk0>(boxes, nbox, blockIdx.x)];
    const int b = (int)(&d - boxes);
    if (skip && skip[b]) return;
    const int r = (blockIdx.x - d.rblk0) * kB + threadIdx.x;
    if (r >= d.N_rigid) return;
    const unsigned size = d.rigid_size[r];
    if (size == 0) return;
    const unsigned* mem = d.rigid_members + (long)r * d.Nmax;
    const double* q = x + d.xoff;
    double acc = 0.0;
    for (unsigned k = 0; k < size; ++k) acc += q[mem[k]];
    y[d.xoff + d.N + d.iso[mem[0]]] = precond ? acc / (double)size : acc;
}

// The reference's diagonal preconditioner (precondition, cppmodule/ConductorRigid.cu:816-859):
// bead rows -4 pi, an isolated bead's body row 1/4pi, a rigid body's row 1/size.
__global__ __launch_bounds__(kB) void k_bprecond_beads(const RowBox* __restrict__ boxes, int nbox,
                                                       const double* __restrict__ x, double* __restrict__ y)
{
    const RowBox& d = boxes[owner<RowBox, &RowBox::blk0>(boxes, nbox, blockIdx.x)];
    const int i = (blockIdx.x - d.blk0) * kB + threadIdx.x;
    if (i >= d.N) return;
    y[d.xoff + i] = -4.0 * kPi * x[d.xoff + i];
    if (d.rigid_membership[i] == NOT_MEMBER) {
        const long r = d.xoff + d.N + d.iso[i];
        y[r] = x[r] / (4.0 * kPi);
    }
}

__global__ __launch_bounds__(kB) void k_bprecond_rigid(const RowBox* __restrict__ boxes, int nbox,
                                                       const double* __restrict__ x, double* __restrict__ y)
{
    const RowBox& d = boxes[owner<RowBox, &RowBox::rblk0>(boxes, nbox, blockIdx.x)];
    const int r = (blockIdx.x - d.rblk0) * kB + threadIdx.x;
    if (r >= d.N_rigid) return;
    const unsigned size = d.rigid_size[r];
    if (size == 0) return;
    const long row = d.xoff + d.N + d.iso[d.rigid_members[(long)r * d.Nmax]];
    y[row] = x[row] / (double)size;
}

// ---- Krylov (single-box forms in k_krylov.hip)

__global__ __launch_bounds__(kB) void k_bdot(const double* __restrict__ V, long ld, int fixed_cols,
                                             const int* __restrict__ step, const double* __restrict__ w,
                                             const BatchSeg* __restrict__ segs, int nbox, double* __restrict__ partial,
                                             const int* __restrict__ skip)
{
    const BatchSeg& s = segs[owner<BatchSeg, &BatchSeg::blk0>(segs, nbox, blockIdx.x)];
    const int b = (int)(&s - segs);
    if (skip && skip[b]) return;
    const int ncols = fixed_cols > 0 ? fixed_cols : *step + 1;
    const int c0 = blockIdx.y * kCols;
    if (c0 >= ncols) return;
    long lo, hi;
    seg_range(s, blockIdx.x, lo, hi);
    double acc[kCols];
#pragma unroll
    for (int k = 0; k < kCols; ++k) acc[k] = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += kB) {
        const double wi = w[i];
#pragma unroll
        for (int k = 0; k < kCols; ++k)
            if (c0 + k < ncols) acc[k] += V[(long)(c0 + k) * ld + i] * wi;
    }
    __shared__ double res[kCols];
    block_sum<kCols>(acc, res);
    if (threadIdx.x < kCols && c0 + (int)threadIdx.x < ncols)
        partial[(long)blockIdx.x * 32 + c0 + threadIdx.x] = res[threadIdx.x];
}

__device__ __forceinline__ double* red_target(BatchGmres& g, int which, int c)
{
    switch (which) {
        case kRedH1: return &g.g.h1[c];
        case kRedH2: return &g.g.h2[c];
        case kRedWnorm: return &g.g.wnorm2;
        default: return &g.norm2;
    }
}

// One workgroup per (column, box); fixed-order sums, so results are deterministic.
__global__ __launch_bounds__(kB) void k_breduce(const double* __restrict__ partial, const BatchSeg* __restrict__ segs,
                                                int fixed_cols, const int* __restrict__ step,
                                                BatchGmres* __restrict__ gs, int which, const int* __restrict__ skip)
{
    const int c = blockIdx.x, b = blockIdx.y;
    if (skip && skip[b]) return;
    const int ncols = fixed_cols > 0 ? fixed_cols : *step + 1;
    if (c >= ncols) return;
    const BatchSeg& s = segs[b];
    double v = 0.0;
    for (int k = threadIdx.x; k < s.nblk; k += kB) v += partial[(long)(s.blk0 + k) * 32 + c];
    double a[1] = {v};
    __shared__ double r[1];
    block_sum<1>(a, r);
    if (threadIdx.x == 0) *red_target(gs[b], which, c) = r[0];
}

__global__ __launch_bounds__(kB) void k_bgs_update(const double* __restrict__ V, long ld, const int* __restrict__ step,
                                                   double* __restrict__ w, const BatchSeg* __restrict__ segs, int nbox,
                                                   const BatchGmres* __restrict__ gs, int which, int mode,
                                                   double* __restrict__ partial, const int* __restrict__ skip)
{
    const BatchSeg& s = segs[owner<BatchSeg, &BatchSeg::blk0>(segs, nbox, blockIdx.x)];
    const int b = (int)(&s - segs);
    if (skip && skip[b]) return;
    const int ncols = *step + 1;
    __shared__ double hs[32];
    const double* h = which == kRedH1 ? gs[b].g.h1 : gs[b].g.h2;
    if (threadIdx.x < 32) hs[threadIdx.x] = (threadIdx.x < (unsigned)ncols) ? h[threadIdx.x] : 0.0;
    __syncthreads();
    long lo, hi;
    seg_range(s, blockIdx.x, lo, hi);
    double acc[1] = {0.0};
    for (long i = lo + threadIdx.x; i < hi; i += kB) {
        double wi = w[i];
        for (int k = 0; k < ncols; ++k) wi -= hs[k] * V[(long)k * ld + i];
        w[i] = wi;
        acc[0] += wi * wi;
    }
    if (mode == 2) {
        __shared__ double r[1];
        block_sum<1>(acc, r);
        if (threadIdx.x == 0) partial[(long)blockIdx.x * 32] = r[0];
    }
}

// Givens plane rotations (Saad, Iterative Methods for Sparse Linear Systems, Alg. 6.9),
// as in the single-box solver.
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

__global__ void k_bgivens(BatchGmres* __restrict__ gs, int nbox, const int* __restrict__ step, int R,
                          int* __restrict__ skip)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox || skip[b]) return;
    GmresState* g = &gs[b].g;
    const int i = *step;
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
    if (!(g->resid > gs[b].tol)) skip[b] = 1;  // converged (or non-finite: the host checks)
}

__global__ __launch_bounds__(kB) void k_bscale_copy(const double* __restrict__ w, const BatchSeg* __restrict__ segs,
                                                    int nbox, const BatchGmres* __restrict__ gs,
                                                    const int* __restrict__ step, int col_fixed, double sign,
                                                    double* __restrict__ V, long ld, const int* __restrict__ skip)
{
    const BatchSeg& s = segs[owner<BatchSeg, &BatchSeg::blk0>(segs, nbox, blockIdx.x)];
    const int b = (int)(&s - segs);
    if (skip && skip[b]) return;
    const int col = col_fixed >= 0 ? col_fixed : *step + 1;
    const double f = sign * (col_fixed >= 0 ? gs[b].g.inv_beta : gs[b].g.inv_norm);
    long lo, hi;
    seg_range(s, blockIdx.x, lo, hi);
    double* dst = V + (long)col * ld;
    for (long i = lo + threadIdx.x; i < hi; i += kB) dst[i] = f * w[i];
}

__global__ void k_bcopy_col(const double* __restrict__ V, long ld, const int* __restrict__ step,
                            double* __restrict__ u, long n)
{
    const double* src = V + (long)(*step) * ld;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        u[i] = src[i];
}

__global__ void k_bcycle_init(BatchGmres* __restrict__ gs, int nbox, const int* __restrict__ finished,
                              int* __restrict__ skip)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    if (finished[b]) {
        skip[b] = 1;
        return;
    }
    GmresState* g = &gs[b].g;
    const double beta = sqrt(gs[b].norm2);
    g->beta = beta;
    g->inv_beta = (beta != 0.0) ? 1.0 / beta : 0.0;
    for (int k = 0; k < 33; ++k) g->s[k] = 0.0;
    g->s[0] = beta;
    g->resid = beta;
    g->ncols = 0;
    skip[b] = !(beta > gs[b].tol);
}

__global__ void k_bbacksub(BatchGmres* __restrict__ gs, int nbox, int R, const int* __restrict__ finished)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox || finished[b]) return;
    GmresState* g = &gs[b].g;
    const int ld = R + 1;
    const int m = g->ncols;
    for (int k = 0; k < m; ++k) g->y[k] = g->s[k];
    for (int j = m - 1; j >= 0; --j) {
        g->y[j] /= g->H[j * ld + j];
        for (int k = j - 1; k >= 0; --k) g->y[k] -= g->H[j * ld + k] * g->y[j];
    }
}

__global__ __launch_bounds__(kB) void k_bcombine(const double* __restrict__ V, long ld,
                                                 const BatchSeg* __restrict__ segs, int nbox,
                                                 const BatchGmres* __restrict__ gs, const int* __restrict__ finished,
                                                 double* __restrict__ x)
{
    const BatchSeg& s = segs[owner<BatchSeg, &BatchSeg::blk0>(segs, nbox, blockIdx.x)];
    const int b = (int)(&s - segs);
    if (finished[b]) return;
    const GmresState& g = gs[b].g;
    const int K = g.ncols;
    long lo, hi;
    seg_range(s, blockIdx.x, lo, hi);
    for (long i = lo + threadIdx.x; i < hi; i += kB) {
        double xi = x[i];
        for (int k = 0; k < K; ++k) xi += g.y[k] * V[(long)k * ld + i];
        x[i] = xi;
    }
}

__global__ void k_bstep_end(int* __restrict__ step, const BatchGmres* __restrict__ gs, const int* __restrict__ skip,
                            int nbox, double* __restrict__ pub, int inc)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) {
        pub[3 * b] = gs[b].g.resid;
        pub[3 * b + 1] = (double)skip[b];
        pub[3 * b + 2] = gs[b].tol;
    }
    if (b == 0 && inc) *step += 1;
}

__global__ void k_bset_tol(BatchGmres* __restrict__ gs, int nbox, const double* __restrict__ rel)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) gs[b].tol = rel[b] * sqrt(gs[b].norm2);
}

__global__ void k_bcollect(const BatchGmres* __restrict__ gs, int nbox, double* __restrict__ out, int stride, int slot)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) out[(long)stride * b + slot] = gs[b].norm2;
}

__global__ void k_bset_step(int* step, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = v;
}

inline unsigned cdiv(long n, long d) { return (unsigned)std::max(1L, (n + d - 1) / d); }

}  // namespace

int batch_blocks(long n)
{
    long b = (n + 4L * kB - 1) / (4L * kB);
    if (b > kMaxSegBlocks) b = kMaxSegBlocks;
    return (int)std::max(1L, b);
}

template <typename T>
hipError_t launch_brows(const RowBox* boxes, int nbox, int bead_blocks, int rigid_blocks, const double* x,
                        const T* xt, double* y, int precond, const int* skip, hipStream_t st)
{
    if (bead_blocks > 0)
        hipLaunchKernelGGL(k_brows<T>, dim3(bead_blocks), dim3(kB), 0, st, boxes, nbox, x, xt, y, precond, skip);
    if (rigid_blocks > 0)
        hipLaunchKernelGGL(k_brigid_rows, dim3(rigid_blocks), dim3(kB), 0, st, boxes, nbox, x, y, precond, skip);
    return hipGetLastError();
}
template hipError_t launch_brows<float>(const RowBox*, int, int, int, const double*, const float*, double*, int,
                                        const int*, hipStream_t);
template hipError_t launch_brows<double>(const RowBox*, int, int, int, const double*, const double*, double*, int,
                                         const int*, hipStream_t);

hipError_t launch_bdot(const double* V, long ld, int fixed_cols, const int* step, const double* w,
                       const BatchSeg* segs, int nbox, int total_blocks, double* partial, const int* skip,
                       hipStream_t st)
{
    // column chunks sized for the widest projection (32 columns) when read from the counter
    const int cols = fixed_cols > 0 ? fixed_cols : 32;
    const unsigned ny = (unsigned)((cols + kCols - 1) / kCols);
    hipLaunchKernelGGL(k_bdot, dim3((unsigned)total_blocks, ny), dim3(kB), 0, st, V, ld, fixed_cols, step, w, segs,
                       nbox, partial, skip);
    return hipGetLastError();
}

hipError_t launch_breduce(const double* partial, const BatchSeg* segs, int nbox, int fixed_cols, const int* step,
                          BatchGmres* gs, int which, const int* skip, hipStream_t st)
{
    const unsigned nx = (unsigned)(fixed_cols > 0 ? fixed_cols : 32);
    hipLaunchKernelGGL(k_breduce, dim3(nx, (unsigned)nbox), dim3(kB), 0, st, partial, segs, fixed_cols, step, gs,
                       which, skip);
    return hipGetLastError();
}

hipError_t launch_bgs_update(const double* V, long ld, const int* step, double* w, const BatchSeg* segs, int nbox,
                             int total_blocks, const BatchGmres* gs, int which, int mode, double* partial,
                             const int* skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_bgs_update, dim3((unsigned)total_blocks), dim3(kB), 0, st, V, ld, step, w, segs, nbox, gs,
                       which, mode, partial, skip);
    return hipGetLastError();
}

hipError_t launch_bgivens(BatchGmres* gs, int nbox, const int* step, int R, int* skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_bgivens, dim3(cdiv(nbox, 64)), dim3(64), 0, st, gs, nbox, step, R, skip);
    return hipGetLastError();
}

hipError_t launch_bscale_copy(const double* w, const BatchSeg* segs, int nbox, int total_blocks,
                              const BatchGmres* gs, const int* step, int col_fixed, double sign, double* V, long ld,
                              const int* skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_bscale_copy, dim3((unsigned)total_blocks), dim3(kB), 0, st, w, segs, nbox, gs, step,
                       col_fixed, sign, V, ld, skip);
    return hipGetLastError();
}

hipError_t launch_bcopy_col(const double* V, long ld, const int* step, double* u, long n, hipStream_t st)
{
    hipLaunchKernelGGL(k_bcopy_col, dim3(std::min(cdiv(n, kB), 2048u)), dim3(kB), 0, st, V, ld, step, u, n);
    return hipGetLastError();
}

hipError_t launch_bcycle_init(BatchGmres* gs, int nbox, const int* finished, int* skip, hipStream_t st)
{
    hipLaunchKernelGGL(k_bcycle_init, dim3(cdiv(nbox, 64)), dim3(64), 0, st, gs, nbox, finished, skip);
    return hipGetLastError();
}

hipError_t launch_bbacksub(BatchGmres* gs, int nbox, int R, const int* finished, hipStream_t st)
{
    hipLaunchKernelGGL(k_bbacksub, dim3(cdiv(nbox, 64)), dim3(64), 0, st, gs, nbox, R, finished);
    return hipGetLastError();
}

hipError_t launch_bcombine(const double* V, long ld, const BatchSeg* segs, int nbox, int total_blocks,
                           const BatchGmres* gs, const int* finished, double* x, hipStream_t st)
{
    hipLaunchKernelGGL(k_bcombine, dim3((unsigned)total_blocks), dim3(kB), 0, st, V, ld, segs, nbox, gs, finished, x);
    return hipGetLastError();
}

hipError_t launch_bstep_end(int* step, const BatchGmres* gs, const int* skip, int nbox, double* pub, int inc,
                            hipStream_t st)
{
    hipLaunchKernelGGL(k_bstep_end, dim3(cdiv(nbox, 64)), dim3(64), 0, st, step, gs, skip, nbox, pub, inc);
    return hipGetLastError();
}

hipError_t launch_bset_tol(BatchGmres* gs, int nbox, const double* rel, hipStream_t st)
{
    hipLaunchKernelGGL(k_bset_tol, dim3(cdiv(nbox, 64)), dim3(64), 0, st, gs, nbox, rel);
    return hipGetLastError();
}

hipError_t launch_bcollect(const BatchGmres* gs, int nbox, double* out, int stride, int slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_bcollect, dim3(cdiv(nbox, 64)), dim3(64), 0, st, gs, nbox, out, stride, slot);
    return hipGetLastError();
}

hipError_t launch_bprecond_diag(const RowBox* boxes, int nbox, int bead_blocks, int rigid_blocks, const double* x,
                                double* y, hipStream_t st)
{
    if (bead_blocks > 0) hipLaunchKernelGGL(k_bprecond_beads, dim3(bead_blocks), dim3(kB), 0, st, boxes, nbox, x, y);
    if (rigid_blocks > 0)
        hipLaunchKernelGGL(k_bprecond_rigid, dim3(rigid_blocks), dim3(kB), 0, st, boxes, nbox, x, y);
    return hipGetLastError();
}

hipError_t launch_bset_step(int* step, int value, hipStream_t st)
{
    hipLaunchKernelGGL(k_bset_step, dim3(1), dim3(64), 0, st, step, value);
    return hipGetLastError();
}

}  // namespace cr
