// k_real.hip -- real-space pair sums, constraint rows, objective, preconditioner and
// membership of the ConductorRigid saddle-point operator.
//
// Reference kernels: real_space (cppmodule/ConductorRigid.cu:469-567),
// real_space_force (:570-664), sum_charges / set_charge (:667-761),
// compute_objective (:764-813), precondition (:816-859),
// initialize/update_membership (:92-141), zeroforce (:56-66).
#include "cr_kernels.h"

namespace cr {

namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double min_image(double d, double L) { return d - L * rint(d / L); }

// Linear interpolation in a (value, next value) table (cppmodule/ConductorRigid.cu:549-554),
// index clamped to [0, N_table-1].
template <typename T>
__device__ __forceinline__ T table_lookup(const RealParams& rp, const T* __restrict__ tab, double d)
{
    int ti = (int)floor(rp.N_table * (d - rp.drtable) / (rp.rc - rp.drtable));
    ti = ti < 0 ? 0 : (ti > rp.N_table - 1 ? rp.N_table - 1 : ti);
    const double lin = d / rp.drtable - ti - 1.0;
    const T a = tab[2 * ti], b = tab[2 * ti + 1];
    return a + (b - a) * (T)lin;
}

template <typename T>
__global__ __launch_bounds__(256) void k_real_space(RealParams rp, SysView s, const double* __restrict__ x,
                                                    const T* __restrict__ phi_recip, const T* __restrict__ table,
                                                    double* __restrict__ y, int precond)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const double* q = x;
    const double* Phi = x + s.N;
    const double4 ri = s.pos[i];
    T acc = phi_recip[i] + (T)rp.selfterm * (T)q[i];
    const unsigned head = s.head_list[i], nn = s.n_neigh[i];
    for (unsigned jj = 0; jj < nn; ++jj) {
        const unsigned j = s.nlist[head + jj];
        const double4 rj = s.pos[j];
        const double dx = min_image(ri.x - rj.x, rp.L[0]);
        const double dy = min_image(ri.y - rj.y, rp.L[1]);
        const double dz = min_image(ri.z - rj.z, rp.L[2]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < rp.rc2 && d2 >= rp.rmin2) acc += table_lookup<T>(rp, table, sqrt(d2)) * (T)q[j];
    }
    const double out = Phi[s.iso[i]] - (double)acc;
    y[i] = precond ? -4.0 * kPi * out : out;
}

// The real-space part of the operator is constant while GMRES runs (positions are
// fixed), so it is assembled once per force evaluation: the entries of bead i are the
// pairs inside [dr, rc) with the same table interpolation as real_space.  Storage is
// column-major ELL -- entry k of row i at k*ldn + i, row_len[i] entries per row -- so
// that a wavefront's 64 consecutive rows read 64 consecutive words per step.
// One atomic per workgroup: atomics on a single address are serialised at the memory
// side, so one per wavefront (4096 at 1M) took ~50 us.
__global__ __launch_bounds__(256) void k_max_u32(const unsigned* __restrict__ a, int n, unsigned* __restrict__ out)
{
    __shared__ unsigned wmax[4];
    unsigned v = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        v = max(v, a[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, off, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}

// The neighbour rows of a workgroup's 256 beads usually form one contiguous span of the
// list (head_list is a running sum); stage that span in LDS with coalesced loads so each
// thread then walks its own row from LDS.  Rows outside the staged span (any other
// head_list layout, or a span beyond kNlStage) are read from global memory.
constexpr int kNlStage = 8192;

struct NlSpan {
    unsigned lo, hi;  // staged entries [lo, hi)
};

// blk: the workgroup's block of beads [blk * blockDim.x, ...) in the system s
__device__ __forceinline__ NlSpan stage_nlist(const SysView& s, unsigned* sh, int blk)
{
    __shared__ unsigned bounds[2];
    const int i0 = blk * blockDim.x;
    const int i1 = min(s.N, i0 + (int)blockDim.x) - 1;
    if (threadIdx.x == 0) {
        const unsigned lo = s.head_list[i0], hi = s.head_list[i1] + s.n_neigh[i1];
        const bool ok = hi >= lo && hi - lo <= (unsigned)kNlStage;
        bounds[0] = lo;
        bounds[1] = ok ? hi : lo;
    }
    __syncthreads();
    NlSpan sp{bounds[0], bounds[1]};
    // a thread's loads go out in batches of 8 before their stores, so that their latencies
    // overlap (one load and wait per entry would serialise ~30 round trips per workgroup)
    constexpr unsigned kBatch = 8;
    for (unsigned k0 = sp.lo + threadIdx.x; k0 < sp.hi; k0 += kBatch * blockDim.x) {
        unsigned v[kBatch];
#pragma unroll
        for (unsigned u = 0; u < kBatch; ++u) {
            const unsigned k = k0 + u * blockDim.x;
            v[u] = k < sp.hi ? s.nlist[k] : 0u;
        }
#pragma unroll
        for (unsigned u = 0; u < kBatch; ++u) {
            const unsigned k = k0 + u * blockDim.x;
            if (k < sp.hi) sh[k - sp.lo] = v[u];
        }
    }
    __syncthreads();
    return sp;
}

// Neighbours of a row are visited in groups of kNbGroup: their indices, then their
// positions, are loaded before any is used (the row's order of visits is unchanged).
constexpr unsigned kNbGroup = 4;

__device__ __forceinline__ unsigned nl_at(const SysView& s, const unsigned* sh, const NlSpan& sp, unsigned k)
{
    return (k >= sp.lo && k < sp.hi) ? sh[k - sp.lo] : s.nlist[k];
}

template <typename T>
__device__ __forceinline__ void csr_build_rows(const RealParams& rp, const SysView& s, const T* __restrict__ table,
                                               long ldn, unsigned* __restrict__ col, T* __restrict__ val,
                                               unsigned* __restrict__ row_len, int blk)
{
    __shared__ unsigned sh[kNlStage];
    const NlSpan sp = stage_nlist(s, sh, blk);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const double4 ri = s.pos[i];
    const unsigned head = s.head_list[i], nn = s.n_neigh[i];
    unsigned cnt = 0;
    auto visit = [&](unsigned j, const double4& rj) {
        const double dx = min_image(ri.x - rj.x, rp.L[0]);
        const double dy = min_image(ri.y - rj.y, rp.L[1]);
        const double dz = min_image(ri.z - rj.z, rp.L[2]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < rp.rc2 && d2 >= rp.rmin2) {
            col[cnt * ldn + i] = j;
            val[cnt * ldn + i] = table_lookup<T>(rp, table, sqrt(d2));
            ++cnt;
        }
    };
    unsigned jj = 0;
    for (; jj + kNbGroup <= nn; jj += kNbGroup) {
        unsigned j[kNbGroup];
        double4 rj[kNbGroup];
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) j[u] = nl_at(s, sh, sp, head + jj + u);
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) rj[u] = s.pos[j[u]];
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) visit(j[u], rj[u]);
    }
    for (; jj < nn; ++jj) {
        const unsigned j = nl_at(s, sh, sp, head + jj);
        visit(j, s.pos[j]);
    }
    row_len[i] = cnt;
}

template <typename T>
__global__ __launch_bounds__(256) void k_real_csr_build(RealParams rp, SysView s, const T* __restrict__ table,
                                                        long ldn, unsigned* __restrict__ col, T* __restrict__ val,
                                                        unsigned* __restrict__ row_len)
{
    csr_build_rows<T>(rp, s, table, ldn, col, val, row_len, blockIdx.x);
}

// Batched boxes: a box owns the workgroups [blk0, blk0 + ceil(N / 256)) of the launch.
__device__ __forceinline__ int pair_box_of(const PairBox* __restrict__ boxes, int nbox, int g)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].blk0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ __launch_bounds__(256) void k_real_csr_build_batch(const PairBox* __restrict__ boxes, int nbox)
{
    const PairBox& d = boxes[pair_box_of(boxes, nbox, blockIdx.x)];
    csr_build_rows<T>(d.rp, d.s, static_cast<const T*>(d.table), d.ldn, d.col, static_cast<T*>(d.val), d.row_len,
                      blockIdx.x - d.blk0);
}

// real_space on the assembled rows: y_i = Phi_{b(i)} - (phi_i + self q_i + sum_j A_ij q_j).
// qt: the charges in the row precision T (the gathered operand; x itself when T = double).
template <typename T>
__global__ __launch_bounds__(256) void k_real_space_csr(RealParams rp, SysView s, const double* __restrict__ x,
                                                        const T* __restrict__ qt, const T* __restrict__ phi_recip,
                                                        long ldn, const unsigned* __restrict__ col,
                                                        const T* __restrict__ val,
                                                        const unsigned* __restrict__ row_len, double* __restrict__ y,
                                                        int precond)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const double* Phi = x + s.N;
    T acc = phi_recip[i] + (T)rp.selfterm * qt[i];
    const unsigned len = row_len[i];
    // groups of 8 entries: their column and value loads, then their gathers, go out
    // together; the sum keeps its order (same result as one entry at a time)
    constexpr unsigned U = 8;
    unsigned k = 0;
    for (; k + U <= len; k += U) {
        unsigned c[U];
        T v[U], g[U];
#pragma unroll
        for (unsigned u = 0; u < U; ++u) {
            c[u] = col[(k + u) * ldn + i];
            v[u] = val[(k + u) * ldn + i];
        }
#pragma unroll
        for (unsigned u = 0; u < U; ++u) g[u] = qt[c[u]];
#pragma unroll
        for (unsigned u = 0; u < U; ++u) acc += v[u] * g[u];
    }
    for (; k < len; ++k) acc += val[k * ldn + i] * qt[col[k * ldn + i]];
    const double out = Phi[s.iso[i]] - (double)acc;
    y[i] = precond ? -4.0 * kPi * out : out;
}

// y[i] += Phi[body of bead i] for the bead rows: A [q; Phi] = A [q; 0] + that.
__global__ void k_add_body_potentials(SysView s, const double* __restrict__ Phi, double* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.N) y[i] += Phi[s.iso[i]];
}

template <typename T>
__device__ __forceinline__ void force_rows(const RealParams& rp, const SysView& s, const double* __restrict__ q,
                                           const T* __restrict__ gtable, double E0x, double E0y, double E0z,
                                           double* __restrict__ force, int blk, bool accum)
{
    __shared__ unsigned sh[kNlStage];
    const NlSpan sp = stage_nlist(s, sh, blk);
    const int i = blk * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const double qi = q[i];
    // accum: force += (the reciprocal part is already there); otherwise force = this part
    double Fx = (accum ? force[4 * i] : 0.0) + qi * E0x;
    double Fy = (accum ? force[4 * i + 1] : 0.0) + qi * E0y;
    double Fz = (accum ? force[4 * i + 2] : 0.0) + qi * E0z;
    T fx = 0, fy = 0, fz = 0;
    const double4 ri = s.pos[i];
    const unsigned head = s.head_list[i], nn = s.n_neigh[i];
    auto visit = [&](const double4& rj, double qj) {
        const double dx = min_image(ri.x - rj.x, rp.L[0]);
        const double dy = min_image(ri.y - rj.y, rp.L[1]);
        const double dz = min_image(ri.z - rj.z, rp.L[2]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < rp.rc2 && d2 >= rp.rmin2) {
            const double d = sqrt(d2);
            const T A = table_lookup<T>(rp, gtable, d);
            const T c = -A * (T)(qi * qj / d);
            fx += c * (T)dx;
            fy += c * (T)dy;
            fz += c * (T)dz;
        }
    };
    unsigned jj = 0;
    for (; jj + kNbGroup <= nn; jj += kNbGroup) {
        unsigned j[kNbGroup];
        double4 rj[kNbGroup];
        double qj[kNbGroup];
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) j[u] = nl_at(s, sh, sp, head + jj + u);
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) {
            rj[u] = s.pos[j[u]];
            qj[u] = q[j[u]];
        }
#pragma unroll
        for (unsigned u = 0; u < kNbGroup; ++u) visit(rj[u], qj[u]);
    }
    for (; jj < nn; ++jj) {
        const unsigned j = nl_at(s, sh, sp, head + jj);
        visit(s.pos[j], q[j]);
    }
    force[4 * i] = Fx + (double)fx;
    force[4 * i + 1] = Fy + (double)fy;
    force[4 * i + 2] = Fz + (double)fz;
    force[4 * i + 3] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_real_space_force(RealParams rp, SysView s, const double* __restrict__ q,
                                                          const T* __restrict__ gtable, double E0x, double E0y,
                                                          double E0z, double* __restrict__ force, int accum)
{
    force_rows<T>(rp, s, q, gtable, E0x, E0y, E0z, force, blockIdx.x, accum != 0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_real_space_force_batch(const PairBox* __restrict__ boxes, int nbox)
{
    const PairBox& d = boxes[pair_box_of(boxes, nbox, blockIdx.x)];
    force_rows<T>(d.rp, d.s, d.q, static_cast<const T*>(d.table), d.E0[0], d.E0[1], d.E0[2], d.force,
                  blockIdx.x - d.blk0, true);
}

// sum_charges: y_bottom[tag(r)] = sum of bead charges of rigid body r (FP64), optionally
// scaled by the preconditioner's 1/size.
// K = 64 / W rigid bodies per wavefront with W = min(Nmax, 64) lanes each; lane m of a
// body sums members m, m + W, ...: the member loads are independent, and the per-body
// sum is a fixed-order pass over LDS.
__global__ __launch_bounds__(256) void k_sum_charges(SysView s, int W, const double* __restrict__ q,
                                                     double* __restrict__ yb, int precond)
{
    __shared__ double part[256];
    const int K = 64 / W;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = lane / W, m = lane - k * W;
    const int r = (blockIdx.x * 4 + wave) * K + k;
    const bool live = k < K && r < s.N_rigid;
    const unsigned* mem = s.rigid_members + (long)r * s.Nmax;
    // member slot m (< W <= Nmax, inside the body's row) is read together with the size;
    // lane 0 also fetches the body's row index before the barrier
    const unsigned first = live ? mem[m] : 0u;
    const unsigned size = live ? s.rigid_size[r] : 0u;
    double acc = 0.0;
    unsigned row = 0;
    if ((unsigned)m < size) {
        acc = q[first];
        if (m == 0) row = s.iso[first];
    }
    for (unsigned j = (unsigned)(m + W); j < size; j += (unsigned)W) acc += q[mem[j]];
    part[threadIdx.x] = acc;
    __syncthreads();
    if (live && m == 0 && size > 0) {
        const double* p = part + wave * 64 + k * W;
        double t = 0.0;
        for (int j = 0; j < W; ++j) t += p[j];
        yb[row] = precond ? t / (double)size : t;
    }
}

// set_charge: isolated beads carry their own body's charge row.
__global__ void k_set_charge(SysView s, const double* __restrict__ q, double* __restrict__ yb, int precond)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    if (s.rigid_membership[i] != NOT_MEMBER) return;
    yb[s.iso[i]] = precond ? q[i] / (4.0 * kPi) : q[i];
}

__global__ void k_precond_diag(int N, int B, const double* __restrict__ x, double* __restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) y[i] = -4.0 * kPi * x[i];
    else if (i < N + B) y[i] = x[i] / (4.0 * kPi);
}

__global__ void k_precond_rigid(SysView s, const double* __restrict__ x, double* __restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.N_rigid) return;
    const unsigned size = s.rigid_size[r];
    if (size == 0) return;
    const unsigned tag = s.iso[s.rigid_members[(long)r * s.Nmax]];
    y[s.N + tag] = x[s.N + tag] / (double)size;
}

// compute_objective with zero_array folded in: rigid beads get -(r_i - R_b).E0 (minimum
// image), beads outside rigid bodies get 0, rows N.. get the body charges.
__device__ __forceinline__ void objective_rows(const SysView& s, double Lx, double Ly, double Lz, double E0x,
                                               double E0y, double E0z, double* __restrict__ b, long t)
{
    const long nslot = (long)s.N_rigid * s.Nmax;
    if (t < nslot) {
        const unsigned idx = s.rigid_members[t];
        if (idx != NOT_MEMBER) {
            const double4 p = s.pos[idx];
            const double4 c = s.rigid_com[t / s.Nmax];
            const double rx = min_image(p.x - c.x, Lx), ry = min_image(p.y - c.y, Ly), rz = min_image(p.z - c.z, Lz);
            b[idx] = -(rx * E0x + ry * E0y + rz * E0z);
        }
    }
    if (t < s.N && s.rigid_membership[t] == NOT_MEMBER) b[t] = 0.0;
    if (t < s.N_isolated) b[s.N + t] = s.body_charge[t];
}

__global__ void k_objective(SysView s, double Lx, double Ly, double Lz, double E0x, double E0y, double E0z,
                            double* __restrict__ b)
{
    objective_rows(s, Lx, Ly, Lz, E0x, E0y, E0z, b, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched boxes (PrepBox layout: a box owns the workgroups [blk0, blk0 + nblk))
__global__ __launch_bounds__(256) void k_objective_batch(const PrepBox* __restrict__ boxes, int nbox)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].blk0 <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const PrepBox& d = boxes[lo];
    const long t = (long)(blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
    objective_rows(d.s, d.mp.L[0], d.mp.L[1], d.mp.L[2], d.E0[0], d.E0[1], d.E0[2], d.b, t);
    if (d.bcat) objective_rows(d.s, d.mp.L[0], d.mp.L[1], d.mp.L[2], d.E0[0], d.E0[1], d.E0[2], d.bcat, t);
}

__global__ void k_membership(int N, const unsigned* __restrict__ tag, const unsigned* __restrict__ body_of_tag,
                             unsigned* __restrict__ iso)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) iso[i] = body_of_tag[tag[i]];
}

__global__ void k_zero_seg(int B, unsigned* counts, double* sums)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    if (counts) counts[b] = 0u;
    if (sums) sums[b] = 0.0;
}

__global__ void k_count_beads(int N, const unsigned* __restrict__ iso, unsigned* __restrict__ counts)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) atomicAdd(&counts[iso[i]], 1u);
}

// sums[b] += v_i - (w ? w_i : 0) over the beads of b
__global__ void k_seg_sum(int N, const unsigned* __restrict__ iso, const double* __restrict__ v,
                          const double* __restrict__ w, double* __restrict__ sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) unsafeAtomicAdd(&sums[iso[i]], w ? v[i] - w[i] : v[i]);
}

__global__ void k_project(int N, const unsigned* __restrict__ iso, const double* __restrict__ Q,
                          const unsigned* __restrict__ counts, const double* __restrict__ sums, double* __restrict__ q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const unsigned b = iso[i];
        q[i] += (Q[b] - sums[b]) / (double)counts[b];
    }
}

__global__ void k_seg_mean(int B, const unsigned* __restrict__ counts, const double* __restrict__ sums,
                           double* __restrict__ out)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) out[b] = counts[b] ? sums[b] / (double)counts[b] : 0.0;
}

__global__ void k_zero_force(int N, double* f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 4 * N) f[i] = 0.0;
}

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

}  // namespace

template <typename T>
hipError_t launch_real_space(const RealParams& rp, const SysView& s, const double* x, const T* phi_recip,
                             const T* table, double* y, int precond, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_space<T>, dim3(blocks(s.N, 256)), dim3(256), 0, st, rp, s, x, phi_recip, table, y,
                       precond);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_real_csr_build(const RealParams& rp, const SysView& s, const T* table, long ldn, unsigned* col, T* val,
                                 unsigned* row_len, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_csr_build<T>, dim3(blocks(s.N, 256)), dim3(256), 0, st, rp, s, table, ldn, col, val,
                       row_len);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_real_space_csr(const RealParams& rp, const SysView& s, const double* x, const T* qt,
                                 const T* phi_recip, long ldn, const unsigned* col, const T* val,
                                 const unsigned* row_len, double* y, int precond, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_space_csr<T>, dim3(blocks(s.N, 256)), dim3(256), 0, st, rp, s, x, qt, phi_recip, ldn,
                       col, val, row_len, y, precond);
    return hipGetLastError();
}

hipError_t launch_add_body_potentials(const SysView& s, const double* Phi, double* y, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_body_potentials, dim3(blocks(s.N, 256)), dim3(256), 0, st, s, Phi, y);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_real_space_force(const RealParams& rp, const SysView& s, const double* q, const T* gtable,
                                   const double E0[3], double* force, hipStream_t st, bool accum)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_space_force<T>, dim3(blocks(s.N, 256)), dim3(256), 0, st, rp, s, q, gtable, E0[0],
                       E0[1], E0[2], force, accum ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_constraint_rows(const SysView& s, const double* q, double* yb, int precond, hipStream_t st)
{
    if (s.N_rigid > 0) {
        const int W = s.Nmax < 64 ? s.Nmax : 64, K = 64 / W;
        hipLaunchKernelGGL(k_sum_charges, dim3((unsigned)((s.N_rigid + 4 * K - 1) / (4 * K))), dim3(256), 0, st, s, W,
                           q, yb, precond);
    }
    // beads outside rigid bodies (one isolated body each) carry their own charge row
    if (s.N > 0 && s.N_isolated > s.N_rigid)
        hipLaunchKernelGGL(k_set_charge, dim3(blocks(s.N, 256)), dim3(256), 0, st, s, q, yb, precond);
    return hipGetLastError();
}

hipError_t launch_precondition(const SysView& s, const double* x, double* y, hipStream_t st)
{
    const long n = (long)s.N + s.N_isolated;
    if (n > 0) hipLaunchKernelGGL(k_precond_diag, dim3(blocks(n, 256)), dim3(256), 0, st, s.N, s.N_isolated, x, y);
    if (s.N_rigid > 0)
        hipLaunchKernelGGL(k_precond_rigid, dim3(blocks(s.N_rigid, 256)), dim3(256), 0, st, s, x, y);
    return hipGetLastError();
}

hipError_t launch_objective(const SysView& s, const double L[3], const double E0[3], double* b, hipStream_t st)
{
    long n = (long)s.N_rigid * s.Nmax;
    if (s.N > n) n = s.N;
    if (s.N_isolated > n) n = s.N_isolated;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_objective, dim3(blocks(n, 256)), dim3(256), 0, st, s, L[0], L[1], L[2], E0[0], E0[1], E0[2],
                       b);
    return hipGetLastError();
}

hipError_t launch_objective_batch(const PrepBox* boxes, int nbox, int total_blocks, hipStream_t st)
{
    if (nbox <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_objective_batch, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    return hipGetLastError();
}

hipError_t launch_membership(int N, const unsigned* tag, const unsigned* body_of_tag, unsigned* iso, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_membership, dim3(blocks(N, 256)), dim3(256), 0, st, N, tag, body_of_tag, iso);
    return hipGetLastError();
}

hipError_t launch_body_counts(const SysView& s, unsigned* counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_seg, dim3(blocks(s.N_isolated, 256)), dim3(256), 0, st, s.N_isolated, counts,
                       (double*)nullptr);
    if (s.N > 0) hipLaunchKernelGGL(k_count_beads, dim3(blocks(s.N, 256)), dim3(256), 0, st, s.N, s.iso, counts);
    return hipGetLastError();
}

hipError_t launch_project_charges(const SysView& s, const unsigned* counts, double* sums, double* q, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_seg, dim3(blocks(s.N_isolated, 256)), dim3(256), 0, st, s.N_isolated, (unsigned*)nullptr,
                       sums);
    hipLaunchKernelGGL(k_seg_sum, dim3(blocks(s.N, 256)), dim3(256), 0, st, s.N, s.iso, (const double*)q,
                       (const double*)nullptr, sums);
    hipLaunchKernelGGL(k_project, dim3(blocks(s.N, 256)), dim3(256), 0, st, s.N, s.iso, s.body_charge, counts,
                       (const double*)sums, q);
    return hipGetLastError();
}

hipError_t launch_fit_potentials(const SysView& s, const unsigned* counts, const double* b, const double* y,
                                 double* sums, double* phi, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_seg, dim3(blocks(s.N_isolated, 256)), dim3(256), 0, st, s.N_isolated, (unsigned*)nullptr,
                       sums);
    if (s.N > 0)
        hipLaunchKernelGGL(k_seg_sum, dim3(blocks(s.N, 256)), dim3(256), 0, st, s.N, s.iso, b, y, sums);
    hipLaunchKernelGGL(k_seg_mean, dim3(blocks(s.N_isolated, 256)), dim3(256), 0, st, s.N_isolated, counts,
                       (const double*)sums, phi);
    return hipGetLastError();
}

hipError_t launch_zero_force(int N, double* force, hipStream_t st)
{
    if (N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_force, dim3(blocks(4L * N, 256)), dim3(256), 0, st, N, force);
    return hipGetLastError();
}

template hipError_t launch_real_space<float>(const RealParams&, const SysView&, const double*, const float*,
                                             const float*, double*, int, hipStream_t);
template hipError_t launch_real_space<double>(const RealParams&, const SysView&, const double*, const double*,
                                              const double*, double*, int, hipStream_t);
template hipError_t launch_real_csr_build<float>(const RealParams&, const SysView&, const float*, long, unsigned*,
                                                 float*, unsigned*, hipStream_t);
template hipError_t launch_real_csr_build<double>(const RealParams&, const SysView&, const double*, long, unsigned*,
                                                  double*, unsigned*, hipStream_t);
template hipError_t launch_real_space_csr<float>(const RealParams&, const SysView&, const double*, const float*,
                                                 const float*, long, const unsigned*, const float*, const unsigned*,
                                                 double*, int, hipStream_t);
template hipError_t launch_real_space_csr<double>(const RealParams&, const SysView&, const double*, const double*,
                                                  const double*, long, const unsigned*, const double*,
                                                  const unsigned*, double*, int, hipStream_t);
template <typename T>
hipError_t launch_real_csr_build_batch(const PairBox* boxes, int nbox, int total_blocks, hipStream_t st)
{
    if (nbox <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_csr_build_batch<T>, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    return hipGetLastError();
}
template hipError_t launch_real_csr_build_batch<float>(const PairBox*, int, int, hipStream_t);
template hipError_t launch_real_csr_build_batch<double>(const PairBox*, int, int, hipStream_t);

template <typename T>
hipError_t launch_real_space_force_batch(const PairBox* boxes, int nbox, int total_blocks, hipStream_t st)
{
    if (nbox <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_real_space_force_batch<T>, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    return hipGetLastError();
}
template hipError_t launch_real_space_force_batch<float>(const PairBox*, int, int, hipStream_t);
template hipError_t launch_real_space_force_batch<double>(const PairBox*, int, int, hipStream_t);

hipError_t launch_max_u32(const unsigned* a, int n, unsigned* out, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_max_u32, dim3(std::min(blocks(n, 256), 256u)), dim3(256), 0, st, a, n, out);
    return hipGetLastError();
}
template hipError_t launch_real_space_force<float>(const RealParams&, const SysView&, const double*, const float*,
                                                   const double*, double*, hipStream_t, bool);
template hipError_t launch_real_space_force<double>(const RealParams&, const SysView&, const double*, const double*,
                                                    const double*, double*, hipStream_t, bool);

}  // namespace cr
