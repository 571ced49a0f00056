// k_nlist.hip -- full neighbour list in HOOMD layout (head_list, n_neigh, nlist).
//
// Stands in for the HOOMD CellListGPU + NeighborListGPUBinned the reference's Python
// layer creates (pymodule/compute.py:38-44): every pair with minimum-image distance
// below rlist = rc + skin, both directions, exclusions cleared (intra-body pairs
// kept).  Consumed by the real-space kernels exactly as the reference consumes
// HOOMD's list (cppmodule/ConductorRigid.cu:509-525).
//
// Cells of edge >= rlist, counting sort of beads into cells (members of a cell in
// increasing index order so the list is deterministic), one count pass and one
// fill pass over the 27 neighbouring cells, exclusive scans for the offsets.
#include "cr_kernels.h"

namespace cr {

namespace {

constexpr int kScanBlock = 1024;

__device__ __forceinline__ double min_image(double d, double L) { return d - L * rint(d / L); }

struct CellGrid {
    int nc[3];
    double L[3];
    double r2;
    int brute;   // fewer than 3 cells along some axis: all pairs
};

__device__ __forceinline__ int cell_coord(double x, double L, int nc)
{
    double f = (x + 0.5 * L) / L;
    int k = (int)floor(f * nc);
    k %= nc;
    if (k < 0) k += nc;
    return k;
}

__global__ void k_cell_of(int N, const double4* __restrict__ pos, CellGrid cg, unsigned* __restrict__ cell,
                          unsigned* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double4 p = pos[i];
    const int cx = cell_coord(p.x, cg.L[0], cg.nc[0]);
    const int cy = cell_coord(p.y, cg.L[1], cg.nc[1]);
    const int cz = cell_coord(p.z, cg.L[2], cg.nc[2]);
    const unsigned c = (unsigned)((cx * cg.nc[1] + cy) * cg.nc[2] + cz);
    cell[i] = c;
    atomicAdd(&count[c], 1u);
}

__global__ void k_cell_fill(int N, const unsigned* __restrict__ cell, const unsigned* __restrict__ start,
                            unsigned* __restrict__ cursor, unsigned* __restrict__ members)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const unsigned c = cell[i];
    const unsigned slot = atomicAdd(&cursor[c], 1u);
    members[start[c] + slot] = (unsigned)i;
}

// insertion sort of each cell's members (cells hold tens of beads)
__global__ void k_cell_sort(long ncell, const unsigned* __restrict__ start, const unsigned* __restrict__ count,
                            unsigned* __restrict__ members)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    unsigned* m = members + start[c];
    const unsigned n = count[c];
    for (unsigned a = 1; a < n; ++a) {
        unsigned v = m[a];
        unsigned b = a;
        while (b > 0 && m[b - 1] > v) {
            m[b] = m[b - 1];
            --b;
        }
        m[b] = v;
    }
}

template <bool FILL>
__global__ void k_pairs(int N, const double4* __restrict__ pos, CellGrid cg, const unsigned* __restrict__ cell,
                        const unsigned* __restrict__ start, const unsigned* __restrict__ count,
                        const unsigned* __restrict__ members, unsigned* __restrict__ n_neigh,
                        const unsigned* __restrict__ head, unsigned* __restrict__ nlist, long capacity)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double4 p = pos[i];
    unsigned cnt = 0;
    const long base = FILL ? (long)head[i] : 0;
    if (cg.brute) {
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            const double4 q = pos[j];
            const double dx = min_image(p.x - q.x, cg.L[0]), dy = min_image(p.y - q.y, cg.L[1]),
                         dz = min_image(p.z - q.z, cg.L[2]);
            if (dx * dx + dy * dy + dz * dz < cg.r2) {
                if (FILL && base + cnt < capacity) nlist[base + cnt] = (unsigned)j;
                ++cnt;
            }
        }
    } else {
        const unsigned c = cell[i];
        const int cz = c % cg.nc[2], cy = (c / cg.nc[2]) % cg.nc[1], cx = c / (cg.nc[2] * cg.nc[1]);
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int d = -1; d <= 1; ++d) {
                    const int x = (cx + a + cg.nc[0]) % cg.nc[0];
                    const int y = (cy + b + cg.nc[1]) % cg.nc[1];
                    const int z = (cz + d + cg.nc[2]) % cg.nc[2];
                    const unsigned nb = (unsigned)((x * cg.nc[1] + y) * cg.nc[2] + z);
                    const unsigned s0 = start[nb], s1 = s0 + count[nb];
                    for (unsigned m = s0; m < s1; ++m) {
                        const unsigned j = members[m];
                        if ((int)j == i) continue;
                        const double4 q = pos[j];
                        const double dx = min_image(p.x - q.x, cg.L[0]), dy = min_image(p.y - q.y, cg.L[1]),
                                     dz = min_image(p.z - q.z, cg.L[2]);
                        if (dx * dx + dy * dy + dz * dz < cg.r2) {
                            if (FILL && base + cnt < capacity) nlist[base + cnt] = j;
                            ++cnt;
                        }
                    }
                }
    }
    if (!FILL) n_neigh[i] = cnt;
}

// Block-level exclusive scan of 32-bit counts into 64-bit-safe offsets (uint32 output,
// the list is indexed with 32-bit offsets as in HOOMD).
__global__ __launch_bounds__(kScanBlock) void k_scan_block(const unsigned* __restrict__ in, unsigned* __restrict__ out,
                                                           long n, unsigned long long* __restrict__ block_sums)
{
    __shared__ unsigned long long buf[kScanBlock];
    const long i = (long)blockIdx.x * kScanBlock + threadIdx.x;
    const unsigned long long v = (i < n) ? in[i] : 0ull;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        unsigned long long t = (threadIdx.x >= (unsigned)off) ? buf[threadIdx.x - off] : 0ull;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = (unsigned)(buf[threadIdx.x] - v);
    if (threadIdx.x == kScanBlock - 1) block_sums[blockIdx.x] = buf[threadIdx.x];
}

// Serial exclusive scan of the block sums (few thousand entries at most), total at [nb].
__global__ void k_scan_sums(unsigned long long* sums, long nb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long acc = 0;
    for (long b = 0; b < nb; ++b) {
        unsigned long long v = sums[b];
        sums[b] = acc;
        acc += v;
    }
    sums[nb] = acc;
}

__global__ void k_scan_add(unsigned* __restrict__ out, long n, const unsigned long long* __restrict__ sums)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += (unsigned)sums[i / kScanBlock];
}

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }

}  // namespace

// out = exclusive scan of in; total (device) lands in sums[nblocks].  sums: >= n/kScanBlock + 2 entries.
hipError_t exclusive_scan_u32(const unsigned* in, unsigned* out, long n, unsigned long long* sums, hipStream_t st)
{
    const long nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb > 0) hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(kScanBlock), 0, st, in, out, n, sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, st, sums, nb);
    if (nb > 0) hipLaunchKernelGGL(k_scan_add, dim3(blocks(n, 256)), dim3(256), 0, st, out, n, sums);
    return hipGetLastError();
}

hipError_t build_nlist(int N, const double4* pos, const double L[3], double rlist, unsigned* head, unsigned* nn,
                       unsigned* nlist, long capacity, long* total, void* /*scratch_pool*/, hipStream_t st)
{
    *total = 0;
    if (N <= 0) return hipSuccess;
    CellGrid cg;
    cg.brute = 0;
    long ncell = 1;
    for (int d = 0; d < 3; ++d) {
        cg.L[d] = L[d];
        int n = (int)floor(L[d] / rlist);
        if (n < 1) n = 1;
        if (n > 1024) n = 1024;
        cg.nc[d] = n;
        if (n < 3) cg.brute = 1;
        ncell *= n;
    }
    cg.r2 = rlist * rlist;
    if (cg.brute) {
        cg.nc[0] = cg.nc[1] = cg.nc[2] = 1;
        ncell = 1;
    }
    unsigned *cell = nullptr, *count = nullptr, *start = nullptr, *cursor = nullptr, *members = nullptr;
    unsigned long long* sums = nullptr;
    const long nsum = ((ncell > N ? ncell : N) + kScanBlock - 1) / kScanBlock + 2;
    hipError_t e = hipSuccess;
#define NL_TRY(x)                 \
    do {                          \
        e = (x);                  \
        if (e != hipSuccess) goto done; \
    } while (0)
    NL_TRY(hipMallocAsync((void**)&cell, sizeof(unsigned) * N, st));
    NL_TRY(hipMallocAsync((void**)&members, sizeof(unsigned) * N, st));
    NL_TRY(hipMallocAsync((void**)&count, sizeof(unsigned) * ncell, st));
    NL_TRY(hipMallocAsync((void**)&start, sizeof(unsigned) * ncell, st));
    NL_TRY(hipMallocAsync((void**)&cursor, sizeof(unsigned) * ncell, st));
    NL_TRY(hipMallocAsync((void**)&sums, sizeof(unsigned long long) * nsum, st));
    NL_TRY(hipMemsetAsync(count, 0, sizeof(unsigned) * ncell, st));
    NL_TRY(hipMemsetAsync(cursor, 0, sizeof(unsigned) * ncell, st));
    hipLaunchKernelGGL(k_cell_of, dim3(blocks(N, 256)), dim3(256), 0, st, N, pos, cg, cell, count);
    NL_TRY(exclusive_scan_u32(count, start, ncell, sums, st));
    hipLaunchKernelGGL(k_cell_fill, dim3(blocks(N, 256)), dim3(256), 0, st, N, cell, start, cursor, members);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks(ncell, 256)), dim3(256), 0, st, ncell, start, count, members);
    hipLaunchKernelGGL((k_pairs<false>), dim3(blocks(N, 128)), dim3(128), 0, st, N, pos, cg, cell, start, count,
                       members, nn, (const unsigned*)nullptr, (unsigned*)nullptr, 0L);
    NL_TRY(exclusive_scan_u32(nn, head, N, sums, st));
    hipLaunchKernelGGL((k_pairs<true>), dim3(blocks(N, 128)), dim3(128), 0, st, N, pos, cg, cell, start, count,
                       members, (unsigned*)nullptr, head, nlist, capacity);
    {
        unsigned long long tot = 0;
        const long nb = (N + kScanBlock - 1) / kScanBlock;
        NL_TRY(hipMemcpyAsync(&tot, sums + nb, sizeof(tot), hipMemcpyDeviceToHost, st));
        NL_TRY(hipStreamSynchronize(st));
        *total = (long)tot;
    }
done:
    if (cell) (void)hipFreeAsync(cell, st);
    if (members) (void)hipFreeAsync(members, st);
    if (count) (void)hipFreeAsync(count, st);
    if (start) (void)hipFreeAsync(start, st);
    if (cursor) (void)hipFreeAsync(cursor, st);
    if (sums) (void)hipFreeAsync(sums, st);
#undef NL_TRY
    if (e == hipSuccess) e = hipGetLastError();
    return e;
}

}  // namespace cr
