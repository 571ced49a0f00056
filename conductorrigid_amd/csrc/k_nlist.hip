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

#include <algorithm>
#include <mutex>

namespace cr {

// Scratch of the allocation-free build (cr_build_nlist_noalloc): file-scope device memory
// that every GPU the library runs on holds until the process ends.  Host code never uses
// this name as an address: the address on each device comes from hipGetSymbolAddress
// (nlist_scratch_setup) and is kept per device.
constexpr size_t kNlistScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kNlistScratchBytes];

namespace {

constexpr int kScanBlock = 1024;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void k_fill_u32(unsigned* __restrict__ p, long n, unsigned v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

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

namespace {

// Cell-list working arrays of one build (device).
struct NlistWork {
    unsigned *cell, *members, *count, *start, *cursor;
    unsigned long long* sums;
};

struct NlistPlan {
    CellGrid cg;
    long ncell, nsum;
};

NlistPlan nlist_plan(int N, const double L[3], double rlist)
{
    NlistPlan p;
    p.cg.brute = 0;
    p.ncell = 1;
    for (int d = 0; d < 3; ++d) {
        p.cg.L[d] = L[d];
        int n = (int)floor(L[d] / rlist);
        if (n < 1) n = 1;
        if (n > 1024) n = 1024;
        p.cg.nc[d] = n;
        if (n < 3) p.cg.brute = 1;
        p.ncell *= n;
    }
    p.cg.r2 = rlist * rlist;
    if (p.cg.brute) {
        p.cg.nc[0] = p.cg.nc[1] = p.cg.nc[2] = 1;
        p.ncell = 1;
    }
    p.nsum = ((p.ncell > N ? p.ncell : N) + kScanBlock - 1) / kScanBlock + 2;
    return p;
}

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of the working arrays, each 256-byte aligned, in the order nlist_carve lays them out.
size_t nlist_work_bytes(int N, const NlistPlan& p)
{
    return 2 * align256(sizeof(unsigned) * (size_t)N) + 3 * align256(sizeof(unsigned) * (size_t)p.ncell) +
           align256(sizeof(unsigned long long) * (size_t)p.nsum);
}

NlistWork nlist_carve(void* base, int N, const NlistPlan& p)
{
    char* b = static_cast<char*>(base);
    NlistWork w;
    w.cell = reinterpret_cast<unsigned*>(b);
    b += align256(sizeof(unsigned) * (size_t)N);
    w.members = reinterpret_cast<unsigned*>(b);
    b += align256(sizeof(unsigned) * (size_t)N);
    w.count = reinterpret_cast<unsigned*>(b);
    b += align256(sizeof(unsigned) * (size_t)p.ncell);
    w.start = reinterpret_cast<unsigned*>(b);
    b += align256(sizeof(unsigned) * (size_t)p.ncell);
    w.cursor = reinterpret_cast<unsigned*>(b);
    b += align256(sizeof(unsigned) * (size_t)p.ncell);
    w.sums = reinterpret_cast<unsigned long long*>(b);
    return w;
}

// The build on given working arrays; counters are cleared with a fill kernel.
hipError_t nlist_run(int N, const double4* pos, const NlistPlan& p, const NlistWork& w, unsigned* head, unsigned* nn,
                     unsigned* nlist, long capacity, long* total, hipStream_t st)
{
    const CellGrid& cg = p.cg;
    const long ncell = p.ncell;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(blocks(ncell, 256), 1024u)), dim3(256), 0, st, w.count, ncell, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(blocks(ncell, 256), 1024u)), dim3(256), 0, st, w.cursor, ncell, 0u);
    hipLaunchKernelGGL(k_cell_of, dim3(blocks(N, 256)), dim3(256), 0, st, N, pos, cg, w.cell, w.count);
    hipError_t e = exclusive_scan_u32(w.count, w.start, ncell, w.sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cell_fill, dim3(blocks(N, 256)), dim3(256), 0, st, N, w.cell, w.start, w.cursor, w.members);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks(ncell, 256)), dim3(256), 0, st, ncell, w.start, w.count, w.members);
    hipLaunchKernelGGL((k_pairs<false>), dim3(blocks(N, 128)), dim3(128), 0, st, N, pos, cg, w.cell, w.start, w.count,
                       w.members, nn, (const unsigned*)nullptr, (unsigned*)nullptr, 0L);
    e = exclusive_scan_u32(nn, head, N, w.sums, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_pairs<true>), dim3(blocks(N, 128)), dim3(128), 0, st, N, pos, cg, w.cell, w.start, w.count,
                       w.members, (unsigned*)nullptr, head, nlist, capacity);
    unsigned long long tot = 0;
    const long nb = (N + kScanBlock - 1) / kScanBlock;
    e = hipMemcpyAsync(&tot, w.sums + nb, sizeof(tot), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    *total = (long)tot;
    return hipGetLastError();
}

}  // namespace

hipError_t build_nlist(int N, const double4* pos, const double L[3], double rlist, unsigned* head, unsigned* nn,
                       unsigned* nlist, long capacity, long* total, void* /*scratch_pool*/, hipStream_t st)
{
    *total = 0;
    if (N <= 0) return hipSuccess;
    const NlistPlan p = nlist_plan(N, L, rlist);
    void* base = nullptr;
    hipError_t e = hipMallocAsync(&base, nlist_work_bytes(N, p), st);
    if (e != hipSuccess) return e;
    e = nlist_run(N, pos, p, nlist_carve(base, N, p), head, nn, nlist, capacity, total, st);
    const hipError_t f = hipFreeAsync(base, st);
    return e != hipSuccess ? e : f;
}

hipError_t nlist_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the device must be current
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

size_t nlist_noalloc_bytes(int N, const double L[3], double rlist)
{
    return N > 0 ? nlist_work_bytes(N, nlist_plan(N, L, rlist)) : 0;
}

size_t nlist_noalloc_capacity() { return kNlistScratchBytes; }

hipError_t build_nlist_noalloc(int N, const double4* pos, const double L[3], double rlist, unsigned* head,
                               unsigned* nn, unsigned* nlist, long capacity, long* total, hipStream_t st)
{
    *total = 0;
    if (N <= 0) return hipSuccess;
    const NlistPlan p = nlist_plan(N, L, rlist);
    if (nlist_work_bytes(N, p) > kNlistScratchBytes) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = nlist_scratch_setup(dev);
    if (e != hipSuccess) return e;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[dev];
    }
    return nlist_run(N, pos, p, nlist_carve(base, N, p), head, nn, nlist, capacity, total, st);
}

}  // namespace cr
