// k_tiled.hip -- binned spreading and contraction on mesh tiles (spectral Ewald).
//
// Reference: spread (cppmodule/ConductorRigid.cu:144-234), contract (:260-360) and
// contract_force (:363-466), which launch one P x P x P thread block per bead and
// spread with one global float atomic per window node.
//
// Here beads are binned by the mesh tile holding the first node of their window.
// Tile edges divide the mesh and are >= P-1 nodes, so a window stays inside the tile
// extended by P-1 nodes on the + side of each axis.  One workgroup owns one tile and
// works through its bin in chunks: the chunk's bead records are loaded together, the
// 3P separable axis weights of every bead are evaluated by all threads into an LDS
// table, and then each wavefront takes one bead at a time, lanes over the (y, z)
// window plane, walking x -- every operand of the inner loop comes from LDS.
//   spread:   window products are added into an LDS copy of the extended tile; the
//             interior is stored straight to the mesh (each node has one owner), the
//             extension to a per-tile scratch block added in by k_spread_halo_add.
//   contract: the extended tile is read into LDS once; each wavefront reduces its
//             bead's window across lanes.
// No global atomics and no mesh clear.
#include "se_device.h"

namespace cr {

namespace {

using namespace se;

__device__ __forceinline__ int tile_of_bead(const MeshParams& mp, const TileGeom& tg, const double4& r)
{
    const int sx = wrap_idx(window_start(r.x, mp.L[0], mp.Nx, mp.P), mp.Nx);
    const int sy = wrap_idx(window_start(r.y, mp.L[1], mp.Ny, mp.P), mp.Ny);
    const int sz = wrap_idx(window_start(r.z, mp.L[2], mp.Nz, mp.P), mp.Nz);
    return ((sx / tg.T[0]) * tg.nt[1] + sy / tg.T[1]) * tg.nt[2] + sz / tg.T[2];
}

__global__ void k_bin_count(MeshParams mp, TileGeom tg, int N, const double4* __restrict__ pos,
                            unsigned* __restrict__ bead_tile, unsigned* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const unsigned t = (unsigned)tile_of_bead(mp, tg, pos[i]);
    bead_tile[i] = t;
    atomicAdd(&count[t], 1u);
}

__global__ void k_bin_fill(int N, const unsigned* __restrict__ bead_tile, const unsigned* __restrict__ start,
                           unsigned* __restrict__ cursor, unsigned* __restrict__ sorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const unsigned t = bead_tile[i];
    sorted[start[t] + atomicAdd(&cursor[t], 1u)] = (unsigned)i;
}

__global__ void k_zero_u32(unsigned* p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Workgroups are dealt round-robin to the 8 XCDs (block b on XCD b % 8): give each XCD
// a contiguous run of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles)
{
    if (ntiles % 8 != 0) return b;
    return (b % 8) * (ntiles / 8) + b / 8;
}

struct TileCoords {
    int t, x0, y0, z0, Ex, Ey, Ez;
};

__device__ __forceinline__ TileCoords tile_coords(const MeshParams& mp, const TileGeom& tg)
{
    TileCoords c;
    c.t = xcd_tile(blockIdx.x, tg.ntiles);
    const int tz = c.t % tg.nt[2], ty = (c.t / tg.nt[2]) % tg.nt[1], tx = c.t / (tg.nt[2] * tg.nt[1]);
    c.x0 = tx * tg.T[0];
    c.y0 = ty * tg.T[1];
    c.z0 = tz * tg.T[2];
    c.Ex = tg.T[0] + mp.P - 1;
    c.Ey = tg.T[1] + mp.P - 1;
    c.Ez = tg.T[2] + mp.P - 1;
    return c;
}

// LDS: [ext Ex*Ey*Ez T][w ch*3P T][u ch*3 T][qs ch T][off ch int][ids ch uint]
template <typename T>
struct Stage {
    T* ext;
    T* w;
    T* u;
    T* qs;
    int* off;
    unsigned* ids;
};

template <typename T>
__device__ __forceinline__ Stage<T> stage_layout(char* smem, int esz, int ch, int P)
{
    Stage<T> s;
    s.ext = reinterpret_cast<T*>(smem);
    s.w = s.ext + esz;
    s.u = s.w + ch * 3 * P;
    s.qs = s.u + ch * 3;
    s.off = reinterpret_cast<int*>(s.qs + ch);
    s.ids = reinterpret_cast<unsigned*>(s.off + ch);
    return s;
}

// Load beads [k0, k0+cnt) of the bin and fill the chunk's LDS tables.  u holds each
// bead's position relative to its window start in node units (node s+p minus bead =
// (p - u) h), so the axis weight is exp(-2 xi^2 h^2 (p - u)^2 / eta).
template <typename T>
__device__ __forceinline__ void stage_chunk(const MeshParams& mp, const TileCoords& tc, const unsigned* __restrict__ sorted,
                                            unsigned k0, int cnt, const double4* __restrict__ pos,
                                            const double* __restrict__ q, T qscale, Stage<T>& s)
{
    const int P = mp.P;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const unsigned bead = sorted[k0 + i];
        const double4 r = pos[bead];
        const int sx = window_start(r.x, mp.L[0], mp.Nx, P);
        const int sy = window_start(r.y, mp.L[1], mp.Ny, P);
        const int sz = window_start(r.z, mp.L[2], mp.Nz, P);
        const int ox = wrap_idx(sx, mp.Nx) - tc.x0, oy = wrap_idx(sy, mp.Ny) - tc.y0, oz = wrap_idx(sz, mp.Nz) - tc.z0;
        s.off[i] = ox | (oy << 10) | (oz << 20);
        s.ids[i] = bead;
        s.u[3 * i + 0] = (T)((r.x + 0.5 * mp.L[0]) / mp.h[0] - sx);
        s.u[3 * i + 1] = (T)((r.y + 0.5 * mp.L[1]) / mp.h[1] - sy);
        s.u[3 * i + 2] = (T)((r.z + 0.5 * mp.L[2]) / mp.h[2] - sz);
        s.qs[i] = q ? qscale * (T)q[bead] : (T)0;
    }
    __syncthreads();
    const int per = 3 * P;
    const T c0 = (T)(mp.xiterm * mp.h[0] * mp.h[0] / mp.eta[0]);
    const T c1 = (T)(mp.xiterm * mp.h[1] * mp.h[1] / mp.eta[1]);
    const T c2 = (T)(mp.xiterm * mp.h[2] * mp.h[2] / mp.eta[2]);
    for (int j = threadIdx.x; j < cnt * per; j += blockDim.x) {
        const int i = j / per, rem = j - i * per;
        const int d = rem / P, p = rem - d * P;
        const T dd = (T)p - s.u[3 * i + d];
        const T cd = d == 0 ? c0 : (d == 1 ? c1 : c2);
        s.w[j] = gexp<T>(-cd * dd * dd);
    }
    __syncthreads();
}

// PT > 0: window support fixed at compile time (loops unroll, P divisions become
// constants, the P node updates of a lane issue back to back); PT == 0: runtime P <= 16.
template <typename T, int PT>
__global__ __launch_bounds__(256) void k_spread_tiled(MeshParams mp, TileGeom tg, int ch,
                                                      const unsigned* __restrict__ tstart,
                                                      const unsigned* __restrict__ sorted,
                                                      const double4* __restrict__ pos, const double* __restrict__ q,
                                                      T* __restrict__ grid, T* __restrict__ halo)
{
    extern __shared__ char smem_raw[];
    const TileCoords tc = tile_coords(mp, tg);
    constexpr int PM = PT > 0 ? PT : 16;
    const int P = PT > 0 ? PT : mp.P;
    const int esz = tc.Ex * tc.Ey * tc.Ez;
    Stage<T> s = stage_layout<T>(smem_raw, esz, ch, P);
    for (int k = threadIdx.x; k < esz; k += blockDim.x) s.ext[k] = (T)0;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int xs = tc.Ey * tc.Ez, ys = tc.Ez, per = 3 * P;
    const unsigned kb = tstart[tc.t], ke = tstart[tc.t + 1];
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        stage_chunk<T>(mp, tc, sorted, k0, cnt, pos, q, (T)mp.prefac, s);
        for (int i = wave; i < cnt; i += nwave) {
            const int o = s.off[i];
            const int ox = o & 1023, oy = (o >> 10) & 1023, oz = o >> 20;
            const T* wi = s.w + i * per;
            const T qs = s.qs[i];
            T wx[PM];
#pragma unroll
            for (int px = 0; px < PM; ++px) wx[px] = (px < P) ? wi[px] : (T)0;
            for (int pair = lane; pair < P * P; pair += kWave) {
                const int py = pair / P, pz = pair - py * P;
                const T c = qs * wi[P + py] * wi[2 * P + pz];
                T* col = s.ext + ox * xs + (oy + py) * ys + (oz + pz);
#pragma unroll
                for (int px = 0; px < PM; ++px)
                    if (px < P) atomic_add(col + px * xs, c * wx[px]);
            }
        }
        __syncthreads();
    }
    const long plane = (long)mp.Ny * mp.Nz;
    T* hb = halo + (long)tc.t * esz;
    const int Tx = tg.T[0], Ty = tg.T[1], Tz = tg.T[2];
    for (int k = threadIdx.x; k < esz; k += blockDim.x) {
        const int lz = k % tc.Ez, ly = (k / tc.Ez) % tc.Ey, lx = k / (tc.Ez * tc.Ey);
        if (lx < Tx && ly < Ty && lz < Tz)
            grid[(long)(tc.x0 + lx) * plane + (long)(tc.y0 + ly) * mp.Nz + (tc.z0 + lz)] = s.ext[k];
        else
            hb[k] = s.ext[k];
    }
}

// Adds the extensions of the (up to 7) tiles below each node's tile into the mesh.
template <typename T>
__global__ __launch_bounds__(256) void k_spread_halo_add(MeshParams mp, TileGeom tg, const T* __restrict__ halo,
                                                         T* __restrict__ grid)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long G = (long)mp.Nx * mp.Ny * mp.Nz;
    if (g >= G) return;
    const int z = (int)(g % mp.Nz), y = (int)((g / mp.Nz) % mp.Ny), x = (int)(g / ((long)mp.Nz * mp.Ny));
    const int P = mp.P;
    const int lx = x % tg.T[0], ly = y % tg.T[1], lz = z % tg.T[2];
    if (lx >= P - 1 && ly >= P - 1 && lz >= P - 1) return;
    const int tx = x / tg.T[0], ty = y / tg.T[1], tz = z / tg.T[2];
    const int Ey = tg.T[1] + P - 1, Ez = tg.T[2] + P - 1;
    const long esz = (long)(tg.T[0] + P - 1) * Ey * Ez;
    T acc = 0;
    for (int a = 0; a < 2; ++a) {
        if (a && lx >= P - 1) continue;
        for (int b = 0; b < 2; ++b) {
            if (b && ly >= P - 1) continue;
            for (int c = 0; c < 2; ++c) {
                if ((c && lz >= P - 1) || (a + b + c == 0)) continue;
                const int sx = (tx - a + tg.nt[0]) % tg.nt[0];
                const int sy = (ty - b + tg.nt[1]) % tg.nt[1];
                const int sz = (tz - c + tg.nt[2]) % tg.nt[2];
                const long src = ((long)sx * tg.nt[1] + sy) * tg.nt[2] + sz;
                const int ex = lx + a * tg.T[0], ey = ly + b * tg.T[1], ez = lz + c * tg.T[2];
                acc += halo[src * esz + ((long)ex * Ey + ey) * Ez + ez];
            }
        }
    }
    grid[g] += acc;
}

template <typename T, bool FORCE, int PT>
__global__ __launch_bounds__(256) void k_contract_tiled(MeshParams mp, TileGeom tg, int ch,
                                                        const unsigned* __restrict__ tstart,
                                                        const unsigned* __restrict__ sorted,
                                                        const double4* __restrict__ pos, const double* __restrict__ q,
                                                        const T* __restrict__ grid, T* __restrict__ phi,
                                                        double* __restrict__ force)
{
    extern __shared__ char smem_raw[];
    const TileCoords tc = tile_coords(mp, tg);
    constexpr int PM = PT > 0 ? PT : 16;
    const int P = PT > 0 ? PT : mp.P;
    const int esz = tc.Ex * tc.Ey * tc.Ez;
    Stage<T> s = stage_layout<T>(smem_raw, esz, ch, P);
    const long plane = (long)mp.Ny * mp.Nz;
    for (int k = threadIdx.x; k < esz; k += blockDim.x) {
        const int lz = k % tc.Ez, ly = (k / tc.Ez) % tc.Ey, lx = k / (tc.Ez * tc.Ey);
        const int gx = (tc.x0 + lx) % mp.Nx, gy = (tc.y0 + ly) % mp.Ny, gz = (tc.z0 + lz) % mp.Nz;
        s.ext[k] = grid[(long)gx * plane + (long)gy * mp.Nz + gz];
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int xs = tc.Ey * tc.Ez, ys = tc.Ez, per = 3 * P;
    const T pc = (T)(mp.quadW * mp.prefac);
    const unsigned kb = tstart[tc.t], ke = tstart[tc.t + 1];
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        stage_chunk<T>(mp, tc, sorted, k0, cnt, pos, FORCE ? q : nullptr, (T)1, s);
        for (int i = wave; i < cnt; i += nwave) {
            const int o = s.off[i];
            const int ox = o & 1023, oy = (o >> 10) & 1023, oz = o >> 20;
            const T* wi = s.w + i * per;
            const T ux = s.u[3 * i], uy = s.u[3 * i + 1], uz = s.u[3 * i + 2];
            T wx[PM];
#pragma unroll
            for (int px = 0; px < PM; ++px) wx[px] = (px < P) ? wi[px] : (T)0;
            T acc = 0, fx = 0, fy = 0, fz = 0;
            for (int pair = lane; pair < P * P; pair += kWave) {
                const int py = pair / P, pz = pair - py * P;
                const T wyz = wi[P + py] * wi[2 * P + pz];
                const T* col = s.ext + ox * xs + (oy + py) * ys + (oz + pz);
                T g[PM];
#pragma unroll
                for (int px = 0; px < PM; ++px) g[px] = (px < P) ? col[px * xs] : (T)0;
                T sx = 0, sxr = 0;
#pragma unroll
                for (int px = 0; px < PM; ++px) {
                    const T wg = wx[px] * g[px];
                    sx += wg;
                    if (FORCE) sxr += wg * (ux - (T)px);
                }
                acc += wyz * sx;
                if (FORCE) {
                    fx += wyz * sxr;
                    fy += wyz * sx * (uy - (T)py);
                    fz += wyz * sx * (uz - (T)pz);
                }
            }
            const unsigned bead = s.ids[i];
            if (!FORCE) {
                acc = wave_sum(acc);
                if (lane == 0) phi[bead] = acc * pc;
            } else {
                fx = wave_sum(fx);
                fy = wave_sum(fy);
                fz = wave_sum(fz);
                if (lane == 0) {
                    // F = 2 (h^3 prefac) (2 xi^2) q_i sum phi_g e^{...} (r/eta), r = bead - node = (u - p) h
                    // (cppmodule/ConductorRigid.cu:449)
                    const double c = 2.0 * (mp.quadW * mp.prefac) * mp.xiterm * (double)s.qs[i];
                    force[4 * bead + 0] = c * mp.h[0] / mp.eta[0] * (double)fx;
                    force[4 * bead + 1] = c * mp.h[1] / mp.eta[1] * (double)fy;
                    force[4 * bead + 2] = c * mp.h[2] / mp.eta[2] * (double)fz;
                    force[4 * bead + 3] = 0.0;
                }
            }
        }
        __syncthreads();
    }
}

inline unsigned blocks(long n, int b)
{
    long v = (n + b - 1) / b;
    return (unsigned)(v > 0 ? v : 1);
}

}  // namespace

TiledLds tiled_lds(const MeshParams& mp, const TileGeom& tg, size_t elem)
{
    TiledLds l;
    l.ext_bytes = (size_t)(tg.T[0] + mp.P - 1) * (tg.T[1] + mp.P - 1) * (tg.T[2] + mp.P - 1) * elem;
    const size_t per_bead = (size_t)(3 * mp.P + 3 + 1) * elem + 2 * sizeof(int);
    long ch = l.ext_bytes < kTiledLdsBudget ? (long)((kTiledLdsBudget - l.ext_bytes) / per_bead) : 0;
    if (ch > 256) ch = 256;
    ch -= ch % 4;
    l.chunk = (int)ch;
    l.total_bytes = l.ext_bytes + (size_t)ch * per_bead;
    return l;
}

hipError_t launch_bin_beads(const MeshParams& mp, const TileGeom& tg, const SysView& s, unsigned* bead_tile,
                            unsigned* tile_count, unsigned* tile_start, unsigned* cursor, unsigned* sorted,
                            unsigned long long* scan_sums, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks(tg.ntiles + 1, 256)), dim3(256), 0, st, tile_count, (long)tg.ntiles + 1);
    hipLaunchKernelGGL(k_zero_u32, dim3(blocks(tg.ntiles, 256)), dim3(256), 0, st, cursor, (long)tg.ntiles);
    if (s.N > 0)
        hipLaunchKernelGGL(k_bin_count, dim3(blocks(s.N, 256)), dim3(256), 0, st, mp, tg, s.N, s.pos, bead_tile,
                           tile_count);
    // exclusive scan over ntiles+1 counts (the last is 0) gives tile_start[ntiles] = N
    hipError_t e = exclusive_scan_u32(tile_count, tile_start, (long)tg.ntiles + 1, scan_sums, st);
    if (e != hipSuccess) return e;
    if (s.N > 0)
        hipLaunchKernelGGL(k_bin_fill, dim3(blocks(s.N, 256)), dim3(256), 0, st, s.N, bead_tile, tile_start, cursor,
                           sorted);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_spread_tiled(const MeshParams& mp, const TileGeom& tg, const SysView& s, const unsigned* tile_start,
                               const unsigned* sorted, const double* q, T* grid, T* halo, hipStream_t st)
{
    const TiledLds l = tiled_lds(mp, tg, sizeof(T));
    if (l.chunk < 4 || mp.P > 16) return hipErrorInvalidValue;
#define CR_SPREAD_P(PV)                                                                                          \
    hipLaunchKernelGGL((k_spread_tiled<T, PV>), dim3(tg.ntiles), dim3(256), l.total_bytes, st, mp, tg, l.chunk,  \
                       tile_start, sorted, s.pos, q, grid, halo)
    switch (mp.P) {
        case 4: CR_SPREAD_P(4); break;
        case 5: CR_SPREAD_P(5); break;
        case 6: CR_SPREAD_P(6); break;
        case 7: CR_SPREAD_P(7); break;
        case 8: CR_SPREAD_P(8); break;
        case 9: CR_SPREAD_P(9); break;
        case 10: CR_SPREAD_P(10); break;
        default: CR_SPREAD_P(0); break;
    }
#undef CR_SPREAD_P
    const long G = (long)mp.Nx * mp.Ny * mp.Nz;
    hipLaunchKernelGGL(k_spread_halo_add<T>, dim3(blocks(G, 256)), dim3(256), 0, st, mp, tg, (const T*)halo, grid);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_contract_tiled(const MeshParams& mp, const TileGeom& tg, const SysView& s, const unsigned* tile_start,
                                 const unsigned* sorted, const double* q, const T* grid, T* phi, double* force,
                                 hipStream_t st)
{
    const TiledLds l = tiled_lds(mp, tg, sizeof(T));
    if (l.chunk < 4 || mp.P > 16) return hipErrorInvalidValue;
#define CR_CONTRACT_P(PV)                                                                                       \
    do {                                                                                                        \
        if (force)                                                                                              \
            hipLaunchKernelGGL((k_contract_tiled<T, true, PV>), dim3(tg.ntiles), dim3(256), l.total_bytes, st,  \
                               mp, tg, l.chunk, tile_start, sorted, s.pos, q, grid, (T*)nullptr, force);        \
        else                                                                                                    \
            hipLaunchKernelGGL((k_contract_tiled<T, false, PV>), dim3(tg.ntiles), dim3(256), l.total_bytes, st, \
                               mp, tg, l.chunk, tile_start, sorted, s.pos, (const double*)nullptr, grid, phi,   \
                               (double*)nullptr);                                                               \
    } while (0)
    switch (mp.P) {
        case 4: CR_CONTRACT_P(4); break;
        case 5: CR_CONTRACT_P(5); break;
        case 6: CR_CONTRACT_P(6); break;
        case 7: CR_CONTRACT_P(7); break;
        case 8: CR_CONTRACT_P(8); break;
        case 9: CR_CONTRACT_P(9); break;
        case 10: CR_CONTRACT_P(10); break;
        default: CR_CONTRACT_P(0); break;
    }
#undef CR_CONTRACT_P
    return hipGetLastError();
}

template hipError_t launch_spread_tiled<float>(const MeshParams&, const TileGeom&, const SysView&, const unsigned*,
                                               const unsigned*, const double*, float*, float*, hipStream_t);
template hipError_t launch_spread_tiled<double>(const MeshParams&, const TileGeom&, const SysView&, const unsigned*,
                                                const unsigned*, const double*, double*, double*, hipStream_t);
template hipError_t launch_contract_tiled<float>(const MeshParams&, const TileGeom&, const SysView&, const unsigned*,
                                                 const unsigned*, const double*, const float*, float*, double*,
                                                 hipStream_t);
template hipError_t launch_contract_tiled<double>(const MeshParams&, const TileGeom&, const SysView&, const unsigned*,
                                                  const unsigned*, const double*, const double*, double*, double*,
                                                  hipStream_t);

}  // namespace cr
