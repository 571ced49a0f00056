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
//   spread:   window products are added into an LDS copy of the extended tile, which
//             is then added into the (cleared) mesh with one global atomic per node:
//             each node receives the sums of the at most 8 tiles whose extension covers
//             it.  Measured against storing the interior and gathering the extensions
//             from per-tile scratch in a second pass, this moves fewer bytes.
//   contract: the extended tile is read into LDS once; each wavefront reduces its
//             bead's window across lanes.
// Per bead, all node updates stay in LDS.
#include "se_device.h"

#include <map>
#include <mutex>
#include <utility>

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

// ctr[key] += 1 for every active lane, with one atomic per distinct key of the wavefront
// (spatially ordered beads mostly share their tile); returns each lane's slot, the
// counter's old value plus the lane's rank among the lanes of its key.
__device__ __forceinline__ unsigned wave_count(unsigned* __restrict__ ctr, unsigned key, bool on)
{
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(on);
    unsigned slot = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const unsigned lk = __shfl(key, leader, 64);
        const unsigned long long same = __ballot(on && key == lk) & todo;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&ctr[lk], (unsigned)__popcll(same));
        base = __shfl(base, leader, 64);
        if (on && key == lk) slot = base + (unsigned)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return slot;
}

__global__ __launch_bounds__(256) void k_bin_count(MeshParams mp, TileGeom tg, int N, const double4* __restrict__ pos,
                                                   unsigned* __restrict__ bead_tile, unsigned* __restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < N;
    const unsigned t = on ? (unsigned)tile_of_bead(mp, tg, pos[i]) : 0u;
    if (on) bead_tile[i] = t;
    (void)wave_count(count, t, on);
}

__global__ __launch_bounds__(256) void k_bin_fill(int N, const unsigned* __restrict__ bead_tile,
                                                  const unsigned* __restrict__ start, unsigned* __restrict__ cursor,
                                                  unsigned* __restrict__ sorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = i < N;
    const unsigned t = on ? bead_tile[i] : 0u;
    const unsigned slot = wave_count(cursor, t, on);
    if (on) sorted[start[t] + slot] = (unsigned)i;
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

// LDS strides of the extended tile: the z stride is Ez rounded up to even, so that the
// 64 lanes of one window update (8 y rows x 8 z nodes of doubles) fall on alternating
// bank halves instead of overlapping ones.
constexpr bool kExtPad = true;
__host__ __device__ __forceinline__ int ext_ys(int Ez) { return kExtPad ? (Ez + 1) & ~1 : Ez; }

struct TileCoords {
    int t, x0, y0, z0, Ex, Ey, Ez;
    int ys, xs, esz;  // LDS z stride, x stride, elements of the (padded) extended tile
};

__device__ __forceinline__ TileCoords tile_coords(const MeshParams& mp, const TileGeom& tg, int t)
{
    TileCoords c;
    c.t = t;
    const int tz = c.t % tg.nt[2], ty = (c.t / tg.nt[2]) % tg.nt[1], tx = c.t / (tg.nt[2] * tg.nt[1]);
    c.x0 = tx * tg.T[0];
    c.y0 = ty * tg.T[1];
    c.z0 = tz * tg.T[2];
    c.Ex = tg.T[0] + mp.P - 1;
    c.Ey = tg.T[1] + mp.P - 1;
    c.Ez = tg.T[2] + mp.P - 1;
    c.ys = ext_ys(c.Ez);
    c.xs = c.Ey * c.ys;
    c.esz = c.Ex * c.xs;
    return c;
}

// LDS: [ext Ex*Ey*Ez (accumulator type)][w ch*3P T][u ch*3 T][qs ch T][off ch int][ids ch uint]
template <typename T>
struct Stage {
    T* w;
    T* u;
    T* qs;
    int* off;
    unsigned* ids;
};

template <typename T>
__device__ __forceinline__ Stage<T> stage_layout(char* smem, size_t ext_bytes, int ch, int P)
{
    Stage<T> s;
    s.w = reinterpret_cast<T*>(smem + ext_bytes);
    s.u = s.w + ch * 3 * P;
    s.qs = s.u + ch * 3;
    s.off = reinterpret_cast<int*>(s.qs + ch);
    s.ids = reinterpret_cast<unsigned*>(s.off + ch);
    return s;
}

// A bead's window: packed offset of its first node in the tile (x | y << 10 | z << 20)
// and u, its position relative to the window start in node units (node s+p minus bead
// = (p - u) h), so the axis weight is exp(-2 xi^2 h^2 (p - u)^2 / eta).
__device__ __forceinline__ int bead_window(const MeshParams& mp, const double4& r, int x0, int y0, int z0,
                                           double u[3])
{
    const int P = mp.P;
    const int sx = window_start(r.x, mp.L[0], mp.Nx, P);
    const int sy = window_start(r.y, mp.L[1], mp.Ny, P);
    const int sz = window_start(r.z, mp.L[2], mp.Nz, P);
    u[0] = (r.x + 0.5 * mp.L[0]) / mp.h[0] - sx;
    u[1] = (r.y + 0.5 * mp.L[1]) / mp.h[1] - sy;
    u[2] = (r.z + 0.5 * mp.L[2]) / mp.h[2] - sz;
    const int ox = wrap_idx(sx, mp.Nx) - x0, oy = wrap_idx(sy, mp.Ny) - y0, oz = wrap_idx(sz, mp.Nz) - z0;
    return ox | (oy << 10) | (oz << 20);
}

template <typename T>
__device__ __forceinline__ T axis_coef(const MeshParams& mp, int d)
{
    return (T)(mp.xiterm * mp.h[d] * mp.h[d] / mp.eta[d]);
}

// Load beads [k0, k0+cnt) of the bin and fill the chunk's LDS tables from positions.
template <typename T>
__device__ __forceinline__ void stage_chunk(const MeshParams& mp, const TileCoords& tc, const unsigned* __restrict__ sorted,
                                            unsigned k0, int cnt, const double4* __restrict__ pos,
                                            const double* __restrict__ q, T qscale, Stage<T>& s)
{
    const int P = mp.P;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const unsigned bead = sorted[k0 + i];
        double u[3];
        s.off[i] = bead_window(mp, pos[bead], tc.x0, tc.y0, tc.z0, u);
        s.ids[i] = bead;
        s.u[3 * i + 0] = (T)u[0];
        s.u[3 * i + 1] = (T)u[1];
        s.u[3 * i + 2] = (T)u[2];
        s.qs[i] = q ? qscale * (T)q[bead] : (T)0;
    }
    __syncthreads();
    const int per = 3 * P;
    const T c0 = axis_coef<T>(mp, 0), c1 = axis_coef<T>(mp, 1), c2 = axis_coef<T>(mp, 2);
    for (int j = threadIdx.x; j < cnt * per; j += blockDim.x) {
        const int i = j / per, rem = j - i * per;
        const int d = rem / P, p = rem - d * P;
        const T dd = (T)p - s.u[3 * i + d];
        const T cd = d == 0 ? c0 : (d == 1 ? c1 : c2);
        s.w[j] = gexp<T>(-cd * dd * dd);
    }
    __syncthreads();
}

// Window records of every bead in binned order, built once per bead configuration: the
// window's packed start in its tile and, in FP32, the bead's position in the window per
// axis (node units; 4 values per bead for aligned loads), from which the kernels evaluate
// the axis weights; in FP64 the 3P axis weights themselves (double exponentials are too
// slow to re-evaluate in every application).
template <typename T>
__global__ __launch_bounds__(256) void k_bead_records(MeshParams mp, TileGeom tg, int N,
                                                      const unsigned* __restrict__ sorted,
                                                      const unsigned* __restrict__ bead_tile,
                                                      const double4* __restrict__ pos, int* __restrict__ rec_off,
                                                      T* __restrict__ rec_win)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const unsigned bead = sorted[k];
    const int t = (int)bead_tile[bead];
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    double u[3];
    rec_off[k] = bead_window(mp, pos[bead], tx * tg.T[0], ty * tg.T[1], tz * tg.T[2], u);
    if constexpr (sizeof(T) == sizeof(double)) {
        const int P = mp.P;
        T* w = rec_win + (long)k * 3 * P;
        for (int d = 0; d < 3; ++d) {
            const T cd = axis_coef<T>(mp, d), ud = (T)u[d];
            for (int p = 0; p < P; ++p) {
                const T dd = (T)p - ud;
                w[d * P + p] = gexp<T>(-cd * dd * dd);
            }
        }
    } else {
        T* uk = rec_win + 4 * (long)k;
        uk[0] = (T)u[0];
        uk[1] = (T)u[1];
        uk[2] = (T)u[2];
        uk[3] = (T)0;
    }
}

// Fill the chunk's LDS tables for beads [k0, k0+cnt) from their window records.  FP32:
// three threads per bead evaluate its 3P axis weights exp(-c_d (p - u_d)^2), one axis
// each, and keep u; FP64: the stored weights are copied.
template <typename T>
__device__ __forceinline__ void stage_records(const MeshParams& mp, int P, const unsigned* __restrict__ sorted,
                                              unsigned k0, int cnt, const int* __restrict__ rec_off,
                                              const T* __restrict__ rec_win, const double* __restrict__ q, T qscale,
                                              Stage<T>& s)
{
    const int per = 3 * P;
    if constexpr (sizeof(T) == sizeof(double)) {
        const T* src = rec_win + (long)k0 * per;
        for (int j = threadIdx.x; j < cnt * per; j += blockDim.x) s.w[j] = src[j];
    } else {
        for (int j = threadIdx.x; j < 3 * cnt; j += blockDim.x) {
            const int i = j / 3, d = j - 3 * i;
            const T ud = rec_win[4 * (long)(k0 + i) + d];
            s.u[j] = ud;  // the force contraction's (u - p) factors
            const T cd = axis_coef<T>(mp, d);
            T* w = s.w + i * per + d * P;
            for (int p = 0; p < P; ++p) {
                const T dd = (T)p - ud;
                w[p] = gexp<T>(-cd * dd * dd);
            }
        }
    }
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const unsigned bead = sorted[k0 + i];
        s.off[i] = rec_off[k0 + i];
        s.ids[i] = bead;
        if (q) s.qs[i] = qscale * (T)q[bead];
    }
    __syncthreads();
}

// Walks a thread's nodes k = tid, tid + stride, ... of an Ex x Ey x Ez box in
// row-major order, keeping (lx, ly, lz) by carries instead of divisions.
struct BoxWalk {
    int lx, ly, lz, dx, dy, dz, Ey, Ez;
    __device__ __forceinline__ BoxWalk(int tid, int stride, int Ey_, int Ez_) : Ey(Ey_), Ez(Ez_)
    {
        lz = tid % Ez;
        const int r = tid / Ez;
        ly = r % Ey;
        lx = r / Ey;
        dz = stride % Ez;
        const int q = stride / Ez;
        dy = q % Ey;
        dx = q / Ey;
    }
    __device__ __forceinline__ void next()
    {
        lz += dz;
        int c = lz >= Ez;
        lz -= c ? Ez : 0;
        ly += dy + c;
        c = ly >= Ey;
        ly -= c ? Ey : 0;
        lx += dx + c;
    }
};

// PT > 0: window support fixed at compile time (loops unroll, P divisions become
// constants, the P node updates of a lane issue back to back); PT == 0: runtime P <= 16.
// The LDS tile accumulates in A = double for both mesh precisions: ds_add_f64 measures
// several times the throughput of ds_add_f32 on gfx950, and the sums gain accuracy.
template <typename T, typename A, int PT>
__device__ __forceinline__ void spread_tile(const MeshParams& mp, const TileGeom& tg, int t, int ch,
                                            const unsigned* __restrict__ tstart,
                                            const unsigned* __restrict__ sorted, const int* __restrict__ rec_off,
                                            const T* __restrict__ rec_win, const double* __restrict__ q,
                                            T* __restrict__ grid, char* smem_raw)
{
    const TileCoords tc = tile_coords(mp, tg, t);
    constexpr int PM = PT > 0 ? PT : 16;
    const int P = PT > 0 ? PT : mp.P;
    const int esz = tc.esz;
    const unsigned kb = tstart[tc.t], ke = tstart[tc.t + 1];
    if (kb == ke) return;  // an empty bin adds nothing to the (cleared) mesh
    A* ext = reinterpret_cast<A*>(smem_raw);
    Stage<T> s = stage_layout<T>(smem_raw, (size_t)esz * sizeof(A), ch, P);
    for (int k = threadIdx.x; k < esz; k += blockDim.x) ext[k] = (A)0;
    // the write-out below reads nodes other threads zeroed
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int xs = tc.xs, ys = tc.ys, per = 3 * P;
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        stage_records<T>(mp, P, sorted, k0, cnt, rec_off, rec_win, q, (T)mp.prefac, s);
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
                A* col = ext + ox * xs + (oy + py) * ys + (oz + pz);
#pragma unroll
                for (int px = 0; px < PM; ++px)
                    if (px < P) atomic_add(col + px * xs, (A)(c * wx[px]));
            }
        }
        __syncthreads();
    }
    const long plane = (long)mp.Ny * mp.Nz;
    const int nodes = tc.Ex * tc.Ey * tc.Ez;
    BoxWalk bw(threadIdx.x, blockDim.x, tc.Ey, tc.Ez);
    for (int k = threadIdx.x; k < nodes; k += blockDim.x, bw.next()) {
        int gx = tc.x0 + bw.lx, gy = tc.y0 + bw.ly, gz = tc.z0 + bw.lz;
        gx -= gx >= mp.Nx ? mp.Nx : 0;
        gy -= gy >= mp.Ny ? mp.Ny : 0;
        gz -= gz >= mp.Nz ? mp.Nz : 0;
        // nodes no window of the bin reached stay exactly zero: adding them changes nothing
        const A v = ext[bw.lx * xs + bw.ly * ys + bw.lz];
        if (v != (A)0) atomic_add(&grid[(long)gx * plane + (long)gy * mp.Nz + gz], (T)v);
    }
}

template <typename T, typename A, int PT>
__global__ __launch_bounds__(kSpreadThreads) void k_spread_tiled(MeshParams mp, TileGeom tg, int ch,
                                                      const unsigned* __restrict__ tstart,
                                                      const unsigned* __restrict__ sorted,
                                                      const int* __restrict__ rec_off, const T* __restrict__ rec_win,
                                                      const double* __restrict__ q, T* __restrict__ grid)
{
    extern __shared__ char smem_raw[];
    spread_tile<T, A, PT>(mp, tg, xcd_tile(blockIdx.x, tg.ntiles), ch, tstart, sorted, rec_off, rec_win, q, grid,
                          smem_raw);
}

template <typename T, bool FORCE, int PT>
__device__ __forceinline__ void contract_tile(const MeshParams& mp, const TileGeom& tg, int t, int ch,
                                              const unsigned* __restrict__ tstart,
                                              const unsigned* __restrict__ sorted, const int* __restrict__ rec_off,
                                              const T* __restrict__ rec_win, const double4* __restrict__ pos,
                                              const double* __restrict__ q, const T* __restrict__ grid,
                                              T* __restrict__ phi, double* __restrict__ force,
                                              const double* __restrict__ fadd, char* smem_raw)
{
    const TileCoords tc = tile_coords(mp, tg, t);
    constexpr int PM = PT > 0 ? PT : 16;
    const int P = PT > 0 ? PT : mp.P;
    const int esz = tc.esz;
    const int xs = tc.xs, ys = tc.ys, per = 3 * P;
    T* ext = reinterpret_cast<T*>(smem_raw);
    Stage<T> s = stage_layout<T>(smem_raw, (size_t)esz * sizeof(T), ch, P);
    const long plane = (long)mp.Ny * mp.Nz;
    const int nodes = tc.Ex * tc.Ey * tc.Ez;
    BoxWalk bw(threadIdx.x, blockDim.x, tc.Ey, tc.Ez);
    // a thread's nodes are loaded in batches before they are stored, so that the loads'
    // latencies overlap instead of adding up
    constexpr int kBatch = 4;
    for (int k0 = threadIdx.x; k0 < nodes; k0 += kBatch * blockDim.x) {
        T v[kBatch];
        int o[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            o[u] = -1;
            v[u] = (T)0;
            if (k0 + u * (int)blockDim.x < nodes) {
                int gx = tc.x0 + bw.lx, gy = tc.y0 + bw.ly, gz = tc.z0 + bw.lz;
                gx -= gx >= mp.Nx ? mp.Nx : 0;
                gy -= gy >= mp.Ny ? mp.Ny : 0;
                gz -= gz >= mp.Nz ? mp.Nz : 0;
                v[u] = grid[(long)gx * plane + (long)gy * mp.Nz + gz];
                o[u] = bw.lx * xs + bw.ly * ys + bw.lz;
            }
            bw.next();
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (o[u] >= 0) ext[o[u]] = v[u];
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const T pc = (T)(mp.quadW * mp.prefac);
    const unsigned kb = tstart[tc.t], ke = tstart[tc.t + 1];
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        if (FORCE && sizeof(T) == sizeof(double))  // FP64 records hold weights, not positions
            stage_chunk<T>(mp, tc, sorted, k0, cnt, pos, q, (T)1, s);
        else
            stage_records<T>(mp, P, sorted, k0, cnt, rec_off, rec_win, FORCE ? q : nullptr, (T)1, s);
        for (int i = wave; i < cnt; i += nwave) {
            const int o = s.off[i];
            const int ox = o & 1023, oy = (o >> 10) & 1023, oz = o >> 20;
            const T* wi = s.w + i * per;
            const T ux = FORCE ? s.u[3 * i] : (T)0, uy = FORCE ? s.u[3 * i + 1] : (T)0;
            const T uz = FORCE ? s.u[3 * i + 2] : (T)0;
            T wx[PM];
#pragma unroll
            for (int px = 0; px < PM; ++px) wx[px] = (px < P) ? wi[px] : (T)0;
            T acc = 0, fx = 0, fy = 0, fz = 0;
            for (int pair = lane; pair < P * P; pair += kWave) {
                const int py = pair / P, pz = pair - py * P;
                const T wyz = wi[P + py] * wi[2 * P + pz];
                const T* col = ext + ox * xs + (oy + py) * ys + (oz + pz);
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
                acc = wave_total(acc);
                if (lane == 0) phi[bead] = acc * pc;
            } else {
                fx = wave_total(fx);
                fy = wave_total(fy);
                fz = wave_total(fz);
                if (lane == 0) {
                    // F = 2 (h^3 prefac) (2 xi^2) q_i sum phi_g e^{...} (r/eta), r = bead - node = (u - p) h
                    // (cppmodule/ConductorRigid.cu:449)
                    const double c = 2.0 * (mp.quadW * mp.prefac) * mp.xiterm * (double)s.qs[i];
                    const long o = 4 * (long)bead;
                    // fadd: the real-space force of the same bead, computed beside this pass
                    const double a0 = fadd ? fadd[o] : 0.0, a1 = fadd ? fadd[o + 1] : 0.0;
                    const double a2 = fadd ? fadd[o + 2] : 0.0;
                    force[o + 0] = c * mp.h[0] / mp.eta[0] * (double)fx + a0;
                    force[o + 1] = c * mp.h[1] / mp.eta[1] * (double)fy + a1;
                    force[o + 2] = c * mp.h[2] / mp.eta[2] * (double)fz + a2;
                    force[o + 3] = 0.0;
                }
            }
        }
        __syncthreads();
    }
}

template <typename T, bool FORCE, int PT>
__global__ __launch_bounds__(contract_threads<T>()) void k_contract_tiled(MeshParams mp, TileGeom tg, int ch,
                                                        const unsigned* __restrict__ tstart,
                                                        const unsigned* __restrict__ sorted,
                                                        const int* __restrict__ rec_off, const T* __restrict__ rec_win,
                                                        const double4* __restrict__ pos, const double* __restrict__ q,
                                                        const T* __restrict__ grid, T* __restrict__ phi,
                                                        double* __restrict__ force, const double* __restrict__ fadd)
{
    extern __shared__ char smem_raw[];
    contract_tile<T, FORCE, PT>(mp, tg, xcd_tile(blockIdx.x, tg.ntiles), ch, tstart, sorted, rec_off, rec_win, pos, q,
                                grid, phi, force, fadd, smem_raw);
}

// Batched launches: one workgroup per tile of any box.  Boxes own consecutive ranges of
// the global tile index (tile0 ascending); the XCD mapping runs over the global index, so
// each XCD takes a contiguous run of tiles that stays within one box where it can.
__device__ __forceinline__ int batch_box(const TiledBox* __restrict__ boxes, int nbox, int g)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].tile0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ void k_fill_batch(const TiledBox* __restrict__ boxes, const int* __restrict__ skip)
{
    const TiledBox& d = boxes[blockIdx.y];
    if (skip && skip[d.box]) return;
    T* g = static_cast<T*>(d.grid);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < d.G; i += (long)gridDim.x * blockDim.x)
        g[i] = (T)0;
}

template <typename T, typename A, int PT>
__global__ __launch_bounds__(kSpreadThreads) void k_spread_tiled_batch(const TiledBox* __restrict__ boxes, int nbox,
                                                                      int total, const int* __restrict__ skip)
{
    extern __shared__ char smem_raw[];
    const int g = xcd_tile(blockIdx.x, total);
    const TiledBox& d = boxes[batch_box(boxes, nbox, g)];
    if (skip && skip[d.box]) return;
    spread_tile<T, A, PT>(d.mp, d.tg, g - d.tile0, d.ch, d.tile_start, d.sorted, d.rec_off,
                          static_cast<const T*>(d.rec_win), d.q, static_cast<T*>(d.grid), smem_raw);
}

template <typename T, bool FORCE, int PT>
__global__ __launch_bounds__(contract_threads<T>()) void k_contract_tiled_batch(const TiledBox* __restrict__ boxes, int nbox,
                                                                        int total, const int* __restrict__ skip)
{
    extern __shared__ char smem_raw[];
    const int g = xcd_tile(blockIdx.x, total);
    const TiledBox& d = boxes[batch_box(boxes, nbox, g)];
    if (skip && skip[d.box]) return;
    contract_tile<T, FORCE, PT>(d.mp, d.tg, g - d.tile0, d.ch, d.tile_start, d.sorted, d.rec_off,
                                static_cast<const T*>(d.rec_win), d.pos, FORCE ? d.q : nullptr,
                                static_cast<const T*>(d.grid), static_cast<T*>(d.phi), FORCE ? d.force : nullptr,
                                nullptr, smem_raw);
}

// ---- batched per-evaluation preparation (see PrepBox) ------------------------
__device__ __forceinline__ int prep_box_of(const PrepBox* __restrict__ boxes, int nbox, int g)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (boxes[mid].blk0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// counters of the box's binning and body grouping, its ELL width and weight bound
__global__ __launch_bounds__(256) void k_prep_zero(const PrepBox* __restrict__ boxes, int nbox)
{
    const PrepBox& d = boxes[prep_box_of(boxes, nbox, blockIdx.x)];
    const long i = (long)(blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
    if (d.bin) {
        if (i <= d.tg.ntiles) d.tile_count[i] = 0u;
        if (i < d.tg.ntiles) d.cursor[i] = 0u;
    }
    if (d.body_count && i <= d.s.N_isolated) d.body_count[i] = 0u;
    if (i == 0) {
        if (d.width) *d.width = 0u;
        if (d.wbits) *d.wbits = 0u;
    }
}

// per bead: membership, tile and tile counts, the ELL width; per body: its bead count
// (a rigid body holds rigid_size beads, a bead outside every rigid body is a body alone)
__global__ __launch_bounds__(256) void k_prep_beads(const PrepBox* __restrict__ boxes, int nbox)
{
    const PrepBox& d = boxes[prep_box_of(boxes, nbox, blockIdx.x)];
    const SysView& s = d.s;
    const long i = (long)(blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
    const bool bead = i < s.N;
    if (bead) {
        const unsigned b = d.body_of_tag[d.tag[i]];
        d.iso[i] = b;
        if (d.body_count && s.rigid_membership[i] == NOT_MEMBER) d.body_count[b] = 1u;
    }
    const bool on = d.bin && bead;
    const unsigned t = on ? (unsigned)tile_of_bead(d.mp, d.tg, s.pos[i]) : 0u;
    if (on) d.bead_tile[i] = t;
    (void)wave_count(d.tile_count, t, on);
    if (d.body_count && i < s.N_rigid && s.rigid_size[i] > 0)
        d.body_count[d.body_of_tag[d.tag[s.rigid_members[i * s.Nmax]]]] = s.rigid_size[i];
    if (d.width) {
        unsigned v = bead ? s.n_neigh[i] : 0u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, off, 64));
        if ((threadIdx.x & 63) == 0 && v) atomicMax(d.width, v);
    }
}

// exclusive scan of n counts into out (out[n] need not exist), one workgroup, chunks of its
// size with a running carry
__device__ __forceinline__ void block_exclusive_scan(const unsigned* __restrict__ in, unsigned* __restrict__ out,
                                                     long n)
{
    __shared__ unsigned wsum[16];
    __shared__ unsigned carry_s;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) carry_s = 0u;
    __syncthreads();
    for (long base = 0; base < n; base += blockDim.x) {
        const long k = base + threadIdx.x;
        const unsigned v = k < n ? in[k] : 0u;
        unsigned x = v;  // inclusive scan within the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned y = (unsigned)__shfl_up((int)x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) wsum[wave] = x;
        __syncthreads();
        unsigned before = carry_s;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (k < n) out[k] = before + x - v;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned tot = 0;
            for (int w = 0; w < nw; ++w) tot += wsum[w];
            carry_s += tot;
        }
        __syncthreads();
    }
}

// blockIdx.x: box; blockIdx.y: 0 the tile counts, 1 the body counts
__global__ __launch_bounds__(1024) void k_prep_scan(const PrepBox* __restrict__ boxes)
{
    const PrepBox& d = boxes[blockIdx.x];
    if (blockIdx.y == 0) {
        if (d.bin) block_exclusive_scan(d.tile_count, d.tile_start, (long)d.tg.ntiles + 1);
    } else if (d.body_count) {
        block_exclusive_scan(d.body_count, d.body_start, (long)d.s.N_isolated + 1);
    }
}

// beads into their tiles' slots; bodies' members in a fixed order (as k_body_fill_*)
__global__ __launch_bounds__(256) void k_prep_fill(const PrepBox* __restrict__ boxes, int nbox)
{
    const PrepBox& d = boxes[prep_box_of(boxes, nbox, blockIdx.x)];
    const SysView& s = d.s;
    const long i = (long)(blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
    const bool on = d.bin && i < s.N;
    const unsigned t = on ? d.bead_tile[i] : 0u;
    const unsigned slot = wave_count(d.cursor, t, on);
    if (on) d.sorted[d.tile_start[t] + slot] = (unsigned)i;
    if (!d.body_beads) return;
    if (i < s.N && s.rigid_membership[i] == NOT_MEMBER) d.body_beads[d.body_start[d.iso[i]]] = (unsigned)i;
    if (i < (long)s.N_rigid * s.Nmax) {
        const long r = i / s.Nmax;
        const int k = (int)(i - r * s.Nmax);
        if (k < (int)s.rigid_size[r]) {
            const unsigned* mem = s.rigid_members + r * s.Nmax;
            d.body_beads[d.body_start[d.iso[mem[0]]] + k] = mem[k];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_prep_records(const PrepBox* __restrict__ boxes, int nbox)
{
    const PrepBox& d = boxes[prep_box_of(boxes, nbox, blockIdx.x)];
    if (!d.bin) return;
    const long k = (long)(blockIdx.x - d.blk0) * blockDim.x + threadIdx.x;
    if (k >= d.s.N) return;
    const MeshParams& mp = d.mp;
    const TileGeom& tg = d.tg;
    const unsigned bead = d.sorted[k];
    const int t = (int)d.bead_tile[bead];
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    double u[3];
    d.rec_off[k] = bead_window(mp, d.s.pos[bead], tx * tg.T[0], ty * tg.T[1], tz * tg.T[2], u);
    T* rec_win = static_cast<T*>(d.rec_win);
    if constexpr (sizeof(T) == sizeof(double)) {
        const int P = mp.P;
        T* w = rec_win + k * 3 * P;
        for (int dd = 0; dd < 3; ++dd) {
            const T cd = axis_coef<T>(mp, dd), ud = (T)u[dd];
            for (int p = 0; p < P; ++p) {
                const T e = (T)p - ud;
                w[dd * P + p] = gexp<T>(-cd * e * e);
            }
        }
    } else {
        T* uk = rec_win + 4 * k;
        uk[0] = (T)u[0];
        uk[1] = (T)u[1];
        uk[2] = (T)u[2];
        uk[3] = (T)0;
    }
}

inline unsigned blocks(long n, int b)
{
    long v = (n + b - 1) / b;
    return (unsigned)(v > 0 ? v : 1);
}

}  // namespace

TiledLds tiled_lds(const MeshParams& mp, const TileGeom& tg, size_t ext_elem, size_t stage_elem)
{
    TiledLds l;
    l.ext_bytes = (size_t)(tg.T[0] + mp.P - 1) * (tg.T[1] + mp.P - 1) * ext_ys(tg.T[2] + mp.P - 1) * ext_elem;
    const size_t per_bead = (size_t)(3 * mp.P + 3 + 1) * stage_elem + 2 * sizeof(int);
    const size_t budget = tg.lds > 0 ? (size_t)tg.lds : kTiledLdsBudget;
    long ch = l.ext_bytes < budget ? (long)((budget - l.ext_bytes) / per_bead) : 0;
    if (ch > kTiledChunk) ch = kTiledChunk;
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
hipError_t launch_bead_records(const MeshParams& mp, const TileGeom& tg, const SysView& s, const unsigned* sorted,
                               const unsigned* bead_tile, int* rec_off, T* rec_win, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bead_records<T>, dim3(blocks(s.N, 256)), dim3(256), 0, st, mp, tg, s.N, sorted, bead_tile,
                       s.pos, rec_off, rec_win);
    return hipGetLastError();
}

// Raise a kernel's dynamic-LDS limit above 64 KB once per device and size.  The attribute
// persists, and setting it costs host time on the launch path of every GMRES step.
static void allow_lds(const void* fn, size_t bytes)
{
    if (bytes <= 65536) return;
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> raised;  // (device, kernel) -> bytes allowed
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = raised[{dev, fn}];
    if (have >= bytes) return;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) have = bytes;
}

template <typename T>
hipError_t launch_spread_tiled(const MeshParams& mp, const TileGeom& tg, const BinView& bv, const double* q, T* grid,
                               hipStream_t st, bool clear)
{
    const TiledLds l = tiled_lds(mp, tg, sizeof(SpreadAcc), sizeof(T));
    if (l.chunk < 4 || mp.P > 16) return hipErrorInvalidValue;
    if (clear) {
        hipError_t e = launch_fill<T>(grid, (T)0, (long)mp.Nx * mp.Ny * mp.Nz, st);
        if (e != hipSuccess) return e;
    }
    const T* rec_win = static_cast<const T*>(bv.rec_win);
#define CR_SPREAD_P(PV)                                                                                        \
    do {                                                                                                       \
        allow_lds(reinterpret_cast<const void*>(&k_spread_tiled<T, SpreadAcc, PV>), l.total_bytes);          \
        hipLaunchKernelGGL((k_spread_tiled<T, SpreadAcc, PV>), dim3(tg.ntiles), dim3(kSpreadThreads),          \
                           l.total_bytes, st, mp, tg, l.chunk, bv.tile_start, bv.sorted, bv.rec_off, rec_win, q, \
                           grid);                                                                              \
    } while (0)
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
    return hipGetLastError();
}

template <typename T>
hipError_t launch_contract_tiled(const MeshParams& mp, const TileGeom& tg, const SysView& s, const BinView& bv,
                                 const double* q, const T* grid, T* phi, double* force, hipStream_t st,
                                 const double* fadd)
{
    const TiledLds l = tiled_lds(mp, tg, sizeof(T), sizeof(T));
    if (l.chunk < 4 || mp.P > 16) return hipErrorInvalidValue;
    const T* rec_win = static_cast<const T*>(bv.rec_win);
#define CR_CONTRACT_P(PV)                                                                                       \
    do {                                                                                                        \
        allow_lds(reinterpret_cast<const void*>(&k_contract_tiled<T, true, PV>), l.total_bytes);              \
        allow_lds(reinterpret_cast<const void*>(&k_contract_tiled<T, false, PV>), l.total_bytes);             \
        if (force)                                                                                              \
            hipLaunchKernelGGL((k_contract_tiled<T, true, PV>), dim3(tg.ntiles), dim3(contract_threads<T>()), l.total_bytes, st,  \
                               mp, tg, l.chunk, bv.tile_start, bv.sorted, bv.rec_off, rec_win, s.pos, q, grid,    \
                               (T*)nullptr, force, fadd);                                                       \
        else                                                                                                    \
            hipLaunchKernelGGL((k_contract_tiled<T, false, PV>), dim3(tg.ntiles), dim3(contract_threads<T>()), l.total_bytes, st, \
                               mp, tg, l.chunk, bv.tile_start, bv.sorted, bv.rec_off, rec_win, s.pos,             \
                               (const double*)nullptr, grid, phi, (double*)nullptr, (const double*)nullptr);    \
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

template <typename T>
hipError_t launch_fill_batch(const TiledBox* boxes, int nbox, long Gmax, const int* skip, hipStream_t st)
{
    if (nbox <= 0) return hipSuccess;
    long nb = (Gmax + 255) / 256;
    if (nb > 2048) nb = 2048;
    hipLaunchKernelGGL(k_fill_batch<T>, dim3((unsigned)std::max(1L, nb), (unsigned)nbox), dim3(256), 0, st, boxes,
                       skip);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_spread_tiled_batch(const TiledBox* boxes, int nbox, int total_tiles, int P, size_t lds_bytes,
                                     const int* skip, hipStream_t st)
{
    if (nbox <= 0 || total_tiles <= 0) return hipSuccess;
    if (P > 16) return hipErrorInvalidValue;
#define CR_SPREAD_B(PV)                                                                                         \
    do {                                                                                                        \
        allow_lds(reinterpret_cast<const void*>(&k_spread_tiled_batch<T, SpreadAcc, PV>), lds_bytes);         \
        hipLaunchKernelGGL((k_spread_tiled_batch<T, SpreadAcc, PV>), dim3(total_tiles), dim3(kSpreadThreads),    \
                           lds_bytes, st, boxes, nbox, total_tiles, skip);                                      \
    } while (0)
    switch (P) {
        case 4: CR_SPREAD_B(4); break;
        case 5: CR_SPREAD_B(5); break;
        case 6: CR_SPREAD_B(6); break;
        case 7: CR_SPREAD_B(7); break;
        case 8: CR_SPREAD_B(8); break;
        case 9: CR_SPREAD_B(9); break;
        case 10: CR_SPREAD_B(10); break;
        default: CR_SPREAD_B(0); break;
    }
#undef CR_SPREAD_B
    return hipGetLastError();
}

template <typename T>
hipError_t launch_contract_tiled_batch(const TiledBox* boxes, int nbox, int total_tiles, int P, size_t lds_bytes,
                                       const int* skip, hipStream_t st, bool force)
{
    if (nbox <= 0 || total_tiles <= 0) return hipSuccess;
    if (P > 16) return hipErrorInvalidValue;
#define CR_CONTRACT_B(PV)                                                                                       \
    do {                                                                                                        \
        allow_lds(reinterpret_cast<const void*>(&k_contract_tiled_batch<T, false, PV>), lds_bytes);           \
        allow_lds(reinterpret_cast<const void*>(&k_contract_tiled_batch<T, true, PV>), lds_bytes);            \
        if (force)                                                                                              \
            hipLaunchKernelGGL((k_contract_tiled_batch<T, true, PV>), dim3(total_tiles), dim3(contract_threads<T>()),   \
                               lds_bytes, st, boxes, nbox, total_tiles, skip);                                  \
        else                                                                                                    \
            hipLaunchKernelGGL((k_contract_tiled_batch<T, false, PV>), dim3(total_tiles), dim3(contract_threads<T>()),  \
                               lds_bytes, st, boxes, nbox, total_tiles, skip);                                  \
    } while (0)
    switch (P) {
        case 4: CR_CONTRACT_B(4); break;
        case 5: CR_CONTRACT_B(5); break;
        case 6: CR_CONTRACT_B(6); break;
        case 7: CR_CONTRACT_B(7); break;
        case 8: CR_CONTRACT_B(8); break;
        case 9: CR_CONTRACT_B(9); break;
        case 10: CR_CONTRACT_B(10); break;
        default: CR_CONTRACT_B(0); break;
    }
#undef CR_CONTRACT_B
    return hipGetLastError();
}

#define CR_INST(T)                                                                                               \
    template hipError_t launch_bead_records<T>(const MeshParams&, const TileGeom&, const SysView&, const unsigned*, \
                                               const unsigned*, int*, T*, hipStream_t);                         \
    template hipError_t launch_spread_tiled<T>(const MeshParams&, const TileGeom&, const BinView&, const double*, \
                                               T*, hipStream_t, bool);                                          \
    template hipError_t launch_contract_tiled<T>(const MeshParams&, const TileGeom&, const SysView&,              \
                                                 const BinView&, const double*, const T*, T*, double*, hipStream_t, \
                                                 const double*);                                                    \
    template hipError_t launch_fill_batch<T>(const TiledBox*, int, long, const int*, hipStream_t);             \
    template hipError_t launch_spread_tiled_batch<T>(const TiledBox*, int, int, int, size_t, const int*,          \
                                                     hipStream_t);                                              \
    template hipError_t launch_contract_tiled_batch<T>(const TiledBox*, int, int, int, size_t, const int*,        \
                                                       hipStream_t, bool);
CR_INST(float)
CR_INST(double)
#undef CR_INST

hipError_t launch_prep_batch(const PrepBox* boxes, int nbox, int total_blocks, bool fp32, hipStream_t st)
{
    if (nbox <= 0 || total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prep_zero, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    hipLaunchKernelGGL(k_prep_beads, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    hipLaunchKernelGGL(k_prep_scan, dim3(nbox, 2), dim3(1024), 0, st, boxes);
    hipLaunchKernelGGL(k_prep_fill, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    if (fp32)
        hipLaunchKernelGGL(k_prep_records<float>, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    else
        hipLaunchKernelGGL(k_prep_records<double>, dim3(total_blocks), dim3(256), 0, st, boxes, nbox);
    return hipGetLastError();
}

}  // namespace cr
