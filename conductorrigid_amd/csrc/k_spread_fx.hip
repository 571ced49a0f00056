// k_spread_fx.hip -- FP32-bulk spreading (cppmodule/ConductorRigid.cu:144-234) in 32-bit
// fixed point.
//
// The reference adds every window product into the mesh with a float atomic.  Here, as in
// the binned kernels of k_tiled.hip, beads are binned by the mesh tile holding their
// window's first node and one workgroup accumulates its tile, extended by P-1 nodes per
// axis, in LDS.  The accumulator is a 32-bit integer: each product c = S q w_x w_y w_z is
// rounded to the nearest integer and added with ds_add_u32.  On gfx950 an LDS float atomic
// add runs an order of magnitude slower than the integer one, and a double accumulator
// takes twice the LDS; the integer tile is half that size, so two workgroups fit a CU.  The
// extended tile then leaves with plain stores as a block of a scratch buffer, and
// k_fx_gather sums, for every mesh node, the (at most 8) blocks that cover it and writes the
// float mesh.  Global atomics would add the extended tiles (3x the mesh at 16^3 tiles) at
// the memory-side atomic rate, about a quarter of the plain-store bandwidth; the blocks
// cost one store and one load each instead, and the mesh needs no clearing.
//
// Integer sums do not depend on the order of the additions, so the spread is
// deterministic: every run gives the same mesh bits.
//
// Scale.  A mesh node's value is prefac * sum_b q_b w_b(node).  With qmax = max |q_b| and
// W a bound on sum_b w_b(node) over all nodes, |node| <= prefac qmax W, so the multiplier
// m = 2^30 / (qmax W) applied to q w keeps every partial and total sum below 2^30 in
// magnitude (the rounding adds at most 1/2 per product), and int32 cannot overflow.  W is
// computed once per bead configuration (k_spread_fx<..., ONES = true>): every tile
// accumulates sum_b w_b for its own beads and reports its largest node; a node is covered
// by at most 8 tiles (two per axis), so W = 8 x the largest tile node.  qmax is computed for
// every spread by k_charges_fx, which also rounds the charges to float for the FP32-bulk
// kernels; k_fx_scale then writes m and the decode factor prefac qmax W / 2^30.
#include "cr_kernels.h"
#include "se_device.h"

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

namespace cr {

namespace {

constexpr int kFxWave = 64;
constexpr int kFxChargeThreads = 256;
constexpr int kFxLoadBatch = 5;  // extended-tile items per thread in flight (16^3 tiles, P = 8: 4232 items)

__device__ __forceinline__ int round_fx(float x)
{
    // floor(x + 0.5) in one instruction
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Workgroups are dealt round-robin to the 8 XCDs: give each XCD a contiguous run of tiles.
__device__ __forceinline__ int fx_tile(int b, int ntiles)
{
    if (ntiles % 8 != 0) return b;
    return (b % 8) * (ntiles / 8) + b / 8;
}

// NaN-propagating maximum of |values| (a NaN charge must reach the mesh as NaN)
__device__ __forceinline__ float nanmax(float m, float a) { return (a > m || a != a) ? a : m; }

__device__ __forceinline__ float wave_nanmax(float m)
{
#pragma unroll
    for (int off = kFxWave / 2; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off, kFxWave));
    return m;
}

// Block blk of nblk rounds its share of x to float and leaves |q|'s NaN-propagating
// maximum in fs->part[blk].
__device__ __forceinline__ void charges_fx_part(const double* __restrict__ x, int N, float* __restrict__ qf,
                                                FxScale* __restrict__ fs, int blk, int nblk)
{
    __shared__ float wmax[kFxChargeThreads / kFxWave];
    float m = 0.0f;
    for (long i = (long)blk * blockDim.x + threadIdx.x; i < N; i += (long)nblk * blockDim.x) {
        const float v = (float)x[i];
        qf[i] = v;
        m = nanmax(m, fabsf(v));
    }
    m = wave_nanmax(m);
    const int lane = threadIdx.x & (kFxWave - 1), wave = threadIdx.x / kFxWave;
    if (lane == 0) wmax[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float bm = wmax[0];
        for (int k = 1; k < kFxChargeThreads / kFxWave; ++k) bm = nanmax(bm, wmax[k]);
        fs->part[blk] = bm;
    }
}

__global__ __launch_bounds__(kFxChargeThreads) void k_charges_fx(const double* __restrict__ x, int N,
                                                                 float* __restrict__ qf, FxScale* __restrict__ fs)
{
    charges_fx_part(x, N, qf, fs, blockIdx.x, gridDim.x);
}

// The multiplier and decode factor for charges of NaN-propagating maximum qm.
__device__ __forceinline__ void fx_scale_values(float qm, unsigned wbits, double prefac, float& mult, float& inv)
{
    const double W = 8.0 * (double)__uint_as_float(wbits);
    const double qW = (double)qm * W;
    if (qm != qm || !(qm <= 3.0e38f)) {  // NaN or infinite charge: decode to NaN
        mult = 0.0f;
        inv = __int_as_float(0x7fc00000);
    } else if (!(qW > 1e-30)) {          // all charges (or all weights) zero
        mult = 0.0f;
        inv = 0.0f;
    } else {
        mult = (float)(1073741824.0 / qW);
        inv = (float)(prefac * qW / 1073741824.0);
    }
}

// One wavefront: qmax from the partial maxima, then the multiplier and decode factor.
__device__ __forceinline__ void fx_scale_wave(int nparts, FxScale* __restrict__ fs, double prefac)
{
    const int lane = threadIdx.x;
    float qm = 0.0f;
    for (int k = lane; k < nparts; k += kFxWave) qm = nanmax(qm, fs->part[k]);
    qm = wave_nanmax(qm);
    if (lane != 0) return;
    float mult, inv;
    fx_scale_values(qm, fs->wbits, prefac, mult, inv);
    fs->mult = mult;
    fs->inv = inv;
    fs->qmax = qm;
}

__global__ __launch_bounds__(kFxWave) void k_fx_scale(int nparts, FxScale* __restrict__ fs, double prefac)
{
    fx_scale_wave(nparts, fs, prefac);
}

// Batched boxes: blockIdx.y is the box (the same block count for every box; blocks past a
// small box's beads report a zero partial).
__global__ __launch_bounds__(kFxChargeThreads) void k_charges_fx_batch(const FxBox* __restrict__ boxes,
                                                                       const int* __restrict__ skip)
{
    const FxBox& d = boxes[blockIdx.y];
    if (skip && skip[d.box]) return;
    charges_fx_part(d.x, d.N, d.qf, d.fs, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kFxWave) void k_fx_scale_batch(const FxBox* __restrict__ boxes, int nparts,
                                                             const int* __restrict__ skip)
{
    const FxBox& d = boxes[blockIdx.x];
    if (skip && skip[d.box]) return;
    fx_scale_wave(nparts, d.fs, d.mp.prefac);
}

// Node walk over an EX x EY x EZ box in row-major order with carries (no divisions).
struct FxWalk {
    int lx, ly, lz, dx, dy, dz, EY, EZ;
    __device__ __forceinline__ FxWalk(int tid, int stride, int EY_, int EZ_) : EY(EY_), EZ(EZ_)
    {
        lz = tid % EZ;
        const int r = tid / EZ;
        ly = r % EY;
        lx = r / EY;
        dz = stride % EZ;
        const int q = stride / EZ;
        dy = q % EY;
        dx = q / EY;
    }
    __device__ __forceinline__ void next()
    {
        lz += dz;
        int c = lz >= EZ;
        lz -= c ? EZ : 0;
        ly += dy + c;
        c = ly >= EY;
        ly -= c ? EY : 0;
        lx += dx + c;
    }
};

// Per-bead staging stride (floats): the 3 P axis weights, padded to whole float4s.
__host__ __device__ constexpr int fx_wstride(int P) { return (3 * P + 3) / 4 * 4; }

// A packed window record (x | y << 10 | z << 20, the window's first node in the extended
// tile) as an element offset in an LDS tile of row pitches xs (x) and ys (y).
__device__ __forceinline__ int fx_ext_off(int rec, int xs, int ys)
{
    return (rec & 1023) * xs + ((rec >> 10) & 1023) * ys + (rec >> 20);
}

// The bead loops run one wavefront per bead over the wavefront's uniform bead index; the
// lane's window node pairs (py, pz) = pair / P, pair % P for pair = lane and lane + 64.
__device__ __forceinline__ int fx_wave() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kFxWave)); }

// Sum over the 64 lanes of a fully active wavefront into lane 63 (DPP row shifts give each
// row's total in its lane 15, then row broadcasts 15 and 31 carry them up).
__device__ __forceinline__ float fx_total63(float v)
{
    v += se::dpp_shift<0x111>(v);
    v += se::dpp_shift<0x112>(v);
    v += se::dpp_shift<0x114>(v);
    v += se::dpp_shift<0x118>(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xa, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xc, 0xf, false));
    return v;
}

// TX = 0: tile edges from tg at run time; otherwise compile-time edges (the LDS strides
// become instruction offsets).  ONES: accumulate sum_b w_b with a per-tile multiplier and
// report the tile's largest node into fs->wbits instead of writing the mesh.
template <int PT, int TX, int TY, int TZ, bool ONES>
__device__ __forceinline__ void spread_fx_tile(const MeshParams& mp, const TileGeom& tg, int t, int ch,
                                               const unsigned* __restrict__ tstart,
                                               const unsigned* __restrict__ sorted, const int* __restrict__ rec_off,
                                               const float4* __restrict__ rec_u, const float* __restrict__ qf,
                                               FxScale* __restrict__ fs, long bstride, int* __restrict__ blocks)
{
    constexpr bool CT = TX > 0;
    constexpr int WS = fx_wstride(PT);
    const int T0 = CT ? TX : tg.T[0], T1 = CT ? TY : tg.T[1], T2 = CT ? TZ : tg.T[2];
    const int EX = T0 + PT - 1, EY = T1 + PT - 1, EZ = T2 + PT - 1;
    const int ys = (EZ + 1) & ~1, xs = EY * ys, esz = EX * xs;
    extern __shared__ __align__(16) char smem[];
    __shared__ unsigned tile_max;
    const unsigned kb = tstart[t], ke = tstart[t + 1];
    if (kb == ke) {
        // an empty bin: a block of zeros (the gather reads every covering block)
        if constexpr (!ONES) {
            int* blk = blocks + (long)t * bstride;
            for (int k = threadIdx.x; k < (int)(bstride / 4); k += kFxThreads)
                reinterpret_cast<uint4*>(blk)[k] = make_uint4(0, 0, 0, 0);
        }
        return;
    }
    unsigned* ext = reinterpret_cast<unsigned*>(smem);
    float* w = reinterpret_cast<float*>(smem + ((size_t)esz * 4 + 15) / 16 * 16);
    int2* oq = reinterpret_cast<int2*>(w + (size_t)ch * WS);
    for (int k = threadIdx.x; k < esz / 4; k += kFxThreads) reinterpret_cast<uint4*>(ext)[k] = make_uint4(0, 0, 0, 0);
    for (int k = (esz / 4) * 4 + threadIdx.x; k < esz; k += kFxThreads) ext[k] = 0u;
    if (threadIdx.x == 0) tile_max = 0u;
    __syncthreads();
    // ONES: per-tile multiplier 2^30 / count keeps the tile's sums of weights <= 1 below 2^30
    const float mult = ONES ? (float)(1073741824.0 / (double)(ke - kb)) : fs->mult;
    const float c0 = (float)(mp.xiterm * mp.h[0] * mp.h[0] / mp.eta[0]);
    const float c1 = (float)(mp.xiterm * mp.h[1] * mp.h[1] / mp.eta[1]);
    const float c2 = (float)(mp.xiterm * mp.h[2] * mp.h[2] / mp.eta[2]);
    const int lane = threadIdx.x & (kFxWave - 1), wave = fx_wave();
    constexpr int NW = kFxThreads / kFxWave;
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        // three threads per bead: the P weights exp(-c_d (p - u_d)^2) of one axis
        for (int j = threadIdx.x; j < 3 * cnt; j += kFxThreads) {
            const int i = j / 3, d = j - 3 * i;
            const float4 u4 = rec_u[k0 + i];
            const float ud = d == 0 ? u4.x : (d == 1 ? u4.y : u4.z);
            const float cd = d == 0 ? c0 : (d == 1 ? c1 : c2);
            float* wp = w + i * WS + d * PT;
#pragma unroll
            for (int p = 0; p < PT; ++p) {
                const float dd = (float)p - ud;
                wp[p] = __expf(-cd * dd * dd);
            }
        }
        for (int i = threadIdx.x; i < cnt; i += kFxThreads) {
            const float qs = ONES ? mult : mult * qf[sorted[k0 + i]];
            oq[i] = make_int2(fx_ext_off(rec_off[k0 + i], xs, ys), __float_as_int(qs));
        }
        __syncthreads();
        // one wavefront per bead: lanes over the (y, z) window plane, each walking x
        for (int i = wave; i < cnt; i += NW) {
            const int2 r = oq[i];
            const float qs = __int_as_float(r.y);
            const float* wi = w + i * WS;
            float wx[PT];
            if constexpr (PT % 4 == 0) {
#pragma unroll
                for (int p = 0; p < PT; p += 4) {
                    const float4 v4 = *reinterpret_cast<const float4*>(wi + p);
                    wx[p] = v4.x;
                    wx[p + 1] = v4.y;
                    wx[p + 2] = v4.z;
                    wx[p + 3] = v4.w;
                }
            } else {
#pragma unroll
                for (int p = 0; p < PT; ++p) wx[p] = wi[p];
            }
#pragma unroll
            for (int pair = lane; pair < PT * PT; pair += kFxWave) {
                const int py = pair / PT, pz = pair - py * PT;
                const float c = qs * wi[PT + py] * wi[2 * PT + pz];
                unsigned* col = ext + r.x + py * ys + pz;
#pragma unroll
                for (int px = 0; px < PT; ++px) atomicAdd(col + px * xs, (unsigned)round_fx(c * wx[px]));
            }
        }
        __syncthreads();
    }
    if constexpr (ONES) {
        // sums of positive weights: the unsigned maximum is the largest node
        unsigned m = 0u;
        for (int k = threadIdx.x; k < esz; k += kFxThreads) m = max(m, ext[k]);
#pragma unroll
        for (int off = kFxWave / 2; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, kFxWave));
        if (lane == 0) atomicMax(&tile_max, m);
        __syncthreads();
        if (threadIdx.x == 0) {
            // each node sums at most (ke - kb) rounded products, each off by <= 1/2
            const double bound = ((double)tile_max + 0.5 * (double)(ke - kb)) / (double)mult * (1.0 + 1e-5);
            atomicMax(&fs->wbits, __float_as_uint((float)bound));
        }
        return;
    }
    // the extended tile leaves as one contiguous block of the scratch (plain 16-byte stores,
    // rows padded as in LDS); k_fx_gather sums the blocks that cover each mesh node
    int* blk = blocks + (long)t * bstride;
    for (int k = threadIdx.x; k < esz / 4; k += kFxThreads)
        reinterpret_cast<uint4*>(blk)[k] = reinterpret_cast<const uint4*>(ext)[k];
    for (int k = (esz / 4) * 4 + threadIdx.x; k < esz; k += kFxThreads) blk[k] = (int)ext[k];
}

// Persistent forms (compile-time tile shapes).  A workgroup walks `per` tiles: with per > 1
// and ntiles a multiple of 8 per, workgroup b (dealt to XCD b mod 8) takes the consecutive
// tiles (b mod 8) ntiles / 8 + (b / 8) per + k, so each XCD still holds a contiguous run.
__device__ __forceinline__ int fx_run_tile(int b, int k, int per, int nt)
{
    if (per == 1) return fx_tile(b, nt);
    return (b % 8) * (nt / 8) + (b / 8) * per + k;
}

// A tile's first records, one bead per thread slot: thread j holds the window record of
// bead j / 3 (axis j % 3 of its weights, j < 3 min(n, npf)) and the offset and bead index
// of bead j (j < min(n, npf)).
struct FxPrefetch {
    float4 u;
    int ro;
    unsigned si;
};

__device__ __forceinline__ FxPrefetch fx_prefetch(unsigned kb, unsigned ke, int npf, const int* __restrict__ rec_off,
                                                  const unsigned* __restrict__ sorted, const float4* __restrict__ rec_u)
{
    FxPrefetch p{make_float4(0.f, 0.f, 0.f, 0.f), 0, 0u};
    const int m = (int)min(ke - kb, (unsigned)npf), j = threadIdx.x;
    if (j < 3 * m) p.u = rec_u[kb + j / 3];
    if (j < m) {
        p.ro = rec_off[kb + j];
        p.si = sorted[kb + j];
    }
    return p;
}

// Staged weights of bead slot i along axis d (3 threads per bead): exp(-c_d (p - u_d)^2).
template <int PT, int WS>
__device__ __forceinline__ void fx_stage_weights(float* __restrict__ w, int i, int d, const float4& u4, float c0,
                                                 float c1, float c2)
{
    const float ud = d == 0 ? u4.x : (d == 1 ? u4.y : u4.z);
    const float cd = d == 0 ? c0 : (d == 1 ? c1 : c2);
    float* wp = w + i * WS + d * PT;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        const float dd = (float)p - ud;
        wp[p] = __expf(-cd * dd * dd);
    }
}

// The spread of spread_fx_tile (ONES = false) as a persistent workgroup over `per` tiles,
// consecutive along z in one (x, y) column of tiles.  While one tile accumulates, the
// registers already hold the next tile's records (and, by the write-out, its charges), so
// that the staging's global-memory round trips overlap the LDS work instead of stalling
// each tile.  The first `NR` rounds of a tile (NR ch beads, 3 threads per bead for the
// weights) come from registers, later rounds from memory.
//
// The +z halo of a tile (its extended nodes z = TZ .. TZ + P - 2) is the low z range of the
// next tile of the run, so it stays in LDS: after a tile's write-out the halo rows move to
// the front of the accumulator and the rest is cleared, and the next tile adds onto them.
// A tile therefore leaves only its TZ z-planes (an EX x EY x TZ block); the last tile of a
// run also leaves its halo (an EX x EY x HS block per run, HS = P - 1 rounded to 4), which
// k_fx_gather_zc adds to the first tile of the next run.  Integer sums: the same mesh bits
// as the per-tile kernel.
template <int PT, int TX, int TY, int TZ>
struct FxZc {
    static constexpr int EX = TX + PT - 1, EY = TY + PT - 1, EZ = TZ + PT - 1;
    static constexpr int ys = (EZ + 1) & ~1;       // the LDS tile's row pitch
    static constexpr int H = PT - 1, HS = (H + 3) & ~3;
    // rows of 16-byte groups, and a halo row that ends inside the zero padding of an LDS row
    static constexpr bool ok = TZ % 4 == 0 && ys % 4 == 0 && TZ + HS <= ys && HS <= TZ;
    static constexpr long interior = (long)EX * EY * TZ, halo = (long)EX * EY * HS;
};

template <int PT, int TX, int TY, int TZ>
__global__ __launch_bounds__(kFxThreads) void k_spread_fx_run(MeshParams mp, TileGeom tg, int ch, int per,
                                                              const unsigned* __restrict__ tstart,
                                                              const unsigned* __restrict__ sorted,
                                                              const int* __restrict__ rec_off,
                                                              const float4* __restrict__ rec_u,
                                                              const float* __restrict__ qf,
                                                              FxScale* __restrict__ fs, int nparts, double prefac,
                                                              int* __restrict__ blocks)
{
    constexpr int WS = fx_wstride(PT);
    constexpr int EX = TX + PT - 1, EY = TY + PT - 1, EZ = TZ + PT - 1;
    constexpr int ys = (EZ + 1) & ~1, xs = EY * ys, esz = EX * xs;
    constexpr int NW = kFxThreads / kFxWave;
    extern __shared__ __align__(16) char smem[];
    unsigned* ext = reinterpret_cast<unsigned*>(smem);
    float* w = reinterpret_cast<float*>(smem + ((size_t)esz * 4 + 15) / 16 * 16);
    int2* oq = reinterpret_cast<int2*>(w + (size_t)ch * WS);
    const int j = threadIdx.x, lane = j & (kFxWave - 1), wave = fx_wave();
    const int NR = max(1, (kFxThreads / 3) / ch), npf = NR * ch;
    const int nt = tg.ntiles;
    const float c0 = (float)(mp.xiterm * mp.h[0] * mp.h[0] / mp.eta[0]);
    const float c1 = (float)(mp.xiterm * mp.h[1] * mp.h[1] / mp.eta[1]);
    const float c2 = (float)(mp.xiterm * mp.h[2] * mp.h[2] / mp.eta[2]);
    using Zc = FxZc<PT, TX, TY, TZ>;
    static_assert(Zc::ok, "tile shape without a z-carried halo");
    constexpr int rows = EX * EY;
    int* hblocks = blocks + (long)nt * Zc::interior;  // one halo block per run, after the tile blocks
    int t = fx_run_tile(blockIdx.x, 0, per, nt);
    unsigned kb = tstart[t], ke = tstart[t + 1];
    FxPrefetch cur = fx_prefetch(kb, ke, npf, rec_off, sorted, rec_u);
    float qv = (j < (int)min(ke - kb, (unsigned)npf)) ? qf[cur.si] : 0.0f;
    for (int q = j; q < esz / 4; q += kFxThreads) reinterpret_cast<uint4*>(ext)[q] = make_uint4(0, 0, 0, 0);
    // the charge scale from the partial maxima (as k_fx_scale; workgroup 0 stores it for the
    // gather), or, nparts == 0, the one already in fs
    float mult;
    if (nparts > 0) {
        __shared__ float qw[NW];
        float m = j < nparts ? fs->part[j] : 0.0f;
        m = wave_nanmax(m);
        if (lane == 0) qw[wave] = m;
        __syncthreads();
        float qm = qw[0];
        for (int k = 1; k < NW; ++k) qm = nanmax(qm, qw[k]);
        float inv;
        fx_scale_values(qm, fs->wbits, prefac, mult, inv);
        if (blockIdx.x == 0 && j == 0) {
            fs->mult = mult;
            fs->inv = inv;
            fs->qmax = qm;
        }
    } else {
        mult = fs->mult;
    }
    for (int k = 0; k < per; ++k) {
        const bool more = k + 1 < per;
        const int tn = more ? fx_run_tile(blockIdx.x, k + 1, per, nt) : t;
        const unsigned kbn = more ? tstart[tn] : 0u, ken = more ? tstart[tn + 1] : 0u;
        const FxPrefetch nxt = fx_prefetch(kbn, ken, npf, rec_off, sorted, rec_u);
        const unsigned n = ke - kb;
        int r = 0;
        for (unsigned r0 = 0; r0 < n; r0 += ch, ++r) {
            const int cnt = (int)min((unsigned)ch, n - r0);
            if (r > 0) __syncthreads();  // the previous round's beads are added
            if (r < NR) {
                const int jw = j - 3 * r * ch, jo = j - r * ch;
                if (jw >= 0 && jw < 3 * cnt) fx_stage_weights<PT, WS>(w, jw / 3, jw - 3 * (jw / 3), cur.u, c0, c1, c2);
                if (jo >= 0 && jo < cnt) oq[jo] = make_int2(fx_ext_off(cur.ro, xs, ys), __float_as_int(mult * qv));
            } else {
                for (int jj = j; jj < 3 * cnt; jj += kFxThreads)
                    fx_stage_weights<PT, WS>(w, jj / 3, jj - 3 * (jj / 3), rec_u[kb + r0 + jj / 3], c0, c1, c2);
                for (int jj = j; jj < cnt; jj += kFxThreads)
                    oq[jj] = make_int2(fx_ext_off(rec_off[kb + r0 + jj], xs, ys),
                                       __float_as_int(mult * qf[sorted[kb + r0 + jj]]));
            }
            __syncthreads();
            for (int i = wave; i < cnt; i += NW) {
                const int2 rr = oq[i];
                const float qs = __int_as_float(rr.y);
                const float* wi = w + i * WS;
                float wx[PT];
                if constexpr (PT % 4 == 0) {
#pragma unroll
                    for (int p = 0; p < PT; p += 4) {
                        const float4 v4 = *reinterpret_cast<const float4*>(wi + p);
                        wx[p] = v4.x;
                        wx[p + 1] = v4.y;
                        wx[p + 2] = v4.z;
                        wx[p + 3] = v4.w;
                    }
                } else {
#pragma unroll
                    for (int p = 0; p < PT; ++p) wx[p] = wi[p];
                }
#pragma unroll
                for (int pair = lane; pair < PT * PT; pair += kFxWave) {
                    const int py = pair / PT, pz = pair - py * PT;
                    const float c = qs * wi[PT + py] * wi[2 * PT + pz];
                    unsigned* col = ext + rr.x + py * ys + pz;
#pragma unroll
                    for (int px = 0; px < PT; ++px) atomicAdd(col + px * xs, (unsigned)round_fx(c * wx[px]));
                }
            }
        }
        // the next tile's charges (its bead indices have arrived), in flight over the write-out
        const float qvn = (more && j < (int)min(ken - kbn, (unsigned)npf)) ? qf[nxt.si] : 0.0f;
        __syncthreads();
        // the tile's TZ z-planes (its carried-in halo already added), rows of TZ ints
        int* blk = blocks + (long)t * Zc::interior;
        for (int q = j; q < rows * (TZ / 4); q += kFxThreads) {
            const int row = q / (TZ / 4), c4 = q - row * (TZ / 4);
            reinterpret_cast<uint4*>(blk + (long)row * TZ)[c4] = reinterpret_cast<const uint4*>(ext + row * ys)[c4];
        }
        if (!more) {
            // the run's last halo, for the first tile of the next run along z
            int* hb = hblocks + (long)(t / per) * Zc::halo;
            for (int q = j; q < rows * (Zc::HS / 4); q += kFxThreads) {
                const int row = q / (Zc::HS / 4), c4 = q - row * (Zc::HS / 4);
                reinterpret_cast<uint4*>(hb + (long)row * Zc::HS)[c4] =
                    reinterpret_cast<const uint4*>(ext + row * ys + TZ)[c4];
            }
            break;
        }
        __syncthreads();
        // carry: each row's halo (and the zero padding after it) to the front, the rest cleared
        for (int row = j; row < rows; row += kFxThreads) {
            uint4* r4 = reinterpret_cast<uint4*>(ext + row * ys);
#pragma unroll
            for (int c4 = 0; c4 < Zc::HS / 4; ++c4) r4[c4] = r4[TZ / 4 + c4];
#pragma unroll
            for (int c4 = Zc::HS / 4; c4 < ys / 4; ++c4) r4[c4] = make_uint4(0, 0, 0, 0);
        }
        t = tn;
        kb = kbn;
        ke = ken;
        cur = nxt;
        qv = qvn;
    }
}

template <int PT, int TX, int TY, int TZ, bool ONES>
__global__ __launch_bounds__(kFxThreads) void k_spread_fx(MeshParams mp, TileGeom tg, int ch,
                                                          const unsigned* __restrict__ tstart,
                                                          const unsigned* __restrict__ sorted,
                                                          const int* __restrict__ rec_off,
                                                          const float4* __restrict__ rec_u,
                                                          const float* __restrict__ qf, FxScale* __restrict__ fs,
                                                          long bstride, int* __restrict__ blocks)
{
    spread_fx_tile<PT, TX, TY, TZ, ONES>(mp, tg, fx_tile(blockIdx.x, tg.ntiles), ch, tstart, sorted, rec_off, rec_u,
                                         qf, fs, bstride, blocks);
}

// Batched boxes: tiles of all boxes in one launch; a box owns the global tiles
// [tile0, tile0 + ntiles) (ascending with the box), found by binary search.
__device__ __forceinline__ int fx_box_of(const FxBox* __restrict__ boxes, int nbox, int g, bool by_tile)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((by_tile ? boxes[mid].tile0 : boxes[mid].gblk0) <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int PT, bool ONES>
__global__ __launch_bounds__(kFxThreads) void k_spread_fx_batch(const FxBox* __restrict__ boxes, int nbox, int total,
                                                                const int* __restrict__ skip)
{
    const int g = fx_tile(blockIdx.x, total);
    const FxBox& d = boxes[fx_box_of(boxes, nbox, g, true)];
    if (skip && skip[d.box]) return;
    spread_fx_tile<PT, 0, 0, 0, ONES>(d.mp, d.tg, g - d.tile0, d.ch, d.tile_start, d.sorted, d.rec_off, d.rec_u,
                                      d.qf, d.fs, d.bstride, d.blocks);
}

// The mesh from the tiles' blocks: a node sums the blocks that cover it -- its own tile's
// and, within P - 1 nodes of a tile's low face, the previous tile's along that axis (two at
// most per axis, as P - 1 <= the tile edge) -- and leaves as float, int sum x inv.  Integer
// sums do not depend on their order, so the mesh is the same as from atomic additions.
// Every node is written, so the mesh needs no clearing.  V = 4: a thread takes 4
// consecutive z nodes of one tile (tile edge, block row pitch and Nz multiples of 4; the
// row padding past EZ holds zeros) with 16-byte loads, all of them issued together.
template <int V>
__device__ __forceinline__ void gather_fx_nodes(const MeshParams& mp, const TileGeom& tg, long bstride,
                                                const int* __restrict__ blocks, const FxScale* __restrict__ fs,
                                                float* __restrict__ mesh, long gid)
{
    const int T0 = tg.T[0], T1 = tg.T[1], T2 = tg.T[2];
    const int EX = T0 + mp.P - 1, EY = T1 + mp.P - 1, EZ = T2 + mp.P - 1;
    const int ys = (EZ + 1) & ~1;
    const int nq = mp.Nz / V;
    const long rowid = gid / nq;
    if (rowid >= (long)mp.Nx * mp.Ny) return;
    const int gz = (int)(gid - rowid * nq) * V;
    const int gy = (int)(rowid % mp.Ny), gx = (int)(rowid / mp.Ny);
    const int tx = gx / T0, lx = gx - tx * T0, ty = gy / T1, ly = gy - ty * T1, tz = gz / T2, lz = gz - tz * T2;
    const int px = tx == 0 ? tg.nt[0] - 1 : tx - 1, py = ty == 0 ? tg.nt[1] - 1 : ty - 1;
    const int pz = tz == 0 ? tg.nt[2] - 1 : tz - 1;
    const bool bx = lx + T0 < EX, by = ly + T1 < EY, bz = lz + T2 < EZ;
    using IV = typename std::conditional<V == 4, int4, int>::type;
    IV v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
        const bool on = (!dx || bx) && (!dy || by) && (!dz || bz);
        const long u = ((long)(dx ? px : tx) * tg.nt[1] + (dy ? py : ty)) * tg.nt[2] + (dz ? pz : tz);
        const long off = u * bstride + ((long)(lx + dx * T0) * EY + (ly + dy * T1)) * ys + lz + dz * T2;
        if constexpr (V == 4)
            v[c] = on ? *reinterpret_cast<const int4*>(blocks + off) : make_int4(0, 0, 0, 0);
        else
            v[c] = on ? blocks[off] : 0;
    }
    const float inv = fs->inv;
    float* out = mesh + rowid * mp.Nz + gz;
    if constexpr (V == 4) {
        int4 a = v[0];
#pragma unroll
        for (int c = 1; c < 8; ++c) {
            a.x += v[c].x;
            a.y += v[c].y;
            a.z += v[c].z;
            a.w += v[c].w;
        }
        *reinterpret_cast<float4*>(out) = make_float4((float)a.x * inv, (float)a.y * inv, (float)a.z * inv, (float)a.w * inv);
    } else {
        int a = v[0];
#pragma unroll
        for (int c = 1; c < 8; ++c) a += v[c];
        out[0] = (float)a * inv;
    }
}

template <int V>
__global__ __launch_bounds__(256) void k_fx_gather(MeshParams mp, TileGeom tg, long bstride,
                                                   const int* __restrict__ blocks, const FxScale* __restrict__ fs,
                                                   float* __restrict__ mesh)
{
    gather_fx_nodes<V>(mp, tg, bstride, blocks, fs, mesh, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Batched boxes: a box owns the workgroups [gblk0, gblk0 + its count) of the launch.
template <int V>
__global__ __launch_bounds__(256) void k_fx_gather_batch(const FxBox* __restrict__ boxes, int nbox,
                                                         const int* __restrict__ skip)
{
    const FxBox& d = boxes[fx_box_of(boxes, nbox, (int)blockIdx.x, false)];
    if (skip && skip[d.box]) return;
    gather_fx_nodes<V>(d.mp, d.tg, d.bstride, d.blocks, d.fs, d.grid,
                       (long)(blockIdx.x - d.gblk0) * blockDim.x + threadIdx.x);
}

// The mesh from k_spread_fx_run's z-carried blocks: a node sums its own tile's block and,
// within P - 1 nodes of the tile's low x and y faces, the previous tiles' blocks along x and
// y (at most 4 blocks; a tile's z halo inside a run was added in LDS).  At the first tile of
// a run (tz a multiple of per), within P - 1 nodes of the low z face, the previous run's halo
// blocks as well (4 more at most).  A thread takes 4 consecutive z nodes with 16-byte loads;
// the halo rows' last ints past P - 1 hold zeros.  Integer sums, so the mesh bits are those
// of k_fx_gather over the per-tile blocks.
template <int PT, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_fx_gather_zc(MeshParams mp, TileGeom tg, int per_log2,
                                                      const int* __restrict__ blocks,
                                                      const FxScale* __restrict__ fs, float* __restrict__ mesh)
{
    using Zc = FxZc<PT, TX, TY, TZ>;
    constexpr int EY = TY + PT - 1;
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nq = mp.Nz >> 2;
    const long rowid = gid / nq;
    if (rowid >= (long)mp.Nx * mp.Ny) return;
    const int gz = (int)(gid - rowid * nq) * 4;
    const int gy = (int)(rowid % mp.Ny), gx = (int)(rowid / mp.Ny);
    const int tx = gx / TX, lx = gx - tx * TX, ty = gy / TY, ly = gy - ty * TY, tz = gz / TZ, lz = gz - tz * TZ;
    const int px = tx == 0 ? tg.nt[0] - 1 : tx - 1, py = ty == 0 ? tg.nt[1] - 1 : ty - 1;
    const int pz = tz == 0 ? tg.nt[2] - 1 : tz - 1;
    const bool bx = lx < PT - 1, by = ly < PT - 1;
    const bool bz = ((tz & ((1 << per_log2) - 1)) == 0) && lz < Zc::HS;
    const int* hblocks = blocks + (long)tg.ntiles * Zc::interior;
    int4 v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
        const bool on = (!dx || bx) && (!dy || by) && (!dz || bz);
        const long u = ((long)(dx ? px : tx) * tg.nt[1] + (dy ? py : ty)) * tg.nt[2] + (dz ? pz : tz);
        const long row = (long)(lx + dx * TX) * EY + (ly + dy * TY);
        const int* src = dz ? hblocks + (u >> per_log2) * Zc::halo + row * Zc::HS + lz
                            : blocks + u * Zc::interior + row * TZ + lz;
        v[c] = on ? *reinterpret_cast<const int4*>(src) : make_int4(0, 0, 0, 0);
    }
    int4 a = v[0];
#pragma unroll
    for (int c = 1; c < 8; ++c) {
        a.x += v[c].x;
        a.y += v[c].y;
        a.z += v[c].z;
        a.w += v[c].w;
    }
    const float inv = fs->inv;
    *reinterpret_cast<float4*>(mesh + rowid * mp.Nz + gz) =
        make_float4((float)a.x * inv, (float)a.y * inv, (float)a.z * inv, (float)a.w * inv);
}

// FP32-bulk potential contraction (cppmodule/ConductorRigid.cu:260-360) on the same
// binned tiles: phi_i = h^3 prefac sum_window w(node) grid(node).  The extended tile is
// read into LDS once (float4 rows where the tile does not wrap in z), the chunk's axis
// weights are staged as in the spread, and each wavefront reduces one bead's window
// (lanes over the (y, z) plane, each walking x with compile-time LDS offsets).
template <int PT, int TX, int TY, int TZ, bool FORCE>
__device__ __forceinline__ void contract_fx_tile(const MeshParams& mp, const TileGeom& tg, int t, int ch,
                                                 const unsigned* __restrict__ tstart,
                                                 const unsigned* __restrict__ sorted, const int* __restrict__ rec_off,
                                                 const float4* __restrict__ rec_u, const float* __restrict__ grid,
                                                 float* __restrict__ phi, const double* __restrict__ q,
                                                 double* __restrict__ force, const double* __restrict__ fadd)
{
    constexpr bool CT = TX > 0;
    constexpr int WS = fx_wstride(PT);
    const int T0 = CT ? TX : tg.T[0], T1 = CT ? TY : tg.T[1], T2 = CT ? TZ : tg.T[2];
    const int EX = T0 + PT - 1, EY = T1 + PT - 1, EZ = T2 + PT - 1;
    const int ys = (EZ + 3) & ~3, xs = EY * ys, esz = EX * xs;
    extern __shared__ __align__(16) char smem[];
    const unsigned kb = tstart[t], ke = tstart[t + 1];
    if (kb == ke) return;
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int x0 = tx * T0, y0 = ty * T1, z0 = tz * T2;
    float* ext = reinterpret_cast<float*>(smem);
    float* w = reinterpret_cast<float*>(smem + ((size_t)esz * 4 + 15) / 16 * 16);
    int2* oi = reinterpret_cast<int2*>(w + (size_t)ch * WS);
    const long plane = (long)mp.Ny * mp.Nz;
    const bool dma = T2 % 4 == 0 && mp.Nz % 4 == 0;
    if (dma) {
        // rows of the tile along z as ys / 4 float4 groups each (the last group may run past
        // the extension into the row's padding; the tile's z start and Nz are multiples of 4,
        // so a group wraps past Nz as a whole), loaded straight into LDS by
        // global_load_lds_dwordx4: the 64 lanes of a wavefront fill 64 consecutive 16-byte
        // slots, so item k lands at float4 slot k of the tile without a register round trip
        const int per_row = ys / 4, items = EX * EY * per_row;
        const int wbase = threadIdx.x & ~(kFxWave - 1);
        for (int k0 = 0; k0 < items; k0 += kFxThreads) {
            const int k = k0 + (int)threadIdx.x;
            if (k < items) {
                const int r = k / per_row, c = k - r * per_row;
                const int ly = r % EY, lx = r / EY;
                int gx = x0 + lx, gy = y0 + ly, gz = z0 + 4 * c;
                gx -= gx >= mp.Nx ? mp.Nx : 0;
                gy -= gy >= mp.Ny ? mp.Ny : 0;
                gz -= gz >= mp.Nz ? mp.Nz : 0;
                __builtin_amdgcn_global_load_lds(grid + (long)gx * plane + (long)gy * mp.Nz + gz,
                                                 ext + 4 * (k0 + wbase), 16, 0, 0);
            }
        }
    } else {
        const int nodes = EX * EY * EZ;
        FxWalk bw(threadIdx.x, kFxThreads, EY, EZ);
        for (int k = threadIdx.x; k < nodes; k += kFxThreads, bw.next()) {
            int gx = x0 + bw.lx, gy = y0 + bw.ly, gz = z0 + bw.lz;
            gx -= gx >= mp.Nx ? mp.Nx : 0;
            gy -= gy >= mp.Ny ? mp.Ny : 0;
            gz -= gz >= mp.Nz ? mp.Nz : 0;
            ext[bw.lx * xs + bw.ly * ys + bw.lz] = grid[(long)gx * plane + (long)gy * mp.Nz + gz];
        }
    }
    const float c0 = (float)(mp.xiterm * mp.h[0] * mp.h[0] / mp.eta[0]);
    const float c1 = (float)(mp.xiterm * mp.h[1] * mp.h[1] / mp.eta[1]);
    const float c2 = (float)(mp.xiterm * mp.h[2] * mp.h[2] / mp.eta[2]);
    const float pc = (float)(mp.quadW * mp.prefac);
    const int lane = threadIdx.x & (kFxWave - 1), wave = fx_wave();
    constexpr int NW = kFxThreads / kFxWave;
    for (unsigned k0 = kb; k0 < ke; k0 += ch) {
        const int cnt = (int)min((unsigned)ch, ke - k0);
        for (int j = threadIdx.x; j < 3 * cnt; j += kFxThreads) {
            const int i = j / 3, d = j - 3 * i;
            const float4 u4 = rec_u[k0 + i];
            const float ud = d == 0 ? u4.x : (d == 1 ? u4.y : u4.z);
            const float cd = d == 0 ? c0 : (d == 1 ? c1 : c2);
            float* wp = w + i * WS + d * PT;
#pragma unroll
            for (int p = 0; p < PT; ++p) {
                const float dd = (float)p - ud;
                wp[p] = __expf(-cd * dd * dd);
            }
        }
        for (int i = threadIdx.x; i < cnt; i += kFxThreads)
            oi[i] = make_int2(fx_ext_off(rec_off[k0 + i], xs, ys), (int)sorted[k0 + i]);
        if (dma && k0 == kb) __builtin_amdgcn_s_waitcnt(0);  // this wavefront's tile loads have landed
        __syncthreads();  // (the first one also orders the extended tile's stores)
        // A bead's totals land in lane 63, which leaves them in LDS (the potential in its
        // record's offset slot, the force sums in its first weight slots, both read before);
        // the chunk's results then go out together.
        for (int i = wave; i < cnt; i += NW) {
            const int2 r = oi[i];
            float* wi = w + i * WS;
            float wx[PT];
            if constexpr (PT % 4 == 0) {
#pragma unroll
                for (int p = 0; p < PT; p += 4) {
                    const float4 v4 = *reinterpret_cast<const float4*>(wi + p);
                    wx[p] = v4.x;
                    wx[p + 1] = v4.y;
                    wx[p + 2] = v4.z;
                    wx[p + 3] = v4.w;
                }
            } else {
#pragma unroll
                for (int p = 0; p < PT; ++p) wx[p] = wi[p];
            }
            if constexpr (!FORCE) {
                float acc = 0.0f;
#pragma unroll
                for (int pair = lane; pair < PT * PT; pair += kFxWave) {
                    const int py = pair / PT, pz = pair - py * PT;
                    const float* col = ext + r.x + py * ys + pz;
                    float sx = 0.0f;
#pragma unroll
                    for (int px = 0; px < PT; ++px) sx += wx[px] * col[px * xs];
                    acc += wi[PT + py] * wi[2 * PT + pz] * sx;
                }
                acc = fx_total63(acc);
                if (lane == kFxWave - 1) oi[i].x = __float_as_int(acc);
            } else {
                // gradient of the bead's window: node offsets r = (u - p) h per axis
                const float4 u4 = rec_u[k0 + i];
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
                for (int pair = lane; pair < PT * PT; pair += kFxWave) {
                    const int py = pair / PT, pz = pair - py * PT;
                    const float* col = ext + r.x + py * ys + pz;
                    float sx = 0.0f, sxr = 0.0f;
#pragma unroll
                    for (int px = 0; px < PT; ++px) {
                        const float wg = wx[px] * col[px * xs];
                        sx += wg;
                        sxr += wg * (u4.x - (float)px);
                    }
                    const float wyz = wi[PT + py] * wi[2 * PT + pz];
                    fx += wyz * sxr;
                    fy += wyz * sx * (u4.y - (float)py);
                    fz += wyz * sx * (u4.z - (float)pz);
                }
                fx = fx_total63(fx);
                fy = fx_total63(fy);
                fz = fx_total63(fz);
                if (lane == kFxWave - 1) {
                    wi[0] = fx;
                    wi[1] = fy;
                    wi[2] = fz;
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += kFxThreads) {
            const int2 r = oi[i];
            if constexpr (!FORCE) {
                phi[r.y] = __int_as_float(r.x) * pc;
            } else {
                // F = 2 (h^3 prefac) (2 xi^2) q_i sum_nodes grid w (r / eta), the reference's
                // contract_force (cppmodule/ConductorRigid.cu:449)
                const float* wi = w + i * WS;
                const double c = 2.0 * (mp.quadW * mp.prefac) * mp.xiterm * (double)(float)q[r.y];
                const long o = 4 * (long)r.y;
                // fadd: the real-space force of the same bead, computed beside this pass
                const double a0 = fadd ? fadd[o] : 0.0, a1 = fadd ? fadd[o + 1] : 0.0;
                const double a2 = fadd ? fadd[o + 2] : 0.0;
                force[o + 0] = c * mp.h[0] / mp.eta[0] * (double)wi[0] + a0;
                force[o + 1] = c * mp.h[1] / mp.eta[1] * (double)wi[1] + a1;
                force[o + 2] = c * mp.h[2] / mp.eta[2] * (double)wi[2] + a2;
                force[o + 3] = 0.0;
            }
        }
        __syncthreads();
    }
}

template <int PT, int TX, int TY, int TZ, bool FORCE>
__global__ __launch_bounds__(kFxThreads) void k_contract_fx(MeshParams mp, TileGeom tg, int ch,
                                                            const unsigned* __restrict__ tstart,
                                                            const unsigned* __restrict__ sorted,
                                                            const int* __restrict__ rec_off,
                                                            const float4* __restrict__ rec_u,
                                                            const float* __restrict__ grid, float* __restrict__ phi,
                                                            const double* __restrict__ q, double* __restrict__ force,
                                                            const double* __restrict__ fadd)
{
    contract_fx_tile<PT, TX, TY, TZ, FORCE>(mp, tg, fx_tile(blockIdx.x, tg.ntiles), ch, tstart, sorted, rec_off, rec_u,
                                            grid, phi, q, force, fadd);
}

template <int PT, bool FORCE>
__global__ __launch_bounds__(kFxThreads) void k_contract_fx_batch(const FxBox* __restrict__ boxes, int nbox, int total,
                                                                  const int* __restrict__ skip)
{
    const int g = fx_tile(blockIdx.x, total);
    const FxBox& d = boxes[fx_box_of(boxes, nbox, g, true)];
    if (skip && skip[d.box]) return;
    contract_fx_tile<PT, 0, 0, 0, FORCE>(d.mp, d.tg, g - d.tile0, d.cch, d.tile_start, d.sorted, d.rec_off, d.rec_u,
                                         d.grid, d.phi, d.q, d.force, nullptr);
}

void allow_lds_fx(const void* fn, size_t bytes)
{
    if (bytes <= 65536) return;
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> raised;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = raised[{dev, fn}];
    if (have >= bytes) return;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess) have = bytes;
}

// Tiles per persistent workgroup: the largest of 8, 4, 2 that divides the tiles into XCD
// runs and each column of tiles into whole runs along z, and still gives 512 workgroups (two
// per CU of an MI355X), else 1.
constexpr int kFxRunMaxPer = 8;
inline int fx_run_per(const TileGeom& tg)
{
    for (int per = kFxRunMaxPer; per > 1; per /= 2)
        if (tg.ntiles % (8 * per) == 0 && tg.nt[2] % per == 0 && tg.ntiles / per >= 512) return per;
    return 1;
}

// Compile-time tile shapes (ONES = false) spread with k_spread_fx_run and gather with
// k_fx_gather_zc here, and set *gathered; other shapes leave the per-tile blocks to the
// caller's k_fx_gather.
template <int PT, int TX, int TY, int TZ, bool ONES>
hipError_t launch_fx_kernel(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                            const float* qf, FxScale* fs, int nparts, long bstride, int* blocks, float* mesh,
                            bool* gathered, hipStream_t st)
{
    if constexpr (!ONES && TX > 0) {
        const int per = fx_run_per(tg);
        auto fr = k_spread_fx_run<PT, TX, TY, TZ>;
        allow_lds_fx(reinterpret_cast<const void*>(fr), pl.lds_bytes);
        hipLaunchKernelGGL(fr, dim3(tg.ntiles / per), dim3(kFxThreads), pl.lds_bytes, st, mp, tg, pl.chunk, per,
                           bv.tile_start, bv.sorted, bv.rec_off, static_cast<const float4*>(bv.rec_win), qf, fs,
                           nparts, mp.prefac, blocks);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const unsigned gb = (unsigned)fx_gather_blocks(mp, true);
        hipLaunchKernelGGL((k_fx_gather_zc<PT, TX, TY, TZ>), dim3(gb), dim3(256), 0, st, mp, tg, __builtin_ctz(per),
                           blocks, fs, mesh);
        *gathered = true;
        return hipGetLastError();
    }
    (void)mesh;
    (void)gathered;
    if (!ONES && nparts > 0) hipLaunchKernelGGL(k_fx_scale, dim3(1), dim3(kFxWave), 0, st, nparts, fs, mp.prefac);
    auto fn = k_spread_fx<PT, TX, TY, TZ, ONES>;
    allow_lds_fx(reinterpret_cast<const void*>(fn), pl.lds_bytes);
    hipLaunchKernelGGL(fn, dim3(tg.ntiles), dim3(kFxThreads), pl.lds_bytes, st, mp, tg, pl.chunk, bv.tile_start,
                       bv.sorted, bv.rec_off, static_cast<const float4*>(bv.rec_win), qf, fs, bstride, blocks);
    return hipGetLastError();
}

template <int PT, bool ONES>
hipError_t dispatch_shape(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                          const float* qf, FxScale* fs, int nparts, long bstride, int* blocks, float* mesh,
                          bool* gathered, hipStream_t st)
{
    const int a = tg.T[0], b = tg.T[1], c = tg.T[2];
#define CR_FX_SHAPE(X, Y, Z) \
    if (a == X && b == Y && c == Z)  \
    return launch_fx_kernel<PT, X, Y, Z, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered, st)
    if constexpr (PT == 8) {
        CR_FX_SHAPE(16, 16, 16);
        CR_FX_SHAPE(16, 8, 16);
        CR_FX_SHAPE(16, 16, 32);
        CR_FX_SHAPE(32, 16, 16);
        CR_FX_SHAPE(16, 32, 32);
        CR_FX_SHAPE(8, 8, 8);
    } else if constexpr (PT == 9) {
        CR_FX_SHAPE(16, 16, 16);
        CR_FX_SHAPE(16, 8, 16);
        CR_FX_SHAPE(8, 8, 8);
    }
#undef CR_FX_SHAPE
    return launch_fx_kernel<PT, 0, 0, 0, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered, st);
}

template <bool ONES>
hipError_t dispatch_p(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv, const float* qf,
                      FxScale* fs, int nparts, long bstride, int* blocks, float* mesh, bool* gathered,
                      hipStream_t st)
{
    switch (mp.P) {
        case 4: return dispatch_shape<4, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 5: return dispatch_shape<5, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 6: return dispatch_shape<6, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 7: return dispatch_shape<7, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 8: return dispatch_shape<8, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 9: return dispatch_shape<9, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        case 10: return dispatch_shape<10, ONES>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, gathered,
                                                  st);
        default: return hipErrorInvalidValue;
    }
}

template <int PT, int TX, int TY, int TZ, bool FORCE>
hipError_t launch_contract_kernel(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                                  const float* grid, float* phi, const double* q, double* force, const double* fadd,
                                  hipStream_t st)
{
    auto fn = k_contract_fx<PT, TX, TY, TZ, FORCE>;
    allow_lds_fx(reinterpret_cast<const void*>(fn), pl.lds_bytes);
    hipLaunchKernelGGL(fn, dim3(tg.ntiles), dim3(kFxThreads), pl.lds_bytes, st, mp, tg, pl.chunk, bv.tile_start,
                       bv.sorted, bv.rec_off, static_cast<const float4*>(bv.rec_win), grid, phi, q, force, fadd);
    return hipGetLastError();
}

template <int PT, bool FORCE>
hipError_t contract_shape(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                          const float* grid, float* phi, const double* q, double* force, const double* fadd,
                          hipStream_t st)
{
    const int a = tg.T[0], b = tg.T[1], c = tg.T[2];
#define CR_FXC_SHAPE(X, Y, Z) \
    if (a == X && b == Y && c == Z)  \
    return launch_contract_kernel<PT, X, Y, Z, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st)
    if constexpr (PT == 8) {
        CR_FXC_SHAPE(16, 16, 16);
        CR_FXC_SHAPE(16, 8, 16);
        CR_FXC_SHAPE(8, 16, 16);
        CR_FXC_SHAPE(16, 16, 32);
        CR_FXC_SHAPE(8, 8, 8);
    } else if constexpr (PT == 9) {
        CR_FXC_SHAPE(16, 16, 16);
        CR_FXC_SHAPE(16, 8, 16);
        CR_FXC_SHAPE(8, 8, 8);
    }
#undef CR_FXC_SHAPE
    return launch_contract_kernel<PT, 0, 0, 0, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
}

inline unsigned nblk(long n, int b)
{
    const long v = (n + b - 1) / b;
    return (unsigned)(v > 0 ? v : 1);
}

}  // namespace

FxPlan fx_plan(const MeshParams& mp, const TileGeom& tg, size_t lds_max)
{
    FxPlan p{};
    if (!tg.valid || mp.P < 4 || mp.P > 10) return p;
    const size_t EX = (size_t)tg.T[0] + mp.P - 1, EY = (size_t)tg.T[1] + mp.P - 1, EZ = (size_t)tg.T[2] + mp.P - 1;
    const size_t ext = (EX * EY * ((EZ + 1) & ~(size_t)1) * 4 + 15) / 16 * 16;
    const size_t per = (size_t)fx_wstride(mp.P) * 4 + 8;
    // two workgroups per CU when the tile leaves room for >= 64 staged beads in half the LDS
    for (const size_t budget : {lds_max / 2, lds_max}) {
        if (ext >= budget) continue;
        long ch = (long)((budget - ext) / per);
        if (ch > kFxChunk) ch = kFxChunk;
        ch -= ch % 4;
        if (ch < (budget == lds_max ? 16 : 64)) continue;
        p.chunk = (int)ch;
        p.lds_bytes = ext + (size_t)ch * per;
        p.valid = 1;
        return p;
    }
    return p;
}

FxPlan fx_contract_plan(const MeshParams& mp, const TileGeom& tg, size_t lds_max)
{
    FxPlan p{};
    if (!tg.valid || mp.P < 4 || mp.P > 10) return p;
    const size_t EX = (size_t)tg.T[0] + mp.P - 1, EY = (size_t)tg.T[1] + mp.P - 1, EZ = (size_t)tg.T[2] + mp.P - 1;
    const size_t ext = (EX * EY * ((EZ + 3) & ~(size_t)3) * 4 + 15) / 16 * 16;
    const size_t per = (size_t)fx_wstride(mp.P) * 4 + 8;
    for (const size_t budget : {lds_max / 2, lds_max}) {
        if (ext >= budget) continue;
        long ch = (long)((budget - ext) / per);
        if (ch > kFxChunk) ch = kFxChunk;
        ch -= ch % 4;
        if (ch < (budget == lds_max ? 16 : 64)) continue;
        p.chunk = (int)ch;
        p.lds_bytes = ext + (size_t)ch * per;
        p.valid = 1;
        return p;
    }
    return p;
}

template <bool FORCE>
hipError_t contract_fx_p(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                         const float* grid, float* phi, const double* q, double* force, const double* fadd,
                         hipStream_t st)
{
    if (!pl.valid) return hipErrorInvalidValue;
    switch (mp.P) {
        case 4: return contract_shape<4, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 5: return contract_shape<5, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 6: return contract_shape<6, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 7: return contract_shape<7, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 8: return contract_shape<8, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 9: return contract_shape<9, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        case 10: return contract_shape<10, FORCE>(mp, tg, pl, bv, grid, phi, q, force, fadd, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_contract_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                              const float* grid, float* phi, hipStream_t st)
{
    return contract_fx_p<false>(mp, tg, pl, bv, grid, phi, nullptr, nullptr, nullptr, st);
}

hipError_t launch_contract_force_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                                    const float* grid, const double* q, double* force, const double* fadd,
                                    hipStream_t st)
{
    return contract_fx_p<true>(mp, tg, pl, bv, grid, nullptr, q, force, fadd, st);
}

hipError_t launch_charges_fx(const double* x, int N, float* qf, FxScale* fs, double prefac, hipStream_t st)
{
    const unsigned blocks = std::min<unsigned>(kFxParts, nblk(std::max(N, 1), kFxChargeThreads));
    hipLaunchKernelGGL(k_charges_fx, dim3(blocks), dim3(kFxChargeThreads), 0, st, x, N, qf, fs);
    hipLaunchKernelGGL(k_fx_scale, dim3(1), dim3(kFxWave), 0, st, (int)blocks, fs, prefac);
    return hipGetLastError();
}

hipError_t launch_charges_fx_parts(const double* x, int N, float* qf, FxScale* fs, int* nparts, hipStream_t st)
{
    const unsigned blocks = std::min<unsigned>(kFxParts, nblk(std::max(N, 1), kFxChargeThreads));
    hipLaunchKernelGGL(k_charges_fx, dim3(blocks), dim3(kFxChargeThreads), 0, st, x, N, qf, fs);
    *nparts = (int)blocks;
    return hipGetLastError();
}

hipError_t launch_fx_scale(int nparts, FxScale* fs, double prefac, hipStream_t st)
{
    if (nparts < 1 || nparts > kFxParts) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_fx_scale, dim3(1), dim3(kFxWave), 0, st, nparts, fs, prefac);
    return hipGetLastError();
}

bool fx_gather_v4(const MeshParams& mp, const TileGeom& tg)
{
    const long ys = ((long)tg.T[2] + mp.P) & ~1L;  // block row pitch, (EZ + 1) & ~1
    return tg.T[2] % 4 == 0 && ys % 4 == 0 && mp.Nz % 4 == 0;
}

long fx_gather_blocks(const MeshParams& mp, bool v4)
{
    return (long)nblk((long)mp.Nx * mp.Ny * (mp.Nz / (v4 ? 4 : 1)), 256);
}

long fx_block_ints(const MeshParams& mp, const TileGeom& tg)
{
    const long EX = tg.T[0] + mp.P - 1, EY = tg.T[1] + mp.P - 1, EZ = tg.T[2] + mp.P - 1;
    return (EX * EY * ((EZ + 1) & ~1L) + 3) & ~3L;
}

hipError_t launch_spread_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                            const float* qf, FxScale* fs, int* blocks, float* mesh, hipStream_t st, int nparts)
{
    if (nparts < 0 || nparts > kFxParts) return hipErrorInvalidValue;
    if (!pl.valid) return hipErrorInvalidValue;
    for (int d = 0; d < 3; ++d)
        if (tg.T[d] < mp.P - 1) return hipErrorInvalidValue;  // at most two blocks per axis cover a node
    const long bstride = fx_block_ints(mp, tg);
    bool gathered = false;
    hipError_t e = dispatch_p<false>(mp, tg, pl, bv, qf, fs, nparts, bstride, blocks, mesh, &gathered, st);
    if (e != hipSuccess || gathered) return e;
    const bool v4 = fx_gather_v4(mp, tg);
    const unsigned gb = (unsigned)fx_gather_blocks(mp, v4);
    if (v4)
        hipLaunchKernelGGL(k_fx_gather<4>, dim3(gb), dim3(256), 0, st, mp, tg, bstride, blocks, fs, mesh);
    else
        hipLaunchKernelGGL(k_fx_gather<1>, dim3(gb), dim3(256), 0, st, mp, tg, bstride, blocks, fs, mesh);
    return hipGetLastError();
}

hipError_t launch_spread_fx_bound(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                                  FxScale* fs, hipStream_t st)
{
    if (!pl.valid) return hipErrorInvalidValue;
    hipError_t e = launch_fill<unsigned>(&fs->wbits, 0u, 1, st);
    if (e != hipSuccess) return e;
    bool gathered = false;
    return dispatch_p<true>(mp, tg, pl, bv, nullptr, fs, 0, 0, nullptr, nullptr, &gathered, st);
}

hipError_t launch_charges_fx_batch(const FxBox* boxes, int nbox, int max_n, const int* skip, hipStream_t st)
{
    if (nbox <= 0) return hipSuccess;
    const int nparts = (int)std::min<unsigned>(kFxParts, nblk(std::max(max_n, 1), kFxChargeThreads));
    hipLaunchKernelGGL(k_charges_fx_batch, dim3(nparts, nbox), dim3(kFxChargeThreads), 0, st, boxes, skip);
    hipLaunchKernelGGL(k_fx_scale_batch, dim3(nbox), dim3(kFxWave), 0, st, boxes, nparts, skip);
    return hipGetLastError();
}

namespace {
template <int PT, bool ONES>
hipError_t spread_fx_batch_p(const FxBox* boxes, int nbox, int total_tiles, size_t lds, const int* skip,
                             hipStream_t st)
{
    auto fn = k_spread_fx_batch<PT, ONES>;
    allow_lds_fx(reinterpret_cast<const void*>(fn), lds);
    hipLaunchKernelGGL(fn, dim3(total_tiles), dim3(kFxThreads), lds, st, boxes, nbox, total_tiles, skip);
    return hipGetLastError();
}

template <int PT>
hipError_t contract_fx_batch_p(const FxBox* boxes, int nbox, int total_tiles, size_t lds, bool force,
                               const int* skip, hipStream_t st)
{
    if (force) {
        auto fn = k_contract_fx_batch<PT, true>;
        allow_lds_fx(reinterpret_cast<const void*>(fn), lds);
        hipLaunchKernelGGL(fn, dim3(total_tiles), dim3(kFxThreads), lds, st, boxes, nbox, total_tiles, skip);
    } else {
        auto fn = k_contract_fx_batch<PT, false>;
        allow_lds_fx(reinterpret_cast<const void*>(fn), lds);
        hipLaunchKernelGGL(fn, dim3(total_tiles), dim3(kFxThreads), lds, st, boxes, nbox, total_tiles, skip);
    }
    return hipGetLastError();
}
}  // namespace

hipError_t launch_spread_fx_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds, long gather_blocks,
                                  bool v4, const int* skip, hipStream_t st)
{
    if (nbox <= 0 || total_tiles <= 0) return hipSuccess;
    hipError_t e;
    switch (P) {
        case 4: e = spread_fx_batch_p<4, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 5: e = spread_fx_batch_p<5, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 6: e = spread_fx_batch_p<6, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 7: e = spread_fx_batch_p<7, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 8: e = spread_fx_batch_p<8, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 9: e = spread_fx_batch_p<9, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        case 10: e = spread_fx_batch_p<10, false>(boxes, nbox, total_tiles, lds, skip, st); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (v4)
        hipLaunchKernelGGL(k_fx_gather_batch<4>, dim3((unsigned)gather_blocks), dim3(256), 0, st, boxes, nbox, skip);
    else
        hipLaunchKernelGGL(k_fx_gather_batch<1>, dim3((unsigned)gather_blocks), dim3(256), 0, st, boxes, nbox, skip);
    return hipGetLastError();
}

hipError_t launch_contract_fx_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds, bool force,
                                    const int* skip, hipStream_t st)
{
    if (nbox <= 0 || total_tiles <= 0) return hipSuccess;
    switch (P) {
        case 4: return contract_fx_batch_p<4>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 5: return contract_fx_batch_p<5>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 6: return contract_fx_batch_p<6>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 7: return contract_fx_batch_p<7>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 8: return contract_fx_batch_p<8>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 9: return contract_fx_batch_p<9>(boxes, nbox, total_tiles, lds, force, skip, st);
        case 10: return contract_fx_batch_p<10>(boxes, nbox, total_tiles, lds, force, skip, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_spread_fx_bound_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds,
                                        hipStream_t st)
{
    if (nbox <= 0 || total_tiles <= 0) return hipSuccess;
    switch (P) {
        case 4: return spread_fx_batch_p<4, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 5: return spread_fx_batch_p<5, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 6: return spread_fx_batch_p<6, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 7: return spread_fx_batch_p<7, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 8: return spread_fx_batch_p<8, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 9: return spread_fx_batch_p<9, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        case 10: return spread_fx_batch_p<10, true>(boxes, nbox, total_tiles, lds, nullptr, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cr
