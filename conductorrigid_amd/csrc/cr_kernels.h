// cr_kernels.h -- internal launcher interface between the host context
// (cr_context.cpp) and the HIP kernels (k_*.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cr {

constexpr unsigned NOT_MEMBER = 0xffffffffu;

// Spectral-Ewald mesh parameters (SetParams, cppmodule/ConductorRigid.cc:199-207;
// kernel constants, cppmodule/ConductorRigid.cu:913-917).
struct MeshParams {
    int Nx, Ny, Nz, P;
    double L[3];
    double h[3];
    double eta[3];
    double prefac;   // (2 xi^2/pi)^{3/2} / sqrt(eta_x eta_y eta_z)
    double xiterm;   // 2 xi^2
    double quadW;    // h_x h_y h_z
};

// Real-space table parameters (cppmodule/ConductorRigid.cc:278-279, .cu:540-554).
struct RealParams {
    double L[3];
    double rc, rc2, rmin2, drtable, selfterm;
    int N_table;
};

// One batched 1-D FFT stage plan along one mesh axis.
struct AxisPlan {
    int n;            // transform length
    int nstages;
    int radix[24];
    const void* tw;   // n twiddles e^{-2 pi i e/n} (float2 or double2)
};

// Device views of the system (borrowed pointers).
struct SysView {
    int N, N_rigid, N_isolated, Nmax;
    const double4* pos;
    const unsigned* iso;             // isolated-body tag per global index (context-owned)
    const double* body_charge;
    const unsigned* rigid_membership;
    const unsigned* rigid_members;
    const unsigned* rigid_size;
    const double4* rigid_com;
    const unsigned* nlist;
    const unsigned* head_list;
    const unsigned* n_neigh;
};

// Batched boxes' pair kernels (ELL build, real-space forces): one descriptor per box,
// boxes in ascending blk0; a box owns the workgroups [blk0, blk0 + ceil(N / 256)).
struct PairBox {
    RealParams rp;
    SysView s;
    const void* table;   // potential table (ELL build) or gradient table (forces), T
    long ldn;            // ELL leading dimension
    unsigned* col;       // ELL column indices
    void* val;           // ELL values (T)
    unsigned* row_len;   // ELL row lengths
    const double* q;     // charges (forces)
    double* force;       // [4N] forces: += the pair forces and q E0
    double E0[3];
    int blk0;
};

// Mesh tiling for the owner-computes spread: tile edges divide the mesh and are at
// least P-1 nodes, so a bead's window touches its own tile and the next one per axis.
struct TileGeom {
    int T[3];      // tile edge per axis (nodes)
    int nt[3];     // tiles per axis
    int ntiles;
    int valid;     // 0: no admissible tiling, use the atomic spread
    int lds;       // LDS bytes a binned workgroup may use (extended tile + staging)
};

// ---- k_se.hip ----
template <typename T>
hipError_t launch_spread(const MeshParams& mp, const SysView& s, const double* q, T* grid, hipStream_t st);

// ---- k_tiled.hip ----
// Bin beads by the tile holding their window's first node (counting sort):
// tile_count[ntiles+1] (zeroed here), tile_start[ntiles+1], cursor[ntiles], sorted[N].
hipError_t launch_bin_beads(const MeshParams& mp, const TileGeom& tg, const SysView& s, unsigned* bead_tile,
                            unsigned* tile_count, unsigned* tile_start, unsigned* cursor, unsigned* sorted,
                            unsigned long long* scan_sums, hipStream_t st);
// Binned bead order and per-bead window records (k_bead_records): rec_off packs the
// window's first node relative to its tile origin (x | y << 10 | z << 20).  rec_win
// holds, in FP32, the bead's position in its window per axis in node units (x, y, z, 0),
// from which the kernels evaluate the axis weights exp(-c_d (p - u_d)^2); in FP64 the 3P
// weights themselves (x, then y, then z).
struct BinView {
    const unsigned* tile_start;   // [ntiles + 1]
    const unsigned* sorted;       // [N] bead index per binned slot
    const int* rec_off;           // [N]
    const void* rec_win;          // window_record_bytes per bead
};
inline size_t window_record_bytes(const MeshParams& mp, bool fp32)
{
    return fp32 ? 4 * sizeof(float) : 3 * (size_t)mp.P * sizeof(double);
}
template <typename T>
hipError_t launch_bead_records(const MeshParams& mp, const TileGeom& tg, const SysView& s, const unsigned* sorted,
                               const unsigned* bead_tile, int* rec_off, T* rec_win, hipStream_t st);
// grid = spread of q; clear = false leaves the mesh clear to the caller (it must be zero)
template <typename T>
hipError_t launch_spread_tiled(const MeshParams& mp, const TileGeom& tg, const BinView& bv, const double* q, T* grid,
                               hipStream_t st, bool clear = true);
// LDS plan of the binned kernels: the tile extended by P-1 nodes per axis (ext_elem bytes
// per node) plus per-bead staging tables (stage_elem bytes per value) for `chunk` beads,
// within kTiledLdsBudget bytes.  The spread accumulates its tile in SpreadAcc.
constexpr size_t kTiledLdsBudget = 64 * 1024;   // default per-workgroup LDS budget
// Workgroups of the binned kernels.  The spread's LDS tile limits a CU to two
// workgroups, so 1024 threads give it twice the waves to hide LDS latency; the FP32
// contraction keeps 512 (its smaller tile admits three workgroups per CU), the FP64 one
// measures faster at 1024.
constexpr int kSpreadThreads = 1024;
template <typename T>
constexpr int contract_threads() { return sizeof(T) == sizeof(double) ? 1024 : 512; }
constexpr int kTiledChunk = 64;      // beads staged per round (bins hold ~60 at 1M / 256^3)
using SpreadAcc = double;
struct TiledLds {
    size_t ext_bytes;
    size_t total_bytes;
    int chunk;
};
TiledLds tiled_lds(const MeshParams& mp, const TileGeom& tg, size_t ext_elem, size_t stage_elem);
// Binned contraction from an LDS copy of the tile extended by P-1 nodes on the + side: potential (force == nullptr)
// or reciprocal force (phi unused).
template <typename T>
hipError_t launch_contract_tiled(const MeshParams& mp, const TileGeom& tg, const SysView& s, const BinView& bv,
                                 const double* q, const T* grid, T* phi, double* force, hipStream_t st,
                                 const double* fadd = nullptr);
// One box of a batched binned spread / contraction (device array, tile0 ascending): its
// mesh, tiling and binned bead records, and the arrays it reads and writes.
struct TiledBox {
    MeshParams mp;
    TileGeom tg;
    int ch;                      // staging chunk within the launch's LDS size
    int tile0;                   // first global tile of this box
    int box;                     // index into the skip flags
    long G;                      // mesh nodes
    const unsigned* tile_start;  // [ntiles + 1]
    const unsigned* sorted;      // [N]
    const int* rec_off;          // [N]
    const void* rec_win;         // window_record_bytes per bead (T)
    const double* q;             // spread input / charges of the force contraction, bead order
    void* grid;                  // real mesh (T)
    void* phi;                   // contraction output (T), bead order
    const double4* pos;          // bead positions (force contraction)
    double* force;               // force contraction output [4N] (overwritten)
};
// skip (device, may be null): boxes with skip[box] != 0 are left untouched.
template <typename T>
hipError_t launch_fill_batch(const TiledBox* boxes, int nbox, long Gmax, const int* skip, hipStream_t st);
template <typename T>
hipError_t launch_spread_tiled_batch(const TiledBox* boxes, int nbox, int total_tiles, int P, size_t lds_bytes,
                                     const int* skip, hipStream_t st);
template <typename T>
hipError_t launch_contract_tiled_batch(const TiledBox* boxes, int nbox, int total_tiles, int P, size_t lds_bytes,
                                       const int* skip, hipStream_t st, bool force = false);
// ---- k_spread_fx.hip: FP32-bulk spread in 32-bit fixed point ----
// Node values are summed as int32 in units of 1 / (mult / prefac) and leave as float
// (int sum x inv).
constexpr int kFxParts = 1024;  // workgroups of k_charges_fx (partial maxima)
struct FxScale {
    float mult;       // charge -> fixed-point units: 2^30 / (qmax W)
    float inv;        // fixed-point units -> mesh value: prefac qmax W / 2^30 (NaN for a NaN charge)
    float qmax;       // max |q| of the last spread's charges
    unsigned wbits;   // float bits of the largest tile node of sum_b w_b (W = 8 x this)
    float part[kFxParts];
};
constexpr int kFxThreads = 1024;
constexpr int kFxChunk = 256;    // beads staged per round
struct FxPlan {
    int valid;
    int chunk;
    size_t lds_bytes;
};
// LDS plan for the context's tiling (lds_max: the device's LDS per workgroup); valid == 0
// when the tile does not fit or P is outside 4..10.
FxPlan fx_plan(const MeshParams& mp, const TileGeom& tg, size_t lds_max);
// qf = (float) x [N]; max |qf|, then fs->mult and fs->inv from it and fs->wbits.
hipError_t launch_charges_fx(const double* x, int N, float* qf, FxScale* fs, double prefac, hipStream_t st);
// The scale alone (fs->mult, fs->inv, fs->qmax) from nparts partial maxima in fs->part.
hipError_t launch_fx_scale(int nparts, FxScale* fs, double prefac, hipStream_t st);
// qf = (float) x [N] and the partial maxima alone (*nparts of them): launch_spread_fx with
// nparts forms the scale.
hipError_t launch_charges_fx_parts(const double* x, int N, float* qf, FxScale* fs, int* nparts, hipStream_t st);
// Scratch ints per tile block of launch_spread_fx (a multiple of 4).
long fx_block_ints(const MeshParams& mp, const TileGeom& tg);
// mesh (float, G nodes) = fixed-point spread of qf at fs's scale: the tiles' integer
// blocks go to the scratch blocks (tg.ntiles x fx_block_ints) and are summed per node.
// nparts > 0: the scale is first formed from the nparts partial maxima in fs->part (and
// stored in fs); nparts == 0: fs already holds it.
hipError_t launch_spread_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                            const float* qf, FxScale* fs, int* blocks, float* mesh, hipStream_t st, int nparts = 0);
// fs->wbits for the binned configuration (once per bead configuration).
hipError_t launch_spread_fx_bound(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                                  FxScale* fs, hipStream_t st);
// Batched boxes' per-evaluation preparation (membership, binning, window records,
// objective, body grouping, ELL width, fixed-point bound zeroing): one descriptor per box,
// boxes in ascending blk0; a box owns the workgroups [blk0, blk0 + nblk) of every phase,
// nblk = ceil(max(N, N_rigid * Nmax, N_isolated, ntiles + 1) / 256).
struct PrepBox {
    MeshParams mp;
    TileGeom tg;
    SysView s;                // s.iso is the context's membership array (written here)
    int blk0;
    int bin;                  // 1: bin the beads on tg
    const unsigned* tag;      // the caller's tags and body-of-tag map (membership)
    const unsigned* body_of_tag;
    unsigned* iso;
    unsigned* bead_tile;      // binning (bin = 1)
    unsigned* tile_count;
    unsigned* tile_start;
    unsigned* cursor;
    unsigned* sorted;
    int* rec_off;
    void* rec_win;
    double* b;                // objective
    double* bcat;             // the same rows in the batch's concatenated vector
    double E0[3];
    unsigned* body_count;     // body grouping (null: none)
    unsigned* body_start;
    unsigned* body_beads;
    unsigned* width;          // max n_neigh (null: none)
    unsigned* wbits;          // fixed-point weight bound to zero (null: none)
};
// The preparation of all boxes: five launches (zero, beads, scans, fills, records), fp32
// selects the window-record layout; then the objective rows of all boxes in one launch.
hipError_t launch_prep_batch(const PrepBox* boxes, int nbox, int total_blocks, bool fp32, hipStream_t st);
hipError_t launch_objective_batch(const PrepBox* boxes, int nbox, int total_blocks, hipStream_t st);

// Batched boxes on the fixed-point kernels (every box FP32 bulk with valid plans, one P):
// one descriptor per box, boxes in ascending tile0 and gblk0.
struct FxBox {
    MeshParams mp;
    TileGeom tg;
    int box;                     // index into the skip flags
    int tile0;                   // first global tile of this box
    int gblk0;                   // first gather workgroup of this box
    int N;                       // beads
    int ch, cch;                 // staging chunks of the spread and contraction plans
    long bstride;                // ints per tile block (fx_block_ints)
    const unsigned* tile_start;  // [ntiles + 1]
    const unsigned* sorted;      // [N]
    const int* rec_off;          // [N]
    const float4* rec_u;         // [N] window records
    const double* x;             // spread input: the box's charges
    float* qf;                   // [N] the charges rounded to float
    FxScale* fs;
    int* blocks;                 // tile blocks (tg.ntiles x bstride)
    float* grid;                 // the box's mesh
    float* phi;                  // potential contraction output [N]
    const double* q;             // charges of the force contraction
    double* force;               // force contraction output [4N] (overwritten)
};
// whether the gather takes 4 nodes per thread for this mesh and tiling
bool fx_gather_v4(const MeshParams& mp, const TileGeom& tg);
// gather workgroups of one box (V = 4 or 1 nodes per thread)
long fx_gather_blocks(const MeshParams& mp, bool v4);
// fixed-point scale of every box not skipped: charges of x rounded to float (the same
// workgroup count for every box, set by the largest, max_n beads), then mult / inv per box
hipError_t launch_charges_fx_batch(const FxBox* boxes, int nbox, int max_n, const int* skip, hipStream_t st);
// every box's mesh from its charges: tile kernel over all boxes' tiles, then the gather
// (gather_blocks workgroups in all)
hipError_t launch_spread_fx_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds, long gather_blocks,
                                  bool v4, const int* skip, hipStream_t st);
// every box's fixed-point weight bound (fs->wbits, zeroed beforehand), once per bead
// configuration
hipError_t launch_spread_fx_bound_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds,
                                        hipStream_t st);
// potential (force = false) or reciprocal-force contraction of every box not skipped
hipError_t launch_contract_fx_batch(const FxBox* boxes, int nbox, int P, int total_tiles, size_t lds, bool force,
                                    const int* skip, hipStream_t st);
// FP32-bulk potential contraction on the binned tiles (k_spread_fx.hip); its own LDS plan.
FxPlan fx_contract_plan(const MeshParams& mp, const TileGeom& tg, size_t lds_max);
hipError_t launch_contract_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                              const float* grid, float* phi, hipStream_t st);
// FP32-bulk reciprocal force (the gradient of each bead's window) on the same tiles and
// plan: force [4N] = the reciprocal-space force of the charges q (+ fadd [4N] when given).
hipError_t launch_contract_force_fx(const MeshParams& mp, const TileGeom& tg, const FxPlan& pl, const BinView& bv,
                                    const float* grid, const double* q, double* force, const double* fadd,
                                    hipStream_t st);

// ---- k_se.hip (unbinned fallbacks) ----
template <typename T>
hipError_t launch_contract(const MeshParams& mp, const SysView& s, const T* grid, T* phi, hipStream_t st);
template <typename T>
hipError_t launch_contract_force(const MeshParams& mp, const SysView& s, const double* q, const T* grid,
                                 double* force, hipStream_t st, const double* fadd = nullptr);
template <typename T>
hipError_t launch_fill(T* p, T v, long n, hipStream_t st);
template <typename Tin, typename Tout>
hipError_t launch_convert(const Tin* in, Tout* out, long n, hipStream_t st);
// k-space factor tables regenerated on the device after a box change (k_se.hip): one
// axis's wavenumbers and Gaussian factors, and the cubic-box radial table.
hipError_t launch_kspace_axis(int n, double L, double xi, double eta, double* k, double* e, hipStream_t st);
// gauss >= 0: each entry also times exp(-gauss m2) (the Gaussian factors of the three axes).
template <typename T>
hipError_t launch_kspace_radial(long m2max, double kappa, double invG, T* rad, hipStream_t st, double gauss = -1.0);

// ---- k_fft.hip ----
// Real mesh [Nx][Ny][Nz] <-> half spectrum [Nx][Ny][Nz/2+1] (interleaved complex) with rows
// zp complex elements apart (zp >= Nz/2+1; the padding is never read or written; the
// generic mixed-radix passes take zp = Nz/2+1 only).
template <typename T>
hipError_t launch_fft_z_r2c(const AxisPlan& pz, int Nx, int Ny, const T* real, T* cplx, int zp, hipStream_t st);
template <typename T>
hipError_t launch_fft_z_c2r(const AxisPlan& pz, int Nx, int Ny, const T* cplx, int zp, T* real, hipStream_t st);
// C2C along y (axis 1) or x (axis 0) of the Zc columns of a complex mesh [Nx][Ny][Zp]; sign -1 / +1.
template <typename T>
hipError_t launch_fft_yx(const AxisPlan& pa, int axis, int Nx, int Ny, int Zc, int Zp, T* cplx, int sign,
                         hipStream_t st);
// C2C along z of a full complex mesh [Nx][Ny][Nz].
template <typename T>
hipError_t launch_fft_z_c2c(const AxisPlan& pz, int Nx, int Ny, T* cplx, int sign, hipStream_t st);
// Forward x FFT, multiply by the Green's function scale S(k), inverse x FFT.
struct ScaleTables {
    const double* k[3];    // wavenumbers per axis
    const double* e[3];    // exp(-(1-eta_d) k_d^2 / (4 xi^2)) per axis
    double invG;
    // Cubic boxes: k^2 = (2 pi / L)^2 m2 with m2 = mx^2 + my^2 + mz^2 an integer, so the
    // non-separable factor j0(|k|)^2 / k^2 / G is a table over m2 (mesh precision), read
    // by the register x pass instead of a square root, sine and two divisions per mode.
    // Null for other boxes.
    const void* radial;
    // radial also holds the Gaussian factor e_x e_y e_z = exp(-(1-eta) k^2 / (4 xi^2)) (eta
    // the same on every axis), so that S = radial[m2] alone
    int radial_gauss;
    int n[3];              // mesh size per axis (signed frequency m of index i)
};
template <typename T>
hipError_t launch_fft_x_scale(const AxisPlan& px, int Nx, int Ny, int Zc, int Zp, T* cplx, const ScaleTables& sc,
                              hipStream_t st);

// ---- k_fft_reg.hip: register four-step passes for n = R * R, R in {8, 16, 32} ----
int fftr_radix(int n);  // n when the register-resident passes take length n, else 0
template <typename T>
hipError_t launch_fftr_z_r2c(const AxisPlan& pz, long nrows, int zp, const T* real, T* cplx, hipStream_t st);
template <typename T>
hipError_t launch_fftr_z_c2r(const AxisPlan& pz, long nrows, int zp, const T* cplx, T* real, hipStream_t st);
template <typename T>
hipError_t launch_fftr_z_c2c(const AxisPlan& pz, long nrows, T* data, int sign, hipStream_t st);
template <typename T>
hipError_t launch_fftr_strided(const AxisPlan& pa, int axis, int Nx, int Ny, int Zc, int Zp, T* data, int sign,
                               hipStream_t st);
template <typename T>
hipError_t launch_fftr_x_scale(const AxisPlan& px, int Ny, int Zc, int Zp, T* data, const ScaleTables& sc,
                               hipStream_t st);
// x pass + scale for nbox meshes stored back to back ([box][Nx][Ny][Zp], Nx = px.n), each
// with its own scale tables sct[box] (device array); register-path lengths only.
template <typename T>
hipError_t launch_fftr_x_scale_batch(const AxisPlan& px, int nbox, int Ny, int Zc, int Zp, T* data,
                                     const ScaleTables* sct, hipStream_t st);

// ---- k_real.hip ----
template <typename T>
hipError_t launch_real_space(const RealParams& rp, const SysView& s, const double* x, const T* phi_recip,
                             const T* table, double* y, int precond, hipStream_t st);
// Real-space operator assembled once per force evaluation (column-major ELL, row stride
// ldn >= N, width >= max n_neigh) and applied per matvec.
template <typename T>
hipError_t launch_real_csr_build(const RealParams& rp, const SysView& s, const T* table, long ldn, unsigned* col, T* val,
                                 unsigned* row_len, hipStream_t st);
template <typename T>
hipError_t launch_real_space_csr(const RealParams& rp, const SysView& s, const double* x, const T* qt,
                                 const T* phi_recip, long ldn, const unsigned* col, const T* val,
                                 const unsigned* row_len, double* y, int precond, hipStream_t st);
// *out = max(*out, a[i]) over i < n (out must be initialised)
hipError_t launch_max_u32(const unsigned* a, int n, unsigned* out, hipStream_t st);
template <typename T>
// force += q E0 + pair forces (accum), or force = q E0 + pair forces
hipError_t launch_real_space_force(const RealParams& rp, const SysView& s, const double* q, const T* gtable,
                                   const double E0[3], double* force, hipStream_t st, bool accum = true);
// The ELL build and the pair forces of all boxes of a batch, one launch each
// (total_blocks = the sum of ceil(N / 256) over the boxes).
template <typename T>
hipError_t launch_real_csr_build_batch(const PairBox* boxes, int nbox, int total_blocks, hipStream_t st);
template <typename T>
hipError_t launch_real_space_force_batch(const PairBox* boxes, int nbox, int total_blocks, hipStream_t st);
hipError_t launch_constraint_rows(const SysView& s, const double* q, double* y_bottom, int precond, hipStream_t st);
hipError_t launch_precondition(const SysView& s, const double* x, double* y, hipStream_t st);
hipError_t launch_objective(const SysView& s, const double L[3], const double E0[3], double* b, hipStream_t st);
hipError_t launch_membership(int N, const unsigned* tag, const unsigned* body_of_tag, unsigned* iso, hipStream_t st);
// Per-body segments (isolated-body tag s.iso): counts[b] = beads of body b.
hipError_t launch_body_counts(const SysView& s, unsigned* counts, hipStream_t st);
// q_i += (Q_b - sum_{j in b} q_j) / n_b, so every body carries its net charge exactly.
// sums: scratch [N_isolated] doubles.
hipError_t launch_project_charges(const SysView& s, const unsigned* counts, double* sums, double* q, hipStream_t st);
// Phi_b = mean over the beads of b of (b_i - y_i), the least-squares equipotential of the
// rows Phi_b - (M q)_i = b_i given y_i = -(M q)_i.  sums: scratch [N_isolated] doubles.
// y[i] += Phi[body of bead i], i < N: A [q; Phi] from A [q; 0] (the body potential enters
// only the bead rows, with coefficient 1)
hipError_t launch_add_body_potentials(const SysView& s, const double* Phi, double* y, hipStream_t st);
hipError_t launch_fit_potentials(const SysView& s, const unsigned* counts, const double* b, const double* y,
                                 double* sums, double* phi, hipStream_t st);
hipError_t launch_zero_force(int N, double* force, hipStream_t st);

// ---- k_precond.hip ----
// Near-field block preconditioner: per isolated body, the inverse of its own saddle block.
// The single-system form keeps its bodies in slot order (ascending smallest bead index):
// order[slot] is the body, count[slot] its bead count and members[slot * nmax + k] its k-th
// bead; ginv, uvec and inv_s are stored by slot.  A slot's inverse is kept from the previous
// setup when the slot holds the same body, beads and shape (cache_key: B entries, cache_sig:
// 2B, cache_ok: they describe the stored inverses).  Scratch: bucket_count and bucket_start
// of block_order_buckets(N) + 1 entries, scan_sums of block_order_buckets(N) / 1024 + 4, key
// and mkey of B.
struct BlockSlots {
    unsigned* order;
    unsigned* count;
    unsigned* members;
    unsigned* bucket_count;
    unsigned* bucket_start;
    unsigned long long* scan_sums;
    uint2* key;
    unsigned* mkey;
    unsigned long long* cache_key;
    double* cache_sig;
    int cache_ok;
};
long block_order_buckets(int N);
// Floats per slot of ginv: a block's inverse is symmetric, so only its lower triangle is
// stored (row-major, entry (a, b) with a >= b at a (a + 1) / 2 + b), padded to whole float4s.
__host__ __device__ constexpr long block_inv_stride(int nmax) { return ((long)nmax * (nmax + 1) / 2 + 3) / 4 * 4; }
// Bodies grouped (body_start, body_beads); with invert, the slots and the per-slot inverses
// (bodies of at most 64 beads).
hipError_t launch_block_setup(const SysView& s, const double L[3], int nmax, unsigned* body_count,
                              unsigned* body_start, unsigned* body_beads,
                              unsigned long long* scan_sums, float* ginv, float* uvec, double* inv_s, hipStream_t st,
                              bool invert = true, const BlockSlots* slots = nullptr);
// One box of a batched block-preconditioner application (device array, body0 ascending):
// the box's per-body inverses and its offset in the batch vectors.
struct BlockBox {
    long xoff;
    int N, nmax, body0;
    const unsigned* start;
    const unsigned* beads;
    const float* ginv;
    const float* uvec;
    const double* inv_s;
    const double4* pos;   // bead positions and box (setup only)
    double L[3];
};
// Per-body inverses of every box's bodies in one launch (the boxes' beads already grouped
// by body); nmax_all = the largest nmax of the batch, at most 64.
hipError_t launch_block_setup_batch(const BlockBox* boxes, int nbox, int total_bodies, int nmax_all, hipStream_t st);
hipError_t launch_block_apply_batch(const BlockBox* boxes, int nbox, int total_bodies, const double* y, double* w,
                                    const int* skip, hipStream_t st);
// w = D y with the near-field block preconditioner (slot-major, from launch_block_setup).
// xq != nullptr: body t's constraint row is taken as the sum of xq over its beads (the
// operator's charge row for input xq) instead of y[N + t], which is then never read.
hipError_t launch_block_apply(const SysView& s, int nmax, const BlockSlots& slots, const float* ginv,
                              const float* uvec, const double* inv_s, const double* y, double* w, hipStream_t st,
                              const double* xq = nullptr);

// ---- k_mlp.hip ----
// Learned warm start: per-bead MLP trunk kMlpTrunk -> kMlpHidden -> kMlpHidden -> kMlpOut
// (SiLU), output y = (E0, 1) . u (the solved charges are linear in E0 at fixed geometry).
constexpr int kMlpIn = 176;       // feature columns per bead (row pitch)
constexpr int kMlpTrunk = 168;    // 32 neighbours x 5, charge share, COM offset and |offset|, zero padded
constexpr int kMlpGate = 168;     // E0 as bf16 pairs: hi at [168, 171), lo at [171, 174)
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 4;
constexpr int kMlpNeigh = 32;
constexpr double kMlpRadius = 3.0;
// network outputs are charges in units of 4 pi (potential = q / (4 pi r) in these units)
constexpr double kMlpChargeUnit = 12.566370614359172;
// Flat FP32 parameter vector: W1 [kMlpIn][H] | b1 [H] | W2 [H][H] | b2 [H] | W3 [H][4] | b3 [4]
// (W1 rows >= kMlpTrunk multiply the gate columns, which are masked, and stay unused)
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpIn * kMlpHidden;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden * kMlpOut;
constexpr int kMlpParams = kMlpOffB3 + kMlpOut;
// bf16 compute copies (raw bits) and FP32 biases
struct MlpDevWeights {
    const unsigned short* W1t;   // [H][kMlpIn]
    const unsigned short* W2t;   // [H][H], W2 transposed
    const unsigned short* W2n;   // [H][H], W2 as stored
    const unsigned short* W3;    // [H][kMlpOut]
    const float* b1;
    const float* b2;
    const float* b3;
};
// activations and gradients of one training batch of M rows
struct MlpTrainBufs {
    float* z1;             // [M][H]
    float* z2;             // [M][H]
    float* du;             // [M][kMlpOut]
    float* part;           // [slices][kMlpIn][H]
    unsigned short* h1;    // [M][H]
    unsigned short* h2;    // [M][H]
    unsigned short* dz1;   // [M][H]
    unsigned short* dz2;   // [M][H]
    int slices;
};
hipError_t launch_mlp_features(const SysView& s, const double L[3], const double E0[3], const unsigned* body_count,
                               unsigned short* X, hipStream_t st);
hipError_t launch_mlp_pack(const float* theta, const MlpDevWeights& w, hipStream_t st);
// y[m] for m < M; tb != nullptr also stores the activations for the backward pass
hipError_t launch_mlp_forward(int M, const unsigned short* X, const MlpDevWeights& w, float* y, const MlpTrainBufs* tb,
                              hipStream_t st);
// grad (flat, kMlpParams) of mean_m (y_m - t_m)^2; y receives the forward outputs
hipError_t launch_mlp_gradient(int M, const unsigned short* X, const float* t, const MlpDevWeights& w,
                               const MlpTrainBufs& tb, float* y, float* grad, hipStream_t st);
hipError_t launch_mlp_adam(float* theta, const float* g, float* m1, float* m2, float lr, int step, hipStream_t st);
// t_i = (q_i - Q_b / n_b) / kMlpChargeUnit (to = 0) or q_i = Q_b / n_b + kMlpChargeUnit y_i (to = 1)
hipError_t launch_mlp_share(const SysView& s, const unsigned* counts, const double* q, const float* y, float* t,
                            double* q_out, int to_charges, hipStream_t st);

// ---- k_dynamics.hip ----
struct BdParams {
    double L[3];
    double dt, kT, gamma_t, gamma_r, k_rep;
};
// Brownian rigid-body step from bead forces plus steric repulsion (force_tmp: [N]
// scratch); noise_rigid [6 N_rigid], noise_beads [3 N] (may be null when every bead is
// in a rigid body).  pos/com are updated in place.
hipError_t launch_bd_step(const BdParams& p, const SysView& s, double4* pos, double4* com, const double4* force_in,
                          double4* force_tmp, const double* noise_rigid, const double* noise_beads, hipStream_t st);

// ---- k_nlist.hip ----
// out = exclusive scan of in[0:n]; sums needs n/1024 + 2 entries; total lands in sums[nblocks].
hipError_t exclusive_scan_u32(const unsigned* in, unsigned* out, long n, unsigned long long* sums, hipStream_t st);
hipError_t build_nlist(int N, const double4* pos, const double L[3], double rlist, unsigned* head,
                       unsigned* nn, unsigned* nlist, long capacity, long* total, void* scratch_pool,
                       hipStream_t st);
// The same build on the file-scope static scratch of k_nlist.hip (no allocation):
// nlist_scratch_setup(device) records that device's scratch address (the device must be
// current); nlist_noalloc_bytes is what a build needs, nlist_noalloc_capacity what the
// scratch holds; build_nlist_noalloc returns hipErrorInvalidValue beyond it.
hipError_t nlist_scratch_setup(int device);
size_t nlist_noalloc_bytes(int N, const double L[3], double rlist);
size_t nlist_noalloc_capacity();
hipError_t build_nlist_noalloc(int N, const double4* pos, const double L[3], double rlist, unsigned* head,
                               unsigned* nn, unsigned* nlist, long capacity, long* total, hipStream_t st);

// ---- k_krylov.hip ----
// Partial dot products of w with columns 0..ncols-1 of V (column pitch ld):
// partial[k * 1024 + b] for workgroup b (column-major, so that a reduction over the blocks
// reads one contiguous run); nblocks returned by krylov_blocks(n), at most 1024.
int krylov_blocks(long n);
hipError_t launch_multidot(const double* V, long ld, int ncols, const double* w, long n, double* partial,
                           hipStream_t st);
// out[k] = sum over blocks of partial[k * 1024 + b], k < ncols (fixed summation order)
hipError_t launch_reduce_partials(const double* partial, int nblocks, int ncols, double* out, hipStream_t st);
// w -= V[:,0:ncols] h; then (mode 1) partial dots V^T w, (mode 2) partial ||w||^2, (mode 0) nothing;
// mode 3: mode 2 unless ||h[0:ncols]|| is below 1e-13 ||w|| with h[ncols] = ||w||^2, in which
// case w is left as it is and the partials carry h[ncols] (CGS2's second correction)
hipError_t launch_gs_update(const double* V, long ld, int ncols, const double* h, double* w, long n, int mode,
                            double* partial, hipStream_t st);
// y = a*x + b*y elementwise helpers
// w -= V[:, 0:ncols] h and partial = the projections of the updated w (layout of launch_multidot);
// with_norm (ncols < 32) adds column ncols: the updated w's squared norm
hipError_t launch_gs_update_dot(const double* V, long ld, int ncols, const double* h, double* w, long n,
                                double* partial, hipStream_t st, bool with_norm = false);
hipError_t launch_axpby(double a, const double* x, double b, double* y, long n, hipStream_t st);
hipError_t launch_scale_copy(const double* x, const double* s_ptr, double sign, double* y, long n, hipStream_t st);
// x += V[:,0:k] c  (c on device)
hipError_t launch_combine(const double* V, long ld, const double* c, const int* kcount, double* x, long n,
                          hipStream_t st);

// Arnoldi step bookkeeping on device (single thread): see k_krylov.hip.
struct GmresState {
    double H[33 * 32];     // column-major (R+1) x R, R <= 32
    double cs[32], sn[32], s[33];
    double h1[32], h2[32];
    double wnorm2;
    double inv_norm;       // 1 / H(i+1,i)
    double beta, inv_beta;
    double y[32];
    int i;                 // current Arnoldi column
    int ncols;             // columns used in the last cycle (i+1)
    double resid;
};
hipError_t launch_givens(GmresState* g, int i, int R, hipStream_t st);
// One device scalar published to host memory that the host polls: value, then (after a
// system-scope fence) the sequence number.  dst must be fine-grained (coherent) pinned
// host memory.
struct HostPub {
    double value;
    unsigned long long seq;
};
// Mailboxes per context: a GMRES step's residual can be read one step late, while the next
// step's is being published.
constexpr int kPubSlots = 2;
hipError_t launch_publish(const double* src, HostPub* dst, unsigned long long seq, hipStream_t st);
// The end of Arnoldi step i: ||w||^2 from the last CGS pass's partials (nblocks of them,
// column 0), vnext = w / ||w||, column i of H and its Givens rotations, then g->resid
// published to pub (when not null) with sequence number seq.  qf != nullptr: also the FP32
// charges of vnext's first nq entries and their partial maxima in fs->part[0:*nparts), as
// launch_charges_fx leaves them before its scale (launch_fx_scale completes it).
hipError_t launch_step_tail(const double* partial, int nblocks, GmresState* g, int i, int R, const double* w,
                            double* vnext, long n, HostPub* pub, unsigned long long seq, hipStream_t st,
                            float* qf = nullptr, FxScale* fs = nullptr, long nq = 0, int* nparts = nullptr,
                            const double* V = nullptr, long ld = 0, int ncols = 0, const double* h2 = nullptr);
// (h2: the norm and CGS2's second correction w -= V[:, 0:ncols] h2 come from h2[0:ncols]
// and h2[ncols] = ||w||^2 (partial unused); w itself is left as it is)
hipError_t launch_cycle_init(GmresState* g, const double* norm2, hipStream_t st);
hipError_t launch_backsub(GmresState* g, int R, hipStream_t st);

// ---- k_batch.hip: a batch of independent boxes solved in lockstep ----
// The unknown vectors [q (N_b) ; Phi (B_b)] of all boxes are concatenated (box b at
// offset xoff); segmented launches give each box a contiguous run of workgroups (blk0
// ascending over boxes) so every dot product stays within one box.  skip[b] != 0 (a
// box that has converged) makes a box's workgroups return at once.
struct BatchSeg {
    long xoff;       // offset of the box's unknowns in the batch vectors
    int n;           // N + B
    int blk0, nblk;  // the box's workgroups in segmented vector launches
};
// One box of the batched operator rows: real-space ELL rows (the reciprocal potential
// already contracted into phi_recip) and the per-body charge rows.
struct RowBox {
    long xoff;
    int N, N_rigid, N_isolated, Nmax;
    int blk0;    // first bead-row workgroup
    int rblk0;   // first rigid-body-row workgroup
    const unsigned* iso;
    const unsigned* rigid_membership;
    const unsigned* rigid_members;
    const unsigned* rigid_size;
    const void* phi_recip;  // T[N]
    long ldn;
    const unsigned* col;
    const void* val;        // T[width * ldn]
    const unsigned* row_len;
    double selfterm;
};
// Per-box Krylov state of the batched solver.
struct BatchGmres {
    GmresState g;
    double tol;      // tol * ||b|| (or ||D b||)
    double norm2;    // scratch: squared norm reduced by launch_breduce(..., kRedNorm)
};
enum BatchRed { kRedH1 = 0, kRedH2 = 1, kRedWnorm = 2, kRedNorm = 3 };
constexpr int kBatchBlock = 256;
// Workgroups a box of n unknowns gets in segmented launches.
int batch_blocks(long n);
// xt: x in the row precision T (x itself for double), the real-space rows' gathered operand
template <typename T>
hipError_t launch_brows(const RowBox* boxes, int nbox, int bead_blocks, int rigid_blocks, const double* x,
                        const T* xt, double* y, int precond, const int* skip, hipStream_t st);
// partial[k * 32 + c] = sum over block k's range of V[c] . w for c < ncols, with
// ncols = fixed_cols if > 0 else *step + 1 (device step counter).
hipError_t launch_bdot(const double* V, long ld, int fixed_cols, const int* step, const double* w,
                       const BatchSeg* segs, int nbox, int total_blocks, double* partial, const int* skip,
                       hipStream_t st);
hipError_t launch_breduce(const double* partial, const BatchSeg* segs, int nbox, int fixed_cols, const int* step,
                          BatchGmres* gs, int which, const int* skip, hipStream_t st);
// w -= V[:, 0:*step+1] h (h1 or h2 by `which`); mode 2 adds ||w||^2 partials in partial[k*32]
hipError_t launch_bgs_update(const double* V, long ld, const int* step, double* w, const BatchSeg* segs, int nbox,
                             int total_blocks, const BatchGmres* gs, int which, int mode, double* partial,
                             const int* skip, hipStream_t st);
// Givens step i = *step for every box not skipped; a box whose estimate meets its tol sets skip.
hipError_t launch_bgivens(BatchGmres* gs, int nbox, const int* step, int R, int* skip, hipStream_t st);
// V[:, col] = sign * scale * w per box; col = *step + 1 with scale 1/||w|| (col_fixed < 0),
// or col_fixed with scale 1/beta.
hipError_t launch_bscale_copy(const double* w, const BatchSeg* segs, int nbox, int total_blocks,
                              const BatchGmres* gs, const int* step, int col_fixed, double sign, double* V, long ld,
                              const int* skip, hipStream_t st);
hipError_t launch_bcopy_col(const double* V, long ld, const int* step, double* u, long n, hipStream_t st);
// Cycle start from the reduced ||w||^2: beta, rotations reset; skip[b] = beta <= tol for
// active boxes (finished[b] != 0 boxes keep skip = 1).
hipError_t launch_bcycle_init(BatchGmres* gs, int nbox, const int* finished, int* skip, hipStream_t st);
hipError_t launch_bbacksub(BatchGmres* gs, int nbox, int R, const int* finished, hipStream_t st);
hipError_t launch_bcombine(const double* V, long ld, const BatchSeg* segs, int nbox, int total_blocks,
                           const BatchGmres* gs, const int* finished, double* x, hipStream_t st);
// tol_b = rel_b * sqrt(norm2_b) from the last kRedNorm reduction.
hipError_t launch_bset_tol(BatchGmres* gs, int nbox, const double* rel, hipStream_t st);
// Publish pub[3b] = resid, pub[3b+1] = skip, pub[3b+2] = tol for the host; inc != 0 ends a
// lockstep step (++*step).
hipError_t launch_bstep_end(int* step, const BatchGmres* gs, const int* skip, int nbox, double* pub, int inc,
                            hipStream_t st);
hipError_t launch_bset_step(int* step, int value, hipStream_t st);
// out[stride * b + slot] = gs[b].norm2 (collects kRedNorm results for one readback)
hipError_t launch_bcollect(const BatchGmres* gs, int nbox, double* out, int stride, int slot, hipStream_t st);
// y = D x with the reference's diagonal preconditioner for every box (rows as launch_brows)
hipError_t launch_bprecond_diag(const RowBox* boxes, int nbox, int bead_blocks, int rigid_blocks, const double* x,
                                double* y, hipStream_t st);

}  // namespace cr
