// cr_context_impl.h -- the context behind the C ABI (cr_context.cpp) and the internal
// operations the batched solver (cr_batch.cpp) composes from it.  Not part of the ABI.
#pragma once

#include "cr_host.h"
#include "cr_kernels.h"

using namespace cr;

struct cr_context {
    int device = 0;
    bool ready = false;
    bool check_residual = true;  // recompute the true residual when stats are requested
    cr_params params{};
    cr_resolved res{};
    MeshParams mp{};
    RealParams rp{};
    int Nzh = 0;  // half-spectrum columns, Nz/2+1
    int Nzp = 0;  // half-spectrum row pitch (>= Nzh)
    long G = 0, Gh = 0;
    AxisPlan plan_d[3]{}, plan_f[3]{};
    DevBuf tw_d[3], tw_f[3];
    DevBuf ktab[3], etab[3];
    DevBuf radtab;  // j0(|k|)^2 / k^2 / G over m2 = |m|^2 (cubic boxes)
    ScaleTables sc{};
    DevBuf tab_d, gtab_d, tab_f, gtab_f;
    DevBuf grid_real, grid_cplx;
    DevBuf phi_recip, iso;
    // real-space operator assembled per force evaluation (CSR over the neighbour list rows)
    bool csr_ready = false;
    // The mesh was zeroed ahead of a host wait inside one GMRES call; the next spread skips
    // its own clear.  Set and reset only within gmres().
    bool mesh_zeroed = false;
    // FP32 charges already in qf (with their partial maxima in fxs) for the operator input
    // qf_for: set by gmres() for one application, from its step tail
    const double* qf_for = nullptr;
    int qf_parts = 0;
    // partial maxima of the charges in qf (fxs->part), from which the next fixed-point
    // spread forms its scale
    int fx_parts = 0;
    long rs_ldn = 0;
    DevBuf rs_col, rs_val, rs_len;
    DevBuf qf;  // charges rounded to float for the FP32-bulk kernels
    // FP32-bulk fixed-point spread (k_spread_fx.hip): its scale state, LDS plan, whether
    // the weight bound matches the current binning, and whether the mesh holds int32
    // fixed-point values (the next FFT converts them)
    DevBuf fxs;
    DevBuf fxblk;  // the fixed-point spread's per-tile integer blocks (tg.ntiles x fx_block_ints)
    FxPlan fxp{}, fxcp{};  // spread and potential-contraction LDS plans
    bool fx_bound = false;
    size_t lds_max = 65536;
    // bead binning for the owner-computes spread
    TileGeom tg{};
    DevBuf bead_tile, tile_count, tile_start, tile_cursor, sorted, scan_sums;
    DevBuf rec_off, rec_win;  // per-bead window records in binned order
    // Krylov workspace
    int R = 10;
    long ld = 0;
    int n_cap = 0;
    DevBuf V, w, b, Mb, x, partial, scal, yraw;
    // near-field block preconditioner: beads grouped by body tag, per-body inverses (its
    // own scan scratch: its setup runs beside the binning, which scans into scan_sums)
    int block_nmax = 0;
    DevBuf body_count, body_start, body_beads, ginv, uvec, inv_s, block_scan;
    // the block preconditioner's slots (processing order of the bodies, their bead counts
    // and member lists) and their counting-sort scratch
    DevBuf body_order, slot_count, slot_members, order_key, order_mkey, order_count, order_start, order_scan;
    // per slot: the body, bead count, bead-index sum and shape signature of the stored inverse
    DevBuf bcache_key, bcache_sig;
    bool block_cache_ok = false;
    // Independent parts of a force evaluation run beside the caller's stream on these
    // (default priority): the real-space ELL build and the block preconditioner's setup
    // next to the binning, and the real-space force next to the force pass's mesh work,
    // which then adds it (force_rs) to the reciprocal force.
    static constexpr int kSide = 2;
    hipStream_t side[kSide] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSide] = {};
    DevBuf force_rs;
    // per-body segment scratch: bead counts and double sums (warm starts, learned features)
    DevBuf seg_count, seg_sum;
    DevBuf bd_force;  // Brownian step: bead forces plus steric repulsion
    GmresState* g_dev = nullptr;
    GmresState* g_host = nullptr;
    HostPub* hpub = nullptr;            // coherent pinned, kPubSlots mailboxes: residuals the host polls for
    unsigned* hwidth = nullptr;         // pinned: the ELL width read back during the setup
    hipEvent_t ev_width = nullptr;      // ... and its arrival
    unsigned long long pub_seq = 0;
    int P_eta = 0;   // window support before the clamp to the mesh (eta uses it)
    // cr_compute_force_f32: double copies of the caller's float arrays
    DevBuf f32_pos, f32_com, f32_bq, f32_q, f32_phi, f32_force;
    double E0[3] = {0, 0, 0};
    unsigned long generation = 0;  // bumped whenever parameters, tables or tiling change

    ~cr_context()
    {
        for (int k = 0; k < kSide; ++k) {
            if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
            if (side[k]) (void)hipStreamDestroy(side[k]);
        }
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_width) (void)hipEventDestroy(ev_width);
        if (hwidth) (void)hipHostFree(hwidth);
        if (g_dev) (void)hipFree(g_dev);
        if (g_host) (void)hipHostFree(g_host);
        if (hpub) (void)hipHostFree(hpub);
    }

    bool fp32() const { return params.precision == CR_PRECISION_FP32_BULK; }
    bool fx_on() const { return fp32() && tg.valid && fxp.valid; }
};

namespace cr::impl {

SysView make_view(cr_context* c, const cr_system* s);
// Take up a new box (cr_set_box): device-side regeneration of the box-dependent quantities.
int set_box(cr_context* c, const double L[3], hipStream_t st);
// Validate the system and refresh the body-tag map; with bin, also rebin beads and rebuild
// window records.
int prepare(cr_context* c, const cr_system* s, hipStream_t st, SysView* out, bool bin = true);
// prepare without its launches (checks, box take-up, buffers, plans): the batched boxes
// run the membership, binning and records of all boxes in a few launches of their own.
int prepare_host(cr_context* c, const cr_system* s, hipStream_t st, SysView* out, bool bin = true);
BinView bin_view(const cr_context* c);
bool contract_tiled_ok(const cr_context* c, size_t elem);
// Real-space ELL operator for the current positions (valid until the next prepare()).
int ensure_csr(cr_context* c, const SysView& v, hipStream_t st);
// The two halves of ensure_csr for callers that prepare many systems: the ELL width into
// dwidth (device), then, with the width read back, the allocation and the build.
int csr_width(const SysView& v, unsigned* dwidth, hipStream_t st);
int csr_build(cr_context* c, const SysView& v, unsigned width, hipStream_t st);
// The batched ELL build's part for one box: its ELL allocated for width and described in
// *pb (the caller queues the build and then sets csr_ready).
int csr_pair_box(cr_context* c, const SysView& v, unsigned width, PairBox* pb);
// The batched pair forces' descriptor of one box (force += pair forces + q E0).
void force_pair_box(const cr_context* c, const SysView& v, const double* q, double* force, PairBox* pb);
// grid <- Re IFFT(S FFT(grid)) on the context's mesh, in its precision.
int recip_any(cr_context* c, hipStream_t st);
// FFT plans of the context's precision (x, y, z).
const AxisPlan* axis_plans(const cr_context* c);
// rows = false leaves out the constraint rows y[N:N+B)
int matvec_any(cr_context* c, const SysView& v, const double* x, double* y, int precond, hipStream_t st,
               bool rows = true);
bool block_precond(const cr_context* c);
// Group beads by body and (invert) form the per-body inverses of the block preconditioner.
// the block preconditioner's buffers for v (block_setup without launches)
int block_alloc(cr_context* c, const SysView& v);
int block_setup(cr_context* c, const SysView& v, hipStream_t st, bool invert = true);
// xq (block preconditioner only): the constraint rows are formed from xq instead of read from in
int apply_precond(cr_context* c, const SysView& v, const double* in, double* out, hipStream_t st,
                  const double* xq = nullptr);
// x (length N + B) from the x0 mode; uses the context's w as scratch for CR_X0_CHARGES.
int init_x0(cr_context* c, const SysView& v, int x0_mode, const double* q_in, const double* phi_in, double* x,
            hipStream_t st);
// FP64 residuals of x against the context's objective b: ||D r|| and ||r|| / ||b||.
int true_residual(cr_context* c, const SysView& v, const double* x, cr_solve_stats* stats, hipStream_t st);
// force += q E0 + real-space pair forces (the last step of the force pass).
int real_space_force_any(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st);
// Final force pass for charges q (spread, k-space, contract_force, real-space force + q E0).
int force_pass_any(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st);

}  // namespace cr::impl
