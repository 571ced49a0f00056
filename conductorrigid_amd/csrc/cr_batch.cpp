// cr_batch.cpp -- many independent periodic boxes solved together (cr_batch_* in
// include/conductorrigid.h).
//
// The reference computes one box per computeForces call (cppmodule/ConductorRigid.cc:403-557),
// and its solve (cppmodule/ConductorRigid.cu:1051-1151) syncs the host once per dot product.
// For a set of small boxes, each box alone is too small to fill the GPU.  Here each box keeps
// its own context (parameters, mesh, FFT plan, tiling chosen by the autotuner) and its own
// system, and the boxes go through GMRES in lockstep:
//   * their unknowns [q ; Phi] are concatenated into one batch vector;
//   * every operator application is one fill, one binned spread and one binned contraction
//     over the tiles of all boxes, one launch each for the real-space rows and the rigid-body
//     rows, and five FFT launches per distinct mesh: boxes that share a mesh keep their
//     meshes back to back in one batch-owned array, so the z and y passes see one taller
//     mesh and the fused x pass + Green's function scale takes the box from blockIdx.y;
//   * every Arnoldi step is one set of segmented Krylov launches for all boxes, taking its
//     column index from a device counter, so the whole step is identical from step to step.
//     It is captured once as a HIP graph and replayed, with one host sync per step for
//     all boxes (the residual estimates and converged flags come back together).
// Each box keeps its own tolerance, stopping rule, restart cycle and iteration count; a box
// that converges is frozen (its workgroups return at once) while the others go on.
#include "cr_context_impl.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

using namespace cr;
using namespace cr::impl;

namespace {

constexpr int kSideStreams = 4;  // FFT chains of different mesh groups overlap on these

// Boxes of one mesh size: their meshes back to back in batch-owned arrays.
struct MeshGroup {
    int N[3] = {0, 0, 0};
    int Zc = 0;
    long G = 0, Gh = 0;
    std::vector<int> boxes;
    cr_context* c0 = nullptr;  // FFT plans (equal lengths share twiddles)
    DevBuf real, cplx, sct;
};

}  // namespace

struct cr_batch {
    int device = 0;
    hipStream_t st = nullptr;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork = nullptr, join[kSideStreams] = {}, done_ev = nullptr;
    // arena of the lockstep solve
    int nbox = 0, R = 0;
    long n_tot = 0, ld = 0;
    int seg_blocks = 0, bead_blocks = 0, rigid_blocks = 0, total_tiles = 0;
    size_t lds_spread = 0, lds_contract = 0;
    long Gmax = 0;
    DevBuf V, u, w, x, b, Mb, yraw, partial, uf;
    DevBuf d_spread, d_contract, d_rows, d_segs, d_gs, d_skip, d_finished, d_step, d_pub, d_rel, d_blocks;
    DevBuf d_width;  // ELL width per box
    DevBuf d_pairs, d_fpairs;  // descriptors of the batched ELL build and pair forces
    std::vector<PairBox> pair_boxes, fpair_boxes;  // their host copies (alive while the copies run)
    DevBuf d_prep;                                 // descriptors of the batched preparation
    std::vector<PrepBox> prep_boxes;
    int total_bodies = 0;
    std::vector<BlockBox> blocks;
    double* h_pub = nullptr;  // pinned: 3 doubles per box
    size_t h_pub_n = 0;
    std::vector<TiledBox> spread_boxes, contract_boxes;
    // the fixed-point path: every box FP32 bulk with valid fixed-point plans
    bool fx = false, fx_v4 = true;
    int fx_max_n = 0;
    long fx_gblocks = 0;
    size_t lds_fx_s = 0, lds_fx_c = 0;
    std::vector<FxBox> fx_boxes;
    DevBuf d_fx;
    std::vector<RowBox> rows;
    std::vector<BatchSeg> segs;
    std::vector<SysView> views;
    std::vector<cr_context*> ctx;
    std::vector<std::unique_ptr<MeshGroup>> groups;
    // captured Arnoldi step and the launch configuration it was captured for
    hipGraphExec_t step_exec = nullptr;
    std::vector<const void*> step_sig;

    ~cr_batch()
    {
        if (step_exec) (void)hipGraphExecDestroy(step_exec);
        if (h_pub) (void)hipHostFree(h_pub);
        for (int k = 0; k < kSideStreams; ++k) {
            if (join[k]) (void)hipEventDestroy(join[k]);
            if (side[k]) (void)hipStreamDestroy(side[k]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (done_ev) (void)hipEventDestroy(done_ev);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

// A box the batched fixed-point kernels can take: FP32 bulk on binned tiles whose
// fixed-point spread and contraction plans fit.
bool batch_fx_capable(const cr_context* c) { return c->fp32() && c->tg.valid && c->fxp.valid && c->fxcp.valid; }

// grid <- Re IFFT(S FFT(grid)) for every box of a mesh group: the z and y passes treat the
// group as one mesh of K Nx planes; the x pass takes each box's own scale tables.
template <typename T>
int group_recip(MeshGroup& g, const std::vector<cr_context*>& ctx, hipStream_t st)
{
    const AxisPlan* pl = axis_plans(g.c0);
    const int K = (int)g.boxes.size();
    const int Nx = g.N[0], Ny = g.N[1];
    T* real = g.real.as<T>();
    T* cp = g.cplx.as<T>();
    CR_HIP(launch_fft_z_r2c<T>(pl[2], K * Nx, Ny, real, cp, g.Zc, st));
    CR_HIP(launch_fft_yx<T>(pl[1], 1, K * Nx, Ny, g.Zc, g.Zc, cp, -1, st));
    if (fftr_radix(Nx)) {
        CR_HIP(launch_fftr_x_scale_batch<T>(pl[0], K, Ny, g.Zc, g.Zc, cp, g.sct.as<ScaleTables>(), st));
    } else {
        for (int k = 0; k < K; ++k)
            CR_HIP(launch_fft_x_scale<T>(pl[0], Nx, Ny, g.Zc, g.Zc, cp + 2 * (long)k * g.Gh, ctx[g.boxes[k]]->sc,
                                         st));
    }
    CR_HIP(launch_fft_yx<T>(pl[1], 1, K * Nx, Ny, g.Zc, g.Zc, cp, +1, st));
    CR_HIP(launch_fft_z_c2r<T>(pl[2], K * Nx, Ny, cp, g.Zc, real, st));
    return CR_OK;
}

// Spread of u (the descriptors' input) and the k-space step for every box not skipped.
// host_skip (may be null) also drops mesh groups whose boxes are all frozen.
int batch_spread_recip(cr_batch* B, const int* skip, const std::vector<char>* host_skip, hipStream_t st)
{
    cr_context* c0 = B->ctx[0];
    const bool f32 = c0->fp32();
    const int P = c0->mp.P;
    const TiledBox* ds = B->d_spread.as<TiledBox>();
    if (B->fx) {
        // fixed-point spread: per-box scale, tile blocks, gather (every node written: no clear)
        const FxBox* dx = B->d_fx.as<FxBox>();
        CR_HIP(launch_charges_fx_batch(dx, B->nbox, B->fx_max_n, skip, st));
        CR_HIP(launch_spread_fx_batch(dx, B->nbox, P, B->total_tiles, B->lds_fx_s, B->fx_gblocks, B->fx_v4, skip, st));
    } else if (f32) {
        CR_HIP(launch_fill_batch<float>(ds, B->nbox, B->Gmax, skip, st));
        CR_HIP(launch_spread_tiled_batch<float>(ds, B->nbox, B->total_tiles, P, B->lds_spread, skip, st));
    } else {
        CR_HIP(launch_fill_batch<double>(ds, B->nbox, B->Gmax, skip, st));
        CR_HIP(launch_spread_tiled_batch<double>(ds, B->nbox, B->total_tiles, P, B->lds_spread, skip, st));
    }
    CR_HIP(hipEventRecord(B->fork, st));
    for (int k = 0; k < kSideStreams; ++k) CR_HIP(hipStreamWaitEvent(B->side[k], B->fork, 0));
    for (size_t g = 0; g < B->groups.size(); ++g) {
        MeshGroup& mg = *B->groups[g];
        bool any = !host_skip;
        for (int b : mg.boxes) any = any || !(*host_skip)[b];
        if (!any) continue;
        const hipStream_t sg = B->side[g % kSideStreams];
        CR_RET(f32 ? group_recip<float>(mg, B->ctx, sg) : group_recip<double>(mg, B->ctx, sg));
    }
    for (int k = 0; k < kSideStreams; ++k) {
        CR_HIP(hipEventRecord(B->join[k], B->side[k]));
        CR_HIP(hipStreamWaitEvent(st, B->join[k], 0));
    }
    return CR_OK;
}

// out = A u, or D A u with precond, for every box not skipped (u = the batch vector the
// spread descriptors read).
int batch_matvec(cr_batch* B, const double* in, double* out, int precond, const int* skip,
                 const std::vector<char>* host_skip, hipStream_t st)
{
    cr_context* c0 = B->ctx[0];
    const bool f32 = c0->fp32();
    const bool blk = precond && block_precond(c0);
    const int P = c0->mp.P;
    const TiledBox* dc = B->d_contract.as<TiledBox>();
    CR_RET(batch_spread_recip(B, skip, host_skip, st));
    if (B->fx)
        CR_HIP(launch_contract_fx_batch(B->d_fx.as<FxBox>(), B->nbox, P, B->total_tiles, B->lds_fx_c, false, skip, st));
    else if (f32)
        CR_HIP(launch_contract_tiled_batch<float>(dc, B->nbox, B->total_tiles, P, B->lds_contract, skip, st));
    else
        CR_HIP(launch_contract_tiled_batch<double>(dc, B->nbox, B->total_tiles, P, B->lds_contract, skip, st));
    double* rows_out = blk ? B->yraw.as<double>() : out;
    if (f32) {
        // the pair sum's gathered charges, rounded to float once per application
        CR_HIP((launch_convert<double, float>(in, B->uf.as<float>(), B->n_tot, st)));
        CR_HIP(launch_brows<float>(B->d_rows.as<RowBox>(), B->nbox, B->bead_blocks, B->rigid_blocks, in,
                                   B->uf.as<float>(), rows_out, (precond && !blk) ? 1 : 0, skip, st));
    } else {
        CR_HIP(launch_brows<double>(B->d_rows.as<RowBox>(), B->nbox, B->bead_blocks, B->rigid_blocks, in, in,
                                    rows_out, (precond && !blk) ? 1 : 0, skip, st));
    }
    if (blk)
        CR_HIP(launch_block_apply_batch(B->d_blocks.as<BlockBox>(), B->nbox, B->total_bodies, B->yraw.as<double>(), out,
                                        skip, st));
    return CR_OK;
}

// One lockstep Arnoldi step of every box not frozen: column *step of V -> w = D A v,
// CGS2 against columns 0..*step, Givens, V[:, *step+1] = w / ||w||, publish, ++*step.
int batch_step(cr_batch* B, const std::vector<char>* host_skip, hipStream_t st)
{
    const int nb = B->nbox;
    int* skip = B->d_skip.as<int>();
    int* step = B->d_step.as<int>();
    double* V = B->V.as<double>();
    double* w = B->w.as<double>();
    double* part = B->partial.as<double>();
    const BatchSeg* sg = B->d_segs.as<BatchSeg>();
    BatchGmres* gs = B->d_gs.as<BatchGmres>();
    CR_HIP(launch_bcopy_col(V, B->ld, step, B->u.as<double>(), B->n_tot, st));
    CR_RET(batch_matvec(B, B->u.as<double>(), w, 1, skip, host_skip, st));
    CR_HIP(launch_bdot(V, B->ld, 0, step, w, sg, nb, B->seg_blocks, part, skip, st));
    CR_HIP(launch_breduce(part, sg, nb, 0, step, gs, kRedH1, skip, st));
    CR_HIP(launch_bgs_update(V, B->ld, step, w, sg, nb, B->seg_blocks, gs, kRedH1, 0, part, skip, st));
    CR_HIP(launch_bdot(V, B->ld, 0, step, w, sg, nb, B->seg_blocks, part, skip, st));
    CR_HIP(launch_breduce(part, sg, nb, 0, step, gs, kRedH2, skip, st));
    CR_HIP(launch_bgs_update(V, B->ld, step, w, sg, nb, B->seg_blocks, gs, kRedH2, 2, part, skip, st));
    CR_HIP(launch_breduce(part, sg, nb, 1, step, gs, kRedWnorm, skip, st));
    CR_HIP(launch_bgivens(gs, nb, step, B->R, skip, st));
    CR_HIP(launch_bscale_copy(w, sg, nb, B->seg_blocks, gs, step, -1, 1.0, V, B->ld, skip, st));
    CR_HIP(launch_bstep_end(step, gs, skip, nb, B->d_pub.as<double>(), 1, st));
    CR_HIP(hipMemcpyAsync(B->h_pub, B->d_pub.as<double>(), sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
    return CR_OK;
}

// Everything the captured step depends on by address or launch shape.
std::vector<const void*> step_signature(const cr_batch* B)
{
    std::vector<const void*> s = {B->V.p, B->u.p, B->uf.p, B->w.p, B->yraw.p, B->partial.p, B->d_spread.p, B->d_contract.p,
                                  B->d_rows.p, B->d_segs.p, B->d_gs.p, B->d_skip.p, B->d_step.p, B->d_pub.p,
                                  B->d_blocks.p, B->h_pub, B->d_fx.p};
    const long shape[] = {B->nbox, B->R, B->n_tot, B->ld, B->seg_blocks, B->bead_blocks, B->rigid_blocks,
                          B->total_tiles, (long)B->lds_spread, (long)B->lds_contract, B->Gmax, B->total_bodies,
                          (long)B->fx, (long)B->fx_v4, (long)B->fx_max_n, B->fx_gblocks, (long)B->lds_fx_s,
                          (long)B->lds_fx_c};
    for (long v : shape) s.push_back(reinterpret_cast<const void*>(v));
    for (const auto& g : B->groups) {
        s.push_back(g->real.p);
        s.push_back(g->cplx.p);
        s.push_back(g->sct.p);
        s.push_back(reinterpret_cast<const void*>((long)g->boxes.size()));
    }
    for (cr_context* c : B->ctx) {
        s.push_back(c);
        s.push_back(reinterpret_cast<const void*>(c->generation));
        s.push_back(c->phi_recip.p);
        s.push_back(c->qf.p);
        s.push_back(c->fxs.p);
        s.push_back(c->fxblk.p);
    }
    return s;
}

int capture_step(cr_batch* B)
{
    if (B->step_exec) {
        (void)hipGraphExecDestroy(B->step_exec);
        B->step_exec = nullptr;
    }
    B->step_sig.clear();
    CR_HIP(hipStreamBeginCapture(B->st, hipStreamCaptureModeRelaxed));
    const int rc = batch_step(B, nullptr, B->st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->st, &g);
    if (rc != CR_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    CR_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&B->step_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    CR_HIP(ei);
    B->step_sig = step_signature(B);
    return CR_OK;
}

int check_batch(int nbox, cr_context* const* ctx, const cr_system* sys)
{
    if (nbox <= 0 || !ctx || !sys) return set_error(CR_ERR_INVALID, "empty batch");
    for (int b = 0; b < nbox; ++b) {
        const cr_context* c = ctx[b];
        if (!c || !c->ready) return set_error(CR_ERR_STATE, "every box needs a context after cr_set_params");
        if (c->device != ctx[0]->device) return set_error(CR_ERR_INVALID, "boxes of one batch share a device");
        if (c->params.precision != ctx[0]->params.precision)
            return set_error(CR_ERR_INVALID, "boxes of one batch share the precision");
        if (c->mp.P != ctx[0]->mp.P) return set_error(CR_ERR_INVALID, "boxes of one batch share the window support P");
        if (c->R != ctx[0]->R) return set_error(CR_ERR_INVALID, "boxes of one batch share the GMRES restart");
        if (c->res.max_iter != ctx[0]->res.max_iter)
            return set_error(CR_ERR_INVALID, "boxes of one batch share the GMRES iteration limit");
        if (c->params.preconditioner != ctx[0]->params.preconditioner)
            return set_error(CR_ERR_INVALID, "boxes of one batch share the preconditioner");
        if (!c->tg.valid) return set_error(CR_ERR_INVALID, "every box of a batch needs a binned tiling of its mesh");
        for (int k = 0; k < b; ++k)
            if (ctx[k] == c) return set_error(CR_ERR_INVALID, "a context appears twice in one batch");
    }
    return CR_OK;
}

}  // namespace

extern "C" {

int cr_batch_create(cr_batch** out, int device)
{
    if (!out) return set_error(CR_ERR_INVALID, "null output pointer");
    *out = nullptr;
    CR_HIP(hipSetDevice(device));
    cr_batch* B = new cr_batch();
    B->device = device;
    bool ok = hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&B->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&B->done_ev, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < kSideStreams; ++k)
        ok = hipStreamCreateWithFlags(&B->side[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&B->join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete B;
        return set_error(CR_ERR_DEVICE, "stream or event creation failed");
    }
    *out = B;
    return CR_OK;
}

void cr_batch_destroy(cr_batch* b) { delete b; }

int cr_batch_compute_force(cr_batch* B, int nbox, cr_context* const* ctx, const cr_system* sys,
                           double* const* q_inout, double* const* phi_inout, double* const* force, int x0_mode,
                           int use_graph, cr_solve_stats* stats, void* stream)
{
    if (!B) return set_error(CR_ERR_INVALID, "null batch");
    CR_RET(check_batch(nbox, ctx, sys));
    if (!q_inout || !phi_inout || !force) return set_error(CR_ERR_INVALID, "null output array list");
    if (x0_mode != CR_X0_ZERO && x0_mode != CR_X0_PREVIOUS && x0_mode != CR_X0_CHARGES)
        return set_error(CR_ERR_INVALID, "unknown x0 mode");
    CR_HIP(hipSetDevice(B->device));
    hipStream_t caller = (hipStream_t)stream;
    hipStream_t st = B->st;
    // order after the caller's earlier work
    CR_HIP(hipEventRecord(B->done_ev, caller));
    CR_HIP(hipStreamWaitEvent(st, B->done_ev, 0));

    cr_context* c0 = ctx[0];
    const bool f32 = c0->fp32();
    const size_t Tsz = f32 ? sizeof(float) : sizeof(double);
    B->nbox = nbox;
    B->R = c0->R;
    B->ctx.assign(ctx, ctx + nbox);
    B->views.resize(nbox);
    B->segs.resize(nbox);
    B->rows.resize(nbox);
    B->spread_boxes.resize(nbox);
    B->contract_boxes.resize(nbox);
    B->blocks.resize(nbox);

    // per-box preparation (binning, records, ELL rows, objective, preconditioner setup)
    long xoff = 0;
    int blk = 0, bblk = 0, rblk = 0, tiles = 0, bodies = 0, nmax_all = 1;
    size_t lds_s = 0, lds_c = 0;
    long Gmax = 0;
    // binning, records, objective, body grouping and ELL widths of every box in a handful of
    // launches over all boxes (each box owns a run of workgroups), then the widths read
    // back with one synchronisation and the ELL rows of all boxes in one launch
    CR_HIP(B->d_width.ensure(sizeof(unsigned) * nbox));
    unsigned* d_width = B->d_width.as<unsigned>();
    {
        std::vector<PrepBox>& pb = B->prep_boxes;
        pb.assign(nbox, PrepBox{});
        int pblk = 0;
        for (int b = 0; b < nbox; ++b) {
            cr_context* c = ctx[b];
            SysView& v = B->views[b];
            CR_RET(prepare_host(c, &sys[b], st, &v));
            const long n = (long)v.N + v.N_isolated;
            if (n > c->n_cap) return set_error(CR_ERR_INVALID, "system larger than the size given to cr_set_params");
            if ((v.N > 0 && (!q_inout[b] || !force[b])) || (v.N_isolated > 0 && !phi_inout[b]))
                return set_error(CR_ERR_INVALID, "null output array");
            PrepBox& d = pb[b];
            d.mp = c->mp;
            d.tg = c->tg;
            d.s = v;
            d.blk0 = pblk;
            d.bin = c->tg.valid ? 1 : 0;
            d.tag = sys[b].tag;
            d.body_of_tag = sys[b].body_of_tag;
            d.iso = c->iso.as<unsigned>();
            if (d.bin) {
                d.bead_tile = c->bead_tile.as<unsigned>();
                d.tile_count = c->tile_count.as<unsigned>();
                d.tile_start = c->tile_start.as<unsigned>();
                d.cursor = c->tile_cursor.as<unsigned>();
                d.sorted = c->sorted.as<unsigned>();
                d.rec_off = c->rec_off.as<int>();
                d.rec_win = c->rec_win.p;
            }
            d.b = c->b.as<double>();
            for (int k = 0; k < 3; ++k) d.E0[k] = c->E0[k];
            if (block_precond(c)) {  // body grouping here, the inverses in one launch below
                CR_RET(block_alloc(c, v));
                c->block_cache_ok = false;  // the batched setup stores the inverses by body, not by slot
                d.body_count = c->body_count.as<unsigned>();
                d.body_start = c->body_start.as<unsigned>();
                d.body_beads = c->body_beads.as<unsigned>();
            }
            d.width = d_width + b;
            if (batch_fx_capable(c)) {
                // the box's charge and block buffers; its weight bound is zeroed here and
                // taken over all boxes once the descriptors are up
                CR_HIP(c->qf.ensure(sizeof(float) * (size_t)std::max(1, v.N)));
                CR_HIP(c->fxblk.ensure(sizeof(int) * (size_t)c->tg.ntiles * (size_t)fx_block_ints(c->mp, c->tg)));
                d.wbits = &c->fxs.as<FxScale>()->wbits;
            }
            long span = std::max<long>({(long)v.N, (long)v.N_rigid * v.Nmax, (long)v.N_isolated + 1, 1L});
            if (d.bin) span = std::max(span, (long)c->tg.ntiles + 1);
            pblk += (int)((span + 255) / 256);
        }
        // the objective also goes straight into the batch vector
        long n_tot = 0;
        for (int b = 0; b < nbox; ++b) n_tot += (long)B->views[b].N + B->views[b].N_isolated;
        CR_HIP(B->b.ensure(sizeof(double) * (size_t)std::max(1L, (n_tot + 63) / 64 * 64)));
        for (long b = 0, o = 0; b < nbox; o += (long)B->views[b].N + B->views[b].N_isolated, ++b)
            pb[b].bcat = B->b.as<double>() + o;
        CR_HIP(B->d_prep.ensure(sizeof(PrepBox) * nbox));
        CR_HIP(hipMemcpyAsync(B->d_prep.p, pb.data(), sizeof(PrepBox) * nbox, hipMemcpyHostToDevice, st));
        CR_HIP(launch_prep_batch(B->d_prep.as<PrepBox>(), nbox, pblk, f32, st));
        CR_HIP(launch_objective_batch(B->d_prep.as<PrepBox>(), nbox, pblk, st));
    }
    {
        std::vector<unsigned> width(nbox);
        CR_HIP(hipMemcpyAsync(width.data(), d_width, sizeof(unsigned) * nbox, hipMemcpyDeviceToHost, st));
        CR_HIP(hipStreamSynchronize(st));
        // the ELL rows of all boxes in one launch (each box's ELL sized by its width)
        std::vector<PairBox>& pb = B->pair_boxes;
        pb.clear();
        std::vector<int> built;
        int pblk = 0;
        for (int b = 0; b < nbox; ++b) {
            cr_context* c = ctx[b];
            const SysView& v = B->views[b];
            if (c->csr_ready || v.N <= 0) continue;
            PairBox d;
            CR_RET(csr_pair_box(c, v, width[b], &d));
            d.blk0 = pblk;
            pblk += (v.N + 255) / 256;
            pb.push_back(d);
            built.push_back(b);
        }
        if (!pb.empty()) {
            CR_HIP(B->d_pairs.ensure(sizeof(PairBox) * pb.size()));
            CR_HIP(hipMemcpyAsync(B->d_pairs.p, pb.data(), sizeof(PairBox) * pb.size(), hipMemcpyHostToDevice, st));
            if (f32)
                CR_HIP(launch_real_csr_build_batch<float>(B->d_pairs.as<PairBox>(), (int)pb.size(), pblk, st));
            else
                CR_HIP(launch_real_csr_build_batch<double>(B->d_pairs.as<PairBox>(), (int)pb.size(), pblk, st));
            for (int b : built) ctx[b]->csr_ready = true;
        }
    }
    // the fixed-point path when every box can take it (one gather shape for the launch)
    B->fx = f32;
    B->fx_v4 = true;
    for (int b = 0; b < nbox; ++b) {
        B->fx = B->fx && batch_fx_capable(ctx[b]);
        B->fx_v4 = B->fx_v4 && fx_gather_v4(ctx[b]->mp, ctx[b]->tg);
    }
    B->fx_boxes.resize(nbox);
    B->fx_max_n = 0;
    B->lds_fx_s = B->lds_fx_c = 0;
    long gblk = 0;
    for (int b = 0; b < nbox; ++b) {
        cr_context* c = ctx[b];
        SysView& v = B->views[b];
        const long n = (long)v.N + v.N_isolated;

        if (B->fx) {
            FxBox& f = B->fx_boxes[b];
            f = FxBox{};
            f.mp = c->mp;
            f.tg = c->tg;
            f.box = b;
            f.tile0 = tiles;
            f.gblk0 = (int)gblk;
            f.N = v.N;
            f.ch = c->fxp.chunk;
            f.cch = c->fxcp.chunk;
            f.bstride = fx_block_ints(c->mp, c->tg);
            const BinView bv = bin_view(c);
            f.tile_start = bv.tile_start;
            f.sorted = bv.sorted;
            f.rec_off = bv.rec_off;
            f.rec_u = static_cast<const float4*>(bv.rec_win);
            f.qf = c->qf.as<float>();
            f.fs = c->fxs.as<FxScale>();
            f.blocks = c->fxblk.as<int>();
            f.phi = c->phi_recip.as<float>();
            gblk += fx_gather_blocks(c->mp, B->fx_v4);
            B->fx_max_n = std::max(B->fx_max_n, v.N);
            B->lds_fx_s = std::max(B->lds_fx_s, c->fxp.lds_bytes);
            B->lds_fx_c = std::max(B->lds_fx_c, c->fxcp.lds_bytes);
        }

        BlockBox& bb = B->blocks[b];
        bb = BlockBox{};
        bb.xoff = xoff;
        bb.N = v.N;
        bb.nmax = c->block_nmax;
        bb.body0 = bodies;
        bb.start = c->body_start.as<unsigned>();
        bb.beads = c->body_beads.as<unsigned>();
        bb.ginv = c->ginv.as<float>();
        bb.uvec = c->uvec.as<float>();
        bb.inv_s = c->inv_s.as<double>();
        bb.pos = v.pos;
        for (int d = 0; d < 3; ++d) bb.L[d] = c->mp.L[d];
        bodies += v.N_isolated;
        nmax_all = std::max(nmax_all, c->block_nmax);

        BatchSeg& s = B->segs[b];
        s.xoff = xoff;
        s.n = (int)n;
        s.blk0 = blk;
        s.nblk = batch_blocks(n);
        blk += s.nblk;

        const TiledLds ls = tiled_lds(c->mp, c->tg, sizeof(SpreadAcc), Tsz);
        const TiledLds lc = tiled_lds(c->mp, c->tg, Tsz, Tsz);
        if (ls.chunk < 4 || lc.chunk < 4) return set_error(CR_ERR_INVALID, "tiling leaves no LDS for bead staging");
        lds_s = std::max(lds_s, ls.total_bytes);
        lds_c = std::max(lds_c, lc.total_bytes);
        TiledBox tb{};
        tb.mp = c->mp;
        tb.tg = c->tg;
        tb.tile0 = tiles;
        tb.box = b;
        tb.G = c->G;
        const BinView bv = bin_view(c);
        tb.tile_start = bv.tile_start;
        tb.sorted = bv.sorted;
        tb.rec_off = bv.rec_off;
        tb.rec_win = bv.rec_win;
        tb.phi = c->phi_recip.p;
        tiles += c->tg.ntiles;
        Gmax = std::max(Gmax, c->G);
        B->spread_boxes[b] = tb;
        B->spread_boxes[b].ch = ls.chunk;
        B->contract_boxes[b] = tb;
        B->contract_boxes[b].ch = lc.chunk;

        RowBox& r = B->rows[b];
        r = RowBox{};
        r.xoff = xoff;
        r.N = v.N;
        r.N_rigid = v.N_rigid;
        r.N_isolated = v.N_isolated;
        r.Nmax = v.Nmax;
        r.blk0 = bblk;
        r.rblk0 = rblk;
        bblk += (v.N + kBatchBlock - 1) / kBatchBlock;
        rblk += (v.N_rigid + kBatchBlock - 1) / kBatchBlock;
        r.iso = v.iso;
        r.rigid_membership = v.rigid_membership;
        r.rigid_members = v.rigid_members;
        r.rigid_size = v.rigid_size;
        r.phi_recip = c->phi_recip.p;
        r.ldn = c->rs_ldn;
        r.col = c->rs_col.as<unsigned>();
        r.val = c->rs_val.p;
        r.row_len = c->rs_len.as<unsigned>();
        r.selfterm = c->rp.selfterm;
        xoff += n;
    }
    B->n_tot = xoff;
    B->ld = (xoff + 63) / 64 * 64;
    B->seg_blocks = blk;
    B->bead_blocks = bblk;
    B->rigid_blocks = rblk;
    B->total_tiles = tiles;
    B->total_bodies = bodies;
    B->lds_spread = lds_s;
    B->lds_contract = lds_c;
    B->fx_gblocks = gblk;
    B->Gmax = Gmax;

    // boxes that share a mesh keep their meshes back to back (buffers reused across calls)
    {
        std::vector<std::unique_ptr<MeshGroup>> old = std::move(B->groups);
        B->groups.clear();
        for (int b = 0; b < nbox; ++b) {
            const MeshParams& mp = ctx[b]->mp;
            MeshGroup* g = nullptr;
            for (auto& e : B->groups)
                if (e->N[0] == mp.Nx && e->N[1] == mp.Ny && e->N[2] == mp.Nz) g = e.get();
            if (!g) {
                std::unique_ptr<MeshGroup> ng;
                for (auto& e : old)
                    if (e && e->N[0] == mp.Nx && e->N[1] == mp.Ny && e->N[2] == mp.Nz) ng = std::move(e);
                if (!ng) ng = std::make_unique<MeshGroup>();
                ng->N[0] = mp.Nx;
                ng->N[1] = mp.Ny;
                ng->N[2] = mp.Nz;
                ng->Zc = mp.Nz / 2 + 1;
                ng->G = (long)mp.Nx * mp.Ny * mp.Nz;
                ng->Gh = (long)mp.Nx * mp.Ny * ng->Zc;
                ng->boxes.clear();
                ng->c0 = ctx[b];
                g = ng.get();
                B->groups.push_back(std::move(ng));
            }
            g->boxes.push_back(b);
        }
        for (auto& g : B->groups) {
            const size_t K = g->boxes.size();
            CR_HIP(g->real.ensure(Tsz * K * (size_t)g->G));
            CR_HIP(g->cplx.ensure(2 * Tsz * K * (size_t)g->Gh));
            CR_HIP(g->sct.ensure(sizeof(ScaleTables) * K));
            std::vector<ScaleTables> sct(K);
            for (size_t k = 0; k < K; ++k) {
                const int b = g->boxes[k];
                sct[k] = ctx[b]->sc;
                char* slice = g->real.as<char>() + Tsz * k * (size_t)g->G;
                B->spread_boxes[b].grid = slice;
                B->contract_boxes[b].grid = slice;
                if (B->fx) B->fx_boxes[b].grid = reinterpret_cast<float*>(slice);
            }
            CR_HIP(hipMemcpy(g->sct.p, sct.data(), sizeof(ScaleTables) * K, hipMemcpyHostToDevice));
        }
    }

    // arena and descriptors
    const size_t nv = sizeof(double) * (size_t)std::max(1L, B->ld);
    CR_HIP(B->V.ensure(nv * (size_t)(B->R + 1)));
    CR_HIP(B->u.ensure(nv));
    CR_HIP(B->w.ensure(nv));
    CR_HIP(B->x.ensure(nv));
    CR_HIP(B->b.ensure(nv));
    CR_HIP(B->Mb.ensure(nv));
    CR_HIP(B->yraw.ensure(nv));
    CR_HIP(B->uf.ensure(sizeof(float) * (size_t)std::max(1L, B->ld)));
    CR_HIP(B->partial.ensure(sizeof(double) * 32 * (size_t)std::max(1, blk)));
    CR_HIP(B->d_spread.ensure(sizeof(TiledBox) * nbox));
    CR_HIP(B->d_contract.ensure(sizeof(TiledBox) * nbox));
    CR_HIP(B->d_rows.ensure(sizeof(RowBox) * nbox));
    CR_HIP(B->d_segs.ensure(sizeof(BatchSeg) * nbox));
    CR_HIP(B->d_gs.ensure(sizeof(BatchGmres) * nbox));
    CR_HIP(B->d_skip.ensure(sizeof(int) * nbox));
    CR_HIP(B->d_finished.ensure(sizeof(int) * nbox));
    CR_HIP(B->d_step.ensure(sizeof(int)));
    CR_HIP(B->d_pub.ensure(sizeof(double) * 3 * nbox));
    CR_HIP(B->d_rel.ensure(sizeof(double) * nbox));
    CR_HIP(B->d_blocks.ensure(sizeof(BlockBox) * nbox));
    CR_HIP(B->d_fx.ensure(sizeof(FxBox) * nbox));
    if (B->h_pub_n < (size_t)(3 * nbox)) {
        if (B->h_pub) (void)hipHostFree(B->h_pub);
        B->h_pub = nullptr;
        B->h_pub_n = 0;
        CR_HIP(hipHostMalloc((void**)&B->h_pub, sizeof(double) * 3 * nbox, hipHostMallocDefault));
        B->h_pub_n = (size_t)(3 * nbox);
    }
    double* u = B->u.as<double>();
    for (int b = 0; b < nbox; ++b) {
        B->spread_boxes[b].q = u + B->segs[b].xoff;
        B->contract_boxes[b].q = u + B->segs[b].xoff;
        B->contract_boxes[b].pos = B->views[b].pos;
        B->contract_boxes[b].force = force[b];
        if (B->fx) {
            B->fx_boxes[b].x = u + B->segs[b].xoff;
            B->fx_boxes[b].q = u + B->segs[b].xoff;
            B->fx_boxes[b].force = force[b];
        }
    }
    std::vector<double> rel(nbox);
    for (int b = 0; b < nbox; ++b) rel[b] = ctx[b]->res.gmres_tol;
    CR_HIP(hipMemcpyAsync(B->d_spread.p, B->spread_boxes.data(), sizeof(TiledBox) * nbox, hipMemcpyHostToDevice, st));
    CR_HIP(hipMemcpyAsync(B->d_contract.p, B->contract_boxes.data(), sizeof(TiledBox) * nbox, hipMemcpyHostToDevice,
                          st));
    CR_HIP(hipMemcpyAsync(B->d_rows.p, B->rows.data(), sizeof(RowBox) * nbox, hipMemcpyHostToDevice, st));
    CR_HIP(hipMemcpyAsync(B->d_segs.p, B->segs.data(), sizeof(BatchSeg) * nbox, hipMemcpyHostToDevice, st));
    CR_HIP(hipMemcpyAsync(B->d_rel.p, rel.data(), sizeof(double) * nbox, hipMemcpyHostToDevice, st));
    CR_HIP(hipMemcpyAsync(B->d_blocks.p, B->blocks.data(), sizeof(BlockBox) * nbox, hipMemcpyHostToDevice, st));
    if (B->fx) {
        CR_HIP(hipMemcpyAsync(B->d_fx.p, B->fx_boxes.data(), sizeof(FxBox) * nbox, hipMemcpyHostToDevice, st));
        // every box's fixed-point weight bound for this configuration (zeroed in the preparation)
        CR_HIP(launch_spread_fx_bound_batch(B->d_fx.as<FxBox>(), nbox, c0->mp.P, tiles, B->lds_fx_s, st));
        for (int b = 0; b < nbox; ++b) ctx[b]->fx_bound = true;
    }

    // the block preconditioner's per-body inverses of all boxes in one launch
    if (block_precond(c0))
        CR_HIP(launch_block_setup_batch(B->d_blocks.as<BlockBox>(), nbox, bodies, nmax_all, st));
    // x0 of every box into the batch vector (the objective is there already), then the preconditioned objective
    double* x = B->x.as<double>();
    if (x0_mode == CR_X0_ZERO)
        CR_HIP(launch_fill<double>(x, 0.0, B->n_tot, st));
    else
        for (int b = 0; b < nbox; ++b)
            CR_RET(init_x0(ctx[b], B->views[b], x0_mode, q_inout[b], phi_inout[b], x + B->segs[b].xoff, st));
    if (block_precond(c0))
        CR_HIP(launch_block_apply_batch(B->d_blocks.as<BlockBox>(), nbox, bodies, B->b.as<double>(),
                                        B->Mb.as<double>(), nullptr, st));
    else
        CR_HIP(launch_bprecond_diag(B->d_rows.as<RowBox>(), nbox, bblk, rblk, B->b.as<double>(), B->Mb.as<double>(),
                                    st));
    CR_HIP(hipStreamSynchronize(st));  // host descriptors and rel tolerances are consumed

    // stopping scale per box: ||b|| (diagonal, CUSP's default_monitor) or ||D b|| (block)
    const int nb = nbox;
    const BatchSeg* sg = B->d_segs.as<BatchSeg>();
    BatchGmres* gs = B->d_gs.as<BatchGmres>();
    int* skip = B->d_skip.as<int>();
    int* finished = B->d_finished.as<int>();
    int* step = B->d_step.as<int>();
    double* V = B->V.as<double>();
    double* w = B->w.as<double>();
    double* part = B->partial.as<double>();
    std::vector<int> h_fin(nb, 0);
    CR_HIP(hipMemcpyAsync(finished, h_fin.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
    const double* bvec = block_precond(c0) ? B->Mb.as<double>() : B->b.as<double>();
    CR_HIP(launch_bdot(bvec, B->ld, 1, step, bvec, sg, nb, blk, part, nullptr, st));
    CR_HIP(launch_breduce(part, sg, nb, 1, step, gs, kRedNorm, nullptr, st));
    CR_HIP(launch_bset_tol(gs, nb, B->d_rel.as<double>(), st));

    std::vector<int> iters(nb, 0), matvecs(nb, 0), restarts(nb, 0), conv(nb, 0);
    std::vector<double> resid(nb, 0.0), tol(nb, 0.0);
    std::vector<char> host_skip(nb, 0);
    const int maxiter = c0->res.max_iter;
    int iter = 0;
    bool x_zero = x0_mode == CR_X0_ZERO;
    for (;;) {
        // cycle start: w = D (A x) - D b for every unfinished box
        if (x_zero) {
            // x0 = 0 in every box: A x0 = 0 exactly, no operator application needed
            CR_HIP(launch_fill<double>(w, 0.0, B->n_tot, st));
            x_zero = false;
        } else {
            CR_HIP(hipMemcpyAsync(u, x, sizeof(double) * B->n_tot, hipMemcpyDeviceToDevice, st));
            for (int b = 0; b < nb; ++b) host_skip[b] = (char)h_fin[b];
            CR_RET(batch_matvec(B, u, w, 1, finished, &host_skip, st));
            for (int b = 0; b < nb; ++b)
                if (!h_fin[b]) ++matvecs[b];
        }
        CR_HIP(launch_axpby(-1.0, B->Mb.as<double>(), 1.0, w, B->n_tot, st));
        CR_HIP(launch_bdot(w, B->ld, 1, step, w, sg, nb, blk, part, finished, st));
        CR_HIP(launch_breduce(part, sg, nb, 1, step, gs, kRedNorm, finished, st));
        CR_HIP(launch_bcycle_init(gs, nb, finished, skip, st));
        CR_HIP(launch_bstep_end(step, gs, skip, nb, B->d_pub.as<double>(), 0, st));
        CR_HIP(hipMemcpyAsync(B->h_pub, B->d_pub.p, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
        CR_HIP(hipStreamSynchronize(st));
        bool any_active = false;
        for (int b = 0; b < nb; ++b) {
            if (h_fin[b]) continue;
            resid[b] = B->h_pub[3 * b];
            tol[b] = B->h_pub[3 * b + 2];
            if (!std::isfinite(resid[b])) return set_error(CR_ERR_NOT_CONVERGED, "non-finite residual in GMRES");
            if (B->h_pub[3 * b + 1] != 0.0) {
                conv[b] = 1;
                h_fin[b] = 1;
            } else {
                any_active = true;
            }
        }
        if (!any_active || iter >= maxiter) break;
        for (int b = 0; b < nb; ++b)
            if (!h_fin[b]) ++restarts[b];
        CR_HIP(launch_bscale_copy(w, sg, nb, blk, gs, step, 0, -1.0, V, B->ld, skip, st));
        CR_HIP(launch_bset_step(step, 0, st));
        if (use_graph && (!B->step_exec || B->step_sig != step_signature(B))) CR_RET(capture_step(B));
        for (int b = 0; b < nb; ++b) host_skip[b] = (char)h_fin[b];
        int i = -1;
        bool all_conv = false;
        do {
            ++i;
            ++iter;
            if (use_graph)
                CR_HIP(hipGraphLaunch(B->step_exec, st));
            else
                CR_RET(batch_step(B, &host_skip, st));
            CR_HIP(hipStreamSynchronize(st));
            all_conv = true;
            for (int b = 0; b < nb; ++b) {
                if (host_skip[b]) continue;
                ++iters[b];
                ++matvecs[b];
                resid[b] = B->h_pub[3 * b];
                if (!std::isfinite(resid[b])) return set_error(CR_ERR_NOT_CONVERGED, "non-finite residual in GMRES");
                if (B->h_pub[3 * b + 1] != 0.0) {
                    conv[b] = 1;
                    host_skip[b] = 1;
                } else {
                    all_conv = false;
                }
            }
        } while (!all_conv && i + 1 < B->R && iter + 1 <= maxiter);
        CR_HIP(launch_bbacksub(gs, nb, B->R, finished, st));
        CR_HIP(launch_bcombine(V, B->ld, sg, nb, blk, gs, finished, x, st));
        bool any_left = false;
        for (int b = 0; b < nb; ++b) {
            if (conv[b]) h_fin[b] = 1;
            any_left = any_left || !h_fin[b];
        }
        CR_HIP(hipMemcpyAsync(finished, h_fin.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
        CR_HIP(hipStreamSynchronize(st));  // h_fin is read by the copy
        if (!any_left || iter >= maxiter) break;
    }

    // outputs
    for (int b = 0; b < nb; ++b) {
        const SysView& v = B->views[b];
        const long o = B->segs[b].xoff;
        CR_HIP(hipMemcpyAsync(q_inout[b], x + o, sizeof(double) * v.N, hipMemcpyDeviceToDevice, st));
        CR_HIP(hipMemcpyAsync(phi_inout[b], x + o + v.N, sizeof(double) * v.N_isolated, hipMemcpyDeviceToDevice, st));
    }
    CR_HIP(hipMemcpyAsync(u, x, sizeof(double) * B->n_tot, hipMemcpyDeviceToDevice, st));
    bool check = stats != nullptr;
    for (int b = 0; b < nb && check; ++b) check = ctx[b]->check_residual;
    if (stats && !check) {
        for (int b = 0; b < nb; ++b) {
            cr_solve_stats& S = stats[b];
            S.iterations = iters[b];
            S.matvecs = matvecs[b];
            S.restarts = restarts[b];
            S.converged = conv[b];
            S.resid = resid[b];
            S.tolerance = tol[b];
            S.true_resid = -1.0;
            S.true_relres = -1.0;
        }
    }
    if (check) {
        // FP64 check of every box from one batched operator application r = A x - b:
        // ||D r||, ||r|| and ||b|| per box, read back together
        double* dr = B->yraw.as<double>();
        CR_RET(batch_matvec(B, u, w, 0, nullptr, nullptr, st));
        CR_HIP(launch_axpby(-1.0, B->b.as<double>(), 1.0, w, B->n_tot, st));
        if (block_precond(c0))
            CR_HIP(launch_block_apply_batch(B->d_blocks.as<BlockBox>(), nb, B->total_bodies, w, dr, nullptr, st));
        else
            CR_HIP(launch_bprecond_diag(B->d_rows.as<RowBox>(), nb, B->bead_blocks, B->rigid_blocks, w, dr, st));
        const double* vecs[3] = {dr, w, B->b.as<double>()};
        for (int k = 0; k < 3; ++k) {
            CR_HIP(launch_bdot(vecs[k], B->ld, 1, step, vecs[k], sg, nb, blk, part, nullptr, st));
            CR_HIP(launch_breduce(part, sg, nb, 1, step, gs, kRedNorm, nullptr, st));
            CR_HIP(launch_bcollect(gs, nb, B->d_pub.as<double>(), 3, k, st));
        }
        CR_HIP(hipMemcpyAsync(B->h_pub, B->d_pub.p, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st));
        CR_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < nb; ++b) {
            cr_solve_stats& S = stats[b];
            S.iterations = iters[b];
            S.matvecs = matvecs[b];
            S.restarts = restarts[b];
            S.converged = conv[b];
            S.resid = resid[b];
            S.tolerance = tol[b];
            const double rD = B->h_pub[3 * b], r2 = B->h_pub[3 * b + 1], b2 = B->h_pub[3 * b + 2];
            S.true_resid = std::sqrt(rD);
            S.true_relres = b2 > 0 ? std::sqrt(r2 / b2) : std::sqrt(r2);
        }
    }
    // force pass: spread of the converged charges, k-space step and reciprocal-space force
    // contraction for all boxes, then the real-space pair forces + q E0 per box
    CR_RET(batch_spread_recip(B, nullptr, nullptr, st));
    if (B->fx)
        CR_HIP(launch_contract_fx_batch(B->d_fx.as<FxBox>(), nb, c0->mp.P, B->total_tiles, B->lds_fx_c, true, nullptr,
                                        st));
    else if (c0->fp32())
        CR_HIP(launch_contract_tiled_batch<float>(B->d_contract.as<TiledBox>(), nb, B->total_tiles, c0->mp.P,
                                                  B->lds_contract, nullptr, st, true));
    else
        CR_HIP(launch_contract_tiled_batch<double>(B->d_contract.as<TiledBox>(), nb, B->total_tiles, c0->mp.P,
                                                   B->lds_contract, nullptr, st, true));
    // the pair forces + q E0 of all boxes in one launch
    {
        std::vector<PairBox>& pb = B->fpair_boxes;
        pb.clear();
        int pblk = 0;
        for (int b = 0; b < nb; ++b) {
            if (B->views[b].N <= 0) continue;
            PairBox d;
            force_pair_box(ctx[b], B->views[b], x + B->segs[b].xoff, force[b], &d);
            d.blk0 = pblk;
            pblk += (B->views[b].N + 255) / 256;
            pb.push_back(d);
        }
        if (!pb.empty()) {
            CR_HIP(B->d_fpairs.ensure(sizeof(PairBox) * pb.size()));
            CR_HIP(hipMemcpyAsync(B->d_fpairs.p, pb.data(), sizeof(PairBox) * pb.size(), hipMemcpyHostToDevice, st));
            if (c0->fp32())
                CR_HIP(launch_real_space_force_batch<float>(B->d_fpairs.as<PairBox>(), (int)pb.size(), pblk, st));
            else
                CR_HIP(launch_real_space_force_batch<double>(B->d_fpairs.as<PairBox>(), (int)pb.size(), pblk, st));
        }
    }
    CR_HIP(hipEventRecord(B->done_ev, st));
    CR_HIP(hipStreamWaitEvent(caller, B->done_ev, 0));
    return CR_OK;
}

}  // extern "C"
