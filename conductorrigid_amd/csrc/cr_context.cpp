// cr_context.cpp -- host side of the MI355X ConductorRigid force compute: the C ABI
// declared in include/conductorrigid.h.
//
// Plays the role of the reference's host class ConductorRigid
// (cppmodule/ConductorRigid.cc) below the HOOMD interface: SetParams' parameter
// and table construction (:116-379), the GMRES driver MatrixVectorSolve
// (cppmodule/ConductorRigid.cu:1051-1151) and the force pass of gpu_ComputeForce
// (:1153-1286).  Everything runs on the caller's stream; the solver keeps its basis,
// Hessenberg matrix and rotations on the device and reads back one scalar per
// Arnoldi step to decide convergence.
#include "../../include/conductorrigid.h"
#include "cr_context_impl.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace cr;

namespace {

constexpr double kPi = 3.14159265358979323846;

bool is_5smooth(int n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5})
        while (n % f == 0) n /= f;
    return n == 1;
}

std::vector<int> radices(int n)
{
    std::vector<int> r;
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    while (n % 3 == 0) { r.push_back(3); n /= 3; }
    while (n % 5 == 0) { r.push_back(5); n /= 5; }
    return r;
}

// 5-smooth integers in [8, 4096] (cppmodule/ConductorRigid.cc:129-158)
std::vector<int> smooth_list()
{
    std::vector<int> v;
    for (long a = 1; a <= 4096; a *= 2)
        for (long b = 1; a * b <= 4096; b *= 3)
            for (long c = 1; a * b * c <= 4096; c *= 5)
                if (a * b * c >= 8) v.push_back((int)(a * b * c));
    std::sort(v.begin(), v.end());
    return v;
}

double phiq_exact(double dist, double xi)
{
    // potential/charge real-space kernel (cppmodule/ConductorRigid.cc:299-327)
    const double xi2 = xi * xi, dist2 = dist * dist, p15 = std::pow(kPi, 1.5);
    const double expp = std::exp(-(dist + 2) * (dist + 2) * xi2), expm = std::exp(-(dist - 2) * (dist - 2) * xi2),
                 exp0 = std::exp(-dist2 * xi2);
    const double erfp = std::erfc((dist + 2) * xi), erfm = std::erfc((dist - 2) * xi), erf0 = std::erfc(dist * xi);
    const double exppolyp = -(dist + 2.) / (32. * p15 * xi * dist);
    const double exppolym = -(dist - 2.) / (32. * p15 * xi * dist);
    const double exppoly0 = 1. / (16. * p15 * xi);
    const double erfpolyp = 1. / (64. * kPi * xi2 * dist) * (2. * xi2 * (dist + 2.) * (dist + 2.) + 1.);
    const double erfpolym = 1. / (64. * kPi * xi2 * dist) * (2. * xi2 * (dist - 2.) * (dist - 2.) + 1.);
    const double erfpoly0 = -(2. * xi2 * dist2 + 1.) / (32. * kPi * xi2 * dist);
    const double reg = (dist < 2) ? (-1. / (4. * kPi * dist) + (4. - dist) / (16. * kPi)) : 0.;
    return exppolyp * expp + exppolym * expm + exppoly0 * exp0 + erfpolyp * erfp + erfpolym * erfm +
           erfpoly0 * erf0 + reg;
}

double gradphiq_exact(double dist, double xi)
{
    // potential/charge gradient kernel (cppmodule/ConductorRigid.cc:329-345)
    const double xi2 = xi * xi, dist2 = dist * dist, p15 = std::pow(kPi, 1.5);
    const double expp = std::exp(-(dist + 2) * (dist + 2) * xi2), expm = std::exp(-(dist - 2) * (dist - 2) * xi2),
                 exp0 = std::exp(-dist2 * xi2);
    const double erfp = std::erfc((dist + 2) * xi), erfm = std::erfc((dist - 2) * xi), erf0 = std::erfc(dist * xi);
    const double exppolyp = -(dist - 2.) / (32. * p15 * xi * dist2);
    const double exppolym = -(dist + 2.) / (32. * p15 * xi * dist2);
    const double exppoly0 = 1. / (16. * p15 * xi * dist);
    const double erfpoly = (2. * xi2 * dist2 - 8. * xi2 - 1.) / (64. * kPi * xi2 * dist2);
    const double erfpoly0 = -(2. * xi2 * dist2 - 1.) / (32. * kPi * xi2 * dist2);
    const double reg = (dist < 2) ? (1. / (4. * kPi * dist2) - 1. / (16. * kPi)) : 0.;
    return exppolyp * expp + exppolym * expm + exppoly0 * exp0 + erfpoly * erfp + erfpoly * erfm + erfpoly0 * erf0 +
           reg;
}

}  // namespace


namespace cr::impl {

SysView make_view(cr_context* c, const cr_system* s)
{
    SysView v;
    v.N = s->N;
    v.N_rigid = s->N_rigid;
    v.N_isolated = s->N_isolated;
    v.Nmax = s->Nmax > 0 ? s->Nmax : 1;
    v.pos = reinterpret_cast<const double4*>(s->pos);
    v.iso = c->iso.as<unsigned>();
    v.body_charge = s->body_charge;
    v.rigid_membership = s->rigid_membership;
    v.rigid_members = s->rigid_members;
    v.rigid_size = s->rigid_size;
    v.rigid_com = reinterpret_cast<const double4*>(s->rigid_com);
    v.nlist = s->nlist;
    v.head_list = s->head_list;
    v.n_neigh = s->n_neigh;
    return v;
}

// Validate the system view, size per-bead buffers and refresh the body-tag map; with bin,
// also sort the beads into the mesh tiles and write their window records (only the mesh
// operators read those).
int prepare_host(cr_context* c, const cr_system* s, hipStream_t st, SysView* out, bool bin)
{
    if (!c) return set_error(CR_ERR_INVALID, "null context");
    if (!c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    if (!s) return set_error(CR_ERR_INVALID, "null system");
    if (s->N < 0 || s->N_rigid < 0 || s->N_isolated < 0 || s->N_isolated > s->N + s->N_rigid)
        return set_error(CR_ERR_INVALID, "inconsistent system sizes");
    if (s->N > 0 && (!s->pos || !s->tag || !s->body_of_tag || !s->rigid_membership || !s->nlist ||
                     !s->head_list || !s->n_neigh))
        return set_error(CR_ERR_INVALID, "missing per-bead array");
    if (s->N_rigid > 0 && (!s->rigid_members || !s->rigid_size || !s->rigid_com || s->Nmax <= 0))
        return set_error(CR_ERR_INVALID, "missing rigid-body array");
    if (reinterpret_cast<uintptr_t>(s->pos) % 16 != 0) return set_error(CR_ERR_INVALID, "pos must be 16-byte aligned");
    if (s->L[0] != 0.0 || s->L[1] != 0.0 || s->L[2] != 0.0) {
        // the caller's box (HOOMD's BoxDim of this step): take up a change first
        if (s->L[0] != c->params.L[0] || s->L[1] != c->params.L[1] || s->L[2] != c->params.L[2])
            CR_RET(set_box(c, s->L, st));
    }
    const size_t Tsz = c->fp32() ? sizeof(float) : sizeof(double);
    CR_HIP(c->phi_recip.ensure(Tsz * (size_t)std::max(1, s->N)));
    CR_HIP(c->iso.ensure(sizeof(unsigned) * (size_t)std::max(1, s->N)));
    *out = make_view(c, s);
    c->csr_ready = false;  // positions or neighbours may have changed
    c->fx_bound = false;  // the fixed-point weight bound follows the binning
    c->fxp = fx_plan(c->mp, c->tg, c->lds_max);
    c->fxcp = fx_contract_plan(c->mp, c->tg, c->lds_max);
    if (bin && c->tg.valid) {
        // positions may have changed since the last call: rebin (cheap, O(N))
        const size_t nt = (size_t)c->tg.ntiles;
        CR_HIP(c->bead_tile.ensure(sizeof(unsigned) * (size_t)std::max(1, s->N)));
        CR_HIP(c->sorted.ensure(sizeof(unsigned) * (size_t)std::max(1, s->N)));
        CR_HIP(c->tile_count.ensure(sizeof(unsigned) * (nt + 1)));
        CR_HIP(c->tile_start.ensure(sizeof(unsigned) * (nt + 1)));
        CR_HIP(c->tile_cursor.ensure(sizeof(unsigned) * nt));
        CR_HIP(c->scan_sums.ensure(sizeof(unsigned long long) * ((nt + 1) / 1024 + 4)));
        const size_t nb = (size_t)std::max(1, s->N);
        CR_HIP(c->rec_off.ensure(sizeof(int) * nb));
        CR_HIP(c->rec_win.ensure(window_record_bytes(c->mp, c->fp32()) * nb));
    }
    return CR_OK;
}

// The beads binned on the context's tiles with their window records (prepare_host sized
// the buffers).
int bin_records(cr_context* c, const SysView& v, hipStream_t st)
{
    if (!c->tg.valid) return CR_OK;
    CR_HIP(launch_bin_beads(c->mp, c->tg, v, c->bead_tile.as<unsigned>(), c->tile_count.as<unsigned>(),
                            c->tile_start.as<unsigned>(), c->tile_cursor.as<unsigned>(), c->sorted.as<unsigned>(),
                            c->scan_sums.as<unsigned long long>(), st));
    if (c->fp32())
        CR_HIP(launch_bead_records<float>(c->mp, c->tg, v, c->sorted.as<unsigned>(), c->bead_tile.as<unsigned>(),
                                          c->rec_off.as<int>(), c->rec_win.as<float>(), st));
    else
        CR_HIP(launch_bead_records<double>(c->mp, c->tg, v, c->sorted.as<unsigned>(), c->bead_tile.as<unsigned>(),
                                           c->rec_off.as<int>(), c->rec_win.as<double>(), st));
    return CR_OK;
}

// Validate the system and take up its box (prepare_host), then the membership and, with
// bin, the beads binned on the context's tiles with their window records.
int prepare(cr_context* c, const cr_system* s, hipStream_t st, SysView* out, bool bin)
{
    CR_RET(prepare_host(c, s, st, out, bin));
    CR_HIP(launch_membership(s->N, s->tag, s->body_of_tag, c->iso.as<unsigned>(), st));
    if (bin) CR_RET(bin_records(c, *out, st));
    return CR_OK;
}

template <typename T>
const T* table(const cr_context* c, bool grad)
{
    if constexpr (sizeof(T) == sizeof(double))
        return grad ? c->gtab_d.as<T>() : c->tab_d.as<T>();
    else
        return grad ? c->gtab_f.as<T>() : c->tab_f.as<T>();
}

template <typename T>
const AxisPlan* plans(const cr_context* c)
{
    if constexpr (sizeof(T) == sizeof(double))
        return c->plan_d;
    else
        return c->plan_f;
}

// Row pitch of the half spectrum (complex elements): with the register FFT on every axis,
// Nz/2+1 rounded up to 16, so that every row starts on a 128-byte line and the strided
// passes read and write whole lines; the generic passes keep Nz/2+1.
static int half_spectrum_pitch(const MeshParams& mp)
{
    const int nh = mp.Nz / 2 + 1;
    if (!fftr_radix(mp.Nx) || !fftr_radix(mp.Ny) || !fftr_radix(mp.Nz)) return nh;
    return (nh + 15) & ~15;
}

// grid <- Re IFFT(S FFT(grid)) on the context mesh (cppmodule/ConductorRigid.cu:927-933).
template <typename T>
int recip_impl(cr_context* c, hipStream_t st)
{
    const AxisPlan* pl = plans<T>(c);
    const int Nx = c->mp.Nx, Ny = c->mp.Ny;
    T* real = c->grid_real.as<T>();
    T* cp = c->grid_cplx.as<T>();
    CR_HIP(launch_fft_z_r2c<T>(pl[2], Nx, Ny, real, cp, c->Nzp, st));
    CR_HIP(launch_fft_yx<T>(pl[1], 1, Nx, Ny, c->Nzh, c->Nzp, cp, -1, st));
    CR_HIP(launch_fft_x_scale<T>(pl[0], Nx, Ny, c->Nzh, c->Nzp, cp, c->sc, st));
    CR_HIP(launch_fft_yx<T>(pl[1], 1, Nx, Ny, c->Nzh, c->Nzp, cp, +1, st));
    CR_HIP(launch_fft_z_c2r<T>(pl[2], Nx, Ny, cp, c->Nzp, real, st));
    return CR_OK;
}

BinView bin_view(const cr_context* c)
{
    BinView b;
    b.tile_start = c->tile_start.as<unsigned>();
    b.sorted = c->sorted.as<unsigned>();
    b.rec_off = c->rec_off.as<int>();
    b.rec_win = c->rec_win.p;
    return b;
}

bool contract_tiled_ok(const cr_context* c, size_t elem)
{
    return c->tg.valid && tiled_lds(c->mp, c->tg, elem, elem).chunk >= 4;
}

// phi <- contraction of the context mesh (binned LDS version when the extended tile fits)
template <typename T>
int contract_impl(cr_context* c, const SysView& v, T* phi, hipStream_t st)
{
    if constexpr (sizeof(T) == sizeof(float)) {
        if (c->fp32() && c->tg.valid && c->fxcp.valid) {
            CR_HIP(launch_contract_fx(c->mp, c->tg, c->fxcp, bin_view(c), c->grid_real.as<float>(), phi, st));
            return CR_OK;
        }
    }
    if (contract_tiled_ok(c, sizeof(T)))
        CR_HIP(launch_contract_tiled<T>(c->mp, c->tg, v, bin_view(c), nullptr, c->grid_real.as<T>(), phi, nullptr,
                                        st));
    else
        CR_HIP(launch_contract<T>(c->mp, v, c->grid_real.as<T>(), phi, st));
    return CR_OK;
}

// force = the reciprocal-space force of q (+ fadd, 4 doubles per bead, when given)
template <typename T>
int contract_force_impl(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st,
                        const double* fadd = nullptr)
{
    if constexpr (sizeof(T) == sizeof(float)) {
        if (c->tg.valid && c->fxcp.valid) {
            CR_HIP(launch_contract_force_fx(c->mp, c->tg, c->fxcp, bin_view(c), c->grid_real.as<float>(), q, force,
                                            fadd, st));
            return CR_OK;
        }
    }
    if (contract_tiled_ok(c, sizeof(T)))
        CR_HIP(launch_contract_tiled<T>(c->mp, c->tg, v, bin_view(c), q, c->grid_real.as<T>(), nullptr, force, st,
                                        fadd));
    else
        CR_HIP(launch_contract_force<T>(c->mp, v, q, c->grid_real.as<T>(), force, st, fadd));
    return CR_OK;
}

// Assemble the real-space operator for the current positions (valid until the next
// prepare()).  The pair arrays share the neighbour list's row offsets.
int csr_width(const SysView& v, unsigned* dwidth, hipStream_t st)
{
    // ELL width: the longest neighbour row bounds the longest in-cutoff row
    CR_HIP(launch_fill<unsigned>(dwidth, 0u, 1, st));
    if (v.N > 0) CR_HIP(launch_max_u32(v.n_neigh, v.N, dwidth, st));
    return CR_OK;
}

int ensure_csr(cr_context* c, const SysView& v, hipStream_t st)
{
    if (c->csr_ready || v.N <= 0) return CR_OK;
    unsigned* dmax = reinterpret_cast<unsigned*>(c->scal.as<double>() + 32);
    CR_RET(csr_width(v, dmax, st));
    unsigned width = 0;
    CR_HIP(hipMemcpyAsync(&width, dmax, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    CR_HIP(hipStreamSynchronize(st));
    return csr_build(c, v, width, st);
}

// the ELL arrays for rows of at most width entries
int csr_alloc(cr_context* c, const SysView& v, unsigned width)
{
    c->rs_ldn = ((long)v.N + 63) / 64 * 64;
    const size_t cells = std::max<size_t>(1, (size_t)width * (size_t)c->rs_ldn);
    const size_t Tsz = c->fp32() ? sizeof(float) : sizeof(double);
    CR_HIP(c->rs_col.ensure(sizeof(unsigned) * cells));
    CR_HIP(c->rs_val.ensure(Tsz * cells));
    CR_HIP(c->rs_len.ensure(sizeof(unsigned) * (size_t)std::max(1, v.N)));
    return CR_OK;
}

int csr_pair_box(cr_context* c, const SysView& v, unsigned width, PairBox* pb)
{
    CR_RET(csr_alloc(c, v, width));
    *pb = PairBox{};
    pb->rp = c->rp;
    pb->s = v;
    pb->table = c->fp32() ? static_cast<const void*>(table<float>(c, false)) : static_cast<const void*>(table<double>(c, false));
    pb->ldn = c->rs_ldn;
    pb->col = c->rs_col.as<unsigned>();
    pb->val = c->rs_val.p;
    pb->row_len = c->rs_len.as<unsigned>();
    return CR_OK;
}

void force_pair_box(const cr_context* c, const SysView& v, const double* q, double* force, PairBox* pb)
{
    *pb = PairBox{};
    pb->rp = c->rp;
    pb->s = v;
    pb->table = c->fp32() ? static_cast<const void*>(table<float>(c, true)) : static_cast<const void*>(table<double>(c, true));
    pb->q = q;
    pb->force = force;
    for (int d = 0; d < 3; ++d) pb->E0[d] = c->E0[d];
}

int csr_build(cr_context* c, const SysView& v, unsigned width, hipStream_t st)
{
    if (c->csr_ready || v.N <= 0) return CR_OK;
    CR_RET(csr_alloc(c, v, width));
    if (c->fp32())
        CR_HIP(launch_real_csr_build<float>(c->rp, v, table<float>(c, false), c->rs_ldn, c->rs_col.as<unsigned>(),
                                            c->rs_val.as<float>(), c->rs_len.as<unsigned>(), st));
    else
        CR_HIP(launch_real_csr_build<double>(c->rp, v, table<double>(c, false), c->rs_ldn, c->rs_col.as<unsigned>(),
                                             c->rs_val.as<double>(), c->rs_len.as<unsigned>(), st));
    c->csr_ready = true;
    return CR_OK;
}

// FP32 bulk: the charges x[0:N] rounded to float (c->qf) and their partial maxima, from
// which the fixed-point spread forms its scale with the weight bound of the current
// binning (computed first, once per bead configuration: the scale depends on it).
int charges_f32(cr_context* c, const SysView& v, const double* x, hipStream_t st)
{
    if (c->fx_on() && !c->fx_bound) {
        CR_HIP(launch_spread_fx_bound(c->mp, c->tg, c->fxp, bin_view(c), c->fxs.as<FxScale>(), st));
        c->fx_bound = true;
    }
    CR_HIP(c->qf.ensure(sizeof(float) * (size_t)std::max(1, v.N)));
    // the charges and their partial maxima; the fixed-point spread forms the scale from these
    if (x && x == c->qf_for)  // the step tail that produced x left them
        c->fx_parts = c->qf_parts;
    else
        CR_HIP(launch_charges_fx_parts(x, v.N, c->qf.as<float>(), c->fxs.as<FxScale>(), &c->fx_parts, st));
    return CR_OK;
}

// Fixed-point spread of c->qf at the scale charges_f32 set; writes every mesh node (float).
int spread_fx(cr_context* c, hipStream_t st)
{
    if (!c->fx_bound) return set_error(CR_ERR_STATE, "fixed-point spread before its scale");
    CR_HIP(c->fxblk.ensure(sizeof(int) * (size_t)c->tg.ntiles * (size_t)fx_block_ints(c->mp, c->tg)));
    CR_HIP(launch_spread_fx(c->mp, c->tg, c->fxp, bin_view(c), c->qf.as<float>(), c->fxs.as<FxScale>(),
                            c->fxblk.as<int>(), c->grid_real.as<float>(), st, c->fx_parts));
    return CR_OK;
}

// the spread kernel alone on a mesh the caller has cleared (in FP32 bulk with the charge
// rounding, which the fixed-point spread needs first)
template <typename T>
int spread_only(cr_context* c, const SysView& v, const double* q, hipStream_t st)
{
    if constexpr (sizeof(T) == sizeof(float)) {
        CR_RET(charges_f32(c, v, q, st));
        if (c->fx_on()) return spread_fx(c, st);
    }
    if (c->tg.valid)
        CR_HIP(launch_spread_tiled<T>(c->mp, c->tg, bin_view(c), q, c->grid_real.as<T>(), st, false));
    else
        CR_HIP(launch_spread<T>(c->mp, v, q, c->grid_real.as<T>(), st));
    return CR_OK;
}

// q_ready: FP32 bulk only, charges_f32 has already run for q
template <typename T>
int spread_impl(cr_context* c, const SysView& v, const double* q, hipStream_t st, bool q_ready = false)
{
    const bool clear = !c->mesh_zeroed;
    c->mesh_zeroed = false;
    if constexpr (sizeof(T) == sizeof(float)) {
        if (c->fx_on()) {
            if (!q_ready) CR_RET(charges_f32(c, v, q, st));
            return spread_fx(c, st);  // writes every node: no clear
        }
    }
    if (c->tg.valid) {
        CR_HIP(launch_spread_tiled<T>(c->mp, c->tg, bin_view(c), q, c->grid_real.as<T>(), st, clear));
    } else {
        if (clear) CR_HIP(launch_fill<T>(c->grid_real.as<T>(), (T)0, c->G, st));
        CR_HIP(launch_spread<T>(c->mp, v, q, c->grid_real.as<T>(), st));
    }
    return CR_OK;
}

// Zero the mesh now for the next operator application, so that the clear runs on the
// device while the host waits for a residual.  Only gmres() calls this: its next mesh
// operation is always a spread, and it drops the mark on every exit.
int preclear_mesh(cr_context* c, hipStream_t st)
{
    if (c->fx_on()) {  // the fixed-point spread writes every node
        c->mesh_zeroed = true;
        return CR_OK;
    }
    if (c->fp32())
        CR_HIP(launch_fill<float>(c->grid_real.as<float>(), 0.0f, c->G, st));
    else
        CR_HIP(launch_fill<double>(c->grid_real.as<double>(), 0.0, c->G, st));
    c->mesh_zeroed = true;
    return CR_OK;
}

// Real-space rows on the assembled operator.  In FP32 bulk the gathered charges are first
// rounded to float once per application, so the pair sum's gather moves 4-byte operands.
template <typename T>
int real_rows(cr_context* c, const SysView& v, const double* x, double* y, int precond, hipStream_t st,
              bool q_ready = false)
{
    const T* qt = nullptr;
    if constexpr (sizeof(T) == sizeof(double)) {
        qt = x;
    } else {
        if (!q_ready) CR_RET(charges_f32(c, v, x, st));
        qt = c->qf.as<float>();
    }
    CR_HIP(launch_real_space_csr<T>(c->rp, v, x, qt, c->phi_recip.as<T>(), c->rs_ldn, c->rs_col.as<unsigned>(),
                                    c->rs_val.as<T>(), c->rs_len.as<unsigned>(), y, precond, st));
    return CR_OK;
}

// y = A x, or y = D A x with precond (PotentialChargeMultiply + sum_charges/set_charge,
// cppmodule/ConductorRigid.cu:862-1020; D from precondition, :816-859).
// rows = false leaves out the constraint rows y[N:N+B) (the block preconditioner's apply
// forms them from x itself).
template <typename T>
int matvec_impl(cr_context* c, const SysView& v, const double* x, double* y, int precond, hipStream_t st,
                bool rows = true)
{
    constexpr bool f32 = sizeof(T) == sizeof(float);
    if (f32) CR_RET(charges_f32(c, v, x, st));  // float charges for the spread and the pair sum
    CR_RET(spread_impl<T>(c, v, x, st, f32));
    CR_RET(recip_impl<T>(c, st));
    CR_RET(contract_impl<T>(c, v, c->phi_recip.as<T>(), st));
    if (c->csr_ready)
        CR_RET(real_rows<T>(c, v, x, y, precond, st, f32));
    else
        CR_HIP(launch_real_space<T>(c->rp, v, x, c->phi_recip.as<T>(), table<T>(c, false), y, precond, st));
    if (rows) CR_HIP(launch_constraint_rows(v, x, y + v.N, precond, st));
    return CR_OK;
}

int matvec_any(cr_context* c, const SysView& v, const double* x, double* y, int precond, hipStream_t st, bool rows)
{
    return c->fp32() ? matvec_impl<float>(c, v, x, y, precond, st, rows)
                     : matvec_impl<double>(c, v, x, y, precond, st, rows);
}

// ||w||^2 into dst (device)
int norm2(cr_context* c, const double* w, long n, double* dst, hipStream_t st)
{
    CR_HIP(launch_multidot(w, n, 1, w, n, c->partial.as<double>(), st));
    CR_HIP(launch_reduce_partials(c->partial.as<double>(), krylov_blocks(n), 1, dst, st));
    return CR_OK;
}

// A step's residual may be read late only while the geometric trend of the last two
// predicts it above kLagMargin times the tolerance.
constexpr double kLagMargin = 10.0;

// Publish a device scalar into mailbox slot `slot` once the work before it on the stream is
// done; *seq identifies this value for wait_scalar.
int publish_scalar(cr_context* c, const double* dev, int slot, hipStream_t st, unsigned long long* seq)
{
    *seq = ++c->pub_seq;
    CR_HIP(launch_publish(dev, c->hpub + slot, *seq, st));
    return CR_OK;
}

// Poll slot `slot` until the value published with `seq` arrives.  The poll gives up to
// hipStreamSynchronize (and its error) if the stream has drained without it.
int wait_scalar(cr_context* c, int slot, unsigned long long seq, double* host, hipStream_t st)
{
    volatile HostPub* h = c->hpub + slot;
    for (unsigned long spin = 1; h->seq != seq; ++spin) {
        if ((spin & 1023) == 0) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady) return set_error(CR_ERR_DEVICE, hipGetErrorString(q));
            if (q == hipSuccess && h->seq != seq) {
                CR_HIP(hipStreamSynchronize(st));
                if (h->seq != seq) return set_error(CR_ERR_DEVICE, "published scalar did not reach the host");
            }
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    *host = h->value;
    return CR_OK;
}

// A device scalar to the host once the work before it on the stream is done.  The value
// is published into coherent pinned memory by a one-thread kernel and the host polls for
// its sequence number, which costs a few microseconds less per GMRES step than a copy and
// a stream synchronisation.
int read_scalar(cr_context* c, const double* dev, double* host, hipStream_t st)
{
    if (!c->hpub) {
        CR_HIP(hipMemcpyAsync(host, dev, sizeof(double), hipMemcpyDeviceToHost, st));
        CR_HIP(hipStreamSynchronize(st));
        return CR_OK;
    }
    unsigned long long seq = 0;
    CR_RET(publish_scalar(c, dev, 0, st, &seq));
    return wait_scalar(c, 0, seq, host, st);
}

bool block_precond(const cr_context* c) { return c->params.preconditioner == CR_PRECOND_NEARFIELD_BLOCK; }

// Build the per-body inverses of the near-field block preconditioner for the current
// positions (once per force evaluation).
int block_alloc(cr_context* c, const SysView& v)
{
    const int B = v.N_isolated;
    const int nmax = std::max(1, v.Nmax);
    if (nmax > 64) return set_error(CR_ERR_INVALID, "near-field block preconditioner supports bodies of <= 64 beads");
    if (nmax != c->block_nmax) c->block_cache_ok = false;
    c->block_nmax = nmax;
    const size_t b1 = (size_t)std::max(1, B);
    const void* kept[3] = {c->ginv.p, c->bcache_key.p, c->bcache_sig.p};
    CR_HIP(c->body_count.ensure(sizeof(unsigned) * (b1 + 1)));
    CR_HIP(c->body_start.ensure(sizeof(unsigned) * (b1 + 1)));
    CR_HIP(c->body_beads.ensure(sizeof(unsigned) * (size_t)std::max(1, v.N)));
    CR_HIP(c->ginv.ensure(sizeof(float) * b1 * (size_t)block_inv_stride(nmax)));
    CR_HIP(c->uvec.ensure(sizeof(float) * b1 * nmax));
    CR_HIP(c->inv_s.ensure(sizeof(double) * b1));
    CR_HIP(c->block_scan.ensure(sizeof(unsigned long long) * ((b1 + 1) / 1024 + 4)));
    const size_t nbk = (size_t)block_order_buckets(std::max(0, v.N)) + 1;
    CR_HIP(c->body_order.ensure(sizeof(unsigned) * b1));
    CR_HIP(c->slot_count.ensure(sizeof(unsigned) * b1));
    CR_HIP(c->slot_members.ensure(sizeof(unsigned) * b1 * nmax));
    CR_HIP(c->order_key.ensure(sizeof(uint2) * b1));
    CR_HIP(c->order_mkey.ensure(sizeof(unsigned) * b1));
    CR_HIP(c->order_count.ensure(sizeof(unsigned) * nbk));
    CR_HIP(c->order_start.ensure(sizeof(unsigned) * nbk));
    CR_HIP(c->order_scan.ensure(sizeof(unsigned long long) * (nbk / 1024 + 4)));
    CR_HIP(c->bcache_key.ensure(sizeof(unsigned long long) * b1));
    CR_HIP(c->bcache_sig.ensure(sizeof(double) * 2 * b1));
    // new storage holds no inverses (the batched path's setup marks them gone as well)
    if (c->ginv.p != kept[0] || c->bcache_key.p != kept[1] || c->bcache_sig.p != kept[2]) c->block_cache_ok = false;
    return CR_OK;
}

BlockSlots block_slots(cr_context* c)
{
    return BlockSlots{c->body_order.as<unsigned>(), c->slot_count.as<unsigned>(), c->slot_members.as<unsigned>(),
                      c->order_count.as<unsigned>(), c->order_start.as<unsigned>(),
                      c->order_scan.as<unsigned long long>(), c->order_key.as<uint2>(), c->order_mkey.as<unsigned>(),
                      c->bcache_key.as<unsigned long long>(), c->bcache_sig.as<double>(),
                      c->block_cache_ok ? 1 : 0};
}

int block_setup(cr_context* c, const SysView& v, hipStream_t st, bool invert)
{
    CR_RET(block_alloc(c, v));
    const int nmax = c->block_nmax;
    const BlockSlots slots = block_slots(c);
    CR_HIP(launch_block_setup(v, c->mp.L, nmax, c->body_count.as<unsigned>(), c->body_start.as<unsigned>(),
                              c->body_beads.as<unsigned>(),
                              c->block_scan.as<unsigned long long>(), c->ginv.as<float>(), c->uvec.as<float>(),
                              c->inv_s.as<double>(), st, invert, &slots));
    if (invert) c->block_cache_ok = true;  // every slot's entry now describes its stored inverse
    return CR_OK;
}

// out = D in with the configured preconditioner (diagonal: the reference's; block: near field)
int apply_precond(cr_context* c, const SysView& v, const double* in, double* out, hipStream_t st, const double* xq)
{
    if (block_precond(c)) {
        const BlockSlots slots = block_slots(c);
        CR_HIP(launch_block_apply(v, c->block_nmax, slots, c->ginv.as<float>(), c->uvec.as<float>(),
                                  c->inv_s.as<double>(), in, out, st, xq));
    } else {
        CR_HIP(launch_precondition(v, in, out, st));
    }
    return CR_OK;
}

// out = D A x (diagonal D is fused into the matvec's last kernels)
int matvec_prec(cr_context* c, const SysView& v, const double* x, double* out, hipStream_t st)
{
    if (block_precond(c)) {
        // the apply forms the constraint rows from x (one kernel fewer per application)
        CR_RET(matvec_any(c, v, x, c->yraw.as<double>(), 0, st, false));
        return apply_precond(c, v, c->yraw.as<double>(), out, st, x);
    }
    return matvec_any(c, v, x, out, 1, st);
}

// Box change: the mesh, window support, cutoff and real-space tables stay (the tables
// depend on xi only); the spacing h, eta, the window prefactor, the quadrature weight and
// the k-space scale tables follow the box (SetParams, cppmodule/ConductorRigid.cc:199-268),
// the tables regenerated on the device.  A derived mesh that the new box would change
// takes the full SetParams path.
int set_box(cr_context* c, const double L[3], hipStream_t st)
{
    if (!c || !c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0) || !std::isfinite(L[d])) return set_error(CR_ERR_INVALID, "box lengths must be positive");
    cr_params& p = c->params;
    cr_resolved& r = c->res;
    const std::vector<int> sm = smooth_list();
    for (int d = 0; d < 3; ++d) {
        if (p.mesh[d] > 0) continue;
        int m = (int)std::ceil(1.0 + L[d] * r.kcut / kPi);
        for (int v : sm)
            if (m <= v) {
                m = v;
                break;
            }
        if (m != r.mesh[d]) {
            cr_params np = p;
            for (int e = 0; e < 3; ++e) np.L[e] = L[e];
            CR_HIP(hipStreamSynchronize(st));
            return cr_set_params(c, &np, c->n_cap, 0);
        }
    }
    for (int d = 0; d < 3; ++d)
        if (!(r.rc < 0.5 * L[d])) return set_error(CR_ERR_INVALID, "real-space cutoff must stay below half the box");
    const double xi = p.xi;
    for (int d = 0; d < 3; ++d) {
        p.L[d] = L[d];
        r.h[d] = L[d] / r.mesh[d];
        r.eta[d] = c->P_eta * r.h[d] * r.h[d] * xi * xi / kPi;
    }
    r.prefac = r.xiterm * xi / kPi * std::sqrt(2.0 / (kPi * r.eta[0] * r.eta[1] * r.eta[2]));
    r.quadW = r.h[0] * r.h[1] * r.h[2];
    MeshParams& mp = c->mp;
    for (int d = 0; d < 3; ++d) {
        mp.L[d] = L[d];
        mp.h[d] = r.h[d];
        mp.eta[d] = r.eta[d];
        c->rp.L[d] = L[d];
    }
    mp.prefac = r.prefac;
    mp.quadW = r.quadW;
    const int Nd[3] = {mp.Nx, mp.Ny, mp.Nz};
    for (int d = 0; d < 3; ++d)
        CR_HIP(launch_kspace_axis(Nd[d], L[d], xi, r.eta[d], c->ktab[d].as<double>(), c->etab[d].as<double>(), st));
    c->sc.radial = nullptr;
    c->sc.radial_gauss = 0;
    if (L[0] == L[1] && L[1] == L[2]) {
        long m2max = 0;
        for (int d = 0; d < 3; ++d) m2max += (long)(Nd[d] / 2) * (Nd[d] / 2);
        const double kappa = 2.0 * kPi / L[0];
        const bool same_eta = r.eta[0] == r.eta[1] && r.eta[1] == r.eta[2];
        const double gauss = same_eta ? (1.0 - r.eta[0]) * kappa * kappa / (4.0 * xi * xi) : -1.0;
        if (c->fp32()) {
            CR_HIP(c->radtab.ensure(sizeof(float) * (size_t)(m2max + 1)));
            CR_HIP(launch_kspace_radial<float>(m2max, kappa, c->sc.invG, c->radtab.as<float>(), st, gauss));
        } else {
            CR_HIP(c->radtab.ensure(sizeof(double) * (size_t)(m2max + 1)));
            CR_HIP(launch_kspace_radial<double>(m2max, kappa, c->sc.invG, c->radtab.as<double>(), st, gauss));
        }
        c->sc.radial = c->radtab.p;
        c->sc.radial_gauss = same_eta ? 1 : 0;
    }
    c->fx_bound = false;
    ++c->generation;
    return CR_OK;
}

// Per-body bead counts and a double scratch of one entry per body.
int segments(cr_context* c, const SysView& v, hipStream_t st, unsigned** counts, double** sums)
{
    const size_t B = (size_t)std::max(1, v.N_isolated);
    CR_HIP(c->seg_count.ensure(sizeof(unsigned) * B));
    CR_HIP(c->seg_sum.ensure(sizeof(double) * B));
    CR_HIP(launch_body_counts(v, c->seg_count.as<unsigned>(), st));
    *counts = c->seg_count.as<unsigned>();
    *sums = c->seg_sum.as<double>();
    return CR_OK;
}

// Restarted, left-preconditioned GMRES(R) following CUSP's krylov::gmres structure
// (cppmodule/ConductorRigid.cu:1134-1136): the cycle starts from w = D(Ax - b), stops
// when the Givens estimate |s_{i+1}| <= tol ||b||_2, x += V y at the end of each cycle.
// A solve from a caller's start is redone from zero when its FP64 true residual exceeds
// this multiple of the solver tolerance.
constexpr double kFallbackSlack = 10.0;

int gmres(cr_context* c, const SysView& v, cr_solve_stats* stats, hipStream_t st, bool x_zero, const double* ax0)
{
    const long n = (long)v.N + v.N_isolated;
    const int R = c->R, maxiter = c->res.max_iter;
    double* V = c->V.as<double>();
    double* w = c->w.as<double>();
    double* x = c->x.as<double>();
    double* Mb = c->Mb.as<double>();
    double* partial = c->partial.as<double>();
    double* scal = c->scal.as<double>();
    GmresState* g = c->g_dev;
    const int nblk = krylov_blocks(n);

    // stopping scale: ||b|| for the reference's diagonal preconditioner (CUSP's
    // default_monitor), ||D b|| for the block preconditioner
    CR_RET(norm2(c, block_precond(c) ? Mb : c->b.as<double>(), n, scal, st));
    double bnorm2 = 0;
    CR_RET(read_scalar(c, scal, &bnorm2, st));
    const double tolerance = c->res.gmres_tol * std::sqrt(bnorm2);

    int iter = 0, matvecs = 0, restarts = 0, converged = 0;
    double resid = 0;
    // a non-finite residual ends the solve; the counts so far go to stats
    auto nonfinite = [&]() {
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->iterations = iter;
            stats->matvecs = matvecs;
            stats->restarts = restarts;
            stats->resid = std::numeric_limits<double>::quiet_NaN();
            stats->tolerance = tolerance;
            stats->true_resid = -1.0;
            stats->true_relres = -1.0;
        }
        return set_error(CR_ERR_NOT_CONVERGED, "non-finite residual in GMRES");
    };
    struct MeshMarkReset {
        cr_context* c;
        ~MeshMarkReset()
        {
            c->mesh_zeroed = false;
            c->qf_for = nullptr;
        }
    } mark_reset{c};
    // FP32 bulk: each step tail also leaves the next application's float charges
    float* tail_qf = nullptr;
    if (c->fp32() && v.N > 0 && nblk <= kFxParts) {
        CR_HIP(c->qf.ensure(sizeof(float) * (size_t)v.N));
        tail_qf = c->qf.as<float>();
    }
    int tail_parts = 0;
    bool tail_charges = false;
    for (;;) {
        if (x_zero) {
            // x0 = 0: A x0 = 0 exactly, so the first cycle's residual needs no operator
            CR_HIP(launch_fill<double>(w, 0.0, n, st));
            x_zero = false;
        } else if (ax0) {
            // A x0 is known from the initial guess's construction
            CR_RET(apply_precond(c, v, ax0, w, st));
            ax0 = nullptr;
        } else {
            CR_RET(matvec_prec(c, v, x, w, st));
            ++matvecs;
        }
        CR_HIP(launch_axpby(-1.0, Mb, 1.0, w, n, st));
        CR_RET(norm2(c, w, n, scal, st));
        CR_HIP(launch_cycle_init(g, scal, st));
        CR_RET(read_scalar(c, &g->beta, &resid, st));
        if (!std::isfinite(resid)) return nonfinite();
        if (resid <= tolerance) {
            converged = 1;
            break;
        }
        if (iter >= maxiter) break;
        ++restarts;
        CR_HIP(launch_scale_copy(w, &g->inv_beta, -1.0, V, n, st));
        tail_charges = false;  // a cycle starts from V[:, 0], not from a step tail's output
        int i = -1;
        // A step's residual is read one step late while the trend of the last two says it is
        // still far above the tolerance, so that the host does not stall the device between
        // steps.  If it had converged after all, the step already under way is kept (GMRES
        // residuals do not increase).
        bool pending = false;
        int pend_slot = 0;
        unsigned long long pend_seq = 0;
        double r1 = -1.0, r2 = -1.0;  // the last two residuals read this cycle, r1 the newer
        do {
            ++i;
            ++iter;
            double* vi = V + (long)i * c->ld;
            if (tail_charges) {
                c->qf_for = vi;
                c->qf_parts = tail_parts;
            }
            const int mv = matvec_prec(c, v, vi, w, st);
            c->qf_for = nullptr;
            tail_charges = false;
            CR_RET(mv);
            ++matvecs;
            // CGS2: two projections onto V[:,0:i+1]; the second correction is applied only
            // when it is not negligible against ||w|| (h2[i+1] = ||w||^2 after the first)
            const bool wn = i + 1 < 32;
            CR_HIP(launch_multidot(V, c->ld, i + 1, w, n, partial, st));
            CR_HIP(launch_reduce_partials(partial, nblk, i + 1, g->h1, st));
            CR_HIP(launch_gs_update_dot(V, c->ld, i + 1, g->h1, w, n, partial, st, wn));
            CR_HIP(launch_reduce_partials(partial, nblk, i + 1 + (wn ? 1 : 0), g->h2, st));
            // with ||w||^2 among the projections the step tail applies the second correction
            // and takes the norm from them; otherwise a separate pass with norm partials
            if (!wn) CR_HIP(launch_gs_update(V, c->ld, i + 1, g->h2, w, n, 2, partial, st));
            const double* tail_h2 = wn ? g->h2 : nullptr;
            // every continuation (this cycle's next step, or the restart) begins with A x;
            // its mesh clear is queued behind the residual's publication
            const bool go_on = iter + 1 <= maxiter;
            double* vnext = V + (long)(i + 1) * c->ld;
            if (!c->hpub) {
                // ||w||, V[:, i+1], the rotations (no mailbox: the residual is copied back)
                CR_HIP(launch_step_tail(partial, nblk, g, i, R, w, vnext, n, nullptr, 0, st, tail_qf,
                                        c->fxs.as<FxScale>(), v.N, &tail_parts, V, c->ld, i + 1, tail_h2));
                tail_charges = tail_qf != nullptr;
                if (go_on) CR_RET(preclear_mesh(c, st));
                CR_RET(read_scalar(c, &g->resid, &resid, st));
                if (!std::isfinite(resid)) return nonfinite();
                if (resid <= tolerance) {
                    converged = 1;
                    break;
                }
                continue;
            }
            const int slot = iter % kPubSlots;
            // ||w||, V[:, i+1], the rotations and the residual's publication in one launch
            const unsigned long long seq = ++c->pub_seq;
            CR_HIP(launch_step_tail(partial, nblk, g, i, R, w, vnext, n, c->hpub + slot, seq, st, tail_qf,
                                    c->fxs.as<FxScale>(), v.N, &tail_parts, V, c->ld, i + 1, tail_h2));
            tail_charges = tail_qf != nullptr;
            if (go_on) CR_RET(preclear_mesh(c, st));
            if (pending) {
                double rp = 0;
                CR_RET(wait_scalar(c, pend_slot, pend_seq, &rp, st));
                pending = false;
                if (!std::isfinite(rp)) return nonfinite();
                r2 = r1;
                r1 = rp;
                if (rp <= tolerance) {
                    CR_RET(wait_scalar(c, slot, seq, &resid, st));
                    if (!std::isfinite(resid)) return nonfinite();
                    converged = 1;
                    break;
                }
            }
            const bool more = i + 1 < R && iter + 1 <= maxiter;
            if (more && r2 > 0.0 && r1 > 0.0 && r1 * (r1 / r2) > kLagMargin * tolerance) {
                pending = true;
                pend_slot = slot;
                pend_seq = seq;
                continue;
            }
            CR_RET(wait_scalar(c, slot, seq, &resid, st));
            if (!std::isfinite(resid)) return nonfinite();
            r2 = r1;
            r1 = resid;
            if (resid <= tolerance) {
                converged = 1;
                break;
            }
        } while (i + 1 < R && iter + 1 <= maxiter);
        CR_HIP(launch_backsub(g, R, st));
        CR_HIP(launch_combine(V, c->ld, g->y, &g->ncols, x, n, st));
        if (converged || iter >= maxiter) break;
    }
    if (stats) {
        stats->iterations = iter;
        stats->matvecs = matvecs;
        stats->restarts = restarts;
        stats->converged = converged;
        stats->resid = resid;
        stats->tolerance = tolerance;
        stats->true_resid = -1.0;
        stats->true_relres = -1.0;
        stats->fallback = 0;
    }
    return CR_OK;
}

// The bead forces for charges q: the real-space pair forces and q E0 (into force_rs, on a
// side stream) beside the spread and k-space step, then the reciprocal force, which adds
// them per bead.
template <typename T>
int force_pass(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st)
{
    if (v.N <= 0) return CR_OK;
    CR_HIP(c->force_rs.ensure(sizeof(double) * 4 * (size_t)v.N));
    double* frs = c->force_rs.as<double>();
    hipStream_t sd = c->side[0];
    CR_HIP(hipEventRecord(c->ev_fork, st));
    CR_HIP(hipStreamWaitEvent(sd, c->ev_fork, 0));
    CR_HIP(launch_real_space_force<T>(c->rp, v, q, table<T>(c, true), c->E0, frs, sd, false));
    CR_HIP(hipEventRecord(c->ev_join[0], sd));
    CR_RET(spread_impl<T>(c, v, q, st));  // FP32 bulk: rounds the charges itself
    CR_RET(recip_impl<T>(c, st));
    CR_HIP(hipStreamWaitEvent(st, c->ev_join[0], 0));
    CR_RET(contract_force_impl<T>(c, v, q, force, st, frs));
    return CR_OK;
}

template <typename T>
hipError_t upload_twiddles(DevBuf& buf, int n)
{
    std::vector<T> h(2 * (size_t)n);
    for (int e = 0; e < n; ++e) {
        const double a = -2.0 * kPi * (double)e / (double)n;
        h[2 * e] = (T)std::cos(a);
        h[2 * e + 1] = (T)std::sin(a);
    }
    hipError_t err = buf.ensure(sizeof(T) * h.size());
    if (err != hipSuccess) return err;
    return hipMemcpy(buf.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
}

template <typename T>
hipError_t upload(DevBuf& buf, const std::vector<double>& h)
{
    std::vector<T> t(h.begin(), h.end());
    hipError_t err = buf.ensure(sizeof(T) * std::max<size_t>(1, t.size()));
    if (err != hipSuccess) return err;
    return hipMemcpy(buf.p, t.data(), sizeof(T) * t.size(), hipMemcpyHostToDevice);
}

int recip_any(cr_context* c, hipStream_t st) { return c->fp32() ? recip_impl<float>(c, st) : recip_impl<double>(c, st); }

const AxisPlan* axis_plans(const cr_context* c) { return c->fp32() ? plans<float>(c) : plans<double>(c); }

int real_space_force_any(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st)
{
    if (c->fp32())
        CR_HIP(launch_real_space_force<float>(c->rp, v, q, table<float>(c, true), c->E0, force, st));
    else
        CR_HIP(launch_real_space_force<double>(c->rp, v, q, table<double>(c, true), c->E0, force, st));
    return CR_OK;
}

int force_pass_any(cr_context* c, const SysView& v, const double* q, double* force, hipStream_t st)
{
    return c->fp32() ? force_pass<float>(c, v, q, force, st) : force_pass<double>(c, v, q, force, st);
}

int init_x0(cr_context* c, const SysView& v, int x0_mode, const double* q_in, const double* phi_in, double* x,
            hipStream_t st)
{
    const long n = (long)v.N + v.N_isolated;
    if (x0_mode == CR_X0_ZERO) {
        CR_HIP(launch_fill<double>(x, 0.0, n, st));
    } else if (x0_mode == CR_X0_PREVIOUS) {
        CR_HIP(hipMemcpyAsync(x, q_in, sizeof(double) * v.N, hipMemcpyDeviceToDevice, st));
        CR_HIP(hipMemcpyAsync(x + v.N, phi_in, sizeof(double) * v.N_isolated, hipMemcpyDeviceToDevice, st));
    } else if (x0_mode == CR_X0_CHARGES) {
        // q from the caller, shifted per body onto its net charge; Phi fitted to the
        // equipotential rows given that q (one operator application)
        unsigned* counts = nullptr;
        double* sums = nullptr;
        CR_RET(segments(c, v, st, &counts, &sums));
        CR_HIP(hipMemcpyAsync(x, q_in, sizeof(double) * v.N, hipMemcpyDeviceToDevice, st));
        CR_HIP(launch_project_charges(v, counts, sums, x, st));
        CR_HIP(launch_fill<double>(x + v.N, 0.0, v.N_isolated, st));
        // A [q; 0] into yraw; with the fitted Phi added to the bead rows it is A x0, which
        // the solver's first cycle then needs no operator application for
        double* y = c->yraw.as<double>();
        CR_RET(matvec_any(c, v, x, y, 0, st));
        CR_HIP(launch_fit_potentials(v, counts, c->b.as<double>(), y, sums, x + v.N, st));
        CR_HIP(launch_add_body_potentials(v, x + v.N, y, st));
    } else {
        return set_error(CR_ERR_INVALID, "unknown x0 mode");
    }
    return CR_OK;
}

int true_residual(cr_context* c, const SysView& v, const double* x, cr_solve_stats* stats, hipStream_t st)
{
    // FP64 check of the converged state from one operator application r = A x - b:
    // ||D r||_2 and the unpreconditioned ||r|| / ||b||
    const long n = (long)v.N + v.N_isolated;
    double* w = c->w.as<double>();
    double* dw = c->yraw.as<double>();
    CR_RET(matvec_any(c, v, x, w, 0, st));
    CR_HIP(launch_axpby(-1.0, c->b.as<double>(), 1.0, w, n, st));
    CR_RET(apply_precond(c, v, w, dw, st));
    CR_RET(norm2(c, dw, n, c->scal.as<double>(), st));
    CR_RET(norm2(c, w, n, c->scal.as<double>() + 1, st));
    CR_RET(norm2(c, c->b.as<double>(), n, c->scal.as<double>() + 2, st));
    double rb[3] = {0, 0, 0};
    CR_HIP(hipMemcpyAsync(rb, c->scal.as<double>(), 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    CR_HIP(hipStreamSynchronize(st));
    stats->true_resid = std::sqrt(rb[0]);
    stats->true_relres = rb[2] > 0 ? std::sqrt(rb[1] / rb[2]) : std::sqrt(rb[1]);
    return CR_OK;
}

}  // namespace cr::impl

using namespace cr::impl;

extern "C" {

int cr_set_box(cr_context* c, const double L[3], void* stream)
{
    if (!L) return set_error(CR_ERR_INVALID, "null box");
    return set_box(c, L, (hipStream_t)stream);
}

int cr_compute_force_f32(cr_context* c, const cr_system_f32* s, float* q_inout, float* phi_inout, float* force,
                         int x0_mode, cr_solve_stats* stats, void* stream)
{
    if (!c || !c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    if (!s) return set_error(CR_ERR_INVALID, "null system");
    hipStream_t st = (hipStream_t)stream;
    const long N = std::max(0, s->N), B = std::max(0, s->N_isolated), R = std::max(0, s->N_rigid);
    if ((N > 0 && (!s->pos || !q_inout || !force)) || (B > 0 && (!s->body_charge || !phi_inout)) ||
        (R > 0 && !s->rigid_com))
        return set_error(CR_ERR_INVALID, "missing array");
    CR_HIP(hipSetDevice(c->device));
    CR_HIP(c->f32_pos.ensure(sizeof(double) * 4 * (size_t)std::max(1L, N)));
    CR_HIP(c->f32_com.ensure(sizeof(double) * 4 * (size_t)std::max(1L, R)));
    CR_HIP(c->f32_bq.ensure(sizeof(double) * (size_t)std::max(1L, B)));
    CR_HIP(c->f32_q.ensure(sizeof(double) * (size_t)std::max(1L, N)));
    CR_HIP(c->f32_phi.ensure(sizeof(double) * (size_t)std::max(1L, B)));
    CR_HIP(c->f32_force.ensure(sizeof(double) * 4 * (size_t)std::max(1L, N)));
    if (N > 0) {
        CR_HIP((launch_convert<float, double>(s->pos, c->f32_pos.as<double>(), 4 * N, st)));
        CR_HIP((launch_convert<float, double>(q_inout, c->f32_q.as<double>(), N, st)));
    }
    if (R > 0) CR_HIP((launch_convert<float, double>(s->rigid_com, c->f32_com.as<double>(), 4 * R, st)));
    if (B > 0) {
        CR_HIP((launch_convert<float, double>(s->body_charge, c->f32_bq.as<double>(), B, st)));
        CR_HIP((launch_convert<float, double>(phi_inout, c->f32_phi.as<double>(), B, st)));
    }
    cr_system d{};
    d.N = s->N;
    d.pos = c->f32_pos.as<double>();
    d.tag = s->tag;
    d.body_of_tag = s->body_of_tag;
    d.body_charge = c->f32_bq.as<double>();
    d.N_rigid = s->N_rigid;
    d.N_isolated = s->N_isolated;
    d.Nmax = s->Nmax;
    d.rigid_membership = s->rigid_membership;
    d.rigid_members = s->rigid_members;
    d.rigid_size = s->rigid_size;
    d.rigid_com = c->f32_com.as<double>();
    d.nlist = s->nlist;
    d.head_list = s->head_list;
    d.n_neigh = s->n_neigh;
    for (int k = 0; k < 3; ++k) d.L[k] = s->L[k];
    CR_RET(cr_compute_force(c, &d, c->f32_q.as<double>(), c->f32_phi.as<double>(), c->f32_force.as<double>(), x0_mode,
                            stats, stream));
    if (N > 0) {
        CR_HIP((launch_convert<double, float>(c->f32_q.as<double>(), q_inout, N, st)));
        CR_HIP((launch_convert<double, float>(c->f32_force.as<double>(), force, 4 * N, st)));
    }
    if (B > 0) CR_HIP((launch_convert<double, float>(c->f32_phi.as<double>(), phi_inout, B, st)));
    return CR_OK;
}

const char* cr_last_error(void) { return g_error.c_str(); }

int cr_create(cr_context** out, int device)
{
    if (!out) return set_error(CR_ERR_INVALID, "null output pointer");
    *out = nullptr;
    CR_HIP(hipSetDevice(device));
    cr_context* c = new cr_context();
    c->device = device;
    // the allocation-free neighbour-list build's static scratch on this device (its
    // address is looked up once per device; a failure here surfaces at that entry point)
    if (nlist_scratch_setup(device) != hipSuccess) (void)hipGetLastError();
    if (hipMalloc((void**)&c->g_dev, sizeof(GmresState)) != hipSuccess ||
        hipHostMalloc((void**)&c->g_host, sizeof(GmresState), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return set_error(CR_ERR_DEVICE, "allocation of solver state failed");
    }
    if (c->fxs.ensure(sizeof(FxScale)) != hipSuccess || hipMemset(c->fxs.p, 0, sizeof(FxScale)) != hipSuccess) {
        delete c;
        return set_error(CR_ERR_DEVICE, "allocation of the spread scale failed");
    }
    bool ok = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_width, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < cr_context::kSide; ++k)
        ok = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete c;
        return set_error(CR_ERR_DEVICE, "creation of the side streams failed");
    }
    // the setup's ELL width read-back; without pinned memory it is a synchronous copy
    if (hipHostMalloc((void**)&c->hwidth, sizeof(unsigned), hipHostMallocDefault) != hipSuccess) {
        c->hwidth = nullptr;
        (void)hipGetLastError();
    }
    // residual mailbox; without coherent host memory the solver reads back by copy
    if (hipHostMalloc((void**)&c->hpub, kPubSlots * sizeof(HostPub), hipHostMallocCoherent) != hipSuccess) {
        c->hpub = nullptr;
        (void)hipGetLastError();
    } else {
        for (int k = 0; k < kPubSlots; ++k) {
            c->hpub[k].value = 0.0;
            c->hpub[k].seq = 0;
        }
    }
    *out = c;
    return CR_OK;
}

void cr_destroy(cr_context* ctx) { delete ctx; }

int cr_set_params(cr_context* c, const cr_params* p, int n_beads, int n_isolated)
{
    if (!c || !p) return set_error(CR_ERR_INVALID, "null argument");
    if (!(p->xi > 0) || !(p->errortol > 0 && p->errortol < 1)) return set_error(CR_ERR_INVALID, "xi > 0 and 0 < errortol < 1 required");
    for (int d = 0; d < 3; ++d)
        if (!(p->L[d] > 0)) return set_error(CR_ERR_INVALID, "box lengths must be positive");
    if (p->precision != CR_PRECISION_FP64 && p->precision != CR_PRECISION_FP32_BULK)
        return set_error(CR_ERR_INVALID, "unknown precision");
    if (p->preconditioner != CR_PRECOND_DIAGONAL && p->preconditioner != CR_PRECOND_NEARFIELD_BLOCK)
        return set_error(CR_ERR_INVALID, "unknown preconditioner");
    if (n_beads < 0 || n_isolated < 0) return set_error(CR_ERR_INVALID, "negative sizes");
    CR_HIP(hipSetDevice(c->device));
    c->ready = false;
    c->params = *p;
    cr_resolved& r = c->res;
    std::memset(&r, 0, sizeof(r));
    const double xi = p->xi, tol = p->errortol;
    // ConductorRigid.cc:123-207
    r.rc = (p->rc > 0) ? p->rc : std::sqrt(-std::log(tol)) / xi;
    r.kcut = 2.0 * xi * std::sqrt(-std::log(tol));
    const std::vector<int> sm = smooth_list();
    for (int d = 0; d < 3; ++d) {
        if (p->mesh[d] > 0) {
            r.mesh[d] = p->mesh[d];
        } else {
            int m = (int)std::ceil(1.0 + p->L[d] * r.kcut / kPi);
            for (int s : sm)
                if (m <= s) {
                    m = s;
                    break;
                }
            r.mesh[d] = m;
        }
        if (r.mesh[d] > 4096) return set_error(CR_ERR_INVALID, "mesh exceeds 4096 nodes per dimension");
        if (!is_5smooth(r.mesh[d])) return set_error(CR_ERR_INVALID, "mesh sizes must be 5-smooth (2^a 3^b 5^c)");
        r.h[d] = p->L[d] / r.mesh[d];
    }
    int P = (p->P > 0) ? p->P : (int)std::ceil(-2.0 * std::log(tol) / kPi);
    for (int d = 0; d < 3; ++d) r.eta[d] = P * r.h[d] * r.h[d] * xi * xi / kPi;
    c->P_eta = P;
    P = std::min(P, std::min(r.mesh[0], std::min(r.mesh[1], r.mesh[2])));
    r.P = P;
    r.drtable = 0.001;
    r.N_table = (int)(r.rc / r.drtable - 1);
    if (r.N_table < 2) return set_error(CR_ERR_INVALID, "real-space cutoff too small");
    r.xiterm = 2.0 * xi * xi;
    r.prefac = r.xiterm * xi / kPi * std::sqrt(2.0 / (kPi * r.eta[0] * r.eta[1] * r.eta[2]));
    r.selfterm = (1.0 - std::exp(-4.0 * xi * xi)) / (8.0 * kPi * std::sqrt(kPi) * xi) + std::erfc(2.0 * xi) / (4.0 * kPi);
    r.quadW = r.h[0] * r.h[1] * r.h[2];
    r.precision = p->precision;
    r.restart = (p->restart > 0) ? p->restart : 10;
    r.max_iter = (p->max_iter > 0) ? p->max_iter : 100;
    r.gmres_tol = (p->gmres_tol > 0) ? p->gmres_tol : tol;
    if (r.restart > 32) return set_error(CR_ERR_INVALID, "restart must be <= 32");
    // one LDS-resident pencil plus the twiddle table must fit 64 KiB
    const int max_n = (p->precision == CR_PRECISION_FP64) ? 1024 : 2048;
    for (int d = 0; d < 3; ++d)
        if (r.mesh[d] > max_n) return set_error(CR_ERR_INVALID, "mesh too large for the LDS-resident FFT in this precision");

    MeshParams& mp = c->mp;
    mp.Nx = r.mesh[0];
    mp.Ny = r.mesh[1];
    mp.Nz = r.mesh[2];
    mp.P = r.P;
    for (int d = 0; d < 3; ++d) {
        mp.L[d] = p->L[d];
        mp.h[d] = r.h[d];
        mp.eta[d] = r.eta[d];
    }
    mp.prefac = r.prefac;
    mp.xiterm = r.xiterm;
    mp.quadW = r.quadW;
    RealParams& rp = c->rp;
    for (int d = 0; d < 3; ++d) rp.L[d] = p->L[d];
    rp.rc = r.rc;
    rp.rc2 = r.rc * r.rc;
    rp.drtable = r.drtable;
    rp.rmin2 = r.drtable * r.drtable;
    rp.selfterm = r.selfterm;
    rp.N_table = r.N_table;

    // real-space tables (ConductorRigid.cc:276-354)
    const int nt = r.N_table;
    std::vector<double> tab(2 * (size_t)(nt + 1), 0.0), gtab(2 * (size_t)(nt + 1), 0.0);
    for (int i = 0; i <= nt; ++i) {
        const double dist = (i + 1) * r.drtable;
        tab[2 * i] = phiq_exact(dist, xi);
        gtab[2 * i] = gradphiq_exact(dist, xi);
    }
    for (int i = 0; i < nt; ++i) {
        tab[2 * i + 1] = tab[2 * (i + 1)];
        gtab[2 * i + 1] = gtab[2 * (i + 1)];
    }
    CR_HIP(upload<double>(c->tab_d, tab));
    CR_HIP(upload<double>(c->gtab_d, gtab));
    CR_HIP(upload<float>(c->tab_f, tab));
    CR_HIP(upload<float>(c->gtab_f, gtab));

    // FFT plans and k-space factor tables (ConductorRigid.cc:245-274)
    c->Nzh = mp.Nz / 2 + 1;
    c->Nzp = half_spectrum_pitch(mp);
    c->G = (long)mp.Nx * mp.Ny * mp.Nz;
    c->Gh = (long)mp.Nx * mp.Ny * c->Nzp;
    const int Nd[3] = {mp.Nx, mp.Ny, mp.Nz};
    for (int d = 0; d < 3; ++d) {
        std::vector<int> rad = radices(Nd[d]);
        if (rad.size() > 24) return set_error(CR_ERR_INVALID, "too many FFT stages");
        CR_HIP(upload_twiddles<double>(c->tw_d[d], Nd[d]));
        CR_HIP(upload_twiddles<float>(c->tw_f[d], Nd[d]));
        for (AxisPlan* ap : {&c->plan_d[d], &c->plan_f[d]}) {
            ap->n = Nd[d];
            ap->nstages = (int)rad.size();
            for (size_t s = 0; s < rad.size(); ++s) ap->radix[s] = rad[s];
        }
        c->plan_d[d].tw = c->tw_d[d].p;
        c->plan_f[d].tw = c->tw_f[d].p;
        std::vector<double> kv(Nd[d]), ev(Nd[d]);
        for (int i = 0; i < Nd[d]; ++i) {
            const double k = ((i < (Nd[d] + 1) / 2) ? i : i - Nd[d]) * 2.0 * kPi / p->L[d];
            kv[i] = k;
            ev[i] = std::exp(-(1.0 - r.eta[d]) * k * k / (4.0 * xi * xi));
        }
        CR_HIP(upload<double>(c->ktab[d], kv));
        CR_HIP(upload<double>(c->etab[d], ev));
        c->sc.k[d] = c->ktab[d].as<double>();
        c->sc.e[d] = c->etab[d].as<double>();
    }
    c->sc.invG = 1.0 / (double)c->G;
    for (int d = 0; d < 3; ++d) c->sc.n[d] = Nd[d];
    c->sc.radial = nullptr;
    c->sc.radial_gauss = 0;
    if (p->L[0] == p->L[1] && p->L[1] == p->L[2]) {
        // S = j0(|k|)^2 / k^2 * e_x e_y e_z / G with k^2 = kappa^2 m2 (ConductorRigid.cc:258-268);
        // with eta the same on every axis e_x e_y e_z = exp(-(1-eta) k^2 / (4 xi^2)) joins the table
        long m2max = 0;
        for (int d = 0; d < 3; ++d) m2max += (long)(Nd[d] / 2) * (Nd[d] / 2);
        const double kappa = 2.0 * kPi / p->L[0];
        const bool same_eta = r.eta[0] == r.eta[1] && r.eta[1] == r.eta[2];
        const double gauss = (1.0 - r.eta[0]) * kappa * kappa / (4.0 * xi * xi);
        std::vector<double> rad((size_t)m2max + 1);
        rad[0] = 0.0;  // S(0) = 0
        for (long m2 = 1; m2 <= m2max; ++m2) {
            const double k2 = kappa * kappa * (double)m2, k = std::sqrt(k2);
            const double j0 = std::sin(k) / k;
            const double g = same_eta ? std::exp(-gauss * (double)m2) : 1.0;
            rad[(size_t)m2] = j0 * j0 / k2 * c->sc.invG * g;
        }
        if (c->fp32())
            CR_HIP(upload<float>(c->radtab, rad));
        else
            CR_HIP(upload<double>(c->radtab, rad));
        c->sc.radial = c->radtab.p;
        c->sc.radial_gauss = same_eta ? 1 : 0;
    }
    // binned spread/contract tiling: per axis the largest divisor of the mesh size in
    // [max(P-1, 4), 16], else the smallest one in (16, 32]; the tile extended by P-1
    // nodes per axis, held in the spread's double accumulator, must fit 3/4 of the
    // 64 KiB LDS budget, shrinking x, then y, then z if needed
    const size_t Tsz = c->fp32() ? sizeof(float) : sizeof(double);
    {
        TileGeom& tg = c->tg;
        tg.valid = 1;
        const int lo = std::max(r.P - 1, 4);
        std::vector<int> opts[3];
        for (int d = 0; d < 3; ++d) {
            const int n = Nd[d];
            for (int t = std::min(n, 32); t >= lo; --t)
                if (n % t == 0) opts[d].push_back(t);      // descending
            int pick = -1;
            for (size_t i = 0; i < opts[d].size(); ++i)
                if (opts[d][i] <= 16) {
                    pick = (int)i;
                    break;
                }
            if (pick < 0 && !opts[d].empty()) pick = (int)opts[d].size() - 1;
            if (pick < 0) {
                tg.valid = 0;
                tg.T[d] = n;
            } else {
                tg.T[d] = opts[d][pick];
                opts[d].erase(opts[d].begin(), opts[d].begin() + pick);  // remaining: smaller choices
            }
        }
        auto ext_bytes = [&]() {
            return (size_t)(tg.T[0] + r.P - 1) * (tg.T[1] + r.P - 1) * (tg.T[2] + r.P - 1) * sizeof(SpreadAcc);
        };
        // LDS budget per binned workgroup: the default, or CR_LDS_KB (tuning override, up to
        // the device's per-workgroup maximum); about a quarter is left for the per-bead
        // staging tables
        size_t budget = kTiledLdsBudget;
        if (const char* env = std::getenv("CR_LDS_KB")) {
            int dev_max = 0;
            (void)hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
            const long kb = std::atol(env);
            if (kb >= 16 && (size_t)kb * 1024 <= (size_t)dev_max) budget = (size_t)kb * 1024;
        }
        tg.lds = (int)budget;
        const size_t ext_cap = budget * 3 / 4;
        // shrink x first: z stays the longest edge, so the tile's mesh rows are long
        // contiguous runs for the write-out and the extended-tile load
        for (int d = 0; tg.valid && ext_bytes() > ext_cap && d <= 2;) {
            if (opts[d].size() > 1) {
                opts[d].erase(opts[d].begin());
                tg.T[d] = opts[d][0];
            } else {
                ++d;
            }
        }
        if (ext_bytes() > ext_cap) tg.valid = 0;
        // tuning override: CR_TILE="tx,ty,tz" (used when each edge divides the mesh, is
        // >= P-1 and the extended tile fits the LDS budget)
        if (const char* env = std::getenv("CR_TILE")) {
            int t3[3];
            if (std::sscanf(env, "%d,%d,%d", &t3[0], &t3[1], &t3[2]) == 3) {
                bool ok = true;
                for (int d = 0; d < 3; ++d) ok = ok && t3[d] >= std::max(r.P - 1, 1) && Nd[d] % t3[d] == 0;
                const size_t eb =
                    (size_t)(t3[0] + r.P - 1) * (t3[1] + r.P - 1) * (t3[2] + r.P - 1) * sizeof(SpreadAcc);
                if (ok && eb <= ext_cap) {
                    for (int d = 0; d < 3; ++d) tg.T[d] = t3[d];
                    tg.valid = 1;
                }
            }
        }
        tg.ntiles = 1;
        for (int d = 0; d < 3; ++d) {
            tg.nt[d] = Nd[d] / tg.T[d];
            tg.ntiles *= tg.nt[d];
        }
    }
    {
        int dev_max = 0;
        CR_HIP(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
        // a little is left for the kernels' static LDS
        c->lds_max = dev_max > 1024 ? (size_t)dev_max - 256 : 65536;
    }
    CR_HIP(c->grid_real.ensure(Tsz * (size_t)c->G));
    CR_HIP(c->grid_cplx.ensure(2 * Tsz * (size_t)c->Gh));

    // Krylov workspace: basis of R+1 columns of n = N + N_isolated doubles
    c->R = r.restart;
    const long n = (long)n_beads + n_isolated;
    c->n_cap = (int)n;
    c->ld = ((std::max<long>(n, 1) + 63) / 64) * 64;
    CR_HIP(c->V.ensure(sizeof(double) * (size_t)c->ld * (c->R + 1)));
    CR_HIP(c->w.ensure(sizeof(double) * (size_t)c->ld));
    CR_HIP(c->b.ensure(sizeof(double) * (size_t)c->ld));
    CR_HIP(c->Mb.ensure(sizeof(double) * (size_t)c->ld));
    CR_HIP(c->x.ensure(sizeof(double) * (size_t)c->ld));
    CR_HIP(c->yraw.ensure(sizeof(double) * (size_t)c->ld));
    CR_HIP(c->partial.ensure(sizeof(double) * 32 * 1024));
    CR_HIP(c->scal.ensure(sizeof(double) * 64));
    c->ready = true;
    ++c->generation;
    return CR_OK;
}

int cr_get_resolved(const cr_context* c, cr_resolved* out)
{
    if (!c || !out) return set_error(CR_ERR_INVALID, "null argument");
    if (!c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    *out = c->res;
    return CR_OK;
}

int cr_update_field(cr_context* c, const double E0[3])
{
    if (!c || !E0) return set_error(CR_ERR_INVALID, "null argument");
    for (int d = 0; d < 3; ++d) c->E0[d] = E0[d];
    return CR_OK;
}

int cr_zero_force(cr_context* c, int N, double* force, void* stream)
{
    if (!c || (N > 0 && !force)) return set_error(CR_ERR_INVALID, "null argument");
    CR_HIP(launch_zero_force(N, force, (hipStream_t)stream));
    return CR_OK;
}

int cr_objective(cr_context* c, const cr_system* s, double* b, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    CR_HIP(launch_objective(v, c->mp.L, c->E0, b, st));
    return CR_OK;
}

int cr_precondition(cr_context* c, const cr_system* s, const double* x, double* y, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (block_precond(c)) CR_RET(block_setup(c, v, st));
    return apply_precond(c, v, x, y, st);
}

int cr_matvec(cr_context* c, const cr_system* s, const double* x, double* y, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    CR_RET(ensure_csr(c, v, st));
    return matvec_any(c, v, x, y, 0, st);
}

int cr_compute_force(cr_context* c, const cr_system* s, double* q_inout, double* phi_inout, double* force,
                     int x0_mode, cr_solve_stats* stats, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare_host(c, s, st, &v, true));
    const long n = (long)v.N + v.N_isolated;
    if (n > c->n_cap) return set_error(CR_ERR_INVALID, "system larger than the size given to cr_set_params");
    if ((v.N > 0 && (!q_inout || !force)) || (v.N_isolated > 0 && !phi_inout))
        return set_error(CR_ERR_INVALID, "null output array");
    if (x0_mode != CR_X0_ZERO && x0_mode != CR_X0_PREVIOUS && x0_mode != CR_X0_CHARGES)
        return set_error(CR_ERR_INVALID, "unknown x0 mode");
    double* x = c->x.as<double>();
    CR_HIP(launch_membership(s->N, s->tag, s->body_of_tag, c->iso.as<unsigned>(), st));
    // Per-configuration setup in three independent parts, side by side: the real-space ELL
    // operator (side 0), the block preconditioner's inverses and the objective (side 1) and,
    // on the caller's stream, the binning and the fixed-point spread's bound.  They are queued
    // in the order of their length on the device (the host's launches take a few
    // microseconds each): the ELL width and the binning first, the ELL build as soon as its
    // width is back on the host, then the preconditioner.
    CR_HIP(hipEventRecord(c->ev_fork, st));
    CR_HIP(hipStreamWaitEvent(c->side[0], c->ev_fork, 0));
    const bool csr = !c->csr_ready && v.N > 0;
    unsigned* dwidth = reinterpret_cast<unsigned*>(c->scal.as<double>() + 32);
    if (csr) {
        CR_RET(csr_width(v, dwidth, c->side[0]));
        if (c->hwidth) {
            CR_HIP(hipMemcpyAsync(c->hwidth, dwidth, sizeof(unsigned), hipMemcpyDeviceToHost, c->side[0]));
            CR_HIP(hipEventRecord(c->ev_width, c->side[0]));
        }
    }
    CR_RET(bin_records(c, v, st));
    if (c->fx_on() && !c->fx_bound) {
        CR_HIP(launch_spread_fx_bound(c->mp, c->tg, c->fxp, bin_view(c), c->fxs.as<FxScale>(), st));
        c->fx_bound = true;
    }
    if (csr) {
        unsigned width = 0;
        if (c->hwidth) {
            CR_HIP(hipEventSynchronize(c->ev_width));
            width = *c->hwidth;
        } else {
            CR_HIP(hipMemcpyAsync(&width, dwidth, sizeof(unsigned), hipMemcpyDeviceToHost, c->side[0]));
            CR_HIP(hipStreamSynchronize(c->side[0]));
        }
        CR_RET(csr_build(c, v, width, c->side[0]));
    }
    const bool blk = block_precond(c);
    hipStream_t sobj = st;
    if (blk) {
        CR_HIP(hipStreamWaitEvent(c->side[1], c->ev_fork, 0));
        CR_RET(block_setup(c, v, c->side[1]));
        sobj = c->side[1];
    }
    // objective (MatrixVectorSolve, cppmodule/ConductorRigid.cu:1099-1107)
    CR_HIP(launch_objective(v, c->mp.L, c->E0, c->b.as<double>(), sobj));
    if (blk) CR_HIP(hipEventRecord(c->ev_join[1], c->side[1]));
    CR_HIP(hipEventRecord(c->ev_join[0], c->side[0]));
    CR_HIP(hipStreamWaitEvent(st, c->ev_join[0], 0));
    if (blk) CR_HIP(hipStreamWaitEvent(st, c->ev_join[1], 0));
    // the objective's preconditioned copy
    CR_RET(apply_precond(c, v, c->b.as<double>(), c->Mb.as<double>(), st));
    CR_RET(init_x0(c, v, x0_mode, q_inout, phi_inout, x, st));
    cr_solve_stats local{};
    cr_solve_stats* sp = stats ? stats : &local;
    const int rc = gmres(c, v, sp, st, x0_mode == CR_X0_ZERO,
                         x0_mode == CR_X0_CHARGES ? c->yraw.as<double>() : nullptr);
    bool redo = false, checked = false;
    if (x0_mode == CR_X0_ZERO) {
        CR_RET(rc);
    } else if (rc == CR_ERR_NOT_CONVERGED) {
        redo = true;  // a non-finite residual from the given start
    } else {
        CR_RET(rc);
        if (!sp->converged) {
            redo = true;  // the iteration limit from the given start
        } else if (stats && c->check_residual) {
            CR_RET(true_residual(c, v, x, sp, st));
            checked = true;
            redo = !(sp->true_resid <= kFallbackSlack * sp->tolerance);
        }
    }
    if (redo) {
        // plain GMRES from x0 = 0 decides (the fallback and judge of the converged residual)
        const cr_solve_stats first = *sp;
        CR_RET(init_x0(c, v, CR_X0_ZERO, q_inout, phi_inout, x, st));
        CR_RET(gmres(c, v, sp, st, true, nullptr));
        sp->iterations += first.iterations;
        sp->matvecs += first.matvecs;
        sp->restarts += first.restarts;
        sp->fallback = 1;
        checked = false;
    }
    CR_HIP(hipMemcpyAsync(q_inout, x, sizeof(double) * v.N, hipMemcpyDeviceToDevice, st));
    CR_HIP(hipMemcpyAsync(phi_inout, x + v.N, sizeof(double) * v.N_isolated, hipMemcpyDeviceToDevice, st));
    if (stats && c->check_residual && !checked) CR_RET(true_residual(c, v, x, stats, st));
    if (c->fp32())
        CR_RET(force_pass<float>(c, v, x, force, st));
    else
        CR_RET(force_pass<double>(c, v, x, force, st));
    return CR_OK;
}

int cr_spread(cr_context* c, const cr_system* s, const double* q, double* grid_out, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (c->fp32()) {
        CR_RET(spread_impl<float>(c, v, q, st));
        if (grid_out) CR_HIP((launch_convert<float, double>(c->grid_real.as<float>(), grid_out, c->G, st)));
    } else {
        CR_RET(spread_impl<double>(c, v, q, st));
        if (grid_out) CR_HIP((launch_convert<double, double>(c->grid_real.as<double>(), grid_out, c->G, st)));
    }
    return CR_OK;
}

int cr_recip_grid(cr_context* c, const double* grid_in, double* grid_out, void* stream)
{
    if (!c || !c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    hipStream_t st = (hipStream_t)stream;
    if (c->fp32()) {
        if (grid_in) CR_HIP((launch_convert<double, float>(grid_in, c->grid_real.as<float>(), c->G, st)));
        CR_RET(recip_impl<float>(c, st));
        if (grid_out) CR_HIP((launch_convert<float, double>(c->grid_real.as<float>(), grid_out, c->G, st)));
    } else {
        if (grid_in) CR_HIP((launch_convert<double, double>(grid_in, c->grid_real.as<double>(), c->G, st)));
        CR_RET(recip_impl<double>(c, st));
        if (grid_out) CR_HIP((launch_convert<double, double>(c->grid_real.as<double>(), grid_out, c->G, st)));
    }
    return CR_OK;
}

int cr_fft_c2c(cr_context* c, const double* in, double* out, int sign, void* stream)
{
    if (!c || !c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    if (!in || !out || (sign != 1 && sign != -1)) return set_error(CR_ERR_INVALID, "bad argument");
    hipStream_t st = (hipStream_t)stream;
    const long n2 = 2 * c->G;
    const int Nx = c->mp.Nx, Ny = c->mp.Ny, Nz = c->mp.Nz;
    DevBuf tmp;
    if (c->fp32()) {
        CR_HIP(tmp.ensure(sizeof(float) * (size_t)n2));
        float* t = tmp.as<float>();
        CR_HIP((launch_convert<double, float>(in, t, n2, st)));
        CR_HIP(launch_fft_z_c2c<float>(c->plan_f[2], Nx, Ny, t, sign, st));
        CR_HIP(launch_fft_yx<float>(c->plan_f[1], 1, Nx, Ny, Nz, Nz, t, sign, st));
        CR_HIP(launch_fft_yx<float>(c->plan_f[0], 0, Nx, Ny, Nz, Nz, t, sign, st));
        CR_HIP((launch_convert<float, double>(t, out, n2, st)));
    } else {
        CR_HIP(tmp.ensure(sizeof(double) * (size_t)n2));
        double* t = tmp.as<double>();
        CR_HIP((launch_convert<double, double>(in, t, n2, st)));
        CR_HIP(launch_fft_z_c2c<double>(c->plan_d[2], Nx, Ny, t, sign, st));
        CR_HIP(launch_fft_yx<double>(c->plan_d[1], 1, Nx, Ny, Nz, Nz, t, sign, st));
        CR_HIP(launch_fft_yx<double>(c->plan_d[0], 0, Nx, Ny, Nz, Nz, t, sign, st));
        CR_HIP((launch_convert<double, double>(t, out, n2, st)));
    }
    CR_HIP(hipStreamSynchronize(st));
    return CR_OK;
}

int cr_contract(cr_context* c, const cr_system* s, double* phi, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (c->fp32()) {
        CR_RET(contract_impl<float>(c, v, c->phi_recip.as<float>(), st));
        CR_HIP((launch_convert<float, double>(c->phi_recip.as<float>(), phi, v.N, st)));
    } else {
        CR_RET(contract_impl<double>(c, v, phi, st));
    }
    return CR_OK;
}

int cr_contract_force(cr_context* c, const cr_system* s, const double* q, double* force, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (c->fp32())
        CR_RET(contract_force_impl<float>(c, v, q, force, st));
    else
        CR_RET(contract_force_impl<double>(c, v, q, force, st));
    return CR_OK;
}

int cr_real_space(cr_context* c, const cr_system* s, const double* x, double* y, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (c->fp32()) {
        CR_HIP((launch_convert<double, float>(y, c->phi_recip.as<float>(), v.N, st)));
        CR_HIP(launch_real_space<float>(c->rp, v, x, c->phi_recip.as<float>(), table<float>(c, false), y, 0, st));
    } else {
        CR_HIP((launch_convert<double, double>(y, c->phi_recip.as<double>(), v.N, st)));
        CR_HIP(launch_real_space<double>(c->rp, v, x, c->phi_recip.as<double>(), table<double>(c, false), y, 0, st));
    }
    return CR_OK;
}

int cr_real_space_force(cr_context* c, const cr_system* s, const double* q, double* force, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (c->fp32())
        CR_HIP(launch_real_space_force<float>(c->rp, v, q, table<float>(c, true), c->E0, force, st));
    else
        CR_HIP(launch_real_space_force<double>(c->rp, v, q, table<double>(c, true), c->E0, force, st));
    return CR_OK;
}

int cr_profile_matvec(cr_context* c, const cr_system* s, const double* x, int reps, double* stage_ms,
                      int max_stages, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v));
    if (reps < 1 || !stage_ms || !x) return set_error(CR_ERR_INVALID, "bad argument");
    CR_RET(ensure_csr(c, v, st));
    const int nst = CR_NUM_STAGES;
    hipEvent_t ev[nst + 1];
    for (auto& e : ev) CR_HIP(hipEventCreate(&e));
    double acc[nst] = {0};
    double* y = c->w.as<double>();
    // FP32 bulk: the fixed-point scale's window bound is set once per configuration, before
    // the timed repetitions (as in a solve, where every matvec after the first reuses it)
    if (c->fp32() && c->fx_on()) CR_RET(charges_f32(c, v, x, st));
    for (int r = 0; r < reps; ++r) {
        CR_HIP(hipEventRecord(ev[0], st));
        if (c->fp32()) {
            using T = float;
            const AxisPlan* pl = plans<T>(c);
            T* real = c->grid_real.as<T>();
            T* cp = c->grid_cplx.as<T>();
            // the fixed-point spread writes every node and needs no clear; its charge
            // rounding takes the clear's place, so that the spread stage times the spread
            // kernels alone
            if (c->fx_on())
                CR_RET(charges_f32(c, v, x, st));
            else
                CR_HIP(launch_fill<T>(real, (T)0, c->G, st));
            CR_HIP(hipEventRecord(ev[1], st));
            if (c->fx_on())
                CR_RET(spread_fx(c, st));
            else
                CR_RET(spread_only<T>(c, v, x, st));
            CR_HIP(hipEventRecord(ev[2], st));
            CR_HIP(launch_fft_z_r2c<T>(pl[2], c->mp.Nx, c->mp.Ny, real, cp, c->Nzp, st));
            CR_HIP(hipEventRecord(ev[3], st));
            CR_HIP(launch_fft_yx<T>(pl[1], 1, c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, -1, st));
            CR_HIP(hipEventRecord(ev[4], st));
            CR_HIP(launch_fft_x_scale<T>(pl[0], c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, c->sc, st));
            CR_HIP(hipEventRecord(ev[5], st));
            CR_HIP(launch_fft_yx<T>(pl[1], 1, c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, +1, st));
            CR_HIP(hipEventRecord(ev[6], st));
            CR_HIP(launch_fft_z_c2r<T>(pl[2], c->mp.Nx, c->mp.Ny, cp, c->Nzp, real, st));
            CR_HIP(hipEventRecord(ev[7], st));
            CR_RET(contract_impl<T>(c, v, c->phi_recip.as<T>(), st));
            CR_HIP(hipEventRecord(ev[8], st));
            CR_RET(real_rows<T>(c, v, x, y, 1, st, true));
        } else {
            using T = double;
            const AxisPlan* pl = plans<T>(c);
            T* real = c->grid_real.as<T>();
            T* cp = c->grid_cplx.as<T>();
            CR_HIP(launch_fill<T>(real, (T)0, c->G, st));
            CR_HIP(hipEventRecord(ev[1], st));
            CR_RET(spread_only<T>(c, v, x, st));
            CR_HIP(hipEventRecord(ev[2], st));
            CR_HIP(launch_fft_z_r2c<T>(pl[2], c->mp.Nx, c->mp.Ny, real, cp, c->Nzp, st));
            CR_HIP(hipEventRecord(ev[3], st));
            CR_HIP(launch_fft_yx<T>(pl[1], 1, c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, -1, st));
            CR_HIP(hipEventRecord(ev[4], st));
            CR_HIP(launch_fft_x_scale<T>(pl[0], c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, c->sc, st));
            CR_HIP(hipEventRecord(ev[5], st));
            CR_HIP(launch_fft_yx<T>(pl[1], 1, c->mp.Nx, c->mp.Ny, c->Nzh, c->Nzp, cp, +1, st));
            CR_HIP(hipEventRecord(ev[6], st));
            CR_HIP(launch_fft_z_c2r<T>(pl[2], c->mp.Nx, c->mp.Ny, cp, c->Nzp, real, st));
            CR_HIP(hipEventRecord(ev[7], st));
            CR_RET(contract_impl<T>(c, v, c->phi_recip.as<T>(), st));
            CR_HIP(hipEventRecord(ev[8], st));
            CR_RET(real_rows<T>(c, v, x, y, 1, st));
        }
        CR_HIP(hipEventRecord(ev[9], st));
        CR_HIP(launch_constraint_rows(v, x, y + v.N, 1, st));
        CR_HIP(hipEventRecord(ev[10], st));
        CR_HIP(hipEventSynchronize(ev[10]));
        for (int k = 0; k < nst; ++k) {
            float ms = 0;
            CR_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < nst && k < max_stages; ++k) stage_ms[k] = acc[k] / reps;
    return CR_OK;
}

int cr_build_nlist_noalloc(cr_context* c, int N, const double* pos, const double L[3], double rlist,
                           unsigned* head_list, unsigned* n_neigh, unsigned* nlist, long capacity, long* total,
                           void* stream)
{
    if (!c || !L || !total || !(rlist > 0)) return set_error(CR_ERR_INVALID, "bad argument");
    if (N > 0 && (!pos || !head_list || !n_neigh || (capacity > 0 && !nlist)))
        return set_error(CR_ERR_INVALID, "null array");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 0)) return set_error(CR_ERR_INVALID, "bad box");
    const size_t need = nlist_noalloc_bytes(N, L, rlist), have = nlist_noalloc_capacity();
    if (need > have)
        return set_error(CR_ERR_INVALID, "cell-list scratch for N = " + std::to_string(N) + " needs " +
                                             std::to_string(need) + " bytes; the static scratch holds " +
                                             std::to_string(have) + " (use cr_build_nlist)");
    CR_HIP(hipSetDevice(c->device));
    CR_HIP(build_nlist_noalloc(N, reinterpret_cast<const double4*>(pos), L, rlist, head_list, n_neigh, nlist, capacity,
                               total, (hipStream_t)stream));
    return CR_OK;
}

int cr_build_nlist(cr_context* c, int N, const double* pos, const double L[3], double rlist, unsigned* head_list,
                   unsigned* n_neigh, unsigned* nlist, long capacity, long* total, void* stream)
{
    if (!c || !L || !total || !(rlist > 0)) return set_error(CR_ERR_INVALID, "bad argument");
    if (N > 0 && (!pos || !head_list || !n_neigh || (capacity > 0 && !nlist)))
        return set_error(CR_ERR_INVALID, "null array");
    for (int d = 0; d < 3; ++d)
        if (!(L[d] > 2 * rlist) && !(L[d] > 0)) return set_error(CR_ERR_INVALID, "bad box");
    CR_HIP(build_nlist(N, reinterpret_cast<const double4*>(pos), L, rlist, head_list, n_neigh, nlist, capacity, total,
                       nullptr, (hipStream_t)stream));
    return CR_OK;
}

// ---- learned warm start: features and the charge-share transform ----

int cr_mlp_features(cr_context* c, const cr_system* s, unsigned short* X, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v, false));
    if (v.N > 0 && !X) return set_error(CR_ERR_INVALID, "null feature array");
    unsigned* counts = nullptr;
    double* sums = nullptr;
    CR_RET(segments(c, v, st, &counts, &sums));
    CR_HIP(launch_mlp_features(v, c->mp.L, c->E0, counts, X, st));
    return CR_OK;
}

int cr_mlp_targets(cr_context* c, const cr_system* s, const double* q, float* t, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v, false));
    if (v.N > 0 && (!q || !t)) return set_error(CR_ERR_INVALID, "null array");
    unsigned* counts = nullptr;
    double* sums = nullptr;
    CR_RET(segments(c, v, st, &counts, &sums));
    CR_HIP(launch_mlp_share(v, counts, q, nullptr, t, nullptr, 0, st));
    return CR_OK;
}

int cr_mlp_charges(cr_context* c, const cr_system* s, const float* y, double* q, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v, false));
    if (v.N > 0 && (!q || !y)) return set_error(CR_ERR_INVALID, "null array");
    unsigned* counts = nullptr;
    double* sums = nullptr;
    CR_RET(segments(c, v, st, &counts, &sums));
    CR_HIP(launch_mlp_share(v, counts, nullptr, y, nullptr, q, 1, st));
    return CR_OK;
}

// ---- Brownian rigid-body step (trajectory driver for warm-start runs) ----

int cr_bd_step(cr_context* c, const cr_system* s, double* pos, double* rigid_com, const double* force,
               const double* noise_rigid, const double* noise_beads, double dt, double kT, double gamma_t,
               double gamma_r, double k_rep, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    SysView v;
    CR_RET(prepare(c, s, st, &v, false));
    if ((v.N > 0 && (!pos || !force)) || (v.N_rigid > 0 && (!rigid_com || !noise_rigid)))
        return set_error(CR_ERR_INVALID, "null array");
    if (!(dt > 0) || !(kT >= 0) || !(gamma_t > 0) || !(gamma_r > 0) || !(k_rep >= 0))
        return set_error(CR_ERR_INVALID, "bad BD parameter");
    CR_HIP(c->bd_force.ensure(4 * sizeof(double) * (size_t)std::max(1, v.N)));
    BdParams p;
    for (int d = 0; d < 3; ++d) p.L[d] = c->mp.L[d];
    p.dt = dt;
    p.kT = kT;
    p.gamma_t = gamma_t;
    p.gamma_r = gamma_r;
    p.k_rep = k_rep;
    CR_HIP(launch_bd_step(p, v, reinterpret_cast<double4*>(pos), reinterpret_cast<double4*>(rigid_com),
                          reinterpret_cast<const double4*>(force), c->bd_force.as<double4>(), noise_rigid,
                          noise_beads, st));
    return CR_OK;
}

// ---- tile autotuner ----

int cr_set_residual_check(cr_context* c, int on)
{
    if (!c) return set_error(CR_ERR_INVALID, "null context");
    c->check_residual = on != 0;
    return CR_OK;
}

}  // extern "C"

namespace {
// A tile shape is usable when the binned spread with FP64 accumulators fits the budget
// tg.lds with at least 16 staged beads.  That is the tightest of the binned kernels (the
// batched boxes run it in both precisions); the FP32-bulk fixed-point kernels, with 4-byte
// nodes and their own LDS plans, fit every such tile.  Larger FP32 tiles that only the
// fixed-point spread would fit were measured no faster at N = 1M.
bool tile_fits(const cr_context* c, const TileGeom& tg)
{
    const size_t Tsz = c->fp32() ? sizeof(float) : sizeof(double);
    return tiled_lds(c->mp, tg, sizeof(SpreadAcc), Tsz).chunk >= 16;
}

// the smallest LDS budget (in 32 KiB steps from 64 KiB) that holds the tile and at least 32
// staged beads (more only lowers occupancy); 0 when none does
int tile_budget(const cr_context* c, const int T3[3], int dev_max)
{
    const size_t Tsz = c->fp32() ? sizeof(float) : sizeof(double);
    for (int b = 64 * 1024; b <= dev_max; b += 32 * 1024) {
        TileGeom g{};
        for (int d = 0; d < 3; ++d) g.T[d] = T3[d];
        g.lds = b;
        if (tiled_lds(c->mp, g, sizeof(SpreadAcc), Tsz).chunk >= 32) return b;
    }
    return 0;
}
}  // namespace

extern "C" {

int cr_set_tiles(cr_context* c, const int tile[3], int lds_bytes)
{
    if (!c || !c->ready || !tile) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    const int P = c->mp.P, Nd[3] = {c->mp.Nx, c->mp.Ny, c->mp.Nz};
    int dev_max = 0;
    CR_HIP(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    if (lds_bytes < 16 * 1024 || lds_bytes > dev_max) return set_error(CR_ERR_INVALID, "LDS budget out of range");
    for (int d = 0; d < 3; ++d)
        if (tile[d] < std::max(P - 1, 1) || Nd[d] % tile[d] != 0)
            return set_error(CR_ERR_INVALID, "tile edges must divide the mesh and be >= P - 1");
    TileGeom tg{};
    for (int d = 0; d < 3; ++d) tg.T[d] = tile[d];
    tg.lds = lds_bytes;
    tg.valid = 1;
    tg.ntiles = 1;
    for (int d = 0; d < 3; ++d) {
        tg.nt[d] = Nd[d] / tg.T[d];
        tg.ntiles *= tg.nt[d];
    }
    if (!tile_fits(c, tg)) return set_error(CR_ERR_INVALID, "extended tile does not fit the LDS budget");
    c->tg = tg;
    ++c->generation;
    return CR_OK;
}

int cr_autotune(cr_context* c, const cr_system* s, cr_tuning* out, void* stream)
{
    if (!c || !c->ready) return set_error(CR_ERR_STATE, "cr_set_params has not been called");
    hipStream_t st = (hipStream_t)stream;
    const int P = c->mp.P, Nd[3] = {c->mp.Nx, c->mp.Ny, c->mp.Nz};
    int dev_max = 0;
    CR_HIP(hipDeviceGetAttribute(&dev_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    std::vector<int> opts[3];
    for (int d = 0; d < 3; ++d)
        for (int t = std::max(P - 1, 4); t <= std::min(Nd[d], 32); ++t)
            if (Nd[d] % t == 0) opts[d].push_back(t);
    const TileGeom keep = c->tg;
    hipEvent_t e0, e1;
    CR_HIP(hipEventCreate(&e0));
    CR_HIP(hipEventCreate(&e1));
    int tried = 0;
    const double* q = nullptr;
    SysView v;
    // spread + contraction of one tiling, average over reps (after one warm run)
    auto time_tiling = [&](int reps, double* ms_out) -> int {
        CR_RET(prepare(c, s, st, &v));
        if (!q) {  // any charges will do for timing: the bead records
            CR_HIP(c->yraw.ensure(sizeof(double) * (size_t)std::max(1, v.N)));
            CR_HIP(launch_fill<double>(c->yraw.as<double>(), 1.0, v.N, st));
            q = c->yraw.as<double>();
        }
        auto run = [&]() -> int {
            if (c->fp32()) {
                CR_RET(spread_impl<float>(c, v, q, st));
                CR_RET(contract_impl<float>(c, v, c->phi_recip.as<float>(), st));
            } else {
                CR_RET(spread_impl<double>(c, v, q, st));
                CR_RET(contract_impl<double>(c, v, c->phi_recip.as<double>(), st));
            }
            return CR_OK;
        };
        CR_RET(run());  // warm
        CR_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < reps; ++r) CR_RET(run());
        CR_HIP(hipEventRecord(e1, st));
        CR_HIP(hipEventSynchronize(e1));
        float ms = 0;
        CR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *ms_out = ms / reps;
        return CR_OK;
    };
    // a short timing of every admissible tiling, then a longer one of those within 10% of
    // the fastest (single short timings are within run-to-run noise of each other)
    std::vector<std::pair<double, TileGeom>> cand;
    for (int tx : opts[0])
        for (int ty : opts[1])
            for (int tz : opts[2]) {
                const int T3[3] = {tx, ty, tz};
                const int budget = tile_budget(c, T3, dev_max);
                if (!budget || cr_set_tiles(c, T3, budget) != CR_OK) continue;
                double ms = 0;
                CR_RET(time_tiling(3, &ms));
                ++tried;
                cand.emplace_back(ms, c->tg);
            }
    double first = 1e300;
    for (const auto& cd : cand) first = std::min(first, cd.first);
    double best = 1e300;
    TileGeom best_tg = keep;
    for (const auto& cd : cand) {
        if (cd.first > 1.1 * first) continue;
        c->tg = cd.second;
        ++c->generation;
        double ms = 0;
        CR_RET(time_tiling(10, &ms));
        if (ms < best) {
            best = ms;
            best_tg = cd.second;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tg = tried ? best_tg : keep;
    ++c->generation;
    if (out) {
        for (int d = 0; d < 3; ++d) out->tile[d] = c->tg.T[d];
        out->lds_bytes = c->tg.lds;
        out->spread_contract_ms = tried ? best : -1.0;
        out->candidates = tried;
    }
    return CR_OK;
}

}  // extern "C"
