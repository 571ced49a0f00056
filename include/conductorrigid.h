/*
 * conductorrigid.h -- C ABI of the MI355X-native ConductorRigid force compute.
 *
 * Drop-in boundary for the reference's device ABI (cppmodule/ConductorRigid.cuh:19-129):
 * the reference's host class ConductorRigid (cppmodule/ConductorRigid.cc) calls four
 * host-callable functions that take raw device pointers -- gpu_ZeroForce,
 * gpu_ComputeForce, MatrixVectorMultiply and Precondition -- and owns a cuFFT plan,
 * the k-space scale table and the real-space tables built by SetParams.  Here the
 * same roles are served by an opaque context that owns the mesh buffers, FFT
 * twiddles, tables and Krylov workspace; every compute entry point takes plain
 * device pointers and sizes plus an explicit HIP stream (passed as void*), and
 * returns an int status (CR_OK == 0).  No torch or HIP types appear here.
 *
 * Precision: positions, charges, potentials and forces cross the boundary as
 * IEEE double (HOOMD's Scalar in double precision).  CR_PRECISION_FP64 computes
 * everything in double ("plain method"); CR_PRECISION_FP32_BULK runs spreading,
 * FFTs, contraction and the pair sum in float and keeps the Krylov solver, the
 * residual check and the per-body constraint sums in double.
 */
#ifndef CONDUCTORRIGID_H
#define CONDUCTORRIGID_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CR_OK 0
#define CR_ERR_INVALID 1     /* bad argument or inconsistent sizes */
#define CR_ERR_DEVICE 2      /* HIP runtime error */
#define CR_ERR_NOT_CONVERGED 3
#define CR_ERR_STATE 4       /* call out of order (e.g. compute before set_params) */

#define CR_NOT_MEMBER 0xffffffffu

#define CR_PRECISION_FP64 0
#define CR_PRECISION_FP32_BULK 1

/* x0 modes for the solve (the reference always uses the previous step's (q, Phi),
 * cppmodule/ConductorRigid.cu:1125-1127) */
#define CR_X0_PREVIOUS 0     /* start from the caller's q_inout / phi_inout */
#define CR_X0_ZERO 1         /* start from zero */
#define CR_X0_CHARGES 2      /* start from the caller's q_inout (e.g. a learned prediction), shifted per
                                body onto its net charge, with the body potentials fitted to the
                                equipotential rows for that q (one extra operator application);
                                phi_inout is output only.  No reference counterpart. */

typedef struct cr_context cr_context;

/* Numerical parameters: the reference derives all of them from (xi, errortol) in
 * SetParams (cppmodule/ConductorRigid.cc:116-207); any field > 0 overrides the derived
 * value.  GMRES defaults follow the reference (restart 10, 100 iterations, tolerance
 * errortol, cppmodule/ConductorRigid.cu:1134-1136). */
typedef struct {
    double L[3];          /* orthorhombic box lengths; box spans [-L/2, L/2) */
    double xi;            /* Ewald splitting parameter */
    double errortol;      /* error tolerance */
    double rc;            /* real-space cutoff override (0: sqrt(-ln tol)/xi) */
    int mesh[3];          /* mesh override (0: derived, rounded up to 5-smooth) */
    int P;                /* window support override (0: ceil(-2 ln tol / pi)) */
    int precision;        /* CR_PRECISION_* */
    int restart;          /* GMRES restart (0: 10) */
    int max_iter;         /* GMRES iteration limit (0: 100) */
    double gmres_tol;     /* relative GMRES tolerance (0: errortol) */
    int preconditioner;   /* CR_PRECOND_* */
} cr_params;

/* Preconditioners.  DIAGONAL is the reference's (cppmodule/ConductorRigid.cu:816-859) and
 * keeps CUSP's stopping rule |s_{i+1}| <= tol ||b||_2.  NEARFIELD_BLOCK inverts each body's
 * own saddle block (free-space shell-shell interactions of its beads plus its constraint
 * row) and stops at ||D r|| <= tol ||D b||_2; it has no counterpart in the reference. */
#define CR_PRECOND_DIAGONAL 0
#define CR_PRECOND_NEARFIELD_BLOCK 1

/* Parameters as resolved by cr_set_params (SetParams' derived quantities). */
typedef struct {
    double rc, kcut;
    int mesh[3];
    int P;
    double h[3], eta[3];
    int N_table;
    double drtable, selfterm, prefac, xiterm, quadW;
    int precision, restart, max_iter;
    double gmres_tol;
} cr_resolved;

/* HOOMD-shaped system view: all pointers are DEVICE pointers (borrowed for the call).
 * Mirrors the arrays ConductorRigid::computeForces hands to gpu_ComputeForce
 * (cppmodule/ConductorRigid.cc:420-544).  The active group is every particle. */
typedef struct {
    int N;                              /* beads */
    const double* pos;                  /* [4N] x, y, z, type by global index */
    const unsigned* tag;                /* [N] tag of each global index */
    const unsigned* body_of_tag;        /* [N] isolated-body tag per particle tag (user `body`) */
    const double* body_charge;          /* [N_isolated] net charge per isolated body (user `charge`) */
    int N_rigid;                        /* rigid bodies */
    int N_isolated;                     /* rigid bodies + isolated beads */
    int Nmax;                           /* pitch of rigid_members */
    const unsigned* rigid_membership;   /* [N] rigid index per global index or CR_NOT_MEMBER */
    const unsigned* rigid_members;      /* [N_rigid*Nmax] global indices, CR_NOT_MEMBER padded */
    const unsigned* rigid_size;         /* [N_rigid] */
    const double* rigid_com;            /* [4*N_rigid] centres of mass */
    const unsigned* nlist;              /* full neighbour list */
    const unsigned* head_list;          /* [N] first neighbour of each bead */
    const unsigned* n_neigh;            /* [N] neighbour count of each bead */
    double L[3];                        /* box lengths of this call (HOOMD passes its BoxDim to
                                           gpu_ComputeForce every step); all zero: the context's box.
                                           A different box is taken up first, as cr_set_box does. */
} cr_system;

typedef struct {
    int iterations;        /* Arnoldi steps */
    int matvecs;           /* operator applications incl. restart residuals */
    int restarts;
    int converged;         /* 1 when the residual estimate met tol * ||b|| */
    double resid;          /* final preconditioned residual estimate */
    double tolerance;      /* tol * ||b||_2 */
    double true_resid;     /* ||D (b - A x)||_2 recomputed in FP64 after the solve */
    double true_relres;    /* ||b - A x||_2 / ||b||_2 (unpreconditioned), FP64 */
    int fallback;          /* 1 when the solve was redone from x0 = 0 (see cr_compute_force) */
} cr_solve_stats;

/* ---- lifetime ---------------------------------------------------------- */
int  cr_create(cr_context** out, int device);
void cr_destroy(cr_context* ctx);
const char* cr_last_error(void);

/* SetParams (cppmodule/ConductorRigid.cc:116-379): resolve parameters, build the
 * real-space tables, FFT plan/twiddles and size the mesh workspace.  n_beads and
 * n_isolated size the Krylov workspace. */
int cr_set_params(cr_context* ctx, const cr_params* params, int n_beads, int n_isolated);
int cr_get_resolved(const cr_context* ctx, cr_resolved* out);

/* UpdateField (cppmodule/ConductorRigid.cc:381-386). */
int cr_update_field(cr_context* ctx, const double E0[3]);

/* Box change (SetParams reads the box, cppmodule/ConductorRigid.cc:116-121; HOOMD hands
 * gpu_ComputeForce the current BoxDim every step, cppmodule/ConductorRigid.cuh:23-73).
 * Keeps the mesh, window support, cutoff and real-space tables (they depend on xi and the
 * tolerance only) and regenerates on the device what depends on the box: the mesh spacing
 * h, eta, the window prefactor, the quadrature weight and the k-space scale tables.  When
 * the mesh was derived from the box (no mesh override) and the new box derives another
 * mesh, the parameters are set again in full (cr_set_params).  Work is queued on stream.
 * Every compute entry point calls this itself when its cr_system carries a box that
 * differs from the context's. */
int cr_set_box(cr_context* ctx, const double L[3], void* stream);

/* ---- the reference's device ABI ---------------------------------------- */

/* gpu_ZeroForce (cppmodule/ConductorRigid.cuh:19-21). force: device double[4N]. */
int cr_zero_force(cr_context* ctx, int N, double* force, void* stream);

/* gpu_ComputeForce (cppmodule/ConductorRigid.cuh:23-73): solve for the bead charges
 * and body potentials (GMRES, x0 per x0_mode), write them back to q_inout [N] and
 * phi_inout [N_isolated], and overwrite force [4N] with
 * F = q E0 + reciprocal + real-space forces.  stats may be NULL.
 * Plain GMRES from x0 = 0 is the fallback: a solve started from CR_X0_PREVIOUS or
 * CR_X0_CHARGES is redone from zero (stats->fallback = 1, iteration counts summed) when it
 * meets a non-finite residual, stops at the iteration limit without converging, or -- with
 * the residual check on and stats given -- ends with an FP64 true residual ||D(b - Ax)|| above
 * 10x the solver tolerance.  The reference accepts any outcome
 * (cppmodule/ConductorRigid.cu:1142-1147). */
int cr_compute_force(cr_context* ctx, const cr_system* sys, double* q_inout, double* phi_inout,
                     double* force, int x0_mode, cr_solve_stats* stats, void* stream);

/* gpu_ComputeForce as a SINGLE_PRECISION HOOMD build hands it over (the only build the
 * reference compiles in, cppmodule/ConductorRigid.h:6-10; Scalar4 / Scalar pointers,
 * cppmodule/ConductorRigid.cuh:23-73): the same view with float arrays for positions
 * (float4 x, y, z, type), body net charges and rigid centres of mass (float4), float bead
 * charges q_inout [N], body potentials phi_inout [N_isolated] and forces [4N] (float4,
 * w = 0).  The arrays are converted to double on the device and the call proceeds as
 * cr_compute_force (in the context's precision); results are rounded back to float. */
typedef struct {
    int N;
    const float* pos;                   /* [4N] x, y, z, type by global index */
    const unsigned* tag;
    const unsigned* body_of_tag;
    const float* body_charge;           /* [N_isolated] */
    int N_rigid;
    int N_isolated;
    int Nmax;
    const unsigned* rigid_membership;
    const unsigned* rigid_members;
    const unsigned* rigid_size;
    const float* rigid_com;             /* [4*N_rigid] */
    const unsigned* nlist;
    const unsigned* head_list;
    const unsigned* n_neigh;
    double L[3];
} cr_system_f32;
int cr_compute_force_f32(cr_context* ctx, const cr_system_f32* sys, float* q_inout, float* phi_inout,
                         float* force, int x0_mode, cr_solve_stats* stats, void* stream);

/* MatrixVectorMultiply (cppmodule/ConductorRigid.cuh:76-114): y = A x with
 * x = [q (N) ; Phi (N_isolated)], y = [Phi_b(i) - (M q)_i ; sum_{i in b} q_i]. */
int cr_matvec(cr_context* ctx, const cr_system* sys, const double* x, double* y, void* stream);

/* Precondition (cppmodule/ConductorRigid.cuh:117-129): y = D x, D diagonal. */
int cr_precondition(cr_context* ctx, const cr_system* sys, const double* x, double* y, void* stream);

/* compute_objective (cppmodule/ConductorRigid.cu:764-813): b = [-(r_i - R_b).E0 ; Q_b]. */
int cr_objective(cr_context* ctx, const cr_system* sys, double* b, void* stream);

/* ---- stages of the matvec, exposed for parity tests and kernel timing ---- */

/* spread (cppmodule/ConductorRigid.cu:144-234) of the charges q [N] onto the real
 * mesh; grid_out (may be NULL) receives a double copy [Nx*Ny*Nz], z fastest. */
int cr_spread(cr_context* ctx, const cr_system* sys, const double* q, double* grid_out, void* stream);

/* grid <- Re IFFT(S * FFT(grid)) on the context's real mesh (cufftExecC2C forward,
 * scale, inverse: cppmodule/ConductorRigid.cu:927-933).  grid_in/grid_out (double,
 * may be NULL) load / read back the real mesh. */
int cr_recip_grid(cr_context* ctx, const double* grid_in, double* grid_out, void* stream);

/* Forward / inverse 3-D FFT of a full complex double mesh (interleaved re,im),
 * unnormalised, sign -1 / +1 (cuFFT C2C conventions).  For tests of the FFT kernels. */
int cr_fft_c2c(cr_context* ctx, const double* in, double* out, int sign, void* stream);

/* contract (cppmodule/ConductorRigid.cu:260-360): phi [N] from the context's mesh. */
int cr_contract(cr_context* ctx, const cr_system* sys, double* phi, void* stream);

/* contract_force (cppmodule/ConductorRigid.cu:363-466): force [4N] (overwritten). */
int cr_contract_force(cr_context* ctx, const cr_system* sys, const double* q, double* force,
                      void* stream);

/* real_space (cppmodule/ConductorRigid.cu:469-567): y[i] = Phi[b(i)] - (y[i] + self q_i
 * + sum_j A(r_ij) q_j) for x = [q ; Phi]. */
int cr_real_space(cr_context* ctx, const cr_system* sys, const double* x, double* y, void* stream);

/* real_space_force (cppmodule/ConductorRigid.cu:570-664): force += q E0 + pair forces. */
int cr_real_space_force(cr_context* ctx, const cr_system* sys, const double* q, double* force,
                        void* stream);

/* ---- kernel timing ------------------------------------------------------
 * Runs each stage of the matvec y = A x `reps` times on `stream`, bracketing every
 * stage with HIP events, and writes the average milliseconds per launch into
 * stage_ms[k] for k < min(max_stages, CR_NUM_STAGES), in this order:
 *   0 mesh clear (in FP32 bulk with the fixed-point charge rounding), 1 spread, 2 FFT z R2C,
 *   3 FFT y forward, 4 FFT x forward+scale+inverse, 5 FFT y inverse, 6 FFT z C2R,
 *   7 contract, 8 real space, 9 constraint rows.
 * No reference counterpart: the reference has no per-kernel timing (HOOMD's
 * Profiler wraps the whole step, cppmodule/ConductorRigid.cc:415-416). */
#define CR_NUM_STAGES 10
int cr_profile_matvec(cr_context* ctx, const cr_system* sys, const double* x, int reps, double* stage_ms,
                      int max_stages, void* stream);

/* ---- neighbour list (the HOOMD NeighborListGPUBinned the reference's compute.py
 * creates, pymodule/compute.py:38-44): full list, r < rlist, intra-body pairs kept.
 * head_list/n_neigh: device uint[N]; nlist: device uint[capacity].  Returns the
 * total pair count in *total; if it exceeds capacity nothing past capacity is
 * written and the caller retries with a larger buffer. */
int cr_build_nlist(cr_context* ctx, int N, const double* pos, const double L[3], double rlist,
                   unsigned* head_list, unsigned* n_neigh, unsigned* nlist, long capacity,
                   long* total, void* stream);

/* No reference counterpart: cr_build_nlist without allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, such as a captured Brownian step
 * that rebuilds the list; tight serving loops).  Same arguments and results as
 * cr_build_nlist.  The cell-list working arrays (8 N bytes, 12 bytes per cell of edge
 * >= rlist and the scan sums, each rounded to 256 bytes) live in a 16 MiB static device
 * array the library holds on every GPU it runs on until the process ends: about 1.6
 * million beads at rlist = 4.4 and volume fraction 0.3.  Larger inputs are rejected with
 * CR_ERR_INVALID.  Calls on one GPU share that array: calls that overlap on one GPU race,
 * and each call clears the counters it needs with kernels (no memset nodes). */
int cr_build_nlist_noalloc(cr_context* ctx, int N, const double* pos, const double L[3], double rlist,
                           unsigned* head_list, unsigned* n_neigh, unsigned* nlist, long capacity, long* total,
                           void* stream);

/* ---- learned warm start -------------------------------------------------
 * No reference counterpart: the reference starts each solve from the previous step's
 * (q, Phi) (cppmodule/ConductorRigid.cu:1125-1127).  A per-bead MLP predicts each bead's
 * charge relative to its body's uniform share Q_b / n_b, in units of 4 pi; the
 * prediction enters the solve through CR_X0_CHARGES, so GMRES (to the same tolerance)
 * still decides the answer.  At fixed geometry the charges are linear in E0, so the
 * network outputs four geometric responses u = (a_x, a_y, a_z, c) and predicts
 * y = E0 . a + c.  Trunk: CR_MLP_TRUNK bf16 features -> 128 -> 128 -> 4, SiLU, bf16 weights
 * with FP32 accumulation on MFMA; FP32 master weights trained with Adam on the device.
 * Feature row (CR_MLP_IN columns, bf16 bits): the 32 nearest neighbours within r < 3 as
 * d = (dx, dy, dz)/3, a same-body flag (+1 / -1) and |d| (empty slots 0) [0, 160);
 * Q_b / n_b [160]; the offset o from the rigid body's centre of mass / 4 and |o|
 * [161, 165); zero [165, 168); E0 as bf16 pairs hi [168, 171), lo [171, 174); zero.
 * Parameter vector (FP32, CR_MLP_PARAMS):
 * W1 [CR_MLP_IN][128] | b1 [128] | W2 [128][128] | b2 [128] | W3 [128][4] | b3 [4]. */
#define CR_MLP_IN 176
#define CR_MLP_TRUNK 168
#define CR_MLP_HIDDEN 128
#define CR_MLP_OUT 4
#define CR_MLP_PARAMS 39684
typedef struct cr_mlp cr_mlp;
int  cr_mlp_create(cr_mlp** out, int device, unsigned long long seed);
void cr_mlp_destroy(cr_mlp* m);
int  cr_mlp_set_params(cr_mlp* m, const float* theta);      /* host [CR_MLP_PARAMS]; resets Adam */
int  cr_mlp_get_params(const cr_mlp* m, float* theta);      /* host [CR_MLP_PARAMS] */
/* X: device bf16 bits [N * CR_MLP_IN] for the system's beads (field from the context) */
int  cr_mlp_features(cr_context* ctx, const cr_system* sys, unsigned short* X, void* stream);
/* t_i = (q_i - Q_b / n_b) / (4 pi)  (training targets from solved charges; device arrays) */
int  cr_mlp_targets(cr_context* ctx, const cr_system* sys, const double* q, float* t, void* stream);
/* q_i = Q_b / n_b + 4 pi y_i  (charges from network outputs; device arrays) */
int  cr_mlp_charges(cr_context* ctx, const cr_system* sys, const float* y, double* q, void* stream);
/* y [M] = network(X [M rows]) */
int  cr_mlp_predict(cr_mlp* m, int M, const unsigned short* X, float* y, void* stream);
/* grad [CR_MLP_PARAMS] (device FP32) of mean_m (y_m - t_m)^2, y [M] optional */
int  cr_mlp_gradient(cr_mlp* m, int M, const unsigned short* X, const float* t, float* y, float* grad, void* stream);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that loss; y [M] optional */
int  cr_mlp_train_step(cr_mlp* m, int M, const unsigned short* X, const float* t, float lr, float* y, void* stream);

/* ---- Brownian rigid-body step -----------------------------------------------
 * The force's downstream consumer (HOOMD integrates the bodies in the reference's
 * setting); used to generate trajectories.  Per rigid body: F = sum of member forces,
 * T = sum (r_i - R) x F_i; dR = F dt/gamma_t + sqrt(2 kT dt/gamma_t) n, rotation vector
 * T dt/gamma_r + sqrt(2 kT dt/gamma_r) n' applied about R.  Beads outside rigid bodies
 * translate.  Bead forces include a soft steric repulsion k_rep (2 - r) r_hat from beads of
 * other bodies closer than contact (r < 2).
 * noise_rigid: device [6 N_rigid] standard normals; noise_beads: [3 N] or NULL.
 * pos [4N] and rigid_com [4 N_rigid] (the arrays sys points to) are updated in place;
 * the neighbour list must be rebuilt by the caller. */
int cr_bd_step(cr_context* ctx, const cr_system* sys, double* pos, double* rigid_com, const double* force,
               const double* noise_rigid, const double* noise_beads, double dt, double kT, double gamma_t,
               double gamma_r, double k_rep, void* stream);

/* ---- spread/contract tiling ---------------------------------------------
 * No reference counterpart (the reference launches one thread block per bead).  The
 * binned spread and contraction work on mesh tiles whose edges divide the mesh and are
 * at least P - 1; each workgroup holds its tile extended by P - 1 nodes per axis in LDS.
 * cr_set_params picks a tiling by rule; cr_autotune times the spread + contraction of
 * the given system for every admissible tile shape (edges 4..32, each with the smallest
 * LDS budget that holds it) and keeps the fastest; cr_set_tiles sets one explicitly. */
typedef struct {
    int tile[3];
    int lds_bytes;
    double spread_contract_ms;   /* best measured time of one spread + contraction */
    int candidates;              /* tile shapes timed */
} cr_tuning;
int cr_autotune(cr_context* ctx, const cr_system* sys, cr_tuning* out, void* stream);
int cr_set_tiles(cr_context* ctx, const int tile[3], int lds_bytes);

/* No reference counterpart: the reference never checks the converged state
 * (cppmodule/ConductorRigid.cu:1142-1147).  With the check on (the default), a force
 * evaluation that is given a stats struct also recomputes the residual from one FP64
 * operator application (true_resid, true_relres); off, those two fields are set to -1 and
 * the evaluation does only the solve and the forces.  Applies to cr_compute_force and, via
 * each box's context, to cr_batch_compute_force (checked when every box has it on). */
int cr_set_residual_check(cr_context* ctx, int on);

/* ---- batched boxes ------------------------------------------------------
 * No reference counterpart: the reference computes one box per computeForces call
 * (cppmodule/ConductorRigid.cc:403-557).  A batch solves many independent periodic boxes
 * together.  Each box keeps its own context (cr_set_params with its own box, mesh and
 * tolerance; cr_autotune for its tiling) and its own system; cr_batch_compute_force gives
 * every box the result cr_compute_force would give it (same outputs, same per-box
 * stopping rule and restart cycle), but the boxes advance through GMRES in lockstep: one
 * binned spread and one binned contraction launch over the tiles of all boxes per operator
 * application, segmented FP64 Krylov launches over all boxes per Arnoldi step, and one
 * host sync per step for the whole batch.  use_graph != 0 captures the Arnoldi step once
 * as a HIP graph (re-captured only when buffers or launch shapes change) and replays it.
 * All boxes of a call must share the device, precision, window support P, restart,
 * iteration limit and preconditioner, and have a binned tiling; a context may appear
 * once per call.  q_inout / phi_inout / force / stats hold one entry per box (stats may
 * be NULL).  Work runs on the batch's own streams, ordered after the caller's stream and
 * before its later work. */
typedef struct cr_batch cr_batch;
int  cr_batch_create(cr_batch** out, int device);
void cr_batch_destroy(cr_batch* batch);
int  cr_batch_compute_force(cr_batch* batch, int nbox, cr_context* const* ctx, const cr_system* sys,
                            double* const* q_inout, double* const* phi_inout, double* const* force, int x0_mode,
                            int use_graph, cr_solve_stats* stats, void* stream);

#ifdef __cplusplus
}
#endif
#endif
