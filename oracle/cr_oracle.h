/*
 * cr_oracle.h -- CPU FP64 restatement of ConductorRigid's force compute.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never links or calls it.
 *
 * Every entry point cites the reference function it restates
 * (cppmodule/ConductorRigid.cc, cppmodule/ConductorRigid.cu).
 * All arithmetic is IEEE double; all pointers are host memory.
 */
#ifndef CR_ORACLE_H
#define CR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define CRO_NOT_MEMBER 0xffffffffu

/* Numerical parameters, as produced by SetParams
 * (cppmodule/ConductorRigid.cc:116-379). */
typedef struct {
    double L[3];      /* box edge lengths (orthorhombic, centred on 0) */
    double xi;        /* Ewald splitting parameter */
    double errortol;  /* error tolerance */
    double rc;        /* real-space cutoff */
    double kcut;      /* wave-space cutoff */
    int    N[3];      /* mesh nodes per dimension */
    int    P;         /* window support (nodes per dimension) */
    double h[3];      /* mesh spacing */
    double eta[3];    /* spectral splitting parameter */
    int    N_table;   /* real-space table: entries 0..N_table */
    double drtable;   /* table spacing (0.001) */
    double selfterm;  /* real-space self coefficient */
    double prefac;    /* Gaussian prefactor (2 xi^2/pi)^{3/2} / sqrt(eta_x eta_y eta_z) */
    double xiterm;    /* 2 xi^2 */
    double quadW;     /* h_x h_y h_z */
} cro_params;

/* HOOMD-shaped system view (host copies of what ConductorRigid.cc hands to
 * gpu_ComputeForce, cppmodule/ConductorRigid.cc:494-544).  The active group is
 * every particle, in global-index order. */
typedef struct {
    int N;                             /* beads */
    const double* pos;                 /* [4N] x,y,z,type by global index */
    const unsigned* tag;               /* [N] particle tag of each global index */
    const unsigned* body_of_tag;       /* [N] isolated-body tag per particle tag (user `body`) */
    const double* body_charge;         /* [N_isolated] net charge per isolated-body tag (user `charge`) */
    int N_rigid;                       /* rigid bodies */
    int N_isolated;                    /* rigid bodies + isolated beads */
    int Nmax;                          /* pitch of rigid_members */
    const unsigned* rigid_membership;  /* [N] rigid index per global index, or NOT_MEMBER */
    const unsigned* rigid_members;     /* [N_rigid*Nmax] global indices, NOT_MEMBER padded */
    const unsigned* rigid_size;        /* [N_rigid] */
    const double* rigid_com;           /* [4*N_rigid] centre of mass */
    const unsigned* nlist;             /* neighbour indices (full list) */
    const unsigned* head_list;         /* [N] start of each bead's neighbours */
    const unsigned* n_neigh;           /* [N] neighbour count */
} cro_system;

/* --- parameters and tables (ConductorRigid.cc:116-379) --- */
int  cro_smooth_list(int* out, int cap);
void cro_set_params(cro_params* p, const double L[3], double xi, double errortol,
                    double rc_override, const int mesh_override[3], int P_override);
void cro_tables(const cro_params* p, double* phiq /*[2*(N_table+1)]*/,
                double* gradphiq /*[2*(N_table+1)]*/);
void cro_kscale(const cro_params* p, double* S /*[G]*/);
double cro_phiq_exact(double r, double xi);
double cro_gradphiq_exact(double r, double xi);
double cro_table_lookup(const cro_params* p, const double* table, double dist);

/* --- spectral Ewald pieces (ConductorRigid.cu:144-466) --- */
void cro_spread(const cro_params* p, int N, const double* pos, const double* q, double* grid);
void cro_fft3d(const int n[3], double* re, double* im, int sign);
void cro_dft3d_brute(const int n[3], const double* re, const double* im,
                     double* ore, double* oim, int sign);
void cro_scale(const cro_params* p, const double* S, double* re, double* im);
void cro_recip_grid(const cro_params* p, const double* S, double* grid_inout);
void cro_contract(const cro_params* p, int N, const double* pos, const double* grid, double* phi);
void cro_contract_force(const cro_params* p, int N, const double* pos, const double* q,
                        const double* grid, double* force4);

/* --- real space, constraints, objective, preconditioner (ConductorRigid.cu:469-859) --- */
void cro_membership(const cro_system* s, unsigned* iso_membership /*[N]*/);
void cro_real_space(const cro_params* p, const cro_system* s, const double* table,
                    const unsigned* iso_membership, const double* q, const double* Phi,
                    double* phiq_inout);
void cro_real_space_force(const cro_params* p, const cro_system* s, const double* gtable,
                          const double* q, const double E0[3], double* force4_inout);
void cro_sum_charges(const cro_system* s, const unsigned* iso_membership, const double* q,
                     double* out /*[N_isolated]*/);
void cro_objective(const cro_system* s, const double L[3], const double E0[3],
                   double* b /*[N+N_isolated]*/);
void cro_precondition(const cro_system* s, const unsigned* iso_membership,
                      const double* x, double* y);

/* --- composite operations --- */
typedef struct cro_ctx cro_ctx;
cro_ctx* cro_create(const cro_params* p, const cro_system* s);
void     cro_destroy(cro_ctx* c);
void     cro_matvec(cro_ctx* c, const double* x, double* y);
void     cro_apply_precond(cro_ctx* c, const double* x, double* y);
int      cro_gmres(cro_ctx* c, double* x, const double* b, int restart, int maxiter,
                   double tol, int* iters, double* resid);
int      cro_compute_forces(cro_ctx* c, const double E0[3], double* q_inout, double* Phi_inout,
                            double* force4, int restart, int maxiter, double tol,
                            int* iters, double* resid);

/* --- direct Ewald oracle: closed-form real space, explicit k-sum, no mesh --- */
void cro_direct_matvec(const cro_params* p, const cro_system* s, double rc_real,
                       double kmax, const double* q, double* phi);
void cro_direct_force(const cro_params* p, const cro_system* s, double rc_real, double kmax,
                      const double* q, const double E0[3], double* force4);

/* --- neighbour list in HOOMD layout (full list, all pairs with r < rlist) --- */
long cro_build_nlist(int N, const double* pos, const double L[3], double rlist,
                     unsigned* head_list, unsigned* n_neigh, unsigned* nlist, long cap);

int cro_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
