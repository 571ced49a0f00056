// k_dynamics.hip -- overdamped (Brownian) rigid-body step driven by the bead forces:
// the downstream consumer of the force compute (HOOMD integrates the rigid bodies in
// the reference's setting; pymodule/compute.py only supplies forces).  Used to generate
// trajectories for warm-start training and evaluation.
//
// Per rigid body b with centre R_b and members i:
//   F_b = sum_i F_i,  T_b = sum_i (r_i - R_b) x F_i   (minimum image)
//   dR = F_b dt / gamma_t + sqrt(2 kT dt / gamma_t) n_t
//   dtheta = T_b dt / gamma_r + sqrt(2 kT dt / gamma_r) n_r
// members are rotated about R_b by the rotation vector dtheta (Rodrigues), then
// translated by dR.  Beads outside rigid bodies translate with their own force.
// n_t, n_r: caller-supplied standard normals (6 per rigid body, 3 per bead).
// F_i includes a soft steric repulsion k_rep (2 - r) r_hat from every bead of another
// body closer than contact (r < 2), so that bodies do not drift into each other.
#include "cr_kernels.h"

namespace cr {

namespace {

__device__ __forceinline__ double wrap1(double x, double L)
{
    double y = x - L * floor((x + 0.5 * L) / L);
    return (y >= 0.5 * L) ? y - L : y;
}

__device__ __forceinline__ double mimg(double d, double L) { return d - L * rint(d / L); }

// f_out[i] = force[i] + sum over neighbours j of other bodies with r < 2 of k (2 - r) r_hat
__global__ void k_bd_forces(SysView s, BdParams p, const double4* __restrict__ force, double4* __restrict__ f_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.N) return;
    const double4 ri = s.pos[i];
    double4 f = force[i];
    const unsigned bi = s.iso[i], head = s.head_list[i], nn = s.n_neigh[i];
    for (unsigned k = 0; k < nn; ++k) {
        const unsigned j = s.nlist[head + k];
        if (s.iso[j] == bi) continue;
        const double4 rj = s.pos[j];
        const double dx = mimg(ri.x - rj.x, p.L[0]), dy = mimg(ri.y - rj.y, p.L[1]), dz = mimg(ri.z - rj.z, p.L[2]);
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r < 2.0 && r > 0.0) {
            const double c = p.k_rep * (2.0 - r) / r;
            f.x += c * dx;
            f.y += c * dy;
            f.z += c * dz;
        }
    }
    f_out[i] = f;
}

__global__ void k_rigid_bd(BdParams p, int N_rigid, int Nmax, const unsigned* __restrict__ members,
                           const unsigned* __restrict__ size, double4* __restrict__ pos, double4* __restrict__ com,
                           const double4* __restrict__ force, const double* __restrict__ noise)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N_rigid) return;
    const double4 R = com[b];
    const unsigned n = size[b];
    double F[3] = {0, 0, 0}, T[3] = {0, 0, 0};
    for (unsigned k = 0; k < n; ++k) {
        const unsigned i = members[(long)b * Nmax + k];
        const double4 r = pos[i];
        const double4 f = force[i];
        const double d[3] = {mimg(r.x - R.x, p.L[0]), mimg(r.y - R.y, p.L[1]), mimg(r.z - R.z, p.L[2])};
        F[0] += f.x;
        F[1] += f.y;
        F[2] += f.z;
        T[0] += d[1] * f.z - d[2] * f.y;
        T[1] += d[2] * f.x - d[0] * f.z;
        T[2] += d[0] * f.y - d[1] * f.x;
    }
    const double st = sqrt(2.0 * p.kT * p.dt / p.gamma_t), sr = sqrt(2.0 * p.kT * p.dt / p.gamma_r);
    const double* nz = noise + 6 * (long)b;
    double dR[3], th[3];
    for (int a = 0; a < 3; ++a) {
        dR[a] = F[a] * p.dt / p.gamma_t + st * nz[a];
        th[a] = T[a] * p.dt / p.gamma_r + sr * nz[3 + a];
    }
    const double ang = sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
    double ax[3] = {0, 0, 1};
    if (ang > 0) {
        ax[0] = th[0] / ang;
        ax[1] = th[1] / ang;
        ax[2] = th[2] / ang;
    }
    const double c = cos(ang), s = sin(ang);
    const double Rn[3] = {R.x + dR[0], R.y + dR[1], R.z + dR[2]};
    for (unsigned k = 0; k < n; ++k) {
        const unsigned i = members[(long)b * Nmax + k];
        double4 r = pos[i];
        const double v[3] = {mimg(r.x - R.x, p.L[0]), mimg(r.y - R.y, p.L[1]), mimg(r.z - R.z, p.L[2])};
        // Rodrigues: v' = v c + (a x v) s + a (a.v)(1 - c)
        const double kv[3] = {ax[1] * v[2] - ax[2] * v[1], ax[2] * v[0] - ax[0] * v[2], ax[0] * v[1] - ax[1] * v[0]};
        const double kd = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
        r.x = wrap1(Rn[0] + v[0] * c + kv[0] * s + ax[0] * kd * (1 - c), p.L[0]);
        r.y = wrap1(Rn[1] + v[1] * c + kv[1] * s + ax[1] * kd * (1 - c), p.L[1]);
        r.z = wrap1(Rn[2] + v[2] * c + kv[2] * s + ax[2] * kd * (1 - c), p.L[2]);
        pos[i] = r;
    }
    double4 Rw = R;
    Rw.x = wrap1(Rn[0], p.L[0]);
    Rw.y = wrap1(Rn[1], p.L[1]);
    Rw.z = wrap1(Rn[2], p.L[2]);
    com[b] = Rw;
}

__global__ void k_bead_bd(BdParams p, int N, const unsigned* __restrict__ rigid_membership, double4* __restrict__ pos,
                          const double4* __restrict__ force, const double* __restrict__ noise)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || rigid_membership[i] != NOT_MEMBER) return;
    const double st = sqrt(2.0 * p.kT * p.dt / p.gamma_t);
    double4 r = pos[i];
    const double4 f = force[i];
    r.x = wrap1(r.x + f.x * p.dt / p.gamma_t + st * noise[3 * (long)i], p.L[0]);
    r.y = wrap1(r.y + f.y * p.dt / p.gamma_t + st * noise[3 * (long)i + 1], p.L[1]);
    r.z = wrap1(r.z + f.z * p.dt / p.gamma_t + st * noise[3 * (long)i + 2], p.L[2]);
    pos[i] = r;
}

inline unsigned nb(long n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace

hipError_t launch_bd_step(const BdParams& p, const SysView& s, double4* pos, double4* com, const double4* force_in,
                          double4* force_tmp, const double* noise_rigid, const double* noise_beads, hipStream_t st)
{
    if (s.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bd_forces, dim3(nb(s.N)), dim3(256), 0, st, s, p, force_in, force_tmp);
    const double4* force = force_tmp;
    if (s.N_rigid > 0)
        hipLaunchKernelGGL(k_rigid_bd, dim3(nb(s.N_rigid)), dim3(256), 0, st, p, s.N_rigid, s.Nmax, s.rigid_members,
                           s.rigid_size, pos, com, force, noise_rigid);
    if (s.N > 0 && noise_beads)
        hipLaunchKernelGGL(k_bead_bd, dim3(nb(s.N)), dim3(256), 0, st, p, s.N, s.rigid_membership, pos, force,
                           noise_beads);
    return hipGetLastError();
}

}  // namespace cr
