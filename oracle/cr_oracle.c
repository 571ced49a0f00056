/*
 * cr_oracle.c -- CPU FP64 restatement of ConductorRigid's force compute.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Nothing in the product path links or calls this file.
 *
 * Reference (zeesherman/ConductorRigid, FP32-only CUDA/HOOMD plugin):
 *   SetParams              cppmodule/ConductorRigid.cc:116-379
 *   spread                 cppmodule/ConductorRigid.cu:144-234
 *   scale                  cppmodule/ConductorRigid.cu:237-257
 *   contract               cppmodule/ConductorRigid.cu:260-360
 *   contract_force         cppmodule/ConductorRigid.cu:363-466
 *   real_space             cppmodule/ConductorRigid.cu:469-567
 *   real_space_force       cppmodule/ConductorRigid.cu:570-664
 *   sum_charges/set_charge cppmodule/ConductorRigid.cu:667-761
 *   compute_objective      cppmodule/ConductorRigid.cu:764-813
 *   precondition           cppmodule/ConductorRigid.cu:816-859
 *   matvec assembly        cppmodule/ConductorRigid.cu:862-1020
 *   solve / force pass     cppmodule/ConductorRigid.cu:1051-1286
 *   GMRES                  CUSP krylov::gmres (header-only, not vendored,
 *                          version unpinned, era 0.4/0.5): restarted, left
 *                          preconditioned, modified Gram-Schmidt Arnoldi,
 *                          Givens rotations; default_monitor stops when the
 *                          Givens residual estimate <= tol * ||b||_2.
 *
 * Deliberate, documented differences from the FP32 reference:
 *   - all arithmetic is double (the reference only builds in single
 *     precision);
 *   - cuFFT's C2C transform is replaced by this file's own mixed-radix
 *     complex FFT (same unnormalised sign conventions);
 *   - the preconditioner's rigid-row write race is resolved as "rigid body
 *     rows divide by the body size", which equals the reference whenever a
 *     rigid body's isolated tag equals its rigid index;
 *   - the real-space table index is clamped to [0, N_table-1] (the reference
 *     can only leave that range through rounding at r = dr or r -> rc).
 */
#include "cr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

int cro_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Parameters: ConductorRigid.cc:116-207                               */
/* ------------------------------------------------------------------ */

static int cmp_int(const void* a, const void* b)
{
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* 5-smooth integers in [8, 4096] in increasing order (ConductorRigid.cc:129-158). */
int cro_smooth_list(int* out, int cap)
{
    int n = 0;
    int tmp[512];
    for (int a = 0, p2 = 1; a < 13; ++a, p2 *= 2)
        for (int b = 0, p3 = 1; b < 8; ++b, p3 *= 3)
            for (int c = 0, p5 = 1; c < 6; ++c, p5 *= 5) {
                long m = (long)p2 * p3 * p5;
                if (m >= 8 && m <= 4096) tmp[n++] = (int)m;
            }
    qsort(tmp, n, sizeof(int), cmp_int);
    for (int i = 0; i < n && i < cap; ++i) out[i] = tmp[i];
    return n;
}

void cro_set_params(cro_params* p, const double L[3], double xi, double errortol,
                    double rc_override, const int mesh_override[3], int P_override)
{
    memset(p, 0, sizeof(*p));
    for (int d = 0; d < 3; ++d) p->L[d] = L[d];
    p->xi = xi;
    p->errortol = errortol;
    /* ConductorRigid.cc:123-124 */
    p->rc = (rc_override > 0) ? rc_override : sqrt(-log(errortol)) / xi;
    p->kcut = 2.0 * xi * sqrt(-log(errortol));
    int smooth[256];
    int ns = cro_smooth_list(smooth, 256);
    for (int d = 0; d < 3; ++d) {
        if (mesh_override && mesh_override[d] > 0) {
            p->N[d] = mesh_override[d];
        } else {
            /* ConductorRigid.cc:125-178: ceil(1 + L kcut / pi), then next 5-smooth */
            int m = (int)ceil(1.0 + L[d] * p->kcut / PI);
            for (int i = 0; i < ns; ++i)
                if (m <= smooth[i]) { m = smooth[i]; break; }
            p->N[d] = m;
        }
        p->h[d] = L[d] / p->N[d];                        /* :200 */
    }
    int P = (P_override > 0) ? P_override : (int)ceil(-2.0 * log(errortol) / PI);  /* :201 */
    for (int d = 0; d < 3; ++d)
        p->eta[d] = P * p->h[d] * p->h[d] * xi * xi / PI; /* :202, before the clamp */
    if (P > p->N[0]) P = p->N[0];                      /* :205-207 */
    if (P > p->N[1]) P = p->N[1];
    if (P > p->N[2]) P = p->N[2];
    p->P = P;
    /* tables: ConductorRigid.cc:278-279 */
    p->drtable = 0.001;
    p->N_table = (int)(p->rc / p->drtable - 1);
    /* kernel constants: ConductorRigid.cu:913-917 */
    double xi2 = xi * xi;
    p->xiterm = 2.0 * xi2;
    p->prefac = p->xiterm * xi / PI * sqrt(2.0 / (PI * p->eta[0] * p->eta[1] * p->eta[2]));
    p->selfterm = (1.0 - exp(-4.0 * xi2)) / (8.0 * PI * sqrt(PI) * xi) + erfc(2.0 * xi) / (4.0 * PI);
    p->quadW = p->h[0] * p->h[1] * p->h[2];
}

/* ------------------------------------------------------------------ */
/* Real-space tables: ConductorRigid.cc:276-354                        */
/* ------------------------------------------------------------------ */

double cro_phiq_exact(double dist, double xi)
{
    double xi2 = xi * xi, dist2 = dist * dist;
    double expp = exp(-(dist + 2) * (dist + 2) * xi2);
    double expm = exp(-(dist - 2) * (dist - 2) * xi2);
    double exp0 = exp(-dist2 * xi2);
    double erfp = erfc((dist + 2) * xi);
    double erfm = erfc((dist - 2) * xi);
    double erf0 = erfc(dist * xi);
    double exppolyp = -(dist + 2.) / (32. * pow(PI, 1.5) * xi * dist);
    double exppolym = -(dist - 2.) / (32. * pow(PI, 1.5) * xi * dist);
    double exppoly0 = 1. / (16. * pow(PI, 1.5) * xi);
    double erfpolyp = 1. / (64. * PI * xi2 * dist) * (2. * xi2 * (dist + 2.) * (dist + 2.) + 1.);
    double erfpolym = 1. / (64. * PI * xi2 * dist) * (2. * xi2 * (dist - 2.) * (dist - 2.) + 1.);
    double erfpoly0 = -(2. * xi2 * dist2 + 1.) / (32. * PI * xi2 * dist);
    double regpoly = (dist < 2) ? (-1. / (4. * PI * dist) + (4. - dist) / (16. * PI)) : 0.;
    return exppolyp * expp + exppolym * expm + exppoly0 * exp0 + erfpolyp * erfp +
           erfpolym * erfm + erfpoly0 * erf0 + regpoly;
}

double cro_gradphiq_exact(double dist, double xi)
{
    double xi2 = xi * xi, dist2 = dist * dist;
    double expp = exp(-(dist + 2) * (dist + 2) * xi2);
    double expm = exp(-(dist - 2) * (dist - 2) * xi2);
    double exp0 = exp(-dist2 * xi2);
    double erfp = erfc((dist + 2) * xi);
    double erfm = erfc((dist - 2) * xi);
    double erf0 = erfc(dist * xi);
    double exppolyp = -(dist - 2.) / (32. * pow(PI, 1.5) * xi * dist2);
    double exppolym = -(dist + 2.) / (32. * pow(PI, 1.5) * xi * dist2);
    double exppoly0 = 1. / (16. * pow(PI, 1.5) * xi * dist);
    double erfpolyp = (2. * xi2 * dist2 - 8. * xi2 - 1.) / (64. * PI * xi2 * dist2);
    double erfpolym = (2. * xi2 * dist2 - 8. * xi2 - 1.) / (64. * PI * xi2 * dist2);
    double erfpoly0 = -(2. * xi2 * dist2 - 1.) / (32. * PI * xi2 * dist2);
    double regpoly = (dist < 2) ? (1. / (4. * PI * dist2) - 1. / (16. * PI)) : 0.;
    return exppolyp * expp + exppolym * expm + exppoly0 * exp0 + erfpolyp * erfp +
           erfpolym * erfm + erfpoly0 * erf0 + regpoly;
}

/* Tables are Scalar2 arrays: entry i holds (value at (i+1) dr, value at (i+2) dr);
 * the last entry's second field stays 0 as in a freshly zeroed GPUArray. */
void cro_tables(const cro_params* p, double* phiq, double* gradphiq)
{
    int nt = p->N_table;
    for (int i = 0; i <= nt; ++i) {
        double dist = (i + 1) * p->drtable;
        phiq[2 * i] = cro_phiq_exact(dist, p->xi);
        gradphiq[2 * i] = cro_gradphiq_exact(dist, p->xi);
        phiq[2 * i + 1] = 0.0;
        gradphiq[2 * i + 1] = 0.0;
    }
    for (int i = 0; i < nt; ++i) {
        phiq[2 * i + 1] = phiq[2 * (i + 1)];
        gradphiq[2 * i + 1] = gradphiq[2 * (i + 1)];
    }
}

/* Table interpolation as in ConductorRigid.cu:549-554 (index clamped). */
double cro_table_lookup(const cro_params* p, const double* table, double dist)
{
    double dr = p->drtable;
    int ti = (int)floor(p->N_table * (dist - dr) / (p->rc - dr));
    if (ti < 0) ti = 0;
    if (ti > p->N_table - 1) ti = p->N_table - 1;
    double lin = dist / dr - ti - 1.0;
    double a = table[2 * ti], b = table[2 * ti + 1];
    return a + (b - a) * lin;
}

/* ------------------------------------------------------------------ */
/* k-space scale: ConductorRigid.cc:245-274                            */
/* ------------------------------------------------------------------ */

static double wavenumber(int i, int n, double L)
{
    return ((i < (n + 1) / 2) ? i : i - n) * 2.0 * PI / L;
}

void cro_kscale(const cro_params* p, double* S)
{
    const int Nx = p->N[0], Ny = p->N[1], Nz = p->N[2];
    const double G = (double)Nx * Ny * Nz;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                long idx = ((long)i * Ny + j) * Nz + k;
                double kx = wavenumber(i, Nx, p->L[0]);
                double ky = wavenumber(j, Ny, p->L[1]);
                double kz = wavenumber(k, Nz, p->L[2]);
                double k2 = kx * kx + ky * ky + kz * kz;
                double etak2 = (1.0 - p->eta[0]) * kx * kx + (1.0 - p->eta[1]) * ky * ky +
                               (1.0 - p->eta[2]) * kz * kz;
                if (i == 0 && j == 0 && k == 0) {
                    S[idx] = 0.0;
                } else {
                    double kk = sqrt(k2);
                    double j0 = sin(kk) / kk;  /* boost sph_bessel(0, k) */
                    S[idx] = j0 * j0 * exp(-etak2 / (4.0 * p->xi * p->xi)) / k2 / G;
                }
            }
}

/* ------------------------------------------------------------------ */
/* Spreading / contraction: ConductorRigid.cu:144-466                  */
/* ------------------------------------------------------------------ */

/* First window node and unwrapped-node offset rule of ConductorRigid.cu:193-210. */
static void window_start(const cro_params* p, const double* r, int start[3])
{
    int halfP = p->P / 2;
    for (int d = 0; d < 3; ++d) {
        double frac = (r[d] - (-0.5 * p->L[d])) / p->L[d];   /* BoxDim::makeFraction */
        double pf = frac * (double)p->N[d];
        int x = (int)pf;
        start[d] = x - halfP + 1 - (p->P % 2) * ((pf - (double)x) < 0.5);
    }
}

static inline int wrap(int x, int n)
{
    return (x < 0) ? x + n : ((x > n - 1) ? x - n : x);
}

/* Gaussian spread of bead charges onto the real grid (zeroed here).
 * Deterministic parallel scatter: beads are binned into x-slabs at least P
 * nodes wide and even/odd slabs are processed in two phases. */
void cro_spread(const cro_params* p, int N, const double* pos, const double* q, double* grid)
{
    const int Nx = p->N[0], Ny = p->N[1], Nz = p->N[2], P = p->P;
    const long G = (long)Nx * Ny * Nz;
    memset(grid, 0, sizeof(double) * G);
    int nslab = Nx / (P > 0 ? P : 1);
    if (nslab % 2) nslab -= 1;
    int* slab = (int*)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (int b = 0; b < N; ++b) {
        int st[3];
        window_start(p, &pos[4 * b], st);
        int x0 = wrap(st[0], Nx);
        slab[b] = (nslab >= 2) ? (int)((long)x0 * nslab / Nx) : 0;
    }
    const double halfL[3] = {0.5 * p->L[0], 0.5 * p->L[1], 0.5 * p->L[2]};
    int nphase = (nslab >= 2) ? 2 : 1;
    int nwork = (nslab >= 2) ? nslab : 1;
    for (int phase = 0; phase < nphase; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int s = phase; s < nwork; s += nphase) {
            for (int b = 0; b < N; ++b) {
                if (slab[b] != s) continue;
                const double* r = &pos[4 * b];
                int st[3];
                window_start(p, r, st);
                for (int tx = 0; tx < P; ++tx)
                    for (int ty = 0; ty < P; ++ty)
                        for (int tz = 0; tz < P; ++tz) {
                            int xi_ = st[0] + tx, yi_ = st[1] + ty, zi_ = st[2] + tz;
                            double gx = p->h[0] * xi_ - halfL[0];
                            double gy = p->h[1] * yi_ - halfL[1];
                            double gz = p->h[2] * zi_ - halfL[2];
                            double rx = gx - r[0], ry = gy - r[1], rz = gz - r[2];
                            double r2eta = rx * rx / p->eta[0] + ry * ry / p->eta[1] +
                                           rz * rz / p->eta[2];
                            double w = p->prefac * exp(-p->xiterm * r2eta) * q[b];
                            long gi = ((long)wrap(xi_, Nx) * Ny + wrap(yi_, Ny)) * Nz + wrap(zi_, Nz);
                            grid[gi] += w;
                        }
            }
        }
    }
    free(slab);
}

void cro_contract(const cro_params* p, int N, const double* pos, const double* grid, double* phi)
{
    const int Nx = p->N[0], Ny = p->N[1], Nz = p->N[2], P = p->P;
    const double halfL[3] = {0.5 * p->L[0], 0.5 * p->L[1], 0.5 * p->L[2]};
    const double pc = p->quadW * p->prefac;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < N; ++b) {
        const double* r = &pos[4 * b];
        int st[3];
        window_start(p, r, st);
        double acc = 0.0;
        for (int tx = 0; tx < P; ++tx)
            for (int ty = 0; ty < P; ++ty)
                for (int tz = 0; tz < P; ++tz) {
                    int xi_ = st[0] + tx, yi_ = st[1] + ty, zi_ = st[2] + tz;
                    double rx = r[0] - (p->h[0] * xi_ - halfL[0]);
                    double ry = r[1] - (p->h[1] * yi_ - halfL[1]);
                    double rz = r[2] - (p->h[2] * zi_ - halfL[2]);
                    double r2eta = rx * rx / p->eta[0] + ry * ry / p->eta[1] + rz * rz / p->eta[2];
                    long gi = ((long)wrap(xi_, Nx) * Ny + wrap(yi_, Ny)) * Nz + wrap(zi_, Nz);
                    acc += pc * exp(-p->xiterm * r2eta) * grid[gi];
                }
        phi[b] = acc;
    }
}

void cro_contract_force(const cro_params* p, int N, const double* pos, const double* q,
                        const double* grid, double* force4)
{
    const int Nx = p->N[0], Ny = p->N[1], Nz = p->N[2], P = p->P;
    const double halfL[3] = {0.5 * p->L[0], 0.5 * p->L[1], 0.5 * p->L[2]};
    const double pc = p->quadW * p->prefac;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < N; ++b) {
        const double* r = &pos[4 * b];
        int st[3];
        window_start(p, r, st);
        double f[3] = {0, 0, 0};
        for (int tx = 0; tx < P; ++tx)
            for (int ty = 0; ty < P; ++ty)
                for (int tz = 0; tz < P; ++tz) {
                    int xi_ = st[0] + tx, yi_ = st[1] + ty, zi_ = st[2] + tz;
                    double rx = r[0] - (p->h[0] * xi_ - halfL[0]);
                    double ry = r[1] - (p->h[1] * yi_ - halfL[1]);
                    double rz = r[2] - (p->h[2] * zi_ - halfL[2]);
                    double ex = rx / p->eta[0], ey = ry / p->eta[1], ez = rz / p->eta[2];
                    double r2eta = rx * ex + ry * ey + rz * ez;
                    long gi = ((long)wrap(xi_, Nx) * Ny + wrap(yi_, Ny)) * Nz + wrap(zi_, Nz);
                    double c = 2.0 * pc * p->xiterm * q[b] * grid[gi] * exp(-p->xiterm * r2eta);
                    f[0] += c * ex;
                    f[1] += c * ey;
                    f[2] += c * ez;
                }
        force4[4 * b + 0] = f[0];
        force4[4 * b + 1] = f[1];
        force4[4 * b + 2] = f[2];
        force4[4 * b + 3] = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Mixed-radix complex FFT (replaces cuFFT's unnormalised C2C).        */
/* ------------------------------------------------------------------ */

/* Recursive decimation in time: x (stride s, length n) -> X (contiguous). */
static void fft1d_rec(int n, const double* xr, const double* xi, long s, double* Xr, double* Xi,
                      int sign, const double* wr, const double* wi, int nfull)
{
    /* wr/wi hold e^{2 pi i e / nfull}; W_n^e = W_nfull^{e nfull/n} */
    const long tw = nfull / n;
    if (n == 1) { Xr[0] = xr[0]; Xi[0] = xi[0]; return; }
    int p = 0;
    for (int f = 2; f * f <= n; ++f) if (n % f == 0) { p = f; break; }
    if (!p) p = n;
    int m = n / p;
    for (int r = 0; r < p; ++r)
        fft1d_rec(m, xr + r * s, xi + r * s, s * p, Xr + (long)r * m, Xi + (long)r * m, sign, wr, wi, nfull);
    /* combine: X[k + t m] = sum_r W_n^{r (k + t m)} Y_r[k] */
    double yr[64], yi[64];
    double* tr = (p <= 64) ? yr : (double*)malloc(sizeof(double) * p);
    double* ti = (p <= 64) ? yi : (double*)malloc(sizeof(double) * p);
    double* outr = (double*)malloc(sizeof(double) * n);
    double* outi = (double*)malloc(sizeof(double) * n);
    for (int k = 0; k < m; ++k) {
        for (int r = 0; r < p; ++r) { tr[r] = Xr[(long)r * m + k]; ti[r] = Xi[(long)r * m + k]; }
        for (int t = 0; t < p; ++t) {
            int kk = k + t * m;
            double ar = 0, ai = 0;
            for (int r = 0; r < p; ++r) {
                long e = (((long)r * kk) % n) * tw;
                double c = wr[e], sn = sign * wi[e];
                ar += tr[r] * c - ti[r] * sn;
                ai += tr[r] * sn + ti[r] * c;
            }
            outr[kk] = ar;
            outi[kk] = ai;
        }
    }
    memcpy(Xr, outr, sizeof(double) * n);
    memcpy(Xi, outi, sizeof(double) * n);
    free(outr);
    free(outi);
    if (tr != yr) { free(tr); free(ti); }
}

/* 1-D FFTs along one axis of a row-major (z fastest) complex 3-D array. */
static void fft_axis(const int n[3], double* re, double* im, int axis, int sign)
{
    int len = n[axis];
    long stride = (axis == 2) ? 1 : (axis == 1 ? n[2] : (long)n[1] * n[2]);
    long count = (long)n[0] * n[1] * n[2] / len;
    double* wr = (double*)malloc(sizeof(double) * len);
    double* wi = (double*)malloc(sizeof(double) * len);
    for (int e = 0; e < len; ++e) {
        wr[e] = cos(2.0 * PI * e / len);
        wi[e] = sin(2.0 * PI * e / len);
    }
#pragma omp parallel
    {
        double* br = (double*)malloc(sizeof(double) * len);
        double* bi = (double*)malloc(sizeof(double) * len);
        double* Xr = (double*)malloc(sizeof(double) * len);
        double* Xi = (double*)malloc(sizeof(double) * len);
#pragma omp for schedule(static)
        for (long c = 0; c < count; ++c) {
            long base;
            if (axis == 2) base = c * len;
            else if (axis == 1) base = (c / n[2]) * ((long)n[1] * n[2]) + (c % n[2]);
            else base = c;
            for (int t = 0; t < len; ++t) { br[t] = re[base + t * stride]; bi[t] = im[base + t * stride]; }
            fft1d_rec(len, br, bi, 1, Xr, Xi, sign, wr, wi, len);
            for (int t = 0; t < len; ++t) { re[base + t * stride] = Xr[t]; im[base + t * stride] = Xi[t]; }
        }
        free(br); free(bi); free(Xr); free(Xi);
    }
    free(wr);
    free(wi);
}

/* In-place unnormalised 3-D DFT, sign -1 forward (e^{-i k x}), +1 inverse. */
void cro_fft3d(const int n[3], double* re, double* im, int sign)
{
    fft_axis(n, re, im, 2, sign);
    fft_axis(n, re, im, 1, sign);
    fft_axis(n, re, im, 0, sign);
}

void cro_dft3d_brute(const int n[3], const double* re, const double* im, double* ore,
                     double* oim, int sign)
{
    long G = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long o = 0; o < G; ++o) {
        int ka = (int)(o / ((long)n[1] * n[2])), kb = (int)((o / n[2]) % n[1]), kc = (int)(o % n[2]);
        double ar = 0, ai = 0;
        for (long g = 0; g < G; ++g) {
            int a = (int)(g / ((long)n[1] * n[2])), b = (int)((g / n[2]) % n[1]), c = (int)(g % n[2]);
            double ph = 2.0 * PI * ((double)((long)ka * a % n[0]) / n[0] +
                                    (double)((long)kb * b % n[1]) / n[1] +
                                    (double)((long)kc * c % n[2]) / n[2]);
            double cs = cos(ph), sn = sign * sin(ph);
            ar += re[g] * cs - im[g] * sn;
            ai += re[g] * sn + im[g] * cs;
        }
        ore[o] = ar;
        oim[o] = ai;
    }
}

void cro_scale(const cro_params* p, const double* S, double* re, double* im)
{
    long G = (long)p->N[0] * p->N[1] * p->N[2];
#pragma omp parallel for schedule(static)
    for (long g = 0; g < G; ++g) { re[g] *= S[g]; im[g] *= S[g]; }
}

/* grid <- Re IFFT( S * FFT(grid) ), as ConductorRigid.cu:927-933. */
void cro_recip_grid(const cro_params* p, const double* S, double* grid)
{
    long G = (long)p->N[0] * p->N[1] * p->N[2];
    double* im = (double*)calloc(G, sizeof(double));
    cro_fft3d(p->N, grid, im, -1);
    cro_scale(p, S, grid, im);
    cro_fft3d(p->N, grid, im, +1);
    free(im);
}

/* ------------------------------------------------------------------ */
/* Real space, constraints, objective, preconditioner                  */
/* ------------------------------------------------------------------ */

static inline void min_image(const double L[3], double r[3])
{
    for (int d = 0; d < 3; ++d) r[d] -= L[d] * rint(r[d] / L[d]);
}

/* update_membership (ConductorRigid.cu:117-141): body tag per global index. */
void cro_membership(const cro_system* s, unsigned* iso)
{
    for (int i = 0; i < s->N; ++i) iso[i] = s->body_of_tag[s->tag[i]];
}

/* real_space (ConductorRigid.cu:469-567):
 * out_i = Phi_{b(i)} - (phiq_i + self q_i + sum_j A(r_ij) q_j). */
void cro_real_space(const cro_params* p, const cro_system* s, const double* table,
                    const unsigned* iso, const double* q, const double* Phi, double* phiq)
{
    const double L[3] = {p->L[0], p->L[1], p->L[2]};
    const double rc2 = p->rc * p->rc, rmin2 = p->drtable * p->drtable;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->N; ++i) {
        double acc = phiq[i] + p->selfterm * q[i];
        const double* ri = &s->pos[4 * i];
        unsigned head = s->head_list[i], nn = s->n_neigh[i];
        for (unsigned jj = 0; jj < nn; ++jj) {
            unsigned j = s->nlist[head + jj];
            double r[3] = {ri[0] - s->pos[4 * j], ri[1] - s->pos[4 * j + 1], ri[2] - s->pos[4 * j + 2]};
            min_image(L, r);
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (d2 < rc2 && d2 >= rmin2) acc += cro_table_lookup(p, table, sqrt(d2)) * q[j];
        }
        phiq[i] = Phi[iso[i]] - acc;
    }
}

/* real_space_force (ConductorRigid.cu:570-664): F_i += q_i E0 - sum_j A'(r) q_i q_j rhat. */
void cro_real_space_force(const cro_params* p, const cro_system* s, const double* gtable,
                          const double* q, const double E0[3], double* f4)
{
    const double L[3] = {p->L[0], p->L[1], p->L[2]};
    const double rc2 = p->rc * p->rc, rmin2 = p->drtable * p->drtable;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->N; ++i) {
        double F[3] = {f4[4 * i] + q[i] * E0[0], f4[4 * i + 1] + q[i] * E0[1], f4[4 * i + 2] + q[i] * E0[2]};
        const double* ri = &s->pos[4 * i];
        unsigned head = s->head_list[i], nn = s->n_neigh[i];
        for (unsigned jj = 0; jj < nn; ++jj) {
            unsigned j = s->nlist[head + jj];
            double r[3] = {ri[0] - s->pos[4 * j], ri[1] - s->pos[4 * j + 1], ri[2] - s->pos[4 * j + 2]};
            min_image(L, r);
            double d2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (d2 < rc2 && d2 >= rmin2) {
                double d = sqrt(d2);
                double A = cro_table_lookup(p, gtable, d);
                for (int k = 0; k < 3; ++k) F[k] += -A * q[i] * q[j] * (r[k] / d);
            }
        }
        f4[4 * i] = F[0];
        f4[4 * i + 1] = F[1];
        f4[4 * i + 2] = F[2];
        f4[4 * i + 3] = 0.0;
    }
}

/* sum_charges + set_charge (ConductorRigid.cu:667-761). */
void cro_sum_charges(const cro_system* s, const unsigned* iso, const double* q, double* out)
{
    for (int r = 0; r < s->N_rigid; ++r) {
        unsigned size = s->rigid_size[r];
        if (size == 0) continue;
        unsigned first = s->rigid_members[(long)r * s->Nmax];
        double acc = 0.0;
        for (unsigned k = 0; k < size; ++k) acc += q[s->rigid_members[(long)r * s->Nmax + k]];
        out[iso[first]] = acc;
    }
    for (int i = 0; i < s->N; ++i)
        if (s->rigid_membership[i] == CRO_NOT_MEMBER) out[iso[i]] = q[i];
}

/* zero_array + compute_objective (ConductorRigid.cu:764-813, 1099-1107). */
void cro_objective(const cro_system* s, const double Lb[3], const double E0[3], double* b)
{
    const int N = s->N;
    memset(b, 0, sizeof(double) * (N + s->N_isolated));
    for (long t = 0; t < (long)s->N_rigid * s->Nmax; ++t) {
        unsigned idx = s->rigid_members[t];
        if (idx == CRO_NOT_MEMBER) continue;
        const double* c = &s->rigid_com[4 * (t / s->Nmax)];
        double r[3] = {s->pos[4 * idx] - c[0], s->pos[4 * idx + 1] - c[1], s->pos[4 * idx + 2] - c[2]};
        min_image(Lb, r);
        b[idx] = -(r[0] * E0[0] + r[1] * E0[1] + r[2] * E0[2]);
    }
    for (int t = 0; t < s->N_isolated; ++t) b[N + t] = s->body_charge[t];
}

/* precondition (ConductorRigid.cu:816-859), rigid-row race resolved as documented above. */
void cro_precondition(const cro_system* s, const unsigned* iso, const double* x, double* y)
{
    const int N = s->N;
    for (int i = 0; i < N; ++i) y[i] = -4.0 * PI * x[i];
    for (int t = 0; t < s->N_isolated; ++t) y[N + t] = x[N + t] / (4.0 * PI);
    for (int r = 0; r < s->N_rigid; ++r) {
        unsigned size = s->rigid_size[r];
        if (size == 0) continue;
        unsigned t = iso[s->rigid_members[(long)r * s->Nmax]];
        y[N + t] = x[N + t] / size;
    }
}

/* ------------------------------------------------------------------ */
/* Composite: matvec, GMRES, force evaluation                          */
/* ------------------------------------------------------------------ */

struct cro_ctx {
    cro_params p;
    cro_system s;
    unsigned* iso;
    double* S;
    double* table;
    double* gtable;
    double* grid;
    int n;
};

cro_ctx* cro_create(const cro_params* p, const cro_system* s)
{
    cro_ctx* c = (cro_ctx*)calloc(1, sizeof(cro_ctx));
    c->p = *p;
    c->s = *s;
    c->n = s->N + s->N_isolated;
    long G = (long)p->N[0] * p->N[1] * p->N[2];
    c->iso = (unsigned*)malloc(sizeof(unsigned) * (s->N > 0 ? s->N : 1));
    cro_membership(s, c->iso);
    c->S = (double*)malloc(sizeof(double) * G);
    cro_kscale(p, c->S);
    c->table = (double*)malloc(sizeof(double) * 2 * (p->N_table + 1));
    c->gtable = (double*)malloc(sizeof(double) * 2 * (p->N_table + 1));
    cro_tables(p, c->table, c->gtable);
    c->grid = (double*)malloc(sizeof(double) * G);
    return c;
}

void cro_destroy(cro_ctx* c)
{
    if (!c) return;
    free(c->iso); free(c->S); free(c->table); free(c->gtable); free(c->grid);
    free(c);
}

/* MatrixVectorMultiply (ConductorRigid.cu:953-1020) with PotentialChargeMultiply (:862-946). */
void cro_matvec(cro_ctx* c, const double* x, double* y)
{
    const int N = c->s.N;
    const double* q = x;
    const double* Phi = x + N;
    cro_spread(&c->p, N, c->s.pos, q, c->grid);
    cro_recip_grid(&c->p, c->S, c->grid);
    cro_contract(&c->p, N, c->s.pos, c->grid, y);
    cro_real_space(&c->p, &c->s, c->table, c->iso, q, Phi, y);
    cro_sum_charges(&c->s, c->iso, q, y + N);
}

void cro_apply_precond(cro_ctx* c, const double* x, double* y)
{
    cro_precondition(&c->s, c->iso, x, y);
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void axpy(int n, double a, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) y[i] += a * x[i];
}

static void apply_rot(double* dx, double* dy, double cs, double sn)
{
    double t = cs * (*dx) + sn * (*dy);
    *dy = -sn * (*dx) + cs * (*dy);
    *dx = t;
}

static void gen_rot(double dx, double dy, double* cs, double* sn)
{
    if (dy == 0.0) { *cs = 1.0; *sn = 0.0; }
    else if (fabs(dy) > fabs(dx)) { double t = dx / dy; *sn = 1.0 / sqrt(1.0 + t * t); *cs = t * (*sn); }
    else { double t = dy / dx; *cs = 1.0 / sqrt(1.0 + t * t); *sn = t * (*cs); }
}

/* Restarted, left-preconditioned GMRES(R) following CUSP krylov::gmres.
 * Returns 1 when converged (estimate <= tol*||b||), 0 at the iteration limit.
 * The preconditioner is applied out of place (into pw): the reference's precondition
 * kernel (ConductorRigid.cu:816-859) reads a rigid body's potential row after the same
 * thread has written it, so an in-place call would divide that row by 4 pi n_b instead
 * of n_b. */
int cro_gmres(cro_ctx* c, double* x, const double* b, int R, int maxiter, double tol,
              int* iters_out, double* resid_out)
{
    const int n = c->n;
    double* V = (double*)malloc(sizeof(double) * (long)n * (R + 1));
    double* w = (double*)malloc(sizeof(double) * n);
    double* pw = (double*)malloc(sizeof(double) * n);
    double* H = (double*)calloc((size_t)(R + 1) * R, sizeof(double));
    double* s = (double*)calloc(R + 1, sizeof(double));
    double* cs = (double*)calloc(R, sizeof(double));
    double* sn = (double*)calloc(R, sizeof(double));
    const double tolerance = tol * sqrt(dot(n, b, b));
    int iter = 0, converged = 0;
    double resid = 0.0;
#define HH(r, col) H[(long)(col) * (R + 1) + (r)]
    for (;;) {
        /* initial residual of this cycle: w = M (A x - b), beta = ||w|| */
        cro_matvec(c, x, pw);
        axpy(n, -1.0, b, pw);
        cro_apply_precond(c, pw, w);
        double beta = sqrt(dot(n, w, w));
        resid = beta;
        if (resid <= tolerance) { converged = 1; break; }
        if (iter >= maxiter) break;
        for (int i = 0; i < n; ++i) V[i] = -w[i] / beta;
        memset(s, 0, sizeof(double) * (R + 1));
        s[0] = beta;
        int i = -1;
        do {
            ++i;
            ++iter;
            cro_matvec(c, &V[(long)i * n], pw);
            cro_apply_precond(c, pw, w);
            for (int k = 0; k <= i; ++k) {
                HH(k, i) = dot(n, w, &V[(long)k * n]);
                axpy(n, -HH(k, i), &V[(long)k * n], w);
            }
            HH(i + 1, i) = sqrt(dot(n, w, w));
            double inv = (HH(i + 1, i) != 0.0) ? 1.0 / HH(i + 1, i) : 0.0;
            for (int t = 0; t < n; ++t) V[(long)(i + 1) * n + t] = w[t] * inv;
            for (int k = 0; k < i; ++k) apply_rot(&HH(k, i), &HH(k + 1, i), cs[k], sn[k]);
            gen_rot(HH(i, i), HH(i + 1, i), &cs[i], &sn[i]);
            apply_rot(&HH(i, i), &HH(i + 1, i), cs[i], sn[i]);
            apply_rot(&s[i], &s[i + 1], cs[i], sn[i]);
            resid = fabs(s[i + 1]);
            if (resid <= tolerance) { converged = 1; break; }
        } while (i + 1 < R && iter + 1 <= maxiter);
        /* back substitution and update x += V y */
        for (int j = i; j >= 0; --j) {
            s[j] /= HH(j, j);
            for (int k = j - 1; k >= 0; --k) s[k] -= HH(k, j) * s[j];
        }
        for (int j = 0; j <= i; ++j) axpy(n, s[j], &V[(long)j * n], x);
        if (converged || iter >= maxiter) break;
    }
#undef HH
    if (iters_out) *iters_out = iter;
    if (resid_out) *resid_out = resid;
    free(V); free(w); free(pw); free(H); free(s); free(cs); free(sn);
    return converged;
}

/* gpu_ComputeForce (ConductorRigid.cu:1153-1286): solve for (q, Phi) starting from the
 * caller's values, then F = contract_force + q E0 + real-space force. */
int cro_compute_forces(cro_ctx* c, const double E0[3], double* q, double* Phi, double* force4,
                       int restart, int maxiter, double tol, int* iters, double* resid)
{
    const int N = c->s.N, n = c->n;
    double* b = (double*)malloc(sizeof(double) * n);
    double* x = (double*)malloc(sizeof(double) * n);
    cro_objective(&c->s, c->p.L, E0, b);
    memcpy(x, q, sizeof(double) * N);
    memcpy(x + N, Phi, sizeof(double) * c->s.N_isolated);
    int conv = cro_gmres(c, x, b, restart, maxiter, tol, iters, resid);
    memcpy(q, x, sizeof(double) * N);
    memcpy(Phi, x + N, sizeof(double) * c->s.N_isolated);
    cro_spread(&c->p, N, c->s.pos, q, c->grid);
    cro_recip_grid(&c->p, c->S, c->grid);
    cro_contract_force(&c->p, N, c->s.pos, q, c->grid, force4);
    cro_real_space_force(&c->p, &c->s, c->gtable, q, E0, force4);
    free(b);
    free(x);
    return conv;
}

/* ------------------------------------------------------------------ */
/* Direct Ewald oracle (closed-form real space + explicit k-sum)       */
/* ------------------------------------------------------------------ */

typedef struct { int nx, ny, nz; double kx, ky, kz, G; } kvec;

static kvec* kvectors(const cro_params* p, double kmax, long* count)
{
    int mx = (int)ceil(kmax * p->L[0] / (2 * PI)), my = (int)ceil(kmax * p->L[1] / (2 * PI)),
        mz = (int)ceil(kmax * p->L[2] / (2 * PI));
    long cap = (long)(2 * mx + 1) * (2 * my + 1) * (mz + 1);
    kvec* ks = (kvec*)malloc(sizeof(kvec) * cap);
    long c = 0;
    const double xi = p->xi;
    for (int a = -mx; a <= mx; ++a)
        for (int b = -my; b <= my; ++b)
            for (int d = 0; d <= mz; ++d) {
                /* half space: d > 0, or d == 0 and (b > 0 or (b == 0 and a > 0)) */
                if (d == 0 && (b < 0 || (b == 0 && a <= 0))) continue;
                double kx = 2 * PI * a / p->L[0], ky = 2 * PI * b / p->L[1], kz = 2 * PI * d / p->L[2];
                double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 > kmax * kmax) continue;
                double kk = sqrt(k2), j0 = sin(kk) / kk;
                kvec v = {a, b, d, kx, ky, kz, j0 * j0 * exp(-k2 / (4 * xi * xi)) / k2};
                ks[c++] = v;
            }
    *count = c;
    return ks;
}

/* Structure factors rho(k) = sum_j q_j e^{-i k.r_j} for every half-space k. */
static void structure_factors(const cro_system* s, const kvec* ks, long nk, const double* q,
                              double* rr, double* ri)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long k = 0; k < nk; ++k) {
        double ar = 0, ai = 0;
        for (int j = 0; j < s->N; ++j) {
            const double* r = &s->pos[4 * j];
            double ph = ks[k].kx * r[0] + ks[k].ky * r[1] + ks[k].kz * r[2];
            ar += q[j] * cos(ph);
            ai -= q[j] * sin(ph);
        }
        rr[k] = ar;
        ri[k] = ai;
    }
}

/* phi_i = self q_i + sum_{j, images, r < rc_real} phi(r) q_j
 *       + (1/V) sum_{k != 0, |k| <= kmax} j0(k)^2 e^{-k^2/4xi^2}/k^2 Re[e^{i k.r_i} rho(k)] */
void cro_direct_matvec(const cro_params* p, const cro_system* s, double rc_real, double kmax,
                       const double* q, double* phi)
{
    const double V = p->L[0] * p->L[1] * p->L[2];
    long nk;
    kvec* ks = kvectors(p, kmax, &nk);
    double* rr = (double*)malloc(sizeof(double) * (nk ? nk : 1));
    double* ri = (double*)malloc(sizeof(double) * (nk ? nk : 1));
    structure_factors(s, ks, nk, q, rr, ri);
    int im[3];
    for (int d = 0; d < 3; ++d) im[d] = (int)ceil(rc_real / p->L[d]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < s->N; ++i) {
        const double* r = &s->pos[4 * i];
        double acc = p->selfterm * q[i];
        for (long k = 0; k < nk; ++k) {
            double ph = ks[k].kx * r[0] + ks[k].ky * r[1] + ks[k].kz * r[2];
            /* full-space sum = 2 x half-space real part */
            acc += 2.0 / V * ks[k].G * (cos(ph) * rr[k] - sin(ph) * ri[k]);
        }
        for (int j = 0; j < s->N; ++j) {
            double d0[3] = {r[0] - s->pos[4 * j], r[1] - s->pos[4 * j + 1], r[2] - s->pos[4 * j + 2]};
            min_image(p->L, d0);
            for (int a = -im[0]; a <= im[0]; ++a)
                for (int b = -im[1]; b <= im[1]; ++b)
                    for (int c = -im[2]; c <= im[2]; ++c) {
                        if (j == i && a == 0 && b == 0 && c == 0) continue;
                        double x = d0[0] + a * p->L[0], y = d0[1] + b * p->L[1], z = d0[2] + c * p->L[2];
                        double d2 = x * x + y * y + z * z;
                        if (d2 < rc_real * rc_real) acc += cro_phiq_exact(sqrt(d2), p->xi) * q[j];
                    }
        }
        phi[i] = acc;
    }
    free(ks); free(rr); free(ri);
}

void cro_direct_force(const cro_params* p, const cro_system* s, double rc_real, double kmax,
                      const double* q, const double E0[3], double* f4)
{
    const double V = p->L[0] * p->L[1] * p->L[2];
    long nk;
    kvec* ks = kvectors(p, kmax, &nk);
    double* rr = (double*)malloc(sizeof(double) * (nk ? nk : 1));
    double* ri = (double*)malloc(sizeof(double) * (nk ? nk : 1));
    structure_factors(s, ks, nk, q, rr, ri);
    int im[3];
    for (int d = 0; d < 3; ++d) im[d] = (int)ceil(rc_real / p->L[d]);
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < s->N; ++i) {
        const double* r = &s->pos[4 * i];
        double F[3] = {q[i] * E0[0], q[i] * E0[1], q[i] * E0[2]};
        for (long k = 0; k < nk; ++k) {
            double ph = ks[k].kx * r[0] + ks[k].ky * r[1] + ks[k].kz * r[2];
            /* F_i = (q_i/V) sum_k G k Im[e^{i k.r_i} rho(k)], full space = 2 x half */
            double imv = sin(ph) * rr[k] + cos(ph) * ri[k];
            double c = 2.0 / V * ks[k].G * q[i] * imv;
            F[0] += c * ks[k].kx;
            F[1] += c * ks[k].ky;
            F[2] += c * ks[k].kz;
        }
        for (int j = 0; j < s->N; ++j) {
            double d0[3] = {r[0] - s->pos[4 * j], r[1] - s->pos[4 * j + 1], r[2] - s->pos[4 * j + 2]};
            min_image(p->L, d0);
            for (int a = -im[0]; a <= im[0]; ++a)
                for (int b = -im[1]; b <= im[1]; ++b)
                    for (int c = -im[2]; c <= im[2]; ++c) {
                        if (j == i && a == 0 && b == 0 && c == 0) continue;
                        double x = d0[0] + a * p->L[0], y = d0[1] + b * p->L[1], z = d0[2] + c * p->L[2];
                        double d2 = x * x + y * y + z * z;
                        if (d2 < rc_real * rc_real) {
                            double d = sqrt(d2);
                            double A = cro_gradphiq_exact(d, p->xi);
                            F[0] += -A * q[i] * q[j] * x / d;
                            F[1] += -A * q[i] * q[j] * y / d;
                            F[2] += -A * q[i] * q[j] * z / d;
                        }
                    }
        }
        f4[4 * i] = F[0];
        f4[4 * i + 1] = F[1];
        f4[4 * i + 2] = F[2];
        f4[4 * i + 3] = 0.0;
    }
    free(ks); free(rr); free(ri);
}

/* ------------------------------------------------------------------ */
/* Neighbour list in HOOMD layout: full list of pairs with r < rlist    */
/* ------------------------------------------------------------------ */

long cro_build_nlist(int N, const double* pos, const double L[3], double rlist,
                     unsigned* head_list, unsigned* n_neigh, unsigned* nlist, long cap)
{
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(L[d] / rlist);
        if (nc[d] < 1) nc[d] = 1;
        if (nc[d] > 1024) nc[d] = 1024;
    }
    long ncell = (long)nc[0] * nc[1] * nc[2];
    int* cell_of = (int*)malloc(sizeof(int) * (N ? N : 1));
    int* count = (int*)calloc((size_t)(ncell > 0 ? ncell : 1) + 1, sizeof(int));
    for (int i = 0; i < N; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            double f = (pos[4 * i + d] + 0.5 * L[d]) / L[d];
            int k = (int)floor(f * nc[d]);
            c[d] = ((k % nc[d]) + nc[d]) % nc[d];
        }
        cell_of[i] = (c[0] * nc[1] + c[1]) * nc[2] + c[2];
        count[cell_of[i] + 1]++;
    }
    for (long c = 0; c < ncell; ++c) count[c + 1] += count[c];
    int* members = (int*)malloc(sizeof(int) * (N ? N : 1));
    int* fill = (int*)calloc((size_t)(ncell > 0 ? ncell : 1), sizeof(int));
    for (int i = 0; i < N; ++i) members[count[cell_of[i]] + fill[cell_of[i]]++] = i;
    /* small boxes: fall back to all pairs so no image is double counted */
    int allpairs = (nc[0] < 3 || nc[1] < 3 || nc[2] < 3);
    long total = 0;
    double r2 = rlist * rlist;
    for (int i = 0; i < N; ++i) {
        head_list[i] = (unsigned)total;
        unsigned cnt = 0;
        if (allpairs) {
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double r[3] = {pos[4 * i] - pos[4 * j], pos[4 * i + 1] - pos[4 * j + 1], pos[4 * i + 2] - pos[4 * j + 2]};
                min_image(L, r);
                if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < r2) {
                    if (total < cap) nlist[total] = (unsigned)j;
                    ++total; ++cnt;
                }
            }
        } else {
            int ci = cell_of[i];
            int cx = ci / (nc[1] * nc[2]), cy = (ci / nc[2]) % nc[1], cz = ci % nc[2];
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b)
                    for (int c = -1; c <= 1; ++c) {
                        int x = (cx + a + nc[0]) % nc[0], y = (cy + b + nc[1]) % nc[1], z = (cz + c + nc[2]) % nc[2];
                        long cell = ((long)x * nc[1] + y) * nc[2] + z;
                        for (int m = count[cell]; m < count[cell + 1]; ++m) {
                            int j = members[m];
                            if (j == i) continue;
                            double r[3] = {pos[4 * i] - pos[4 * j], pos[4 * i + 1] - pos[4 * j + 1], pos[4 * i + 2] - pos[4 * j + 2]};
                            min_image(L, r);
                            if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < r2) {
                                if (total < cap) nlist[total] = (unsigned)j;
                                ++total; ++cnt;
                            }
                        }
                    }
        }
        n_neigh[i] = cnt;
    }
    free(cell_of); free(count); free(members); free(fill);
    return total;
}
