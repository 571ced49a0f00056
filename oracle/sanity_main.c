/*
 * sanity_main.c -- drives every CPU oracle entry point on a small system, for the
 * AddressSanitizer / UndefinedBehaviorSanitizer build (make -C oracle sanitize).
 * TEST INFRASTRUCTURE ONLY (the oracle is the checker, never the product).
 *
 * System: 4 rigid bodies of 3 touching beads plus 2 isolated beads in a 14^3 box,
 * E0 = (0, 0, 1), net charges +1, -1, +1, -1 and 0, 0.  Exits non-zero on a failed
 * check; the sanitizers abort on the first error they find.
 */
#include "cr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NB 14
#define NR 4
#define NI 6  /* rigid bodies + isolated beads */

static int fail(const char* what)
{
    fprintf(stderr, "sanity: %s\n", what);
    return 1;
}

int main(void)
{
    const double L[3] = {14.0, 14.0, 14.0};
    double pos[4 * NB];
    unsigned tag[NB], body_of_tag[NB], membership[NB], members[NR * 3], size[NR];
    double com[4 * NR], bq[NI] = {1.0, -1.0, 1.0, -1.0, 0.0, 0.0};
    for (int r = 0; r < NR; ++r) {
        const double c[3] = {-4.0 + 5.0 * (r % 2), -4.0 + 5.0 * (r / 2), -1.0 + 0.5 * r};
        for (int k = 0; k < 3; ++k) {
            const int i = 3 * r + k;
            pos[4 * i + 0] = c[0] + 2.0 * (k - 1);
            pos[4 * i + 1] = c[1];
            pos[4 * i + 2] = c[2];
            pos[4 * i + 3] = 0.0;
            membership[i] = (unsigned)r;
            members[3 * r + k] = (unsigned)i;
        }
        size[r] = 3;
        com[4 * r + 0] = c[0];
        com[4 * r + 1] = c[1];
        com[4 * r + 2] = c[2];
        com[4 * r + 3] = 0.0;
    }
    const double iso[2][3] = {{3.0, 3.0, 4.5}, {-3.5, 2.5, -4.0}};
    for (int k = 0; k < 2; ++k) {
        const int i = 3 * NR + k;
        for (int d = 0; d < 3; ++d) pos[4 * i + d] = iso[k][d];
        pos[4 * i + 3] = 0.0;
        membership[i] = CRO_NOT_MEMBER;
    }
    for (int i = 0; i < NB; ++i) {
        tag[i] = (unsigned)i;
        body_of_tag[i] = i < 3 * NR ? (unsigned)(i / 3) : (unsigned)(NR + i - 3 * NR);
    }

    unsigned head[NB], nn[NB];
    const long pairs = cro_build_nlist(NB, pos, L, 4.0, head, nn, NULL, 0);
    unsigned* nl = (unsigned*)malloc(sizeof(unsigned) * (size_t)(pairs > 0 ? pairs : 1));
    if (cro_build_nlist(NB, pos, L, 4.0, head, nn, nl, pairs) != pairs) return fail("nlist");

    cro_system s = {NB, pos, tag, body_of_tag, bq, NR, NI, 3, membership, members, size, com, nl, head, nn};
    cro_params p;
    const int mesh[3] = {16, 18, 20};
    cro_set_params(&p, L, 0.9, 1e-5, 3.6, mesh, 6);
    if (p.N[0] != 16 || p.N[2] != 20 || p.P != 6) return fail("params");

    int sm[200];
    if (cro_smooth_list(sm, 200) != 131) return fail("5-smooth list");
    double* tab = (double*)malloc(sizeof(double) * 2 * (size_t)(p.N_table + 1));
    double* gtab = (double*)malloc(sizeof(double) * 2 * (size_t)(p.N_table + 1));
    cro_tables(&p, tab, gtab);
    if (!isfinite(cro_table_lookup(&p, tab, 2.5)) || !isfinite(cro_phiq_exact(1.0, 0.9)) ||
        !isfinite(cro_gradphiq_exact(3.0, 0.9)))
        return fail("tables");

    const long G = (long)p.N[0] * p.N[1] * p.N[2];
    double* S = (double*)malloc(sizeof(double) * (size_t)G);
    double* grid = (double*)calloc((size_t)G, sizeof(double));
    double* re = (double*)calloc((size_t)G, sizeof(double));
    double* im = (double*)calloc((size_t)G, sizeof(double));
    double* ore = (double*)calloc((size_t)G, sizeof(double));
    double* oim = (double*)calloc((size_t)G, sizeof(double));
    cro_kscale(&p, S);
    double q[NB], phi[NB], F[4 * NB];
    for (int i = 0; i < NB; ++i) q[i] = 0.1 * (i % 5) - 0.2;
    cro_spread(&p, NB, pos, q, grid);
    const int small[3] = {4, 6, 5};
    for (int k = 0; k < 120; ++k) {
        re[k] = sin(0.3 * k);
        im[k] = cos(0.7 * k);
    }
    cro_dft3d_brute(small, re, im, ore, oim, -1);
    cro_fft3d(small, re, im, -1);
    for (int k = 0; k < 120; ++k)
        if (fabs(re[k] - ore[k]) > 1e-9 || fabs(im[k] - oim[k]) > 1e-9) return fail("fft vs dft");
    cro_recip_grid(&p, S, grid);
    cro_contract(&p, NB, pos, grid, phi);
    cro_contract_force(&p, NB, pos, q, grid, F);

    unsigned isom[NB];
    cro_membership(&s, isom);
    double Phi[NI] = {0}, sums[NI];
    cro_real_space(&p, &s, tab, isom, q, Phi, phi);
    const double E0[3] = {0.0, 0.0, 1.0};
    cro_real_space_force(&p, &s, gtab, q, E0, F);
    cro_sum_charges(&s, isom, q, sums);
    const int n = NB + NI;
    double b[NB + NI], x[NB + NI], y[NB + NI];
    cro_objective(&s, L, E0, b);
    cro_precondition(&s, isom, b, y);

    cro_ctx* c = cro_create(&p, &s);
    for (int i = 0; i < n; ++i) x[i] = 0.01 * (i % 7);
    cro_matvec(c, x, y);
    cro_apply_precond(c, y, x);
    int iters = 0;
    double resid = 0.0;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    if (!cro_gmres(c, x, b, 10, 500, 1e-10, &iters, &resid)) return fail("gmres");
    double qq[NB] = {0}, PP[NI] = {0}, FF[4 * NB];
    if (!cro_compute_forces(c, E0, qq, PP, FF, 10, 500, 1e-10, &iters, &resid)) return fail("compute_forces");
    for (int i = 0; i < 4 * NB; ++i)
        if (!isfinite(FF[i])) return fail("forces");
    cro_direct_matvec(&p, &s, 6.0, 3.0, qq, phi);
    cro_direct_force(&p, &s, 6.0, 3.0, qq, E0, F);
    cro_destroy(c);
    printf("sanity: ok (%ld pairs, %d GMRES iterations, %d threads)\n", pairs, iters, cro_num_threads());
    free(nl);
    free(tab);
    free(gtab);
    free(S);
    free(grid);
    free(re);
    free(im);
    free(ore);
    free(oim);
    return 0;
}
