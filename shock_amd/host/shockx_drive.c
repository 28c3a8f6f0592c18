/*
 * shockx_drive.c -- C host loop over libshockx.so with the semantics of SHOCK's
 * startSimulation step loop (src/SHOCK.c:557-666): fixed tau, one CalcRungeKutta per step,
 * checkNAN after every step, U_lastStep <- U_RK at the step end (inside the library).
 *
 * Case: 2D periodic unit square, n x n nodes, the manufactured field
 *   rho = 1 + 0.1 sin(2 pi x) sin(2 pi y), u = 0.5 + 0.1 cos(2 pi x) sin(2 pi y),
 *   v = 0.1 sin(2 pi x) cos(2 pi y), p = 1/gamma + 0.05 cos(2 pi (x + y))   (acoustic units)
 * WENO5 + viscous, RK4, gamma 1.4, Ma 0.5, Re 1000, Pr 0.72; tau from CFL 0.5.
 *
 * usage: shockx_drive [n=64] [steps=100]   prints mass and energy sums every 10 steps.
 * Build: gcc -O2 -I include shock_amd/host/shockx_drive.c -L shock_amd/lib -lshockx -lm
 */
#include "shockx.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#define CHECK(call)                                                                         \
    do {                                                                                    \
        int rc_ = (call);                                                                   \
        if (rc_ != SHX_OK) {                                                                \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, shx_last_error());         \
            exit(1);                                                                        \
        }                                                                                   \
    } while (0)

int main(int argc, char **argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 64;
    const int steps = argc > 2 ? atoi(argv[2]) : 100;
    const double gamma = 1.4, mach = 0.5, pi2 = 2.0 * M_PI;
    shx_desc d = {0};
    d.dims = 2; d.space_order = 5; d.time_order = 4;
    d.n[0] = n; d.n[1] = n; d.n[2] = 1;
    d.gamma = gamma; d.mach = mach; d.reynolds = 1000.0; d.prandtl = 0.72; d.T0 = 300.0;
    d.quirks = SHX_QUIRKS_REFERENCE;
    shx_case *c = NULL;
    CHECK(shx_create(&d, &c));
    int ng[3];
    CHECK(shx_padded_extent(c, ng));
    const int g = 3;
    const long ntot = (long)ng[0] * ng[1] * ng[2];

    /* extended coordinates: padded extent + one layer, periodic copies shifted by the period */
    const int e0 = ng[0] + 2, e1 = ng[1] + 2;
    double *xe = calloc((size_t)e0 * e1, sizeof(double)), *ye = calloc((size_t)e0 * e1, sizeof(double));
    double *ze = calloc((size_t)e0 * e1, sizeof(double));
    for (int i = 0; i < e0; i++)
        for (int j = 0; j < e1; j++) {
            int ri = i - 1 - g, rj = j - 1 - g;
            int si = (int)floor((double)ri / n), sj = (int)floor((double)rj / n);
            xe[(long)i * e1 + j] = (double)(ri - si * n) / n + si;
            ye[(long)i * e1 + j] = (double)(rj - sj * n) / n + sj;
        }
    CHECK(shx_build_metric(c, xe, ye, ze));

    double *prim[5];
    for (int m = 0; m < 5; m++) prim[m] = calloc((size_t)ntot, sizeof(double));
    const double ups = 1.0 / (gamma * mach * mach);
    for (int i = 0; i < ng[0]; i++)
        for (int j = 0; j < ng[1]; j++) {
            long q = (long)i * ng[1] + j;
            double x = xe[(long)(i + 1) * e1 + (j + 1)], y = ye[(long)(i + 1) * e1 + (j + 1)];
            double rho = 1.0 + 0.1 * sin(pi2 * x) * sin(pi2 * y);
            double u = (0.5 + 0.1 * cos(pi2 * x) * sin(pi2 * y)) / mach;
            double v = (0.1 * sin(pi2 * x) * cos(pi2 * y)) / mach;
            double p = gamma * (1.0 / gamma + 0.05 * cos(pi2 * (x + y)));
            prim[0][q] = rho; prim[1][q] = u; prim[2][q] = v; prim[3][q] = 0.0;
            prim[4][q] = 0.5 * (u * u + v * v) + ups * p / ((gamma - 1.0) * rho);
        }
    CHECK(shx_upload_state(c, (const double *const *)prim));
    double tau;
    CHECK(shx_cfl_dt(c, 0.5, &tau));
    printf("n=%d steps=%d tau=%.17g\n", n, steps, tau);

    for (int it = 1; it <= steps; it++) {
        CHECK(shx_calc_runge_kutta(c, tau));          /* CalcRungeKutta, src/SHOCK.c:565 */
        int bad = 0;
        CHECK(shx_check_nan(c, &bad));                 /* checkNAN, src/SHOCK.c:577 */
        if (bad) { fprintf(stderr, "non-finite density at %d nodes in step %d\n", bad, it); return 2; }
        if (it % 10 == 0 || it == steps) {
            CHECK(shx_download_state(c, prim));
            double mass = 0.0, energy = 0.0;
            for (int i = g; i < g + n; i++)
                for (int j = g; j < g + n; j++) {
                    long q = (long)i * ng[1] + j;
                    mass += prim[0][q];
                    energy += prim[0][q] * prim[4][q];
                }
            printf("step %4d  mass %.17g  energy %.17g\n", it, mass / (n * n), energy / (n * n));
        }
    }
    CHECK(shx_destroy(c));
    for (int m = 0; m < 5; m++) free(prim[m]);
    free(xe); free(ye); free(ze);
    return 0;
}
