/*
 * shx_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the SHOCK RHS / RK path.
 *
 * Plain-C FP64 restatement of the reference algorithm.  Every routine cites the
 * reference file:line it follows.  Operation order inside each expression follows
 * the reference so that the oracle reproduces the reference's rounding where the
 * compiler does not contract (build with -ffp-contract=off).
 *
 * Not linked into the product.  See shx_oracle.h for the pinning status.
 */
#include "shx_oracle.h"
#include <float.h>
#include <omp.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NEQ 5

long orc_idx(const orc_cfg *c, int i, int j, int k)
{
    return ((long)i * c->ng[1] + j) * c->ng[2] + k;
}

/* DefineParameters, src/Functions.c:1758-2032; ghost/index setup src/Functions.c:56-103;
 * Gamma (heat-flux factor) src/Functions.c:141-152. */
int orc_define(orc_cfg *c)
{
    int d;
    if (c->dims != 2 && c->dims != 3) return -1;
    if (c->S != 5 && c->S != 9) return -1;
    if (c->time_order != 3 && c->time_order != 4) return -1;
    for (d = 0; d < 6; d++) {
        if (c->bc[d] < ORC_BC_PERIODIC || c->bc[d] > ORC_BC_WALL_ISOTHERMAL) return -1;
        if ((c->bc[d] == ORC_BC_PERIODIC) != (c->bc[d ^ 1] == ORC_BC_PERIODIC)) return -1;
    }
    c->g = (c->S + 1) / 2;
    if (c->dims == 2) c->n[2] = 1;
    for (d = 0; d < 3; d++) c->ng[d] = c->n[d] + c->S + 1;
    if (c->dims == 2) c->ng[2] = 1;
    c->ntot = (long)c->ng[0] * c->ng[1] * c->ng[2];

    c->upsilon = 1.0 / (c->gamma * (c->mach * c->mach));
    c->psi = 1.0 / c->reynolds;
    c->suth = 100.4 / c->T0;
    c->gam_heat = 1.0 / ((c->gamma - 1.0) * pow(c->mach, 2.0) * c->reynolds * c->prandtl);
    c->eps = (c->eps_override > 0.0) ? c->eps_override : 10. * pow(DBL_MIN, 0.5);

    memset(c->rk_a, 0, sizeof c->rk_a); memset(c->rk_b, 0, sizeof c->rk_b);
    memset(c->rk_q, 0, sizeof c->rk_q); memset(c->rk_f, 0, sizeof c->rk_f);
    memset(c->rk_m, 0, sizeof c->rk_m);
    if (c->time_order == 4) {
        /* classic RK4, src/Functions.c:1961-1989 */
        double a[4] = {1., 1., 1., 1.}, q[4] = {0.5, 0.5, 1., 1. / 6.};
        double f[4] = {0., 0., 0., 1. / 6.}, m[4] = {1., 2., 2., 0.};
        memcpy(c->rk_a, a, sizeof a); memcpy(c->rk_q, q, sizeof q);
        memcpy(c->rk_f, f, sizeof f); memcpy(c->rk_m, m, sizeof m);
    } else {
        /* SSP-RK3, src/Functions.c:1990-2009 */
        double a[3] = {1., 3. / 4., 1. / 3.}, b[3] = {0., 1. / 4., 2. / 3.};
        double q[3] = {1., 1. / 4., 2. / 3.};
        memcpy(c->rk_a, a, sizeof a); memcpy(c->rk_b, b, sizeof b); memcpy(c->rk_q, q, sizeof q);
    }

    memset(c->zd_mid, 0, sizeof c->zd_mid); memset(c->zd_int, 0, sizeof c->zd_int);
    memset(c->zd_der, 0, sizeof c->zd_der);
    if (c->S == 9) {
        /* src/Functions.c:1858-1894 */
        static const double in9[10] = {1. / 1260., -23. / 2520., 127. / 2520., -473. / 2520., 1627. / 2520.,
                                       1627. / 2520., -473. / 2520., 127. / 2520., -23. / 2520., 1. / 1260.};
        static const double mi9[10] = {-3. / 9450., 351. / 75600., -2649. / 75600., 15351. / 75600., -110649. / 75600.,
                                       110649. / 75600., -15351. / 75600., 2649. / 75600., -351. / 75600., 3. / 9450.};
        static const double de9[11] = {-1. / 1260., 25. / 2520., -150. / 2520., 600. / 2520., -2100. / 2520., 0.,
                                       2100. / 2520., -600. / 2520., 150. / 2520., -25. / 2520., 1. / 1260.};
        memcpy(c->zd_int, in9, sizeof in9); memcpy(c->zd_mid, mi9, sizeof mi9); memcpy(c->zd_der, de9, sizeof de9);
    } else {
        /* src/Functions.c:1915-1942 */
        static const double in5[6] = {1. / 60., -8. / 60., 37. / 60., 37. / 60., -8. / 60., 1. / 60.};
        static const double mi5[6] = {-1. / 90., 25. / 180., -245. / 180., 245. / 180., -25. / 180., 1. / 90.};
        static const double de5[7] = {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.};
        memcpy(c->zd_int, in5, sizeof in5); memcpy(c->zd_mid, mi5, sizeof mi5); memcpy(c->zd_der, de5, sizeof de5);
    }
    return 0;
}

/* range helpers: interior [g, g+n-1]; all [0, ng-1]; 2D k range is [0,0] */
static int rs(const orc_cfg *c, int d) { return (c->dims == 2 && d == 2) ? 0 : c->g; }
static int re(const orc_cfg *c, int d) { return (c->dims == 2 && d == 2) ? 0 : c->g + c->n[d] - 1; }
static int ge(const orc_cfg *c, int d) { return c->ng[d] - 1; }
static long stride(const orc_cfg *c, int d)
{
    return d == 0 ? (long)c->ng[1] * c->ng[2] : (d == 1 ? (long)c->ng[2] : 1L);
}

/* CalcValues, src/Functions.c:2392-2481 (3D form; in 2D w = 0 and the z-metric terms are 0,
 * which makes the 3D expressions reduce exactly to the 2D ones). */
void orc_calc_values(const orc_cfg *c, double *const met[10], double *const prim[5], double *const der[8])
{
    long n;
    for (n = 0; n < c->ntot; n++) {
        double rho = prim[0][n], u = prim[1][n], v = prim[2][n], w = prim[3][n], e = prim[4][n];
        double p;
        der[4][n] = u * met[0][n] + v * met[1][n] + w * met[2][n];
        der[5][n] = u * met[3][n] + v * met[4][n] + w * met[5][n];
        der[6][n] = u * met[6][n] + v * met[7][n] + w * met[8][n];
        p = fabs(rho * (c->gamma - 1) / c->upsilon * (e - 0.5 * ((u * u) + (v * v) + (w * w))));
        der[0][n] = p;
        der[1][n] = fabs(p / rho);
        der[2][n] = sqrt(c->upsilon * c->gamma * p / rho);
        der[3][n] = ((1.0 + c->suth) * pow(p / rho, 1.5) / (p / rho + c->suth));
        if (c->mms) der[3][n] = 1.0;
    }
}

/* CalcEigenVectorsIn{Xi,Eta,Zeta}Direction, src/WENO.c:1085-1548 (identical formulas for the three
 * directions with the averaged normalized metric row nh).  L and R row-major: L[c*5+m], R[m*5+c]. */
void orc_eigvec(const double nh[3], double u, double v, double w, double rho, double cs,
                double gamma, double L[25], double R[25])
{
    double nx = nh[0], ny = nh[1], nz = nh[2];
    double s2 = sqrt(2.);
    double M = sqrt(u * u + v * v + w * w) / cs;
    double q2 = (u * u + v * v + w * w);
    double gc = ((gamma - 1.0) / (cs * cs));
    double ics = (1. / (cs * s2));
    double un = (u * nx + v * ny + w * nz);

    R[0 * 5 + 0] = nx; R[0 * 5 + 1] = ny; R[0 * 5 + 2] = nz; R[0 * 5 + 3] = ics; R[0 * 5 + 4] = ics;
    R[1 * 5 + 0] = u * nx; R[1 * 5 + 1] = u * ny - rho * nz; R[1 * 5 + 2] = u * nz + rho * ny;
    R[1 * 5 + 3] = ics * (u + nx * cs); R[1 * 5 + 4] = ics * (u - nx * cs);
    R[2 * 5 + 0] = v * nx + rho * nz; R[2 * 5 + 1] = v * ny; R[2 * 5 + 2] = v * nz - rho * nx;
    R[2 * 5 + 3] = ics * (v + ny * cs); R[2 * 5 + 4] = ics * (v - ny * cs);
    R[3 * 5 + 0] = w * nx - rho * ny; R[3 * 5 + 1] = w * ny + rho * nx; R[3 * 5 + 2] = w * nz;
    R[3 * 5 + 3] = ics * (w + nz * cs); R[3 * 5 + 4] = ics * (w - nz * cs);
    R[4 * 5 + 0] = q2 / 2. * nx + rho * (v * nz - w * ny);
    R[4 * 5 + 1] = q2 / 2. * ny + rho * (w * nx - u * nz);
    R[4 * 5 + 2] = q2 / 2. * nz + rho * (u * ny - v * nx);
    R[4 * 5 + 3] = ics * (q2 / 2. + cs * cs / (gamma - 1.) + cs * un);
    R[4 * 5 + 4] = ics * (q2 / 2. + cs * cs / (gamma - 1.) - cs * un);

    L[0 * 5 + 0] = nx * (1.0 - 0.5 * (gamma - 1.0) * (M * M)) - 1. / rho * (v * nz - w * ny);
    L[0 * 5 + 1] = gc * u * nx;
    L[0 * 5 + 2] = gc * v * nx + (nz) / rho;
    L[0 * 5 + 3] = gc * w * nx - (ny) / rho;
    L[0 * 5 + 4] = -gc * nx;
    L[1 * 5 + 0] = ny * (1.0 - 0.5 * (gamma - 1.0) * (M * M)) - 1. / rho * (w * nx - u * nz);
    L[1 * 5 + 1] = gc * u * ny - (nz) / rho;
    L[1 * 5 + 2] = gc * v * ny;
    L[1 * 5 + 3] = gc * w * ny + (nx) / rho;
    L[1 * 5 + 4] = -gc * ny;
    L[2 * 5 + 0] = nz * (1.0 - 0.5 * (gamma - 1.0) * (M * M)) - 1. / rho * (u * ny - v * nx);
    L[2 * 5 + 1] = gc * u * nz + (ny) / rho;
    L[2 * 5 + 2] = gc * v * nz - (nx) / rho;
    L[2 * 5 + 3] = gc * w * nz;
    L[2 * 5 + 4] = -gc * nz;
    L[3 * 5 + 0] = 1. / (s2 * cs) * ((gamma - 1.0) * q2 / 2. - cs * un);
    L[3 * 5 + 1] = (-(gamma - 1.0) * u / cs + nx) / s2;
    L[3 * 5 + 2] = (-(gamma - 1.0) * v / cs + ny) / s2;
    L[3 * 5 + 3] = (-(gamma - 1.0) * w / cs + nz) / s2;
    L[3 * 5 + 4] = (gamma - 1.) / (s2 * cs);
    L[4 * 5 + 0] = 1. / (s2 * cs) * ((gamma - 1.0) * q2 / 2. + cs * un);
    L[4 * 5 + 1] = (-(gamma - 1.) * u / cs - nx) / s2;
    L[4 * 5 + 2] = (-(gamma - 1.) * v / cs - ny) / s2;
    L[4 * 5 + 3] = (-(gamma - 1.) * w / cs - nz) / s2;
    L[4 * 5 + 4] = (gamma - 1.) / (s2 * cs);
}

/* Phi_Function_W5, src/WENO.c:118-204: Jiang-Shu WENO5 deviation from the central flux,
 * smoothness indicators from the flux differences d[0..3] (scaled by 12), weights 1:6:3. */
void orc_is_w5(const double *d, double is[3])
{
    is[0] = 13. * (d[0] - d[1]) * (d[0] - d[1]) + 3. * (d[0] - 3. * d[1]) * (d[0] - 3. * d[1]);
    is[1] = 13. * (d[1] - d[2]) * (d[1] - d[2]) + 3. * (d[1] + d[2]) * (d[1] + d[2]);
    is[2] = 13. * (d[2] - d[3]) * (d[2] - d[3]) + 3. * (3. * d[2] - d[3]) * (3. * d[2] - d[3]);
}

double orc_phi_w5(double eps, const double *f, const double *d)
{
    double is0, is1, is2, a0, a1, a2, w0, w2, q1, q2, isv[3];
    (void)f;
    orc_is_w5(d, isv);
    is0 = isv[0]; is1 = isv[1]; is2 = isv[2];
    a0 = 1. / ((is0 + eps) * (is0 + eps));
    a1 = 6. / ((is1 + eps) * (is1 + eps));
    a2 = 3. / ((is2 + eps) * (is2 + eps));
    w0 = a0 / (a0 + a1 + a2);
    w2 = a2 / (a0 + a1 + a2);
    q1 = d[0] - 2. * d[1] + d[2];
    q2 = d[1] - 2. * d[2] + d[3];
    return 1. / 3. * w0 * q1 + 1. / 6. * (w2 - 0.5) * q2;
}

/* Phi_Function_W9, src/WENO.c:11-116: Balsara-Shu WENO9 smoothness indicators from the point
 * fluxes f[0..8], optimal weights {1/126, 10/63, 10/21, 20/63, 5/126}, deviation from
 * 4th differences of the flux differences d[0..7]. */
void orc_is_w9(const double *f, double is[5])
{
    double is0, is1, is2, is3, is4;
    is0 = f[0] * (22658. * f[0] - 208501. * f[1] + 364863. * f[2] - 288007. * f[3] + 86329. * f[4])
        + f[1] * (482963. * f[1] - 1704396. * f[2] + 1358458. * f[3] - 411487. * f[4])
        + f[2] * (1521393. * f[2] - 2462076. * f[3] + 758823. * f[4])
        + f[3] * (1020563. * f[3] - 649501. * f[4])
        + f[4] * (107918. * f[4]);
    is1 = f[1] * (6908. * f[1] - 60871. * f[2] + 99213. * f[3] - 70237. * f[4] + 18079. * f[5])
        + f[2] * (138563. * f[2] - 464976. * f[3] + 337018. * f[4] - 88297. * f[5])
        + f[3] * (406293. * f[3] - 611976. * f[4] + 165153. * f[5])
        + f[4] * (242723. * f[4] - 140251. * f[5])
        + f[5] * (22658. * f[5]);
    is2 = f[2] * (6908. * f[2] - 51001. * f[3] + 67923. * f[4] - 38947. * f[5] + 8209. * f[6])
        + f[3] * (104963. * f[3] - 299076. * f[4] + 179098. * f[5] - 38947. * f[6])
        + f[4] * (231153. * f[4] - 299076. * f[5] + 67923. * f[6])
        + f[5] * (104963. * f[5] - 51001. * f[6])
        + f[6] * (6908. * f[6]);
    is3 = f[3] * (22658. * f[3] - 140251. * f[4] + 165153. * f[5] - 88297. * f[6] + 18079. * f[7])
        + f[4] * (242723. * f[4] - 611976. * f[5] + 337018. * f[6] - 70237. * f[7])
        + f[5] * (406293. * f[5] - 464976. * f[6] + 99213. * f[7])
        + f[6] * (138563. * f[6] - 60871. * f[7])
        + f[7] * (6908. * f[7]);
    is4 = f[4] * (107918. * f[4] - 649501. * f[5] + 758823. * f[6] - 411487. * f[7] + 86329. * f[8])
        + f[5] * (1020563. * f[5] - 2462076. * f[6] + 1358458. * f[7] - 288007. * f[8])
        + f[6] * (1521393. * f[6] - 1704396. * f[7] + 364863. * f[8])
        + f[7] * (482963. * f[7] - 208501. * f[8])
        + f[8] * (22658. * f[8]);
    is[0] = is0; is[1] = is1; is[2] = is2; is[3] = is3; is[4] = is4;
}

double orc_phi_w9(double eps, const double *f, const double *d)
{
    static const double dk[5] = {1. / 126., 10. / 63., 10. / 21., 20. / 63., 5. / 126.};
    double is0, is1, is2, is3, is4, a0, a1, a2, a3, a4, w0, w2, w3, w4, q1, q2, q3, q4, isv[5];
    orc_is_w9(f, isv);
    is0 = isv[0]; is1 = isv[1]; is2 = isv[2]; is3 = isv[3]; is4 = isv[4];
    a0 = dk[0] / ((is0 + eps) * (is0 + eps));
    a1 = dk[1] / ((is1 + eps) * (is1 + eps));
    a2 = dk[2] / ((is2 + eps) * (is2 + eps));
    a3 = dk[3] / ((is3 + eps) * (is3 + eps));
    a4 = dk[4] / ((is4 + eps) * (is4 + eps));
    w0 = a0 / (a0 + a1 + a2 + a3 + a4);
    w2 = a2 / (a0 + a1 + a2 + a3 + a4);
    w3 = a3 / (a0 + a1 + a2 + a3 + a4);
    w4 = a4 / (a0 + a1 + a2 + a3 + a4);
    q1 = d[0] - 4. * d[1] + 6. * d[2] - 4. * d[3] + d[4];
    q2 = d[1] - 4. * d[2] + 6. * d[3] - 4. * d[4] + d[5];
    q3 = d[2] - 4. * d[3] + 6. * d[4] - 4. * d[5] + d[6];
    q4 = d[3] - 4. * d[4] + 6. * d[5] - 4. * d[6] + d[7];
    return 1. / 5. * w0 * q1 + 1. / 20. * w2 * q2 + 1. / 60. * w3 * (3. * q2 - 2. * q3)
         + 1. / 60. * w4 * (3. * q2 - 2. * q3 + 3. * q4) - 1. / 840. * (39. * q2 - 14. * q3 + 3. * q4);
}

/* pointwise flux in direction dir (src/WENO.c:251-290); the convective momentum terms carry the
 * BC_Corrector factors (1 away from wall ghosts) */
static void point_flux(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                       double *const der[8], long n, double F[5])
{
    double J = met[9][n], rho = prim[0][n], th = der[4 + dir][n], p = der[0][n];
    double c0 = c->corr[0] ? c->corr[0][n] : 1.0, c1 = c->corr[1] ? c->corr[1][n] : 1.0,
           c2 = c->corr[2] ? c->corr[2][n] : 1.0;
    F[0] = J * (rho * th);
    F[1] = J * (rho * th * prim[1][n] * c0 + met[3 * dir + 0][n] * p * c->upsilon);
    F[2] = J * (rho * th * prim[2][n] * c1 + met[3 * dir + 1][n] * p * c->upsilon);
    F[3] = J * (rho * th * prim[3][n] * c2 + met[3 * dir + 2][n] * p * c->upsilon);
    F[4] = J * (rho * th * (prim[4][n] + p / rho * c->upsilon));
}

static double metric_norm(double *const met[10], int dir, long n)
{
    double a = met[3 * dir][n], b = met[3 * dir + 1][n], d = met[3 * dir + 2][n];
    return sqrt(pow(a, 2) + pow(b, 2) + pow(d, 2.));
}

/* One interface i+1/2 along dir at node n (CalcFluxesIn*Direction inner body,
 * src/WENO.c:296-464; GetLambdaMax* src/WENO.c:1550-1752; Theta_Function_W5/W9
 * src/WENO.c:1754-1816).  F is the pointwise flux array set for this direction. */
static void interface_flux(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                           double *const der[8], double *const F[5], long n, double out[5])
{
    long st = stride(c, dir);
    long np1 = n + st;
    int r = (c->S - 1) / 2, cc, m, s;
    double Jh, lam = 0.0, nh[3], L[25], R[25];
    double fp[9], fm[9], dp[8], dm[8], phis[5];
    double u, v, w, rho, cs;

    Jh = (met[9][n] + met[9][np1]) / 2.0;
    u = 0.5 * (prim[1][n] + prim[1][np1]);
    v = 0.5 * (prim[2][n] + prim[2][np1]);
    w = 0.5 * (prim[3][n] + prim[3][np1]);
    rho = 0.5 * (prim[0][n] + prim[0][np1]);
    cs = 0.5 * (der[2][n] + der[2][np1]);
    for (m = 0; m < 3; m++) {
        nh[m] = 0.5 * (met[3 * dir + m][n] / metric_norm(met, dir, n) +
                       met[3 * dir + m][np1] / metric_norm(met, dir, np1));
    }
    orc_eigvec(nh, u, v, w, rho, cs, c->gamma, L, R);

    for (s = -r; s <= r + 1; s++) {
        long ns = n + s * st;
        double a = met[3 * dir][ns], b = met[3 * dir + 1][ns], d = met[3 * dir + 2][ns];
        double l = fabs(der[4 + dir][ns]) + der[2][ns] * sqrt(a * a + b * b + d * d);
        if (lam < l) lam = l;
    }

    for (cc = 0; cc < NEQ; cc++) {
        for (s = 0; s < c->S; s++) {
            long npl = n + (s - r) * st;          /* plus stencil i-r .. i+r      */
            long nmi = n + (r + 1 - s) * st;      /* minus stencil i+r+1 .. i-r+1 */
            fp[s] = 0.5 * (L[cc * 5 + 0] * (F[0][npl] + lam * prim[0][npl] * Jh)
                         + L[cc * 5 + 1] * (F[1][npl] + lam * prim[0][npl] * prim[1][npl] * Jh)
                         + L[cc * 5 + 2] * (F[2][npl] + lam * prim[0][npl] * prim[2][npl] * Jh)
                         + L[cc * 5 + 3] * (F[3][npl] + lam * prim[0][npl] * prim[3][npl] * Jh)
                         + L[cc * 5 + 4] * (F[4][npl] + lam * prim[0][npl] * prim[4][npl] * Jh));
            fm[s] = 0.5 * (L[cc * 5 + 0] * (F[0][nmi] - lam * prim[0][nmi] * Jh)
                         + L[cc * 5 + 1] * (F[1][nmi] - lam * prim[0][nmi] * prim[1][nmi] * Jh)
                         + L[cc * 5 + 2] * (F[2][nmi] - lam * prim[0][nmi] * prim[2][nmi] * Jh)
                         + L[cc * 5 + 3] * (F[3][nmi] - lam * prim[0][nmi] * prim[3][nmi] * Jh)
                         + L[cc * 5 + 4] * (F[4][nmi] - lam * prim[0][nmi] * prim[4][nmi] * Jh));
        }
        for (s = 0; s <= c->S - 2; s++) {
            dp[s] = fp[s + 1] - fp[s];
            dm[s] = fm[s] - fm[s + 1];
        }
        if (c->S == 9)
            phis[cc] = (-orc_phi_w9(c->eps, fp, dp) + orc_phi_w9(c->eps, fm, dm));
        else
            phis[cc] = (-orc_phi_w5(c->eps, fp, dp) + orc_phi_w5(c->eps, fm, dm));
    }
    for (m = 0; m < NEQ; m++) {
        double pr = phis[0] * R[m * 5 + 0] + phis[1] * R[m * 5 + 1] + phis[2] * R[m * 5 + 2]
                  + phis[3] * R[m * 5 + 3] + phis[4] * R[m * 5 + 4];
        double th;
        const double *Fm = F[m];
        if (c->S == 9) {
            th = 1. / 840. * (533. * (Fm[np1] + Fm[n]) - 139. * (Fm[n + 2 * st] + Fm[n - st])
                              + 29. * (Fm[n + 3 * st] + Fm[n - 2 * st]) - 3. * (Fm[n + 4 * st] + Fm[n - 3 * st]));
        } else {
            th = 1. / 12. * (-1. * (Fm[n - st] + Fm[n + 2 * st]) + 7. * (Fm[n] + Fm[np1]));
        }
        out[m] = pr + th;
    }
}

static void fill_point_flux(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                            double *const der[8], double *const F[5])
{
    int lo[3], hi[3], d, i, j, k;
    for (d = 0; d < 3; d++) { lo[d] = rs(c, d); hi[d] = re(c, d); }
    lo[dir] = 0; hi[dir] = ge(c, dir);
    for (i = lo[0]; i <= hi[0]; i++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (k = lo[2]; k <= hi[2]; k++) {
                long n = orc_idx(c, i, j, k);
                double Fv[5];
                int m;
                point_flux(c, dir, met, prim, der, n, Fv);
                for (m = 0; m < 5; m++) F[m][n] = Fv[m];
            }
}

static double **alloc_fields(const orc_cfg *c, int nf)
{
    double **f = (double **)malloc(sizeof(double *) * nf);
    int m;
    for (m = 0; m < nf; m++) f[m] = (double *)calloc((size_t)c->ntot, sizeof(double));
    return f;
}
static void free_fields(double **f, int nf)
{
    int m;
    for (m = 0; m < nf; m++) free(f[m]);
    free(f);
}

/* CalcFluxesIn{Xi,Eta,Zeta}Direction, src/WENO.c:206-1083: interface fluxes for interfaces
 * g-1 .. g+n-1 along dir (interior transverse), then Q -= F(i+1/2) - F(i-1/2). */
void orc_weno_dir(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                  double *const der[8], double *const q[5])
{
    double **F = alloc_fields(c, 5), **Fh = alloc_fields(c, 5);
    int lo[3], hi[3], d, i, j, k, m;
    long st = stride(c, dir);
    fill_point_flux(c, dir, met, prim, der, F);
    for (d = 0; d < 3; d++) { lo[d] = rs(c, d); hi[d] = re(c, d); }
    lo[dir] = rs(c, dir) - 1;
#pragma omp parallel for private(j, k, m) schedule(static)
    for (i = lo[0]; i <= hi[0]; i++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (k = lo[2]; k <= hi[2]; k++) {
                long n = orc_idx(c, i, j, k);
                double out[5];
                interface_flux(c, dir, met, prim, der, F, n, out);
                for (m = 0; m < 5; m++) Fh[m][n] = out[m];
            }
    lo[dir] = rs(c, dir);
    for (i = lo[0]; i <= hi[0]; i++)
        for (j = lo[1]; j <= hi[1]; j++)
            for (k = lo[2]; k <= hi[2]; k++) {
                long n = orc_idx(c, i, j, k);
                for (m = 0; m < 5; m++) q[m][n] = q[m][n] - (Fh[m][n] - Fh[m][n - st]);
            }
    free_fields(F, 5);
    free_fields(Fh, 5);
}

void orc_weno_interface(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                        double *const der[8], int i, int j, int k, double out[5])
{
    double **F = alloc_fields(c, 5);
    fill_point_flux(c, dir, met, prim, der, F);
    interface_flux(c, dir, met, prim, der, F, orc_idx(c, i, j, k), out);
    free_fields(F, 5);
}

void orc_point_flux_at(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                       double *const der[8], int i, int j, int k, double F[5])
{
    point_flux(c, dir, met, prim, der, orc_idx(c, i, j, k), F);
}

/* CreateViscidMetric, src/Functions.c:6245-6386: 30 factors of one node for direction dir.
 * m = (xi_x,xi_y,xi_z,eta_x,eta_y,eta_z,zeta_x,zeta_y,zeta_z,J).  The table is the contraction
 * of the Newtonian stress (4/3, 1, -2/3 weights) with metric rows a (= dir) and b in {xi,eta,zeta}. */
void orc_visc_fac(int dir, const double m[10], double fac[30])
{
    const double J = m[9];
    const double *a = &m[3 * dir];      /* own-direction row (ax, ay, az) */
    int b, k;
    /* For each component block (x: 0..8, y: 9..17, z: 18..26) and each gradient slot
     * (u_b, v_b, w_b for b = xi, eta, zeta) the factor multiplies d(vel)/d(b). */
    for (b = 0; b < 3; b++) {
        const double *r = &m[3 * b];    /* row of direction b (rx, ry, rz) */
        /* x-momentum block */
        fac[0 + b]  = J * (4.0 / 3.0 * r[0] * a[0] + r[1] * a[1] + r[2] * a[2]);
        fac[3 + b]  = J * (-2.0 / 3.0 * r[1] * a[0] + r[0] * a[1]);
        fac[6 + b]  = J * (-2.0 / 3.0 * r[2] * a[0] + r[0] * a[2]);
        /* y-momentum block */
        fac[9 + b]  = J * (r[1] * a[0] - 2.0 / 3.0 * r[0] * a[1]);
        fac[12 + b] = J * (r[0] * a[0] + 4.0 / 3.0 * r[1] * a[1] + r[2] * a[2]);
        fac[15 + b] = J * (-2.0 / 3.0 * r[2] * a[1] + r[1] * a[2]);
        /* z-momentum block */
        fac[18 + b] = J * (r[2] * a[0] - 2.0 / 3.0 * r[0] * a[2]);
        fac[21 + b] = J * (r[2] * a[1] - 2.0 / 3.0 * r[1] * a[2]);
        fac[24 + b] = J * (r[0] * a[0] + r[1] * a[1] + 4.0 / 3.0 * r[2] * a[2]);
        /* heat flux */
        fac[27 + b] = J * (r[0] * a[0] + r[1] * a[1] + r[2] * a[2]);
    }
    (void)k;
}

/* CalcDeviationsForDirectViscidFluxes, src/ZD.c:697-810: nodal central derivatives of u,v,w,T
 * along each direction; range: own direction interior, transverse directions all nodes. */
void orc_visc_derivs(const orc_cfg *c, double *const prim[5], double *const der[8], double *const grad[12])
{
    int dir, lo[3], hi[3], d, i, j, k, l;
    for (dir = 0; dir < c->dims; dir++) {
        long st = stride(c, dir);
        int half = (c->S + 1) / 2;
        for (d = 0; d < 3; d++) { lo[d] = 0; hi[d] = ge(c, d); }
        lo[dir] = rs(c, dir); hi[dir] = re(c, dir);
        for (i = lo[0]; i <= hi[0]; i++)
            for (j = lo[1]; j <= hi[1]; j++)
                for (k = lo[2]; k <= hi[2]; k++) {
                    long n = orc_idx(c, i, j, k);
                    double du = 0.0, dv = 0.0, dw = 0.0, dT = 0.0;
                    for (l = 0; l <= c->S + 1; l++) {
                        long nl = n + (l - half) * st;
                        du += c->zd_der[l] * prim[1][nl];
                        dv += c->zd_der[l] * prim[2][nl];
                        dT += c->zd_der[l] * der[1][nl];
                        if (c->dims == 3) dw += c->zd_der[l] * prim[3][nl];
                    }
                    grad[0 + dir][n] = du;
                    grad[3 + dir][n] = dv;
                    if (c->dims == 3) grad[6 + dir][n] = dw;
                    grad[9 + dir][n] = dT;
                }
    }
}

/* CalcViscidFluxesIn{Xi,Eta,Zeta}DirectionDirectly, src/ZD.c:47-694.  With
 * ORC_QUIRK_ZETA_VISC_J the zeta stencil's k offset is taken from j (src/ZD.c:547-548). */
void orc_visc_dir(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                  double *const der[8], double *const grad[12], double *const q[5])
{
    int i, j, k, l, s, b;
    int half = (c->S + 1) / 2;
    long st = stride(c, dir);
    int three = (c->dims == 3);
#pragma omp parallel for private(j, k, l, s, b) schedule(static)
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                /* d[side][slot]: slot 0..8 = u_b,v_b,w_b blocks (u_xi,u_eta,u_zeta,v_xi,...), 9..11 = T_b */
                double dm[12], dp[12], Km[30], Kp[30], fm[30], f0[30], fp[30], mv[10];
                long nm = n - st, np = n + st;
                for (s = 0; s < 12; s++) { dm[s] = 0.; dp[s] = 0.; }
                for (l = 0; l <= c->S; l++) {
                    long im, ip;
                    if (dir == 2 && (c->quirks & ORC_QUIRK_ZETA_VISC_J)) {
                        long base = ((long)i * c->ng[1] + j) * c->ng[2];
                        im = base + (j + l - half);
                        ip = base + (j + l - half) + 1;
                    } else {
                        im = n + (l - half) * st;
                        ip = n + (l - half + 1) * st;
                    }
                    /* own-direction midpoint derivatives of u, v, (w), T */
                    dm[0 * 3 + dir] += c->zd_mid[l] * prim[1][im]; dp[0 * 3 + dir] += c->zd_mid[l] * prim[1][ip];
                    dm[1 * 3 + dir] += c->zd_mid[l] * prim[2][im]; dp[1 * 3 + dir] += c->zd_mid[l] * prim[2][ip];
                    if (three) { dm[2 * 3 + dir] += c->zd_mid[l] * prim[3][im]; dp[2 * 3 + dir] += c->zd_mid[l] * prim[3][ip]; }
                    dm[9 + dir] += c->zd_mid[l] * der[1][im]; dp[9 + dir] += c->zd_mid[l] * der[1][ip];
                    /* cross derivatives interpolated to the midpoints */
                    for (b = 0; b < 3; b++) {
                        if (b == dir) continue;
                        if (b == 2 && !three) continue;
                        dm[0 * 3 + b] += c->zd_int[l] * grad[0 * 3 + b][im]; dp[0 * 3 + b] += c->zd_int[l] * grad[0 * 3 + b][ip];
                        dm[1 * 3 + b] += c->zd_int[l] * grad[1 * 3 + b][im]; dp[1 * 3 + b] += c->zd_int[l] * grad[1 * 3 + b][ip];
                        if (three) { dm[2 * 3 + b] += c->zd_int[l] * grad[2 * 3 + b][im]; dp[2 * 3 + b] += c->zd_int[l] * grad[2 * 3 + b][ip]; }
                        dm[9 + b] += c->zd_int[l] * grad[9 + b][im]; dp[9 + b] += c->zd_int[l] * grad[9 + b][ip];
                    }
                }
                for (s = 0; s < 10; s++) mv[s] = met[s][nm];
                orc_visc_fac(dir, mv, fm);
                for (s = 0; s < 10; s++) mv[s] = met[s][n];
                orc_visc_fac(dir, mv, f0);
                for (s = 0; s < 10; s++) mv[s] = met[s][np];
                orc_visc_fac(dir, mv, fp);
                for (l = 0; l < 30; l++) {
                    Km[l] = 0.5 * (der[3][nm] * fm[l] + der[3][n] * f0[l]);
                    Kp[l] = 0.5 * (der[3][n] * f0[l] + der[3][np] * fp[l]);
                }
                {
                    double mom[3], heat, prod, work;
                    int blk;
                    for (blk = 0; blk < 3; blk++) {
                        double acc = dp[0] * Kp[blk * 9 + 0] - dm[0] * Km[blk * 9 + 0];
                        for (s = 1; s < 9; s++) acc = acc + dp[s] * Kp[blk * 9 + s] - dm[s] * Km[blk * 9 + s];
                        mom[blk] = acc;
                    }
                    q[1][n] = q[1][n] + c->psi * mom[0];
                    q[2][n] = q[2][n] + c->psi * mom[1];
                    if (three) q[3][n] = q[3][n] + c->psi * mom[2];
                    heat = dp[9] * Kp[27] - dm[9] * Km[27] + dp[10] * Kp[28] - dm[10] * Km[28]
                         + dp[11] * Kp[29] - dm[11] * Km[29];
                    /* u*(block x) + v*(block y) + w*(block z), each term separately as in src/ZD.c:200-229 */
                    work = 0.0;
                    {
                        double vel[3] = {prim[1][n], prim[2][n], prim[3][n]};
                        int first = 1;
                        for (blk = 0; blk < 3; blk++)
                            for (s = 0; s < 9; s++) {
                                double t = vel[blk] * (dp[s] * Kp[blk * 9 + s] - dm[s] * Km[blk * 9 + s]);
                                work = first ? t : work + t;
                                first = 0;
                            }
                    }
                    /* nodal product term: d(vel_blk)/d(dir) * grad_s * mu * fac */
                    prod = 0.0;
                    {
                        int first = 1;
                        for (blk = 0; blk < 3; blk++) {
                            double gd = grad[blk * 3 + dir][n];
                            for (s = 0; s < 9; s++) {
                                double t = gd * grad[s][n] * der[3][n] * f0[blk * 9 + s];
                                prod = first ? t : prod + t;
                                first = 0;
                            }
                        }
                    }
                    /* the factor of the node and its upper neighbour (src/ZD.c:193-194) */
                    double gh = c->gam ? 0.5 * (c->gam[np] + c->gam[n]) : c->gam_heat;
                    q[4][n] = q[4][n] + gh * heat + c->psi * (work + prod);
                }
            }
}

void orc_zero(const orc_cfg *c, double *const f[], int nf)
{
    int m;
    for (m = 0; m < nf; m++) memset(f[m], 0, sizeof(double) * (size_t)c->ntot);
}

/* Periodic ghost fill: the fixed point of the reference's ghost-inclusive halo passes
 * (TransferFlowParameterWithGhosts, src/Functions.c:3580-3684, repeated 1-3 times per stage,
 * src/Functions.c:2117-2136) when every face is a periodic self-interface: each ghost node takes
 * the value of the interior node at the wrapped index in every direction. */
void orc_ghost_periodic(const orc_cfg *c, double *const f[], int nf)
{
    int i, j, k, m, dd;
    for (i = 0; i < c->ng[0]; i++)
        for (j = 0; j < c->ng[1]; j++)
            for (k = 0; k < c->ng[2]; k++) {
                int idx[3] = {i, j, k}, src[3], ghost = 0;
                for (dd = 0; dd < 3; dd++) {
                    src[dd] = idx[dd];
                    if (dd == 2 && c->dims == 2) continue;
                    if (c->bc[2 * dd] != ORC_BC_PERIODIC) continue;
                    if (idx[dd] < c->g) { src[dd] = idx[dd] + c->n[dd]; ghost = 1; }
                    else if (idx[dd] > c->g + c->n[dd] - 1) { src[dd] = idx[dd] - c->n[dd]; ghost = 1; }
                }
                if (ghost) {
                    long to = orc_idx(c, i, j, k), from = orc_idx(c, src[0], src[1], src[2]);
                    for (m = 0; m < nf; m++) f[m][to] = f[m][from];
                }
            }
}

/* ---- boundary writers ---------------------------------------------------------------- */
/* WriteManufacturedSolution at a ghost node (src/ManufacturedSolution.c:566-648): the exact
 * rho,u,v,w,e,p (bs[0..5], evaluated by the caller) plus T = |p/rho|, c and mu = 1.  theta is not
 * written there, so it keeps the CalcValues value. */
static void bc_mms(const orc_cfg *c, double *const prim[5], double *const der[8], double *const bs[6], long n)
{
    int m;
    for (m = 0; m < 5; m++) prim[m][n] = bs[m][n];
    der[0][n] = bs[5][n];
    der[1][n] = fabs(der[0][n] / prim[0][n]);
    der[2][n] = (double)sqrtl((long double)(c->upsilon * c->gamma * der[0][n] / prim[0][n]));
    der[3][n] = 1.0;
}

/* WriteInflowSupersonicNormalBoundary (src/BC.c:2505-2542) */
static void bc_inflow_super(const orc_cfg *c, double *const met[10], double *const prim[5],
                            double *const der[8], long n)
{
    double p = c->p_in, rho = c->rho_in, u = c->u_in, v = c->v_in, w = c->w_in;
    prim[0][n] = rho; prim[1][n] = u; prim[2][n] = v; prim[3][n] = w;
    der[0][n] = p;
    der[4][n] = u * met[0][n] + v * met[1][n] + w * met[2][n];
    der[5][n] = u * met[3][n] + v * met[4][n] + w * met[5][n];
    der[6][n] = u * met[6][n] + v * met[7][n] + w * met[8][n];
    prim[4][n] = (0.5 * ((u * u) + (v * v) + (w * w)) + p / rho / (c->gamma - 1.0) * c->upsilon);
    der[2][n] = sqrt(c->upsilon * c->gamma * p / rho);
    der[1][n] = c->p_in / c->rho_in;
    der[3][n] = ((1.0 + c->suth) * pow(p / rho, 1.5) / (p / rho + c->suth));
}

/* WriteOutflowSubsonicRiemannBoundary (src/BC.c:2572-2622), mirror node s */
static void bc_outflow_riemann(const orc_cfg *c, double *const met[10], double *const prim[5],
                               double *const der[8], long n, long s)
{
    double cinf = 1.0 / c->mach, cii = der[2][s], d1, d2, rrr, u, v, w, rho, p;
    u = 0.5 * (prim[1][s] + c->u_in) - (cinf - cii) / (c->gamma - 1.);
    v = prim[2][s];
    w = prim[3][s];
    d1 = (0.5 * (cii + cinf) - 0.25 * (c->gamma - 1.) * (c->u_in - prim[1][s])) / cii;
    d2 = 2. / (c->gamma - 1.);
    rrr = pow(d1, d2);
    rho = prim[0][s] * rrr;
    d2 = 2. * c->gamma / (c->gamma - 1.);
    rrr = pow(d1, d2);
    p = der[0][s] * rrr;
    prim[0][n] = rho; prim[1][n] = u; prim[2][n] = v; prim[3][n] = w;
    der[0][n] = p;
    der[4][n] = u * met[0][n] + v * met[1][n] + w * met[2][n];
    der[5][n] = u * met[3][n] + v * met[4][n] + w * met[5][n];
    der[6][n] = u * met[6][n] + v * met[7][n] + w * met[8][n];
    der[2][n] = sqrt(c->upsilon * c->gamma * p / rho);
    prim[4][n] = (0.5 * ((u * u) + (v * v) + (w * w)) + p / rho / (c->gamma - 1.0) * c->upsilon);
    der[3][n] = der[3][s];
    der[1][n] = der[1][s];
}

/* One face: the ghost layers outside the face; tangential range all nodes for inflow/outflow
 * (src/BC.c:1344-1370, 2111-2140), interior for the manufactured solution
 * (src/ManufacturedSolution.c:202-338).  Mirror index: ghost StartReal-1-l <-> StartReal+l. */
/* Cartesian velocity at ghost node n from the contravariant velocities th of its mirror node and
   the ghost node's own metric rows: Cramer's rule for the 3x3 system, terms in the reference's
   order (src/BC.c:2556-2558, 2633-2635) */
static void velocity_from_theta(double *const met[10], long n, const double th[3], double *u, double *v,
                                double *w)
{
    double xx = met[0][n], xy = met[1][n], xz = met[2][n], ex = met[3][n], ey = met[4][n], ez = met[5][n];
    double zx = met[6][n], zy = met[7][n], zz = met[8][n];
    double den = -xx * ey * zz + xx * ez * zy + xy * ex * zz - xy * ez * zx - xz * ex * zy + xz * ey * zx;
    *u = -(th[0] * ey * zz - th[0] * ez * zy - th[1] * xy * zz + th[1] * xz * zy + th[2] * xy * ez - th[2] * xz * ey) / den;
    *v = -(-th[0] * ex * zz + th[0] * ez * zx + th[1] * xx * zz - th[1] * xz * zx - th[2] * xx * ez + th[2] * xz * ex) / den;
    *w = -(-th[0] * ex * zy + th[0] * ey * zx + th[1] * xx * zy - th[1] * xy * zx - th[2] * xx * ey + th[2] * xy * ex) / -den;
}
/* WriteOutflowSubsonicNormalBoundary (src/BC.c:2544-2570) and WriteOutflowSubsonicRudyBoundary
   (src/BC.c:2624-2655): theta, rho, e, c, mu and T from the mirror node s; p = p_out (normal) or
   relaxed from the step-start ghost state (Rudy); rho scaled by p/p_s (Rudy) */
static void bc_outflow_subsonic(const orc_cfg *c, int rudy, double *const met[10], double *const prim[5],
                                double *const der[8], long n, long s)
{
    double th[3] = {der[4][s], der[5][s], der[6][s]}, u, v, w, p;
    velocity_from_theta(met, n, th, &u, &v, &w);
    der[4][n] = th[0]; der[5][n] = th[1]; der[6][n] = th[2];
    prim[1][n] = u; prim[2][n] = v; prim[3][n] = w;
    if (rudy) {
        double at = c->alpha_nonref * c->tau_step;
        p = (c->p_last[n] + at * c->p_out + c->rho_last[n] * c->c_last[s] * (u - c->u_last[n]) / c->upsilon) / (1.0 + at);
        der[0][n] = p;
        prim[0][n] = p / der[0][s] * prim[0][s];
    } else {
        der[0][n] = c->p_out;
        prim[0][n] = prim[0][s];
    }
    prim[4][n] = prim[4][s];
    der[2][n] = der[2][s];
    der[3][n] = der[3][s];
    der[1][n] = der[1][s];
}
/* WriteWallSlipBoundary (src/BC.c:2688-2719): the mirror's contravariant velocities times the
 * face's corrector (the normal one reversed), the Cartesian velocity that has them under the ghost
 * node's metric rows, the corrector stored for the momentum fluxes; rho, p, e, c, mu, T copied.
 * WriteWallNoSlipBoundary (src/BC.c:2657-2686): velocities and contravariant velocities reversed,
 * corrector -1 on all three momentum fluxes; rho, p, e, c, mu, T copied. */
static void bc_wall(const orc_cfg *c, int slip, const double corr[3], double *const met[10],
                    double *const prim[5], double *const der[8], long n, long s, int iso)
{
    int m;
    if (slip) {
        double th[3] = {corr[0] * der[4][s], corr[1] * der[5][s], corr[2] * der[6][s]}, u, v, w;
        for (m = 0; m < 3; m++) der[4 + m][n] = th[m];
        velocity_from_theta(met, n, th, &u, &v, &w);
        prim[1][n] = u; prim[2][n] = v; prim[3][n] = w;
        for (m = 0; m < 3; m++) if (c->corr[m]) c->corr[m][n] = corr[m];
    } else {
        for (m = 0; m < 3; m++) der[4 + m][n] = -der[4 + m][s];
        for (m = 1; m < 4; m++) prim[m][n] = -prim[m][s];
        for (m = 0; m < 3; m++) if (c->corr[m]) c->corr[m][n] = -1.0;
    }
    prim[0][n] = prim[0][s];
    der[0][n] = der[0][s];
    prim[4][n] = prim[4][s];
    der[2][n] = der[2][s];
    der[3][n] = der[3][s];
    der[1][n] = der[1][s];
    if (iso) {
        /* WriteWallNoSlipIsothermalBoundary (src/BC.c:2943-2978): T = T_wall and the heat
           conduction factor of Prandtl 0.72 on the ghost */
        der[1][n] = c->T_wall;
        if (c->gam) c->gam[n] = 1.0 / ((c->gamma - 1.0) * pow(c->mach, 2.0) * c->reynolds * 0.72);
    }
}

/* Subsonic inflow (src/BC.c:2363-2503), mirror node s.  Normal (WriteInflowSubsonicNormalBoundary):
 * the inflow velocity, p and rho.  Riemann (WriteInflowSubsonicRiemannBoundary): the velocity
 * normal part from the far field's and the mirror's invariants, v and w of the inflow, rho and p
 * on the far field's isentrope.  Isentropic (WriteInflowSubsonicIsentropBoundary): the mirror's
 * contravariant velocities under the ghost metric, rho from the isentropic Bernoulli relation in
 * u, p on the far field's isentrope.  e, c and theta follow; T and mu come from the mirror. */
static void bc_inflow_sub(const orc_cfg *c, int kind, double *const met[10], double *const prim[5],
                          double *const der[8], long n, long s)
{
    double gm = c->gamma, u, v, w, rho, p;
    if (kind == ORC_BC_INFLOW_SUB_NORMAL) {
        u = c->u_in; v = c->v_in; w = c->w_in;
        p = c->p_in; rho = c->rho_in;
    } else if (kind == ORC_BC_INFLOW_SUB_RIEMANN) {
        double cinf = 1.0 / c->mach, cii = der[2][s], d1, d2, rrr;
        u = 0.5 * (prim[1][s] + c->u_in) + (cinf - cii) / (gm - 1.);
        v = c->v_in; w = c->w_in;
        d1 = (0.5 * (cii + cinf) + 0.25 * (gm - 1.) * (c->u_in - prim[1][s])) / cinf;
        d2 = 2. / (gm - 1.);
        rrr = pow(d1, d2);
        rho = c->rho_in * rrr;
        d2 = 2. * gm / (gm - 1.);
        rrr = pow(d1, d2);
        p = c->p_in * rrr;
    } else {
        double th[3] = {der[4][s], der[5][s], der[6][s]};
        velocity_from_theta(met, n, th, &u, &v, &w);
        rho = pow(pow(c->rho_in, (gm - 1.0)) - pow(c->rho_in, gm) / c->p_in * (gm - 1.0) / gm * gm * c->mach * c->mach
                  * (u * u / 2.0), 1.0 / (gm - 1.0));
        p = c->p_in * pow(rho / c->rho_in, gm);
    }
    prim[0][n] = rho; prim[1][n] = u; prim[2][n] = v; prim[3][n] = w;
    der[0][n] = p;
    if (kind == ORC_BC_INFLOW_SUB_ISENTROPIC) {
        der[4][n] = der[4][s]; der[5][n] = der[5][s]; der[6][n] = der[6][s];
    } else {
        der[4][n] = u * met[0][n] + v * met[1][n] + w * met[2][n];
        der[5][n] = u * met[3][n] + v * met[4][n] + w * met[5][n];
        der[6][n] = u * met[6][n] + v * met[7][n] + w * met[8][n];
    }
    prim[4][n] = (0.5 * ((u * u) + (v * v) + (w * w)) + p / rho / (gm - 1.0) * c->upsilon);
    der[2][n] = sqrt(c->upsilon * gm * p / rho);
    der[3][n] = der[3][s];
    der[1][n] = der[1][s];
}

static void apply_face(const orc_cfg *c, int face, double *const met[10], double *const prim[5],
                       double *const der[8], double *const bs[6])
{
    int d = face / 2, upper = face % 2, kind = c->bc[face];
    int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2, l, a, b;
    int lo1, hi1, lo2, hi2;
    /* the wall writers' corrector: the momentum component along the face's index direction */
    double corr[3] = {1.0, 1.0, 1.0};
    corr[d] = -1.0;
    if (kind == ORC_BC_PERIODIC) return;
    if (kind == ORC_BC_MMS || kind == ORC_BC_WALL_SLIP || kind == ORC_BC_WALL_NOSLIP || kind == ORC_BC_FARFIELD ||
        kind == ORC_BC_WALL_ISOTHERMAL) {
        lo1 = rs(c, t1); hi1 = re(c, t1); lo2 = rs(c, t2); hi2 = re(c, t2);
    }
    else { lo1 = 0; hi1 = ge(c, t1); lo2 = 0; hi2 = ge(c, t2); }
    for (l = 0; l < c->g; l++) {
        int gi = upper ? re(c, d) + 1 + l : rs(c, d) - 1 - l;
        int si = upper ? re(c, d) - l : rs(c, d) + l;
        for (a = lo1; a <= hi1; a++)
            for (b = lo2; b <= hi2; b++) {
                int idx[3], sdx[3];
                long n, s;
                idx[d] = gi; sdx[d] = si;
                idx[t1] = a; sdx[t1] = a;
                idx[t2] = b; sdx[t2] = b;
                n = orc_idx(c, idx[0], idx[1], idx[2]);
                s = orc_idx(c, sdx[0], sdx[1], sdx[2]);
                if (kind == ORC_BC_MMS) bc_mms(c, prim, der, bs, n);
                else if (kind == ORC_BC_INFLOW_SUPERSONIC) bc_inflow_super(c, met, prim, der, n);
                else if (kind == ORC_BC_OUTFLOW_RIEMANN) bc_outflow_riemann(c, met, prim, der, n, s);
                else if (kind == ORC_BC_OUTFLOW_NORMAL) bc_outflow_subsonic(c, 0, met, prim, der, n, s);
                else if (kind == ORC_BC_OUTFLOW_RUDY) bc_outflow_subsonic(c, 1, met, prim, der, n, s);
                else if (kind == ORC_BC_WALL_SLIP) bc_wall(c, 1, corr, met, prim, der, n, s, 0);
                else if (kind == ORC_BC_WALL_NOSLIP) bc_wall(c, 0, corr, met, prim, der, n, s, 0);
                else if (kind == ORC_BC_WALL_ISOTHERMAL) bc_wall(c, 0, corr, met, prim, der, n, s, 1);
                else if (kind == ORC_BC_FARFIELD) {
                    /* WriteFarfieldBoundary (src/BC.c:2817-2852): every value of the mirror node */
                    int m;
                    for (m = 0; m < 5; m++) prim[m][n] = prim[m][s];
                    for (m = 0; m < 7; m++) der[m][n] = der[m][s];
                }
                else bc_inflow_sub(c, kind, met, prim, der, n, s);
            }
    }
}

void orc_mirror_metric(const orc_cfg *c, double *const met[10])
{
    int face;
    for (face = 0; face < 2 * c->dims; face++) {
        int d = face / 2, upper = face % 2, kind = c->bc[face];
        int t1 = (d == 0) ? 1 : 0, t2 = (d == 2) ? 1 : 2, l, a, b, m;
        if (kind != ORC_BC_WALL_SLIP && kind != ORC_BC_WALL_NOSLIP && kind != ORC_BC_FARFIELD &&
            kind != ORC_BC_WALL_ISOTHERMAL) continue;
        for (l = 0; l < c->g; l++) {
            int gi = upper ? re(c, d) + 1 + l : rs(c, d) - 1 - l;
            int si = upper ? re(c, d) - l : rs(c, d) + l;
            for (a = 0; a <= ge(c, t1); a++)
                for (b = 0; b <= ge(c, t2); b++) {
                    int idx[3], sdx[3];
                    long n, s;
                    idx[d] = gi; sdx[d] = si;
                    idx[t1] = a; sdx[t1] = a;
                    idx[t2] = b; sdx[t2] = b;
                    n = orc_idx(c, idx[0], idx[1], idx[2]);
                    s = orc_idx(c, sdx[0], sdx[1], sdx[2]);
                    for (m = 0; m < 10; m++) met[m][n] = met[m][s];
                }
        }
    }
}

void orc_apply_bc(const orc_cfg *c, double *const met[10], double *const prim[5], double *const der[8],
                  double *const bstate[6])
{
    int face;
    for (face = 0; face < 2 * c->dims; face++) apply_face(c, face, met, prim, der, bstate);
}

/* RK stage update over interior nodes, src/Functions.c:2300-2365 (in place on urk). */
void orc_rk_update(const orc_cfg *c, int s, double tau, double *const met[10],
                   double *const un[5], double *const urk[5], double *const q[5], double *const qsum[5])
{
    int i, j, k, m;
    double a = c->rk_a[s], b = c->rk_b[s], qf = c->rk_q[s], ff = c->rk_f[s], mf = c->rk_m[s];
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                double rho_old, J = met[9][n];
                for (m = 0; m < 5; m++) qsum[m][n] = qsum[m][n] + mf * q[m][n];
                rho_old = urk[0][n];
                urk[0][n] = a * un[0][n] + b * urk[0][n] + tau / J * (qf * q[0][n] + ff * qsum[0][n]);
                for (m = 1; m < 5; m++) {
                    if (m == 3 && c->dims == 2) { urk[3][n] = 0.0; continue; }
                    urk[m][n] = (a * un[m][n] * un[0][n] + b * urk[m][n] * rho_old
                                 + tau / J * (qf * q[m][n] + ff * qsum[m][n])) / urk[0][n];
                }
            }
}

/* AddManufacturedSolutionSource, src/ManufacturedSolution.c:25-175: Q += J * source. */
void orc_add_source(const orc_cfg *c, double *const met[10], double *const src[5], double *const q[5])
{
    int i, j, k, m;
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                for (m = 0; m < 5; m++) {
                    if (m == 3 && c->dims == 2) continue;
                    q[m][n] += met[9][n] * src[m][n];
                }
            }
}

void orc_rhs(const orc_cfg *c, double *const met[10], double *const prim[5],
             double *const scratch[30], double *const src[5])
{
    orc_rhs_b(c, met, prim, scratch, src, NULL);
}

void orc_rhs_b(const orc_cfg *c, double *const met[10], double *const prim[5],
               double *const scratch[30], double *const src[5], double *const bstate[6])
{
    double *const *der = scratch, *const *grad = scratch + 8, *const *q = scratch + 20;
    int d;
    orc_zero(c, q, 5);
    orc_zero(c, grad, 12);
    orc_ghost_periodic(c, prim, 5);
    orc_calc_values(c, met, prim, der);
    orc_apply_bc(c, met, prim, der, bstate);
    for (d = 0; d < c->dims; d++) orc_weno_dir(c, d, met, prim, der, q);
    if (!c->inviscid) {
        orc_visc_derivs(c, prim, der, grad);
        for (d = 0; d < c->dims; d++) orc_visc_dir(c, d, met, prim, der, grad, q);
    }
    if (src) orc_add_source(c, met, src, q);
}

/* CalcRungeKutta, src/Functions.c:2034-2390, for an all-periodic single block. */
void orc_step(const orc_cfg *c, double tau, double *const met[10], double *const un[5],
              double *const urk[5], double *const scratch[30], double *const src[5])
{
    orc_step_b(c, tau, met, un, urk, scratch, src, NULL);
}

void orc_step_b(const orc_cfg *c, double tau, double *const met[10], double *const un[5],
                double *const urk[5], double *const scratch[30], double *const src[5],
                double *const bstate[6])
{
    double *const *q = scratch + 20, *const *qsum = scratch + 25;
    int s;
    orc_zero(c, qsum, 5);
    for (s = 0; s < c->time_order; s++) {
        orc_rhs_b(c, met, urk, scratch, src, bstate);
        orc_rk_update(c, s, tau, met, un, urk, q, qsum);
    }
}

/* checkNAN, src/Functions.c:5419-5467: count non-finite interior rho, overwrite with 1e6. */
int orc_check_nan(const orc_cfg *c, double *const prim[5])
{
    int i, j, k, cnt = 0;
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                if (isinf(prim[0][n]) || isnan(prim[0][n])) {
                    cnt++;
                    prim[1][n] = prim[2][n] = prim[3][n] = prim[0][n] = 1000000;
                }
            }
    return cnt;
}

/* CFL time step (absent in the reference, which runs with a fixed tau, src/Import.c:49):
 * dt = CFL / max_interior sum_d (|theta_d| + c |grad xi_d|), the per-node spectral radius of
 * GetLambdaMax* (src/WENO.c:1560-1615) summed over directions, unit index spacing. */
double orc_cfl_dt(const orc_cfg *c, double cfl, double *const met[10], double *const prim[5])
{
    int i, j, k, d;
    double mx = 0.0;
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                double rho = prim[0][n], u = prim[1][n], v = prim[2][n], w = prim[3][n], e = prim[4][n];
                double p = fabs(rho * (c->gamma - 1) / c->upsilon * (e - 0.5 * ((u * u) + (v * v) + (w * w))));
                double cs = sqrt(c->upsilon * c->gamma * p / rho), sum = 0.0;
                for (d = 0; d < c->dims; d++) {
                    double a = met[3 * d][n], b = met[3 * d + 1][n], z = met[3 * d + 2][n];
                    double th = u * a + v * b + w * z;
                    sum += fabs(th) + cs * sqrt(a * a + b * b + z * z);
                }
                if (sum > mx) mx = sum;
            }
    return mx > 0.0 ? cfl / mx : 0.0;
}

/* CreateMetric, src/Functions.c:6388-6509 (2nd-order central differences in index space on the
 * extended coordinate arrays; 3D off-diagonal terms zeroed under ORC_QUIRK_METRIC_DIAG3D). */
void orc_metric_from_coords(const orc_cfg *c, const double *xe, const double *ye,
                            const double *ze, double *const met[10])
{
    int ne1 = c->ng[1] + 2, ne2 = (c->dims == 3) ? c->ng[2] + 2 : 1;
    int i, j, k;
    for (i = 0; i < c->ng[0]; i++)
        for (j = 0; j < c->ng[1]; j++)
            for (k = 0; k < c->ng[2]; k++) {
                int ie = i + 1, je = j + 1, ke = (c->dims == 3) ? k + 1 : 0;
                long ce = ((long)ie * ne1 + je) * ne2 + ke;
                long ip = ce + (long)ne1 * ne2, im = ce - (long)ne1 * ne2;
                long jp = ce + ne2, jm = ce - ne2, kp = ce + 1, km = ce - 1;
                double x_xi, y_xi, z_xi, x_eta, y_eta, z_eta, x_zeta, y_zeta, z_zeta, J;
                long n = orc_idx(c, i, j, k);
                if (c->dims == 2) {
                    x_xi = (-0.5 * xe[im] + 0.5 * xe[ip]);
                    y_xi = (-0.5 * ye[im] + 0.5 * ye[ip]);
                    z_xi = 0.0;
                    x_eta = (-0.5 * xe[jm] + 0.5 * xe[jp]);
                    y_eta = (-0.5 * ye[jm] + 0.5 * ye[jp]);
                    z_eta = 0.0;
                    x_zeta = 0.0; y_zeta = 0.0; z_zeta = 1.0;
                } else {
                    x_xi = (-0.5 * xe[im] + 0.5 * xe[ip]);
                    y_xi = (-0.5 * ye[im] + 0.5 * ye[ip]);
                    z_xi = (-0.5 * ze[im] + 0.5 * ze[ip]);
                    x_eta = (-0.5 * xe[jm] + 0.5 * xe[jp]);
                    y_eta = (-0.5 * ye[jm] + 0.5 * ye[jp]);
                    z_eta = (-0.5 * ze[jm] + 0.5 * ze[jp]);
                    x_zeta = (-0.5 * xe[km] + 0.5 * xe[kp]);
                    y_zeta = (-0.5 * ye[km] + 0.5 * ye[kp]);
                    z_zeta = (-0.5 * ze[km] + 0.5 * ze[kp]);
                    if (c->quirks & ORC_QUIRK_METRIC_DIAG3D) {
                        y_xi = 0.0; z_xi = 0.0; x_eta = 0.0; z_eta = 0.0; x_zeta = 0.0; y_zeta = 0.0;
                    }
                }
                J = x_xi * y_eta * z_zeta + x_eta * y_zeta * z_xi + x_zeta * y_xi * z_eta
                  - x_zeta * y_eta * z_xi - x_eta * y_xi * z_zeta - x_xi * y_zeta * z_eta;
                met[9][n] = J;
                met[0][n] = (y_eta * z_zeta - y_zeta * z_eta) / J;
                met[1][n] = (x_zeta * z_eta - x_eta * z_zeta) / J;
                met[2][n] = (x_eta * y_zeta - x_zeta * y_eta) / J;
                met[3][n] = (y_zeta * z_xi - y_xi * z_zeta) / J;
                met[4][n] = (x_xi * z_zeta - x_zeta * z_xi) / J;
                met[5][n] = (x_zeta * y_xi - x_xi * y_zeta) / J;
                met[6][n] = (y_xi * z_eta - y_eta * z_xi) / J;
                met[7][n] = (x_eta * z_xi - x_xi * z_eta) / J;
                met[8][n] = (x_xi * y_eta - x_eta * y_xi) / J;
            }
}

/* check_Metric (src/Functions.c:6198-6243): padded nodes whose Jacobian is negative */
int orc_check_metric(const orc_cfg *c, double *const met[10])
{
    long n;
    int bad = 0;
    for (n = 0; n < c->ntot; n++)
        if (met[9][n] < 0.0) bad++;
    return bad;
}

/* lambda2 of node n (CalcLambda2, src/Functions.c:5846-6070): velocity derivatives as differences
 * of midpoint-interpolated values (the viscous interpolation table), Cartesian gradient through
 * the metric, the middle eigenvalue of S^2 + Omega^2 by the trigonometric cubic solution, capped
 * at 0.  A degenerate case where no root lies strictly between the others keeps 0. */
static double lambda2_at(const orc_cfg *c, double *const met[10], double *const prim[5], int i, int j, int k)
{
    long n = orc_idx(c, i, j, k), st[3] = {stride(c, 0), stride(c, 1), stride(c, 2)};
    int half = (c->S + 1) / 2, m, d, q;
    double dv[3][3];           /* dv[q][d]: velocity component q along index direction d */
    double lo[3][3], hi[3][3];
    for (q = 0; q < 3; q++)
        for (d = 0; d < 3; d++) { lo[q][d] = 0.; hi[q][d] = 0.; }
    for (m = 0; m <= c->S; m++)
        for (d = 0; d < 3; d++) {
            long i0 = n + (long)(m - half) * st[d], i1 = n + (long)(m - half + 1) * st[d];
            for (q = 0; q < 3; q++) {
                lo[q][d] += c->zd_int[m] * prim[1 + q][i0];
                hi[q][d] += c->zd_int[m] * prim[1 + q][i1];
            }
        }
    for (q = 0; q < 3; q++)
        for (d = 0; d < 3; d++) dv[q][d] = hi[q][d] - lo[q][d];
    {
        double xx = met[0][n], xy = met[1][n], xz = met[2][n], ex = met[3][n], ey = met[4][n], ez = met[5][n];
        double zx = met[6][n], zy = met[7][n], zz = met[8][n];
        double udx = dv[0][0] * xx + dv[0][1] * ex + dv[0][2] * zx, udy = dv[0][0] * xy + dv[0][1] * ey + dv[0][2] * zy;
        double udz = dv[0][0] * xz + dv[0][1] * ez + dv[0][2] * zz;
        double vdx = dv[1][0] * xx + dv[1][1] * ex + dv[1][2] * zx, vdy = dv[1][0] * xy + dv[1][1] * ey + dv[1][2] * zy;
        double vdz = dv[1][0] * xz + dv[1][1] * ez + dv[1][2] * zz;
        double wdx = dv[2][0] * xx + dv[2][1] * ex + dv[2][2] * zx, wdy = dv[2][0] * xy + dv[2][1] * ey + dv[2][2] * zy;
        double wdz = dv[2][0] * xz + dv[2][1] * ez + dv[2][2] * zz;
        double om12 = 0.5 * (udy - vdx), om13 = 0.5 * (udz - wdx), om23 = 0.5 * (vdz - wdy);
        double s11 = 0.5 * (udx), s12 = 0.5 * (udy + vdx), s13 = 0.5 * (udz + wdx);
        double s22 = 0.5 * (vdy), s23 = 0.5 * (vdz + wdy), s33 = 0.5 * (wdz);
        double oo11 = -om12 * om12 - om13 * om13, oo12 = -om13 * om23, oo13 = om12 * om23;
        double oo22 = -om12 * om12 - om23 * om23, oo23 = -om12 * om13, oo33 = -om13 * om13 - om23 * om23;
        double ss11 = s11 * s11 + s12 * s12 + s13 * s13, ss12 = s11 * s12 + s12 * s22 + s13 * s23;
        double ss13 = s11 * s13 + s12 * s23 + s13 * s33, ss22 = s12 * s12 + s22 * s22 + s23 * s23;
        double ss23 = s12 * s13 + s22 * s23 + s23 * s33, ss33 = s13 * s13 + s23 * s23 + s33 * s33;
        double a11 = ss11 + oo11, a12 = ss12 + oo12, a13 = ss13 + oo13, a22 = ss22 + oo22, a23 = ss23 + oo23;
        double a33 = ss33 + oo33;
        double f1 = -a11 - a22 - a33;
        double f2 = -a23 * a23 - a12 * a12 - a13 * a13 + a11 * a33 + a11 * a22 + a22 * a33;
        double f3 = -a11 * a22 * a33 + a11 * a23 * a23 + a22 * a13 * a13 + a33 * a12 * a12 - 2. * a12 * a13 * a23;
        double p = -f1 * f1 / 3. + f2, qq = 2. * f1 * f1 * f1 / 27. - f1 * f2 / 3. + f3;
        double rrr = sqrt((p < 0. ? -p : p) / 3.), al, l1, l2, l3, lmax, lmin, lam = 0.;
        const double pi = 3.14159265358979323846264338327950288;
        al = acos(qq / (2. * rrr * rrr * rrr));
        l1 = -2. * rrr * cos(al / 3.) - f1 / 3;
        l2 = 2. * rrr * cos(pi / 3. - al / 3.) - f1 / 3;
        l3 = 2. * rrr * cos(pi / 3. + al / 3.) - f1 / 3;
        lmax = (l2 >= l1) ? l2 : l1;
        if (l3 >= lmax) lmax = l3;
        lmin = (l2 <= l1) ? l2 : l1;
        if (l3 <= lmin) lmin = l3;
        if (l1 == l2 || l2 == l3) lam = l2;
        if (l1 > lmin && l1 < lmax) lam = l1;
        else if (l2 > lmin && l2 < lmax) lam = l2;
        else if (l3 > lmin && l3 < lmax) lam = l3;
        if (lam >= 0.) lam = 0.;
        return lam;
    }
}

/* CalcValuesForPost (src/Functions.c:3687-3761) on interior nodes: Mach number from the stored
 * pressure p (U_lastStep's, one RK stage behind the state: the RK update does not renew it), the
 * density "gradient" as the root of the sum of squares of every (central difference x metric
 * entry) product, and lambda2 (3D; untouched in 2D).  out = Mach, gradRho, Lambda2. */
void orc_calc_post(const orc_cfg *c, double *const met[10], double *const prim[5], const double *p,
                   double *const out[3])
{
    int i, j, k;
    long si = stride(c, 0), sj = stride(c, 1), sk = stride(c, 2);
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++)
            for (k = rs(c, 2); k <= re(c, 2); k++) {
                long n = orc_idx(c, i, j, k);
                const double *r = prim[0];
                double u = prim[1][n], v = prim[2][n], w = prim[3][n];
                double dxi = (0.5 * r[n + si] - 0.5 * r[n - si]) / 1.0;
                double deta = (0.5 * r[n + sj] - 0.5 * r[n - sj]) / 1.0;
                out[0][n] = (sqrt(u * u + v * v + w * w)) / sqrt(p[n] / r[n]) * c->mach;
                if (c->dims == 2) {
                    out[1][n] = sqrt(pow(dxi * met[0][n], 2.0) + pow(deta * met[3][n], 2.0)
                                     + pow(dxi * met[1][n], 2.0) + pow(deta * met[4][n], 2.0));
                } else {
                    double dze = (0.5 * r[n + sk] - 0.5 * r[n - sk]) / 1.0;
                    out[1][n] = sqrt(pow(dxi * met[0][n], 2.0) + pow(deta * met[3][n], 2.0) + pow(dze * met[6][n], 2.0)
                                     + pow(dxi * met[1][n], 2.0) + pow(deta * met[4][n], 2.0) + pow(dze * met[7][n], 2.0)
                                     + pow(dxi * met[2][n], 2.0) + pow(deta * met[5][n], 2.0) + pow(dze * met[8][n], 2.0));
                    out[2][n] = lambda2_at(c, met, prim, i, j, k);
                }
            }
}

/* ErrorManufacturedSolution (src/ManufacturedSolution.c:745-790): long double sums of the squared
 * deviations of rho and of the stored p from the exact values, each divided by the interior node
 * count, over the interior nodes of the last k plane (the reference's k loop starts at kEndReal),
 * and the largest deviations.  ex_rho, ex_p: exact values on the padded grid.
 * out = L2(rho), Linf(rho), L2(p), Linf(p). */
void orc_mms_error(const orc_cfg *c, double *const prim[5], const double *p, const double *ex_rho,
                   const double *ex_p, double out[4])
{
    long double N = (long double)c->n[0] * c->n[1] * c->n[2], l2r = 0.0L, lir = 0.0L, l2p = 0.0L, lip = 0.0L;
    int i, j, k = re(c, 2);
    for (i = rs(c, 0); i <= re(c, 0); i++)
        for (j = rs(c, 1); j <= re(c, 1); j++) {
            long n = orc_idx(c, i, j, k);
            long double dr = (long double)ex_rho[n] - prim[0][n], dp = (long double)ex_p[n] - p[n];
            l2r += powl(dr, 2.0L) / N;
            if (lir < fabsl(dr)) lir = fabsl(dr);
            l2p += powl(dp, 2.0L) / N;
            if (lip < fabsl(dp)) lip = fabsl(dp);
        }
    out[0] = (double)sqrtl(l2r);
    out[1] = (double)lir;
    out[2] = (double)sqrtl(l2p);
    out[3] = (double)lip;
}

void orc_set_threads(int n)
{
    if (n > 0)
        omp_set_num_threads(n);
}
