/*
 * compat_run.c -- the reference's step loop (src/SHOCK.c:557-666: CalcRungeKutta, checkNAN,
 * WriteValuesFromU1ToU2) written against shock_compat.h, on a case read from a file.
 *
 * usage: compat_run <case file> <output file> [donor rank of the interface faces]
 * Case file (native endianness): int32 dims, SPACEORDER, n[3], TimeOrder, Inviscid, steps,
 * face[6] (0 neighbour, 1 Inflow + supersonic option, 2 Outflow + Riemann option,
 * 3 WallInviscid, 4 WallViscous, 5 Farfield); float64 gamma, Mach, Reynolds, Prandtl, T0, tau,
 * rho, u, v, w, p of the inflow, p_out, AlphaNonRef, T_wall; then the metric (xi_x .. zeta_z, J)
 * and rho, u, v, w, e, each on the padded grid.  Output: U_lastStep's rho, u, v, w, e.
 * Prints the time of the first step (it uploads the metric and the state) and the mean time of
 * the others, host-to-host: the loop's own state copies and compares in host memory included.
 */
#include "shock_compat.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (fread(p, sz, n, f) != n) { fprintf(stderr, "short case file\n"); exit(2); }
}

static void alloc_U(struct strct_U *U, long ntot)
{
    FLT **f[12] = {&U->rho, &U->u, &U->v, &U->w, &U->p, &U->e, &U->c, &U->T, &U->mue, &U->theta1, &U->theta2, &U->theta3};
    for (int m = 0; m < 12; m++) *f[m] = (FLT *)calloc(ntot, sizeof(FLT));
}

int main(int argc, char **argv)
{
    if (argc < 3) { fprintf(stderr, "usage: compat_run <case> <out>\n"); return 2; }
    FILE *in = fopen(argv[1], "rb");
    if (!in) { perror(argv[1]); return 2; }
    int hdr[14];
    double par[14];
    rd(hdr, sizeof(int), 14, in);
    rd(par, sizeof(double), 14, in);
    struct strct_configuration cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.MESHDIMENSIONS = hdr[0]; cfg.SPACEORDER = hdr[1];
    cfg.int_iMeshPoints = hdr[2]; cfg.int_jMeshPoints = hdr[3]; cfg.int_kMeshPoints = hdr[4];
    cfg.int_TimeOrder = hdr[5]; cfg.flag_Inviscid = hdr[6];
    const int steps = hdr[7];
    cfg.gammaNumber = par[0]; cfg.machNumber = par[1]; cfg.reynoldsNumber = par[2]; cfg.prandtlNumber = par[3];
    cfg.T0_dim = par[4]; cfg.numericalTau = par[5];
    cfg.rho_inflow = par[6]; cfg.u_inflow = par[7]; cfg.v_inflow = par[8]; cfg.w_inflow = par[9]; cfg.p_inflow = par[10];
    cfg.p_out = par[11]; cfg.AlphaNonRef = par[12]; cfg.T_wall = par[13];
    /* the reference's BC names (config.ini) */
    strcpy(cfg.BCInflow, "Inflow"); strcpy(cfg.BCOutflow, "Outflow"); strcpy(cfg.BCWallInviscid, "WallInviscid");
    strcpy(cfg.BCWallViscous, "WallViscous"); strcpy(cfg.BCFarfield, "Farfield");
    strcpy(cfg.BCOutflowSubsonic, "OutflowSubsonic"); strcpy(cfg.BCInflowSupersonic, "InflowSupersonic");
    strcpy(cfg.BCInflowSubsonic, "InflowSubsonic"); strcpy(cfg.BCWallViscousIsothermal, "WallViscousIsothermal");
    strcpy(cfg.BCManufacturedSolution, "ManufacturedSolution");
    cfg.flag_BC_option_inflow_normal_super = 1;
    cfg.flag_BC_option_outflow_riemann_sub = 1;
    static const char *names[6] = {"", "Inflow", "Outflow", "WallInviscid", "WallViscous", "Farfield"};
    char *bc[6] = {cfg.BC_Left, cfg.BC_Right, cfg.BC_Bottom, cfg.BC_Top, cfg.BC_Behind, cfg.BC_InFront};
    int *nb[6] = {&cfg.InterfaceNeighbourLeft, &cfg.InterfaceNeighbourRight, &cfg.InterfaceNeighbourBottom,
                  &cfg.InterfaceNeighbourTop, &cfg.InterfaceNeighbourBehind, &cfg.InterfaceNeighbourInFront};
    /* face f with code 0 is interface f, whose donor is this rank (rank 0) unless argv[3] names
       another; transforms are the identity */
    static int donor[6], tr_rows[6][3];
    static int *tr[6];
    cfg.MPI_rank = 0;
    for (int f = 0; f < 6; f++) {
        const int code = hdr[8 + f];
        strcpy(bc[f], code ? names[code] : "Interface");
        *nb[f] = code ? NO_NEIGHBOUR : f;
        donor[f] = argc > 3 ? atoi(argv[3]) : cfg.MPI_rank;
        for (int d = 0; d < 3; d++) tr_rows[f][d] = d + 1;
        tr[f] = tr_rows[f];
    }
    cfg.MPI_rankNeighbours = donor;
    cfg.TransformMatrixOfInterface = tr;
    compat_define(&cfg);
    const long ntot = (long)cfg.int_iMeshPointsGhostCells * cfg.int_jMeshPointsGhostCells * cfg.int_kMeshPointsGhostCells;
    struct strct_mesh mesh;
    FLT **mf[10] = {&mesh.xi_x, &mesh.xi_y, &mesh.xi_z, &mesh.eta_x, &mesh.eta_y, &mesh.eta_z,
                    &mesh.zeta_x, &mesh.zeta_y, &mesh.zeta_z, &mesh.jacobian};
    for (int m = 0; m < 10; m++) { *mf[m] = (FLT *)malloc(sizeof(FLT) * ntot); rd(*mf[m], sizeof(FLT), ntot, in); }
    struct strct_U U_lastStep, U_RK;
    alloc_U(&U_lastStep, ntot);
    alloc_U(&U_RK, ntot);
    FLT *sf[5] = {U_lastStep.rho, U_lastStep.u, U_lastStep.v, U_lastStep.w, U_lastStep.e};
    for (int m = 0; m < 5; m++) rd(sf[m], sizeof(FLT), ntot, in);
    fclose(in);
    WriteValuesFromU1ToU2(&cfg, &U_lastStep, &U_RK);
    struct strct_Flux Q = {0}, Q_sum = {0}, Flux = {0}, Flux_PlusHalf = {0};
    struct timespec t0, t1, t2;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    t1 = t0;
    int done = 0;
    for (int it = 1; it <= steps; it++) {
        CalcRungeKutta(&cfg, &mesh, &U_lastStep, &U_RK, &Q, &Q_sum, &Flux, &Flux_PlusHalf);
        if (checkNAN(&cfg, &U_RK, &mesh) == 1) { printf("SHOCK: NAN detected at iteration %d\n", it); break; }
        WriteValuesFromU1ToU2(&cfg, &U_RK, &U_lastStep);
        done = it;
        if (it == 1) clock_gettime(CLOCK_MONOTONIC, &t1);
    }
    clock_gettime(CLOCK_MONOTONIC, &t2);
    const double first = 1e3 * ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec));
    const double rest = 1e3 * ((double)(t2.tv_sec - t1.tv_sec) + 1e-9 * (double)(t2.tv_nsec - t1.tv_nsec));
    if (done > 1)
        printf("compat_run: first step %.3f ms, then %.3f ms per step over %d steps\n", first, rest / (done - 1),
               done - 1);
    FILE *out = fopen(argv[2], "wb");
    if (!out) { perror(argv[2]); return 2; }
    for (int m = 0; m < 5; m++) fwrite(sf[m], sizeof(FLT), ntot, out);
    fclose(out);
    compat_release(&cfg);
    printf("compat_run: %d steps done\n", steps);
    return 0;
}
