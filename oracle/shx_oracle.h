/*
 * shx_oracle.h -- CPU oracle for the SHOCK RHS / RK-stage hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference algorithm (mag1000/SHOCK, src/WENO.c, src/ZD.c, src/Functions.c,
 * src/BC.c) used as the parity checker for the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path (shock_amd/, libshockx.so) never links, loads or calls it.
 *
 * Layout follows the reference exactly: ghost-padded SoA arrays of doubles,
 * linear index ijk = i*ng[1]*ng[2] + j*ng[2] + k (k fastest), g = (S+1)/2
 * ghost layers, interior i in [g, g+n-1] (src/Functions.c:56-103).  In 2D
 * ng[2] = 1 and k = 0.
 *
 * Pinning status per function (how each one is checked, tests/test_oracle_*.py):
 *   orc_define            rationals vs exact Taylor-moment solutions; RK tableaux vs textbook
 *   orc_calc_values       closed form (ideal gas, Sutherland)
 *   orc_eigvec            L*R = I and R*diag(lambda)*L = flux Jacobian (exact algebra)
 *   orc_phi_w5/w9         eps -> inf reduces to the linear 5th/9th-order upwind flux
 *                         (coefficients derived by exact rational solve); smoothness
 *                         indicators vs Jiang-Shu / Balsara-Shu integrals (sympy)
 *   orc_weno_dir          free-stream preservation, telescoping conservation,
 *                         mirror symmetry, order of accuracy on smooth data; the LF weight
 *                         J_{i+1/2} by free stream on separably stretched grids; the
 *                         arithmetic-mean eigen-state by 1D data against numpy's eigenvectors
 *                         of the Euler Jacobian (tests/test_oracle_pins.py, with altered copies
 *                         of this file that the pins must reject)
 *   orc_visc_*            polynomial exactness vs analytic Navier-Stokes viscous
 *                         terms, order of accuracy; the face coefficient (mu fac)_i/2 +
 *                         (mu fac)_{i+1}/2 by the locality of the response to one node's mu
 *   orc_rk_*              closed form (linear ODE per node reproduces RK3/RK4 polynomials)
 *   orc_metric_*          exact on affine grids; 2nd-order central differences
 *   orc_ghost_periodic    permutation identity
 *   wall writers          ghost state against an independent linear solve (slip) and sign
 *                         flips (no-slip); quiescent gas and wall-parallel stream stay exact
 *   orc_mirror_metric     copy identity on every ghost node of the slab
 *   farfield writer       copy identity on the interior tangential range
 *   orc_calc_post         Mach closed form; gradRho exact on an anisotropic affine grid for a
 *                         linear density; lambda2 against numpy's eigenvalues of S^2 + Omega^2
 *                         for a linear velocity field on a sheared affine grid
 *   orc_mms_error         brute force in numpy over the same plane and node count
 *   isothermal wall       no-slip ghost with T = T_wall and the Prandtl-0.72 factor; the energy
 *                         term's mean factor checked against a hand-set factor array
 *   subsonic inflow writers  normal: inflow state; Riemann: the far field's u + 2c/(gamma-1) and
 *                         entropy, the mirror's u - 2c/(gamma-1); isentropic: linear solve for
 *                         the velocity, far-field entropy and Bernoulli's h + u^2/2 = h_inflow
 *   orc_cfl_dt, orc_check_nan   brute force
 * The reference itself cannot be compiled here (it needs mpi.h and cgnslib.h,
 * which the image lacks), so no compiled-reference fixture exists.
 */
#ifndef SHX_ORACLE_H
#define SHX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* quirk bits (reference behaviours that a "fixed" mode may switch off) */
#define ORC_QUIRK_ZETA_VISC_J   1u  /* src/ZD.c:547-548: zeta viscous stencil offset uses j, not k */
#define ORC_QUIRK_METRIC_DIAG3D 2u  /* src/Functions.c:6479-6484: 3D metric drops off-diagonal terms */
#define ORC_QUIRKS_REFERENCE (ORC_QUIRK_ZETA_VISC_J | ORC_QUIRK_METRIC_DIAG3D)

/* boundary kinds per face (order: left, right, bottom, top, behind, in front = i-, i+, j-, j+, k-, k+) */
#define ORC_BC_PERIODIC 0          /* self-interface (src/Functions.c:3580-3684) */
#define ORC_BC_MMS 1               /* manufactured-solution Dirichlet (src/ManufacturedSolution.c:202-338) */
#define ORC_BC_INFLOW_SUPERSONIC 2 /* src/BC.c:2505-2542 */
#define ORC_BC_OUTFLOW_RIEMANN 3   /* src/BC.c:2572-2622 */
#define ORC_BC_OUTFLOW_NORMAL 4    /* src/BC.c:2544-2570: p = p_out, rho, e, c from the mirror node */
#define ORC_BC_OUTFLOW_RUDY 5      /* src/BC.c:2624-2655: non-reflecting relaxation towards p_out */
#define ORC_BC_WALL_SLIP 6         /* WallInviscid, src/BC.c:2688-2719 (face loops :988-1172) */
#define ORC_BC_WALL_NOSLIP 7       /* WallViscous, src/BC.c:2657-2686 (face loops :804-986) */
#define ORC_BC_INFLOW_SUB_NORMAL 8     /* src/BC.c:2457-2503: inflow state, T and mu from the mirror */
#define ORC_BC_INFLOW_SUB_RIEMANN 9    /* src/BC.c:2404-2455: Riemann invariants, far-field entropy */
#define ORC_BC_INFLOW_SUB_ISENTROPIC 10 /* src/BC.c:2363-2402: mirror theta, isentropic Bernoulli density */
#define ORC_BC_FARFIELD 11         /* src/BC.c:2817-2852 (face loops :638-803): the mirror node copied */
#define ORC_BC_WALL_ISOTHERMAL 12  /* WallViscousIsothermal, src/BC.c:2943-2978 (loops :2980-3151) */

typedef struct orc_cfg {
    /* inputs */
    int dims;          /* 2 or 3 (MESHDIMENSIONS) */
    int S;             /* 5 or 9 (SPACEORDER) */
    int n[3];          /* interior points per direction (n[2] = 1 in 2D) */
    int time_order;    /* 3 (SSP-RK3) or 4 (classic RK4) */
    int inviscid;      /* options:Inviscid */
    int mms;           /* options:ManufacturedSolution: mu = 1, sources added */
    unsigned quirks;   /* ORC_QUIRK_* */
    double gamma, mach, reynolds, prandtl, T0;
    double eps_override; /* 0 -> reference epsilon 10*sqrt(DBL_MIN) */
    int bc[6];           /* ORC_BC_* per face; both faces of a direction periodic or neither */
    double rho_in, u_in, v_in, w_in, p_in;   /* boundaryconditions:{rho,p}_inflow, InitializeValues u0 (AoA) */
    double p_out;        /* boundaryconditions:p_outflow (src/Import.c:94) */
    double alpha_nonref; /* AlphaNonRef = 0.25 L0/u0 (src/Functions.c:1763) */
    /* state at the start of the step, read by the Rudy writer (src/BC.c:2643-2646): the step's tau,
       U_lastStep's stored p, rho and u, and U_lastStep's c (never recomputed after the
       initialisation: CalcValues only runs on U_RK).  Unused unless a face is ORC_BC_OUTFLOW_RUDY. */
    double tau_step;
    const double *p_last, *rho_last, *u_last, *c_last;
    /* BC_Corrector_{xi,eta,zeta}Momentum (src/Functions.c:1128-1130): padded arrays, 1 everywhere
       after check_Metric (src/Functions.c:6212-6214), set on wall ghosts by the wall writers and
       multiplying the convective momentum fluxes (src/WENO.c:265-277).  NULL: 1 everywhere. */
    double *corr[3];
    /* boundaryconditions:T_wall (src/Import.c:96) and the per-node heat-conduction factor
       pnt_config->Gamma (src/Functions.c:141-152; the isothermal writer sets it on its ghosts,
       src/BC.c:2974-2976).  gam NULL: gam_heat everywhere. */
    double T_wall;
    double *gam;
    /* derived by orc_define */
    int g;
    int ng[3];
    long ntot;
    double upsilon, psi, gam_heat, suth, eps;
    double rk_a[4], rk_b[4], rk_q[4], rk_f[4], rk_m[4];
    double zd_mid[10];   /* midpoint-derivative coefficients, S+1 used */
    double zd_int[10];   /* midpoint-interpolation coefficients, S+1 used */
    double zd_der[11];   /* nodal first-derivative coefficients, S+2 used */
} orc_cfg;

/* Field bundles are arrays of padded FP64 arrays:
 *   prim[5]    = rho, u, v, w, e          (stored state, strct_U)
 *   der[8]     = p, T, c, mu, theta1, theta2, theta3, (unused)
 *   met[10]    = xi_x, xi_y, xi_z, eta_x, eta_y, eta_z, zeta_x, zeta_y, zeta_z, J
 *   q[5]       = Mass, xiMomentum, etaMomentum, zetaMomentum, Energy (strct_Flux)
 *   grad[12]   = u_xi,u_eta,u_zeta, v_xi,v_eta,v_zeta, w_xi,w_eta,w_zeta, T_xi,T_eta,T_zeta
 */

int  orc_define(orc_cfg *c);
long orc_idx(const orc_cfg *c, int i, int j, int k);

void orc_calc_values(const orc_cfg *c, double *const met[10], double *const prim[5], double *const der[8]);

void orc_eigvec(const double nh[3], double u, double v, double w, double rho, double cs,
                double gamma, double L[25], double R[25]);
double orc_phi_w5(double eps, const double *f, const double *d);
double orc_phi_w9(double eps, const double *f, const double *d);
/* smoothness indicators as used inside orc_phi_w5 (from 4 differences) and orc_phi_w9 (9 values) */
void orc_is_w5(const double *d, double is[3]);
void orc_is_w9(const double *f, double is[5]);

void orc_weno_dir(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                  double *const der[8], double *const q[5]);
/* single-interface probe used by the tests: interface between node (i,j,k) and its +1 neighbour */
void orc_weno_interface(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                        double *const der[8], int i, int j, int k, double out[5]);

/* pointwise flux of node (i,j,k) in direction dir, as the WENO sweep forms it */
void orc_point_flux_at(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                       double *const der[8], int i, int j, int k, double F[5]);

void orc_visc_fac(int dir, const double m[10], double fac[30]);
void orc_visc_derivs(const orc_cfg *c, double *const prim[5], double *const der[8], double *const grad[12]);
void orc_visc_dir(const orc_cfg *c, int dir, double *const met[10], double *const prim[5],
                  double *const der[8], double *const grad[12], double *const q[5]);

/* mirrorGhostCellsMetric (src/Functions.c:6849-7014): on wall and farfield faces, the metric of every ghost
 * node of the face slab (full tangential range) is copied from its mirror node, faces in order. */
void orc_mirror_metric(const orc_cfg *c, double *const met[10]);

/* CalcValuesForPost with CalcLambda2 (src/Functions.c:3687-3761, 5846-6070) on interior nodes;
 * p is the stored pressure of U_lastStep; out = Mach, gradRho, Lambda2 (3D only). */
void orc_calc_post(const orc_cfg *c, double *const met[10], double *const prim[5], const double *p,
                   double *const out[3]);

/* ErrorManufacturedSolution's norms (src/ManufacturedSolution.c:745-790): out = L2(rho),
 * Linf(rho), L2(p), Linf(p) over the last k plane, p the stored pressure. */
void orc_mms_error(const orc_cfg *c, double *const prim[5], const double *p, const double *ex_rho,
                   const double *ex_p, double out[4]);

/* Thread count of the OpenMP loops (timing helper; results do not depend on it). */
void orc_set_threads(int n);
void orc_zero(const orc_cfg *c, double *const f[], int nf);
void orc_ghost_periodic(const orc_cfg *c, double *const f[], int nf);
/* Face boundary writers in the reference's order (SetAllBoundaryConditions, src/BC.c:13-74),
 * applied after CalcValues: they overwrite primitive and derived values of ghost nodes.
 * bstate: boundary state (rho,u,v,w,e,p on the padded grid) read by ORC_BC_MMS faces (may be NULL
 * otherwise). */
void orc_apply_bc(const orc_cfg *c, double *const met[10], double *const prim[5], double *const der[8],
                  double *const bstate[6]);
void orc_rk_update(const orc_cfg *c, int s, double tau, double *const met[10],
                   double *const un[5], double *const urk[5], double *const q[5], double *const qsum[5]);
void orc_add_source(const orc_cfg *c, double *const met[10], double *const src[5], double *const q[5]);

/* One full time step (CalcRungeKutta, src/Functions.c:2034-2390) for an all-periodic block.
 * scratch: der[8] + grad[12] + q[5] + qsum[5] = 30 padded arrays supplied by the caller.
 * src may be NULL (no manufactured source). */
void orc_step(const orc_cfg *c, double tau, double *const met[10], double *const un[5],
              double *const urk[5], double *const scratch[30], double *const src[5]);

/* Full RHS of one stage without the update: Q for state prim (ghosts refilled periodically). */
void orc_rhs(const orc_cfg *c, double *const met[10], double *const prim[5],
             double *const scratch[30], double *const src[5]);
/* variants with a boundary state for ORC_BC_MMS faces */
void orc_rhs_b(const orc_cfg *c, double *const met[10], double *const prim[5],
               double *const scratch[30], double *const src[5], double *const bstate[6]);
void orc_step_b(const orc_cfg *c, double tau, double *const met[10], double *const un[5],
                double *const urk[5], double *const scratch[30], double *const src[5],
                double *const bstate[6]);

int    orc_check_nan(const orc_cfg *c, double *const prim[5]);
double orc_cfl_dt(const orc_cfg *c, double cfl, double *const met[10], double *const prim[5]);

/* Metric from the extended coordinate arrays (extent ng+2 per direction, 1 in 2D z),
 * CreateMetric src/Functions.c:6388-6509.  Writes all padded nodes. */
void orc_metric_from_coords(const orc_cfg *c, const double *xe, const double *ye,
                            const double *ze, double *const met[10]);

/* check_Metric (src/Functions.c:6198-6243): number of padded nodes with a negative Jacobian. */
int orc_check_metric(const orc_cfg *c, double *const met[10]);

#ifdef __cplusplus
}
#endif
#endif
