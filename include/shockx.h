/*
 * shockx.h -- C ABI of libshockx.so, the MI355X (gfx950) implementation of SHOCK's
 * right-hand-side and time-step hot path.
 *
 * Drop-in boundary.  SHOCK's driver calls the path through plain C functions that take
 * its global structs (src/SHOCK.h:69-642); those structs embed MPI and CGNS types, so this
 * library exposes the same operations over an opaque handle and plain pointers instead.
 * Each entry point below names the reference function it replaces.  Field arrays use the
 * reference's own layout: ghost-padded FP64 SoA, linear index i*ng1*ng2 + j*ng2 + k
 * (k fastest), ng = n + S + 1 per direction, ng2 = 1 in 2D (src/Functions.c:56-103).
 *
 * Conventions: every function returns SHX_OK (0) or a negative SHX_ERR_* code and never
 * aborts; shx_last_error() gives the message.  All work is enqueued on the handle's own
 * HIP stream; functions that return host data synchronize that stream.  A handle is not
 * thread-safe (the reference is single-threaded per rank as well).
 */
#ifndef SHOCKX_H
#define SHOCKX_H

#ifdef __cplusplus
extern "C" {
#endif

#define SHX_OK               0
#define SHX_ERR_ARG         -1  /* invalid argument or configuration */
#define SHX_ERR_HIP         -2  /* HIP runtime error (message in shx_last_error) */
#define SHX_ERR_UNSUPPORTED -3  /* configuration not implemented */
#define SHX_ERR_STATE       -4  /* call out of order (e.g. stepping before a metric is set) */
#define SHX_ERR_NAN         -5  /* shx_run / shx_run_cfl: checkNAN found a non-finite density after a
                                   step; the case holds that step's state, marked as checkNAN marks it */

/* Reference behaviours that can be switched off for verification runs. */
#define SHX_QUIRK_ZETA_VISC_J   1u /* src/ZD.c:547-548: zeta viscous stencil k offset taken from j */
#define SHX_QUIRK_METRIC_DIAG3D 2u /* src/Functions.c:6479-6484: 3D metric keeps only x_xi, y_eta, z_zeta */
#define SHX_QUIRKS_REFERENCE (SHX_QUIRK_ZETA_VISC_J | SHX_QUIRK_METRIC_DIAG3D)

/* Boundary kinds per face (left, right, bottom, top, behind, in front = i-, i+, j-, j+, k-, k+).
 * The two faces of a direction are both periodic or both not.  Face writers run after the
 * periodic transfer and CalcValues, in face order (SetAllBoundaryConditions, src/BC.c:13-74). */
#define SHX_BC_PERIODIC           0 /* self-interface, src/Functions.c:3580-3684 */
#define SHX_BC_MANUFACTURED       1 /* exact state on the face's ghost layers, interior tangential
                                       range (WriteBCManufacturedSolution*, src/ManufacturedSolution.c:202-338) */
#define SHX_BC_INFLOW_SUPERSONIC  2 /* inflow state on every ghost node of the face slab
                                       (WriteInflowSupersonicNormalBoundary, src/BC.c:2505-2542) */
#define SHX_BC_OUTFLOW_RIEMANN    3 /* characteristic outflow from the mirror node
                                       (WriteOutflowSubsonicRiemannBoundary, src/BC.c:2572-2622) */
#define SHX_BC_OUTFLOW_NORMAL     4 /* p = p_out; rho, e, c, contravariant velocity from the mirror
                                       node (WriteOutflowSubsonicNormalBoundary, src/BC.c:2544-2570) */
#define SHX_BC_OUTFLOW_RUDY       5 /* non-reflecting: ghost p relaxed from the step-start ghost state
                                       towards p_out (WriteOutflowSubsonicRudyBoundary, src/BC.c:2624-2655).
                                       The step-start sound speed is the one of the state uploaded
                                       last: the reference never recomputes U_lastStep's c. */
#define SHX_BC_WALL_SLIP          6 /* WallInviscid: the mirror's contravariant velocity normal to
                                       the face reversed; rho, e, p, c, T, mu from the mirror node;
                                       interior tangential range (WriteWallSlipBoundary,
                                       src/BC.c:2688-2719, face loops :988-1172) */
#define SHX_BC_WALL_NOSLIP        7 /* WallViscous: velocities reversed; rho, e, p, c, T, mu from the
                                       mirror node; interior tangential range
                                       (WriteWallNoSlipBoundary, src/BC.c:2657-2686, loops :804-986).
                                       On the ghost nodes of both wall kinds the WENO sweeps apply the
                                       reference's BC_Corrector to the convective momentum terms
                                       (src/WENO.c:265-277), and shx_build_metric mirrors the ghost
                                       metric of the face slab (src/Functions.c:6849-7014). */
#define SHX_BC_INFLOW_SUB_NORMAL  8 /* subsonic inflow: inflow velocity, rho and p; T, mu from the
                                       mirror (WriteInflowSubsonicNormalBoundary, src/BC.c:2457-2503) */
#define SHX_BC_INFLOW_SUB_RIEMANN 9 /* subsonic inflow from the far field's incoming invariant and
                                       entropy and the mirror's outgoing invariant
                                       (WriteInflowSubsonicRiemannBoundary, src/BC.c:2404-2455) */
#define SHX_BC_INFLOW_SUB_ISENTROPIC 10 /* subsonic inflow: the mirror's contravariant velocity, rho
                                       from Bernoulli along the far field's isentrope
                                       (WriteInflowSubsonicIsentropBoundary, src/BC.c:2363-2402).
                                       All three cover every ghost node of the face slab. */
#define SHX_BC_WALL_ISOTHERMAL   12 /* WallViscousIsothermal: as the no-slip wall with T = T_wall on
                                       the ghosts, whose heat-conduction factor is the one of
                                       Prandtl 0.72; the viscous energy term of a node takes the
                                       mean factor of the node and its upper neighbour, so the last
                                       node below an upper isothermal face sees it
                                       (WriteWallNoSlipIsothermalBoundary, src/BC.c:2943-2978;
                                       src/ZD.c:193-194, 422-423, 617-618) */
#define SHX_BC_INTERFACE         13 /* face joined to a face of another block of the same case on
                                       this device (a zone interface of the reference's division,
                                       src/Load.c:191-250); its ghost layers are gathered from that
                                       block before every stage by shx_blocks_run
                                       (TransferFlowParameterWithGhosts, src/Functions.c:3580-3684) */
#define SHX_BC_FARFIELD          11 /* every value of the mirror node, interior tangential range;
                                       the ghost metric is mirrored as at walls
                                       (WriteFarfieldBoundary, src/BC.c:2817-2852, loops :638-803) */

/* Arithmetic forms. */
#define SHX_FP64 0   /* every operation in FP64 (the reference's PRECISION=2) */
#define SHX_MIXED 1  /* FP32 WENO nonlinear part (projected differences, smoothness indicators,
                        weights, deviation); state, fluxes, central part, viscous terms, RK
                        update, metric and CFL in FP64 */

/* Case description: the subset of strct_configuration (src/SHOCK.h:69-544) the path reads. */
typedef struct shx_desc {
    int dims;          /* MESHDIMENSIONS: 2 or 3 */
    int space_order;   /* SPACEORDER: 5 (WENO5) or 9 (WENO9) */
    int time_order;    /* numerics:TimeOrder: 3 (SSP-RK3) or 4 (RK4) */
    int n[3];          /* interior points (zonesize); n[2] ignored in 2D */
    double gamma;      /* fluidproperties:Gamma */
    double mach;       /* fluidproperties:Mach */
    double reynolds;   /* fluidproperties:Reynolds */
    double prandtl;    /* fluidproperties:Prandtl */
    double T0;         /* fluidproperties:T, Sutherland constant = 100.4 / T0 */
    int inviscid;      /* options:Inviscid */
    int manufactured;  /* options:ManufacturedSolution (mu = 1, source added each stage) */
    unsigned quirks;   /* SHX_QUIRK_* bits; SHX_QUIRKS_REFERENCE reproduces the reference */
    int bc[6];         /* SHX_BC_* per face */
    int precision;     /* SHX_FP64 or SHX_MIXED */
    double eps_override; /* 0: WENO epsilon = 10*sqrt(DBL_MIN) as src/Functions.c:1829 */
    /* inflow state (boundaryconditions:rho_inflow, p_inflow; u,v,w_inflow from the angle of
     * attack, src/Import.c:90-91,233-235); the Riemann outflow uses u_inflow as its far field */
    double rho_inflow, u_inflow, v_inflow, w_inflow, p_inflow;
    /* subsonic outflow (SHX_BC_OUTFLOW_NORMAL / _RUDY): boundaryconditions:p_outflow
     * (src/Import.c:94) and the Rudy relaxation rate AlphaNonRef = 0.25 L0/u0
     * (src/Functions.c:1763), multiplied by the step's tau as in src/BC.c:2643-2646 */
    double p_out, alpha_nonref;
    /* isothermal wall (SHX_BC_WALL_ISOTHERMAL): boundaryconditions:T_wall (src/Import.c:96) */
    double T_wall;
} shx_desc;

typedef struct shx_case shx_case;

const char *shx_last_error(void);
int shx_device_count(int *count);

/* Allocates the device arrays of one block (replaces the path's part of AllocMemory,
 * src/Functions.c:900-1290, and DefineParameters, src/Functions.c:1758-2032). */
int shx_create(const shx_desc *desc, shx_case **out);
int shx_destroy(shx_case *c);
int shx_padded_extent(const shx_case *c, int ng[3]);

/* Geometry (once per run).  shx_build_metric: CreateMetric (src/Functions.c:6388-6509) on the
 * extended coordinate arrays (extent ng+2 per direction, 1 in 2D z; the reference's
 * x_extrapolate), followed by the periodic ghost transfer of the metric
 * (TransferMeshParameter, src/Functions.c:3481-3578) and, on wall and farfield faces, the mirrored ghost
 * metric (mirrorGhostCellsMetric, src/Functions.c:6849-7014).  shx_upload_metric takes metric
 * arrays a caller already has (strct_mesh xi_x ... zeta_z, jacobian, mirrored at walls), order
 * xi_x,xi_y,xi_z,eta_x,eta_y,eta_z,zeta_x,zeta_y,zeta_z,J. */
int shx_build_metric(shx_case *c, const double *xe, const double *ye, const double *ze);
int shx_upload_metric(shx_case *c, const double *const met[10]);
int shx_download_metric(shx_case *c, double *const met[10]);

/* State: primitive rho, u, v, w, e (strct_U, src/SHOCK.h:590-620).  Upload sets U_lastStep
 * (and therefore U_RK at the next step start); download returns U_lastStep.  Field transfers
 * (state, metric, source, boundary state, post-processing fields, residual) go through two
 * pinned staging buffers of up to 8 MB the handle allocates on its first transfer (the copies
 * between them and the caller's arrays run on up to four short-lived host threads); an upload
 * returns once the caller's arrays have been read. */
int shx_upload_state(shx_case *c, const double *const prim[5]);
int shx_download_state(shx_case *c, double *const prim[5]);

/* Exact state for SHX_BC_MANUFACTURED faces: rho, u, v, w, e, p on the padded grid (the
 * values WriteManufacturedSolution evaluates, src/ManufacturedSolution.c:566-648; only ghost
 * nodes of manufactured faces are read).  Required before stepping when such a face exists. */
int shx_upload_boundary_state(shx_case *c, const double *const bs[6]);

/* Manufactured-solution source per node (Q += J * src each stage,
 * AddManufacturedSolutionSource src/ManufacturedSolution.c:25-175); NULL clears it. */
int shx_upload_source(shx_case *c, const double *const src[5]);

/* One time step with fixed tau: CalcRungeKutta (src/Functions.c:2034-2390) followed by the
 * step-end copy U_lastStep <- U_RK (WriteValuesFromU1ToU2, src/SHOCK.c:663-666). */
int shx_calc_runge_kutta(shx_case *c, double tau);
/* nsteps of the above with checkNAN after every step (the startSimulation loop body,
 * src/SHOCK.c:557-666, without I/O; checkNAN src/Functions.c:5419-5467).  The check is fused into
 * the last stage's update: non-finite interior densities are counted and those nodes get
 * rho = u = v = w (and the stored p) = 1e6.  The first step that finds one ends the run with
 * SHX_ERR_NAN (the reference's rollback is shx_simulate's).  The stream is synchronized after
 * every step, as the reference's loop reads its flag every step. */
int shx_run(shx_case *c, double tau, int nsteps);
/* As shx_run with the time step of every step from the CFL condition on the state it starts from
 * (shx_cfl_dt's formula, new: the reference uses a fixed tau): the first from the current state,
 * later ones from the maximum the previous step's last stage reduced while writing the state, so
 * no separate reduction pass runs between steps.  *time (may be NULL) receives the sum of the
 * steps taken.  Not for cases with Rudy outflow faces (SHX_ERR_UNSUPPORTED). */
int shx_run_cfl(shx_case *c, double cfl, int nsteps, double *time);

/* startSimulation's iteration loop with its NaN rollback and time-step control
 * (src/SHOCK.c:531-691, without sampling and I/O).  Per iteration: time += tau L0/u0, one step,
 * checkNAN on the new state.  On NaN: stop if number_resets rollbacks were taken (the case then
 * holds the failed step's state, NaN nodes marked as checkNAN marks them); otherwise
 * tau *= decel_factor, distance_forward restarts, and the older of two device-resident backups
 * (state, stored p, iteration and time) is reloaded.  Without NaN and with the accelerator on, the
 * tau resetter (back to the starting tau after distance_nan steps' worth of time) and the tau
 * accelerator (tau *= accel_factor after distance_forward steps' worth, which then quadruples) act
 * as in the reference.  Backups: slot 1 every distance_backward iterations, slot 2 half-way
 * between; slot 1 starts as the initial state at iteration 1, slot 2 empty (iteration 999999). */
typedef struct shx_sim {
    int start_iteration, end_iteration;   /* int_StartIteration, int_EndIteration */
    double tau;                           /* numerics:tau (numericalTauStart) */
    double time, time_last_action;        /* time_dim, time_dim_lastAction at the start */
    double L0, u0;                        /* L0_dim, u0_dim (time_dim += tau L0 / u0) */
    int accelerator;                      /* general:TauAccelerator */
    double accel_factor, decel_factor;    /* tau:tauaccelerator_factor, tau:taudecelerator_factor */
    int number_resets;                    /* tau:NumberResets */
    int distance_nan, distance_forward, distance_backward;  /* tau:DistanceNAN, DistanceForward, DistanceBackward */
    /* sampling into the film (shx_film_open first): int_Samples, IterationsBetweenSamples,
       StartSampling; samples = 0: none.  Each sample is CalcValuesForPost + WriteValuesFromUToFilm
       of the step's state; a rollback rewinds the sample index with the state. */
    int samples, iterations_between_samples, start_sampling;
    /* probe histories (WriteValuesForPressureHistory / WriteValuesForVelocityHistory,
       src/Functions.c:5381-5417): the stored p and u, v, w of nprobes interior nodes (padded
       linear indices) every iteration, kept on the device under index iteration - (start + 1),
       which a rollback rewinds; read them with shx_history_download.  nprobes = 0: none. */
    const long *probes;
    int nprobes;
} shx_sim;
typedef struct shx_sim_report {
    int iteration;           /* last iteration the loop ran (the failing one after a stop) */
    double time, tau;        /* time_dim and numericalTau at the end */
    int resets;              /* rollbacks taken */
    int nan_steps;           /* steps whose checkNAN found a NaN or inf density */
    int accelerations, tau_resets;
    int aborted;             /* 1: stopped at number_resets */
    int samples;             /* samples in the film at the end */
} shx_sim_report;
int shx_simulate(shx_case *c, const shx_sim *sim, shx_sim_report *report);

/* Post-processing on the device.  shx_post: CalcValuesForPost with CalcLambda2
 * (src/Functions.c:3687-3761, 5846-6070) on the interior nodes of the current state: Mach number,
 * density gradient and lambda2 (3D) into host arrays on the padded grid (NULL entries skipped).
 * The Mach number uses U_lastStep's stored pressure, which the reference's RK update leaves one
 * stage behind the state; the library keeps that pressure from the first shx_post or
 * shx_film_open call on (from the start for cases with subsonic outflow or slip wall faces), so
 * call one of them before stepping for the reference's values, and before shx_batch_create for a
 * case that a batch will advance (a batch records its launches when it is created).
 * The film keeps samples on the device (WriteValuesFromUToFilm, src/Functions.c:2510-2553):
 * shx_film_open allocates room for `samples`, shx_film_sample appends one sample of the current
 * state with its time, shx_film_download copies sample s out as rho, u, v, w, p, gradRho, Mach,
 * Lambda2 over the interior nodes (n0 x n1 x n2, k fastest). */
int shx_post(shx_case *c, double *const out[3]);
int shx_film_open(shx_case *c, int samples);
int shx_film_sample(shx_case *c, double time);
int shx_film_count(const shx_case *c, int *count);
int shx_film_download(shx_case *c, int sample, double *const out[8], double *time);
/* probe histories of the last shx_simulate: time[iterations], then p, u, v, w each
 * [nprobes][iterations] (iterations = end - start); NULL entries skipped */
int shx_history_download(shx_case *c, double *time, double *p, double *u, double *v, double *w);
/* ErrorManufacturedSolution (src/ManufacturedSolution.c:745-790) of the current state against the
 * exact state of shx_upload_boundary_state: out = L2(rho), Linf(rho), L2(p), Linf(p), as the
 * reference forms them -- over the interior nodes of the last k plane only, each squared deviation
 * divided by the full interior node count, p the stored pressure (as in shx_post).  FP64 sums where
 * the reference accumulates in long double, against exact values rounded to FP64. */
int shx_mms_error(shx_case *c, double out[4]);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): shx_mms_error with its
 * partial sums in a static device array of the library (256 KiB per device, held until the
 * process ends) instead of per-call allocations, so the call allocates nothing.  Limit: at most
 * 4096 * 256 = 1048576 interior nodes per k plane (n0 * n1), else SHX_ERR_ARG.  The case must
 * already keep its stored pressure (after one shx_mms_error, shx_post or shx_film_open call, or
 * from creation for cases with subsonic outflow or slip wall faces), else SHX_ERR_STATE.  Every
 * call on a device shares that device's array: calls on one device must not overlap (one host
 * thread, or streams ordered against each other). */
int shx_mms_error_static(shx_case *c, double out[4]);

/* Stage right-hand side of the current U_lastStep (ghost fill, CalcValues, all WENO and
 * viscous fluxes and the source) returned in q[5] (strct_Flux Q after one stage's flux
 * assembly, src/Functions.c:2111-2293).  For parity tests. */
int shx_rhs(shx_case *c, double *const q[5]);

/* check_Metric (src/Functions.c:6198-6243): number of padded nodes whose Jacobian is negative
 * (the reference prints it and continues; a left-handed grid or crossing ghost cells). */
int shx_check_metric(shx_case *c, int *negative);

/* checkNAN (src/Functions.c:5419-5467) on U_lastStep: count of non-finite interior rho;
 * those nodes get rho = u = v = w = 1e6 as in the reference. */
int shx_check_nan(shx_case *c, int *count);

/* CFL time step (new; the reference uses a fixed tau): dt = cfl / max over interior nodes of
 * sum_d (|theta_d| + c |grad xi_d|), computed on U_lastStep. */
int shx_cfl_dt(shx_case *c, double cfl, double *dt);

int shx_synchronize(shx_case *c);

/* Autotuner (no reference counterpart): times the WENO sweep tile variants on the current state
 * (4 x 64 or 2 x 128 along the contiguous axis, 16 x 16 or 8 x 32 otherwise) and keeps the fastest
 * per direction; choice[3] (may be NULL) receives the variant indices.  The state is unchanged. */
int shx_tune(shx_case *c, int reps, int choice[3]);

/* Measurement aid (no reference counterpart): `reps` launches of a plain device copy of one
 * padded field (8 * ng0*ng1*ng2 bytes read and written per launch), used to calibrate
 * hardware byte counters against a kernel of known traffic. */
int shx_copy_probe(shx_case *c, int reps);

/* Per-kernel timing with HIP events on the handle's stream.  When enabled, every launch is
 * bracketed by events; shx_kernel_times synchronizes and returns, per kernel class, the total
 * milliseconds and the launch count since the last reset.  names[k] points to static strings. */
#define SHX_NUM_KERNELS 16
int shx_set_timing(shx_case *c, int enable);
int shx_kernel_times(shx_case *c, const char *names[SHX_NUM_KERNELS], double ms[SHX_NUM_KERNELS],
                     int launches[SHX_NUM_KERNELS]);

/* Multi-block cases on one device: the reference's division of a case into zones, one MPI rank
 * per zone with a halo transfer per stage (src/Load.c:191-250, src/Functions.c:3580-3684), here with
 * every zone a block (a shx_case) on the same GPU and the halo an on-device gather.  A connection
 * joins face `face` of block `block` to face `donor_face` of block `donor` with the CGNS transform
 * of the interface (TransformMatrixOfInterface, src/SHOCK.h:285; src/Functions.c:3902-4287): index
 * direction d of the block runs along the donor's direction |transform[d]| - 1, reversed where it
 * is negative; across the faces the layers count outwards on one side and inwards on the other
 * (the reference derives that sign from the point ranges, src/Functions.c:3902-3990).  Faces join
 * whole (tangential extents equal under the transform); both sides of a connection are listed; a
 * block may join itself (a periodic pair with a transform).  Velocities are Cartesian, so the
 * transfer moves values without rotating them (src/Functions.c:2809-2863); the ghost metric takes
 * the donor's rows permuted and signed by the transform (src/Functions.c:2640-2690).  Ghost nodes
 * beyond two or three faces (edges, corners) get the value the reference's repeated transfers
 * settle to: the composition of the connections and periodic wraps of the faces they lie beyond.
 * Limits: at most SHX_BLOCKS_MAX blocks; all blocks share dims, scheme, RK order and viscosity. */
#define SHX_BLOCKS_MAX 8
typedef struct shx_interface {
    int block, face;        /* receiving block (index into the block array) and its face 0..5 */
    int donor, donor_face;  /* the block across the face and its face */
    int transform[3];       /* transform[2] ignored in 2D */
} shx_interface;
typedef struct shx_blocks shx_blocks;
int shx_blocks_create(shx_case *const *blocks, int nblocks, const shx_interface *ifc, int nifc,
                      shx_blocks **out);
/* ghost metric of the interface faces from the donors (TransferMeshParameter,
 * src/Functions.c:3481-3578): after every block's shx_build_metric or shx_upload_metric */
int shx_blocks_transfer_metric(shx_blocks *g);
/* nsteps steps of fixed tau: per stage the interface halo of every block, then each block's stage
 * (blocks on their own streams, ordered by events); checkNAN over every block after each step
 * (SHX_ERR_NAN as shx_run).  While a block belongs to a group it cannot be stepped alone. */
int shx_blocks_run(shx_blocks *g, double tau, int nsteps);
int shx_blocks_destroy(shx_blocks *g);

/* Batches (no reference counterpart: SHOCK runs one case per job).  A batch advances many
 * independent cases, each created, uploaded and owned as above (any face kinds), with one launch
 * per kernel class and group of like cases (same dims, scheme, metric kind, precision, viscosity,
 * RK order, WENO tiles) per stage over a tile work list, replayed from a captured HIP graph; the
 * face writers run as one launch per face index.  tau holds one
 * time step per case.  The cases keep their handles: their state can be read or replaced between
 * shx_batch_run calls (work on a case's stream is ordered after the batch).  A case belongs to at
 * most one batch; while it does, shx_calc_runge_kutta, shx_run and shx_simulate on it (which
 * would move its state to another buffer than the batch's captured step uses) and shx_destroy
 * return SHX_ERR_STATE.  Destroy the batch before its cases. */
typedef struct shx_batch shx_batch;
int shx_batch_create(shx_case *const *cases, int ncases, const double *tau, shx_batch **out);
int shx_batch_set_tau(shx_batch *b, const double *tau);
int shx_batch_run(shx_batch *b, int nsteps);
/* Picks the resident blocks per CU of each kernel class (ghost, values, grad, weno, visc, rk) by
 * timing replayed steps; the case states are restored.  factors_out may be NULL. */
int shx_batch_tune(shx_batch *b, int steps, int factors_out[6]);
/* Waits for the batch's steps.  checkNAN (src/Functions.c:5419-5467) is fused into every case's
 * last stage as in shx_run, but read here rather than after every step, so that shx_batch_run
 * returns at once: SHX_ERR_NAN when any case's new state held a non-finite density in the steps
 * since the last call (those nodes hold 1e6; the message names the cases). */
int shx_batch_synchronize(shx_batch *b);
/* Waits for the batch and writes, per case (in creation order), the running count of non-finite
 * densities its fused checkNAN has found. */
int shx_batch_nan_counts(shx_batch *b, int *counts);
int shx_batch_destroy(shx_batch *b);

#ifdef __cplusplus
}
#endif
#endif
