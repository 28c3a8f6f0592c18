/*
 * shock_compat.h -- the reference's entry points of the RHS / time-step path under their own
 * names, argument order and meaning, over restated structs that hold only the fields the path
 * reads (field names as in src/SHOCK.h).  The work runs in libshockx.so on the GPU: every
 * CalcRungeKutta call uploads U_lastStep, takes one step on the device and writes U_RK back, so
 * a caller written against the reference's calls runs unchanged (at PCIe cost per step; callers
 * that keep the state on the device use shockx.h directly).
 *
 *   CalcRungeKutta         src/Functions.h:206-214 (called at src/SHOCK.c:565)
 *   checkNAN               src/Functions.h:361-364
 *   WriteValuesFromU1ToU2  src/Functions.h:216-219
 *
 * The reference's compile-time MESHDIMENSIONS and SPACEORDER are fields of the configuration
 * here; compat_define() fills the ghost-padded extents DefineParameters derives
 * (src/Functions.c:56-103).  Errors in the configuration print a message and end the process
 * with status 4242, where the reference calls MPI_Abort.
 */
#ifndef SHOCK_COMPAT_H
#define SHOCK_COMPAT_H

#ifdef __cplusplus
extern "C" {
#endif

#define NO_NEIGHBOUR -1
typedef double FLT;

struct strct_configuration {
    int MESHDIMENSIONS;                       /* 2 or 3 */
    int SPACEORDER;                           /* 5 or 9 */
    int int_iMeshPoints, int_jMeshPoints, int_kMeshPoints;
    int int_iMeshPointsGhostCells, int_jMeshPointsGhostCells, int_kMeshPointsGhostCells;  /* compat_define */
    int int_iStartReal, int_iEndReal, int_jStartReal, int_jEndReal, int_kStartReal, int_kEndReal;
    int int_TimeOrder;
    FLT numericalTau;
    FLT machNumber, reynoldsNumber, prandtlNumber, gammaNumber, T0_dim;
    int flag_Inviscid, flag_ManufacturedSolution, flag_NAN;
    char BC_Left[30], BC_Right[30], BC_Top[30], BC_Bottom[30], BC_Behind[30], BC_InFront[30];
    char BCFarfield[30], BCInflow[30], BCOutflow[30], BCOutflowSubsonic[30], BCWallInviscid[30];
    char BCWallViscous[30], BCInflowSupersonic[30], BCInflowSubsonic[30], BCWallViscousIsothermal[30];
    char BCManufacturedSolution[30];
    int flag_BC_option_inflow_normal_sub, flag_BC_option_inflow_riemann_sub;
    int flag_BC_option_inflow_isentrop_sub, flag_BC_option_inflow_normal_super;
    int flag_BC_option_outflow_normal_sub, flag_BC_option_outflow_riemann_sub, flag_BC_option_outflow_rudy_sub;
    FLT rho_inflow, p_inflow, u_inflow, v_inflow, w_inflow;
    FLT p_out, AlphaNonRef, T_wall;
    /* interface index of each face, or NO_NEIGHBOUR (getNeighbour, src/Functions.c:3795-3850);
       the donor rank and the CGNS transform of interface i (src/SHOCK.h:257, 278, 285).  The
       GPU path takes a face with an interface only as half of a periodic self-interface: donor
       == MPI_rank, transform (1, 2, 3), and the opposite face of the direction the same. */
    int InterfaceNeighbourLeft, InterfaceNeighbourRight, InterfaceNeighbourBottom;
    int InterfaceNeighbourTop, InterfaceNeighbourBehind, InterfaceNeighbourInFront;
    int MPI_rank;
    int *MPI_rankNeighbours;
    int **TransformMatrixOfInterface;
    unsigned compat_quirks;                   /* SHX_QUIRK_* bits; compat_define sets the reference's */
    void *compat_case;                        /* the device case, created by the first step */
};

/* metric on the padded grid, already mirrored at walls as the reference leaves it */
struct strct_mesh {
    FLT *xi_x, *xi_y, *xi_z, *eta_x, *eta_y, *eta_z, *zeta_x, *zeta_y, *zeta_z, *jacobian;
};

/* state on the padded grid: rho, u, v, w, e are the path's; the derived fields are not written */
struct strct_U {
    FLT *rho, *u, *v, *w, *p, *e, *c, *T, *mue, *theta1, *theta2, *theta3;
};

/* the reference's flux scratch; accepted and unused (the device keeps its own) */
struct strct_Flux {
    FLT *Mass, *xiMomentum, *etaMomentum, *zetaMomentum, *Energy;
};

/* loadDivision (src/Load.c:191-250): record `divzone` (1-based) of a division file written for
 * `proccount` processes, DIM = dims.  Each record: zone name (32 bytes), corner indices
 * (2 x DIM cgsize_t, 1-based, lower then upper), then per face (i-, i+, j-, j+, k-, k+) the
 * neighbour zone (0: none), transform (DIM ints), range and donor range (2 x DIM cgsize_t each)
 * and three PHYSDIM = 3 float vectors.  cgsize_t is 8 bytes; the bytes are the host's
 * (little-endian) unless swap is 2 (always swap) -- the reference's flag_swapDivisionFile.
 * Returns 0, or -1 (size mismatch: the file must hold proccount records) / -2 (unreadable). */
struct compat_interface {
    int neighbour;
    int transform[3];
    long long range[2][3], drange[2][3];
    float rc[3], ra[3], tr[3];
};
struct compat_division {
    char zonename[33];
    long long corners[2][3];
    struct compat_interface face[6];
    int ni;                                   /* faces with a neighbour */
};
int compat_load_division(const char *path, int dims, int proccount, unsigned divzone, int swap,
                         struct compat_division *out);

/* The face-to-face connections of a whole division (zones[z] = record z + 1), for
 * shx_blocks_create (shockx.h): every face with a neighbour is joined to the neighbour's face on
 * which the donor range lies (its constant index: 1 is the lower face, as
 * check_TransformationMatrix reads it, src/Functions.c:3902-3990), with the record's transform;
 * zone z is block z.  Writes at most 2 * dims * nzones entries and returns their number, or -1
 * when a donor range lies on no face. */
struct shx_interface;
int compat_division_connections(const struct compat_division *zones, int nzones, int dims,
                                struct shx_interface *out);

/* The mesh part of loadFile (src/Load.c:252-744) for CGNS files stored in HDF5, read without the
 * CGNS and HDF5 libraries (the HDF5 subset is described in cgns_mesh.c).  The block is division
 * record `div`'s corner range of the zone it names; x, y, z receive its nodes, i fastest
 * (i2 + ni (j2 + nj k2)), as doubles (z is zeroed in 2D; any of them may be NULL).  For every face
 * without a neighbour, info->bc holds the string assignBC writes for the zone's boundary
 * condition on that face (FamilySpecified resolved through the family); faces with a neighbour
 * keep an empty string.  Returns 0, or -1 with a message in err (a wrong cell dimension, a zone
 * or a face boundary condition that cannot be found, an unsupported BC type, or an HDF5 feature
 * outside the subset). */
struct compat_mesh_info {
    int celldim, physdim;
    long long zone_size[3];                   /* vertex counts of the whole zone */
    long start_iteration;                     /* the base's "Iterations" descriptor, else 0 */
    int coord_real8;                          /* the coordinates are stored as RealDouble */
    char bc[6][30];                           /* i-, i+, j-, j+, k-, k+ */
};
int compat_load_mesh(const char *path, int dims, const struct compat_division *div, double *x, double *y,
                     double *z, struct compat_mesh_info *info, char *err, int errlen);

/* ExtrapolateGhostCells (src/Functions.c:6526-6670) for a block read by compat_load_mesh: from the
 * node coordinates x, y, z (n[0] x n[1] x n[2] nodes, i fastest) the extended arrays
 * shx_build_metric takes (extent n + space_order + 3 per direction, k fastest; 1 in 2D z, where
 * ze is zeroed): layer l beyond a face is x[-1-l] = (x[-l] + x[l]) - x[l+1], direction by
 * direction (i, then j over the extended i, then k).  At a face with a neighbour the reference
 * extrapolates as well, and the metric ghosts then come from the neighbour
 * (shx_blocks_transfer_metric).  Returns 0, or -1 (bad arguments or no memory). */
int compat_extend_coords(int dims, int space_order, const long long n[3], const double *x, const double *y,
                         const double *z, double *xe, double *ye, double *ze);

void compat_define(struct strct_configuration *cfg);
void compat_release(struct strct_configuration *cfg);

void CalcRungeKutta(struct strct_configuration *pnt_config, struct strct_mesh *pnt_mesh,
                    struct strct_U *pnt_U_lastStep, struct strct_U *pnt_U_RK, struct strct_Flux *pnt_Q,
                    struct strct_Flux *pnt_Q_sum, struct strct_Flux *pnt_Flux,
                    struct strct_Flux *pnt_Flux_PlusHalf);
int checkNAN(struct strct_configuration *pnt_config, struct strct_U *pnt_U, struct strct_mesh *pnt_mesh);
void WriteValuesFromU1ToU2(struct strct_configuration *pnt_config, struct strct_U *pnt_U1, struct strct_U *pnt_U2);

#ifdef __cplusplus
}
#endif
#endif
