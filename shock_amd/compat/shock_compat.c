/*
 * shock_compat.c -- the reference's RHS / time-step entry points over libshockx.so (see
 * shock_compat.h).  The device case is created by the first CalcRungeKutta of a configuration:
 * face kinds from the reference's BC strings and option flags (SetAllBoundaryConditions,
 * src/BC.c:75-637), the metric from the mesh arrays.  The state is uploaded when U_lastStep
 * differs from what the previous call wrote back (a shadow copy is kept), so the usual loop
 * -- CalcRungeKutta, checkNAN, WriteValuesFromU1ToU2 -- moves each state over PCIe once per step
 * and the Rudy writer keeps the uploaded sound speed as the reference keeps U_lastStep's.
 */
#include "shock_compat.h"
#include "shockx.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Host-memory passes over whole fields (the shadow compare and copy, WriteValuesFromU1ToU2) on up
 * to PAR_THREADS threads: one thread reaches a fraction of the host's memory bandwidth.  Fields
 * under PAR_MIN bytes, or a thread that cannot be started, are done on the calling thread. */
#define PAR_THREADS 8
#define PAR_MIN ((size_t)4 << 20)
struct par_job {
    const char *a, *b;   /* compare a with b, or copy a to dst */
    char *dst;
    size_t n;
    int diff;
};
static void *par_cmp(void *p)
{
    struct par_job *j = (struct par_job *)p;
    j->diff = memcmp(j->a, j->b, j->n) != 0;
    return NULL;
}
static void *par_cpy(void *p)
{
    struct par_job *j = (struct par_job *)p;
    memcpy(j->dst, j->a, j->n);
    return NULL;
}
/* returns 1 when any compared part differs (par_cmp), 0 otherwise */
static int par_run(void *(*fn)(void *), const void *a, const void *b, void *dst, size_t n)
{
    struct par_job job[PAR_THREADS];
    pthread_t th[PAR_THREADS];
    int started[PAR_THREADS] = {0};
    const int nt = n < PAR_MIN ? 1 : PAR_THREADS;
    const size_t part = ((n + nt - 1) / nt + 63) & ~(size_t)63;
    int parts = 0;
    for (int t = 0; t < nt; t++) {
        const size_t off = (size_t)t * part;
        if (off >= n) break;
        job[t].a = (const char *)a + off;
        job[t].b = b ? (const char *)b + off : NULL;
        job[t].dst = dst ? (char *)dst + off : NULL;
        job[t].n = (n - off < part) ? n - off : part;
        job[t].diff = 0;
        parts = t + 1;
        if (t > 0) started[t] = pthread_create(&th[t], NULL, fn, &job[t]) == 0;
    }
    fn(&job[0]);
    int diff = job[0].diff;
    for (int t = 1; t < parts; t++) {
        if (started[t]) pthread_join(th[t], NULL);
        else fn(&job[t]);
        diff |= job[t].diff;
    }
    return diff;
}
static int fields_differ(const double *a, const double *b, size_t bytes) { return par_run(par_cmp, a, b, NULL, bytes); }
static void copy_field(double *dst, const double *src, size_t bytes) { par_run(par_cpy, src, NULL, dst, bytes); }

struct compat_ctx {
    shx_case *c;
    long ntot;
    double *shadow[5];   /* rho, u, v, w, e as last written to the host */
    int have_shadow;
};

static void compat_abort(const char *what, int rc)
{
    printf("SHOCK: GPU path: %s failed (%d): %s\n", what, rc, shx_last_error());
    exit(4242);
}
#define CALL(x)                          \
    do {                                 \
        int rc_ = (x);                   \
        if (rc_ != SHX_OK) compat_abort(#x, rc_); \
    } while (0)

void compat_define(struct strct_configuration *cfg)
{
    const int g = (cfg->SPACEORDER + 1) / 2;
    cfg->int_iMeshPointsGhostCells = cfg->int_iMeshPoints + cfg->SPACEORDER + 1;
    cfg->int_jMeshPointsGhostCells = cfg->int_jMeshPoints + cfg->SPACEORDER + 1;
    cfg->int_iStartReal = g; cfg->int_iEndReal = g + cfg->int_iMeshPoints - 1;
    cfg->int_jStartReal = g; cfg->int_jEndReal = g + cfg->int_jMeshPoints - 1;
    if (cfg->MESHDIMENSIONS == 3) {
        cfg->int_kMeshPointsGhostCells = cfg->int_kMeshPoints + cfg->SPACEORDER + 1;
        cfg->int_kStartReal = g; cfg->int_kEndReal = g + cfg->int_kMeshPoints - 1;
    } else {
        cfg->int_kMeshPoints = 1;
        cfg->int_kMeshPointsGhostCells = 1;
        cfg->int_kStartReal = 0; cfg->int_kEndReal = 0;
    }
    cfg->compat_quirks = SHX_QUIRKS_REFERENCE;
}

static const char *const face_name[6] = {"left", "right", "bottom", "top", "behind", "in front"};

/* an interface face is one half of a periodic self-interface: its donor is this rank, its
   transform the identity, and the opposite face of the direction is such a face too (the
   single-rank self-transfer, src/Functions.c:3580-3684).  Anything else needs the halo exchange
   of another block, which this path does not have. */
static int self_interface(const struct strct_configuration *cfg, const int nbr[6], int f)
{
    const int i = nbr[f];
    if (i == NO_NEIGHBOUR) return 0;
    if (i < 0 || !cfg->MPI_rankNeighbours || !cfg->TransformMatrixOfInterface ||
        !cfg->TransformMatrixOfInterface[i]) {
        printf("SHOCK: GPU path: face %s: interface %d without donor rank and transform\n", face_name[f], i);
        exit(4242);
    }
    if (cfg->MPI_rankNeighbours[i] != cfg->MPI_rank) {
        printf("SHOCK: GPU path: face %s: interface to rank %d; only periodic self-interfaces of rank %d "
               "are supported\n", face_name[f], cfg->MPI_rankNeighbours[i], cfg->MPI_rank);
        exit(4242);
    }
    for (int d = 0; d < cfg->MESHDIMENSIONS; d++)
        if (cfg->TransformMatrixOfInterface[i][d] != d + 1) {
            printf("SHOCK: GPU path: face %s: interface transform is not the identity; only periodic "
                   "self-interfaces are supported\n", face_name[f]);
            exit(4242);
        }
    return 1;
}

/* face kind of one face (SetAllBoundaryConditions' selection, src/BC.c:75-637) */
static int face_kind(const struct strct_configuration *cfg, const char *bc, const int nbr[6], int f)
{
    const int inflow = !strcmp(bc, cfg->BCInflow) || !strcmp(bc, cfg->BCInflowSubsonic) ||
                       !strcmp(bc, cfg->BCInflowSupersonic);
    const int outflow = !strcmp(bc, cfg->BCOutflow) || !strcmp(bc, cfg->BCOutflowSubsonic);
    if (self_interface(cfg, nbr, f)) {
        if (!self_interface(cfg, nbr, f ^ 1)) {
            printf("SHOCK: GPU path: face %s is an interface but face %s is not; only periodic "
                   "self-interfaces are supported\n", face_name[f], face_name[f ^ 1]);
            exit(4242);
        }
        return SHX_BC_PERIODIC;
    }
    if (!strcmp(bc, cfg->BCManufacturedSolution)) return SHX_BC_MANUFACTURED;
    if (!strcmp(bc, cfg->BCWallInviscid)) return SHX_BC_WALL_SLIP;
    if (!strcmp(bc, cfg->BCWallViscous)) return SHX_BC_WALL_NOSLIP;
    if (!strcmp(bc, cfg->BCWallViscousIsothermal)) return SHX_BC_WALL_ISOTHERMAL;
    if (!strcmp(bc, cfg->BCFarfield)) return SHX_BC_FARFIELD;
    if (inflow && cfg->flag_BC_option_inflow_normal_super) return SHX_BC_INFLOW_SUPERSONIC;
    if (inflow && cfg->flag_BC_option_inflow_normal_sub) return SHX_BC_INFLOW_SUB_NORMAL;
    if (inflow && cfg->flag_BC_option_inflow_riemann_sub) return SHX_BC_INFLOW_SUB_RIEMANN;
    if (inflow && cfg->flag_BC_option_inflow_isentrop_sub) return SHX_BC_INFLOW_SUB_ISENTROPIC;
    if (outflow && cfg->flag_BC_option_outflow_riemann_sub) return SHX_BC_OUTFLOW_RIEMANN;
    if (outflow && cfg->flag_BC_option_outflow_normal_sub) return SHX_BC_OUTFLOW_NORMAL;
    if (outflow && cfg->flag_BC_option_outflow_rudy_sub) return SHX_BC_OUTFLOW_RUDY;
    printf("SHOCK: GPU path: boundary %s not available\n", bc);
    exit(4242);
}

static struct compat_ctx *ctx_of(struct strct_configuration *cfg, struct strct_mesh *mesh)
{
    struct compat_ctx *x = (struct compat_ctx *)cfg->compat_case;
    if (x) return x;
    if (cfg->flag_ManufacturedSolution) {
        printf("SHOCK: GPU path: the manufactured solution needs shx_upload_source / shx_upload_boundary_state\n");
        exit(4242);
    }
    shx_desc d;
    memset(&d, 0, sizeof d);
    d.dims = cfg->MESHDIMENSIONS; d.space_order = cfg->SPACEORDER; d.time_order = cfg->int_TimeOrder;
    d.n[0] = cfg->int_iMeshPoints; d.n[1] = cfg->int_jMeshPoints; d.n[2] = cfg->int_kMeshPoints;
    d.gamma = cfg->gammaNumber; d.mach = cfg->machNumber; d.reynolds = cfg->reynoldsNumber;
    d.prandtl = cfg->prandtlNumber; d.T0 = cfg->T0_dim;
    d.inviscid = cfg->flag_Inviscid; d.manufactured = 0; d.quirks = cfg->compat_quirks;
    const char *bcs[6] = {cfg->BC_Left, cfg->BC_Right, cfg->BC_Bottom, cfg->BC_Top, cfg->BC_Behind, cfg->BC_InFront};
    const int nbr[6] = {cfg->InterfaceNeighbourLeft, cfg->InterfaceNeighbourRight, cfg->InterfaceNeighbourBottom,
                        cfg->InterfaceNeighbourTop, cfg->InterfaceNeighbourBehind, cfg->InterfaceNeighbourInFront};
    for (int f = 0; f < 6; f++)
        d.bc[f] = (f < 2 * cfg->MESHDIMENSIONS) ? face_kind(cfg, bcs[f], nbr, f) : SHX_BC_PERIODIC;
    d.precision = SHX_FP64;
    d.rho_inflow = cfg->rho_inflow; d.u_inflow = cfg->u_inflow; d.v_inflow = cfg->v_inflow;
    d.w_inflow = cfg->w_inflow; d.p_inflow = cfg->p_inflow;
    d.p_out = cfg->p_out; d.alpha_nonref = cfg->AlphaNonRef; d.T_wall = cfg->T_wall;
    x = (struct compat_ctx *)calloc(1, sizeof *x);
    CALL(shx_create(&d, &x->c));
    const double *met[10] = {mesh->xi_x, mesh->xi_y, mesh->xi_z, mesh->eta_x, mesh->eta_y,
                             mesh->eta_z, mesh->zeta_x, mesh->zeta_y, mesh->zeta_z, mesh->jacobian};
    CALL(shx_upload_metric(x->c, met));
    x->ntot = (long)cfg->int_iMeshPointsGhostCells * cfg->int_jMeshPointsGhostCells * cfg->int_kMeshPointsGhostCells;
    for (int m = 0; m < 5; m++) x->shadow[m] = (double *)malloc(sizeof(double) * x->ntot);
    cfg->compat_case = x;
    return x;
}

void compat_release(struct strct_configuration *cfg)
{
    struct compat_ctx *x = (struct compat_ctx *)cfg->compat_case;
    if (!x) return;
    shx_destroy(x->c);
    for (int m = 0; m < 5; m++) free(x->shadow[m]);
    free(x);
    cfg->compat_case = NULL;
}

static void fields(struct strct_U *U, double *f[5])
{
    f[0] = U->rho; f[1] = U->u; f[2] = U->v; f[3] = U->w; f[4] = U->e;
}

/* the device holds U unless U differs from what was last written back to the host */
static void ensure_on_device(struct compat_ctx *x, struct strct_U *U)
{
    double *f[5];
    fields(U, f);
    int same = x->have_shadow;
    for (int m = 0; m < 5 && same; m++) same = !fields_differ(f[m], x->shadow[m], sizeof(double) * x->ntot);
    if (same) return;
    const double *cf[5] = {f[0], f[1], f[2], f[3], f[4]};
    CALL(shx_upload_state(x->c, cf));
    for (int m = 0; m < 5; m++) copy_field(x->shadow[m], f[m], sizeof(double) * x->ntot);
    x->have_shadow = 1;
}

static void write_back(struct compat_ctx *x, struct strct_U *U)
{
    double *f[5];
    fields(U, f);
    CALL(shx_download_state(x->c, f));
    for (int m = 0; m < 5; m++) copy_field(x->shadow[m], f[m], sizeof(double) * x->ntot);
    x->have_shadow = 1;
}

void CalcRungeKutta(struct strct_configuration *pnt_config, struct strct_mesh *pnt_mesh,
                    struct strct_U *pnt_U_lastStep, struct strct_U *pnt_U_RK, struct strct_Flux *pnt_Q,
                    struct strct_Flux *pnt_Q_sum, struct strct_Flux *pnt_Flux,
                    struct strct_Flux *pnt_Flux_PlusHalf)
{
    (void)pnt_Q; (void)pnt_Q_sum; (void)pnt_Flux; (void)pnt_Flux_PlusHalf;
    struct compat_ctx *x = ctx_of(pnt_config, pnt_mesh);
    ensure_on_device(x, pnt_U_lastStep);
    CALL(shx_calc_runge_kutta(x->c, pnt_config->numericalTau));
    write_back(x, pnt_U_RK);
}

int checkNAN(struct strct_configuration *pnt_config, struct strct_U *pnt_U, struct strct_mesh *pnt_mesh)
{
    struct compat_ctx *x = ctx_of(pnt_config, pnt_mesh);
    ensure_on_device(x, pnt_U);
    int count = 0;
    CALL(shx_check_nan(x->c, &count));
    if (count > 0) {
        /* the marks checkNAN leaves in the state (src/Functions.c:5443-5445) */
        write_back(x, pnt_U);
        for (long n = 0; n < x->ntot; n++)
            if (pnt_U->rho[n] == 1000000 && pnt_U->p) pnt_U->p[n] = 1000000;
        pnt_config->flag_NAN = 1;
        return 1;
    }
    return 0;
}

void WriteValuesFromU1ToU2(struct strct_configuration *pnt_config, struct strct_U *pnt_U1, struct strct_U *pnt_U2)
{
    const long ntot = (long)pnt_config->int_iMeshPointsGhostCells * pnt_config->int_jMeshPointsGhostCells *
                      pnt_config->int_kMeshPointsGhostCells;
    const size_t bytes = sizeof(double) * (size_t)ntot;
    copy_field(pnt_U2->rho, pnt_U1->rho, bytes);
    copy_field(pnt_U2->u, pnt_U1->u, bytes);
    copy_field(pnt_U2->v, pnt_U1->v, bytes);
    copy_field(pnt_U2->w, pnt_U1->w, bytes);
    copy_field(pnt_U2->e, pnt_U1->e, bytes);
    if (pnt_U1->p && pnt_U2->p) copy_field(pnt_U2->p, pnt_U1->p, bytes);
}

/* ---- division file -------------------------------------------------------------------- */
static void swap_bytes(void *mem, size_t size, int count)
{
    for (int k = 0; k < count; k++) {
        unsigned char *b = (unsigned char *)mem + (size_t)k * size;
        for (size_t i = 0; i < size / 2; i++) {
            unsigned char t = b[i];
            b[i] = b[size - 1 - i];
            b[size - 1 - i] = t;
        }
    }
}

int compat_division_connections(const struct compat_division *zones, int nzones, int dims,
                                struct shx_interface *out)
{
    int n = 0;
    for (int z = 0; z < nzones; z++)
        for (int f = 0; f < 2 * dims; f++) {
            const struct compat_interface *it = &zones[z].face[f];
            if (!it->neighbour) continue;
            int df = -1;
            for (int k = 0; k < dims && df < 0; k++)
                if (it->drange[0][k] == it->drange[1][k]) df = 2 * k + (it->drange[0][k] == 1 ? 0 : 1);
            if (df < 0) return -1;
            struct shx_interface *e = &out[n++];
            e->block = z;
            e->face = f;
            e->donor = it->neighbour - 1;
            e->donor_face = df;
            for (int k = 0; k < 3; k++) e->transform[k] = (k < dims) ? it->transform[k] : k + 1;
        }
    return n;
}

int compat_load_division(const char *path, int dims, int proccount, unsigned divzone, int swap,
                         struct compat_division *out)
{
    const int D = dims, PHYS = 3;
    const size_t cg = 8;   /* cgsize_t */
    const size_t ifblock = sizeof(int) + D * sizeof(int) + 2 * D * cg * 2 + PHYS * sizeof(float) * 3;
    const size_t datalen = 32 + cg * D * 2 + ifblock * 2 * D;
    FILE *f = fopen(path, "rb");
    if (!f) return -2;
    fseek(f, 0, SEEK_END);
    const long size = ftell(f);
    if (size != (long)(proccount * datalen)) {
        printf("(EE) Division file '%s' has size %ld but should be processors (%d) times data (%d)\n", path, size,
               proccount, (int)datalen);
        fclose(f);
        return -1;
    }
    unsigned char *rec = (unsigned char *)malloc(datalen);
    fseek(f, (long)((divzone - 1) * datalen), SEEK_SET);
    const size_t got = fread(rec, 1, datalen, f);
    fclose(f);
    if (got != datalen) { free(rec); return -2; }
    /* the reference swaps when asked to (2) or on a big-endian host; files are written little-endian */
    const int one = 1;
    const int do_swap = (swap == 2) || (swap == 0 && *(const char *)&one != 1);
    memset(out, 0, sizeof *out);
    memcpy(out->zonename, rec, 32);
    out->zonename[32] = '\0';
    long long c[6];
    memcpy(c, rec + 32, 2 * D * cg);
    if (do_swap) swap_bytes(c, cg, 2 * D);
    for (int h = 0; h < 2; h++)
        for (int d = 0; d < D; d++) out->corners[h][d] = c[h * D + d];
    for (int fidx = 0; fidx < 2 * D; fidx++) {
        const unsigned char *b = rec + 32 + 2 * D * cg + fidx * ifblock;
        int nb;
        memcpy(&nb, b, sizeof(int));
        if (do_swap) swap_bytes(&nb, sizeof(int), 1);
        if (!nb) continue;
        struct compat_interface *it = &out->face[fidx];
        out->ni++;
        it->neighbour = nb;
        int tr[3];
        memcpy(tr, b + 4, D * sizeof(int));
        long long rg[12];
        memcpy(rg, b + 4 + 4 * D, 4 * D * cg);
        float fl[9];
        memcpy(fl, b + 4 + 4 * D + 4 * D * cg, 9 * sizeof(float));
        if (do_swap) {
            swap_bytes(tr, sizeof(int), D);
            swap_bytes(rg, cg, 4 * D);
            swap_bytes(fl, sizeof(float), 9);
        }
        for (int d = 0; d < D; d++) {
            it->transform[d] = tr[d];
            it->range[0][d] = rg[d];
            it->range[1][d] = rg[D + d];
            it->drange[0][d] = rg[2 * D + d];
            it->drange[1][d] = rg[3 * D + d];
        }
        for (int q = 0; q < 3; q++) { it->rc[q] = fl[q]; it->ra[q] = fl[3 + q]; it->tr[q] = fl[6 + q]; }
    }
    free(rec);
    return 0;
}
