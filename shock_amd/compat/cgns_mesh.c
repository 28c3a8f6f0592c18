/*
 * cgns_mesh.c -- the mesh part of loadFile (src/Load.c:252-744) for CGNS files stored in HDF5,
 * read without the CGNS and HDF5 libraries.
 *
 * CGNS on HDF5 (the layout the CGNS library writes): every CGNS node is an HDF5 group with string
 * attributes "name", "label" (the SIDS type, e.g. Zone_t) and "type" (data type: MT, I4, I8, R4,
 * R8, C1); a node with data holds it in a child dataset named " data", whose HDF5 dimensions are
 * the node's Fortran dimensions reversed.  Children are iterated in creation order, as the CGNS
 * library does.
 *
 * The HDF5 subset read here: superblock version 2 or 3 with 8-byte offsets and lengths; version-2
 * object headers ("OHDR") with their continuation blocks ("OCHK"); links and attributes stored in
 * the object header (compact storage); datasets with contiguous or compact layout and no filter;
 * little-endian integers and IEEE floats, fixed-length strings.  Anything else is refused with a
 * message that names it.  Checksums are not verified.  The whole file is mapped read-only.
 *
 * What loadFile does with the mesh, restated:
 *   - the base is the file's first CGNSBase_t; its cell dimension must equal the compiled one
 *     (src/Load.c:297-302); an "Iterations" descriptor of the base gives the start iteration
 *     (src/Load.c:304-318);
 *   - the zone is the one named in the division record (src/Load.c:322-335);
 *   - each face without a neighbour takes the first boundary condition of the zone (ZoneBC_t
 *     children in order) whose PointRange has a constant index along the face normal, equal to 1
 *     exactly when the face is the lower one, and which overlaps the block's face in every
 *     tangential direction by the reference's strict-inequality test (src/Load.c:455-476);
 *     FamilySpecified resolves through the FamilyName child to the base's Family_t node of that
 *     name and its first FamilyBC_t (src/Load.c:478-491); the BC type becomes the string
 *     assignBC writes (src/Load.c:46-97);
 *   - the coordinates CoordinateX/Y/Z of the block's corner range, i fastest
 *     (ijk2 = k2 nj ni + j2 ni + i2, src/Functions.c:746-783), widened to double.
 * Flow solutions (src/Load.c:590-735) are not read: the reference's testcase meshes hold none.
 */
#include <fcntl.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include "shock_compat.h"

#define UNDEF_ADDR 0xffffffffffffffffull

typedef struct {
    const unsigned char *p;
    size_t n;
    char *err;
    int errlen;
} H5;

static int fail(H5 *h, const char *fmt, ...)
{
    if (h->err && h->errlen > 0 && !h->err[0]) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(h->err, (size_t)h->errlen, fmt, ap);
        va_end(ap);
    }
    return -1;
}

static uint64_t le(const unsigned char *p, int nbytes)
{
    uint64_t v = 0;
    for (int b = nbytes - 1; b >= 0; b--) v = (v << 8) | p[b];
    return v;
}

/* one parsed object header: its hard links, the CGNS attributes and, for a dataset, its shape,
   element type and where its bytes are */
typedef struct {
    int nlink, cap;
    char (*lname)[64];
    uint64_t *laddr, *lorder;
    int dense_links, dense_attrs, v1_group;
    char name[33], label[33], type[3];
    int has_space, rank;
    uint64_t dims[8];
    int has_type, tclass, tsize, tsigned, tbig;
    int layout;                      /* -1 none, 0 compact, 1 contiguous, 2 other */
    uint64_t daddr, dsize;
    const unsigned char *compact;
    int filtered;
} Obj;

static void obj_free(Obj *o)
{
    free(o->lname);
    free(o->laddr);
    free(o->lorder);
    memset(o, 0, sizeof *o);
}

static int add_link(H5 *h, Obj *o, const char *name, int len, uint64_t addr, uint64_t order)
{
    if (o->nlink == o->cap) {
        const int cap = o->cap ? 2 * o->cap : 16;
        void *a = realloc(o->lname, (size_t)cap * 64), *b = realloc(o->laddr, (size_t)cap * 8),
             *c = realloc(o->lorder, (size_t)cap * 8);
        if (a) o->lname = a;
        if (b) o->laddr = b;
        if (c) o->lorder = c;
        if (!a || !b || !c) return fail(h, "out of memory");
        o->cap = cap;
    }
    if (len > 63) len = 63;
    memcpy(o->lname[o->nlink], name, (size_t)len);
    o->lname[o->nlink][len] = '\0';
    o->laddr[o->nlink] = addr;
    o->lorder[o->nlink] = order;
    o->nlink++;
    return 0;
}

/* dataspace message: rank and dimensions */
static int parse_space(H5 *h, const unsigned char *d, size_t sz, int *rank, uint64_t dims[8])
{
    if (sz < 4) return fail(h, "short dataspace message");
    const int ver = d[0], r = d[1];
    if (r > 8) return fail(h, "dataspace of rank %d", r);
    const unsigned char *q = (ver == 1) ? d + 8 : d + 4;
    if (ver != 1 && ver != 2) return fail(h, "dataspace message version %d", ver);
    if ((size_t)(q - d) + 8u * (size_t)r > sz) return fail(h, "short dataspace message");
    for (int i = 0; i < r; i++) dims[i] = le(q + 8 * i, 8);
    *rank = r;
    return 0;
}

/* attribute message: keep the CGNS string attributes name, label and type */
static int parse_attr(H5 *h, Obj *o, const unsigned char *d, size_t sz)
{
    if (sz < 8) return fail(h, "short attribute message");
    const int ver = d[0];
    const size_t nsz = le(d + 2, 2), tsz = le(d + 4, 2), ssz = le(d + 6, 2);
    const unsigned char *q;
    size_t tq, sq, dq;
    if (ver == 1) {
        q = d + 8;
        tq = 8 + ((nsz + 7) & ~(size_t)7);
        sq = tq + ((tsz + 7) & ~(size_t)7);
        dq = sq + ((ssz + 7) & ~(size_t)7);
    } else if (ver == 2 || ver == 3) {
        const size_t base = (ver == 3) ? 9 : 8;
        q = d + base;
        tq = base + nsz;
        sq = tq + tsz;
        dq = sq + ssz;
    } else {
        return fail(h, "attribute message version %d", ver);
    }
    if (dq > sz) return fail(h, "short attribute message");
    char an[64];
    size_t nl = strnlen((const char *)q, nsz);
    if (nl > 63) nl = 63;
    memcpy(an, q, nl);
    an[nl] = '\0';
    const int tclass = d[tq] & 0x0f;
    const size_t esize = le(d + tq + 4, 4);
    if (tclass != 3) return 0;                 /* only the string attributes matter */
    int rank = 0;
    uint64_t dims[8];
    if (parse_space(h, d + sq, ssz, &rank, dims)) return -1;
    size_t count = 1;
    for (int i = 0; i < rank; i++) count *= (size_t)dims[i];
    const size_t bytes = count * esize;
    if (dq + bytes > sz) return fail(h, "attribute '%s' data beyond its message", an);
    char *dst = NULL;
    size_t cap = 0;
    if (!strcmp(an, "name")) { dst = o->name; cap = sizeof o->name; }
    else if (!strcmp(an, "label")) { dst = o->label; cap = sizeof o->label; }
    else if (!strcmp(an, "type")) { dst = o->type; cap = sizeof o->type; }
    if (dst) {
        size_t l = strnlen((const char *)d + dq, bytes);
        if (l > cap - 1) l = cap - 1;
        memcpy(dst, d + dq, l);
        dst[l] = '\0';
    }
    return 0;
}

static int parse_link(H5 *h, Obj *o, const unsigned char *d, size_t sz)
{
    if (sz < 2 || d[0] != 1) return fail(h, "link message version %d", sz ? d[0] : -1);
    const int fl = d[1];
    size_t q = 2;
    int ltype = 0;
    uint64_t order = 0;
    if (fl & 0x08) ltype = d[q++];
    if (fl & 0x04) { order = le(d + q, 8); q += 8; }
    if (fl & 0x10) q++;
    const int lsz = 1 << (fl & 3);
    const size_t ln = le(d + q, lsz);
    q += (size_t)lsz;
    if (q + ln > sz) return fail(h, "short link message");
    const char *nm = (const char *)d + q;
    q += ln;
    if (ltype != 0) return 0;                  /* soft and external links: not part of a mesh read */
    if (q + 8 > sz) return fail(h, "short link message");
    return add_link(h, o, nm, (int)ln, le(d + q, 8), order);
}

/* the object header at addr (parse frees what a failed parse_header collected) */
static int parse_header(H5 *h, uint64_t addr, Obj *o);
static int parse(H5 *h, uint64_t addr, Obj *o)
{
    if (parse_header(h, addr, o)) {
        obj_free(o);
        return -1;
    }
    return 0;
}

static int parse_header(H5 *h, uint64_t addr, Obj *o)
{
    memset(o, 0, sizeof *o);
    o->layout = -1;
    if (addr > h->n || h->n - addr < 8 || memcmp(h->p + addr, "OHDR", 4))
        return fail(h, "no version-2 object header at offset %llu (files written with HDF5's "
                       "oldest format versions are not read)", (unsigned long long)addr);
    if (h->p[addr + 4] != 2) return fail(h, "object header version %d", h->p[addr + 4]);
    const int fl = h->p[addr + 5];
    size_t q = (size_t)addr + 6;
    if (fl & 0x20) q += 16;
    if (fl & 0x10) q += 4;
    const int szb = 1 << (fl & 3);
    if (q + (size_t)szb > h->n) return fail(h, "truncated object header");
    const uint64_t csize = le(h->p + q, szb);
    q += (size_t)szb;
    const int crt = (fl & 0x04) != 0;
    size_t bs[64], be[64];
    int nb = 0, bi = 0;
    bs[nb] = q;
    be[nb++] = q + (size_t)csize;
    while (bi < nb) {
        size_t p = bs[bi], e = be[bi++];
        if (e > h->n) return fail(h, "object header block beyond the end of the file");
        const size_t mh = 4 + (crt ? 2 : 0);
        while (p + mh <= e) {
            const int type = h->p[p];
            const size_t sz = le(h->p + p + 1, 2);
            p += mh;
            if (p + sz > e) return fail(h, "object header message beyond its block");
            const unsigned char *d = h->p + p;
            int rc = 0;
            switch (type) {
            case 0x01: rc = parse_space(h, d, sz, &o->rank, o->dims); o->has_space = 1; break;
            case 0x02: {                                              /* link info */
                if (sz < 18) { rc = fail(h, "short link info message"); break; }
                const size_t off = 2 + ((d[1] & 1) ? 8 : 0);
                if (le(d + off, 8) != UNDEF_ADDR) o->dense_links = 1;
                break;
            }
            case 0x03:                                                /* datatype */
                if (sz < 8) { rc = fail(h, "short datatype message"); break; }
                o->has_type = 1;
                o->tclass = d[0] & 0x0f;
                o->tbig = d[1] & 1;
                o->tsigned = (d[1] >> 3) & 1;
                o->tsize = (int)le(d + 4, 4);
                break;
            case 0x06: rc = parse_link(h, o, d, sz); break;
            case 0x08:                                                /* data layout */
                if (sz < 2) { rc = fail(h, "short layout message"); break; }
                if ((d[0] == 3 || d[0] == 4) && d[1] == 0 && sz >= 4) {
                    o->layout = 0;
                    o->dsize = le(d + 2, 2);
                    o->compact = d + 4;
                    if (4 + o->dsize > sz) rc = fail(h, "compact data beyond its message");
                } else if ((d[0] == 3 || d[0] == 4) && d[1] == 1 && sz >= 18) {
                    o->layout = 1;
                    o->daddr = le(d + 2, 8);
                    o->dsize = le(d + 10, 8);
                } else {
                    o->layout = 2;
                }
                break;
            case 0x0B: o->filtered = 1; break;                        /* filter pipeline */
            case 0x0C: rc = parse_attr(h, o, d, sz); break;
            case 0x10: {                                              /* continuation */
                if (sz < 16) { rc = fail(h, "short continuation message"); break; }
                const uint64_t ca = le(d, 8), cl = le(d + 8, 8);
                if (ca > h->n || cl < 8 || h->n - ca < cl || memcmp(h->p + ca, "OCHK", 4)) {
                    rc = fail(h, "bad object header continuation at %llu", (unsigned long long)ca);
                    break;
                }
                if (nb == 64) { rc = fail(h, "more than 64 object header blocks"); break; }
                bs[nb] = (size_t)ca + 4;
                be[nb++] = (size_t)(ca + cl) - 4;
                break;
            }
            case 0x11: o->v1_group = 1; break;                        /* symbol table */
            case 0x15: {                                              /* attribute info */
                if (sz < 10) { rc = fail(h, "short attribute info message"); break; }
                const size_t off = 2 + ((d[1] & 1) ? 2 : 0);
                if (le(d + off, 8) != UNDEF_ADDR) o->dense_attrs = 1;
                break;
            }
            default: break;
            }
            if (rc) return -1;
            p += sz;
        }
    }
    /* creation order, as the CGNS library iterates children */
    for (int a = 1; a < o->nlink; a++)
        for (int b = a; b > 0 && o->lorder[b] < o->lorder[b - 1]; b--) {
            char t[64];
            memcpy(t, o->lname[b], 64); memcpy(o->lname[b], o->lname[b - 1], 64); memcpy(o->lname[b - 1], t, 64);
            uint64_t x = o->laddr[b]; o->laddr[b] = o->laddr[b - 1]; o->laddr[b - 1] = x;
            x = o->lorder[b]; o->lorder[b] = o->lorder[b - 1]; o->lorder[b - 1] = x;
        }
    return 0;
}

/* a CGNS node: a group whose children can be listed */
static int node(H5 *h, uint64_t addr, Obj *o, const char *what)
{
    if (parse(h, addr, o)) return -1;
    if (o->v1_group || o->dense_links) {
        const int v1 = o->v1_group;
        obj_free(o);
        return fail(h, "%s: group stored with %s, which this reader does not read", what,
                    v1 ? "a symbol table (old-style group)" : "dense link storage");
    }
    if (o->dense_attrs && !o->label[0]) {
        obj_free(o);
        return fail(h, "%s: attributes in dense storage", what);
    }
    return 0;
}

/* the node's data: rank, reversed (HDF5) dimensions and a pointer to its bytes */
typedef struct {
    int rank;
    uint64_t dims[8];
    int tclass, tsize, tsigned;
    const unsigned char *bytes;
    size_t nbytes;
} Data;

static int node_data(H5 *h, const Obj *o, Data *out, const char *what)
{
    for (int c = 0; c < o->nlink; c++) {
        if (strcmp(o->lname[c], " data")) continue;
        Obj ds;
        if (parse(h, o->laddr[c], &ds)) return -1;
        int rc = 0;
        if (!ds.has_space || !ds.has_type) rc = fail(h, "%s: data without dataspace or datatype", what);
        else if (ds.filtered) rc = fail(h, "%s: data stored with a filter (compression), not read here", what);
        else if (ds.tbig) rc = fail(h, "%s: big-endian data", what);
        else if (ds.layout != 0 && ds.layout != 1) rc = fail(h, "%s: data layout other than contiguous or compact (chunked)", what);
        if (!rc) {
            memset(out, 0, sizeof *out);
            out->rank = ds.rank;
            memcpy(out->dims, ds.dims, sizeof ds.dims);
            out->tclass = ds.tclass;
            out->tsize = ds.tsize;
            out->tsigned = ds.tsigned;
            size_t count = 1;
            for (int i = 0; i < ds.rank; i++) count *= (size_t)ds.dims[i];
            out->nbytes = count * (size_t)ds.tsize;
            if (ds.layout == 0) {
                out->bytes = ds.compact;
                if (out->nbytes > ds.dsize) rc = fail(h, "%s: compact data shorter than its shape", what);
            } else if (ds.daddr == UNDEF_ADDR) {
                rc = fail(h, "%s: data never written", what);
            } else if (ds.daddr > h->n || h->n - ds.daddr < out->nbytes) {
                rc = fail(h, "%s: data beyond the end of the file", what);
            } else {
                out->bytes = h->p + ds.daddr;
            }
        }
        obj_free(&ds);
        return rc;
    }
    return fail(h, "%s: node has no data", what);
}

static int data_string(H5 *h, const Obj *o, char *dst, size_t cap, const char *what)
{
    Data d;
    if (node_data(h, o, &d, what)) return -1;
    if (d.tclass != 0 && d.tclass != 3) return fail(h, "%s: character data expected", what);
    size_t l = strnlen((const char *)d.bytes, d.nbytes);
    if (l > cap - 1) l = cap - 1;
    memcpy(dst, d.bytes, l);
    dst[l] = '\0';
    return 0;
}

/* integer element k of a node's data (I4 or I8) */
static long long data_int(const Data *d, size_t k)
{
    return (d->tsize == 8) ? (long long)le(d->bytes + 8 * k, 8) : (long long)(int32_t)le(d->bytes + 4 * k, 4);
}

static int data_ints(H5 *h, const Obj *o, long long *dst, size_t count, const char *what)
{
    Data d;
    if (node_data(h, o, &d, what)) return -1;
    if (d.tclass != 0 || (d.tsize != 4 && d.tsize != 8)) return fail(h, "%s: integer data expected", what);
    if (d.nbytes / (size_t)d.tsize < count) return fail(h, "%s: %zu integers expected", what, count);
    for (size_t k = 0; k < count; k++) dst[k] = data_int(&d, k);
    return 0;
}

/* child c of o parsed as a node */
static int child(H5 *h, const Obj *o, int c, Obj *out)
{
    return node(h, o->laddr[c], out, o->lname[c]);
}

/* the first child with the given label (and name, unless NULL) */
static int find_child(H5 *h, const Obj *o, const char *label, const char *name, Obj *out)
{
    for (int c = 0; c < o->nlink; c++) {
        if (!strcmp(o->lname[c], " data")) continue;
        if (name && strcmp(o->lname[c], name)) continue;
        Obj k;
        if (child(h, o, c, &k)) return -1;
        if (!label || !strcmp(k.label, label)) { *out = k; return 0; }
        obj_free(&k);
    }
    return 1;
}

/* BCType_t names as assignBC writes them (src/Load.c:46-97) */
static const char *const bc_names[] = {"BCFarfield", "BCInflow", "BCOutflow", "BCOutflowSubsonic",
                                       "BCWallInviscid", "BCWallViscous", "BCInflowSupersonic",
                                       "BCInflowSubsonic", "BCWallViscousIsothermal"};

static int face_bc(H5 *h, const Obj *base, const Obj *zbc, int dims, int i, int j,
                   const long long offset[3], const long long size[3], char out[30])
{
    for (int c = 0; c < zbc->nlink; c++) {
        if (!strcmp(zbc->lname[c], " data")) continue;
        Obj bc;
        if (child(h, zbc, c, &bc)) return -1;
        if (strcmp(bc.label, "BC_t")) { obj_free(&bc); continue; }
        Obj pr;
        int rc = find_child(h, &bc, "IndexRange_t", "PointRange", &pr);
        if (rc) {
            obj_free(&bc);
            return rc < 0 ? -1 : fail(h, "boundary condition '%s' has no PointRange (the reference "
                                         "reads PointRange boundaries only)", zbc->lname[c]);
        }
        long long pts[6];
        rc = data_ints(h, &pr, pts, (size_t)(2 * dims), "PointRange");
        obj_free(&pr);
        if (rc) { obj_free(&bc); return -1; }
        const long long *lo = pts, *hi = pts + dims;
        int found = lo[i] == hi[i];                                    /* orientation */
        found = found && ((lo[i] == 1) == (j == 0));                   /* plane along the normal */
        if (found)
            for (int dir = (i + 1) % dims; dir != i; dir = (dir + 1) % dims) {   /* tangent overlap */
                const long long low = 1 + offset[dir], high = 1 + size[dir] + offset[dir];
                found = found && ((low > lo[dir] && low < hi[dir]) || (high > lo[dir] && high < hi[dir]) ||
                                  (lo[dir] > low && lo[dir] < high) || (hi[dir] > low && hi[dir] < high));
            }
        if (!found) { obj_free(&bc); continue; }
        char bct[40];
        if (data_string(h, &bc, bct, sizeof bct, zbc->lname[c])) { obj_free(&bc); return -1; }
        if (!strcmp(bct, "FamilySpecified")) {
            Obj fn, fam, fbc;
            char famname[40];
            rc = find_child(h, &bc, "FamilyName_t", NULL, &fn);
            if (rc) { obj_free(&bc); return rc < 0 ? -1 : fail(h, "BC '%s' names no family", zbc->lname[c]); }
            rc = data_string(h, &fn, famname, sizeof famname, "FamilyName");
            obj_free(&fn);
            if (rc) { obj_free(&bc); return -1; }
            rc = find_child(h, base, "Family_t", famname, &fam);
            if (rc) { obj_free(&bc); return rc < 0 ? -1 : fail(h, "(EE) Family '%s' could not be found in file", famname); }
            rc = find_child(h, &fam, "FamilyBC_t", NULL, &fbc);
            obj_free(&fam);
            if (rc) { obj_free(&bc); return rc < 0 ? -1 : fail(h, "family '%s' has no FamilyBC", famname); }
            rc = data_string(h, &fbc, bct, sizeof bct, "FamilyBC");
            obj_free(&fbc);
            if (rc) { obj_free(&bc); return -1; }
        }
        obj_free(&bc);
        for (size_t b = 0; b < sizeof bc_names / sizeof bc_names[0]; b++)
            if (!strcmp(bct, bc_names[b])) {
                strcpy(out, bc_names[b]);
                return 0;
            }
        return fail(h, "(EE) Unsupported BC %s", bct);
    }
    return fail(h, "(EE) BC for side %i (%s end) could not be found", i, j ? "higher" : "lower");
}

static int load(H5 *h, int dims, const struct compat_division *div, double *x, double *y, double *z,
                struct compat_mesh_info *info)
{
    if (h->n < 48 || memcmp(h->p, "\x89HDF\r\n\x1a\n", 8)) return fail(h, "not an HDF5 file");
    if (h->p[8] != 2 && h->p[8] != 3) return fail(h, "HDF5 superblock version %d (2 or 3 read)", h->p[8]);
    if (h->p[9] != 8 || h->p[10] != 8) return fail(h, "HDF5 offsets/lengths of %d/%d bytes", h->p[9], h->p[10]);
    Obj root, base, zone, gc, zbc;
    if (node(h, le(h->p + 36, 8), &root, "root")) return -1;
    int rc = find_child(h, &root, "CGNSBase_t", NULL, &base);
    obj_free(&root);
    if (rc) return rc < 0 ? -1 : fail(h, "no CGNSBase_t node");
    long long bd[2];
    if (data_ints(h, &base, bd, 2, "base")) { obj_free(&base); return -1; }
    info->celldim = (int)bd[0];
    info->physdim = (int)bd[1];
    if (bd[0] != dims) {
        obj_free(&base);
        return fail(h, "(EE) Mesh file dimension %i does not match compile time dimension %i", (int)bd[0], dims);
    }
    info->start_iteration = 0;
    Obj it;
    rc = find_child(h, &base, "Descriptor_t", "Iterations", &it);
    if (rc < 0) { obj_free(&base); return -1; }
    if (rc == 0) {
        char txt[64];
        rc = data_string(h, &it, txt, sizeof txt, "Iterations");
        obj_free(&it);
        if (rc) { obj_free(&base); return -1; }
        info->start_iteration = atol(txt);
    }
    rc = find_child(h, &base, "Zone_t", div->zonename, &zone);
    if (rc) {
        obj_free(&base);
        return rc < 0 ? -1 : fail(h, "(EE) Could not find associated zone '%s' in the mesh file", div->zonename);
    }
    long long zs[9];
    if (data_ints(h, &zone, zs, (size_t)(3 * dims), "zone")) goto bad_zone;
    long long size[3] = {1, 1, 1}, offset[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) info->zone_size[d] = (d < dims) ? zs[d] : 1;
    for (int d = 0; d < dims; d++) {
        size[d] = div->corners[1][d] - div->corners[0][d] + 1;
        offset[d] = div->corners[0][d] - 1;
        if (offset[d] < 0 || size[d] < 1 || offset[d] + size[d] > zs[d]) {
            fail(h, "division block [%lld, %lld] outside the zone's %lld nodes along direction %d",
                 div->corners[0][d], div->corners[1][d], zs[d], d);
            goto bad_zone;
        }
    }
    /* boundary conditions of the faces without a neighbour */
    rc = find_child(h, &zone, "ZoneBC_t", NULL, &zbc);
    if (rc < 0) goto bad_zone;
    const int have_zbc = rc == 0;
    for (int f = 0; f < 6; f++) info->bc[f][0] = '\0';
    for (int i = 0; i < dims; i++)
        for (int j = 0; j < 2; j++) {
            if (div->face[2 * i + j].neighbour) continue;
            if (!have_zbc) {
                fail(h, "(EE) BC for side %i (%s end) could not be found", i, j ? "higher" : "lower");
                goto bad_zone;
            }
            if (face_bc(h, &base, &zbc, dims, i, j, offset, size, info->bc[2 * i + j])) {
                obj_free(&zbc);
                goto bad_zone;
            }
        }
    if (have_zbc) obj_free(&zbc);
    /* coordinates of the block, i fastest */
    rc = find_child(h, &zone, "GridCoordinates_t", "GridCoordinates", &gc);
    if (rc) {
        if (rc > 0) fail(h, "zone '%s' has no GridCoordinates", div->zonename);
        goto bad_zone;
    }
    static const char *const cname[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
    double *const dst[3] = {x, y, z};
    info->coord_real8 = 0;
    for (int a = 0; a < 3; a++) {
        if (a >= dims) {                      /* 2D: z = 0 (src/Functions.c:775-778) */
            if (dst[a]) for (long long t = 0; t < size[0] * size[1] * size[2]; t++) dst[a][t] = 0.0;
            continue;
        }
        if (!dst[a]) continue;
        Obj ca;
        rc = find_child(h, &gc, "DataArray_t", cname[a], &ca);
        if (rc) {
            if (rc > 0) fail(h, "zone '%s' has no %s", div->zonename, cname[a]);
            obj_free(&gc);
            goto bad_zone;
        }
        Data d;
        rc = node_data(h, &ca, &d, cname[a]);
        obj_free(&ca);
        if (!rc && (d.tclass != 1 || (d.tsize != 4 && d.tsize != 8))) rc = fail(h, "%s: RealSingle or RealDouble expected", cname[a]);
        if (!rc) {
            /* HDF5 dimensions are the zone's, reversed */
            int ok = d.rank == dims;
            for (int q = 0; ok && q < dims; q++) ok = (long long)d.dims[dims - 1 - q] == zs[q];
            if (!ok) rc = fail(h, "%s: shape does not match the zone size", cname[a]);
        }
        if (rc) { obj_free(&gc); goto bad_zone; }
        if (d.tsize == 8) info->coord_real8 = 1;
        const long long NI = zs[0], NJ = (dims > 1) ? zs[1] : 1;
        for (long long k2 = 0; k2 < size[2]; k2++)
            for (long long j2 = 0; j2 < size[1]; j2++)
                for (long long i2 = 0; i2 < size[0]; i2++) {
                    const long long src = (i2 + offset[0]) + NI * ((j2 + offset[1]) + NJ * (k2 + offset[2]));
                    const long long t = i2 + size[0] * (j2 + size[1] * k2);
                    if (d.tsize == 4) {
                        float v;
                        memcpy(&v, d.bytes + 4 * src, 4);
                        dst[a][t] = (double)v;
                    } else {
                        memcpy(&dst[a][t], d.bytes + 8 * src, 8);
                    }
                }
    }
    obj_free(&gc);
    obj_free(&zone);
    obj_free(&base);
    return 0;
bad_zone:
    obj_free(&zone);
    obj_free(&base);
    return -1;
}

int compat_load_mesh(const char *path, int dims, const struct compat_division *div, double *x, double *y,
                     double *z, struct compat_mesh_info *info, char *err, int errlen)
{
    H5 h = {NULL, 0, err, errlen};
    if (err && errlen > 0) err[0] = '\0';
    if (dims != 2 && dims != 3) return fail(&h, "dims must be 2 or 3");
    memset(info, 0, sizeof *info);
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(&h, "cannot open '%s'", path);
    struct stat st;
    if (fstat(fd, &st) || st.st_size <= 0) {
        close(fd);
        return fail(&h, "cannot read '%s'", path);
    }
    void *m = mmap(NULL, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) return fail(&h, "cannot map '%s'", path);
    h.p = (const unsigned char *)m;
    h.n = (size_t)st.st_size;
    const int rc = load(&h, dims, div, x, y, z, info);
    munmap(m, (size_t)st.st_size);
    return rc;
}

/* ExtrapolateGhostCells (src/Functions.c:6526-6670) for a block of node coordinates: along each
 * direction in turn (i, then j over the extended i, then k), layer l beyond a face mirrors the
 * spacing of the interior node at the same distance from it, x[-1-l] = (x[-l] + x[l]) - x[l+1]. */
static void extend_axis(const double *src, double *dst, const long long sh[3], int axis, int gx)
{
    long long dsh[3] = {sh[0], sh[1], sh[2]};
    dsh[axis] += 2 * gx;
    const long long n = sh[axis];
    for (long long a = 0; a < dsh[0]; a++)
        for (long long b = 0; b < dsh[1]; b++)
            for (long long c = 0; c < dsh[2]; c++) {
                const long long idx[3] = {a, b, c};
                const long long p = idx[axis] - gx;          /* position along the axis, source frame */
                if (p < 0 || p >= n) continue;
                long long s[3] = {a, b, c};
                s[axis] = p;
                dst[(a * dsh[1] + b) * dsh[2] + c] = src[(s[0] * sh[1] + s[1]) * sh[2] + s[2]];
            }
    const long long st = (axis == 0) ? dsh[1] * dsh[2] : (axis == 1) ? dsh[2] : 1;
    for (long long a = 0; a < dsh[0]; a++)
        for (long long b = 0; b < dsh[1]; b++)
            for (long long c = 0; c < dsh[2]; c++) {
                const long long idx[3] = {a, b, c};
                if (idx[axis] != 0) continue;                 /* one pass per line along the axis */
                double *L = dst + (a * dsh[1] + b) * dsh[2] + c;
                for (int l = 0; l < gx; l++) {
                    const long long lo = gx - 1 - l, hi = gx + n + l;
                    L[lo * st] = (L[(lo + 1) * st] + L[(gx + l) * st]) - L[(gx + l + 1) * st];
                    L[hi * st] = (L[(hi - 1) * st] + L[(gx + n - 1 - l) * st]) - L[(gx + n - 2 - l) * st];
                }
            }
}

int compat_extend_coords(int dims, int space_order, const long long n[3], const double *x, const double *y,
                         const double *z, double *xe, double *ye, double *ze)
{
    if ((dims != 2 && dims != 3) || (space_order != 5 && space_order != 9)) return -1;
    const int gx = (space_order + 1) / 2 + 1;
    const long long ni = n[0], nj = n[1], nk = (dims == 3) ? n[2] : 1;
    if (ni < 2 || nj < 2 || nk < 1 || (dims == 3 && nk < 2)) return -1;
    const long long E[3] = {ni + 2 * gx, nj + 2 * gx, (dims == 3) ? nk + 2 * gx : 1};
    const size_t tot = (size_t)(E[0] * E[1] * E[2]);
    const double *in[3] = {x, y, z};
    double *out[3] = {xe, ye, ze};
    double *t0 = (double *)malloc(tot * sizeof(double)), *t1 = (double *)malloc(tot * sizeof(double));
    if (!t0 || !t1) { free(t0); free(t1); return -1; }
    for (int a = 0; a < 3; a++) {
        if (a == 2 && dims == 2) {
            for (size_t q = 0; q < tot; q++) out[a][q] = 0.0;
            continue;
        }
        /* i-fastest input (as compat_load_mesh gives it) into k-fastest (i, j, k) order */
        long long sh[3] = {ni, nj, nk};
        for (long long i = 0; i < ni; i++)
            for (long long j = 0; j < nj; j++)
                for (long long k = 0; k < nk; k++) t0[(i * nj + j) * nk + k] = in[a][i + ni * (j + nj * k)];
        double *cur = t0, *nxt = t1;
        for (int d = 0; d < dims; d++) {
            extend_axis(cur, nxt, sh, d, gx);
            sh[d] += 2 * gx;
            double *sw = cur; cur = nxt; nxt = sw;
        }
        memcpy(out[a], cur, tot * sizeof(double));
    }
    free(t0);
    free(t1);
    return 0;
}
