/*
 * ust_h5.c -- FeVariable checkpoint files in HDF5 (FeVariable_SaveToFile and FeVariable_ReadFromFile,
 * StgFEM/Discretisation/src/FeVariable.c:2136-2389 and :2452-2730).
 *
 * The file: root-group attributes "checkpoint file version" (2), "dimensions" and -- for a field on the
 * Cartesian linear mesh -- "mesh resolution" (elements per axis), each a 32-bit big-endian integer array;
 * one dataset "/data" of little-endian doubles with a row per mesh node in global order, holding the node's
 * coordinates first when they are saved, then the field's components.  The velocity and temperature live on
 * the linear mesh (vertices as CartesianGenerator places them, x fastest); the pressure lives on the
 * constant mesh, one node per element at its centre, which is not a Cartesian-generator mesh and so carries no
 * "mesh resolution".  A file without the version attribute is version 1, whose rows start with one extra
 * column before the values (FeVariable.c:2654-2667).
 *
 * The HDF5 C library (1.10 or later: 64-bit hid_t) is loaded with dlopen at the first call, so that the rest
 * of the library does not depend on it: UST_HDF5_LIB names the shared object, else libhdf5.so.103 or
 * libhdf5.so is looked up on the loader's path, then under /opt/conda/lib.  Writes go in one H5Dwrite of the
 * whole row block (the reference writes the same rows one hyperslab at a time); the file holds the same
 * values either way.
 */
#include <dlfcn.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ust_internal.h"

typedef int64_t h5id;
typedef int h5err;
typedef unsigned long long h5size;

#define H5_DEFAULT ((h5id)0) /* H5P_DEFAULT, H5S_ALL and H5E_DEFAULT */
#define H5_ACC_RDONLY 0x0000u
#define H5_ACC_TRUNC 0x0002u

static struct {
    int tried, ok;
    void *so;
    h5err (*open)(void);
    h5err (*libversion)(unsigned *, unsigned *, unsigned *);
    h5err (*eset_auto)(h5id, void *, void *);
    h5id (*fcreate)(const char *, unsigned, h5id, h5id);
    h5id (*fopen)(const char *, unsigned, h5id);
    h5err (*fclose)(h5id);
    h5id (*screate_simple)(int, const h5size *, const h5size *);
    int (*sget_ndims)(h5id);
    int (*sget_dims)(h5id, h5size *, h5size *);
    h5err (*sclose)(h5id);
    h5id (*gopen)(h5id, const char *, h5id);
    h5err (*gclose)(h5id);
    int (*aexists)(h5id, const char *);
    h5id (*acreate)(h5id, const char *, h5id, h5id, h5id, h5id);
    h5id (*aopen)(h5id, const char *, h5id);
    h5id (*aget_space)(h5id);
    h5err (*awrite)(h5id, h5id, const void *);
    h5err (*aread)(h5id, h5id, void *);
    h5err (*aclose)(h5id);
    h5id (*dcreate)(h5id, const char *, h5id, h5id, h5id, h5id, h5id);
    h5id (*dopen)(h5id, const char *, h5id);
    h5id (*dget_space)(h5id);
    h5err (*dwrite)(h5id, h5id, h5id, h5id, h5id, const void *);
    h5err (*dread)(h5id, h5id, h5id, h5id, h5id, void *);
    h5err (*dclose)(h5id);
    const h5id *t_native_double, *t_native_int, *t_std_i32be, *t_ieee_f64le;
} H5;

static char h5_err[512];
static char h5_load_err[256]; /* why the library did not load: reported again by every later call */

static int h5_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(h5_err, sizeof(h5_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *ust_hdf5_error(void) { return h5_err; }

static void *sym(const char *name) {
    void *p = dlsym(H5.so, name);
    if (!p) H5.ok = 0;
    return p;
}

static int h5_open_library(void);

static int h5_load(void) {
    if (H5.tried) return H5.ok ? UST_OK : h5_fail(UST_EINVAL, "%s", h5_load_err);
    H5.tried = 1;
    int rc = h5_open_library();
    if (rc) snprintf(h5_load_err, sizeof(h5_load_err), "%s", h5_err);
    return rc;
}

static int h5_open_library(void) {
    const char *env = getenv("UST_HDF5_LIB");
    const char *names[] = {env, "libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so"};
    for (size_t k = 0; k < sizeof(names) / sizeof(names[0]) && !H5.so; k++)
        if (names[k] && names[k][0]) H5.so = dlopen(names[k], RTLD_NOW | RTLD_LOCAL);
    if (!H5.so) return h5_fail(UST_EINVAL, "no HDF5 C library found (set UST_HDF5_LIB to libhdf5.so)");
    H5.ok = 1;
    *(void **)&H5.open = sym("H5open");
    *(void **)&H5.libversion = sym("H5get_libversion");
    *(void **)&H5.eset_auto = sym("H5Eset_auto2");
    *(void **)&H5.fcreate = sym("H5Fcreate");
    *(void **)&H5.fopen = sym("H5Fopen");
    *(void **)&H5.fclose = sym("H5Fclose");
    *(void **)&H5.screate_simple = sym("H5Screate_simple");
    *(void **)&H5.sget_ndims = sym("H5Sget_simple_extent_ndims");
    *(void **)&H5.sget_dims = sym("H5Sget_simple_extent_dims");
    *(void **)&H5.sclose = sym("H5Sclose");
    *(void **)&H5.gopen = sym("H5Gopen2");
    *(void **)&H5.gclose = sym("H5Gclose");
    *(void **)&H5.aexists = sym("H5Aexists");
    *(void **)&H5.acreate = sym("H5Acreate2");
    *(void **)&H5.aopen = sym("H5Aopen");
    *(void **)&H5.aget_space = sym("H5Aget_space");
    *(void **)&H5.awrite = sym("H5Awrite");
    *(void **)&H5.aread = sym("H5Aread");
    *(void **)&H5.aclose = sym("H5Aclose");
    *(void **)&H5.dcreate = sym("H5Dcreate2");
    *(void **)&H5.dopen = sym("H5Dopen2");
    *(void **)&H5.dget_space = sym("H5Dget_space");
    *(void **)&H5.dwrite = sym("H5Dwrite");
    *(void **)&H5.dread = sym("H5Dread");
    *(void **)&H5.dclose = sym("H5Dclose");
    H5.t_native_double = (const h5id *)sym("H5T_NATIVE_DOUBLE_g");
    H5.t_native_int = (const h5id *)sym("H5T_NATIVE_INT_g");
    H5.t_std_i32be = (const h5id *)sym("H5T_STD_I32BE_g");
    H5.t_ieee_f64le = (const h5id *)sym("H5T_IEEE_F64LE_g");
    if (!H5.ok) return h5_fail(UST_EINVAL, "the HDF5 library found lacks a symbol of the 1.10 C API");
    unsigned maj = 0, mi = 0, rel = 0;
    /* hid_t became 64-bit in 1.10: older libraries have another ABI */
    if (H5.libversion(&maj, &mi, &rel) < 0 || maj != 1 || mi < 10) {
        H5.ok = 0;
        return h5_fail(UST_EINVAL, "HDF5 %u.%u.%u found; 1.10 or later is needed", maj, mi, rel);
    }
    if (H5.open() < 0) {
        H5.ok = 0;
        return h5_fail(UST_EINVAL, "H5open failed");
    }
    H5.eset_auto(H5_DEFAULT, NULL, NULL); /* failures are reported through the return codes */
    return UST_OK;
}

/* one 32-bit big-endian integer attribute of n values on group g */
static int put_int_attr(h5id g, const char *name, const int *v, int n) {
    h5size d = (h5size)n;
    h5id sp = H5.screate_simple(1, &d, NULL);
    if (sp < 0) return -1;
    h5id a = H5.acreate(g, name, *H5.t_std_i32be, sp, H5_DEFAULT, H5_DEFAULT);
    int rc = (a < 0 || H5.awrite(a, *H5.t_native_int, v) < 0) ? -1 : 0;
    if (a >= 0) H5.aclose(a);
    H5.sclose(sp);
    return rc;
}

/* integer attribute into v[cap]: count of values, 0 when absent, -1 on error */
static int get_int_attr(h5id g, const char *name, int *v, int cap) {
    int ex = H5.aexists(g, name);
    if (ex <= 0) return ex == 0 ? 0 : -1;
    h5id a = H5.aopen(g, name, H5_DEFAULT);
    if (a < 0) return -1;
    h5id sp = H5.aget_space(a);
    h5size d[4] = {0, 0, 0, 0};
    int nd = sp >= 0 ? H5.sget_ndims(sp) : -1;
    int rc = -1;
    if (nd >= 0 && nd <= 1) {
        if (nd == 1) H5.sget_dims(sp, d, NULL);
        const int n = nd == 0 ? 1 : (int)d[0];
        if (n >= 1 && n <= cap && H5.aread(a, *H5.t_native_int, v) >= 0) rc = n;
    }
    if (sp >= 0) H5.sclose(sp);
    H5.aclose(a);
    return rc;
}

int ust_hdf5_field_write(const char *path, int dim, const int *res, long rows, int cols, const double *data) {
    int rc = h5_load();
    if (rc) return rc;
    if (!path || (dim != 2 && dim != 3) || rows < 1 || cols < 1 || !data)
        return h5_fail(UST_EINVAL, "ust_hdf5_field_write: bad arguments");
    h5id f = H5.fcreate(path, H5_ACC_TRUNC, H5_DEFAULT, H5_DEFAULT);
    if (f < 0) return h5_fail(UST_EINVAL, "cannot create %s", path);
    const int version = 2;
    h5id g = H5.gopen(f, "/", H5_DEFAULT);
    int bad = g < 0 || put_int_attr(g, "checkpoint file version", &version, 1) || put_int_attr(g, "dimensions", &dim, 1) ||
              (res && put_int_attr(g, "mesh resolution", res, dim));
    if (g >= 0) H5.gclose(g);
    if (!bad) {
        const h5size d[2] = {(h5size)rows, (h5size)cols};
        h5id sp = H5.screate_simple(2, d, NULL);
        h5id ds = sp >= 0 ? H5.dcreate(f, "/data", *H5.t_ieee_f64le, sp, H5_DEFAULT, H5_DEFAULT, H5_DEFAULT) : -1;
        bad = ds < 0 || H5.dwrite(ds, *H5.t_native_double, H5_DEFAULT, H5_DEFAULT, H5_DEFAULT, data) < 0;
        if (ds >= 0) H5.dclose(ds);
        if (sp >= 0) H5.sclose(sp);
    }
    if (H5.fclose(f) < 0) bad = 1;
    return bad ? h5_fail(UST_EINVAL, "writing %s failed", path) : UST_OK;
}

int ust_hdf5_field_info(const char *path, int *version, int *dim, int *res, long *rows, int *cols) {
    int rc = h5_load();
    if (rc) return rc;
    h5id f = H5.fopen(path, H5_ACC_RDONLY, H5_DEFAULT);
    if (f < 0) return h5_fail(UST_EINVAL, "cannot open %s", path);
    int ver = 1, nd = 0, r[3] = {0, 0, 0}, bad = 0;
    h5id g = H5.gopen(f, "/", H5_DEFAULT);
    if (g < 0) {
        bad = 1;
    } else {
        int n = get_int_attr(g, "checkpoint file version", &ver, 1);
        if (n < 0) bad = 1;
        if (n == 0) ver = 1;
        if (!bad && get_int_attr(g, "dimensions", &nd, 1) < 0) bad = 1;
        if (!bad && get_int_attr(g, "mesh resolution", r, 3) < 0) bad = 1;
        H5.gclose(g);
    }
    h5size d[2] = {0, 0};
    if (!bad) {
        h5id ds = H5.dopen(f, "/data", H5_DEFAULT);
        h5id sp = ds >= 0 ? H5.dget_space(ds) : -1;
        if (sp < 0 || H5.sget_ndims(sp) != 2 || H5.sget_dims(sp, d, NULL) < 0) bad = 1;
        if (sp >= 0) H5.sclose(sp);
        if (ds >= 0) H5.dclose(ds);
    }
    H5.fclose(f);
    if (bad) return h5_fail(UST_EINVAL, "%s is not a checkpoint file (root attributes and a 2-D /data)", path);
    if (version) *version = ver;
    if (dim) *dim = nd;
    if (res) memcpy(res, r, sizeof(r));
    if (rows) *rows = (long)d[0];
    if (cols) *cols = (int)d[1];
    return UST_OK;
}

int ust_hdf5_field_read(const char *path, double *data, long rows, int cols) {
    int version, dim, res[3], c;
    long r;
    int rc = ust_hdf5_field_info(path, &version, &dim, res, &r, &c);
    if (rc) return rc;
    if (r != rows || c != cols) return h5_fail(UST_EINVAL, "%s holds %ld x %d values, not %ld x %d", path, r, c, rows, cols);
    h5id f = H5.fopen(path, H5_ACC_RDONLY, H5_DEFAULT);
    if (f < 0) return h5_fail(UST_EINVAL, "cannot open %s", path);
    h5id ds = H5.dopen(f, "/data", H5_DEFAULT);
    int bad = ds < 0 || H5.dread(ds, *H5.t_native_double, H5_DEFAULT, H5_DEFAULT, H5_DEFAULT, data) < 0;
    if (ds >= 0) H5.dclose(ds);
    H5.fclose(f);
    return bad ? h5_fail(UST_EINVAL, "reading /data of %s failed", path) : UST_OK;
}

/* ---- the context's fields ---- */

/* mesh of a field: nodes, components, whether it is the Cartesian linear mesh */
static int field_shape(const ust_context *ctx, int field, long *nodes, int *ncomp, int *cartesian) {
    switch (field) {
        case UST_FIELD_VELOCITY: *nodes = ctx->nnodes; *ncomp = ctx->grid.dim; *cartesian = 1; return 0;
        case UST_FIELD_PRESSURE: *nodes = ctx->nel; *ncomp = 1; *cartesian = 0; return 0;
        case UST_FIELD_TEMPERATURE: *nodes = ctx->nnodes; *ncomp = 1; *cartesian = 1; return 0;
        default: return -1;
    }
}

/* coordinate i of n + 1 vertices on [lo, hi] as CartesianGenerator places them (CartesianGenerator.c:2286-2310,
   no contact depth): the end vertices exactly, the others lo + (i / n) (hi - lo) */
static double vertex_coord(double lo, double hi, int n, long i) {
    if (i <= 0) return lo;
    if (i >= n) return hi;
    return lo + ((double)i / (double)n) * (hi - lo);
}

/* coordinates of node n of the field's mesh: a vertex of the linear mesh, or an element centre, which
   C0Generator takes as the mean of the element's vertices (_Mesh_ElementType_GetCentroid,
   Mesh_ElementType.c:96-125: summed in the element's vertex order, x fastest, times 1 / count) */
static void node_coord(const ust_grid *g, int cartesian, long n, double *x) {
    const int nx = g->nel[0] + (cartesian ? 1 : 0), ny = g->nel[1] + (cartesian ? 1 : 0);
    const long idx[3] = {n % nx, (n / nx) % ny, n / ((long)nx * ny)};
    const int nv = 1 << g->dim;
    for (int d = 0; d < g->dim; d++) {
        if (cartesian) {
            x[d] = vertex_coord(g->xmin[d], g->xmax[d], g->nel[d], idx[d]);
            continue;
        }
        const double lo = vertex_coord(g->xmin[d], g->xmax[d], g->nel[d], idx[d]);
        const double hi = vertex_coord(g->xmin[d], g->xmax[d], g->nel[d], idx[d] + 1);
        double s = 0.0;
        for (int v = 0; v < nv; v++) s = v ? s + (((v >> d) & 1) ? hi : lo) : lo;
        x[d] = s * (1.0 / (double)nv);
    }
}

int ust_save_field_hdf5(ust_context *ctx, int field, int save_coords, const char *path) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    long nodes;
    int ncomp, cart;
    if (field_shape(ctx, field, &nodes, &ncomp, &cart)) UST_FAIL(ctx, UST_EINVAL, "unknown field %d", field);
    const int dim = ctx->grid.dim, cols = ncomp + (save_coords ? dim : 0);
    double *vals = (double *)malloc(sizeof(double) * (size_t)nodes * (size_t)ncomp);
    double *buf = (double *)malloc(sizeof(double) * (size_t)nodes * (size_t)cols);
    if (!vals || !buf) {
        free(vals);
        free(buf);
        UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    const double *src = field == UST_FIELD_VELOCITY ? ctx->u : field == UST_FIELD_PRESSURE ? ctx->p : ctx->T;
    int rc = ust_memcpy_d2h(vals, src, sizeof(double) * (size_t)nodes * (size_t)ncomp, ctx->stream);
    if (!rc) rc = ust_stream_sync(ctx->stream);
    if (rc) {
        free(vals);
        free(buf);
        UST_FAIL(ctx, rc, "device copy of the field failed: %s", ust_dev_error_string());
    }
    for (long n = 0; n < nodes; n++) {
        double *row = buf + (size_t)n * cols;
        if (save_coords) node_coord(&ctx->grid, cart, n, row);
        memcpy(row + (save_coords ? dim : 0), vals + (size_t)n * ncomp, sizeof(double) * (size_t)ncomp);
    }
    free(vals);
    const int res[3] = {ctx->grid.nel[0], ctx->grid.nel[1], ctx->grid.nel[2]};
    rc = ust_hdf5_field_write(path, dim, cart ? res : NULL, nodes, cols, buf);
    free(buf);
    if (rc) UST_FAIL(ctx, rc, "%s", h5_err);
    return UST_OK;
}

int ust_load_field_hdf5(ust_context *ctx, int field, const char *path) {
    if (!ctx->has_mesh) UST_FAIL(ctx, UST_ESTATE, "no mesh");
    long nodes, rows;
    int ncomp, cart, version, dim, res[3], cols;
    if (field_shape(ctx, field, &nodes, &ncomp, &cart)) UST_FAIL(ctx, UST_EINVAL, "unknown field %d", field);
    int rc = ust_hdf5_field_info(path, &version, &dim, res, &rows, &cols);
    if (rc) UST_FAIL(ctx, rc, "%s", h5_err);
    const int gd = ctx->grid.dim;
    /* the checks of FeVariable_ReadFromFile: a known version, the dimension, the resolution of a Cartesian mesh
       (version 2 only), the node count; coordinates are present when the rows are wider than the values */
    if (version != 1 && version != 2) UST_FAIL(ctx, UST_EINVAL, "unknown checkpoint version %d in %s", version, path);
    const int noffset = version == 1 ? 1 : 0;
    if (version == 2) {
        if (dim != gd) UST_FAIL(ctx, UST_EINVAL, "%s has %d dimensions, the mesh %d", path, dim, gd);
        if (cart && res[0] == 0)
            UST_FAIL(ctx, UST_EINVAL, "%s has no mesh resolution attribute, which a linear-mesh field needs", path);
        if (cart)
            for (int d = 0; d < gd; d++)
                if (res[d] != ctx->grid.nel[d])
                    UST_FAIL(ctx, UST_EINVAL, "mesh resolution of %s does not match the mesh (axis %d: %d, %d)", path,
                             d, res[d], ctx->grid.nel[d]);
    }
    if (rows != nodes) UST_FAIL(ctx, UST_EINVAL, "%s holds %ld nodes, the field's mesh %ld", path, rows, nodes);
    const int coords = cols > ncomp + noffset;
    const int first = noffset + (coords ? gd : 0);
    if (first + ncomp > cols) UST_FAIL(ctx, UST_EINVAL, "%s has %d columns, too few for %d components", path, cols, ncomp);
    double *buf = (double *)malloc(sizeof(double) * (size_t)rows * (size_t)cols);
    double *vals = (double *)malloc(sizeof(double) * (size_t)nodes * (size_t)ncomp);
    if (!buf || !vals) {
        free(buf);
        free(vals);
        UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    rc = ust_hdf5_field_read(path, buf, rows, cols);
    if (rc) {
        free(buf);
        free(vals);
        UST_FAIL(ctx, rc, "%s", h5_err);
    }
    for (long n = 0; n < nodes; n++) memcpy(vals + (size_t)n * ncomp, buf + (size_t)n * cols + first, sizeof(double) * (size_t)ncomp);
    free(buf);
    double *dst = field == UST_FIELD_VELOCITY ? ctx->u : field == UST_FIELD_PRESSURE ? ctx->p : ctx->T;
    rc = ust_memcpy_h2d(dst, vals, sizeof(double) * (size_t)nodes * (size_t)ncomp, ctx->stream);
    if (!rc) rc = ust_stream_sync(ctx->stream);
    free(vals);
    if (rc) UST_FAIL(ctx, rc, "device copy of the field failed: %s", ust_dev_error_string());
    if (field == UST_FIELD_TEMPERATURE) ctx->assembled = 0;
    return UST_OK;
}
