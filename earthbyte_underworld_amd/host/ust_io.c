/*
 * ust_io.c -- PETSc binary dumps of the assembled operators and vectors (the reference's -matdumpdir /
 * writeMatVec.c role, auglag-driver-DGTGD.c:101-113; the format of the reference's .matrix fixtures).
 *
 * PETSc binary, big-endian: a matrix is MAT_FILE_CLASSID (1211216), rows, columns, nonzeros, the row lengths,
 * the column indices and the values (doubles); a vector is VEC_FILE_CLASSID (1211214), length, values.  The
 * matrices are formed on the host from the device's element matrices in the structural pattern of the
 * Q1/P0 stencil (every coupling of two dofs that share an element, stored even where its value is zero):
 * K rows/columns are velocity dofs (node-major, components interleaved), G rows are velocity dofs and its
 * columns the elements.  `free_only` drops the Dirichlet dofs and numbers the rest in order, as the
 * reference's equation numbering does; otherwise every dof is kept (the unit-test fixtures' form).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ust_internal.h"

static void put32(unsigned char *p, int v) {
    const unsigned u = (unsigned)v;
    p[0] = (unsigned char)(u >> 24);
    p[1] = (unsigned char)(u >> 16);
    p[2] = (unsigned char)(u >> 8);
    p[3] = (unsigned char)u;
}

static void put64(unsigned char *p, double d) {
    unsigned long long u;
    memcpy(&u, &d, 8);
    for (int k = 0; k < 8; k++) p[k] = (unsigned char)(u >> (56 - 8 * k));
}

static int write_ints(FILE *fp, const int *v, size_t n) {
    unsigned char b[4];
    for (size_t i = 0; i < n; i++) {
        put32(b, v[i]);
        if (fwrite(b, 1, 4, fp) != 4) return -1;
    }
    return 0;
}

static int write_doubles(FILE *fp, const double *v, size_t n) {
    unsigned char b[8];
    for (size_t i = 0; i < n; i++) {
        put64(b, v[i]);
        if (fwrite(b, 1, 8, fp) != 8) return -1;
    }
    return 0;
}

/* equation numbers of the velocity dofs: -1 for a dropped Dirichlet dof */
static int eq_numbers(const ust_context *ctx, int free_only, int *eq) {
    int c = 0;
    for (int i = 0; i < ctx->ndof; i++) eq[i] = (free_only && ctx->mask_h[i]) ? -1 : c++;
    return c;
}

static int write_csr(ust_context *ctx, const char *path, int rows, int cols, const int *rowlen, const int *col,
                     const double *val, size_t nnz) {
    FILE *fp = fopen(path, "wb");
    if (!fp) UST_FAIL(ctx, UST_EINVAL, "cannot open %s for writing", path);
    int head[4] = {1211216, rows, cols, (int)nnz};
    int rc = write_ints(fp, head, 4);
    if (!rc) rc = write_ints(fp, rowlen, (size_t)rows);
    if (!rc) rc = write_ints(fp, col, nnz);
    if (!rc) rc = write_doubles(fp, val, nnz);
    if (fclose(fp) != 0) rc = -1;
    if (rc) UST_FAIL(ctx, UST_EINVAL, "write to %s failed", path);
    return UST_OK;
}

/* K: one row per kept velocity dof; columns: the kept dofs of the 3^dim node neighbourhood, increasing */
static int dump_K(ust_context *ctx, int free_only, const char *path) {
    const int dim = ctx->grid.dim, nd = (1 << dim) * dim;
    const int ex = ctx->grid.nel[0], ey = ctx->grid.nel[1], ez = dim == 3 ? ctx->grid.nel[2] : 1;
    const int nx = ex + 1, ny = ey + 1, nz = dim == 3 ? ez + 1 : 1;
    const size_t nel = (size_t)ctx->nel;
    int *eq = (int *)malloc((size_t)ctx->ndof * sizeof(int));
    double *Ke = (double *)malloc(nel * nd * nd * sizeof(double));
    size_t cap = (size_t)ctx->ndof * (size_t)(dim == 3 ? 81 : 18);
    int *rowlen = (int *)calloc((size_t)ctx->ndof, sizeof(int));
    int *col = (int *)malloc(cap * sizeof(int));
    double *val = (double *)malloc(cap * sizeof(double));
    int rc = (!eq || !Ke || !rowlen || !col || !val) ? UST_ENOMEM : UST_OK;
    if (!rc) rc = ust_get_element_matrices(ctx, 0, ctx->nel, Ke);
    size_t nnz = 0;
    int rows = 0;
    if (!rc) {
        const int neq = eq_numbers(ctx, free_only, eq);
        for (int n = 0; n < ctx->nnodes; n++) {
            const int ix = n % nx, iy = (n / nx) % ny, iz = n / (nx * ny);
            for (int i = 0; i < dim; i++) {
                const int R = n * dim + i;
                if (eq[R] < 0) continue;
                size_t start = nnz;
                for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int mx = ix + dx, my = iy + dy, mz = iz + dz;
                            if (mx < 0 || mx >= nx || my < 0 || my >= ny || mz < 0 || mz >= nz) continue;
                            const int m = mx + nx * (my + ny * mz);
                            for (int j = 0; j < dim; j++) {
                                const int C = m * dim + j;
                                if (eq[C] < 0) continue;
                                double v = 0.0;
                                /* elements holding both nodes, in element order */
                                for (int ez0 = (dim == 3 ? (iz > mz ? iz : mz) - 1 : 0);
                                     ez0 <= (dim == 3 ? (iz < mz ? iz : mz) : 0); ez0++)
                                    for (int ey0 = (iy > my ? iy : my) - 1; ey0 <= (iy < my ? iy : my); ey0++)
                                        for (int ex0 = (ix > mx ? ix : mx) - 1; ex0 <= (ix < mx ? ix : mx); ex0++) {
                                            if (ex0 < 0 || ex0 >= ex || ey0 < 0 || ey0 >= ey || ez0 < 0 || ez0 >= ez)
                                                continue;
                                            const size_t e = (size_t)ex0 + (size_t)ex * (ey0 + (size_t)ey * ez0);
                                            const int a = (ix - ex0) + 2 * (iy - ey0) + 4 * (iz - ez0);
                                            const int b = (mx - ex0) + 2 * (my - ey0) + 4 * (mz - ez0);
                                            v += Ke[e * nd * nd + (size_t)(a * dim + i) * nd + (b * dim + j)];
                                        }
                                col[nnz] = eq[C];
                                val[nnz] = v;
                                nnz++;
                            }
                        }
                rowlen[rows++] = (int)(nnz - start);
            }
        }
        rc = write_csr(ctx, path, neq, neq, rowlen, col, val, nnz);
    } else if (rc == UST_ENOMEM) {
        snprintf(ctx->err, sizeof(ctx->err), "host allocation failed");
    }
    free(eq);
    free(Ke);
    free(rowlen);
    free(col);
    free(val);
    return rc;
}

/* G: one row per kept velocity dof; columns: the (1 to 2^dim) elements holding its node, increasing */
static int dump_G(ust_context *ctx, int free_only, const char *path) {
    const int dim = ctx->grid.dim, nd = (1 << dim) * dim;
    const int ex = ctx->grid.nel[0], ey = ctx->grid.nel[1], ez = dim == 3 ? ctx->grid.nel[2] : 1;
    const int nx = ex + 1, ny = ey + 1;
    const size_t nel = (size_t)ctx->nel;
    int *eq = (int *)malloc((size_t)ctx->ndof * sizeof(int));
    double *Ge = (double *)malloc(nel * nd * sizeof(double));
    size_t cap = (size_t)ctx->ndof * (size_t)(1 << dim);
    int *rowlen = (int *)calloc((size_t)ctx->ndof, sizeof(int));
    int *col = (int *)malloc(cap * sizeof(int));
    double *val = (double *)malloc(cap * sizeof(double));
    int rc = (!eq || !Ge || !rowlen || !col || !val) ? UST_ENOMEM : UST_OK;
    if (!rc) rc = ust_get_element_G(ctx, 0, ctx->nel, Ge);
    if (!rc) {
        const int neq = eq_numbers(ctx, free_only, eq);
        size_t nnz = 0;
        int rows = 0;
        for (int n = 0; n < ctx->nnodes; n++) {
            const int ix = n % nx, iy = (n / nx) % ny, iz = n / (nx * ny);
            for (int i = 0; i < dim; i++) {
                if (eq[n * dim + i] < 0) continue;
                size_t start = nnz;
                for (int ez0 = (dim == 3 ? iz - 1 : 0); ez0 <= (dim == 3 ? iz : 0); ez0++)
                    for (int ey0 = iy - 1; ey0 <= iy; ey0++)
                        for (int ex0 = ix - 1; ex0 <= ix; ex0++) {
                            if (ex0 < 0 || ex0 >= ex || ey0 < 0 || ey0 >= ey || ez0 < 0 || ez0 >= ez) continue;
                            const size_t e = (size_t)ex0 + (size_t)ex * (ey0 + (size_t)ey * ez0);
                            const int a = (ix - ex0) + 2 * (iy - ey0) + 4 * (iz - ez0);
                            col[nnz] = (int)e;
                            val[nnz] = Ge[e * nd + (size_t)(a * dim + i)];
                            nnz++;
                        }
                rowlen[rows++] = (int)(nnz - start);
            }
        }
        rc = write_csr(ctx, path, neq, ctx->nel, rowlen, col, val, nnz);
    } else if (rc == UST_ENOMEM) {
        snprintf(ctx->err, sizeof(ctx->err), "host allocation failed");
    }
    free(eq);
    free(Ge);
    free(rowlen);
    free(col);
    free(val);
    return rc;
}

int ust_write_petsc_matrix(ust_context *ctx, int which, int free_only, const char *path) {
    if (!ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "assemble first");
    if (!path) UST_FAIL(ctx, UST_EINVAL, "no path");
    if (which == UST_MAT_K) return dump_K(ctx, free_only, path);
    if (which == UST_MAT_G) return dump_G(ctx, free_only, path);
    UST_FAIL(ctx, UST_EINVAL, "unknown matrix %d", which);
}

int ust_write_petsc_vector(ust_context *ctx, int which, int free_only, const char *path) {
    if (!path) UST_FAIL(ctx, UST_EINVAL, "no path");
    if ((which == UST_VEC_F || which == UST_VEC_H) && !ctx->assembled) UST_FAIL(ctx, UST_ESTATE, "assemble first");
    const int velocity = which == UST_VEC_F || which == UST_VEC_U;
    if (!velocity && which != UST_VEC_H && which != UST_VEC_P) UST_FAIL(ctx, UST_EINVAL, "unknown vector %d", which);
    const size_t n = velocity ? (size_t)ctx->ndof : (size_t)ctx->nel;
    const double *src = which == UST_VEC_F ? ctx->f : which == UST_VEC_H ? ctx->h : which == UST_VEC_U ? ctx->u : ctx->p;
    double *v = (double *)malloc(n * sizeof(double));
    int *eq = velocity ? (int *)malloc(n * sizeof(int)) : NULL;
    if (!v || (velocity && !eq)) {
        free(v);
        free(eq);
        UST_FAIL(ctx, UST_ENOMEM, "host allocation failed");
    }
    int rc = ust_memcpy_d2h(v, src, n * sizeof(double), ctx->stream);
    size_t m = n;
    if (!rc && velocity && free_only) {
        eq_numbers(ctx, 1, eq);
        m = 0;
        for (size_t i = 0; i < n; i++)
            if (eq[i] >= 0) v[m++] = v[i];
    }
    if (!rc) {
        FILE *fp = fopen(path, "wb");
        if (!fp) {
            rc = UST_EINVAL;
        } else {
            int head[2] = {1211214, (int)m};
            int w = write_ints(fp, head, 2);
            if (!w) w = write_doubles(fp, v, m);
            if (fclose(fp) != 0 || w) rc = UST_EINVAL;
        }
        if (rc) snprintf(ctx->err, sizeof(ctx->err), "write to %s failed", path);
    } else {
        snprintf(ctx->err, sizeof(ctx->err), "download failed: %s", ust_dev_error_string());
    }
    free(v);
    free(eq);
    return rc;
}
