/* vecfile.c -- eigenvector file: "DOCI_VECTOR L N n\n" then n FP64 values
 * (the host's little-endian byte order), in the basis's determinant order. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

int doci_write_vector(const char *path, unsigned L, unsigned N, int64_t n, const double *x)
{
    if (!path || !x || n < 1) return doci_fail(DOCI_E_ARG, "doci_write_vector arguments");
    FILE *fp = fopen(path, "wb");
    if (!fp) return doci_fail(DOCI_E_IO, "cannot open %s for writing", path);
    const int ok = fprintf(fp, "DOCI_VECTOR %u %u %lld\n", L, N, (long long)n) > 0 &&
                   fwrite(x, sizeof(double), (size_t)n, fp) == (size_t)n;
    if (fclose(fp) != 0 || !ok) return doci_fail(DOCI_E_IO, "writing %s failed", path);
    return DOCI_OK;
}

int doci_read_vector(const char *path, unsigned *L, unsigned *N, int64_t *n, double **x)
{
    if (!path || !L || !N || !n || !x) return doci_fail(DOCI_E_ARG, "doci_read_vector arguments");
    *x = NULL;
    FILE *fp = fopen(path, "rb");
    if (!fp) return doci_fail(DOCI_E_IO, "cannot open %s", path);
    long long nn = 0;
    if (fscanf(fp, "DOCI_VECTOR %u %u %lld", L, N, &nn) != 3 || nn < 1 || fgetc(fp) != '\n') {
        fclose(fp);
        return doci_fail(DOCI_E_FORMAT, "%s: not a DOCI_VECTOR file", path);
    }
    double *v = malloc(sizeof(double) * (size_t)nn);
    if (!v) {
        fclose(fp);
        return doci_fail(DOCI_E_NOMEM, "vector of %lld values", nn);
    }
    const size_t got = fread(v, sizeof(double), (size_t)nn, fp);
    fclose(fp);
    if (got != (size_t)nn) {
        free(v);
        return doci_fail(DOCI_E_FORMAT, "%s: %zu of %lld values", path, got, nn);
    }
    *n = nn;
    *x = v;
    return DOCI_OK;
}
