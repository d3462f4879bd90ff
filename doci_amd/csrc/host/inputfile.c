/* inputfile.c -- the reference's key = value input file (src/executable.c:201-230).
 *
 * Each line is tried against every key with the same scanf patterns, so
 * whitespace around '=' is optional, '#' lines match nothing and an empty
 * value keeps the default.  Two keys of this library are added
 * (DOCI_PRECISION = FP64 | MIXED, and EIGENVECTOR_FILE = path: where the
 * CLI writes the converged eigenvector); the reference's parser ignores
 * them. */
#include <stdio.h>
#include <string.h>

#include "internal.h"

int doci_read_inputfile(const char *path, doci_input *in)
{
    char line[255], prec[32] = "";
    FILE *fp = fopen(path, "r");
    if (!fp) return doci_fail(DOCI_E_IO, "Error reading input file");
    memset(in, 0, sizeof *in);
    in->max_its = 200;
    strcpy(in->solver, "D");
    in->tol = 1e-8;
    in->davidson_keep = 2;
    in->davidson_max_vec = 30;
    in->hf_init = 1;
    in->precision = 0;
    while (fgets(line, sizeof line, fp)) {
        sscanf(line, " FCIDUMP = %254s ", in->fcidump);
        sscanf(line, " SOLVER = %19s ", in->solver);
        sscanf(line, " MAX_ITS = %d ", &in->max_its);
        sscanf(line, " TOL = %lf ", &in->tol);
        sscanf(line, " DAVIDSON_KEEP = %d ", &in->davidson_keep);
        sscanf(line, " DAVIDSON_MAX_VEC = %d ", &in->davidson_max_vec);
        sscanf(line, " HF_INIT = %d ", &in->hf_init);
        sscanf(line, " EIGENVECTOR_FILE = %254s ", in->eigvec);
        if (sscanf(line, " DOCI_PRECISION = %31s ", prec) == 1)
            in->precision = !strcmp(prec, "MIXED") || !strcmp(prec, "mixed");
    }
    fclose(fp);
    return DOCI_OK;
}
