/* dropin_caller.c -- the call sequence of the reference's main()
 * (src/executable.c:62-81) made through include/doci_amd.h: fill_doci_data,
 * doci_get_basis_size, doci_get_diagonal, the Hartree-Fock guess
 * (src/executable.c:276-288: unit vector at the first minimum, scanned from
 * D[0] + 1), sparse_eigensolve with doci_matvec, doci_cleanup_data.
 *
 *   dropin_caller FCIDUMP SOLVER TOL MAX_ITS KEEP MAX_VEC
 *
 * Prints "basis n", "code rc", "energy E" (%.15e) and "norm |x|" after
 * whatever the solver itself prints. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/doci_amd.h"

int main(int argc, char **argv)
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s FCIDUMP SOLVER TOL MAX_ITS KEEP MAX_VEC\n", argv[0]);
        return 2;
    }
    setvbuf(stdout, NULL, _IOLBF, BUFSIZ);
    void *dat = NULL;
    fill_doci_data(argv[1], &dat);
    const int n = doci_get_basis_size(dat);
    double *diag = doci_get_diagonal(dat);
    double *x = calloc((size_t)n, sizeof(double));
    if (!x) return 3;
    int ind = 0;
    double low = diag[0] + 1;
    for (int i = 0; i < n; ++i)
        if (low > diag[i]) {
            low = diag[i];
            ind = i;
        }
    x[ind] = 1;
    double energy = 0;
    const int rc = sparse_eigensolve(x, n, &energy, doci_matvec, diag, NULL, atof(argv[3]),
                                     atoi(argv[4]), argv[2], atoi(argv[5]), atoi(argv[6]), dat);
    double s = 0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    printf("basis %d\ncode %d\nenergy %.15e\nnorm %.15e\n", n, rc, energy, sqrt(s));
    doci_cleanup_data(&dat);
    if (dat != NULL) return 4;
    free(x);
    return 0;
}
