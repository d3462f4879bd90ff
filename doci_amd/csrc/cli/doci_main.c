/* doci_main.c -- the `doci [-m MODE] INPUTFILE` program (src/executable.c:41-323),
 * with the same input keys, echo and result lines, on the MI355X library. */
#include <argp.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <time.h>

#include "../../../include/doci_amd.h"

const char *argp_program_version = "doci-amd 0.1";

static char doc[] =
    "doci -- DOCI (Doubly Occupied Configuration Interaction) on one AMD MI355X"
    "\n\n"
    "input-file\n"
    "----------\n"
    "\n"
    "FCIDUMP          = path to the FCIDUMP file.\n"
    "\n"
    "SOLVER           = D (Davidson), CG (conjugate gradient) or CGP (CG with\n"
    "                   diagonal preconditioner). (default: D)\n"
    "\n"
    "MAX_ITS          = Maximum number of iterations. (default: 200)\n"
    "\n"
    "TOL              = The tolerance for convergence. (default: 1e-8)\n"
    "\n"
    "DAVIDSON_KEEP    = Vectors kept after deflation. (default: 2)\n"
    "\n"
    "DAVIDSON_MAX_VEC = Subspace size at which Davidson deflates. (default: 30)\n"
    "\n"
    "HF_INIT          = 1: Hartree-Fock initial guess, 0: random. (default: 1)\n"
    "\n"
    "DOCI_PRECISION   = FP64 or MIXED (FP32 storage, FP64 finish). (default: FP64)\n"
    "\n"
    "EIGENVECTOR_FILE = path: the converged eigenvector is written there\n"
    "                   (header line, then FP64 values in determinant order).\n";

static char args_doc[] = "INPUTFILE";

static struct argp_option options[] = {
    {"mode", 'm', "MODE", 0, "0: DOCI (default); other modes are not implemented.", 0},
    {0}};

struct arguments {
    char *inputfile;
    int mode;
};

static error_t parse_opt(int key, char *arg, struct argp_state *state)
{
    struct arguments *a = state->input;
    switch (key) {
    case 'm': a->mode = atoi(arg); break;
    case ARGP_KEY_ARG:
        if (state->arg_num >= 1) argp_usage(state);
        a->inputfile = arg;
        break;
    case ARGP_KEY_END:
        if (state->arg_num < 1) argp_usage(state);
        break;
    default: return ARGP_ERR_UNKNOWN;
    }
    return 0;
}

static struct argp argp = {options, parse_opt, args_doc, doc, 0, 0, 0};

int main(int argc, char *argv[])
{
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    setvbuf(stdout, NULL, _IOLBF, BUFSIZ);
    struct arguments args = {NULL, 0};
    argp_parse(&argp, argc, argv, 0, 0, &args);
    if (args.mode != 0) {
        fprintf(stderr, "Only DOCI works at this moment!\n");
        exit(EXIT_FAILURE);
    }
    doci_input in;
    if (doci_read_inputfile(args.inputfile, &in) != DOCI_OK) {
        fprintf(stderr, "Error reading input file\n");
        exit(EXIT_FAILURE);
    }
    const char *name = !strcmp(in.solver, "D")     ? "DAVIDSON"
                       : !strcmp(in.solver, "CG")  ? "CONJUGATE GRADIENT"
                       : !strcmp(in.solver, "CGP") ? "CONJUGATE GRADIENT WITH DIAGONAL PRECONDITIONER"
                                                   : NULL;
    if (!name) {
        printf("no correct solver inputted!\n");
        exit(EXIT_FAILURE);
    }
    printf("doci -- An implementation of DOCI (Doubly Occupied Configuration Interaction)"
           "\n\n"
           "input-file\n"
           "----------\n"
           "\n"
           "FCIDUMP          = %s\n"
           "\n"
           "SOLVER           = %s\n"
           "\n"
           "MAX_ITS          = %d\n"
           "\n"
           "TOL              = %e\n"
           "\n"
           "DAVIDSON_KEEP    = %d\n"
           "\n"
           "DAVIDSON_MAX_VEC = %d\n"
           "\n"
           "INITIAL          = %s\n\n",
           in.fcidump, name, in.max_its, in.tol, in.davidson_keep, in.davidson_max_vec,
           in.hf_init ? "Hartree-Fock" : "random");
    if (in.precision) setenv("DOCI_PRECISION", "mixed", 1);

    void *dat = NULL;
    fill_doci_data(in.fcidump, &dat);
    const int n = doci_get_basis_size(dat);
    double *diag = doci_get_diagonal(dat);
    double *x = calloc((size_t)n, sizeof(double));
    if (!x) {
        fprintf(stderr, "Failed to allocate the guess vector\n");
        exit(EXIT_FAILURE);
    }
    if (in.hf_init) { /* executable.c:276-288 */
        int ind = 0;
        double low = diag[0] + 1;
        for (int i = 0; i < n; ++i)
            if (low > diag[i]) {
                low = diag[i];
                ind = i;
            }
        x[ind] = 1;
    } else { /* executable.c:290-297 */
        double s = 0;
        srand((unsigned)time(NULL));
        for (int i = 0; i < n; ++i) {
            x[i] = rand();
            s += x[i] * x[i];
        }
        s = 1 / sqrt(s);
        for (int i = 0; i < n; ++i) x[i] *= s;
    }
    printf("--------------------------------------------------\n");
    double energy = 0;
    const int info = sparse_eigensolve(x, n, &energy, doci_matvec, diag, NULL, in.tol, in.max_its,
                                       in.solver, in.davidson_keep, in.davidson_max_vec, dat);
    if (in.eigvec[0]) { /* EIGENVECTOR_FILE: a key of this library */
        const int rc = doci_write_vector(in.eigvec, doci_op_L(dat), doci_op_N(dat), n, x);
        if (rc != DOCI_OK) {
            fprintf(stderr, "%s\n", doci_last_error());
            exit(EXIT_FAILURE);
        }
    }
    gettimeofday(&t1, NULL);
    const double el = (double)(t1.tv_sec - t0.tv_sec) + 1e-6 * (double)(t1.tv_usec - t0.tv_usec);
    printf("\t\t ** EXIT CODE SOLVER : %d\n"
           "\t\t ** %s ENERGY : %.10lf\n"
           "\t\t ** ELAPSED TIME : %f\n"
           "--------------------------------------------------\n"
           "exiting...\n",
           info, "DOCI", energy, el);
    doci_cleanup_data(&dat);
    free(x);
    return EXIT_SUCCESS;
}
