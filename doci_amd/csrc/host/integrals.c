/* integrals.c -- integral transform and binomials (src/doci.c:630-665, 687-701).
 *
 * The transform keeps the reference's exact loop and summation order: the
 * diagonal is required to be bit-identical, and it depends on h'_ii. */
#include <stddef.h>

#include "internal.h"

void dh_transform_integrals(double *one, double *two, unsigned Lu)
{
    const size_t L = Lu, L2 = L * L, L3 = L2 * L;
    /* h_ij from the lower triangle of the file (doci.c:636-638) */
    for (size_t i = 0; i < L; ++i)
        for (size_t j = 0; j <= i; ++j) one[i * L + j] = one[j * L + i];
    /* eight-fold unpacking of each stored (ij|kl), i>=j, i>=k, k>=l (doci.c:640-655) */
    for (size_t i = 0; i < L; ++i)
        for (size_t j = 0; j <= i; ++j)
            for (size_t k = 0; k <= i; ++k)
                for (size_t l = 0; l <= k; ++l) {
                    const double v = two[i + L * j + L2 * k + L3 * l];
                    if (v == 0) continue;
                    two[k + L * l + L2 * i + L3 * j] = v;
                    two[j + L * i + L2 * l + L3 * k] = v;
                    two[l + L * k + L2 * j + L3 * i] = v;
                    two[j + L * i + L2 * k + L3 * l] = v;
                    two[l + L * k + L2 * i + L3 * j] = v;
                    two[i + L * j + L2 * l + L3 * k] = v;
                    two[k + L * l + L2 * j + L3 * i] = v;
                }
    /* h'_kl = h_kl - 1/2 sum_j (kj|jl), j ascending (doci.c:657-664) */
    for (size_t k = 0; k < L; ++k)
        for (size_t l = 0; l < L; ++l) {
            double acc = 0;
            for (size_t j = 0; j < L; ++j) acc += two[k + j * L + j * L2 + l * L3];
            one[k + l * L] -= 0.5 * acc;
        }
}

uint64_t dh_binom(unsigned k, unsigned n)
{
    if (k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (unsigned d = 1; d <= k; ++d) r = r * (n - k + d) / d;
    return r > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)r;
}
