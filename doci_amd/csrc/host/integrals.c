/* integrals.c -- binomials (src/doci.c:687-701).  The integral transform's
 * results are read without unpacking the tensor: dh_build_tables in
 * operator.c. */
#include <stddef.h>

#include "internal.h"

uint64_t dh_binom(unsigned k, unsigned n)
{
    if (k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (unsigned d = 1; d <= k; ++d) r = r * (n - k + d) / d;
    return r > (unsigned __int128)UINT64_MAX ? UINT64_MAX : (uint64_t)r;
}
