// sigma_star_exact_lo.hip -- the column star pass instantiated at the exact
// row counts AMAX = a (a = 12..16), BMAX = 13 of the default star splits
// (kernels in star_kernels.hpp).
#include <hip/hip_runtime.h>

#include "star_kernels.hpp"

namespace {

template <int AM, int BM>
int launch_col(int prec, int checked, int buf, int a, int b, dim3 grid, size_t ls, hipStream_t s,
               const dk_tables &t, const dk_stars &st, const void *c, void *Z, uint32_t *bad,
               const dk_sigma_io *io, double *kp)
{
    hipLaunchKernelGGL((k_kpack<AM, BM>), dim3(1), dim3(kBlock), 0, s, t, a, b, kp);
    if (buf) {
        if (prec == 0 && !checked)
            hipLaunchKernelGGL((k_star_colb<double, false, AM, BM, true>), grid, dim3(kBlock), ls, s, t, st,
                               (const double *)c, (double *)Z, bad, io, (const double *)kp);
        else if (prec == 0)
            hipLaunchKernelGGL((k_star_colb<double, true, AM, BM, true>), grid, dim3(kBlock), ls, s, t, st,
                               (const double *)c, (double *)Z, bad, io, (const double *)kp);
        else if (!checked)
            hipLaunchKernelGGL((k_star_colb<float, false, AM, BM, true>), grid, dim3(kBlock), ls, s, t, st,
                               (const float *)c, (float *)Z, bad, io, (const double *)kp);
        else
            hipLaunchKernelGGL((k_star_colb<float, true, AM, BM, true>), grid, dim3(kBlock), ls, s, t, st,
                               (const float *)c, (float *)Z, bad, io, (const double *)kp);
        return 0;
    }
    if (prec == 0 && !checked)
        hipLaunchKernelGGL((k_star_col<double, false, AM, BM>), grid, dim3(kBlock), ls, s, t, st,
                           (const double *)c, (double *)Z, bad, io, (const double *)kp);
    else if (prec == 0)
        hipLaunchKernelGGL((k_star_col<double, true, AM, BM>), grid, dim3(kBlock), ls, s, t, st,
                           (const double *)c, (double *)Z, bad, io, (const double *)kp);
    else if (!checked)
        hipLaunchKernelGGL((k_star_col<float, false, AM, BM>), grid, dim3(kBlock), ls, s, t, st,
                           (const float *)c, (float *)Z, bad, io, (const double *)kp);
    else
        hipLaunchKernelGGL((k_star_col<float, true, AM, BM>), grid, dim3(kBlock), ls, s, t, st,
                           (const float *)c, (float *)Z, bad, io, (const double *)kp);
    return 0;
}

} // namespace

int dk_col_exact_lo(int prec, int checked, int buf, int a, int b, dim3 grid, size_t lds, hipStream_t s,
                    const dk_tables &t, const dk_stars &st, const void *c, void *Z, uint32_t *bad,
                    const dk_sigma_io *io, double *kp)
{
    if (b != 13) return -1;
    switch (a) {
    case 12: return launch_col<12, 13>(prec, checked, buf, a, b, grid, lds, s, t, st, c, Z, bad, io, kp);
    case 13: return launch_col<13, 13>(prec, checked, buf, a, b, grid, lds, s, t, st, c, Z, bad, io, kp);
    case 14: return launch_col<14, 13>(prec, checked, buf, a, b, grid, lds, s, t, st, c, Z, bad, io, kp);
    case 15: return launch_col<15, 13>(prec, checked, buf, a, b, grid, lds, s, t, st, c, Z, bad, io, kp);
    case 16: return launch_col<16, 13>(prec, checked, buf, a, b, grid, lds, s, t, st, c, Z, bad, io, kp);
    default: return -1;
    }
}
