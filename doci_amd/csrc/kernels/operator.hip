// operator.hip -- DOCI diagonal and sigma = H c on gfx950.
//
// Diagonal: one lane per determinant, the reference's exact addition order
// (src/doci.c:589-624) so the result is bit-identical.
// Sigma: pull form (each lane owns one output determinant and gathers its
// N(L-N) neighbours), so there are no atomics and no races; the reference's
// own precedent for this mapping is the full-row walk at src/doci.c:310-313.
#include <hip/hip_runtime.h>

#include "../dk.h"
#include "addressing.hpp"

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void stage_u32(uint32_t *dst, const uint32_t *src, int cnt)
{
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
}
__device__ __forceinline__ void stage_f64(double *dst, const double *src, int cnt)
{
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_diagonal(dk_tables t, double *__restrict__ D)
{
    extern __shared__ double lds[];
    const int L = (int)t.L, N = (int)t.N, N2 = N + 2, LL = L * L;
    // XJ[2 (o L + q) + is]: X[o][q] for is = 0, 2 J[o][q] (exact) for is = 1
    double *XJ = lds, *h = XJ + 2 * LL;
    uint32_t *bt = (uint32_t *)(h + L);
    for (int i = threadIdx.x; i < LL; i += blockDim.x) {
        XJ[2 * i] = t.X[i];
        XJ[2 * i + 1] = 2 * t.J[i];
    }
    stage_f64(h, t.h, L);
    stage_u32(bt, t.binom, (L + 1) * N2);
    __syncthreads();
    const uint32_t I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= t.n) return;
    const dk::Binom C{bt, N2};
    D[I] = dk::diag_of(dk::unrank(I, L, N, C), L, t.core, XJ, h);
}

// kSep: the separable rank form (N <= 12, L - N <= 16), else the branch-free walk
template <typename T, bool kSep>
__global__ __launch_bounds__(kBlock) void k_sigma_lane(dk_tables t, const double *__restrict__ D,
                                                       double shift, const T *__restrict__ c,
                                                       T *__restrict__ sigma)
{
    extern __shared__ double lds[];
    const int L = (int)t.L, N = (int)t.N, N2 = N + 2;
    double *K = lds;
    uint32_t *bt = (uint32_t *)(K + L * L);
    stage_f64(K, t.K, L * L);
    stage_u32(bt, t.binom, (L + 1) * N2);
    __syncthreads();
    const uint32_t I = blockIdx.x * kBlock + threadIdx.x;
    if (I >= t.n) return;
    const dk::Binom C{bt, N2};
    const uint64_t m = dk::unrank(I, L, N, C);
    double acc = (D[I] - shift) * (double)c[I];
    if (kSep) {
        // four gathers in flight per group of the walk, through a buffer
        // resource (the discarded steps of a group need no in-range index)
        const __amdgpu_buffer_rsrc_t rc = dk::vec_rsrc(c, (uint64_t)t.n * sizeof(T));
        dk::for_each_hop_sep<12, 16, 4>(I, m, L, N, C, [&](int o, int tt, uint32_t Jr, bool hop) {
            const double v = K[o * L + tt] * dk::vec_load<T>(rc, Jr);
            acc += hop ? v : 0.0;
        });
    } else {
        // branch-free walk: a step whose t is occupied reads c[I] and adds 0
        dk::for_each_hop_pred(I, m, L, N, C, [&](int o, int tt, uint32_t Jr, bool hop) {
            const double v = K[o * L + tt] * (double)c[hop ? Jr : I];
            acc += hop ? v : 0.0;
        });
    }
    sigma[I] = (T)acc;
}

} // namespace

extern "C" int dk_diagonal(const dk_tables *t, double *D, hipStream_t s)
{
    const size_t lds = sizeof(double) * (2 * t->L * t->L + t->L) +
                       sizeof(uint32_t) * (t->L + 1) * (t->N + 2);
    const unsigned blocks = (t->n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_diagonal, dim3(blocks), dim3(kBlock), lds, s, *t, D);
    return (int)hipGetLastError();
}

extern "C" int dk_sigma(const dk_tables *t, const double *D, double shift, const void *c,
                        void *sigma, int prec, hipStream_t s)
{
    const size_t lds = sizeof(double) * t->L * t->L + sizeof(uint32_t) * (t->L + 1) * (t->N + 2);
    const unsigned blocks = (t->n + kBlock - 1) / kBlock;
    const bool sep = t->N <= 12 && t->L - t->N <= 16;
    if (prec == 0 && sep)
        hipLaunchKernelGGL((k_sigma_lane<double, true>), dim3(blocks), dim3(kBlock), lds, s, *t, D,
                           shift, (const double *)c, (double *)sigma);
    else if (prec == 0)
        hipLaunchKernelGGL((k_sigma_lane<double, false>), dim3(blocks), dim3(kBlock), lds, s, *t, D,
                           shift, (const double *)c, (double *)sigma);
    else if (sep)
        hipLaunchKernelGGL((k_sigma_lane<float, true>), dim3(blocks), dim3(kBlock), lds, s, *t, D,
                           shift, (const float *)c, (float *)sigma);
    else
        hipLaunchKernelGGL((k_sigma_lane<float, false>), dim3(blocks), dim3(kBlock), lds, s, *t, D,
                           shift, (const float *)c, (float *)sigma);
    return (int)hipGetLastError();
}
