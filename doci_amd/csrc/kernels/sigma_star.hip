// sigma_star.hip -- sigma = (H - shift) c in two passes (star pass, main
// pass): the launcher and the C API; the kernels are in star_kernels.hpp.
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "star_kernels.hpp"

namespace {

template <typename T, bool kChecked>
int launch(const dk_tables *t, const dk_stars *st, const double *D, double shift, const T *c,
           T *sigma, T *Z, uint32_t *bad, const dk_sigma_io *io, hipStream_t s)
{
    const size_t ls = star_lds<T>(t->L, t->N, *st, nullptr, nullptr);
    if (st->b > 16 || st->a > 64) return (int)hipErrorInvalidValue;
    const int a = (int)st->a, b = (int)st->b, N = (int)t->N;
    int j = 0;
    if (sizeof(T) == 4 && st->pass == 3 && a >= 1 && a <= 32) {
        // FP32: the column pass on the FP32 matrix cores, one launch
        const int yk = (a + 3) / 4;
        const dim3 grid(st->nsgroup);
        switch (yk) {
#define DK_M32_CASE(Y)                                                                             \
    case Y:                                                                                        \
        hipLaunchKernelGGL((k_star_m32<kChecked, Y>), grid, dim3(kBlockM), ls, s, *t, *st,        \
                           (const float *)c, (float *)Z, bad, io);                                 \
        break;
            DK_M32_CASE(1)
            DK_M32_CASE(2)
            DK_M32_CASE(3)
            DK_M32_CASE(4)
            DK_M32_CASE(5)
            DK_M32_CASE(6)
            DK_M32_CASE(7)
            DK_M32_CASE(8)
#undef DK_M32_CASE
        }
        const int e = (int)hipGetLastError();
        if (e) return e;
        j = N; // the other star passes below are skipped
    }
    if ((st->pass == 0 || st->pass == 3 || st->pass == 4) && j < N && a >= 1 && a <= kColAMax) {
        // one launch for every star group: K packed behind the slots, then the column pass
        // (buffer-addressed where the vector and every members' Z sector fit 32-bit offsets)
        double *kp = (double *)((char *)Z + sizeof(double) * (size_t)st->zoff[a + 1]);
        const dim3 grid(st->nsgroup);
        const size_t ls = star_lds<T, false>(t->L, t->N, *st, nullptr, nullptr); // no K in LDS
        const int buf = st->pass != 4 && (uint64_t)t->n * sizeof(T) < kBufLimit &&
                        (uint64_t)st->zsec_max * sizeof(T) < kBufLimit;
        int done = -1;
        if (col_exact_shape(a, b)) {
            done = dk_col_exact_lo(sizeof(T) == 4, kChecked, buf, a, b, grid, ls, s, *t, *st, c, Z,
                                   bad, io, kp);
            if (done < 0)
                done = dk_col_exact_hi(sizeof(T) == 4, kChecked, buf, a, b, grid, ls, s, *t, *st, c,
                                       Z, bad, io, kp);
        }
        const int am = done < 0 ? col_amax(a) : 0, bm = col_bmax(b);
#define DK_COL_CASE(AM, BM)                                                                        \
    if (am == AM && bm == BM) {                                                                    \
        hipLaunchKernelGGL((k_kpack<AM, BM>), dim3(1), dim3(kBlock), 0, s, *t, a, b, kp);          \
        if (buf)                                                                                   \
            hipLaunchKernelGGL((k_star_colb<T, kChecked, AM, BM, false>), grid, dim3(kBlock), ls, s, *t,  \
                               *st, c, Z, bad, io, (const double *)kp);                            \
        else                                                                                       \
            hipLaunchKernelGGL((k_star_col<T, kChecked, AM, BM>), grid, dim3(kBlock), ls, s, *t,   \
                               *st, c, Z, bad, io, (const double *)kp);                            \
    }
        DK_COL_CASE(8, 12)
        DK_COL_CASE(8, 16)
        DK_COL_CASE(16, 12)
        DK_COL_CASE(16, 16)
        DK_COL_CASE(24, 12)
        DK_COL_CASE(24, 16)
        DK_COL_CASE(32, 12)
        DK_COL_CASE(32, 16)
#undef DK_COL_CASE
        const int e = (int)hipGetLastError();
        if (e) return e;
        j = N; // the per-star paths below are skipped
    }
    if (st->pass == 2) {
        // one launch per run of j sectors sharing the row-tile count ceil((a - j) / 16)
        while (j < N) {
            const int xt = (a - j + 15) / 16;
            int j1 = j + 1;
            while (j1 < N && (a - j1 + 15) / 16 == xt) ++j1;
            const uint32_t g0 = st->jgroup[j], g1 = st->jgroup[j1];
            j = j1;
            if (g1 <= g0) continue;
            const dim3 grid(g1 - g0);
            if (xt <= 1)
                hipLaunchKernelGGL((k_star_mfma<T, kChecked, 1>), grid, dim3(kBlock), ls, s, *t, *st,
                                   g0, c, Z, bad, io);
            else if (xt <= 2)
                hipLaunchKernelGGL((k_star_mfma<T, kChecked, 2>), grid, dim3(kBlock), ls, s, *t, *st,
                                   g0, c, Z, bad, io);
            else
                hipLaunchKernelGGL((k_star_mfma<T, kChecked, 4>), grid, dim3(kBlock), ls, s, *t, *st,
                                   g0, c, Z, bad, io);
            const int e = (int)hipGetLastError();
            if (e) return e;
        }
        j = N; // the vector-FMA path below is skipped
    }
    // one launch per run of j sectors sharing a bucket YMAX = 8 * ceil((a - j) / 8)
    while (j < N) {
        const int bucket = (a - j + 7) / 8;
        int j1 = j + 1;
        while (j1 < N && (a - j1 + 7) / 8 == bucket) ++j1;
        const uint32_t g0 = st->jgroup[j], g1 = st->jgroup[j1];
        j = j1;
        if (g1 <= g0) continue;
        const dim3 grid(g1 - g0);
        switch (bucket) {
#define DK_STAR_CASE(B)                                                                            \
    case B:                                                                                        \
        hipLaunchKernelGGL((k_star<T, kChecked, 8 * B, 16>), grid, dim3(kBlock), ls, s, *t, *st,  \
                           g0, c, Z, bad, io);                                                     \
        break;
            DK_STAR_CASE(1)
            DK_STAR_CASE(2)
            DK_STAR_CASE(3)
            DK_STAR_CASE(4)
            DK_STAR_CASE(5)
            DK_STAR_CASE(6)
            DK_STAR_CASE(7)
            DK_STAR_CASE(8)
#undef DK_STAR_CASE
        default:
            return (int)hipErrorInvalidValue;
        }
        const int e = (int)hipGetLastError();
        if (e) return e;
    }
    const size_t lm = main_lds<T>(t->L, t->N, *st);
    hipLaunchKernelGGL((k_star_main<T, kChecked>), dim3(st->ngroup), dim3(kBlock), lm, s, *t, *st,
                       D, shift, c, (const T *)Z, sigma, bad, io);
    return (int)hipGetLastError();
}

} // namespace

extern "C" size_t dk_star_kp_doubles(uint32_t a, uint32_t b)
{
    if (a < 1 || a > (uint32_t)kColAMax) return 0;
    if (col_exact_shape((int)a, (int)b)) return col_kp_doubles((int)a, (int)b);
    return col_kp_doubles(col_amax((int)a), col_bmax((int)b));
}

extern "C" size_t dk_sigma_star_lds(const dk_tables *t, const dk_stars *st, int prec, int pass)
{
    if (pass == 0)
        return prec ? star_lds<float>(t->L, t->N, *st, nullptr, nullptr)
                    : star_lds<double>(t->L, t->N, *st, nullptr, nullptr);
    return prec ? main_lds<float>(t->L, t->N, *st) : main_lds<double>(t->L, t->N, *st);
}

extern "C" int dk_sigma_star(const dk_tables *t, const dk_stars *st, const double *D, double shift,
                             const void *c, void *sigma, void *Z, int prec, uint32_t *bad,
                             hipStream_t s)
{
    if (st->ngroup == 0) return (int)hipErrorInvalidValue;
    if (prec == 0)
        return bad ? launch<double, true>(t, st, D, shift, (const double *)c, (double *)sigma,
                                          (double *)Z, bad, nullptr, s)
                   : launch<double, false>(t, st, D, shift, (const double *)c, (double *)sigma,
                                           (double *)Z, bad, nullptr, s);
    return bad ? launch<float, true>(t, st, D, shift, (const float *)c, (float *)sigma, (float *)Z,
                                     bad, nullptr, s)
               : launch<float, false>(t, st, D, shift, (const float *)c, (float *)sigma, (float *)Z,
                                      bad, nullptr, s);
}

// ---- the slot buffer of the entry point that allocates nothing ----------
// doci_op_sigma_dev_noalloc runs this file's star sigma with its slot buffer
// in `scratch`, a file-scope device array instead of device memory the
// library allocates: every GPU the library runs on holds one copy (64 MiB)
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not a device address, so the address of each device's copy is looked up
// once per device (hipGetSymbolAddress, that device current) and only that
// address is passed to the kernels.  The column star pass writes every slot
// the main pass reads, so the array needs no clearing between calls; calls
// that overlap on one GPU share it and race.
constexpr size_t kStaticSlotDoubles = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kStaticSlotDoubles];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices];
pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;
} // namespace

extern "C" size_t dk_static_slot_doubles(void) { return kStaticSlotDoubles; }

extern "C" int dk_static_slot_buffer(void **addr)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return (int)hipErrorInvalidDevice;
    pthread_mutex_lock(&g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p) g_scratch_addr[dev] = p;
        else if (e == hipSuccess) e = hipErrorInvalidSymbol;
    }
    *addr = g_scratch_addr[dev];
    pthread_mutex_unlock(&g_scratch_mu);
    return (int)e;
}

extern "C" int dk_sigma_star_io(const dk_tables *t, const dk_stars *st, const double *D,
                                const dk_sigma_io *io, void *Z, int prec, hipStream_t s)
{
    if (st->ngroup == 0 || !io) return (int)hipErrorInvalidValue;
    if (prec == 0)
        return launch<double, false>(t, st, D, 0.0, nullptr, nullptr, (double *)Z, nullptr, io, s);
    return launch<float, false>(t, st, D, 0.0, nullptr, nullptr, (float *)Z, nullptr, io, s);
}
