// vec.hip -- fused streaming kernels for the Davidson and CG iterations.
//
// Every kernel streams its vectors once with a grid-stride loop, keeps FP64
// accumulators in registers and writes one partial per block; a second
// launch reduces the partials in a fixed order, so every reduction is
// deterministic for a given n (independent of scheduling).
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#include "../dk.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
// three workgroups per CU of the 256: streams of several vectors reach a
// higher HBM rate with these twelve waves per CU than with more of them in
// flight (measured at 2^27 elements: FP64 multidot of 8 vectors 1.55 ms at
// 768 workgroups, 1.58 at 2048; residual with W dots 3.9 vs 4.3 ms)
constexpr int kMaxBlocks = 768;
// the residual's 2 + 2m values fit the reduction buffers the host sizes for
// DK_MAX_VEC + 8 values
static_assert(2 + 2 * DK_RESIDUAL_WDOTS_MAX <= DK_MAX_VEC + 8, "residual dots");

struct Ptrs {
    const void *p[DK_MAX_VEC];
};
struct Coef {
    double g[DK_MAX_VEC];
};

inline int grid_for(int64_t n)
{
    int64_t b = (n + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > kMaxBlocks) b = kMaxBlocks;
    return (int)b;
}

// Reduce acc[0..nv) over the block and store one partial per value at
// red[blockIdx.x * nv + k].
template <int NV>
__device__ __forceinline__ void block_partials(double (&acc)[NV], int nv, double *red)
{
    __shared__ double sh[NV][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (k < nv) {
            double v = acc[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) sh[k][wave] = v;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += sh[k][w];
        red[(size_t)blockIdx.x * nv + k] = s;
    }
}

// The same for the residual with W dots: acc holds <r,r>, <t,t>, M slots of
// V dots and M of W dots; the partials are stored packed, 2 + 2m per block
// (the W dots right after the m V dots).
template <int M>
__device__ __forceinline__ void block_partials_vw(double (&acc)[2 + 2 * M], int m, double *red)
{
    __shared__ double sh[2 + 2 * M][kWaves];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 2 + 2 * M; ++k) {
        const int j = k < 2 + M ? k - 2 : k - 2 - M;
        if (j < m) {
            double v = acc[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0) sh[k][wave] = v;
        }
    }
    __syncthreads();
    const int nv = 2 + 2 * m;
    for (int k = threadIdx.x; k < nv; k += blockDim.x) {
        const int kin = k < 2 + m ? k : k + (M - m);
        double s = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += sh[kin][w];
        red[(size_t)blockIdx.x * nv + k] = s;
    }
}

// One block per value: sum the per-block partials in a fixed tree order.
__global__ __launch_bounds__(kBlock) void k_finalize(const double *__restrict__ red, int nblocks,
                                                     int nv, double *__restrict__ out)
{
    __shared__ double sh[kBlock];
    const int k = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) s += red[(size_t)b * nv + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
}

int finalize(const double *red, int nblocks, int nv, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(nv), dim3(kBlock), 0, s, red, nblocks, nv, out);
    return (int)hipGetLastError();
}

// Each lane streams VW consecutive elements per step with 16-byte (or, when
// the per-element registers of an M-vector kernel would not fit, narrower)
// loads; the last n % VW elements take a scalar tail loop.  All vectors are
// 256-byte aligned (device allocations and subspace slots).
template <typename T, int VW> struct Vec {
    typedef T type __attribute__((ext_vector_type(VW)));
};
template <typename T> struct Vec<T, 1> {
    typedef T type;
};
template <typename T, int VW>
__device__ __forceinline__ double el(const typename Vec<T, VW>::type &v, int e)
{
    if constexpr (VW == 1) return (double)v;
    else return (double)v[e];
}
template <typename T, int VW>
__device__ __forceinline__ void set_el(typename Vec<T, VW>::type &v, int e, T x)
{
    if constexpr (VW == 1) v = x;
    else v[e] = x;
}
template <typename T, int VW>
__device__ __forceinline__ typename Vec<T, VW>::type ldv(const void *p, int64_t q)
{
    return ((const typename Vec<T, VW>::type *)p)[q];
}
// elements per lane step: 16 bytes, fewer when 32 or more values per element are kept
template <typename T, int M>
constexpr int vw_for()
{
    constexpr int w = 16 / (int)sizeof(T);
    return M <= 16 ? w : (M <= 32 ? (w >= 2 ? w / 2 : 1) : (w >= 4 ? w / 4 : 1));
}

// The M-vector kernels below run at a compile-time M (a bucket >= m): the
// host fills the vector slots j >= m with the last real vector (or y) and
// their coefficients with 0, so every load and FMA is unconditional and all
// M loads of a step are issued before the first product (a load under a
// `j < m` branch waits for the one before it); the repeated slots read lines
// the same lane has just loaded, and their sums are never stored.
//
// Addressing: the vectors are walked in windows of kWin lane steps; per
// window the base addresses advance once (scalar), and inside it every load
// is base + a 32-bit byte offset shared by all vectors, which the hardware
// takes as an SGPR base plus one VGPR offset (no 64-bit address per vector).

// lane steps per window: 16-byte groups, the FP64 diagonal of an FP32 step
// twice that; 2^26 steps keep every offset below 2^31
constexpr int64_t kWin = int64_t(1) << 26;

template <typename VT>
__device__ __forceinline__ VT ldb(const void *p, uint32_t off)
{
    return *(const VT *)((const char *)p + off);
}
template <typename VT>
__device__ __forceinline__ void stb(void *p, uint32_t off, const VT &v)
{
    *(VT *)((char *)p + off) = v;
}
template <int M>
__device__ __forceinline__ void advance(Ptrs &X, size_t bytes)
{
#pragma unroll
    for (int j = 0; j < M; ++j) X.p[j] = (const char *)X.p[j] + bytes;
}
// an offset the compiler cannot prove equal to its input (a second load of
// the same element is issued again instead of being kept in registers)
__device__ __forceinline__ uint32_t launder(uint32_t v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Runs step(off, VW-tag) over all n elements: windows of VW-element groups
// at byte offsets off from the current bases, adv(bytes) moving the bases
// (one set of them, carried from window to window) past each window, then
// the n % VW tail one element per lane.
template <typename T, int VW, typename Step, typename Adv>
__device__ __forceinline__ void stream(int64_t n, Step &&step, Adv &&adv)
{
    typedef typename Vec<T, VW>::type VT;
    const int64_t nq = n / VW;
    const uint32_t stride = gridDim.x * kBlock, tid = blockIdx.x * kBlock + threadIdx.x;
    for (int64_t w0 = 0; w0 < nq; w0 += kWin) {
        const uint32_t wn = (uint32_t)(nq - w0 < kWin ? nq - w0 : kWin);
        for (uint32_t q = tid; q < wn; q += stride)
            step(q * (uint32_t)sizeof(VT), std::integral_constant<int, VW>());
        adv((size_t)wn * sizeof(VT));
    }
    const uint32_t nt = (uint32_t)(n - nq * VW);
    for (uint32_t i = tid; i < nt; i += stride)
        step(i * (uint32_t)sizeof(T), std::integral_constant<int, 1>());
}

// vectors loaded together before their products: all of them up to 16,
// groups of 8 beyond (the registers of more would cost occupancy)
template <int M>
constexpr int kChunk = M <= 16 ? M : 8;
static_assert(24 % kChunk<24> == 0 && 32 % kChunk<32> == 0 && 64 % kChunk<64> == 0, "whole groups");

// ---- multidot: out[0] = <y,y>, out[1+j] = <X_j, y> -------------------------
template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_multidot(Ptrs X, int m, const T *__restrict__ y,
                                                     int64_t n, double *__restrict__ red)
{
    constexpr int C = kChunk<M>;
    double acc[M + 1];
#pragma unroll
    for (int k = 0; k <= M; ++k) acc[k] = 0;
    const char *yb = (const char *)y;
    stream<T, vw_for<T, M>()>(n, [&](uint32_t off, auto vwt) {
        constexpr int VW = decltype(vwt)::value;
        typedef typename Vec<T, VW>::type VT;
        const VT yv = ldb<VT>(yb, off);
#pragma unroll
        for (int e = 0; e < VW; ++e) acc[0] += el<T, VW>(yv, e) * el<T, VW>(yv, e);
#pragma unroll
        for (int j0 = 0; j0 < M; j0 += C) {
            VT xv[C];
#pragma unroll
            for (int j = 0; j < C; ++j) xv[j] = ldb<VT>(X.p[j0 + j], off);
#pragma unroll
            for (int j = 0; j < C; ++j)
#pragma unroll
                for (int e = 0; e < VW; ++e)
                    acc[1 + j0 + j] += el<T, VW>(xv[j], e) * el<T, VW>(yv, e);
        }
    }, [&](size_t b) { advance<M>(X, b); yb += b; });
    block_partials<M + 1>(acc, m + 1, red);
}

// ---- ortho: y -= sum g_j X_j; out[0] = <y,y>, out[1+j] = <X_j,y> -----------
template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_ortho(Ptrs X, int m, Coef g, T *__restrict__ y,
                                                  int64_t n, double *__restrict__ red)
{
    double acc[M + 1];
#pragma unroll
    for (int k = 0; k <= M; ++k) acc[k] = 0;
    char *yw = (char *)y;
    stream<T, vw_for<T, M>()>(n, [&](uint32_t off, auto vwt) {
        constexpr int VW = decltype(vwt)::value;
        typedef typename Vec<T, VW>::type VT;
        VT xv[M];
        VT yv = ldb<VT>(yw, off);
        double yd[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) yd[e] = el<T, VW>(yv, e);
#pragma unroll
        for (int j = 0; j < M; ++j) xv[j] = ldb<VT>(X.p[j], off);
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int e = 0; e < VW; ++e) yd[e] -= g.g[j] * el<T, VW>(xv[j], e);
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            set_el<T, VW>(yv, e, (T)yd[e]);
            yd[e] = el<T, VW>(yv, e);
            acc[0] += yd[e] * yd[e];
        }
        stb<VT>(yw, off, yv);
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int e = 0; e < VW; ++e) acc[1 + j] += el<T, VW>(xv[j], e) * yd[e];
    }, [&](size_t b) { advance<M>(X, b); yw += b; });
    block_partials<M + 1>(acc, m + 1, red);
}

// ---- Davidson residual (src/davidson.c:195-227 fused) ---------------------
// out[0] = <r,r>, out[1] = <t,t>, out[2+j] = <V_j,t>; with WB also
// <W_j,t> (acc[2+M+j]), from W_j loaded a second time once t is known: the
// lines are still in L2, and holding the W values in registers instead
// costs more occupancy than the second loads cost (measured at L = 34: m = 8
// residual 4.26 ms with the second loads, 4.60 ms with the values held).
template <typename T, int M, bool WB>
__global__ __launch_bounds__(kBlock) void k_residual(Ptrs V, Ptrs W, int m, Coef a, double theta,
                                                     double shift, const double *__restrict__ D,
                                                     T *__restrict__ t, T *__restrict__ xo,
                                                     int64_t n, double *__restrict__ red)
{
    constexpr int C = M <= 8 ? M : (M % 8 == 0 ? 8 : M / 2);
    static_assert(M % C == 0, "whole groups");
    constexpr int NA = 2 + (WB ? 2 : 1) * M;
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; ++k) acc[k] = 0;
    constexpr uint32_t kDs = sizeof(double) / sizeof(T);  // D bytes per T byte
    // 16-byte groups up to 8 vectors, 8-byte groups up to 16, one element
    // beyond (registers)
    constexpr int VWR = M <= 8 ? 16 / (int)sizeof(T) : (M <= 16 ? 8 / (int)sizeof(T) : 1);
    const char *Db = (const char *)D;
    char *tb = (char *)t, *xb = (char *)xo;
    stream<T, VWR>(n, [&](uint32_t off, auto vwt) {
        constexpr int VW = decltype(vwt)::value;
        typedef typename Vec<T, VW>::type VT;
        VT vv[M];
        double x[VW], w[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) x[e] = w[e] = 0;
#pragma unroll
        for (int j0 = 0; j0 < M; j0 += C) {
            VT wv[C];
#pragma unroll
            for (int j = 0; j < C; ++j) {
                vv[j0 + j] = ldb<VT>(V.p[j0 + j], off);
                wv[j] = ldb<VT>(W.p[j0 + j], off);
            }
#pragma unroll
            for (int j = 0; j < C; ++j)
#pragma unroll
                for (int e = 0; e < VW; ++e) {
                    x[e] += a.g[j0 + j] * el<T, VW>(vv[j0 + j], e);
                    w[e] += a.g[j0 + j] * el<T, VW>(wv[j], e);
                }
        }
        typedef typename Vec<double, VW>::type DT;
        const DT dv = ldb<DT>(Db, off * kDs);
        VT tv, xv;
        double td[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            const double r = w[e] - (theta - shift) * x[e];
            const double d = fabs(el<double, VW>(dv, e) - theta);
            const double tval = d > 1e-12 ? r / d : r / 1e-12;
            set_el<T, VW>(tv, e, (T)tval);
            set_el<T, VW>(xv, e, (T)x[e]);
            td[e] = el<T, VW>(tv, e);
            acc[0] += r * r;
            acc[1] += td[e] * td[e];
        }
        stb<VT>(tb, off, tv);
        if (xo) stb<VT>(xb, off, xv);
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int e = 0; e < VW; ++e) acc[2 + j] += el<T, VW>(vv[j], e) * td[e];
        if constexpr (WB) {
            const uint32_t off2 = launder(off);
#pragma unroll
            for (int j0 = 0; j0 < M; j0 += C) {
                VT wv[C];
#pragma unroll
                for (int j = 0; j < C; ++j) wv[j] = ldb<VT>(W.p[j0 + j], off2);
#pragma unroll
                for (int j = 0; j < C; ++j)
#pragma unroll
                    for (int e = 0; e < VW; ++e) acc[2 + M + j0 + j] += el<T, VW>(wv[j], e) * td[e];
            }
        }
    }, [&](size_t b) {
        advance<M>(V, b);
        advance<M>(W, b);
        Db += b * kDs;
        tb += b;
        if (xo) xb += b;
    });
    if constexpr (WB) block_partials_vw<M>(acc, m, red);
    else block_partials<M + 2>(acc, m + 2, red);
}

// ---- lincomb: out = sum g_j X_j ---------------------------------------------
template <typename T, typename To, int M>
__global__ __launch_bounds__(kBlock) void k_lincomb(Ptrs X, int m, Coef g, To *__restrict__ out,
                                                    int64_t n)
{
    constexpr int C = kChunk<M>;
    constexpr uint32_t kNum = sizeof(To), kDen = sizeof(T);  // output bytes per input bytes
    char *ob = (char *)out;
    stream<T, 16 / (int)sizeof(T)>(n, [&](uint32_t off, auto vwt) {
        constexpr int VW = decltype(vwt)::value;
        typedef typename Vec<T, VW>::type VT;
        typedef typename Vec<To, VW>::type OT;
        double s[VW];
#pragma unroll
        for (int e = 0; e < VW; ++e) s[e] = 0;
#pragma unroll
        for (int j0 = 0; j0 < M; j0 += C) {
            VT xv[C];
#pragma unroll
            for (int j = 0; j < C; ++j) xv[j] = ldb<VT>(X.p[j0 + j], off);
#pragma unroll
            for (int j = 0; j < C; ++j)
#pragma unroll
                for (int e = 0; e < VW; ++e) s[e] += g.g[j0 + j] * el<T, VW>(xv[j], e);
        }
        OT o;
#pragma unroll
        for (int e = 0; e < VW; ++e) set_el<To, VW>(o, e, (To)s[e]);
        stb<OT>(ob, off / kDen * kNum, o);
    }, [&](size_t b) { advance<M>(X, b); ob += b / kDen * kNum; });
}

template <typename Ti, typename To>
__global__ __launch_bounds__(kBlock) void k_convert(const Ti *__restrict__ in, To *__restrict__ out,
                                                    int64_t n, double scale)
{
    constexpr int VW = 16 / (int)sizeof(Ti);
    const int64_t stride = (int64_t)gridDim.x * kBlock, nq = n / VW;
    const int64_t tid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (int64_t q = tid; q < nq; q += stride) {
        const auto v = ldv<Ti, VW>(in, q);
#pragma unroll
        for (int e = 0; e < VW; ++e) out[q * VW + e] = (To)(scale * el<Ti, VW>(v, e));
    }
    for (int64_t i = nq * VW + tid; i < n; i += stride) out[i] = (To)(scale * (double)in[i]);
}

// ---- CG (src/CG.c) ---------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cg_grad_elem(double x, double ax, double rho, const double *pinv,
                                             int64_t i, T *g, double (&acc)[2])
{
    double gv = 2 * (ax + (-rho) * x);
    acc[0] += gv * gv;
    if (pinv) gv *= pinv[i];
    const T gq = (T)gv;
    g[i] = gq;
    acc[1] += (double)gq * (double)gq;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_grad(const T *__restrict__ x, const T *__restrict__ Ax,
                                                    double rho, const double *__restrict__ pinv,
                                                    T *__restrict__ g, int64_t n,
                                                    double *__restrict__ red)
{
    double acc[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        cg_grad_elem<T>((double)x[i], (double)Ax[i], rho, pinv, i, g, acc);
    block_partials<2>(acc, 2, red);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_step(T *__restrict__ x, const T *__restrict__ p,
                                                    T *__restrict__ Ax, const T *__restrict__ Ap,
                                                    double z0, double z1, double rho,
                                                    const double *__restrict__ pinv,
                                                    T *__restrict__ g, int64_t n,
                                                    double *__restrict__ red)
{
    double acc[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const T xn = (T)(z0 * (double)x[i] + z1 * (double)p[i]);
        const T axn = (T)(z0 * (double)Ax[i] + z1 * (double)Ap[i]);
        x[i] = xn;
        Ax[i] = axn;
        cg_grad_elem<T>((double)xn, (double)axn, rho, pinv, i, g, acc);
    }
    block_partials<2>(acc, 2, red);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dir(T *__restrict__ p, const T *__restrict__ g,
                                                   double q, int first, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        p[i] = first ? (T)(-(double)g[i]) : (T)(q * (double)p[i] + (-1.0) * (double)g[i]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_gram(const T *__restrict__ x, const T *__restrict__ p,
                                                    const T *__restrict__ Ax,
                                                    const T *__restrict__ Ap, int64_t n,
                                                    double *__restrict__ red)
{
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const double xv = (double)x[i], pv = (double)p[i], ax = (double)Ax[i], ap = (double)Ap[i];
        acc[0] += xv * ax;
        acc[1] += pv * ax;
        acc[2] += xv * ap;
        acc[3] += pv * ap;
        acc[4] += xv * xv;
        acc[5] += pv * xv;
        acc[6] += pv * pv;
    }
    block_partials<7>(acc, 7, red);
}

__global__ __launch_bounds__(kBlock) void k_precond_inv(const double *__restrict__ D,
                                                        double *__restrict__ pinv, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        pinv[i] = fabs(D[i]) > 1e-12 ? 1 / D[i] : 1 / 1e-12;
}

// First index of the minimum: (value, index) pairs, ties to the lower index.
__device__ __forceinline__ void amin_merge(double &v, double &ix, double v2, double ix2)
{
    if (v2 < v || (v2 == v && ix2 < ix)) {
        v = v2;
        ix = ix2;
    }
}

__global__ __launch_bounds__(kBlock) void k_argmin_part(const double *__restrict__ D, int64_t n,
                                                        double *__restrict__ red)
{
    __shared__ double sv[kBlock], si[kBlock];
    double v = INFINITY, ix = 1e300;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        amin_merge(v, ix, D[i], (double)i);
    sv[threadIdx.x] = v;
    si[threadIdx.x] = ix;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double a = sv[threadIdx.x], b = si[threadIdx.x];
            amin_merge(a, b, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        red[2 * blockIdx.x] = sv[0];
        red[2 * blockIdx.x + 1] = si[0];
    }
}

// (value, index) pairs are totally ordered, so the result does not depend on
// the order of the merges.
__global__ __launch_bounds__(kBlock) void k_argmin_fin(const double *__restrict__ red, int nb,
                                                       int64_t *__restrict__ idx)
{
    __shared__ double sv[kBlock], si[kBlock];
    double v = INFINITY, ix = 1e300;
    for (int b = threadIdx.x; b < nb; b += kBlock) amin_merge(v, ix, red[2 * b], red[2 * b + 1]);
    sv[threadIdx.x] = v;
    si[threadIdx.x] = ix;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            double a = sv[threadIdx.x], b = si[threadIdx.x];
            amin_merge(a, b, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *idx = (int64_t)si[0];
}

} // namespace

// compile-time vector counts: m rounded up to the next of 1, 2, 4, 6, 8, 12,
// 16, 24, 32 (at most 32 with W dots) and 64
#define DK_BUCKET32(m, CALL)                                                     \
    do {                                                                         \
        if ((m) <= 1) { constexpr int MM = 1; CALL; }                            \
        else if ((m) <= 2) { constexpr int MM = 2; CALL; }                       \
        else if ((m) <= 4) { constexpr int MM = 4; CALL; }                       \
        else if ((m) <= 6) { constexpr int MM = 6; CALL; }                       \
        else if ((m) <= 8) { constexpr int MM = 8; CALL; }                       \
        else if ((m) <= 12) { constexpr int MM = 12; CALL; }                     \
        else if ((m) <= 16) { constexpr int MM = 16; CALL; }                     \
        else if ((m) <= 24) { constexpr int MM = 24; CALL; }                     \
        else { constexpr int MM = 32; CALL; }                                    \
    } while (0)
#define DK_BUCKET(m, CALL)                                                       \
    do {                                                                         \
        if ((m) <= 1) { constexpr int MM = 1; CALL; }                            \
        else if ((m) <= 2) { constexpr int MM = 2; CALL; }                       \
        else if ((m) <= 4) { constexpr int MM = 4; CALL; }                       \
        else if ((m) <= 6) { constexpr int MM = 6; CALL; }                       \
        else if ((m) <= 8) { constexpr int MM = 8; CALL; }                       \
        else if ((m) <= 12) { constexpr int MM = 12; CALL; }                     \
        else if ((m) <= 16) { constexpr int MM = 16; CALL; }                     \
        else if ((m) <= 24) { constexpr int MM = 24; CALL; }                     \
        else if ((m) <= 32) { constexpr int MM = 32; CALL; }                     \
        else { constexpr int MM = 64; CALL; }                                    \
    } while (0)

// vector slots j >= m repeat the last vector (fill when m = 0): their
// coefficients are 0 and their sums are not stored
static Ptrs make_ptrs(int m, void *const *X, const void *fill = nullptr)
{
    Ptrs p;
    for (int j = 0; j < DK_MAX_VEC; ++j) p.p[j] = j < m ? X[j] : (m > 0 ? X[m - 1] : fill);
    return p;
}
static Coef make_coef(int m, const double *g)
{
    Coef c;
    for (int j = 0; j < DK_MAX_VEC; ++j) c.g[j] = j < m ? g[j] : 0.0;
    return c;
}

extern "C" size_t dk_red_scratch_doubles(int nvals)
{
    return (size_t)kMaxBlocks * (size_t)(nvals < 2 ? 2 : nvals);
}

extern "C" int dk_multidot(int m, void *const *X, const void *y, int64_t n, int prec,
                           double *red, double *out, hipStream_t s)
{
    if (m < 0 || m > DK_MAX_VEC) return (int)hipErrorInvalidValue;
    const int nb = grid_for(n);
    const Ptrs P = make_ptrs(m, X, y);
    if (prec == 0)
        DK_BUCKET(m, hipLaunchKernelGGL((k_multidot<double, MM>), dim3(nb), dim3(kBlock), 0, s, P, m,
                                        (const double *)y, n, red));
    else
        DK_BUCKET(m, hipLaunchKernelGGL((k_multidot<float, MM>), dim3(nb), dim3(kBlock), 0, s, P, m,
                                        (const float *)y, n, red));
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, m + 1, out, s);
}

extern "C" int dk_ortho(int m, void *const *X, const double *g, void *y, int64_t n, int prec,
                        double *red, double *out, hipStream_t s)
{
    if (m < 0 || m > DK_MAX_VEC) return (int)hipErrorInvalidValue;
    const int nb = grid_for(n);
    const Ptrs P = make_ptrs(m, X, y);
    const Coef G = make_coef(m, g);
    if (prec == 0)
        DK_BUCKET(m, hipLaunchKernelGGL((k_ortho<double, MM>), dim3(nb), dim3(kBlock), 0, s, P, m, G,
                                        (double *)y, n, red));
    else
        DK_BUCKET(m, hipLaunchKernelGGL((k_ortho<float, MM>), dim3(nb), dim3(kBlock), 0, s, P, m, G,
                                        (float *)y, n, red));
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, m + 1, out, s);
}

extern "C" int dk_residual(int m, void *const *V, void *const *W, const double *a, double theta,
                           double shift, const double *D, void *t, void *x_out, int wdots, int64_t n,
                           int prec, double *red, double *out, hipStream_t s)
{
    if (m < 1 || m > DK_MAX_VEC || (wdots && m > DK_RESIDUAL_WDOTS_MAX))
        return (int)hipErrorInvalidValue;
    const int nb = grid_for(n);
    const Ptrs PV = make_ptrs(m, V), PW = make_ptrs(m, W);
    const Coef A = make_coef(m, a);
#define DK_RES_LAUNCH(T, WB)                                                                  \
    hipLaunchKernelGGL((k_residual<T, MM, WB>), dim3(nb), dim3(kBlock), 0, s, PV, PW, m, A, theta, \
                       shift, D, (T *)t, (T *)x_out, n, red)
    if (wdots) {
        if (prec == 0) DK_BUCKET32(m, DK_RES_LAUNCH(double, true));
        else DK_BUCKET32(m, DK_RES_LAUNCH(float, true));
    } else {
        if (prec == 0) DK_BUCKET(m, DK_RES_LAUNCH(double, false));
        else DK_BUCKET(m, DK_RES_LAUNCH(float, false));
    }
#undef DK_RES_LAUNCH
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, wdots ? 2 + 2 * m : m + 2, out, s);
}

extern "C" int dk_lincomb(int m, void *const *X, const double *g, void *out, int64_t n, int prec,
                          int prec_out, hipStream_t s)
{
    if (m < 1 || m > DK_MAX_VEC) return (int)hipErrorInvalidValue;
    const int nb = grid_for(n);
    const Ptrs P = make_ptrs(m, X);
    const Coef G = make_coef(m, g);
    if (prec == 0 && prec_out == 0)
        DK_BUCKET(m, hipLaunchKernelGGL((k_lincomb<double, double, MM>), dim3(nb), dim3(kBlock), 0, s,
                                        P, m, G, (double *)out, n));
    else if (prec == 0)
        DK_BUCKET(m, hipLaunchKernelGGL((k_lincomb<double, float, MM>), dim3(nb), dim3(kBlock), 0, s,
                                        P, m, G, (float *)out, n));
    else if (prec_out == 0)
        DK_BUCKET(m, hipLaunchKernelGGL((k_lincomb<float, double, MM>), dim3(nb), dim3(kBlock), 0, s,
                                        P, m, G, (double *)out, n));
    else
        DK_BUCKET(m, hipLaunchKernelGGL((k_lincomb<float, float, MM>), dim3(nb), dim3(kBlock), 0, s,
                                        P, m, G, (float *)out, n));
    return (int)hipGetLastError();
}

extern "C" int dk_convert(const void *in, int pin, void *out, int pout, int64_t n, double scale,
                          hipStream_t s)
{
    const int nb = grid_for(n);
    if (pin == 0 && pout == 0)
        hipLaunchKernelGGL((k_convert<double, double>), dim3(nb), dim3(kBlock), 0, s,
                           (const double *)in, (double *)out, n, scale);
    else if (pin == 0 && pout == 1)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(kBlock), 0, s,
                           (const double *)in, (float *)out, n, scale);
    else if (pin == 1 && pout == 0)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(kBlock), 0, s,
                           (const float *)in, (double *)out, n, scale);
    else
        hipLaunchKernelGGL((k_convert<float, float>), dim3(nb), dim3(kBlock), 0, s,
                           (const float *)in, (float *)out, n, scale);
    return (int)hipGetLastError();
}

extern "C" int dk_cg_grad(const void *x, const void *Ax, double rho, const double *pinv, void *g,
                          int64_t n, int prec, double *red, double *out, hipStream_t s)
{
    const int nb = grid_for(n);
    if (prec == 0)
        hipLaunchKernelGGL(k_cg_grad<double>, dim3(nb), dim3(kBlock), 0, s, (const double *)x,
                           (const double *)Ax, rho, pinv, (double *)g, n, red);
    else
        hipLaunchKernelGGL(k_cg_grad<float>, dim3(nb), dim3(kBlock), 0, s, (const float *)x,
                           (const float *)Ax, rho, pinv, (float *)g, n, red);
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, 2, out, s);
}

extern "C" int dk_cg_dir(void *p, const void *g, double q, int first, int64_t n, int prec,
                         hipStream_t s)
{
    const int nb = grid_for(n);
    if (prec == 0)
        hipLaunchKernelGGL(k_cg_dir<double>, dim3(nb), dim3(kBlock), 0, s, (double *)p,
                           (const double *)g, q, first, n);
    else
        hipLaunchKernelGGL(k_cg_dir<float>, dim3(nb), dim3(kBlock), 0, s, (float *)p,
                           (const float *)g, q, first, n);
    return (int)hipGetLastError();
}

extern "C" int dk_cg_gram(const void *x, const void *p, const void *Ax, const void *Ap, int64_t n,
                          int prec, double *red, double *out, hipStream_t s)
{
    const int nb = grid_for(n);
    if (prec == 0)
        hipLaunchKernelGGL(k_cg_gram<double>, dim3(nb), dim3(kBlock), 0, s, (const double *)x,
                           (const double *)p, (const double *)Ax, (const double *)Ap, n, red);
    else
        hipLaunchKernelGGL(k_cg_gram<float>, dim3(nb), dim3(kBlock), 0, s, (const float *)x,
                           (const float *)p, (const float *)Ax, (const float *)Ap, n, red);
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, 7, out, s);
}

extern "C" int dk_cg_step(void *x, const void *p, void *Ax, const void *Ap, double z0, double z1,
                          double rho, const double *pinv, void *g, int64_t n, int prec,
                          double *red, double *out, hipStream_t s)
{
    const int nb = grid_for(n);
    if (prec == 0)
        hipLaunchKernelGGL(k_cg_step<double>, dim3(nb), dim3(kBlock), 0, s, (double *)x,
                           (const double *)p, (double *)Ax, (const double *)Ap, z0, z1, rho, pinv,
                           (double *)g, n, red);
    else
        hipLaunchKernelGGL(k_cg_step<float>, dim3(nb), dim3(kBlock), 0, s, (float *)x,
                           (const float *)p, (float *)Ax, (const float *)Ap, z0, z1, rho, pinv,
                           (float *)g, n, red);
    int e = (int)hipGetLastError();
    return e ? e : finalize(red, nb, 2, out, s);
}

extern "C" int dk_precond_inv(const double *D, double *pinv, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_precond_inv, dim3(grid_for(n)), dim3(kBlock), 0, s, D, pinv, n);
    return (int)hipGetLastError();
}

extern "C" int dk_argmin(const double *D, int64_t n, double *red, int64_t *idx, hipStream_t s)
{
    const int nb = grid_for(n);
    hipLaunchKernelGGL(k_argmin_part, dim3(nb), dim3(kBlock), 0, s, D, n, red);
    hipLaunchKernelGGL(k_argmin_fin, dim3(1), dim3(kBlock), 0, s, red, nb, idx);
    return (int)hipGetLastError();
}
