/*
 * volume.hip - lambda over a voxel grid or a point list.
 *
 * Replaces kernel_calc_volume (kernel.cu:518-532) + pointcalc
 * (kernel.cu:488-498). One lane per voxel; consecutive lanes take
 * consecutive x, so the FP32 lambda writes are coalesced (index
 * x + (y + z*ny)*nx as kernel.cu:525). The coordinate of index i on an axis is
 * lo + ((hi - lo) * i) / n in FP32, which for lo = 0, hi = 4 is the
 * reference's 4*i/n (kernel.cu:527-529).
 *
 * With a re-evaluation rule the FP32 pass appends the voxels it is unsure of
 * to a compact list (one atomic per wave) and a second kernel walks that list
 * with the FP64 form, so the FP64 work runs at full lane occupancy.
 */
#include "internal.hpp"

namespace lyap {
namespace {

struct VolumeArgs {
    float *out;
    const float *xyz;          /* point list, or null for the grid */
    uint64_t n;                /* voxels or points */
    uint32_t nx, ny, nz, z0;   /* z0: first z slice of this launch */
    float lo[3], span[3];
    uint32_t *list;            /* FP64 re-evaluation list */
    uint32_t *count;
    float margin, tau, band;
    uint32_t list_cap;
    EvalDesc e;
};

__device__ __forceinline__ void voxel_coords(const VolumeArgs &a, uint64_t idx,
                                             float &x, float &y, float &z)
{
    if (a.xyz) {
        x = a.xyz[3 * idx];
        y = a.xyz[3 * idx + 1];
        z = a.xyz[3 * idx + 2];
        return;
    }
    const uint64_t plane = (uint64_t)a.nx * a.ny;
    const uint32_t lz = (uint32_t)(idx / plane);
    const uint32_t rem = (uint32_t)(idx - (uint64_t)lz * plane);
    const uint32_t iz = a.z0 + lz;
    const uint32_t iy = rem / a.nx;
    const uint32_t ix = rem - iy * a.nx;
    x = a.lo[0] + (a.span[0] * (float)ix) / (float)a.nx;
    y = a.lo[1] + (a.span[1] * (float)iy) / (float)a.ny;
    z = a.lo[2] + (a.span[2] * (float)iz) / (float)a.nz;
}

/* Appends idx to the list for every lane with `want` set: one atomic per
 * wave, ranks from the ballot. */
__device__ __forceinline__ void append_list(const VolumeArgs &a, bool want,
                                            uint32_t idx)
{
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(a.count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (want) {
        const uint32_t slot = base + (uint32_t)__popcll(mask & __lanemask_lt());
        if (slot < a.list_cap) a.list[slot] = idx;
    }
}

template <int L>
__global__ __launch_bounds__(256) void volume_plain(VolumeArgs a)
{
    const uint64_t idx = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.n) return;
    float x, y, z;
    voxel_coords(a, idx, x, y, z);
    a.out[idx] = lambda_plain<L>(x, y, z, a.e);
}

template <int L, bool HALVES>
__global__ __launch_bounds__(256) void volume_product(VolumeArgs a)
{
    const uint64_t idx = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.n) return;
    float x, y, z;
    voxel_coords(a, idx, x, y, z);
    const ProductResult r = lambda_product<L, HALVES>(x, y, z, a.e);
    a.out[idx] = r.lam;
    if (a.list) {
        bool want = !(__builtin_fabsf(r.lam) >= a.margin);   /* NaN included */
        if (HALVES) {
            const float lo_half = r.half1 < r.half2 ? r.half1 : r.half2;
            want = want || (lo_half < a.band && __builtin_fabsf(r.half1 - r.half2) > a.tau);
        }
        append_list(a, want, (uint32_t)idx);
    }
}

/* FP64 pass: over the list when there is one, else over every voxel. */
template <int L>
__global__ __launch_bounds__(256) void volume_fp64(VolumeArgs a)
{
    const uint64_t total = a.list ? (uint64_t)min(*a.count, a.list_cap) : a.n;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * 256u) {
        const uint64_t idx = a.list ? a.list[i] : i;
        float x, y, z;
        voxel_coords(a, idx, x, y, z);
        a.out[idx] = (float)lambda_fp64<L>(x, y, z, a.e);
    }
}

/* Product form with the two window halves, per point. */
template <int L>
__global__ __launch_bounds__(256) void points_halves(VolumeArgs a, float *h1, float *h2)
{
    const uint64_t idx = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (idx >= a.n) return;
    float x, y, z;
    voxel_coords(a, idx, x, y, z);
    const ProductResult r = lambda_product<L, true>(x, y, z, a.e);
    a.out[idx] = r.lam;
    h1[idx] = r.half1;
    h2[idx] = r.half2;
}

template <int L>
struct HalvesLaunch {
    static int run(lyap_ctx *ctx, VolumeArgs &a, float *h1, float *h2)
    {
        const uint64_t blocks = (a.n + 255) / 256;
        hipLaunchKernelGGL(points_halves<L>, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a, h1, h2);
        LYAP_HIP(ctx, hipGetLastError());
        return LYAP_OK;
    }
};

__global__ void zero_u32(uint32_t *p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <int L>
struct VolumeLaunch {
    static int run(lyap_ctx *ctx, VolumeArgs &a, int precision, bool halves,
                   bool fp64_pass)
    {
        const uint64_t blocks = (a.n + 255) / 256;
        if (precision == LYAP_PREC_PLAIN) {
            hipLaunchKernelGGL(volume_plain<L>, dim3((uint32_t)blocks), dim3(256), 0,
                               ctx->stream, a);
        } else if (precision == LYAP_PREC_PRODUCT) {
            if (halves)
                hipLaunchKernelGGL((volume_product<L, true>), dim3((uint32_t)blocks),
                                   dim3(256), 0, ctx->stream, a);
            else
                hipLaunchKernelGGL((volume_product<L, false>), dim3((uint32_t)blocks),
                                   dim3(256), 0, ctx->stream, a);
        }
        LYAP_HIP(ctx, hipGetLastError());
        if (precision == LYAP_PREC_FP64 || fp64_pass) {
            /* grid-stride: enough waves to fill every CU several times */
            uint64_t g = (uint64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256) * 16;
            if (precision == LYAP_PREC_FP64 && blocks < g) g = blocks;
            if (g == 0) g = 1;
            hipLaunchKernelGGL(volume_fp64<L>, dim3((uint32_t)g), dim3(256), 0,
                               ctx->stream, a);
            LYAP_HIP(ctx, hipGetLastError());
        }
        return LYAP_OK;
    }
};

#define LYAP_L_CASES(F, ...)                                                   \
    case 1: return F<1>::run(__VA_ARGS__);   case 2: return F<2>::run(__VA_ARGS__);   \
    case 3: return F<3>::run(__VA_ARGS__);   case 4: return F<4>::run(__VA_ARGS__);   \
    case 5: return F<5>::run(__VA_ARGS__);   case 6: return F<6>::run(__VA_ARGS__);   \
    case 7: return F<7>::run(__VA_ARGS__);   case 8: return F<8>::run(__VA_ARGS__);   \
    case 9: return F<9>::run(__VA_ARGS__);   case 10: return F<10>::run(__VA_ARGS__); \
    case 11: return F<11>::run(__VA_ARGS__); case 12: return F<12>::run(__VA_ARGS__); \
    case 13: return F<13>::run(__VA_ARGS__); case 14: return F<14>::run(__VA_ARGS__); \
    case 15: return F<15>::run(__VA_ARGS__); case 16: return F<16>::run(__VA_ARGS__); \
    case 17: return F<17>::run(__VA_ARGS__); case 18: return F<18>::run(__VA_ARGS__); \
    case 19: return F<19>::run(__VA_ARGS__); case 20: return F<20>::run(__VA_ARGS__); \
    case 21: return F<21>::run(__VA_ARGS__); case 22: return F<22>::run(__VA_ARGS__); \
    case 23: return F<23>::run(__VA_ARGS__); case 24: return F<24>::run(__VA_ARGS__); \
    case 25: return F<25>::run(__VA_ARGS__); case 26: return F<26>::run(__VA_ARGS__); \
    case 27: return F<27>::run(__VA_ARGS__); case 28: return F<28>::run(__VA_ARGS__); \
    case 29: return F<29>::run(__VA_ARGS__); case 30: return F<30>::run(__VA_ARGS__); \
    case 31: return F<31>::run(__VA_ARGS__); case 32: return F<32>::run(__VA_ARGS__);

} /* namespace */

int launch_zero_u32(lyap_ctx *ctx, uint32_t *p, uint32_t n)
{
    hipLaunchKernelGGL(zero_u32, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, p, n);
    LYAP_HIP(ctx, hipGetLastError());
    return LYAP_OK;
}

int launch_volume(lyap_ctx *ctx, float *out, const EvalDesc &e,
                  const float lo[3], const float hi[3], unsigned nx,
                  unsigned ny, unsigned nz, const float *xyz, size_t npoints,
                  const LyapVolumeMode &mode)
{
    return launch_volume_slab(ctx, out, e, lo, hi, nx, ny, nz, 0, nz, xyz, npoints, mode);
}

int launch_points_halves(lyap_ctx *ctx, float *lam, float *h1, float *h2, const EvalDesc &e,
                         const float *xyz, size_t n)
{
    if (n == 0) return LYAP_OK;
    if (n > 0xFFFFFFFFull) return fail(ctx, LYAP_ERR_UNSUPPORTED, "more than 2^32-1 points");
    VolumeArgs a{};
    a.out = lam;
    a.xyz = xyz;
    a.n = n;
    a.e = e;
    switch (e.L) {
        LYAP_L_CASES(HalvesLaunch, ctx, a, h1, h2)
    default:
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence period %d outside 1..%d", e.L, LYAP_MAX_SEQ);
    }
}

/* z slices [z_begin, z_end) of the grid; out points at slice z_begin. */
int launch_volume_slab(lyap_ctx *ctx, float *out, const EvalDesc &e,
                       const float lo[3], const float hi[3], unsigned nx,
                       unsigned ny, unsigned nz, unsigned z_begin, unsigned z_end,
                       const float *xyz, size_t npoints, const LyapVolumeMode &mode)
{
    VolumeArgs a{};
    a.out = out;
    a.xyz = xyz;
    if (xyz) {
        a.n = npoints;
    } else {
        a.n = (uint64_t)nx * ny * (z_end - z_begin);
        a.z0 = z_begin;
        a.nx = nx;
        a.ny = ny;
        a.nz = nz;
        for (int k = 0; k < 3; ++k) {
            a.lo[k] = lo[k];
            a.span[k] = hi[k] - lo[k];
        }
    }
    if (a.n == 0) return LYAP_OK;
    if (a.n > 0xFFFFFFFFull)
        return fail(ctx, LYAP_ERR_UNSUPPORTED, "more than 2^32-1 voxels in one call");
    a.e = e;
    a.margin = mode.fp64_margin;
    a.tau = mode.converge_tau;
    a.band = mode.converge_band;
    const bool reeval = mode.precision == LYAP_PREC_PRODUCT &&
                        (mode.fp64_margin > 0.0f || mode.converge_tau > 0.0f);
    const bool halves = reeval && mode.converge_tau > 0.0f;
    a.count = ctx->counters;
    int rc = launch_zero_u32(ctx, ctx->counters, 1);
    if (rc) return rc;
    if (reeval) {
        rc = ensure_reeval_list(ctx, a.n);
        if (rc) return rc;
        a.list = ctx->reeval_list;
        a.list_cap = (uint32_t)ctx->reeval_cap;
    }
    switch (e.L) {
        LYAP_L_CASES(VolumeLaunch, ctx, a, mode.precision, halves, reeval)
    default:
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence period %d outside 1..%d",
                    e.L, LYAP_MAX_SEQ);
    }
}

} /* namespace lyap */
