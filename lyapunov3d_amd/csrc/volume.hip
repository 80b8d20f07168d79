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
#include <mutex>

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

/* z slices [z_begin, z_end) of the grid; out points at slice z_begin. The
 * FP64 re-evaluation list (list_cap entries) and its counter are given:
 * nothing here allocates. The counter is cleared by a kernel first. */
static int launch_volume_slab_in(lyap_ctx *ctx, float *out, const EvalDesc &e,
                                 const float lo[3], const float hi[3], unsigned nx,
                                 unsigned ny, unsigned nz, unsigned z_begin, unsigned z_end,
                                 const float *xyz, size_t npoints, const LyapVolumeMode &mode,
                                 uint32_t *list, uint32_t list_cap, uint32_t *count)
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
    a.count = count;
    int rc = launch_zero_u32(ctx, count, 1);
    if (rc) return rc;
    if (reeval) {
        a.list = list;
        a.list_cap = list_cap;
    }
    switch (e.L) {
        LYAP_L_CASES(VolumeLaunch, ctx, a, mode.precision, halves, reeval)
    default:
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence period %d outside 1..%d",
                    e.L, LYAP_MAX_SEQ);
    }
}

int launch_volume_slab(lyap_ctx *ctx, float *out, const EvalDesc &e,
                       const float lo[3], const float hi[3], unsigned nx,
                       unsigned ny, unsigned nz, unsigned z_begin, unsigned z_end,
                       const float *xyz, size_t npoints, const LyapVolumeMode &mode)
{
    const uint64_t n = xyz ? (uint64_t)npoints : (uint64_t)nx * ny * (z_end - z_begin);
    const bool reeval = mode.precision == LYAP_PREC_PRODUCT &&
                        (mode.fp64_margin > 0.0f || mode.converge_tau > 0.0f);
    if (reeval && n > 0 && n <= 0xFFFFFFFFull) {
        int rc = ensure_reeval_list(ctx, n);
        if (rc) return rc;
    }
    return launch_volume_slab_in(ctx, out, e, lo, hi, nx, ny, nz, z_begin, z_end, xyz, npoints,
                                 mode, ctx->reeval_list, (uint32_t)ctx->reeval_cap, ctx->counters);
}

/* ---- the allocation-free volume entry ------------------------------------- */

/* Scratch of lyap_calc_volume_noalloc: word 0 is the FP64 re-evaluation
 * counter (alone on its 256-byte line), the list starts at word
 * kNoAllocListOffset. One copy per device, held until the process ends. In
 * host code this name is only a stand-in: kernels receive the device
 * address that hipGetSymbolAddress returned for the device
 * (noalloc_scratch_setup). */
constexpr uint32_t kNoAllocListOffset = 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocListOffset + kNoAllocListCap];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch[kMaxScratchDevices];       /* device address, per device */
int g_scratch_status[kMaxScratchDevices];      /* 0: not looked up, 1: ok, < 0: failed */
}

int noalloc_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return LYAP_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_status[device] != 0) return g_scratch_status[device] > 0 ? LYAP_OK : LYAP_ERR_DEVICE;
    /* the caller has made `device` current */
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        g_scratch_status[device] = -1;
        return LYAP_ERR_DEVICE;
    }
    g_scratch[device] = static_cast<uint32_t *>(p);
    g_scratch_status[device] = 1;
    return LYAP_OK;
}

int launch_volume_noalloc(lyap_ctx *ctx, float *out, const EvalDesc &e, const float lo[3],
                          const float hi[3], unsigned nx, unsigned ny, unsigned nz,
                          const LyapVolumeMode &mode)
{
    uint32_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (ctx->device >= 0 && ctx->device < kMaxScratchDevices && g_scratch_status[ctx->device] > 0)
            base = g_scratch[ctx->device];
    }
    if (!base)
        return fail(ctx, LYAP_ERR_DEVICE, "the scratch array's device address is unknown "
                    "(hipGetSymbolAddress failed when the context was created)");
    const uint64_t plane = (uint64_t)nx * ny;
    if (plane > kNoAllocListCap)
        return fail(ctx, LYAP_ERR_UNSUPPORTED,
                    "a z slice of %llu voxels exceeds the %u-entry scratch list",
                    (unsigned long long)plane, kNoAllocListCap);
    /* slabs of whole z slices that the list can hold entirely, so every voxel
     * the rule selects is re-evaluated, as in lyap_calc_volume */
    const uint32_t per = (uint32_t)(kNoAllocListCap / plane);
    for (uint32_t z = 0; z < nz; z += per) {
        const uint32_t z1 = nz - z < per ? nz : z + per;
        int rc = launch_volume_slab_in(ctx, out + plane * z, e, lo, hi, nx, ny, nz, z, z1, nullptr, 0,
                                       mode, base + kNoAllocListOffset, kNoAllocListCap, base);
        if (rc) return rc;
    }
    return LYAP_OK;
}

} /* namespace lyap */
