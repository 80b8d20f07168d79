/*
 * render.hip - primary-ray raymarch, refinement, normal and shading.
 *
 * Replaces kernel_calc_render (kernel.cu:500-516) with raymarch
 * (kernel.cu:156-486) and shade (kernel.cu:25-106) fused in.
 *
 * Schedule. The reference runs one thread per pixel, so a wave lasts as long
 * as its slowest ray. Here every lane runs a small state machine
 * (march -> refine -> six normal samples -> shade) and each trip round the
 * kernel's main loop performs exactly one lambda evaluation per busy lane.
 * A lane whose ray finishes (hit, miss or exit) takes the next ray of its
 * wave's tile at once; the wave claims a new tile of rays from a device
 * counter (one atomic per claim) when its tile runs out. The grid is sized
 * to the machine (persistent), every wave exits when the queue is empty and
 * its lanes are idle.
 *
 * Batched views: a tile belongs to one view; the wave keeps all its lanes on
 * one view so the sequence and iteration counts stay wave-uniform. When the
 * next tile is from another view the wave lets its busy lanes finish first.
 *
 * Ray order inside a view is 8x8 pixel blocks, so the 64 rays a wave starts
 * together are neighbours with similar march lengths.
 */
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace lyap {

/* One view of a render launch, in device memory. */
struct ViewDesc {
    EvalDesc e;
    LyapCam cam;
    LyapParams prm;
    uint32_t w, h;
    uint32_t blocks_x;         /* 8x8 pixel blocks per row */
    uint32_t n_rays;           /* 64 * blocks, padded past the image edge */
    uint32_t tile_begin;       /* first tile of this view in the launch */
    uint32_t n_tiles;
    LyapRGBA *rgba;
    LyapPoint *points;
    float *depth;
    float *lambda;
    uint32_t *evals;
    const LyapLight *lights;
    uint32_t num_lights;
    int32_t refine;            /* LYAP_REFINE_* */
    int32_t bisect_steps;
    float normal_h;
    float margin;              /* FP64 re-evaluation margin during bisection */
    uint32_t max_samples;
    uint32_t ray_begin;        /* first ray of this launch (a view may be split) */
    uint32_t cont_base;        /* this view's region of the continuation buffers */
    uint32_t cont_cap;
    uint32_t refine_base;      /* this view's region of the refinement records */
    uint32_t refine_cap;       /* records per sub-region (kRefineSubs of them) */
    uint32_t ray_base;         /* march passes: first ray of this view in ray_start's grid */
    /* Items of a launch are views or volume blocks (kItemVolume). A volume
     * block is z slices [vz0, vz0 + n_rays / (vnx * vny)) of an vnx * vny * vnz
     * grid over [vlo, vlo + vspan), lambda written to vol_out (slice vz0) in
     * the x-fastest layout of kernel.cu:525; voxels the FP32 pass is unsure of
     * go to the block's region of the FP64 list. */
    uint32_t kind;
    uint32_t tile_len;         /* rays or voxels per level-0 work-queue tile */
    float *vol_out;
    uint32_t vnx, vny, vnz, vz0;
    float vlo[3], vspan[3];
    float vmargin, vtau, vband;
    uint32_t vlist_base, vlist_cap;
};

constexpr uint32_t kItemView = 0;
constexpr uint32_t kItemVolume = 1;

/* March state of a ray handed to the next launch (96 bytes; 16-byte
 * aligned, so a record moves in six 16-byte loads or stores). */
struct alignas(16) ContEntry {
    uint32_t pix, evals, index, flags;
    float Vx, Vy, Vz, Px, Py, Pz;
    float t, t1, dt, Fdt, Ndt, last_step;
    float a, c, l;
    float tb, dtb;
    uint32_t m;
    uint32_t pad[2];
};
static_assert(sizeof(ContEntry) == 96, "continuation record size");

/* A ray whose march found its opaque sample, waiting for the refinement
 * launch: the bisection interval [lo, hi] with lambda at both ends, the
 * march step (the normal offset of the reference's method), the march sums
 * and the evaluations so far (48 bytes). */
struct alignas(16) RefineRec {
    uint32_t pix, evals;
    float Vx, Vy, Vz, lo, hi, l_hi, dt, a, c, l_lo;
};
static_assert(sizeof(RefineRec) == 48, "refinement record size");

/* Refinement records of a view are kept in kRefineSubs sub-regions, one per
 * block index modulo kRefineSubs, each with its own counter on a 128-byte
 * line: waves of one XCD share a counter, and device-scope atomics on a line
 * that waves of several XCDs update are slow. */
constexpr uint32_t kRefineSubs = 8;

/* The march passes list their rays the same way: a view's continuation
 * region is kContSubs sub-lists of cont_cap / kContSubs records, one counter
 * per sub-list on its own 128-byte line, taken by block index modulo
 * kContSubs. */
constexpr uint32_t kContSubs = 8;

struct RenderArgs {
    const ViewDesc *views;
    uint32_t n_views;
    uint32_t total_tiles;
    uint32_t tile_rays;
    uint32_t lanes_per_ray;    /* lanes marching one ray while rays remain */
    uint32_t slots;            /* samples each lane evaluates per round (1..4) */
    /* kTileHeads claim counters, kHeadStride words apart (own cache lines) */
    uint32_t *tile_heads;
    unsigned long long *eval_counter;
    /* optional per-wave record of kWaveLogWords words (see the kernel's end) */
    unsigned long long *wave_log;
    /* Launch level: 0 takes rays from pixel tiles; a later level takes the
     * rays the previous launch handed over. A ray that has taken `budget`
     * march samples in this launch is handed to the next (0: never). */
    uint32_t level;
    uint32_t budget;
    uint32_t resident_lanes;
    uint32_t direct;           /* one lane per ray, no refill until the wave is idle */
    const ContEntry *cont_in;
    const uint32_t *in_counts;         /* per view */
    ContEntry *cont_out;
    uint32_t *out_counts;              /* per view */
    /* bisection rays park here when their march ends (0: refine in place);
     * counters per view and sub-region, 32 words apart */
    RefineRec *refine_out;
    uint32_t *refine_counts;
    /* volume blocks: FP64 re-evaluation list (voxel indices within a block)
     * and its counters, per item, kHeadStride words apart */
    uint32_t *vox_list;
    uint32_t *vox_counts;
};

namespace {

enum : int { PH_IDLE = 0, PH_MARCH = 1, PH_REFINE = 2, PH_NORMAL = 3 };

constexpr uint32_t kMaxWalk = 64;

struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(const LyapVec &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 divv(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float mag2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* Vec::normalize (vec3.hpp:141-154) */
__device__ __forceinline__ F3 normalized(F3 a)
{
    const float m2 = mag2(a);
    if ((double)m2 < 1e-12) return {0.f, 0.f, 0.f};
    if (m2 == 1.0f) return a;
    return divv(a, sqrtf(m2));
}
__device__ __forceinline__ float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

struct C4 { float x, y, z, w; };
__device__ __forceinline__ C4 col(const LyapColor &c) { return {c.x, c.y, c.z, c.w}; }
__device__ __forceinline__ C4 cadd(C4 a, C4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
__device__ __forceinline__ C4 cmul(C4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

/* shade (kernel.cu:25-106) */
__device__ C4 shade(F3 P, F3 N, float a, float c, const LyapCam &cam,
                    const LyapLight *lights, uint32_t n_lights)
{
    C4 color = {0.f, 0.f, 0.f, 0.f};
    if (__builtin_isnan(a)) return color;
    const F3 camC = f3(cam.C);
    for (uint32_t li = 0; li < n_lights; ++li) {
        const LyapLight &L = lights[li];
        const F3 camV = sub(camC, P);
        F3 lightV = sub(f3(L.C), P);
        const float d2 = mag2(lightV);
        lightV = normalized(lightV);
        float i = dot(lightV, N);
        const float j = -dot(lightV, f3(L.V));
        C4 phong;
        if (j > L.lightOuterCone) {
            i = clamp01(i);
            const C4 diffuse = cmul(col(L.diffuseColor), i * L.diffusePower);
            const F3 halfV = normalized(add(camV, lightV));
            float s = clamp01(dot(N, halfV));
            s = powf(s, L.specularHardness);
            const C4 specular = cmul(col(L.specularColor), s * L.specularPower);
            phong = cmul(cadd(specular, diffuse), L.lightRange / d2);
            if (j < L.lightInnerCone)
                phong = cmul(phong, (j - L.lightOuterCone) / (L.lightInnerCone - L.lightOuterCone));
            phong = cadd(phong, col(L.ambient));
        } else {
            phong = col(L.ambient);
        }
        if (c > 0.0f) {
            const float k = (float)(0.1125 / (double)logf(c));
            phong = cadd(phong, cmul(col(L.chaosColor), k));
        }
        color = cadd(color, phong);
    }
    return color;
}

/* to_rgba (color.hpp:169-175), saturating. */
__device__ __forceinline__ unsigned char to_u8(float x)
{
    const double v = 255.0 * (double)x;
    if (!(v > 0.0)) return 0;
    if (v >= 255.0) return 255;
    return (unsigned char)v;
}

/* Box entry and exit (kernel.cu:181-285), plane parameters formed in double
 * and stored as float like the reference's double literals make them. */
__device__ bool clip_box(F3 C, F3 V, float lo, float hi, float &t0_out, float &t1_out)
{
    const double dlo = lo, dhi = hi;
    float ts[6];
    const float Ca[3] = {C.x, C.y, C.z};
    const float Va[3] = {V.x, V.y, V.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ts[2 * a] = (Va[a] != 0.0f) ? (float)((dlo - (double)Ca[a]) / (double)Va[a]) : INFINITY;
        ts[2 * a + 1] = (Va[a] != 0.0f) ? (float)((dhi - (double)Ca[a]) / (double)Va[a]) : INFINITY;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (ts[k] == INFINITY) continue;
        const F3 P = add(C, mul(V, ts[k]));
        const float Pa[3] = {P.x, P.y, P.z};
        const int ax = k / 2;
        bool out = false;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (b == ax) continue;
            if ((double)Pa[b] < dlo || (double)Pa[b] > dhi) out = true;
        }
        if (out) ts[k] = NAN;
    }
    float t0 = 3.40282347e+38F, t1 = 0.0f;
    int i0 = -1, i1 = -1;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (!__builtin_isfinite(ts[k])) continue;
        if (i0 == -1 || ts[k] < t0) { t0 = ts[k]; i0 = k; }
        if (i1 == -1 || ts[k] > t1) { t1 = ts[k]; i1 = k; }
    }
    if (i0 == -1 && i1 == -1) return false;
    if (i1 == -1 || i0 == i1) { t1 = t0; t0 = 0.0f; }
    if (t0 < 0) t0 = 0.0f;
    t0_out = t0;
    t1_out = t1;
    return true;
}

/* ---- lambda in the kernel's window ------------------------------------- */

template <int W>
__device__ __noinline__ float eval_fp64_call(const EvalDesc &e, float x, float y, float z)
{
    return (float)lambda_fp64_w<W>(x, y, z, e);
}

template <int W, int PREC>
__device__ __forceinline__ float eval_lambda(const EvalDesc &e, float x, float y, float z)
{
    if (PREC == LYAP_PREC_FP64) return (float)lambda_fp64_w<W>(x, y, z, e);
    if (PREC == LYAP_PREC_PLAIN) return lambda_plain_w<W>(x, y, z, e);
    return lambda_product_w<W>(x, y, z, e);
}

/* Per-lane ray state. Every lane of a group holds the same copy; the group's
 * lanes differ only in which point they evaluate. */
struct Ray {
    int phase;
    uint32_t pix;          /* pixel index x + y*w */
    uint32_t evals;        /* lambda evaluations the method needed */
    uint32_t index;        /* march sample index of (t, P) */
    uint32_t k;            /* refine or normal step counter */
    uint32_t flags;        /* bit 0: near step, bit 1: walk sign, bit 2: hit to shade */
    F3 V, P;
    float t, t1, dt, Fdt, Ndt, last_step;
    float a, c, l;
    /* phase storage, one set live at a time: the reversing walk's
     * {Qdt, Qt0, Qt1, minQdt, QdV} in q[0..6], the bisection's {lo, hi, l_hi,
     * l_lo} in q[0..3], the normal's {h, gx, gy, gz} (central differences) in
     * q[3..6] */
    float q[7];
    __device__ float &lo() { return q[0]; }
    __device__ float lo() const { return q[0]; }
    __device__ float &hi() { return q[1]; }
    __device__ float hi() const { return q[1]; }
    __device__ float &l_hi() { return q[2]; }
    __device__ float l_hi() const { return q[2]; }
    __device__ float &l_lo() { return q[3]; }
    __device__ float l_lo() const { return q[3]; }
    __device__ float &Qdt() { return q[0]; }
    __device__ float Qdt() const { return q[0]; }
    __device__ float &Qt0() { return q[1]; }
    __device__ float Qt0() const { return q[1]; }
    __device__ float &Qt1() { return q[2]; }
    __device__ float Qt1() const { return q[2]; }
    __device__ float &minQdt() { return q[3]; }
    __device__ float minQdt() const { return q[3]; }
    __device__ F3 QdV() const { return {q[4], q[5], q[6]}; }
    __device__ void set_QdV(F3 d) { q[4] = d.x; q[5] = d.y; q[6] = d.z; }
    __device__ float &h() { return q[3]; }
    __device__ float h() const { return q[3]; }
    __device__ float &gx() { return q[4]; }
    __device__ float gx() const { return q[4]; }
    __device__ float &gy() { return q[5]; }
    __device__ float gy() const { return q[5]; }
    __device__ float &gz() { return q[6]; }
    __device__ float gz() const { return q[6]; }
    uint32_t launch_samples;   /* march samples taken in this launch */
    /* direct sampling (bisection method): t = tb + m * dtb, rebased when the
     * step length changes */
    float tb, dtb;
    uint32_t m;
};

__device__ __forceinline__ float shf(float v, int src) { return __shfl(v, src); }
__device__ __forceinline__ uint32_t shu(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src); }

/* Copy of lane `src`'s ray state. */
__device__ __forceinline__ Ray shuffle_ray(const Ray &r, int src)
{
    Ray o;
    o.phase = __shfl(r.phase, src);
    o.pix = shu(r.pix, src);
    o.evals = shu(r.evals, src);
    o.index = shu(r.index, src);
    o.k = shu(r.k, src);
    o.flags = shu(r.flags, src);
    o.V = {shf(r.V.x, src), shf(r.V.y, src), shf(r.V.z, src)};
    o.P = {shf(r.P.x, src), shf(r.P.y, src), shf(r.P.z, src)};
    o.t = shf(r.t, src); o.t1 = shf(r.t1, src); o.dt = shf(r.dt, src);
    o.Fdt = shf(r.Fdt, src); o.Ndt = shf(r.Ndt, src); o.last_step = shf(r.last_step, src);
    o.a = shf(r.a, src); o.c = shf(r.c, src); o.l = shf(r.l, src);
#pragma unroll
    for (int i = 0; i < 7; ++i) o.q[i] = shf(r.q[i], src);
    o.launch_samples = shu(r.launch_samples, src);
    o.tb = shf(r.tb, src); o.dtb = shf(r.dtb, src); o.m = shu(r.m, src);
    return o;
}

/* The view's output pointers, read together (one wait for all of them
 * rather than one per optional output). */
struct OutPtrs {
    LyapRGBA *rgba;
    LyapPoint *points;
    float *depth;
    float *lambda;
    uint32_t *evals;
    uint32_t present;      /* bit 0 points, 1 depth, 2 lambda, 3 evals */
};

__device__ __forceinline__ OutPtrs out_ptrs(const ViewDesc &v)
{
    OutPtrs o;
    o.rgba = v.rgba;
    o.points = v.points;
    o.depth = v.depth;
    o.lambda = v.lambda;
    o.evals = v.evals;
    o.present = (o.points ? 1u : 0u) | (o.depth ? 2u : 0u) | (o.lambda ? 4u : 0u) |
                (o.evals ? 8u : 0u);
    return o;
}

__device__ __forceinline__ void write_miss(const ViewDesc &v, const Ray &r)
{
    const OutPtrs o = out_ptrs(v);
    if (o.present & 1u) {
        LyapPoint p;
        p.P = {0.f, 0.f, 0.f};
        p.N = {0.f, 0.f, 0.f};
        p.a = NAN;
        p.c = 0.f;
        p.l = NAN;
        o.points[r.pix] = p;
    }
    if (o.present & 2u) o.depth[r.pix] = NAN;
    if (o.present & 4u) o.lambda[r.pix] = NAN;
    if (o.present & 8u) o.evals[r.pix] = r.evals;
    LyapRGBA z = {0, 0, 0, 0};
    o.rgba[r.pix] = z;
}

/* A NaN normal component (a normal sample's lambda was NaN) as the one
 * quiet NaN: the refinement launch and refinement in place form the
 * differences in orders that agree on every number but not on a NaN's sign
 * bit. */
__device__ __forceinline__ float canonical(float x) { return __builtin_isnan(x) ? NAN : x; }

__device__ void write_hit(const ViewDesc &v, const Ray &r, F3 N)
{
    const OutPtrs o = out_ptrs(v);
    if (o.present & 1u) {
        LyapPoint p;
        p.P = {r.P.x, r.P.y, r.P.z};
        p.N = {canonical(N.x), canonical(N.y), canonical(N.z)};
        p.a = r.a;
        p.c = r.c;
        p.l = r.l;
        o.points[r.pix] = p;
    }
    if (o.present & 2u) o.depth[r.pix] = r.t;
    if (o.present & 4u) o.lambda[r.pix] = r.l;
    if (o.present & 8u) o.evals[r.pix] = r.evals;
}

/* Hits waiting for their colour, per wave, in LDS. Shading a hit as its
 * ray ends runs the lights loop for one lane of the wave at a time; the
 * wave instead collects up to 64 hits and shades them with all its lanes at
 * once (the same shade and RGBA conversion, so the same bytes). */
struct ShadeList {
    float px[64], py[64], pz[64], nx[64], ny[64], nz[64], a[64], c[64];
    uint32_t pix[64];
};

/* Rays of a tile that meet the box, set up by all lanes of the wave at once
 * (one lane per ray) and then handed to groups as they free up; lanes that
 * set up a miss write it at once. Setting up each group's ray as it frees
 * up ran the box clip (six double divisions) for one group at a time. */
struct SetupList {
    float vx[64], vy[64], vz[64], t0[64], t1[64];
    uint32_t pix[64];
};

/* Shades entries [0, n) of the wave's list into view v's RGBA. */
__device__ void shade_list(const ViewDesc &v, const ShadeList &sl, uint32_t n)
{
    const uint32_t i = __lane_id();
    if (i < n) {
        const F3 P = {sl.px[i], sl.py[i], sl.pz[i]};
        const F3 N = {sl.nx[i], sl.ny[i], sl.nz[i]};
        const C4 cc = shade(P, N, sl.a[i], sl.c[i], v.cam, v.lights, v.num_lights);
        LyapRGBA px;
        px.r = to_u8(cc.x);
        px.g = to_u8(cc.y);
        px.b = to_u8(cc.z);
        px.a = to_u8(cc.w);
        v.rgba[sl.pix[i]] = px;
    }
}

/* Starts ray `id` of view v; returns false (after the leader wrote a miss,
 * or nothing for a padding ray past the image edge) when no evaluation is
 * needed. Every lane of the ray's group runs it and gets the same state. */
enum : int { RAY_PAD = 0, RAY_MISS = 1, RAY_HIT = 2 };

/* Pixel, direction and box interval of ray `id` of view v (pixel mapping
 * kernel.cu:500-516, direction :160-163, box clip :180-279). RAY_PAD for a
 * padding ray past the image edge. */
__device__ __forceinline__ int ray_setup(const ViewDesc &v, uint32_t id, uint32_t &pix, F3 &V,
                                         float &t0, float &t1)
{
    id += v.ray_begin;
    const uint32_t blk = id >> 6, in = id & 63u;
    const uint32_t px = (blk % v.blocks_x) * 8u + (in & 7u);
    const uint32_t py = (blk / v.blocks_x) * 8u + (in >> 3);
    if (px >= v.w || py >= v.h) return RAY_PAD;
    pix = px + py * v.w;
    V = add(add(f3(v.cam.S0), mul(f3(v.cam.SDX), (float)px)), mul(f3(v.cam.SDY), (float)py));
    V = normalized(V);
    V = divv(V, v.cam.M);
    return clip_box(f3(v.cam.C), V, v.prm.lMin, v.prm.lMax, t0, t1) ? RAY_HIT : RAY_MISS;
}
/* The march state of a ray that meets the box over [t0, t1]
 * (kernel.cu:281-314). Every lane of the ray's group runs it. */
__device__ __forceinline__ void start_hit(const ViewDesc &v, uint32_t pix, F3 V, float t0,
                                          float t1, Ray &r)
{
    r.pix = pix;
    r.evals = 0;
    r.V = V;
    const F3 C = f3(v.cam.C);
    r.t = t0;
    r.t1 = t1;
    r.P = add(C, mul(V, t0));
    r.a = 0.f;
    r.c = 0.f;
    r.l = 0.f;
    if (v.prm.stepMethod == 1) r.Fdt = (t1 - t0) / v.prm.depth;
    else r.Fdt = sqrtf(mag2(V)) / v.prm.depth;
    r.dt = r.Fdt;
    r.Ndt = r.dt / v.prm.nearMultiplier;
    r.flags = 0;
    r.last_step = r.dt;
    r.index = 0;
    r.k = 0;
    r.launch_samples = 0;
    r.tb = t0;
    r.dtb = r.dt;
    r.m = 0;
    r.phase = PH_MARCH;
}

/* Resumes a ray handed over by the previous launch. */
__device__ __forceinline__ void resume_ray(const ContEntry &c, Ray &r)
{
    r.pix = c.pix;
    r.evals = c.evals;
    r.index = c.index;
    r.flags = c.flags;
    r.V = {c.Vx, c.Vy, c.Vz};
    r.P = {c.Px, c.Py, c.Pz};
    r.t = c.t; r.t1 = c.t1; r.dt = c.dt; r.Fdt = c.Fdt; r.Ndt = c.Ndt;
    r.last_step = c.last_step;
    r.a = c.a; r.c = c.c; r.l = c.l;
    r.tb = c.tb; r.dtb = c.dtb; r.m = c.m;
    r.k = 0;
    r.launch_samples = 0;
    r.phase = PH_MARCH;
}

__device__ __forceinline__ void suspend_ray(const Ray &r, ContEntry &c)
{
    c.pix = r.pix; c.evals = r.evals; c.index = r.index; c.flags = r.flags;
    c.Vx = r.V.x; c.Vy = r.V.y; c.Vz = r.V.z;
    c.Px = r.P.x; c.Py = r.P.y; c.Pz = r.P.z;
    c.t = r.t; c.t1 = r.t1; c.dt = r.dt; c.Fdt = r.Fdt; c.Ndt = r.Ndt;
    c.last_step = r.last_step;
    c.a = r.a; c.c = r.c; c.l = r.l;
    c.tb = r.tb; c.dtb = r.dtb; c.m = r.m;
    c.pad[0] = c.pad[1] = 0;
}

/* Position of the next march sample after a step of `step` (the next t and,
 * in the reference walk method, P += V*step). The bisection method samples
 * t = tb + m*dtb directly and rebases (tb = t, dtb = step, m = 0) when the
 * step length changes. Returns false when the step does not change t (a
 * grazing ray with t1 - t0 near 0, or depth = inf), which would never leave
 * the box. */
__device__ __forceinline__ bool march_step(const ViewDesc &v, Ray &r, float step)
{
    if (v.refine == LYAP_REFINE_BISECT) {
        if (step != r.dtb) {
            r.tb = r.t;
            r.dtb = step;
            r.m = 0;
        }
        const float tn = r.tb + (float)(r.m + 1u) * r.dtb;
        if (!(tn > r.t)) return false;
        r.t = tn;
        ++r.m;
        r.P = add(f3(v.cam.C), mul(r.V, r.t));
        return true;
    }
    if (!(r.t + step > r.t)) return false;
    r.t += step;
    r.P = add(r.P, mul(r.V, step));
    return true;
}

__device__ __forceinline__ void enter_normal(const ViewDesc &v, Ray &r)
{
    r.phase = PH_NORMAL;
    r.k = 0;
    r.h() = (v.normal_h > 0.0f) ? v.normal_h : r.dt * v.prm.gradient;
}

/* walk loop test (kernel.cu:420) */
__device__ __forceinline__ bool walk_continues(const Ray &r)
{
    return r.k < kMaxWalk && r.t <= r.Qt1() && r.t >= r.Qt0() &&
           (r.Qdt() <= -r.minQdt() || r.Qdt() >= r.minQdt());
}

__device__ __forceinline__ void enter_refine(const ViewDesc &v, Ray &r, float l_prev)
{
    r.phase = PH_REFINE;
    r.k = 0;
    if (v.refine == LYAP_REFINE_BISECT) {
        r.lo() = r.t - r.last_step;
        r.hi() = r.t;
        r.l_hi() = r.l;
        r.l_lo() = l_prev;
        if (v.bisect_steps <= 0) enter_normal(v, r);
    } else {
        r.flags &= ~2u;
        r.Qdt() = r.dt * -0.5f;
        r.set_QdV(mul(r.V, r.Qdt()));
        r.Qt1() = r.t;
        r.Qt0() = r.t - r.dt;
        r.minQdt() = r.dt / v.prm.refine;
        if (!walk_continues(r)) enter_normal(v, r);
    }
}

/* Bisection levels a group of K lanes resolves per trip: the 2^m - 1 nodes
 * of the next m levels of the bisection tree. */
__device__ __forceinline__ uint32_t bisect_levels(uint32_t K, uint32_t left)
{
    uint32_t m = 31u - (uint32_t)__builtin_clz(K + 1u);   /* floor(log2(K+1)) */
    return m < left ? m : left;
}

/* Point `rank` of the group's K samples this trip (a group of lanes with
 * several slots each takes K = lanes x slots samples); false when it has
 * nothing to evaluate. Positions repeat the sequential method's float
 * operations exactly. */
__device__ __forceinline__ bool sample_point(const ViewDesc &v, const Ray &r,
                                             uint32_t rank, uint32_t K, F3 &s,
                                             float &ts)
{
    if (r.phase == PH_MARCH) {
        ts = r.t;
        if (rank == 0) { s = r.P; return true; }
        if (v.prm.jitter != 0.0f) return false;       /* next step needs this lambda */
        if (r.index + rank >= v.max_samples) return false;
        if (v.refine == LYAP_REFINE_BISECT) {
            /* rank steps ahead at the current step length */
            const float tp = r.tb + (float)(r.m + rank - 1u) * r.dt;
            const float t = r.tb + (float)(r.m + rank) * r.dt;
            if (r.dt != r.dtb || !(t > tp) || t > r.t1) return false;
            ts = t;
            s = add(f3(v.cam.C), mul(r.V, t));
            return true;
        }
        /* reference walk method: rank increments of dt */
        float t = r.t;
        F3 P = r.P;
        for (uint32_t j = 0; j < rank; ++j) {
            if (!(t + r.dt > t)) return false;
            t += r.dt;
            P = add(P, mul(r.V, r.dt));
        }
        if (t > r.t1) return false;
        ts = t;
        s = P;
        return true;
    }
    if (r.phase == PH_REFINE) {
        if (v.refine == LYAP_REFINE_BISECT) {
            const uint32_t m = bisect_levels(K, (uint32_t)v.bisect_steps - r.k);
            if (rank >= (1u << m) - 1u) return false;
            /* node rank of the next m levels: depth d, path bits below the top */
            const uint32_t n1 = rank + 1u;
            const uint32_t d = 31u - (uint32_t)__builtin_clz(n1);
            float lo = r.lo(), hi = r.hi();
            for (int i = (int)d - 1; i >= 0; --i) {
                const float mid = lo + (hi - lo) * 0.5f;
                if ((n1 >> i) & 1u) lo = mid;   /* transparent branch */
                else hi = mid;                  /* opaque branch */
            }
            const float mid = lo + (hi - lo) * 0.5f;
            s = add(f3(v.cam.C), mul(r.V, mid));
            return true;
        }
        if (rank != 0) return false;
        s = add(r.P, r.QdV());                            /* walk: t += Qdt; P += QdV */
        return true;
    }
    /* PH_NORMAL: samples k.. in the order x-, x+, y-, y+, z-, z+ */
    const uint32_t n = r.k + rank;
    if (n >= 6) return false;
    s = r.P;
    const float off = (n & 1u) ? r.h() : -r.h();
    if (n < 2) s.x = r.P.x + off;
    else if (n < 4) s.y = r.P.y + off;
    else s.z = r.P.z + off;
    return true;
}

/* Consumes one march sample (the lambda at (t, P)). Returns 0 to go on,
 * 1 when the ray left the box or hit (state updated), 2 when the step length
 * changed so later speculative samples no longer apply. */
__device__ __forceinline__ int march_consume(const ViewDesc &v, Ray &r, float l,
                                             bool leader)
{
    const LyapParams &prm = v.prm;
    ++r.evals;
    ++r.launch_samples;
    const float l_prev = r.l;
    r.l = l;
    bool toggled = false;
    if (r.index > 0) {
        /* c += l above the chaos threshold, else a += l above the opaque one
         * (kernel.cu:370-375); adding +0 leaves either sum unchanged */
        const bool to_c = l > prm.chaosThreshold;
        const bool to_a = !to_c && l > prm.opaqueThreshold;
        r.c += to_c ? l : 0.0f;
        r.a += to_a ? l : 0.0f;
        const bool near_ = (r.flags & 1u) != 0;
        if (l <= prm.nearThreshold && !near_) { r.flags |= 1u; r.dt = r.Ndt; toggled = r.Ndt != r.Fdt; }
        else if (l > prm.nearThreshold && near_) { r.flags &= ~1u; r.dt = r.Fdt; toggled = r.Ndt != r.Fdt; }
    }
    if (l > prm.opaqueThreshold) {
        float step = r.dt;
        if (prm.jitter != 0.0f) {
            float jit = l - truncf(l);
            if (jit < 0) jit = (float)(1.0 - (double)(jit * prm.jitter));
            else jit = (float)(1.0 + (double)(jit * prm.jitter));
            if (__builtin_isfinite(jit)) step = r.dt * jit;
        }
        if (!march_step(v, r, step)) {
            if (leader) write_miss(v, r);
            r.phase = PH_IDLE;
            return 1;
        }
        r.last_step = step;
        ++r.index;
        if (r.t > r.t1 || r.index >= v.max_samples) {
            if (leader) write_miss(v, r);
            r.phase = PH_IDLE;
            return 1;
        }
        /* under jitter only the first lane evaluated (its step needs lambda) */
        return (toggled || prm.jitter != 0.0f) ? 2 : 0;
    }
    enter_refine(v, r, l_prev);
    return 1;
}

/* Samples of one round, per wave, in LDS: sample j of the group whose
 * first lane is gbase sits at index gbase * slots + j. */
constexpr uint32_t kMaxSlots = 4;

/* Tile claiming. Every wave's first tile is its wave index; the tiles after
 * the first n_waves are dealt round-robin to kTileHeads counters (tile
 * n_waves + kTileHeads * c + h is the c-th claim on head h), so claims spread
 * over kTileHeads addresses instead of queueing on one. A block starts on head
 * blockIdx % kTileHeads and moves on when its head runs dry. */
constexpr uint32_t kTileHeads = 8;
constexpr uint32_t kHeadStride = 32;          /* 128 bytes */
constexpr uint32_t kWaveLogWords = 16;

struct RoundStash {
    float lam[64 * kMaxSlots];
    float t[64 * kMaxSlots];
    float x[64 * kMaxSlots], y[64 * kMaxSlots], z[64 * kMaxSlots];
    uint32_t ok[64 * kMaxSlots];
    /* a march sample's share of the chaos and attractor sums (0 when it adds
     * to neither), formed by the lane that evaluated it */
    float ca[64 * kMaxSlots], aa[64 * kMaxSlots];
};

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* The march part of combine: the group's samples j = 0..K-1 (read from the
 * stash at base + j, N entries) were taken at the positions the sequential
 * method would reach if none of them stopped it. Find the first that does,
 * account the transparent ones before it in order, then consume it. Sets
 * `finished` when this lane's ray ended (a miss, written by the leader). */
template <class Stash, uint32_t N>
__device__ __forceinline__ void combine_march(const ViewDesc &v, Ray &r, uint64_t evm, const Stash &st,
                                              uint32_t base, uint32_t K, int ph, bool leader,
                                              bool &finished)
{
    const bool marching = ph == PH_MARCH;
    const uint32_t e = evm ? (uint32_t)__ffsll((unsigned long long)evm) - 1u : K;
    /* transparent samples 0..e-1 added in order (kernel.cu:370-375). Each
     * sample's share was formed when it was evaluated; a sample that adds
     * to neither sum holds +0, and adding +0 leaves a sum unchanged (the
     * sums start at +0 and so are never -0), so the bits are those of the
     * sequential method. Four loads are issued ahead of their adds. */
    for (uint32_t j0 = 0; __ballot(marching && j0 < e) != 0; j0 += 4u) {
        float c4[4], a4[4];
#pragma unroll
        for (uint32_t q = 0; q < 4u; ++q) {
            const uint32_t i = min(base + j0 + q, N - 1u);
            c4[q] = st.ca[i];
            a4[q] = st.aa[i];
        }
#pragma unroll
        for (uint32_t q = 0; q < 4u; ++q) {
            const bool take = marching && j0 + q < e;
            r.c = take ? r.c + c4[q] : r.c;
            r.a = take ? r.a + a4[q] : r.a;
        }
    }
    if (marching && e > 0) {
        r.evals += e;
        r.launch_samples += e;
        r.index += e;
        r.l = st.lam[min(base + e - 1u, N - 1u)];
    }
    /* position and lambda of sample e */
    const uint32_t src = min(base + (e < K ? e : K - 1u), N - 1u);
    const float te = st.t[src];
    const F3 Pe = {st.x[src], st.y[src], st.z[src]};
    const float le = st.lam[src];
    const uint32_t ve = st.ok[src];
    if (marching) {
        if (e > 0) {
            r.last_step = r.dt;
            if (v.refine == LYAP_REFINE_BISECT) r.m += (e < K ? e : K - 1u);
        }
        if (e < K) {
            r.t = te;
            r.P = Pe;
            if (!ve) {
                /* the step to sample e left the box or could not advance */
                if (leader) write_miss(v, r);
                r.phase = PH_IDLE;
                finished = true;
            } else {
                march_consume(v, r, le, leader);
                if (r.phase == PH_IDLE) finished = true;
            }
        } else {
            /* all K transparent: step from sample K-1 to sample K */
            r.t = te;
            r.P = Pe;
            if (r.index >= v.max_samples || !march_step(v, r, r.dt) || r.t > r.t1) {
                if (leader) write_miss(v, r);
                r.phase = PH_IDLE;
                finished = true;
            }
        }
    }
}

/* Combines the group's K samples (read from the round's stash at
 * base + j) and advances the replicated ray state in the sequential
 * method's order. A group's lanes share one phase, and `evm` holds one bit
 * per sample j for it: marching, set when sample j stops the march (opaque,
 * invalid, near/far toggle, jitter); refining, set when sample j's lambda is
 * above the opaque threshold. All lanes run the same uniform loops. Returns
 * true when this lane's ray finished this round. */
__device__ bool combine(const ViewDesc &v, Ray &r, uint64_t evm, const RoundStash &st,
                        uint32_t base, uint32_t K, bool busy, bool leader)
{
    const int ph = busy ? r.phase : PH_IDLE;
    const float thr = v.prm.opaqueThreshold;
    const bool bisect = v.refine == LYAP_REFINE_BISECT;
    bool finished = false;

    if (__ballot(ph == PH_MARCH) != 0)
        combine_march<RoundStash, 64u * kMaxSlots>(v, r, evm, st, base, K, ph, leader, finished);

    /* bisection: follow the actual path through this round's tree levels,
     * reading each node's side of the threshold from `evm`; the lambda of the
     * last node on the opaque side is the one read from the stash */
    if (__ballot(ph == PH_REFINE && bisect) != 0) {
        const uint32_t m = (ph == PH_REFINE && bisect)
                               ? bisect_levels(K, (uint32_t)v.bisect_steps - r.k) : 0u;
        uint32_t node = 0, hi_node = 0xFFFFFFFFu;
        for (uint32_t i = 0; __ballot(i < m) != 0; ++i) {
            if (i < m) {
                const float mid = r.lo() + (r.hi() - r.lo()) * 0.5f;
                ++r.evals;
                ++r.k;
                if (((evm >> node) & 1ull) == 0) { r.hi() = mid; hi_node = node; node = 2u * node + 1u; }
                else { r.lo() = mid; node = 2u * node + 2u; }
            }
        }
        if (__ballot(hi_node != 0xFFFFFFFFu) != 0) {
            const float ln = st.lam[min(base + (hi_node & 63u), 64u * kMaxSlots - 1u)];
            if (hi_node != 0xFFFFFFFFu) r.l_hi() = ln;
        }
        if (ph == PH_REFINE && bisect && (int)r.k >= v.bisect_steps) {
            r.t = r.hi();
            r.P = add(f3(v.cam.C), mul(r.V, r.t));
            r.l = r.l_hi();
            enter_normal(v, r);
        }
    }

    /* reversing walk (kernel.cu:420-441): one sample, the group's first */
    if (__ballot(ph == PH_REFINE && !bisect) != 0) {
        const float l0 = st.lam[min(base, 64u * kMaxSlots - 1u)];
        if (ph == PH_REFINE && !bisect) {
            r.t += r.Qdt();
            r.P = add(r.P, r.QdV());
            ++r.evals;
            ++r.k;
            r.l = l0;
            if (l0 == thr) {
                enter_normal(v, r);
            } else {
                const bool osign = (r.flags & 2u) != 0;
                const bool sign = (l0 < thr) ? false : true;
                if (sign) r.flags |= 2u;
                else r.flags &= ~2u;
                if (sign != osign) {
                    r.Qdt() *= -0.5f;
                    r.set_QdV(mul(r.QdV(), -0.5f));
                }
                if (!walk_continues(r)) enter_normal(v, r);
            }
        }
    }

    /* normal: the six samples in order (kernel.cu:456-473) */
    if (__ballot(ph == PH_NORMAL) != 0) {
        const uint32_t n6 = K < 6u ? K : 6u;
        float l6[6];
#pragma unroll
        for (uint32_t j = 0; j < 6u; ++j) l6[j] = st.lam[min(base + j, 64u * kMaxSlots - 1u)];
#pragma unroll
        for (uint32_t j = 0; j < 6u; ++j) {
            const float lj = l6[j];
            if (j < n6 && ph == PH_NORMAL && r.k < 6u) {
                const uint32_t n = r.k;
                const float sl = (n & 1u) ? lj : -lj;
                if (n < 2) r.gx() = (n == 0) ? sl : r.gx() + sl;
                else if (n < 4) r.gy() = (n == 2) ? sl : r.gy() + sl;
                else r.gz() = (n == 4) ? sl : r.gz() + sl;
                ++r.k;
                ++r.evals;
            }
        }
        if (ph == PH_NORMAL && r.k >= 6u) {
            const F3 N = normalized({r.gx(), r.gy(), r.gz()});
            if (leader) write_hit(v, r, N);
            /* the colour is computed later, from the normal kept here */
            r.gx() = N.x;
            r.gy() = N.y;
            r.gz() = N.z;
            r.flags |= 4u;
            r.phase = PH_IDLE;
            finished = true;
        }
    }
    return finished;
}

__device__ __forceinline__ uint32_t uniform(uint32_t x)
{
    return __builtin_amdgcn_readfirstlane(x);
}

/* The chunks of up to 64 lists, held across a wave: lane j holds list j's
 * entry count and the chunks of lists 0..j (an inclusive scan), so a chunk
 * finds its list with one ballot. The counts come in one vector load rather
 * than a walk of dependent scalar loads at every wave's start. */
struct ChunkMap {
    uint32_t cnt;      /* lane j: entries of list j (0 past the lists) */
    uint32_t end;      /* lane j: chunks of lists 0..j */
    uint32_t total;    /* chunks of all lists */
};

__device__ __forceinline__ ChunkMap chunk_map(uint32_t cnt, uint32_t per, uint32_t lane)
{
    ChunkMap m;
    m.cnt = cnt;
    uint32_t x = (cnt + per - 1u) / per;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off) x += y;
    }
    m.end = x;
    m.total = uniform((uint32_t)__shfl((int)x, 63));
    return m;
}

/* List p holding chunk c (< m.total), the chunks before p and p's entries.
 * Every lane of the wave takes part. */
__device__ __forceinline__ void chunk_find(const ChunkMap &m, uint32_t c, uint32_t per, uint32_t &p,
                                           uint32_t &acc, uint32_t &cnt)
{
    p = uniform((uint32_t)__popcll(__ballot(m.end <= c)));
    cnt = uniform((uint32_t)__shfl((int)m.cnt, (int)p));
    acc = uniform((uint32_t)__shfl((int)m.end, (int)p)) - (cnt + per - 1u) / per;
}

/* Rays [first, last) of view `view` for tile t. Level 0 tiles cover the
 * views' pixel blocks; later levels cover each view's handed-over rays,
 * `tile` entries per tile. Level 0 with up to 64 views: lane v reads view
 * v's first tile and one ballot finds the view; otherwise scalar scans (a
 * held scan here costs the kernel spills). Every lane of the wave takes
 * part. Returns false when t is past the last tile. */
__device__ __forceinline__ bool tile_lookup(const ViewDesc *__restrict__ views,
                                            const RenderArgs &a, uint32_t t,
                                            uint32_t tile, uint32_t &view,
                                            uint32_t &first, uint32_t &last)
{
    const uint32_t lane = __lane_id();
    const bool held = a.n_views <= 64u;
    if (a.level == 0) {
        if (t >= a.total_tiles) return false;
        uint32_t v = 0;
        if (held) {
            /* tile_begin rises with the view index and view 0 begins at 0 */
            const bool before = lane < a.n_views && views[lane].tile_begin <= t;
            v = (uint32_t)__popcll(__ballot(before)) - 1u;
        } else {
            while (v + 1 < a.n_views && views[uniform(v + 1)].tile_begin <= t) ++v;
        }
        view = uniform(v);
        const ViewDesc &d = views[view];
        const uint32_t len = uniform(d.tile_len);
        first = uniform((t - d.tile_begin) * len);
        last = uniform(min(first + len, d.n_rays));
        return true;
    }
    uint32_t acc = 0;
    for (uint32_t v = 0; v < a.n_views; ++v) {
        const uint32_t cnt = uniform(min(a.in_counts[v], views[uniform(v)].cont_cap));
        const uint32_t nt = (cnt + tile - 1) / tile;
        if (t < acc + nt) {
            view = uniform(v);
            first = uniform((t - acc) * tile);
            last = uniform(min(first + tile, cnt));
            return true;
        }
        acc += nt;
    }
    return false;
}

/* Next tile from the claim counters (wave-uniform; every lane calls it), or
 * 0xFFFFFFFF when every counter has run dry. `head` and `tried` persist
 * across calls: the counter in use and how many were found dry. Once off its
 * first counter a wave reads a counter before adding to it, so dry counters
 * cost a load, not another queued atomic. */
__device__ __forceinline__ uint32_t claim_tile(uint32_t *heads, uint32_t n_waves,
                                               uint32_t level_tiles, uint32_t &head,
                                               uint32_t &tried)
{
    while (tried < kTileHeads) {
        const uint32_t first = n_waves + head;
        const uint32_t cnt =
            level_tiles > first ? (level_tiles - first + kTileHeads - 1u) / kTileHeads : 0u;
        if (cnt > 0) {
            uint32_t *h = heads + head * kHeadStride;
            uint32_t c = cnt;
            if (__lane_id() == 0) {
                const uint32_t seen =
                    tried == 0 ? 0u
                               : __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (seen < cnt) c = atomicAdd(h, 1u);
            }
            c = uniform((uint32_t)__shfl((int)c, 0));
            if (c < cnt) return first + kTileHeads * c;
        }
        head = (head + 1u) % kTileHeads;
        ++tried;
    }
    return 0xFFFFFFFFu;
}

/* Coordinates of voxel i of a volume block: the volume kernels' formula
 * lo + (span * index) / n per axis (kernel.cu:527-529 for the box [0, 4)). */
__device__ __forceinline__ void block_voxel(const ViewDesc &v, uint32_t i, float &x, float &y,
                                            float &z)
{
    const uint32_t plane = v.vnx * v.vny;
    const uint32_t lz = i / plane;
    const uint32_t rem = i - lz * plane;
    const uint32_t iy = rem / v.vnx;
    const uint32_t ix = rem - iy * v.vnx;
    x = v.vlo[0] + (v.vspan[0] * (float)ix) / (float)v.vnx;
    y = v.vlo[1] + (v.vspan[1] * (float)iy) / (float)v.vny;
    z = v.vlo[2] + (v.vspan[2] * (float)(v.vz0 + lz)) / (float)v.vnz;
}

/* Voxels [first, last) of volume block `item`, one lane per voxel, 64 at a
 * time, in the precision of the launch. The product form gives the same
 * bits as the volume kernels (lambda_product<L, true>) and lists the voxels
 * its rule selects (|lambda| < margin, or window halves that disagree by
 * more than tau below band) for the FP64 pass, one atomic per wave. Returns
 * the voxels this lane evaluated. */
template <int W, int PREC>
__device__ uint32_t voxel_tile(const ViewDesc &v, uint32_t item, uint32_t first, uint32_t last,
                               const RenderArgs &a)
{
    const uint32_t lane = __lane_id();
    uint32_t done = 0;
    for (uint32_t c = first; c < last; c += 64u) {
        const uint32_t i = c + lane;
        bool want = false;
        if (i < last) {
            float x, y, z;
            block_voxel(v, i, x, y, z);
            float lam;
            if (PREC == LYAP_PREC_PRODUCT) {
                const ProductResult r = lambda_product_halves_w<W>(x, y, z, v.e);
                lam = r.lam;
                want = !(__builtin_fabsf(lam) >= v.vmargin);     /* NaN included */
                if (v.vtau > 0.0f) {
                    const float lo_half = r.half1 < r.half2 ? r.half1 : r.half2;
                    want = want || (lo_half < v.vband && __builtin_fabsf(r.half1 - r.half2) > v.vtau);
                }
            } else {
                lam = eval_lambda<W, PREC>(v.e, x, y, z);
            }
            v.vol_out[i] = lam;
            ++done;
        }
        if (PREC == LYAP_PREC_PRODUCT && v.vlist_cap != 0u) {
            const uint64_t mask = __ballot(want);
            if (mask != 0) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(a.vox_counts + item * kHeadStride, (uint32_t)__popcll(mask));
                base = uniform((uint32_t)__shfl((int)base, 0));
                if (want) {
                    const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
                    if (slot < v.vlist_cap) a.vox_list[v.vlist_base + slot] = i;
                }
            }
        }
    }
    return done;
}

/* Three waves per SIMD (at most 168 VGPRs): the 16-slot FP32 kernels fit
 * without spilling; the 32-slot, FP64 and margin kernels would otherwise
 * take 200+ VGPRs and run two waves, and measure 20-30% slower that way than
 * with the few spills this target costs them. */
template <int W, int PREC, bool MARGIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
void render_persistent(const ViewDesc *__restrict__ views, RenderArgs a)
{
    __shared__ RoundStash stash[4];   /* each wave's samples of the round */
    __shared__ ShadeList shades[4];   /* each wave's hits waiting for their colour */
    const uint32_t lane = __lane_id();
    Ray r;
    r.phase = PH_IDLE;
    r.pix = 0;
    r.evals = 0;
    uint32_t done_evals = 0;
    ShadeList &sh = shades[threadIdx.x >> 6];
    uint32_t sh_count = 0;            /* entries of sh, all of view cur_view */
    __shared__ SetupList setups[4];   /* each wave's set-up rays of its tile */
    SetupList &su = setups[threadIdx.x >> 6];
    uint32_t hpos = 0, hcount = 0;    /* untaken entries [hpos, hcount) of su */
    /* pixel rays are set up a tile at a time, except in the reference's
     * shape, where each lane keeps its own ray of the block */
    const bool batched = a.level == 0 && !a.direct;

    /* samples per lane and round; a group of K lanes takes K * S samples */
    const uint32_t S = uniform(a.slots >= 1u && a.slots <= kMaxSlots ? a.slots : 1u);
    uint32_t K0 = a.lanes_per_ray;    /* lanes per ray while rays remain */
    uint32_t tile = a.tile_rays;
    if (a.level > 0) {
        /* size the groups to the number of rays handed over: samples per
         * ray and round as lanes would be if each took one */
        uint32_t total = 0;
        if (a.n_views <= 64u) {
            uint32_t n = 0;
            if (__lane_id() < a.n_views) n = min(a.in_counts[__lane_id()], views[__lane_id()].cont_cap);
            for (int off = 32; off > 0; off >>= 1) n += (uint32_t)__shfl_xor((int)n, off);
            total = uniform(n);
        } else {
            for (uint32_t v = 0; v < a.n_views; ++v)
                total += uniform(min(a.in_counts[v], views[uniform(v)].cont_cap));
        }
        if (total == 0) return;
        K0 = 1;
        while (K0 < 64u && (uint64_t)a.resident_lanes >= (uint64_t)total * (2u * K0)) K0 *= 2;
        /* one more doubling while the rays still fill two thirds of the
         * resident lanes: handed-over rays are long, so shorter trips pay */
        if (K0 < 64u && 2ull * a.resident_lanes >= 3ull * total * K0) K0 *= 2;
        K0 = K0 / S > 0u ? K0 / S : 1u;
        tile = uniform(64u / K0);
    }
    if (K0 * S > 64u) K0 = 64u / S;
    K0 = uniform(K0);
    /* tiles of this launch: pixel tiles, or each view's handed-over rays */
    uint32_t level_tiles = a.total_tiles;
    if (a.level > 0) {
        level_tiles = 0;
        if (a.n_views <= 64u) {
            uint32_t n = 0;
            if (__lane_id() < a.n_views) n = min(a.in_counts[__lane_id()], views[__lane_id()].cont_cap);
            n = (n + tile - 1u) / tile;
            for (int off = 32; off > 0; off >>= 1) n += (uint32_t)__shfl_xor((int)n, off);
            level_tiles = uniform(n);
        } else {
            for (uint32_t v = 0; v < a.n_views; ++v) {
                const uint32_t cnt = uniform(min(a.in_counts[v], views[uniform(v)].cont_cap));
                level_tiles += (cnt + tile - 1u) / tile;
            }
        }
    }
    uint32_t K = K0;                  /* lanes per ray; wave-uniform */
    const unsigned long long t_start = a.wave_log ? clock64() : 0ull;
    /* the constant-rate clock, common to every XCD, for the launch's timeline */
    const unsigned long long w_start = a.wave_log ? wall_clock64() : 0ull;
    unsigned long long c_eval = 0, c_refill = 0, c_combine = 0, c_claim = 0, c_hand = 0;
    unsigned long long c_start = 0;
    unsigned long long c_mark = t_start;
    uint32_t rounds = 0, rays_started = 0;
    const uint32_t n_waves = gridDim.x * 4u;
    const uint32_t wid = uniform(blockIdx.x * 4u + threadIdx.x / 64u);
    uint32_t head = blockIdx.x % kTileHeads;  /* claim counter in use */
    uint32_t heads_tried = 0;                 /* counters found dry */
    bool first_tile = true;
    uint32_t cur_view = 0xFFFFFFFFu;  /* view the lanes work on */
    uint32_t pos = 0, end = 0;        /* unstarted rays of the current tile */
    uint32_t pend_tile = 0xFFFFFFFFu; /* claimed tile of another view */
    bool queue_empty = false;

    for (;;) {
        if (a.wave_log) {
            const unsigned long long now = clock64();
            c_combine += now - c_mark;
            c_mark = now;
        }
        /* ---- refill idle groups (one ray per group of K lanes) ---------- */
        /* groups are handed rays first and start them after the loop, so
         * the ray state is written once per refill */
        bool started = false;
        bool pend = false;            /* the group was given a ray this refill */
        uint32_t p_id = 0, p_pix = 0; /* its handed-over record, or its pixel */
        F3 p_V = {0.f, 0.f, 0.f};
        float p_t0 = 0.f, p_t1 = 0.f;
        for (;;) {
            /* the reference's shape: a wave's rays all end before new ones start */
            if (a.direct && __ballot(r.phase != PH_IDLE || pend) != 0) break;
            const uint32_t rank = lane % K;
            const bool need = rank == 0 && lane + K <= 64u && r.phase == PH_IDLE && !pend;
            const uint64_t mask = __ballot(need);
            if (mask == 0) break;
            if (batched && hpos >= hcount && pos < end) {
                /* set up the tile's next (up to) 64 rays, one per lane */
                const unsigned long long s0 = a.wave_log ? clock64() : 0ull;
                const ViewDesc &v = views[uniform(cur_view)];
                const uint32_t n = min(end - pos, 64u);
                uint32_t pix = 0;
                F3 V = {0.f, 0.f, 0.f};
                float t0 = 0.f, t1 = 0.f;
                const int k = lane < n ? ray_setup(v, pos + lane, pix, V, t0, t1) : RAY_PAD;
                if (k == RAY_MISS) {
                    Ray m;
                    m.pix = pix;
                    m.evals = 0;
                    write_miss(v, m);
                }
                const uint64_t hits = __ballot(k == RAY_HIT);
                if (k == RAY_HIT) {
                    const uint32_t i = (uint32_t)__popcll(hits & ((1ull << lane) - 1ull));
                    su.vx[i] = V.x;
                    su.vy[i] = V.y;
                    su.vz[i] = V.z;
                    su.t0[i] = t0;
                    su.t1[i] = t1;
                    su.pix[i] = pix;
                }
                wave_sync_lds();
                hpos = 0;
                hcount = (uint32_t)__popcll(hits);
                pos += n;
                if (a.wave_log) c_start += clock64() - s0;
                continue;
            }
            if (pos >= end && hpos >= hcount) {
                if (pend_tile != 0xFFFFFFFFu) {
                    if (__ballot(r.phase != PH_IDLE || pend) != 0) break;   /* drain first */
                    uint32_t tv, f, l;
                    tile_lookup(views, a, pend_tile, tile, tv, f, l);
                    pend_tile = 0xFFFFFFFFu;
                    if (sh_count) {
                        shade_list(views[uniform(cur_view)], sh, sh_count);
                        sh_count = 0;
                    }
                    if (uniform(views[tv].kind) == kItemVolume) {
                        /* a tile of a volume block: every lane is idle, so the
                         * wave evaluates its voxels and claims again */
                        const unsigned long long s0 = a.wave_log ? clock64() : 0ull;
                        done_evals += voxel_tile<W, PREC>(views[tv], tv, f, l, a);
                        if (a.wave_log) c_eval += clock64() - s0;
                        cur_view = 0xFFFFFFFFu;
                        continue;
                    }
                    cur_view = tv;
                    pos = f;
                    end = l;
                    K = K0;
                    continue;
                }
                if (queue_empty) break;
                const unsigned long long c0 = a.wave_log ? clock64() : 0ull;
                uint32_t t;
                if (first_tile) {
                    first_tile = false;
                    t = wid;
                } else {
                    t = claim_tile(a.tile_heads, n_waves, level_tiles, head, heads_tried);
                }
                if (a.wave_log) c_claim += clock64() - c0;
                uint32_t tv, f, l;
                if (t >= level_tiles || !tile_lookup(views, a, t, tile, tv, f, l)) {
                    queue_empty = true;
                    break;
                }
                if (tv != cur_view) { pend_tile = t; continue; }
                pos = f;
                end = l;
                continue;
            }
            const uint32_t avail = batched ? hcount - hpos : end - pos;
            /* every lane of a needing group takes the group's ray: the
             * group's rank among the needing leaders below it */
            const uint32_t gb = lane - rank;
            const bool gneed = ((mask >> gb) & 1ull) != 0;
            const uint32_t rk = (uint32_t)__popcll(mask & ((1ull << gb) - 1ull));
            const unsigned long long s0 = a.wave_log ? clock64() : 0ull;
            if (gneed && rk < avail) {
                const ViewDesc &v = views[uniform(cur_view)];
                if (batched) {
                    const uint32_t i = hpos + rk;
                    p_pix = su.pix[i];
                    p_V = {su.vx[i], su.vy[i], su.vz[i]};
                    p_t0 = su.t0[i];
                    p_t1 = su.t1[i];
                    pend = true;
                } else if (a.level == 0) {
                    const int k = ray_setup(v, pos + rk, p_pix, p_V, p_t0, p_t1);
                    if (k == RAY_HIT) pend = true;
                    else if (k == RAY_MISS && rank == 0) {
                        Ray m;
                        m.pix = p_pix;
                        m.evals = 0;
                        write_miss(v, m);
                    }
                } else {
                    p_id = v.cont_base + pos + rk;
                    pend = true;
                }
            }
            if (a.wave_log) c_start += clock64() - s0;
            const uint32_t taken = (uint32_t)__popcll(mask);
            if (batched) hpos += taken < avail ? taken : avail;
            else pos += taken < avail ? taken : avail;
        }
        if (pend) {
            if (a.level == 0) start_hit(views[uniform(cur_view)], p_pix, p_V, p_t0, p_t1, r);
            else resume_ray(a.cont_in[p_id], r);
            started = true;
        }
        rays_started += (uint32_t)__popcll(__ballot(started && lane % K == 0u));
        const uint64_t busy_mask = __ballot(r.phase != PH_IDLE);
        if (busy_mask == 0) {
            /* a refill that left every group idle (the wave's last one, or
             * tiles whose rays all missed the box) is refill time too */
            if (a.wave_log) {
                const unsigned long long now = clock64();
                c_refill += now - c_mark;
                c_mark = now;
            }
            if (queue_empty && pend_tile == 0xFFFFFFFFu) break;
            continue;
        }
        /* ---- no more rays for this view: spread idle lanes over busy rays */
        const bool starved = !a.direct && pos >= end && hpos >= hcount &&
                             (queue_empty || pend_tile != 0xFFFFFFFFu);
        if (starved) {
            const uint64_t leaders = __ballot(r.phase != PH_IDLE && lane % K == 0u);
            const uint32_t nb = (uint32_t)__popcll(leaders);
            uint32_t K2 = 64u / nb;
            if (K2 * S > 64u) K2 = 64u / S;
            if (K2 > K) {
                /* group i of the new layout takes the i-th busy ray */
                uint32_t src = lane;
                bool keep = false;
                uint64_t m = leaders;
                for (uint32_t i = 0; i < nb; ++i) {
                    const uint32_t b = (uint32_t)__ffsll((unsigned long long)m) - 1u;
                    m &= m - 1;
                    if (lane / K2 == i) { src = b; keep = true; }
                }
                Ray c = shuffle_ray(r, (int)src);
                r = c;
                if (!keep) r.phase = PH_IDLE;
                K = K2;
            }
        }
        ++rounds;
        if (a.wave_log) {
            const unsigned long long now = clock64();
            c_refill += now - c_mark;
            c_mark = now;
        }
        /* ---- S lambda evaluations per lane ------------------------------ */
        /* The round's control (combine, hand-over, refill) is short, branchy
         * and waits on LDS and scalar results; the wave runs it at a raised
         * issue priority and its evaluations at the normal one, so the
         * SIMD serves the control first and fills the gaps with the other
         * waves' evaluations (4 % on the 512x512 frame). */
        __builtin_amdgcn_s_setprio(0);
        const ViewDesc &v = views[uniform(cur_view)];
        const uint32_t rank = lane % K;
        const uint32_t gbase = lane - rank;
        const uint32_t KS = K * S;            /* samples per group and round */
        const bool busy = r.phase != PH_IDLE;
        RoundStash &st = stash[threadIdx.x >> 6];
        const uint64_t kmask = K >= 64u ? ~0ull : ((1ull << K) - 1ull);
        uint64_t evm = 0;                     /* the group's sample bits (see combine) */
        for (uint32_t sl = 0; sl < S; ++sl) {
            const uint32_t q = sl * K + rank;
            float lam = NAN;
            F3 sp = {0.f, 0.f, 0.f};
            float ts = 0.f;
            const bool valid = busy && sample_point(v, r, q, KS, sp, ts);
            if (valid) {
                lam = eval_lambda<W, PREC>(v.e, sp.x, sp.y, sp.z);
                if (MARGIN && PREC != LYAP_PREC_FP64 && r.phase == PH_REFINE &&
                    v.refine == LYAP_REFINE_BISECT &&
                    !(__builtin_fabsf(lam - v.prm.opaqueThreshold) >= v.margin))
                    lam = eval_fp64_call<W>(v.e, sp.x, sp.y, sp.z);
            }
            /* marching: does sample q stop the march (opaque, invalid, a
             * near/far toggle, or jitter, whose step needs this lambda)?
             * refining: is its lambda above the opaque threshold? */
            bool ev = true;
            if (busy && r.phase == PH_REFINE) ev = lam > v.prm.opaqueThreshold;
            if (busy && r.phase == PH_MARCH && valid) {
                const LyapParams &prm = v.prm;
                const bool near_ = (r.flags & 1u) != 0;
                const bool toggles = (r.index + q) > 0 && r.Ndt != r.Fdt &&
                                     ((lam <= prm.nearThreshold) != near_);
                ev = !(lam > prm.opaqueThreshold) || toggles || prm.jitter != 0.0f;
            }
            evm |= ((__ballot(ev) >> gbase) & kmask) << (sl * K);
            /* lanes past the last whole group (after a regroup to a K that
             * does not divide 64) hold no ray and have no stash slots */
            if (gbase + K <= 64u) {
                const uint32_t idx = gbase * S + q;
                st.lam[idx] = lam;
                st.t[idx] = ts;
                st.x[idx] = sp.x;
                st.y[idx] = sp.y;
                st.z[idx] = sp.z;
                st.ok[idx] = valid ? 1u : 0u;
                /* march sample q has march index r.index + q; the first
                 * sample of a ray adds to neither sum (kernel.cu:370-375) */
                float sc = 0.0f, sa = 0.0f;
                if (busy && r.phase == PH_MARCH && r.index + q > 0u) {
                    if (lam > v.prm.chaosThreshold) sc = lam;
                    else if (lam > v.prm.opaqueThreshold) sa = lam;
                }
                st.ca[idx] = sc;
                st.aa[idx] = sa;
            }
        }
        wave_sync_lds();
        __builtin_amdgcn_s_setprio(2);
        if (a.wave_log) {
            const unsigned long long now = clock64();
            c_eval += now - c_mark;
            c_mark = now;
        }
        const int ph0 = r.phase;
        const bool fin = combine(v, r, evm, st, gbase * S, KS, busy, rank == 0) && rank == 0;
        if (fin) done_evals += r.evals;
        /* a ray whose march ended this round leaves for the refinement
         * launch, which bisects it on two lanes: here a bisection trip would
         * use K lanes to resolve log2(K+1) levels. A wave at one lane per ray
         * refines in place, as efficiently and without the record. */
        if (a.refine_out != nullptr && K >= 2u) {
            const bool pk = busy && ph0 == PH_MARCH && r.phase == PH_REFINE;
            const uint64_t pm = __ballot(pk && rank == 0);
            if (pm != 0) {
                const uint32_t sub = blockIdx.x % kRefineSubs;
                uint32_t base = 0;
                if (lane == 0)
                    base = atomicAdd(a.refine_counts + (uniform(cur_view) * kRefineSubs + sub) * kHeadStride,
                                     (uint32_t)__popcll(pm));
                base = uniform((uint32_t)__shfl((int)base, 0));
                uint32_t ok = 0;
                if (pk && rank == 0) {
                    const uint32_t slot = base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
                    if (slot < v.refine_cap) {
                        RefineRec q;
                        q.pix = r.pix;
                        q.evals = r.evals;
                        q.Vx = r.V.x;
                        q.Vy = r.V.y;
                        q.Vz = r.V.z;
                        q.lo = r.lo();
                        q.hi = r.hi();
                        q.l_hi = r.l_hi();
                        q.dt = r.dt;
                        q.a = r.a;
                        q.c = r.c;
                        q.l_lo = r.l_lo();
                        a.refine_out[v.refine_base + sub * v.refine_cap + slot] = q;
                        ok = 1;
                    }
                }
                ok = (uint32_t)__shfl((int)ok, (int)gbase);
                if (pk && ok) r.phase = PH_IDLE;
            }
        }
        /* queue this round's hits for shading; shade the queue first when
         * they would not fit */
        const uint64_t hm = __ballot(fin && (r.flags & 4u) != 0u);
        if (hm != 0) {
            const uint32_t nh = (uint32_t)__popcll(hm);
            if (sh_count + nh > 64u) {
                shade_list(v, sh, sh_count);
                sh_count = 0;
                wave_sync_lds();
            }
            if ((hm >> lane) & 1ull) {
                const uint32_t i = sh_count + (uint32_t)__popcll(hm & ((1ull << lane) - 1ull));
                sh.px[i] = r.P.x;
                sh.py[i] = r.P.y;
                sh.pz[i] = r.P.z;
                sh.nx[i] = r.gx();
                sh.ny[i] = r.gy();
                sh.nz[i] = r.gz();
                sh.a[i] = r.a;
                sh.c[i] = r.c;
                sh.pix[i] = r.pix;
            }
            sh_count += nh;
            wave_sync_lds();
        }
        /* hand rays that used this launch's sample budget to the next launch:
         * one atomic per wave reserves the records of all its leaders */
        if (a.budget != 0) {
            const bool over = r.phase == PH_MARCH && r.launch_samples >= a.budget;
            const uint64_t om = __ballot(over && rank == 0);
            if (om != 0) {
                const unsigned long long h0 = a.wave_log ? clock64() : 0ull;
                uint32_t base = 0;
                if (lane == 0)
                    base = atomicAdd(&a.out_counts[uniform(cur_view)], (uint32_t)__popcll(om));
                base = uniform((uint32_t)__shfl((int)base, 0));
                uint32_t ok = 0;
                if (over && rank == 0) {
                    const uint32_t slot = base + (uint32_t)__popcll(om & ((1ull << lane) - 1ull));
                    if (slot < v.cont_cap) {
                        ContEntry c;
                        suspend_ray(r, c);
                        a.cont_out[v.cont_base + slot] = c;
                        ok = 1;
                    }
                }
                ok = (uint32_t)__shfl((int)ok, (int)gbase);
                if (over && ok) r.phase = PH_IDLE;
                if (a.wave_log) c_hand += clock64() - h0;
            }
        }
    }
    if (sh_count) shade_list(views[uniform(cur_view)], sh, sh_count);
    /* one 64-bit atomic per wave for the evaluation count */
    unsigned long long sum = done_evals;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
    if (lane == 0 && sum) atomicAdd(a.eval_counter, sum);
    if (a.wave_log && lane == 0 && wid < 8192u) {
        unsigned long long *o = a.wave_log + kWaveLogWords * (wid + 8192ull * a.level);
        o[0] = rounds;
        o[1] = rays_started;
        o[2] = t_start;
        o[3] = clock64();
        o[4] = c_eval;
        o[5] = c_refill;
        o[6] = c_combine;
        o[7] = c_claim;
        o[8] = c_hand;
        o[9] = c_start;
        o[10] = w_start;
        o[11] = wall_clock64();
        for (uint32_t k = 12; k < kWaveLogWords; ++k) o[k] = 0;
    }
}

/* ---- march passes ------------------------------------------------------- */

/* A small frame has fewer rays that meet the box than the chip has lanes, so
 * its march speculates: K lanes of a ray evaluate the next K samples at once.
 * In the persistent launch the waves also carry the refill, hand-over and
 * shading state, which holds them to three per SIMD. March passes split the
 * march instead into launches of one round each: march_pass gives each ray
 * K lanes for one round (K samples, the persistent launch's sample
 * positions, evaluation and combine) and lists the rays still marching for
 * the next pass, K rising from pass to pass as the list shrinks; ray_start
 * sets up the pixels' rays (one lane each) beforehand and lists those that
 * meet the box, and the last pass runs its rays to the end of their march.
 * Rays whose march ended park for the refinement launch as before. A ray's
 * arithmetic is the same in every schedule, so the bytes are too. (Setting
 * the rays up in the first pass instead, every lane of a group the same ray,
 * raised the pass kernel's registers and made the frame 8 % slower.) */

/* A marching ray between passes (64 bytes). The passes refine by bisection,
 * where a ray's position is always C + V t, its step Fdt or Fdt /
 * nearMultiplier by the near flag, and both were formed by those same
 * expressions: the record keeps t and Fdt and the resumed ray recomputes P,
 * Ndt and dt bit for bit. */
struct alignas(16) PassRec {
    uint32_t pix, evals;
    uint32_t index_flags;              /* march sample index (< 2^24), flags in the top byte */
    uint32_t m;
    float Vx, Vy, Vz, t, t1, Fdt, last_step, a, c, l, tb, dtb;
};
static_assert(sizeof(PassRec) == 64, "pass record size");

__device__ __forceinline__ void store_pass(const Ray &r, PassRec &q)
{
    q.pix = r.pix;
    q.evals = r.evals;
    q.index_flags = (r.index & 0x00FFFFFFu) | (r.flags << 24);
    q.m = r.m;
    q.Vx = r.V.x; q.Vy = r.V.y; q.Vz = r.V.z;
    q.t = r.t; q.t1 = r.t1; q.Fdt = r.Fdt; q.last_step = r.last_step;
    q.a = r.a; q.c = r.c; q.l = r.l;
    q.tb = r.tb; q.dtb = r.dtb;
}

__device__ __forceinline__ void load_pass(const ViewDesc &v, const PassRec &q, Ray &r)
{
    r.pix = q.pix;
    r.evals = q.evals;
    r.index = q.index_flags & 0x00FFFFFFu;
    r.flags = q.index_flags >> 24;
    r.m = q.m;
    r.V = {q.Vx, q.Vy, q.Vz};
    r.t = q.t; r.t1 = q.t1; r.Fdt = q.Fdt; r.last_step = q.last_step;
    r.a = q.a; r.c = q.c; r.l = q.l;
    r.tb = q.tb; r.dtb = q.dtb;
    r.Ndt = r.Fdt / v.prm.nearMultiplier;          /* start_hit's expression */
    r.dt = (r.flags & 1u) ? r.Ndt : r.Fdt;
    r.P = add(f3(v.cam.C), mul(r.V, r.t));          /* march_step's expression */
    r.k = 0;
    r.launch_samples = 0;
    r.phase = PH_MARCH;
}

struct PassArgs {
    uint32_t n_views;
    uint32_t K;                        /* lanes (samples) per ray in this pass */
    const PassRec *cont_in;            /* march_pass: the rays to advance */
    const uint32_t *in_counts;         /* per view and sub-list, kHeadStride words apart */
    PassRec *cont_out;                 /* the rays still marching */
    uint32_t *out_counts;
    RefineRec *refine_out;
    uint32_t *refine_counts;           /* per view and sub-region, kHeadStride words apart */
    unsigned long long *eval_counter;
    uint32_t rounds;                   /* rounds per ray (0: until the march ends) */
};

/* One wave's samples of a round (one per lane). */
struct PassStash {
    float lam[64], t[64], x[64], y[64], z[64];
    uint32_t ok[64];
    float ca[64], aa[64];
};

/* Listed rays of view v's sub-list s. */
__device__ __forceinline__ uint32_t cont_count(const ViewDesc *__restrict__ views, const uint32_t *counts,
                                               uint32_t v, uint32_t s)
{
    return uniform(min(counts[(v * kContSubs + s) * kHeadStride], views[uniform(v)].cont_cap / kContSubs));
}

/* Appends the rays of this wave's lanes with `want` (leaders) to view vi's
 * sub-list of this block's XCD (block index modulo kContSubs), one atomic
 * per wave; rays that do not fit go on to the next sub-list. A view's
 * sub-lists hold all its rays together, and a full sub-list has every one of
 * its slots written, so each ray finds a slot within kContSubs tries. */
__device__ __forceinline__ void append_cont(const ViewDesc &v, uint32_t vi, const Ray &r, bool want,
                                            PassRec *out, uint32_t *counts)
{
    const uint32_t lane = __lane_id();
    const uint32_t sub_cap = v.cont_cap / kContSubs;
    for (uint32_t tries = 0; tries < kContSubs && __ballot(want) != 0; ++tries) {
        const uint64_t m = __ballot(want);
        const uint32_t sub = (blockIdx.x + tries) % kContSubs;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(counts + (vi * kContSubs + sub) * kHeadStride, (uint32_t)__popcll(m));
        base = uniform((uint32_t)__shfl((int)base, 0));
        if (want) {
            const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (slot < sub_cap) {
                PassRec q;
                store_pass(r, q);
                out[v.cont_base + sub * sub_cap + slot] = q;
                want = false;
            }
        }
    }
}

/* Every pixel ray of every view, one lane each (a block's 256 rays belong
 * to one view: views start at multiples of 256 in this grid): misses of the
 * box are written, the others listed for the first pass with one atomic per
 * block on the sub-list of its XCD (rays that do not fit go on to the next
 * sub-lists, as append_cont). */
__global__ __launch_bounds__(256) void ray_start(const ViewDesc *__restrict__ views, PassArgs a)
{
    __shared__ uint32_t wave_hits[4];
    __shared__ uint32_t block_base;
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    const uint32_t gb = uniform(blockIdx.x * 256u);
    uint32_t vi = 0;
    while (vi + 1u < a.n_views && views[uniform(vi + 1u)].ray_base <= gb) ++vi;
    vi = uniform(vi);
    const ViewDesc &v = views[vi];
    if (v.kind != kItemView || gb >= v.ray_base + v.n_rays) return;   /* the whole block */
    const uint32_t lane = __lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t id = g - v.ray_base;
    uint32_t pix = 0;
    F3 V = {0.f, 0.f, 0.f};
    float t0 = 0.f, t1 = 0.f;
    const int k = id < v.n_rays ? ray_setup(v, id, pix, V, t0, t1) : RAY_PAD;
    Ray r;
    r.phase = PH_IDLE;
    if (k == RAY_MISS) {
        r.pix = pix;
        r.evals = 0;
        write_miss(v, r);
    } else if (k == RAY_HIT) {
        start_hit(v, pix, V, t0, t1, r);
    }
    bool want = k == RAY_HIT;
    const uint64_t m = __ballot(want);
    if (lane == 0) wave_hits[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = wave_hits[0] + wave_hits[1] + wave_hits[2] + wave_hits[3];
        block_base = total ? atomicAdd(a.out_counts + (vi * kContSubs + blockIdx.x % kContSubs) * kHeadStride,
                                       total)
                           : 0u;
    }
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k2 = 0; k2 < w; ++k2) before += wave_hits[k2];
    if (want) {
        const uint32_t sub_cap = v.cont_cap / kContSubs;
        const uint32_t slot = block_base + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (slot < sub_cap) {
            PassRec q;
            store_pass(r, q);
            a.cont_out[v.cont_base + (blockIdx.x % kContSubs) * sub_cap + slot] = q;
            want = false;
        }
    }
    append_cont(v, vi, r, want, a.cont_out, a.out_counts);
}

/* One round for every listed ray, K lanes each (the samples of sample_point
 * ranks 0..K-1), then: rays whose march ended on a miss are written, rays
 * that found their opaque sample park for the refinement launch, the others
 * are listed for the next pass. DRAIN: `rounds` rounds (0: until the march
 * of every ray of the wave has ended) instead of one. */
template <int W, int PREC, bool DRAIN>
__global__ __launch_bounds__(256) void march_pass(const ViewDesc *__restrict__ views, PassArgs a)
{
    __shared__ PassStash stash[4];
    PassStash &st = stash[threadIdx.x >> 6];
    const uint32_t lane = __lane_id();
    const uint32_t K = uniform(a.K);
    const uint32_t per = 64u / K;                 /* rays per wave and round */
    const uint32_t rank = lane % K;
    const uint32_t gbase = lane - rank;
    const uint64_t kmask = K >= 64u ? ~0ull : ((1ull << K) - 1ull);
    const uint32_t n_waves = gridDim.x * 4u;
    const uint32_t wid = uniform(blockIdx.x * 4u + threadIdx.x / 64u);
    const uint32_t pairs = a.n_views * kContSubs;
    const bool held = pairs <= 64u;
    ChunkMap cm{0u, 0u, 0u};
    uint32_t chunks = 0;
    if (held) {
        uint32_t n = 0;
        if (lane < pairs) n = min(a.in_counts[lane * kHeadStride], views[lane / kContSubs].cont_cap / kContSubs);
        cm = chunk_map(n, per, lane);
        chunks = cm.total;
    } else {
        for (uint32_t p = 0; p < pairs; ++p)
            chunks += (cont_count(views, a.in_counts, p / kContSubs, p % kContSubs) + per - 1u) / per;
    }
    uint32_t done_evals = 0;
    for (uint32_t c = wid; c < chunks; c += n_waves) {
        uint32_t acc = 0, p = 0, cnt = 0;
        if (held) {
            chunk_find(cm, c, per, p, acc, cnt);
        } else {
            for (; p < pairs; ++p) {
                cnt = cont_count(views, a.in_counts, p / kContSubs, p % kContSubs);
                const uint32_t nc = (cnt + per - 1u) / per;
                if (c < acc + nc) break;
                acc += nc;
            }
        }
        const uint32_t vi = uniform(p / kContSubs);
        const ViewDesc &v = views[vi];
        const uint32_t i = (c - acc) * per + lane / K;
        bool busy = i < cnt;
        Ray r;
        r.phase = PH_IDLE;
        if (busy) load_pass(v, a.cont_in[v.cont_base + (p % kContSubs) * (v.cont_cap / kContSubs) + i], r);
        /* `rounds` rounds (0: until every ray of the wave has ended); a ray
         * whose march ends leaves its group idle */
        for (uint32_t round = 0;; ++round) {
            float lam = NAN;
            F3 sp = {0.f, 0.f, 0.f};
            float ts = 0.f;
            const bool valid = busy && sample_point(v, r, rank, K, sp, ts);
            if (valid) lam = eval_lambda<W, PREC>(v.e, sp.x, sp.y, sp.z);
            /* does sample `rank` stop the march (opaque, invalid, a near/far
             * toggle, or jitter, whose step needs this lambda)? */
            bool ev = true;
            if (valid) {
                const LyapParams &prm = v.prm;
                const bool near_ = (r.flags & 1u) != 0;
                const bool toggles = (r.index + rank) > 0 && r.Ndt != r.Fdt &&
                                     ((lam <= prm.nearThreshold) != near_);
                ev = !(lam > prm.opaqueThreshold) || toggles || prm.jitter != 0.0f;
            }
            const uint64_t evm = (__ballot(ev) >> gbase) & kmask;
            /* the previous round's combine has read the stash */
            wave_sync_lds();
            st.lam[lane] = lam;
            st.t[lane] = ts;
            st.x[lane] = sp.x;
            st.y[lane] = sp.y;
            st.z[lane] = sp.z;
            st.ok[lane] = valid ? 1u : 0u;
            float sc = 0.0f, sa = 0.0f;
            if (busy && r.index + rank > 0u) {
                if (lam > v.prm.chaosThreshold) sc = lam;
                else if (lam > v.prm.opaqueThreshold) sa = lam;
            }
            st.ca[lane] = sc;
            st.aa[lane] = sa;
            wave_sync_lds();
            bool finished = false;
            combine_march<PassStash, 64u>(v, r, evm, st, gbase, K, busy ? (int)PH_MARCH : (int)PH_IDLE,
                                          rank == 0, finished);
            if (finished && rank == 0) done_evals += r.evals;
            /* park the rays that found their opaque sample: sub-region by
             * block, the next ones while a sub-region is full (the regions
             * hold twice the view's rays, so one has room) */
            bool pk = busy && rank == 0 && r.phase == PH_REFINE;
            for (uint32_t tries = 0; tries < kRefineSubs && __ballot(pk) != 0; ++tries) {
                const uint64_t pm = __ballot(pk);
                const uint32_t sub = (blockIdx.x + tries) % kRefineSubs;
                uint32_t base = 0;
                if (lane == 0)
                    base = atomicAdd(a.refine_counts + (vi * kRefineSubs + sub) * kHeadStride,
                                     (uint32_t)__popcll(pm));
                base = uniform((uint32_t)__shfl((int)base, 0));
                if (pk) {
                    const uint32_t slot = base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull));
                    if (slot < v.refine_cap) {
                        RefineRec q;
                        q.pix = r.pix;
                        q.evals = r.evals;
                        q.Vx = r.V.x;
                        q.Vy = r.V.y;
                        q.Vz = r.V.z;
                        q.lo = r.lo();
                        q.hi = r.hi();
                        q.l_hi = r.l_hi();
                        q.dt = r.dt;
                        q.a = r.a;
                        q.c = r.c;
                        q.l_lo = r.l_lo();
                        a.refine_out[v.refine_base + sub * v.refine_cap + slot] = q;
                        pk = false;
                    }
                }
            }
            busy = busy && r.phase == PH_MARCH;
            if (!DRAIN || (a.rounds != 0 && round + 1u >= a.rounds) || __ballot(busy) == 0) break;
        }
        append_cont(v, vi, r, busy && rank == 0, a.cont_out, a.out_counts);
    }
    unsigned long long sum = done_evals;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
    if (lane == 0 && sum) atomicAdd(a.eval_counter, sum);
}

/* ---- refinement launch -------------------------------------------------- */

struct RefineArgs {
    uint32_t n_views;
    const RefineRec *recs;
    const uint32_t *counts;            /* per view and sub-region, kHeadStride words apart */
    unsigned long long *eval_counter;
};

/* Records of view-sub-region pair p that the march launches wrote. */
__device__ __forceinline__ uint32_t refine_count(const ViewDesc *__restrict__ views,
                                                 const RefineArgs &a, uint32_t p)
{
    return uniform(min(a.counts[p * kHeadStride], views[uniform(p / kRefineSubs)].refine_cap));
}

/* Finishes the parked rays, two lanes per ray and 32 rays of one view per
 * wave: the bisection of [lo, hi] (the midpoint rule), the six normal
 * samples in the order x-, x+, y-, y+, z-, z+ (kernel.cu:456-473), the hit
 * record and its colour. Per bisection trip the first lane evaluates the
 * midpoint and the second the midpoint of the half the crossing of the
 * line through lambda at lo and hi falls in; when that half is the one the
 * first sample chooses, the trip resolves two levels. The normal takes three
 * trips, the minus samples on the first lane and the plus samples on the
 * second. The arithmetic is the march kernel's, step for step, so the bytes
 * are those of refining in place. */
template <int W, int PREC, bool MARGIN>
__device__ __forceinline__ float refine_eval(const ViewDesc &v, F3 sp, float thr)
{
    float lam = eval_lambda<W, PREC>(v.e, sp.x, sp.y, sp.z);
    if (MARGIN && PREC != LYAP_PREC_FP64 && !(__builtin_fabsf(lam - thr) >= v.margin))
        lam = eval_fp64_call<W>(v.e, sp.x, sp.y, sp.z);
    return lam;
}

template <int W, int PREC, bool MARGIN>
__global__ __launch_bounds__(256) void refine_rays(const ViewDesc *__restrict__ views, RefineArgs a)
{
    const uint32_t lane = __lane_id();
    const uint32_t side = lane & 1u;           /* 0: midpoints, minus normal samples */
    const uint32_t n_waves = gridDim.x * 4u;
    const uint32_t wid = uniform(blockIdx.x * 4u + threadIdx.x / 64u);
    const uint32_t pairs = a.n_views * kRefineSubs;
    const bool held = pairs <= 64u;
    ChunkMap cm{0u, 0u, 0u};
    uint32_t total = 0;
    if (held) {
        uint32_t n = 0;
        if (lane < pairs) n = min(a.counts[lane * kHeadStride], views[lane / kRefineSubs].refine_cap);
        cm = chunk_map(n, 32u, lane);
        total = cm.total;
    } else {
        for (uint32_t p = 0; p < pairs; ++p) total += (refine_count(views, a, p) + 31u) / 32u;
    }
    unsigned long long done = 0;
    for (uint32_t t = wid; t < total; t += n_waves) {
        uint32_t acc = 0, p = 0, n = 0;
        if (held) {
            chunk_find(cm, t, 32u, p, acc, n);
        } else {
            for (; p < pairs; ++p) {
                n = refine_count(views, a, p);
                const uint32_t nt = (n + 31u) / 32u;
                if (t < acc + nt) break;
                acc += nt;
            }
        }
        const ViewDesc &v = views[uniform(p / kRefineSubs)];
        const uint32_t i = (t - acc) * 32u + (lane >> 1);
        const bool has = i < n;
        RefineRec q{};
        if (has) q = a.recs[v.refine_base + (p % kRefineSubs) * v.refine_cap + i];
        const F3 C = f3(v.cam.C);
        const F3 V = {q.Vx, q.Vy, q.Vz};
        const float thr = v.prm.opaqueThreshold;
        float lo = q.lo, hi = q.hi, l_hi = q.l_hi, l_lo = q.l_lo;
        uint32_t evals = q.evals;
        /* both lanes of a ray hold the same interval; trips are wave-uniform */
        for (int k = 0; __ballot(has && k < v.bisect_steps) != 0;) {
            const bool go = has && k < v.bisect_steps;
            const float mid = lo + (hi - lo) * 0.5f;
            /* the half the line's crossing falls in, and its midpoint */
            const float f = (l_lo - thr) / (l_lo - l_hi);
            const bool guess_lo = !(lo + (hi - lo) * f > mid);
            const float mid2 = guess_lo ? lo + (mid - lo) * 0.5f : mid + (hi - mid) * 0.5f;
            float lam = 0.f;
            if (go) lam = refine_eval<W, PREC, MARGIN>(v, add(C, mul(V, side ? mid2 : mid)), thr);
            const float lm = __shfl(lam, (int)(lane & ~1u));
            const float lg = __shfl(lam, (int)(lane | 1u));
            if (go) {
                ++evals;
                ++k;
                const bool opaque = !(lm > thr);
                if (opaque) { hi = mid; l_hi = lm; }
                else { lo = mid; l_lo = lm; }
                if (k < v.bisect_steps && opaque == guess_lo) {
                    /* the guessed half is the one taken: its midpoint, formed
                     * from the same interval, is the next level's */
                    ++evals;
                    ++k;
                    const float m2 = lo + (hi - lo) * 0.5f;
                    if (!(lg > thr)) { hi = m2; l_hi = lg; }
                    else { lo = m2; l_lo = lg; }
                }
            }
        }
        Ray r;
        r.pix = q.pix;
        r.t = hi;
        r.P = add(C, mul(V, hi));
        r.l = l_hi;
        r.a = q.a;
        r.c = q.c;
        const float h = (v.normal_h > 0.0f) ? v.normal_h : q.dt * v.prm.gradient;
        float g[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t ax = 0; ax < 3u; ++ax) {
            F3 sp = r.P;
            const float off = side ? h : -h;
            if (ax == 0) sp.x = r.P.x + off;
            else if (ax == 1) sp.y = r.P.y + off;
            else sp.z = r.P.z + off;
            float lam = 0.f;
            if (has) lam = eval_lambda<W, PREC>(v.e, sp.x, sp.y, sp.z);
            const float lminus = __shfl(lam, (int)(lane & ~1u));
            const float lplus = __shfl(lam, (int)(lane | 1u));
            g[ax] = -lminus;
            g[ax] = g[ax] + lplus;
        }
        if (has && side == 0) {
            r.evals = evals + 6u;
            const F3 N = normalized({g[0], g[1], g[2]});
            write_hit(v, r, N);
            const C4 cc = shade(r.P, N, r.a, r.c, v.cam, v.lights, v.num_lights);
            LyapRGBA px;
            px.r = to_u8(cc.x);
            px.g = to_u8(cc.y);
            px.b = to_u8(cc.z);
            px.a = to_u8(cc.w);
            v.rgba[r.pix] = px;
            done += evals + 6u;
        }
    }
    for (int off = 32; off > 0; off >>= 1) done += __shfl_down(done, off);
    if (lane == 0 && done) atomicAdd(a.eval_counter, done);
}
/* ---- FP64 pass over the volume blocks' lists ---------------------------- */

struct VoxelFp64Args {
    uint32_t n_items;
    const uint32_t *list;
    const uint32_t *counts;            /* per item, kHeadStride words apart */
};

__device__ __forceinline__ uint32_t listed(const ViewDesc *__restrict__ views,
                                           const VoxelFp64Args &a, uint32_t p)
{
    return uniform(min(a.counts[p * kHeadStride], views[uniform(p)].vlist_cap));
}

/* Re-evaluates the listed voxels of every volume block in FP64 (the same
 * lambda_fp64 arithmetic as the volume kernels' FP64 pass), 64 entries of
 * one block per wave. */
template <int W>
__global__ __launch_bounds__(256) void voxel_fp64(const ViewDesc *__restrict__ views, VoxelFp64Args a)
{
    const uint32_t lane = __lane_id();
    const uint32_t n_waves = gridDim.x * 4u;
    const uint32_t wid = uniform(blockIdx.x * 4u + threadIdx.x / 64u);
    uint32_t total = 0;
    for (uint32_t p = 0; p < a.n_items; ++p) total += (listed(views, a, p) + 63u) / 64u;
    for (uint32_t t = wid; t < total; t += n_waves) {
        uint32_t acc = 0, p = 0, n = 0;
        for (; p < a.n_items; ++p) {
            n = listed(views, a, p);
            const uint32_t nt = (n + 63u) / 64u;
            if (t < acc + nt) break;
            acc += nt;
        }
        const ViewDesc &v = views[uniform(p)];
        const uint32_t j = (t - acc) * 64u + lane;
        if (j < n) {
            const uint32_t i = a.list[v.vlist_base + j];
            float x, y, z;
            block_voxel(v, i, x, y, z);
            v.vol_out[i] = (float)lambda_fp64_w<W>(x, y, z, v.e);
        }
    }
}

/* Up to four word arrays cleared by one launch: blockIdx.y picks the array. */
struct ZeroRegions {
    uint32_t *p[4];
    uint32_t n[4];
};

__global__ void zero_regions(ZeroRegions z)
{
    const uint32_t r = blockIdx.y;
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (z.p[r] != nullptr && i < z.n[r]) z.p[r][i] = 0u;
}

} /* namespace */

/* Fills a ViewDesc from the API arguments. */
static void fill_view(ViewDesc &v, const EvalDesc &e, const LyapCam &cam,
                      const LyapParams &prm, unsigned w, unsigned h,
                      LyapRGBA *rgba, LyapPoint *points, float *depth,
                      float *lambda, uint32_t *evals, const LyapLight *lights,
                      unsigned num_lights, const LyapRenderMode &mode)
{
    std::memset(&v, 0, sizeof(v));
    v.e = e;
    v.cam = cam;
    v.prm = prm;
    v.w = w;
    v.h = h;
    v.blocks_x = (w + 7) / 8;
    v.n_rays = 64u * v.blocks_x * ((h + 7) / 8);
    v.rgba = rgba;
    v.points = points;
    v.depth = depth;
    v.lambda = lambda;
    v.evals = evals;
    v.lights = lights;
    v.num_lights = num_lights;
    v.refine = mode.refine;
    v.bisect_steps = mode.bisect_steps;
    v.normal_h = mode.normal_h;
    v.margin = mode.fp64_margin;
    v.max_samples = 1u << 20;
    v.kind = kItemView;
}

/* Fills a volume-block item: z slices [z_begin, z_end) of the nx*ny*nz grid
 * over [lo, hi) with output `out` (slice z_begin), under `mode`'s rule. */
static void fill_block(ViewDesc &v, const EvalDesc &e, const LyapParams &prm, const float lo[3],
                       const float hi[3], const uint32_t n[3], uint32_t z_begin, uint32_t z_end,
                       float *out, const LyapVolumeMode &mode)
{
    std::memset(&v, 0, sizeof(v));
    v.kind = kItemVolume;
    v.e = e;
    v.prm = prm;
    v.vol_out = out;
    v.vnx = n[0];
    v.vny = n[1];
    v.vnz = n[2];
    v.vz0 = z_begin;
    for (int k = 0; k < 3; ++k) {
        v.vlo[k] = lo[k];
        v.vspan[k] = hi[k] - lo[k];
    }
    v.n_rays = n[0] * n[1] * (z_end - z_begin);
    const bool reeval = mode.precision == LYAP_PREC_PRODUCT &&
                        (mode.fp64_margin > 0.0f || mode.converge_tau > 0.0f);
    v.vmargin = reeval ? mode.fp64_margin : 0.0f;
    v.vtau = reeval ? mode.converge_tau : 0.0f;
    v.vband = mode.converge_band;
    v.vlist_cap = reeval ? v.n_rays : 0u;
}

/* ---- launch planning ----------------------------------------------------- */

/* Lanes per ray while the queue still has rays: one when the launch has
 * rays for at least four generations of the resident lanes (every sample a
 * lane evaluates is then one the ray needs, and the control work is not
 * replicated), otherwise the power of two that gives about four
 * generations, at most 16 (a small frame's rays finish sooner with several
 * lanes speculating ahead). */
static uint32_t default_lanes_per_ray(uint32_t resident_lanes, uint64_t rays)
{
    uint32_t k = 1;
    while (k < 16u && rays * (uint64_t)k < 4ull * resident_lanes) k *= 2;
    return k;
}

/* Launch shape for `rays` rays: persistent grid, lanes per ray and the rays
 * per work-queue claim. A claim is one generation of a wave's groups, so
 * claims are many times the number of waves and the queue balances. */
struct LaunchPlan {
    uint32_t blocks, lanes_per_ray, tile_rays, resident_lanes, slots;
    uint32_t voxel_tile;          /* voxels per claim of a volume block */
    uint32_t budget[3];           /* level budgets chosen by the tuning, 0 = built-in */
    bool direct;
    bool levels;                  /* the persistent levels only: no march passes */
};

/* Voxels per work-queue claim of a volume block unless tuned: four waves'
 * worth, so claims are cheap next to the evaluations they hand out. */
static const uint32_t kDefaultVoxelTile = 256u;

/* Samples each lane evaluates per round unless tuned. More than one halves
 * the lanes a ray's control work is replicated on, but measured on the
 * 512x512 frame one slot with four lanes per ray is fastest. */
static const uint32_t kDefaultSlots = 1u;

/* March-sample budgets of the launch levels; the last level has none. With
 * fewer than four lanes per ray (a large batch) level 0 keeps a ray for 64
 * samples: 16 made the 64-view batch 17 % slower, 32 2 %, 128 no faster. */
static const uint32_t kLevelBudget[4] = {64u, 64u, 128u, 0u};

/* Level 0 hands rays over sooner when several lanes march each ray (a small
 * launch, where the long rays decide the frame time); with one lane per ray
 * the launch is large and handing over costs more than it saves. */
static uint32_t level_budget(uint32_t level, const LaunchPlan &plan)
{
    if (plan.direct || level >= 3) return 0u;
    if (plan.budget[level]) return plan.budget[level];
    if (level == 0 && plan.lanes_per_ray >= 4u) return 24u;
    return kLevelBudget[level];
}

template <int W, int PREC, bool MARGIN>
static uint32_t resident_blocks_per_cu()
{
    static int cached = -1;
    if (cached < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&render_persistent<W, PREC, MARGIN>), 256, 0) != hipSuccess ||
            n <= 0)
            n = 2;
        cached = n;
    }
    return (uint32_t)cached;
}

static uint32_t kernel_blocks_per_cu(uint32_t W, int precision, bool margin)
{
    if (W == 16) {
        if (precision == LYAP_PREC_PLAIN) return resident_blocks_per_cu<16, LYAP_PREC_PLAIN, false>();
        if (precision == LYAP_PREC_PRODUCT)
            return margin ? resident_blocks_per_cu<16, LYAP_PREC_PRODUCT, true>()
                          : resident_blocks_per_cu<16, LYAP_PREC_PRODUCT, false>();
        return resident_blocks_per_cu<16, LYAP_PREC_FP64, false>();
    }
    if (precision == LYAP_PREC_PLAIN) return resident_blocks_per_cu<32, LYAP_PREC_PLAIN, false>();
    if (precision == LYAP_PREC_PRODUCT)
        return margin ? resident_blocks_per_cu<32, LYAP_PREC_PRODUCT, true>()
                      : resident_blocks_per_cu<32, LYAP_PREC_PRODUCT, false>();
    return resident_blocks_per_cu<32, LYAP_PREC_FP64, false>();
}

static LaunchPlan plan_launch(const lyap_ctx *ctx, uint64_t rays, const LyapTuning &tune,
                              uint32_t W, int precision, bool margin, bool direct = false,
                              bool levels = false)
{
    LaunchPlan p;
    p.direct = direct;
    /* the per-wave record describes the persistent levels */
    p.levels = levels || ctx->debug_wave_log != nullptr;
    const uint32_t cus = ctx->num_cus > 0 ? (uint32_t)ctx->num_cus : 256u;
    uint32_t per_cu = kernel_blocks_per_cu(W, precision, margin);
    if (tune.waves_per_cu && (tune.waves_per_cu + 3u) / 4u < per_cu) per_cu = (tune.waves_per_cu + 3u) / 4u;
    p.blocks = cus * per_cu;
    p.resident_lanes = p.blocks * 256u;
    p.slots = tune.slots ? tune.slots : kDefaultSlots;
    if (p.slots > kMaxSlots) p.slots = kMaxSlots;
    if (tune.lanes_per_ray) {
        p.lanes_per_ray = tune.lanes_per_ray;
    } else {
        /* the default samples per ray and round, spread over the slots */
        const uint32_t ks = default_lanes_per_ray(p.resident_lanes, rays);
        p.lanes_per_ray = ks / p.slots > 0u ? ks / p.slots : 1u;
    }
    if (p.lanes_per_ray > 64u) p.lanes_per_ray = 64u;
    if (p.lanes_per_ray == 0) p.lanes_per_ray = 1;
    if (p.lanes_per_ray * p.slots > 64u) p.lanes_per_ray = 64u / p.slots;
    p.tile_rays = tune.tile_rays ? tune.tile_rays : 64u / p.lanes_per_ray;
    if (p.tile_rays == 0) p.tile_rays = 1;
    p.voxel_tile = tune.voxel_tile ? (tune.voxel_tile + 63u) / 64u * 64u : kDefaultVoxelTile;
    for (int k = 0; k < 3; ++k) p.budget[k] = tune.level_budget[k];
    if (direct) {
        p.lanes_per_ray = 1;
        p.slots = 1;
        p.tile_rays = 64;
    }
    return p;
}

/* ---- launch resources ---------------------------------------------------- */

/* Device and pinned memory one launch sequence uses. The context owns one
 * set for its direct calls; a captured frame owns its own, so a replay never
 * sees memory that a later call reallocated. */
/* Words of the counter block: the evaluation count in the first cache line,
 * then kTileHeads claim counters per launch level. */
constexpr uint32_t kCounterWords = kHeadStride + 4u * kTileHeads * kHeadStride;

struct RenderResources {
    void *view_table = nullptr;      /* device ViewDesc[] */
    size_t view_table_bytes = 0;
    /* pinned host copies of the table, used in turn so a frame's host work
     * overlaps the previous frame on the device; staged[k] is recorded after
     * the launches that read staging[k] */
    void *staging[2] = {nullptr, nullptr};
    size_t staging_bytes[2] = {0, 0};
    hipEvent_t staged[2] = {nullptr, nullptr};
    uint32_t stage = 0;              /* staging buffer of the latest launch */
    uint32_t *counters = nullptr;    /* kCounterWords: [2..3] evaluations, then tile heads */
    void *cont_buf = nullptr;        /* rays handed between levels */
    size_t cont_buf_bytes = 0;
    void *cont_counts = nullptr;     /* per view and level */
    size_t cont_counts_bytes = 0;
    void *refine_buf = nullptr;      /* rays waiting for the refinement launch */
    size_t refine_buf_bytes = 0;
    void *refine_counts = nullptr;   /* per view and sub-region, one 128 B line each */
    size_t refine_counts_bytes = 0;
    void *vox_list = nullptr;        /* volume blocks' FP64 re-evaluation lists */
    size_t vox_list_bytes = 0;
    void *vox_counts = nullptr;      /* per item, one 128 B line each */
    size_t vox_counts_bytes = 0;
};

static int ensure_buffer(lyap_ctx *ctx, void **buf, size_t *cap, size_t bytes, bool may_sync)
{
    if (bytes <= *cap) return LYAP_OK;
    if (may_sync) LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, *buf);
    *buf = nullptr;
    *cap = 0;
    int rc = ctx_alloc(ctx, buf, bytes);
    if (rc) return rc;
    *cap = bytes;
    return LYAP_OK;
}

static void release_resources(lyap_ctx *ctx, RenderResources &r)
{
    ctx_free(ctx, r.view_table);
    ctx_free(ctx, r.cont_buf);
    ctx_free(ctx, r.cont_counts);
    ctx_free(ctx, r.refine_buf);
    ctx_free(ctx, r.refine_counts);
    ctx_free(ctx, r.vox_list);
    ctx_free(ctx, r.vox_counts);
    ctx_free(ctx, r.counters);
    for (int k = 0; k < 2; ++k) {
        if (r.staging[k]) (void)hipHostFree(r.staging[k]);
        if (r.staged[k]) (void)hipEventDestroy(r.staged[k]);
    }
    r = RenderResources{};
}

/* Whether the launch parks bisection rays for the refinement launch: the
 * batched schedule with bisection (the walk refines in place, and the
 * reference's launch shape keeps every ray on its lane). */
static bool parks(const LaunchPlan &plan, const ViewDesc *views)
{
    /* views come before volume blocks in a launch's table */
    return !plan.direct && views[0].kind == kItemView && views[0].refine == LYAP_REFINE_BISECT &&
           views[0].bisect_steps > 0;
}

/* March passes (march_pass): the lanes per ray of each pass, at most
 * kMaxPasses of them; rays still marching after the last pass finish in a
 * persistent launch. */
constexpr uint32_t kMaxPasses = 12;

struct PassPlan {
    uint32_t n = 0;                   /* passes; 0: the persistent levels */
    uint32_t K[kMaxPasses] = {};
    uint32_t R[kMaxPasses] = {};      /* rounds of each pass (0: until the rays end) */
};

/* March passes for small launches of views refined by bisection (the
 * batched schedule's plan gives four or more lanes per ray): K = 4, 8, 16,
 * 32, 64, 64 lanes per ray for one round each, then 64 lanes until the rays
 * end. Measured on the 512x512 frame against other sequences: starting at 8
 * lanes, fewer or more 64-lane passes, or no draining pass (a persistent
 * launch finishing the rays) were 1-12 % slower. LYAP_MARCH_PASSES
 * overrides the sequence ("K" or "KxR" per pass, R rounds, 0 = until the
 * rays end; "0" switches the passes off). */
static PassPlan pass_plan(const LaunchPlan &plan, const ViewDesc *views, unsigned n_views)
{
    PassPlan pp;
    /* a pass finds each chunk's view and sub-list by a scan over them, so
     * launches of many views keep the persistent levels */
    if (plan.direct || plan.levels || !parks(plan, views) || plan.lanes_per_ray < 4u || n_views > 16u)
        return pp;
    for (unsigned i = 0; i < n_views; ++i)
        if (views[i].kind != kItemView) return pp;
    /* "K" or "KxR" per pass, comma separated */
    static const char *env = std::getenv("LYAP_MARCH_PASSES");
    const char *c = env ? env : "4,8,16,32,64,64,64x0";
    /* (the last pass always runs its rays to the end: its records are the
     * passes' own, which the persistent launches do not read) */
    if (c[0] == '0' && c[1] == 0) return pp;
    while (*c && pp.n < kMaxPasses) {
        char *e = nullptr;
        const unsigned long k = std::strtoul(c, &e, 10);
        if (k == 0 || k > 64 || (k & (k - 1))) return PassPlan{};
        unsigned long r = 1;
        if (*e == 'x') r = std::strtoul(e + 1, &e, 10);
        pp.K[pp.n] = (uint32_t)k;
        pp.R[pp.n] = (uint32_t)r;
        ++pp.n;
        if (*e != ',') break;
        c = e + 1;
    }
    if (pp.n) pp.R[pp.n - 1] = 0;
    return pp;
}

/* Prepares `views` for a launch: tile ranges, continuation regions, the
 * common coefficient window; sizes the resources and fills the pinned
 * staging copy. `may_sync` is false while a stream capture is open. */
static int prepare_views(lyap_ctx *ctx, RenderResources &res, ViewDesc *views,
                         unsigned n_views, const LaunchPlan &plan, uint32_t W,
                         bool may_sync, uint32_t *total_tiles)
{
    uint32_t tiles = 0;
    uint64_t cap_total = 0, refine_total = 0, list_total = 0, ray_total = 0;
    const bool passes = pass_plan(plan, views, n_views).n > 0;
    for (unsigned i = 0; i < n_views; ++i) {
        if (views[i].e.W != W) rebuild_blocks(&views[i].e, W);
        views[i].tile_begin = tiles;
        if (views[i].kind == kItemVolume) {
            /* a volume block: voxel tiles, its list region, no ray buffers */
            views[i].tile_len = plan.voxel_tile;
            views[i].n_tiles = (views[i].n_rays + plan.voxel_tile - 1) / plan.voxel_tile;
            tiles += views[i].n_tiles;
            views[i].cont_base = (uint32_t)cap_total;
            views[i].cont_cap = 0;
            views[i].refine_base = (uint32_t)refine_total;
            views[i].refine_cap = 0;
            views[i].vlist_base = (uint32_t)list_total;
            list_total += views[i].vlist_cap;
            continue;
        }
        views[i].tile_len = plan.tile_rays;
        views[i].n_tiles = (views[i].n_rays + plan.tile_rays - 1) / plan.tile_rays;
        tiles += views[i].n_tiles;
        /* continuation region: a quarter of the view's rays, at least 256;
         * march passes list every ray that meets the box */
        uint32_t cap = views[i].n_rays / 4u;
        if (cap < 256u) cap = 256u;
        if (passes) cap = views[i].n_rays;
        /* ray_start's grid: every view starts a block of its own */
        views[i].ray_base = (uint32_t)ray_total;
        ray_total += (views[i].n_rays + 255u) / 256u * 256u;
        views[i].cont_base = (uint32_t)cap_total;
        views[i].cont_cap = cap;
        cap_total += cap;
        /* refinement records: each sub-region holds twice its share of the
         * view's rays before a ray has to be refined in place (or, in the
         * march passes, go to the next sub-region) */
        uint32_t rcap = (2u * views[i].n_rays + kRefineSubs - 1u) / kRefineSubs;
        if (rcap < 64u) rcap = 64u;
        views[i].refine_base = (uint32_t)refine_total;
        views[i].refine_cap = rcap;
        refine_total += (uint64_t)rcap * kRefineSubs;
    }
    if (cap_total > 0xFFFFFFFFull || refine_total > 0xFFFFFFFFull || list_total > 0xFFFFFFFFull ||
        ray_total > 0xFFFFFFFFull)
        return fail(ctx, LYAP_ERR_UNSUPPORTED, "too many rays or voxels in one launch");
    *total_tiles = tiles;
    const size_t table = sizeof(ViewDesc) * n_views;
    int rc = ensure_buffer(ctx, &res.cont_buf, &res.cont_buf_bytes, 2 * cap_total * sizeof(ContEntry), may_sync);
    if (rc) return rc;
    rc = ensure_buffer(ctx, &res.cont_counts, &res.cont_counts_bytes,
                       (size_t)(kMaxPasses + 1) * n_views * kContSubs * kHeadStride * sizeof(uint32_t), may_sync);
    if (rc) return rc;
    if (parks(plan, views)) {
        rc = ensure_buffer(ctx, &res.refine_buf, &res.refine_buf_bytes, refine_total * sizeof(RefineRec),
                           may_sync);
        if (rc) return rc;
        rc = ensure_buffer(ctx, &res.refine_counts, &res.refine_counts_bytes,
                           (size_t)n_views * kRefineSubs * kHeadStride * sizeof(uint32_t), may_sync);
        if (rc) return rc;
    }
    if (list_total > 0) {
        rc = ensure_buffer(ctx, &res.vox_list, &res.vox_list_bytes, list_total * sizeof(uint32_t), may_sync);
        if (rc) return rc;
    }
    rc = ensure_buffer(ctx, &res.vox_counts, &res.vox_counts_bytes,
                       (size_t)n_views * kHeadStride * sizeof(uint32_t), may_sync);
    if (rc) return rc;
    rc = ensure_buffer(ctx, &res.view_table, &res.view_table_bytes, table, may_sync);
    if (rc) return rc;
    if (!res.counters) {
        void *c = nullptr;
        rc = ctx_alloc(ctx, &c, sizeof(uint32_t) * kCounterWords);
        if (rc) return rc;
        res.counters = (uint32_t *)c;
    }
    res.stage ^= 1u;
    const uint32_t k = res.stage;
    if (table > res.staging_bytes[k]) {
        if (may_sync) LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (res.staging[k]) (void)hipHostFree(res.staging[k]);
        res.staging[k] = nullptr;
        res.staging_bytes[k] = 0;
        LYAP_HIP(ctx, hipHostMalloc(&res.staging[k], table, hipHostMallocDefault));
        res.staging_bytes[k] = table;
    }
    if (!res.staged[k]) LYAP_HIP(ctx, hipEventCreateWithFlags(&res.staged[k], hipEventDisableTiming));
    /* the launch before last, which read this buffer, must have copied it
     * (an event never recorded is complete) */
    if (may_sync) LYAP_HIP(ctx, hipEventSynchronize(res.staged[k]));
    std::memcpy(res.staging[k], views, table);
    return LYAP_OK;
}

/* The refinement launch over one record set, on `stream`. */
static int launch_refine(lyap_ctx *ctx, const RenderResources &res, unsigned n_views, uint32_t W,
                         int precision, bool margin)
{
    RefineArgs ra{};
    ra.n_views = n_views;
    ra.recs = (const RefineRec *)res.refine_buf;
    ra.counts = (const uint32_t *)res.refine_counts;
    ra.eval_counter = (unsigned long long *)(res.counters + 2);
    const uint32_t cus = ctx->num_cus > 0 ? (uint32_t)ctx->num_cus : 256u;
    const uint32_t blocks = cus * 8u;
#define LYAP_LAUNCH(WW, PP, MM)                                                                     \
    hipLaunchKernelGGL((refine_rays<WW, PP, MM>), dim3(blocks), dim3(256), 0, ctx->stream,          \
                       (const ViewDesc *)res.view_table, ra)
    if (W == 16) {
        if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH(16, LYAP_PREC_PLAIN, false);
        else if (precision == LYAP_PREC_PRODUCT && margin) LYAP_LAUNCH(16, LYAP_PREC_PRODUCT, true);
        else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH(16, LYAP_PREC_PRODUCT, false);
        else LYAP_LAUNCH(16, LYAP_PREC_FP64, false);
    } else {
        if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH(32, LYAP_PREC_PLAIN, false);
        else if (precision == LYAP_PREC_PRODUCT && margin) LYAP_LAUNCH(32, LYAP_PREC_PRODUCT, true);
        else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH(32, LYAP_PREC_PRODUCT, false);
        else LYAP_LAUNCH(32, LYAP_PREC_FP64, false);
    }
#undef LYAP_LAUNCH
    LYAP_HIP(ctx, hipGetLastError());
    return LYAP_OK;
}

template <int W, int PREC, bool DRAIN>
static uint32_t pass_blocks_per_cu()
{
    static int cached = -1;
    if (cached < 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&march_pass<W, PREC, DRAIN>), 256, 0) != hipSuccess ||
            n <= 0)
            n = 2;
        cached = n;
    }
    return (uint32_t)cached;
}

/* Enqueues ray_start and the march passes: ray_start lists the pixels' rays
 * that meet the box under counts[0], pass p reads the rays listed under
 * counts[p] and lists the rays still marching under counts[p + 1] (the two
 * buffers alternate). */
static int enqueue_passes(lyap_ctx *ctx, RenderResources &res, const ViewDesc *views, unsigned n_views,
                          const PassPlan &pp, uint32_t W, int precision)
{
    uint64_t cap_total = 0, rays = 0;
    for (unsigned i = 0; i < n_views; ++i) {
        cap_total += views[i].cont_cap;
        rays += (views[i].n_rays + 255u) / 256u * 256u;
    }
    PassRec *buf[2] = {(PassRec *)res.cont_buf, (PassRec *)res.cont_buf + cap_total};
    uint32_t *counts = (uint32_t *)res.cont_counts;
    const uint32_t pass_words = n_views * kContSubs * kHeadStride;
    const uint32_t cus = ctx->num_cus > 0 ? (uint32_t)ctx->num_cus : 256u;
    PassArgs a{};
    a.n_views = n_views;
    a.refine_out = (RefineRec *)res.refine_buf;
    a.refine_counts = (uint32_t *)res.refine_counts;
    a.eval_counter = (unsigned long long *)(res.counters + 2);
    a.cont_out = buf[0];
    a.out_counts = counts;
    int rc = timing_mark(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(ray_start, dim3((uint32_t)((rays + 255u) / 256u)), dim3(256), 0, ctx->stream,
                       (const ViewDesc *)res.view_table, a);
    LYAP_HIP(ctx, hipGetLastError());
    rc = timing_mark(ctx);
    if (rc) return rc;
    for (uint32_t p = 0; p < pp.n; ++p) {
        a.K = pp.K[p];
        a.rounds = pp.R[p];
        a.cont_in = buf[p & 1u];
        a.in_counts = counts + p * pass_words;
        a.cont_out = buf[(p + 1u) & 1u];
        a.out_counts = counts + (p + 1u) * pass_words;
        rc = timing_mark(ctx);
        if (rc) return rc;
#define LYAP_LAUNCH(WW, PP, DD)                                                                      \
    hipLaunchKernelGGL((march_pass<WW, PP, DD>), dim3(cus * pass_blocks_per_cu<WW, PP, DD>()), dim3(256), \
                       0, ctx->stream, (const ViewDesc *)res.view_table, a)
#define LYAP_LAUNCH2(WW, PP)                                                                         \
    do {                                                                                             \
        if (a.rounds == 1u) LYAP_LAUNCH(WW, PP, false);                                              \
        else LYAP_LAUNCH(WW, PP, true);                                                              \
    } while (0)
        if (W == 16) {
            if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH2(16, LYAP_PREC_PLAIN);
            else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH2(16, LYAP_PREC_PRODUCT);
            else LYAP_LAUNCH2(16, LYAP_PREC_FP64);
        } else {
            if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH2(32, LYAP_PREC_PLAIN);
            else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH2(32, LYAP_PREC_PRODUCT);
            else LYAP_LAUNCH2(32, LYAP_PREC_FP64);
        }
#undef LYAP_LAUNCH2
#undef LYAP_LAUNCH
        LYAP_HIP(ctx, hipGetLastError());
        rc = timing_mark(ctx);
        if (rc) return rc;
    }
    return LYAP_OK;
}

/* Enqueues the upload and the level launches on the context's stream. No
 * call here synchronises, so it may be recorded into a graph. */
static int enqueue_views(lyap_ctx *ctx, RenderResources &res, const ViewDesc *views,
                         unsigned n_views, uint32_t total_tiles, const LaunchPlan &plan,
                         uint32_t W, int precision)
{
    LYAP_HIP(ctx, hipMemcpyAsync(res.view_table, res.staging[res.stage], sizeof(ViewDesc) * n_views,
                                 hipMemcpyHostToDevice, ctx->stream));
    /* the frame's counters, cleared by one kernel (a graph may hold no
     * memset nodes) */
    const bool park = parks(plan, views);
    ZeroRegions z{};
    z.p[0] = res.counters;
    z.n[0] = kCounterWords;
    const PassPlan pp = pass_plan(plan, views, n_views);
    z.p[1] = (uint32_t *)res.cont_counts;
    z.n[1] = pp.n ? (pp.n + 1u) * n_views * kContSubs * kHeadStride : 3u * n_views;
    z.p[2] = park ? (uint32_t *)res.refine_counts : nullptr;
    z.n[2] = park ? n_views * kRefineSubs * kHeadStride : 0u;
    z.p[3] = (uint32_t *)res.vox_counts;
    z.n[3] = n_views * kHeadStride;
    const uint32_t zmax = std::max(std::max(z.n[0], z.n[1]), std::max(z.n[2], z.n[3]));
    hipLaunchKernelGGL(zero_regions, dim3((zmax + 255u) / 256u, 4), dim3(256), 0, ctx->stream, z);
    LYAP_HIP(ctx, hipGetLastError());
    int rc = LYAP_OK;
    uint64_t cap_total = 0, list_total = 0, lanes0 = 0;
    bool has_rays = false;
    for (unsigned i = 0; i < n_views; ++i) {
        cap_total += views[i].cont_cap;
        list_total += views[i].vlist_cap;
        if (views[i].kind == kItemView) {
            has_rays = true;
            lanes0 += (uint64_t)views[i].n_tiles * plan.tile_rays * plan.lanes_per_ray;
        } else {
            lanes0 += (uint64_t)views[i].n_tiles * 64u;
        }
    }
    ContEntry *bufA = (ContEntry *)res.cont_buf;
    ContEntry *bufB = bufA + cap_total;
    uint32_t *counts = (uint32_t *)res.cont_counts;
    bool margin = false;
    for (unsigned i = 0; i < n_views; ++i)
        if (views[i].margin > 0.0f && views[i].refine == LYAP_REFINE_BISECT) margin = true;
    const uint32_t cus = ctx->num_cus > 0 ? (uint32_t)ctx->num_cus : 256u;
    (void)total_tiles;
    const uint64_t needed0 = (lanes0 + 255u) / 256u;
    /* level l reads the rays level l-1 handed over (buffers alternate) */
    ContEntry *out_buf[4] = {bufA, bufB, bufA, nullptr};
    const ContEntry *in_buf[4] = {nullptr, bufA, bufB, bufA};
    /* volume blocks finish at level 0; later levels carry rays only */
    uint32_t levels = (plan.direct || !has_rays) ? 1u : 4u;
    uint32_t first_level = 0;
    if (pp.n > 0) {
        /* the passes march every ray to its end: no level launch follows */
        rc = enqueue_passes(ctx, res, views, n_views, pp, W, precision);
        if (rc) return rc;
        first_level = levels;
    }
    for (uint32_t level = first_level; level < levels; ++level) {
        RenderArgs a{};
        a.n_views = n_views;
        a.total_tiles = total_tiles;
        a.tile_rays = plan.tile_rays;
        a.lanes_per_ray = plan.lanes_per_ray;
        a.slots = plan.slots;
        a.tile_heads = res.counters + kHeadStride + level * kTileHeads * kHeadStride;
        a.eval_counter = (unsigned long long *)(res.counters + 2);
        a.wave_log = (unsigned long long *)ctx->debug_wave_log;
        a.level = level;
        a.budget = level_budget(level, plan);
        a.direct = plan.direct ? 1u : 0u;
        a.resident_lanes = plan.resident_lanes;
        a.cont_in = in_buf[level];
        a.in_counts = level ? counts + (level - 1) * n_views : nullptr;
        a.cont_out = out_buf[level];
        a.out_counts = level < 3 ? counts + level * n_views : nullptr;
        a.refine_out = park ? (RefineRec *)res.refine_buf : nullptr;
        a.refine_counts = park ? (uint32_t *)res.refine_counts : nullptr;
        a.vox_list = (uint32_t *)res.vox_list;
        a.vox_counts = (uint32_t *)res.vox_counts;
        uint32_t blocks = plan.blocks;
        if (level == 0 && blocks > needed0) blocks = (uint32_t)needed0;
        if (blocks == 0) blocks = 1;
        if (blocks > cus * 8u) blocks = cus * 8u;
        rc = timing_mark(ctx);
        if (rc) return rc;
#define LYAP_LAUNCH(WW, PP, MM)                                                       \
    hipLaunchKernelGGL((render_persistent<WW, PP, MM>), dim3(blocks), dim3(256), 0, ctx->stream, \
                       (const ViewDesc *)res.view_table, a)
        if (W == 16) {
            if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH(16, LYAP_PREC_PLAIN, false);
            else if (precision == LYAP_PREC_PRODUCT && margin) LYAP_LAUNCH(16, LYAP_PREC_PRODUCT, true);
            else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH(16, LYAP_PREC_PRODUCT, false);
            else LYAP_LAUNCH(16, LYAP_PREC_FP64, false);
        } else {
            if (precision == LYAP_PREC_PLAIN) LYAP_LAUNCH(32, LYAP_PREC_PLAIN, false);
            else if (precision == LYAP_PREC_PRODUCT && margin) LYAP_LAUNCH(32, LYAP_PREC_PRODUCT, true);
            else if (precision == LYAP_PREC_PRODUCT) LYAP_LAUNCH(32, LYAP_PREC_PRODUCT, false);
            else LYAP_LAUNCH(32, LYAP_PREC_FP64, false);
        }
#undef LYAP_LAUNCH
        LYAP_HIP(ctx, hipGetLastError());
        rc = timing_mark(ctx);
        if (rc) return rc;
        if (level == 0 && list_total > 0) {
            /* FP64 re-evaluation of the voxels the blocks listed */
            VoxelFp64Args va{};
            va.n_items = n_views;
            va.list = (const uint32_t *)res.vox_list;
            va.counts = (const uint32_t *)res.vox_counts;
            const uint32_t vb = (uint32_t)std::min<uint64_t>(cus * 8u, (list_total + 255u) / 256u);
            if (W == 16)
                hipLaunchKernelGGL(voxel_fp64<16>, dim3(vb), dim3(256), 0, ctx->stream,
                                   (const ViewDesc *)res.view_table, va);
            else
                hipLaunchKernelGGL(voxel_fp64<32>, dim3(vb), dim3(256), 0, ctx->stream,
                                   (const ViewDesc *)res.view_table, va);
            LYAP_HIP(ctx, hipGetLastError());
        }
    }
    if (park && has_rays) {
        rc = timing_mark(ctx);
        if (rc) return rc;
        rc = launch_refine(ctx, res, n_views, W, precision, margin);
        if (rc) return rc;
        rc = timing_mark(ctx);
        if (rc) return rc;
    }
    return LYAP_OK;
}

static RenderResources &ctx_resources(lyap_ctx *ctx)
{
    if (!ctx->render_res) ctx->render_res = new RenderResources();
    return *static_cast<RenderResources *>(ctx->render_res);
}

void free_render_resources(lyap_ctx *ctx)
{
    if (!ctx->render_res) return;
    RenderResources *r = static_cast<RenderResources *>(ctx->render_res);
    release_resources(ctx, *r);
    delete r;
    ctx->render_res = nullptr;
}

/* ---- validation shared by the entries ------------------------------------ */

static int check_mode(lyap_ctx *ctx, const LyapRenderMode *mode_in, LyapRenderMode &mode)
{
    mode = LyapRenderMode{};
    mode.refine = LYAP_REFINE_WALK;
    mode.precision = LYAP_PREC_PLAIN;
    if (mode_in) mode = *mode_in;
    if (mode.refine != LYAP_REFINE_WALK && mode.refine != LYAP_REFINE_BISECT)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown refine mode %d", mode.refine);
    if (mode.precision < LYAP_PREC_PLAIN || mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision %d", mode.precision);
    if (mode.bisect_steps < 0 || mode.bisect_steps > 64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "bisect_steps outside 0..64");
    if (mode.lanes_per_ray > 64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "lanes_per_ray above 64");
    if (mode.schedule != LYAP_SCHEDULE_BATCHED && mode.schedule != LYAP_SCHEDULE_DIRECT &&
        mode.schedule != LYAP_SCHEDULE_LEVELS)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown schedule %d", mode.schedule);
    return LYAP_OK;
}

static int view_from_api(lyap_ctx *ctx, ViewDesc &v, const LyapView &in, const LyapRenderMode &mode)
{
    if (in.width == 0 || in.height == 0) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "empty view");
    if ((uint64_t)in.width * in.height > (1ull << 31))
        return fail(ctx, LYAP_ERR_UNSUPPORTED, "view larger than 2^31 pixels");
    if (!in.rgba) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null rgba output");
    if (in.num_lights > LYAP_MAX_LIGHTS) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "more than %d lights", LYAP_MAX_LIGHTS);
    if (in.num_lights > 0 && !in.lights) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null lights");
    EvalDesc e;
    if (!make_eval_desc(&e, in.prm, in.seq, 16))
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence must hold 1..%d letter indices 0..3 before the -1", LYAP_MAX_SEQ);
    fill_view(v, e, in.cam, in.prm, in.width, in.height, in.rgba, in.points, in.depth,
              in.lambda, in.evals, in.lights, in.num_lights, mode);
    return LYAP_OK;
}

/* Builds the launch of n views: validated descriptors ordered by the cost
 * of one of their rays (iterations per sample), then by size, most costly
 * first: the end of the queue then holds the cheapest rays, and no costly
 * ray starts when the others are done (on 64 mixed views this takes 12% off
 * the batch against ordering by whole-view cost). */
static int build_batch(lyap_ctx *ctx, const LyapView *in, unsigned n, const LyapRenderMode &mode,
                       std::vector<ViewDesc> &views, uint32_t &W)
{
    views.resize(n);
    std::vector<std::pair<double, unsigned>> order(n);
    uint32_t maxL = 1;
    for (unsigned i = 0; i < n; ++i) {
        ViewDesc v;
        int rc = view_from_api(ctx, v, in[i], mode);
        if (rc) return rc;
        if ((uint32_t)v.e.L > maxL) maxL = (uint32_t)v.e.L;
        order[i] = {-(double)(v.prm.settle + 2u * v.prm.accum) * 1e9 - (double)v.n_rays, i};
        views[i] = v;
    }
    std::stable_sort(order.begin(), order.end());
    std::vector<ViewDesc> sorted(n);
    for (unsigned i = 0; i < n; ++i) sorted[i] = views[order[i].second];
    views.swap(sorted);
    W = maxL <= 16 ? 16u : 32u;
    return LYAP_OK;
}

/* Appends the validated volume blocks to a launch's items (after its views)
 * and widens the coefficient window if a block's sequence needs it. */
static int append_blocks(lyap_ctx *ctx, const LyapVolumeBlock *in, unsigned n,
                         const LyapVolumeMode &vmode, std::vector<ViewDesc> &items, uint32_t &W)
{
    for (unsigned i = 0; i < n; ++i) {
        const LyapVolumeBlock &b = in[i];
        const uint32_t z1 = b.z_end ? b.z_end : b.n[2];
        if (b.n[0] == 0 || b.n[1] == 0 || b.n[2] == 0 || b.z_begin >= z1 || z1 > b.n[2])
            return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "volume block %u: empty grid or bad z range", i);
        if ((uint64_t)b.n[0] * b.n[1] * (z1 - b.z_begin) > 0xFFFFFFFFull)
            return fail(ctx, LYAP_ERR_UNSUPPORTED, "volume block %u: more than 2^32-1 voxels", i);
        if (!b.out) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "volume block %u: null output", i);
        EvalDesc e;
        if (!make_eval_desc(&e, b.prm, b.seq, vmode.renorm))
            return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "volume block %u: sequence must hold 1..%d letter "
                        "indices 0..3 before the -1", i, LYAP_MAX_SEQ);
        ViewDesc v;
        fill_block(v, e, b.prm, b.lo, b.hi, b.n, b.z_begin, z1, b.out, vmode);
        items.push_back(v);
        if (e.L > 16) W = 32;
    }
    return LYAP_OK;
}

static int render_views(lyap_ctx *ctx, std::vector<ViewDesc> &views, uint32_t W,
                        const LyapRenderMode &mode, const LyapTuning &tune)
{
    uint64_t rays = 0;
    for (auto &v : views)
        if (v.kind == kItemView) rays += v.n_rays;
    const bool margin = mode.fp64_margin > 0.0f && mode.refine == LYAP_REFINE_BISECT;
    LyapTuning t = tune;
    if (!t.lanes_per_ray && mode.lanes_per_ray) t.lanes_per_ray = mode.lanes_per_ray;
    const LaunchPlan plan = plan_launch(ctx, rays, t, W, mode.precision, margin,
                                        mode.schedule == LYAP_SCHEDULE_DIRECT,
                                        mode.schedule == LYAP_SCHEDULE_LEVELS);
    RenderResources &res = ctx_resources(ctx);
    uint32_t total_tiles = 0;
    int rc = prepare_views(ctx, res, views.data(), (unsigned)views.size(), plan, W, true, &total_tiles);
    if (rc) return rc;
    rc = enqueue_views(ctx, res, views.data(), (unsigned)views.size(), total_tiles, plan, W, mode.precision);
    if (rc) return rc;
    /* lyap_last_evals reads this frame's counters (a capture, which
     * enqueues without running, leaves the previous call's in place) */
    ctx->last_counters = res.counters;
    LYAP_HIP(ctx, hipEventRecord(res.staged[res.stage], ctx->stream));
    return LYAP_OK;
}

static LyapTuning merged_tuning(const lyap_ctx *ctx, const LyapTuning *t)
{
    LyapTuning m = ctx->tuned;
    if (t) {
        if (t->tile_rays) m.tile_rays = t->tile_rays;
        if (t->waves_per_cu) m.waves_per_cu = t->waves_per_cu;
        if (t->lanes_per_ray) m.lanes_per_ray = t->lanes_per_ray;
        if (t->slots) m.slots = t->slots;
        for (int k = 0; k < 3; ++k)
            if (t->level_budget[k]) m.level_budget[k] = t->level_budget[k];
        if (t->voxel_tile) m.voxel_tile = t->voxel_tile;
    }
    return m;
}

int launch_render(lyap_ctx *ctx, LyapRGBA *rgba, LyapPoint *points,
                  const LyapCam &cam, const LyapParams &prm, const EvalDesc &e,
                  const LyapLight *lights, unsigned num_lights, unsigned w,
                  unsigned h, float *depth, float *lambda, uint32_t *evals,
                  const LyapRenderMode &mode)
{
    std::vector<ViewDesc> views(1);
    fill_view(views[0], e, cam, prm, w, h, rgba, points, depth, lambda, evals, lights,
              num_lights, mode);
    const uint32_t W = e.L <= 16 ? 16u : 32u;
    return render_views(ctx, views, W, mode, merged_tuning(ctx, nullptr));
}

/* ---- launch slots for the server ----------------------------------------- */

struct RenderSlot {
    RenderResources res;
};

RenderSlot *render_slot_create() { return new RenderSlot(); }

void render_slot_destroy(lyap_ctx *ctx, RenderSlot *slot)
{
    if (!slot) return;
    if (ctx->last_counters == slot->res.counters) ctx->last_counters = nullptr;
    release_resources(ctx, slot->res);
    delete slot;
}

const uint32_t *render_slot_counters(const RenderSlot *slot) { return slot->res.counters; }

const uint32_t *render_slot_vox_counts(const RenderSlot *slot)
{
    return static_cast<const uint32_t *>(slot->res.vox_counts);
}

static_assert(kItemCountStride == kHeadStride, "per-item counter spacing");

/* Renders rays [ray_begin[i], ray_begin[i] + ray_count[i]) of views[i] for
 * every i, in that order of priority, then the volume blocks (items n ..
 * n + nb - 1 of the launch), with the slot's own resources. The blocks must
 * use the views' precision (or there are no views). The caller guarantees
 * the slot's previous launch has finished, so nothing here waits on the
 * device. ray_begin is a multiple of 64 (whole 8x8 blocks). */
int render_slot_launch(lyap_ctx *ctx, RenderSlot *slot, const LyapView *in,
                       const uint32_t *ray_begin, const uint32_t *ray_count, unsigned n,
                       const LyapVolumeBlock *blocks, unsigned nb, const LyapVolumeMode &vmode,
                       const LyapRenderMode &mode_in, const LyapTuning &tune)
{
    LyapRenderMode mode = mode_in;
    if (n == 0) mode.precision = vmode.precision;
    if (nb && mode.precision != vmode.precision)
        return fail(ctx, LYAP_ERR_UNSUPPORTED, "volume blocks need the views' precision in one launch");
    std::vector<ViewDesc> views(n);
    uint32_t maxL = 1;
    uint64_t rays = 0;
    for (unsigned i = 0; i < n; ++i) {
        int rc = view_from_api(ctx, views[i], in[i], mode);
        if (rc) return rc;
        views[i].ray_begin = ray_begin[i];
        views[i].n_rays = ray_count[i];
        rays += ray_count[i];
        if ((uint32_t)views[i].e.L > maxL) maxL = (uint32_t)views[i].e.L;
    }
    uint32_t W = maxL <= 16 ? 16u : 32u;
    int rc = append_blocks(ctx, blocks, nb, vmode, views, W);
    if (rc) return rc;
    const bool margin = mode.fp64_margin > 0.0f && mode.refine == LYAP_REFINE_BISECT;
    LyapTuning t = tune;
    if (!t.lanes_per_ray && mode.lanes_per_ray) t.lanes_per_ray = mode.lanes_per_ray;
    const LaunchPlan plan = plan_launch(ctx, rays, t, W, mode.precision, margin,
                                        mode.schedule == LYAP_SCHEDULE_DIRECT,
                                        mode.schedule == LYAP_SCHEDULE_LEVELS);
    uint32_t total_tiles = 0;
    const unsigned items = (unsigned)views.size();
    rc = prepare_views(ctx, slot->res, views.data(), items, plan, W, false, &total_tiles);
    if (rc) return rc;
    rc = enqueue_views(ctx, slot->res, views.data(), items, total_tiles, plan, W, mode.precision);
    if (rc) return rc;
    ctx->last_counters = slot->res.counters;
    return LYAP_OK;
}

} /* namespace lyap */

using namespace lyap;

/* A captured frame: its own resources and the instantiated graph. */
struct lyap_frame {
    lyap_ctx *ctx = nullptr;
    RenderResources res;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

/* Diagnostic: device buffer of 4 x uint64 per wave that the next render
 * launches fill with {rounds, rays started, start clock, end clock}; NULL
 * turns it off. */
extern "C" int lyap_debug_wave_log(lyap_ctx *ctx, void *d_log)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    ctx->debug_wave_log = d_log;
    return LYAP_OK;
}

extern "C" int lyap_calc_render(lyap_ctx *ctx, LyapRGBA *d_rgba, LyapPoint *d_points,
                                const LyapCam *cam, const LyapParams *prm,
                                const int32_t *seq, const LyapLight *d_lights,
                                unsigned num_lights, unsigned w, unsigned h,
                                float *d_depth, float *d_lambda, uint32_t *d_evals,
                                const LyapRenderMode *mode_in)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (!cam || !prm || !seq) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null camera, parameters or sequence");
    if ((uint64_t)w * h == 0) return LYAP_OK;
    if (!d_rgba) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null rgba output");
    if (num_lights > LYAP_MAX_LIGHTS) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "more than %d lights", LYAP_MAX_LIGHTS);
    if (num_lights > 0 && !d_lights) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null lights");
    if ((uint64_t)w * h > (1ull << 31)) return fail(ctx, LYAP_ERR_UNSUPPORTED, "image larger than 2^31 pixels");
    LyapRenderMode mode;
    int rc = check_mode(ctx, mode_in, mode);
    if (rc) return rc;
    EvalDesc e;
    if (!make_eval_desc(&e, *prm, seq, 16))
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence must hold 1..%d letter indices 0..3 before the -1", LYAP_MAX_SEQ);
    return launch_render(ctx, d_rgba, d_points, *cam, *prm, e, d_lights, num_lights,
                         w, h, d_depth, d_lambda, d_evals, mode);
}

namespace lyap {
namespace {

/* Shading of stored hit records (kernel.cu:25-106, 511-514): one lane per
 * pixel, the same shade and RGBA conversion the render applies when a ray
 * ends, so a frame re-shaded for new lights is byte-identical to one
 * rendered with them. A miss record (a = NaN) stays transparent black. */
__global__ __launch_bounds__(256) void shade_points(LyapRGBA *__restrict__ rgba,
                                                    const LyapPoint *__restrict__ points,
                                                    LyapCam cam, const LyapLight *lights,
                                                    uint32_t num_lights, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const LyapPoint p = points[i];
    const C4 cc = shade(f3(p.P), f3(p.N), p.a, p.c, cam, lights, num_lights);
    LyapRGBA px;
    px.r = to_u8(cc.x);
    px.g = to_u8(cc.y);
    px.b = to_u8(cc.z);
    px.a = to_u8(cc.w);
    rgba[i] = px;
}

} /* namespace */
} /* namespace lyap */

extern "C" int lyap_shade_points(lyap_ctx *ctx, LyapRGBA *d_rgba, const LyapPoint *d_points,
                                 const LyapCam *cam, const LyapLight *d_lights,
                                 unsigned num_lights, unsigned w, unsigned h)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (!cam) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null camera");
    const uint64_t n = (uint64_t)w * h;
    if (n == 0) return LYAP_OK;
    if (!d_rgba || !d_points) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null rgba or points");
    if (num_lights > LYAP_MAX_LIGHTS) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "more than %d lights", LYAP_MAX_LIGHTS);
    if (num_lights > 0 && !d_lights) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null lights");
    if (n > (1ull << 31)) return fail(ctx, LYAP_ERR_UNSUPPORTED, "image larger than 2^31 pixels");
    hipLaunchKernelGGL(shade_points, dim3((uint32_t)((n + 255u) / 256u)), dim3(256), 0, ctx->stream,
                       d_rgba, d_points, *cam, d_lights, (uint32_t)num_lights, (uint32_t)n);
    LYAP_HIP(ctx, hipGetLastError());
    return LYAP_OK;
}

extern "C" int lyap_render_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                                 const LyapRenderMode *mode_in, const LyapTuning *tuning)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (n_views == 0) return LYAP_OK;
    if (!views) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null views");
    LyapRenderMode mode;
    int rc = check_mode(ctx, mode_in, mode);
    if (rc) return rc;
    std::vector<ViewDesc> vd;
    uint32_t W;
    rc = build_batch(ctx, views, n_views, mode, vd, W);
    if (rc) return rc;
    return render_views(ctx, vd, W, mode, merged_tuning(ctx, tuning));
}

namespace lyap {

/* Validates a volume mode (NULL: plain FP32, no re-evaluation). */
static int check_volume_mode(lyap_ctx *ctx, const LyapVolumeMode *in, LyapVolumeMode &m)
{
    m = LyapVolumeMode{};
    if (in) m = *in;
    if (m.precision < LYAP_PREC_PLAIN || m.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown volume precision %d", m.precision);
    return LYAP_OK;
}

/* The items of a mixed launch (views sorted, then blocks) and the render
 * mode the launch runs in: the views' precision, or the blocks' when there
 * are no views. `fused` is false when the blocks need another precision than
 * the views; they are then left out of `items`. */
static int build_mixed(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                       const LyapVolumeBlock *blocks, unsigned n_blocks, const LyapRenderMode &mode,
                       const LyapVolumeMode &vmode, std::vector<ViewDesc> &items, uint32_t &W,
                       LyapRenderMode &launch_mode, bool &fused)
{
    int rc = build_batch(ctx, views, n_views, mode, items, W);
    if (rc) return rc;
    launch_mode = mode;
    fused = n_blocks == 0 || n_views == 0 || vmode.precision == mode.precision;
    if (n_views == 0) launch_mode.precision = vmode.precision;
    if (fused) return append_blocks(ctx, blocks, n_blocks, vmode, items, W);
    return LYAP_OK;
}

/* The blocks of a batch whose precision differs from the views': through
 * the volume kernels, on the context's stream after the views. */
static int run_blocks_apart(lyap_ctx *ctx, const LyapVolumeBlock *blocks, unsigned n,
                            const LyapVolumeMode &vmode)
{
    for (unsigned i = 0; i < n; ++i) {
        const LyapVolumeBlock &b = blocks[i];
        EvalDesc e;
        if (!make_eval_desc(&e, b.prm, b.seq, vmode.renorm))
            return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "volume block %u: bad sequence", i);
        const uint32_t z1 = b.z_end ? b.z_end : b.n[2];
        int rc = launch_volume_slab(ctx, b.out, e, b.lo, b.hi, b.n[0], b.n[1], b.n[2], b.z_begin, z1,
                                    nullptr, 0, vmode);
        if (rc) return rc;
    }
    return LYAP_OK;
}

} /* namespace lyap */

extern "C" int lyap_run_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                              const LyapVolumeBlock *blocks, unsigned n_blocks,
                              const LyapRenderMode *mode_in, const LyapVolumeMode *vol_mode_in,
                              const LyapTuning *tuning)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (n_views + n_blocks == 0) return LYAP_OK;
    if ((n_views && !views) || (n_blocks && !blocks))
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "null views or blocks");
    LyapRenderMode mode;
    int rc = check_mode(ctx, mode_in, mode);
    if (rc) return rc;
    LyapVolumeMode vmode;
    rc = check_volume_mode(ctx, vol_mode_in, vmode);
    if (rc) return rc;
    std::vector<ViewDesc> items;
    uint32_t W;
    LyapRenderMode lmode;
    bool fused;
    rc = build_mixed(ctx, views, n_views, blocks, n_blocks, mode, vmode, items, W, lmode, fused);
    if (rc) return rc;
    rc = render_views(ctx, items, W, lmode, merged_tuning(ctx, tuning));
    if (rc) return rc;
    if (!fused) return run_blocks_apart(ctx, blocks, n_blocks, vmode);
    return LYAP_OK;
}

namespace lyap {

/* Milliseconds of `reps` runs of the launch sequence (after one untimed). */
static int time_items(lyap_ctx *ctx, const std::vector<ViewDesc> &items, uint32_t W,
                      const LyapRenderMode &mode, const LyapTuning &t, hipEvent_t e0, hipEvent_t e1,
                      int reps, float &ms)
{
    std::vector<ViewDesc> trial = items;
    int rc = render_views(ctx, trial, W, mode, t);
    if (rc) return rc;
    LYAP_HIP(ctx, hipEventRecord(e0, ctx->stream));
    for (int rep = 0; rep < reps; ++rep) {
        trial = items;
        rc = render_views(ctx, trial, W, mode, t);
        if (rc) return rc;
    }
    LYAP_HIP(ctx, hipEventRecord(e1, ctx->stream));
    LYAP_HIP(ctx, hipEventSynchronize(e1));
    LYAP_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    return LYAP_OK;
}

/* The staged search of lyap_autotune_batch: launch shape, then the level
 * budgets one level at a time, then the voxel tile. */
static int autotune_items(lyap_ctx *ctx, const std::vector<ViewDesc> &items, uint32_t W,
                          const LyapRenderMode &mode, bool has_rays, bool has_blocks, LyapTuning &best_t)
{
    hipEvent_t e0, e1;
    LYAP_HIP(ctx, hipEventCreate(&e0));
    LYAP_HIP(ctx, hipEventCreate(&e1));
    float best_ms = 1e30f;
    int rc = LYAP_OK;
    auto trial = [&](const LyapTuning &t) {
        float ms = 0.f;
        if (rc == LYAP_OK) rc = time_items(ctx, items, W, mode, t, e0, e1, 2, ms);
        if (rc == LYAP_OK && ms < best_ms) {
            best_ms = ms;
            best_t = t;
        }
    };
    best_t = LyapTuning{};
    trial(best_t);                                  /* the built-in shape */
    if (has_rays) {
        /* samples per ray and round K x S from 1 (one lane per ray) to 16,
         * tiles of one or two generations of a wave's groups */
        const uint32_t shapes[][2] = {{1, 1}, {2, 1}, {4, 1}, {8, 1}, {16, 1}, {1, 2}, {2, 2}, {4, 2}, {1, 4}};
        const LyapTuning base = best_t;
        for (const auto &sh : shapes) {
            for (uint32_t g : {1u, 2u}) {
                LyapTuning t = base;
                t.lanes_per_ray = sh[0];
                t.slots = sh[1];
                t.tile_rays = g * (64u / sh[0]);
                trial(t);
            }
        }
        /* march-sample budgets of levels 0, 1 and 2, one level at a time */
        const uint32_t cand[3][4] = {{16, 24, 32, 64}, {32, 48, 64, 96}, {96, 128, 192, 256}};
        for (int lv = 0; lv < 3 && rc == LYAP_OK; ++lv) {
            const LyapTuning base2 = best_t;
            for (uint32_t b : cand[lv]) {
                LyapTuning t = base2;
                t.level_budget[lv] = b;
                trial(t);
            }
        }
    }
    if (has_blocks) {
        const LyapTuning base = best_t;
        for (uint32_t vt : {64u, 128u, 512u, 1024u}) {
            LyapTuning t = base;
            t.voxel_tile = vt;
            trial(t);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

} /* namespace lyap */

extern "C" int lyap_autotune_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                                   const LyapVolumeBlock *blocks, unsigned n_blocks,
                                   const LyapRenderMode *mode_in, const LyapVolumeMode *vol_mode_in,
                                   LyapTuning *best)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (n_views + n_blocks == 0 || (n_views && !views) || (n_blocks && !blocks))
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "nothing to tune on");
    LyapRenderMode mode;
    int rc = check_mode(ctx, mode_in, mode);
    if (rc) return rc;
    LyapVolumeMode vmode;
    rc = check_volume_mode(ctx, vol_mode_in, vmode);
    if (rc) return rc;
    std::vector<ViewDesc> items;
    uint32_t W;
    LyapRenderMode lmode;
    bool fused;
    rc = build_mixed(ctx, views, n_views, blocks, n_blocks, mode, vmode, items, W, lmode, fused);
    if (rc) return rc;
    LyapTuning t{};
    rc = autotune_items(ctx, items, W, lmode, n_views > 0, fused && n_blocks > 0, t);
    if (rc) return rc;
    ctx->tuned = t;
    if (best) *best = t;
    return LYAP_OK;
}

extern "C" int lyap_set_tuning(lyap_ctx *ctx, const LyapTuning *tuning)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    ctx->tuned = tuning ? *tuning : LyapTuning{};
    return LYAP_OK;
}

extern "C" int lyap_get_tuning(lyap_ctx *ctx, LyapTuning *out)
{
    if (!ctx || !out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = ctx->tuned;
    return LYAP_OK;
}

extern "C" int lyap_autotune(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                             const LyapRenderMode *mode_in, LyapTuning *best)
{
    if (!ctx) return LYAP_ERR_INVALID_ARGUMENT;
    if (n_views == 0 || !views) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "no views to tune on");
    return lyap_autotune_batch(ctx, views, n_views, nullptr, 0, mode_in, nullptr, best);
}

extern "C" int lyap_frame_capture(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                                  const LyapRenderMode *mode_in, const LyapTuning *tuning,
                                  lyap_frame **out)
{
    if (!ctx || !out) return LYAP_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (n_views == 0 || !views) return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "no views");
    LyapRenderMode mode;
    int rc = check_mode(ctx, mode_in, mode);
    if (rc) return rc;
    std::vector<ViewDesc> vd;
    uint32_t W;
    rc = build_batch(ctx, views, n_views, mode, vd, W);
    if (rc) return rc;
    uint64_t rays = 0;
    for (auto &v : vd) rays += v.n_rays;
    const bool margin = mode.fp64_margin > 0.0f && mode.refine == LYAP_REFINE_BISECT;
    const LaunchPlan plan = plan_launch(ctx, rays, merged_tuning(ctx, tuning), W, mode.precision, margin,
                                        mode.schedule == LYAP_SCHEDULE_DIRECT,
                                        mode.schedule == LYAP_SCHEDULE_LEVELS);
    lyap_frame *f = new lyap_frame();
    f->ctx = ctx;
    uint32_t total_tiles = 0;
    rc = prepare_views(ctx, f->res, vd.data(), n_views, plan, W, true, &total_tiles);
    if (rc == LYAP_OK) {
        /* everything the frame needs exists now; record without syncing */
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            rc = fail(ctx, LYAP_ERR_DEVICE, "stream capture: %s", hipGetErrorString(e));
        } else {
            /* no timing events inside a graph */
            const bool timing = ctx->timing;
            ctx->timing = false;
            rc = enqueue_views(ctx, f->res, vd.data(), n_views, total_tiles, plan, W, mode.precision);
            ctx->timing = timing;
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(ctx->stream, &g);
            if (rc == LYAP_OK && e != hipSuccess)
                rc = fail(ctx, LYAP_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
            f->graph = g;
            if (rc == LYAP_OK) {
                e = hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0);
                if (e != hipSuccess)
                    rc = fail(ctx, LYAP_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
            }
        }
    }
    if (rc != LYAP_OK) {
        lyap_frame_destroy(f);
        return rc;
    }
    *out = f;
    return LYAP_OK;
}

extern "C" int lyap_frame_replay(lyap_ctx *ctx, lyap_frame *f)
{
    if (!ctx || !f || f->ctx != ctx || !f->exec) return LYAP_ERR_INVALID_ARGUMENT;
    LYAP_HIP(ctx, hipGraphLaunch(f->exec, ctx->stream));
    ctx->last_counters = f->res.counters;
    return LYAP_OK;
}

extern "C" void lyap_frame_destroy(lyap_frame *f)
{
    if (!f) return;
    if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    if (f->ctx) {
        /* a later lyap_last_evals must not read the freed counters */
        if (f->ctx->last_counters == f->res.counters) f->ctx->last_counters = nullptr;
        release_resources(f->ctx, f->res);
    }
    delete f;
}
