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
#include <cmath>
#include <cstring>

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
};

struct RenderArgs {
    const ViewDesc *views;
    uint32_t n_views;
    uint32_t total_tiles;
    uint32_t tile_rays;
    uint32_t pad;
    uint32_t *tile_counter;
    unsigned long long *eval_counter;
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

/* ---- lambda dispatch over the sequence period --------------------------- */

#define LYAP_L_SWITCH(EXPR)                                                    \
    switch (e.L) {                                                             \
    case 1: { constexpr int L = 1; EXPR; }   case 2: { constexpr int L = 2; EXPR; }   \
    case 3: { constexpr int L = 3; EXPR; }   case 4: { constexpr int L = 4; EXPR; }   \
    case 5: { constexpr int L = 5; EXPR; }   case 6: { constexpr int L = 6; EXPR; }   \
    case 7: { constexpr int L = 7; EXPR; }   case 8: { constexpr int L = 8; EXPR; }   \
    case 9: { constexpr int L = 9; EXPR; }   case 10: { constexpr int L = 10; EXPR; } \
    case 11: { constexpr int L = 11; EXPR; } case 12: { constexpr int L = 12; EXPR; } \
    case 13: { constexpr int L = 13; EXPR; } case 14: { constexpr int L = 14; EXPR; } \
    case 15: { constexpr int L = 15; EXPR; } case 16: { constexpr int L = 16; EXPR; } \
    case 17: { constexpr int L = 17; EXPR; } case 18: { constexpr int L = 18; EXPR; } \
    case 19: { constexpr int L = 19; EXPR; } case 20: { constexpr int L = 20; EXPR; } \
    case 21: { constexpr int L = 21; EXPR; } case 22: { constexpr int L = 22; EXPR; } \
    case 23: { constexpr int L = 23; EXPR; } case 24: { constexpr int L = 24; EXPR; } \
    case 25: { constexpr int L = 25; EXPR; } case 26: { constexpr int L = 26; EXPR; } \
    case 27: { constexpr int L = 27; EXPR; } case 28: { constexpr int L = 28; EXPR; } \
    case 29: { constexpr int L = 29; EXPR; } case 30: { constexpr int L = 30; EXPR; } \
    case 31: { constexpr int L = 31; EXPR; } default: { constexpr int L = 32; EXPR; } \
    }

__device__ __noinline__ float eval_fp64(const EvalDesc &e, float x, float y, float z)
{
    LYAP_L_SWITCH(return (float)lambda_fp64<L>(x, y, z, e))
}

template <int PREC>
__device__ __forceinline__ float eval_lambda(const EvalDesc &e, float x, float y, float z)
{
    if (PREC == LYAP_PREC_FP64) return eval_fp64(e, x, y, z);
    if (PREC == LYAP_PREC_PLAIN) {
        LYAP_L_SWITCH(return lambda_plain<L>(x, y, z, e))
    } else {
        LYAP_L_SWITCH(return (lambda_product<L, false>(x, y, z, e)).lam)
    }
}

/* Per-lane ray state. */
struct Ray {
    int phase;
    uint32_t pix;          /* pixel index x + y*w */
    uint32_t evals;
    uint32_t index;        /* march sample index */
    uint32_t k;            /* refine or normal step counter */
    bool near_, sign;
    F3 V, P;
    float t, t1, dt, Fdt, Ndt, last_step;
    float a, c, l;
    float lo, hi, l_hi;    /* bisection */
    float Qdt, Qt0, Qt1, minQdt;
    F3 QdV;
    float h;
    float gx, gy, gz;      /* central differences */
};

__device__ __forceinline__ void write_miss(const ViewDesc &v, const Ray &r)
{
    if (v.points) {
        LyapPoint p;
        p.P = {0.f, 0.f, 0.f};
        p.N = {0.f, 0.f, 0.f};
        p.a = NAN;
        p.c = 0.f;
        p.l = NAN;
        v.points[r.pix] = p;
    }
    if (v.depth) v.depth[r.pix] = NAN;
    if (v.lambda) v.lambda[r.pix] = NAN;
    if (v.evals) v.evals[r.pix] = r.evals;
    LyapRGBA z = {0, 0, 0, 0};
    v.rgba[r.pix] = z;
}

__device__ void write_hit(const ViewDesc &v, const Ray &r, F3 N)
{
    if (v.points) {
        LyapPoint p;
        p.P = {r.P.x, r.P.y, r.P.z};
        p.N = {N.x, N.y, N.z};
        p.a = r.a;
        p.c = r.c;
        p.l = r.l;
        v.points[r.pix] = p;
    }
    if (v.depth) v.depth[r.pix] = r.t;
    if (v.lambda) v.lambda[r.pix] = r.l;
    if (v.evals) v.evals[r.pix] = r.evals;
    const C4 cc = shade(r.P, N, r.a, r.c, v.cam, v.lights, v.num_lights);
    LyapRGBA o;
    o.r = to_u8(cc.x);
    o.g = to_u8(cc.y);
    o.b = to_u8(cc.z);
    o.a = to_u8(cc.w);
    v.rgba[r.pix] = o;
}

/* Starts ray `id` of view v; returns false (after writing a miss or nothing
 * for a padding ray) when no lambda evaluation is needed. */
__device__ bool start_ray(const ViewDesc &v, uint32_t id, Ray &r)
{
    const uint32_t blk = id >> 6, in = id & 63u;
    const uint32_t px = (blk % v.blocks_x) * 8u + (in & 7u);
    const uint32_t py = (blk / v.blocks_x) * 8u + (in >> 3);
    if (px >= v.w || py >= v.h) return false;
    r.pix = px + py * v.w;
    r.evals = 0;
    const F3 C = f3(v.cam.C);
    F3 V = add(add(f3(v.cam.S0), mul(f3(v.cam.SDX), (float)px)), mul(f3(v.cam.SDY), (float)py));
    V = normalized(V);
    V = divv(V, v.cam.M);
    r.V = V;
    float t0, t1;
    if (!clip_box(C, V, v.prm.lMin, v.prm.lMax, t0, t1)) {
        write_miss(v, r);
        return false;
    }
    r.t = t0;
    r.t1 = t1;
    r.P = add(C, mul(V, t0));
    r.a = 0.f;
    r.c = 0.f;
    if (v.prm.stepMethod == 1) r.Fdt = (t1 - t0) / v.prm.depth;
    else r.Fdt = sqrtf(mag2(V)) / v.prm.depth;
    r.dt = r.Fdt;
    r.Ndt = r.dt / v.prm.nearMultiplier;
    r.near_ = false;
    r.last_step = r.dt;
    r.index = 0;
    r.phase = PH_MARCH;
    return true;
}

/* Point the lane evaluates this trip (advances the walk first). */
__device__ __forceinline__ F3 sample_point(const ViewDesc &v, Ray &r)
{
    if (r.phase == PH_MARCH) return r.P;
    if (r.phase == PH_REFINE) {
        if (v.refine == LYAP_REFINE_BISECT) {
            const float mid = r.lo + (r.hi - r.lo) * 0.5f;
            return add(f3(v.cam.C), mul(r.V, mid));
        }
        r.t += r.Qdt;
        r.P = add(r.P, r.QdV);
        return r.P;
    }
    /* PH_NORMAL: x-, x+, y-, y+, z-, z+ (P + (-h) equals P - h exactly) */
    F3 s = r.P;
    const float off = (r.k & 1u) ? r.h : -r.h;
    if (r.k < 2) s.x = r.P.x + off;
    else if (r.k < 4) s.y = r.P.y + off;
    else s.z = r.P.z + off;
    return s;
}

__device__ __forceinline__ void enter_normal(const ViewDesc &v, Ray &r)
{
    r.phase = PH_NORMAL;
    r.k = 0;
    r.h = (v.normal_h > 0.0f) ? v.normal_h : r.dt * v.prm.gradient;
}

/* walk loop test (kernel.cu:420) */
__device__ __forceinline__ bool walk_continues(const Ray &r)
{
    return r.k < kMaxWalk && r.t <= r.Qt1 && r.t >= r.Qt0 &&
           (r.Qdt <= -r.minQdt || r.Qdt >= r.minQdt);
}

__device__ __forceinline__ void enter_refine(const ViewDesc &v, Ray &r)
{
    r.phase = PH_REFINE;
    r.k = 0;
    if (v.refine == LYAP_REFINE_BISECT) {
        r.lo = r.t - r.last_step;
        r.hi = r.t;
        r.l_hi = r.l;
        if (v.bisect_steps <= 0) enter_normal(v, r);
    } else {
        r.sign = false;
        r.Qdt = r.dt * -0.5f;
        r.QdV = mul(r.V, r.Qdt);
        r.Qt1 = r.t;
        r.Qt0 = r.t - r.dt;
        r.minQdt = r.dt / v.prm.refine;
        if (!walk_continues(r)) enter_normal(v, r);
    }
}

/* Consumes the lambda of this trip's sample. Returns true when the ray is
 * finished (outputs written). */
__device__ bool advance(const ViewDesc &v, Ray &r, float l)
{
    ++r.evals;
    const LyapParams &prm = v.prm;
    if (r.phase == PH_MARCH) {
        r.l = l;
        if (r.index > 0) {
            if (l > prm.chaosThreshold) r.c += l;
            else if (l > prm.opaqueThreshold) r.a += l;
            if (l <= prm.nearThreshold && !r.near_) { r.near_ = true; r.dt = r.Ndt; }
            else if (l > prm.nearThreshold && r.near_) { r.near_ = false; r.dt = r.Fdt; }
        }
        if (l > prm.opaqueThreshold) {
            float step = r.dt;
            if (prm.jitter != 0.0f) {
                float jit = l - truncf(l);
                if (jit < 0) jit = (float)(1.0 - (double)(jit * prm.jitter));
                else jit = (float)(1.0 + (double)(jit * prm.jitter));
                if (__builtin_isfinite(jit)) step = r.dt * jit;
            }
            r.t += step;
            if (v.refine == LYAP_REFINE_BISECT) r.P = add(f3(v.cam.C), mul(r.V, r.t));
            else r.P = add(r.P, mul(r.V, step));
            r.last_step = step;
            ++r.index;
            if (r.t > r.t1 || r.index >= v.max_samples) {
                write_miss(v, r);
                return true;
            }
            return false;
        }
        enter_refine(v, r);
        return false;
    }
    if (r.phase == PH_REFINE) {
        if (v.refine == LYAP_REFINE_BISECT) {
            const float mid = r.lo + (r.hi - r.lo) * 0.5f;
            if (!(l > prm.opaqueThreshold)) { r.hi = mid; r.l_hi = l; }
            else r.lo = mid;
            if ((int)++r.k >= v.bisect_steps) {
                r.t = r.hi;
                r.P = add(f3(v.cam.C), mul(r.V, r.t));
                r.l = r.l_hi;
                enter_normal(v, r);
            }
            return false;
        }
        r.l = l;
        ++r.k;
        if (l == prm.opaqueThreshold) { enter_normal(v, r); return false; }
        const bool osign = r.sign;
        r.sign = (l < prm.opaqueThreshold) ? false : true;
        if (r.sign != osign) {
            r.Qdt *= -0.5f;
            r.QdV = mul(r.QdV, -0.5f);
        }
        if (!walk_continues(r)) enter_normal(v, r);
        return false;
    }
    /* PH_NORMAL: g = l(+h) - l(-h) per axis, formed as (-l(-h)) + l(+h) */
    const float sl = (r.k & 1u) ? l : -l;
    if (r.k < 2) r.gx = (r.k == 0) ? sl : r.gx + sl;
    else if (r.k < 4) r.gy = (r.k == 2) ? sl : r.gy + sl;
    else r.gz = (r.k == 4) ? sl : r.gz + sl;
    if (++r.k < 6) return false;
    const F3 N = normalized({r.gx, r.gy, r.gz});
    write_hit(v, r, N);
    return true;
}

__device__ __forceinline__ uint32_t uniform(uint32_t x)
{
    return __builtin_amdgcn_readfirstlane(x);
}

/* View owning tile t (views are few; scalar scan). */
__device__ __forceinline__ uint32_t view_of_tile(const RenderArgs &a, uint32_t t)
{
    uint32_t v = 0;
    while (v + 1 < a.n_views && a.views[v + 1].tile_begin <= t) ++v;
    return uniform(v);
}

template <int PREC>
__global__ __launch_bounds__(256) void render_persistent(RenderArgs a)
{
    Ray r;
    r.phase = PH_IDLE;
    uint32_t done_evals = 0;

    uint32_t cur_view = 0xFFFFFFFFu;  /* view the lanes work on */
    uint32_t pos = 0, end = 0;        /* unstarted rays of the current tile */
    uint32_t pend_tile = 0xFFFFFFFFu; /* claimed tile of another view */
    bool queue_empty = false;

    for (;;) {
        /* refill idle lanes */
        for (;;) {
            const bool need = r.phase == PH_IDLE;
            const uint64_t mask = __ballot(need);
            if (mask == 0) break;
            if (pos >= end) {
                if (pend_tile != 0xFFFFFFFFu) {
                    /* switch views only once every lane is idle */
                    if (__ballot(r.phase != PH_IDLE) != 0) break;
                    const uint32_t t = pend_tile;
                    pend_tile = 0xFFFFFFFFu;
                    cur_view = view_of_tile(a, t);
                    const ViewDesc &v = a.views[cur_view];
                    pos = (t - v.tile_begin) * a.tile_rays;
                    end = uniform(min(pos + a.tile_rays, v.n_rays));
                    continue;
                }
                if (queue_empty) break;
                uint32_t t = 0;
                if (__lane_id() == 0) t = atomicAdd(a.tile_counter, 1u);
                t = uniform(__shfl(t, 0));
                if (t >= a.total_tiles) { queue_empty = true; break; }
                const uint32_t tv = view_of_tile(a, t);
                if (tv != cur_view) { pend_tile = t; continue; }
                const ViewDesc &v = a.views[cur_view];
                pos = (t - v.tile_begin) * a.tile_rays;
                end = uniform(min(pos + a.tile_rays, v.n_rays));
                continue;
            }
            const uint32_t avail = end - pos;
            const uint32_t rank = (uint32_t)__popcll(mask & __lanemask_lt());
            const ViewDesc &v = a.views[cur_view];
            if (need && rank < avail) {
                if (!start_ray(v, pos + rank, r)) r.phase = PH_IDLE;
            }
            const uint32_t taken = (uint32_t)__popcll(mask);
            pos += taken < avail ? taken : avail;
        }
        if (__ballot(r.phase != PH_IDLE) == 0) {
            if (queue_empty && pend_tile == 0xFFFFFFFFu) break;
            continue;
        }
        const ViewDesc &v = a.views[cur_view];
        if (r.phase != PH_IDLE) {
            const F3 s = sample_point(v, r);
            float l = eval_lambda<PREC>(v.e, s.x, s.y, s.z);
            if (PREC != LYAP_PREC_FP64 && v.margin > 0.0f && r.phase == PH_REFINE &&
                v.refine == LYAP_REFINE_BISECT &&
                !(__builtin_fabsf(l - v.prm.opaqueThreshold) >= v.margin))
                l = eval_fp64(v.e, s.x, s.y, s.z);
            if (advance(v, r, l)) {
                done_evals += r.evals;
                r.phase = PH_IDLE;
            }
        }
    }
    /* one 64-bit atomic per wave for the evaluation count */
    unsigned long long sum = done_evals;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off);
    if (__lane_id() == 0 && sum) atomicAdd(a.eval_counter, sum);
}

} /* namespace */

/* Fills a ViewDesc from the API arguments. */
static void fill_view(ViewDesc &v, const EvalDesc &e, const LyapCam &cam,
                      const LyapParams &prm, unsigned w, unsigned h,
                      LyapRGBA *rgba, LyapPoint *points, float *depth,
                      float *lambda, uint32_t *evals, const LyapLight *lights,
                      unsigned num_lights, const LyapRenderMode &mode,
                      uint32_t tile_rays, uint32_t tile_begin)
{
    std::memset(&v, 0, sizeof(v));
    v.e = e;
    v.cam = cam;
    v.prm = prm;
    v.w = w;
    v.h = h;
    v.blocks_x = (w + 7) / 8;
    v.n_rays = 64u * v.blocks_x * ((h + 7) / 8);
    v.tile_begin = tile_begin;
    v.n_tiles = (v.n_rays + tile_rays - 1) / tile_rays;
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
}

static int upload_views(lyap_ctx *ctx, const ViewDesc *views, unsigned n)
{
    const size_t bytes = sizeof(ViewDesc) * n;
    if (bytes > ctx->view_table_bytes) {
        LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ctx_free(ctx, ctx->view_table);
        ctx->view_table = nullptr;
        ctx->view_table_bytes = 0;
        int rc = ctx_alloc(ctx, &ctx->view_table, bytes);
        if (rc) return rc;
        ctx->view_table_bytes = bytes;
    }
    if (bytes > ctx->view_staging_bytes) {
        LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->view_staging) (void)hipHostFree(ctx->view_staging);
        ctx->view_staging = nullptr;
        ctx->view_staging_bytes = 0;
        LYAP_HIP(ctx, hipHostMalloc(&ctx->view_staging, bytes, hipHostMallocDefault));
        ctx->view_staging_bytes = bytes;
    }
    /* the staging copy of the previous launch must have been consumed */
    LYAP_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::memcpy(ctx->view_staging, views, bytes);
    LYAP_HIP(ctx, hipMemcpyAsync(ctx->view_table, ctx->view_staging, bytes,
                                 hipMemcpyHostToDevice, ctx->stream));
    return LYAP_OK;
}

int launch_views(lyap_ctx *ctx, ViewDesc *views, unsigned n_views,
                 uint32_t total_tiles, uint32_t tile_rays, uint32_t waves_per_cu,
                 int precision)
{
    int rc = upload_views(ctx, views, n_views);
    if (rc) return rc;
    rc = launch_zero_u32(ctx, ctx->counters, 8);
    if (rc) return rc;
    RenderArgs a{};
    a.views = (const ViewDesc *)ctx->view_table;
    a.n_views = n_views;
    a.total_tiles = total_tiles;
    a.tile_rays = tile_rays;
    a.tile_counter = ctx->counters + 4;
    a.eval_counter = (unsigned long long *)(ctx->counters + 2);
    const uint32_t cus = ctx->num_cus > 0 ? (uint32_t)ctx->num_cus : 256u;
    uint32_t blocks = cus * waves_per_cu / 4u;
    const uint32_t needed = (total_tiles * tile_rays + 255u) / 256u;
    if (blocks > needed) blocks = needed;
    if (blocks == 0) blocks = 1;
    if (precision == LYAP_PREC_PLAIN)
        hipLaunchKernelGGL(render_persistent<LYAP_PREC_PLAIN>, dim3(blocks), dim3(256), 0, ctx->stream, a);
    else if (precision == LYAP_PREC_PRODUCT)
        hipLaunchKernelGGL(render_persistent<LYAP_PREC_PRODUCT>, dim3(blocks), dim3(256), 0, ctx->stream, a);
    else
        hipLaunchKernelGGL(render_persistent<LYAP_PREC_FP64>, dim3(blocks), dim3(256), 0, ctx->stream, a);
    LYAP_HIP(ctx, hipGetLastError());
    return LYAP_OK;
}

int launch_render(lyap_ctx *ctx, LyapRGBA *rgba, LyapPoint *points,
                  const LyapCam &cam, const LyapParams &prm, const EvalDesc &e,
                  const LyapLight *lights, unsigned num_lights, unsigned w,
                  unsigned h, float *depth, float *lambda, uint32_t *evals,
                  const LyapRenderMode &mode)
{
    uint32_t tile_rays = ctx->tuned.tile_rays ? ctx->tuned.tile_rays : 256u;
    uint32_t waves = ctx->tuned.waves_per_cu ? ctx->tuned.waves_per_cu : 32u;
    ViewDesc v;
    fill_view(v, e, cam, prm, w, h, rgba, points, depth, lambda, evals, lights,
              num_lights, mode, tile_rays, 0);
    return launch_views(ctx, &v, 1, v.n_tiles, tile_rays, waves, mode.precision);
}

} /* namespace lyap */

using namespace lyap;

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
    LyapRenderMode mode{};
    mode.refine = LYAP_REFINE_WALK;
    mode.precision = LYAP_PREC_PLAIN;
    if (mode_in) mode = *mode_in;
    if (mode.refine != LYAP_REFINE_WALK && mode.refine != LYAP_REFINE_BISECT)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown refine mode %d", mode.refine);
    if (mode.precision < LYAP_PREC_PLAIN || mode.precision > LYAP_PREC_FP64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "unknown precision %d", mode.precision);
    if (mode.bisect_steps < 0 || mode.bisect_steps > 64)
        return fail(ctx, LYAP_ERR_INVALID_ARGUMENT, "bisect_steps outside 0..64");
    EvalDesc e;
    if (!make_eval_desc(&e, *prm, seq, 16))
        return fail(ctx, LYAP_ERR_BAD_SEQUENCE, "sequence must hold 1..%d letter indices 0..3 before the -1", LYAP_MAX_SEQ);
    return launch_render(ctx, d_rgba, d_points, *cam, *prm, e, d_lights, num_lights,
                         w, h, d_depth, d_lambda, d_evals, mode);
}
