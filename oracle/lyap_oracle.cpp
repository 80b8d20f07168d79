/*
 * lyap_oracle.cpp - CPU restatement of the reference's Lyapunov path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: the parity tests, the
 * smoke check and the bench's CPU-baseline leg load liblyap_oracle.so to
 * compare against or to time. The product (liblyap_amd.so) never links,
 * loads or calls it, and has no CPU fallback.
 *
 * It restates, from the reference sources (read as text; nothing is copied):
 *   lyap4d              kernel.cu:108-154
 *   raymarch            kernel.cu:156-486
 *   shade               kernel.cu:25-106
 *   pointcalc / volume  kernel.cu:488-498, 518-532
 *   to_rgba             color.hpp:169-175 (saturating, see below)
 *   Vec helpers         vec3.hpp:97-154, 190-193
 * in two arithmetic forms:
 *   FP64  - Real = double, the DOUBLE_REALS build (structs.hpp:10-13);
 *   FP32W - Real = float "as written": the unsuffixed 1.0 / 2.0 literals of
 *           kernel.cu:135,143-144 promote the map and the derivative to double
 *           and round to float on assignment; logf per step.
 * Built with -ffp-contract=off so no multiply-add is fused (x86-64 SSE2 double
 * and float arithmetic is IEEE round-to-nearest).
 *
 * Deliberate, documented differences from the reference:
 *  - the box [lMin, lMax]^3 is honoured at run time (the reference hard-codes
 *    0..4 because USE_LMINMAX does not compile, kernel.cu:12-18);
 *  - a miss writes an explicit record (P = N = 0, a = l = NaN, c = 0) instead
 *    of leaving the caller's buffer unwritten (kernel.cu:262-264, 361-363);
 *  - to_rgba saturates to [0, 255] (NaN -> 0) where the reference's
 *    float->unsigned char conversion is undefined;
 *  - the refinement walk stops after 64 steps, the march after 2^20
 *    samples, and a march step that does not change t in float (a grazing
 *    ray whose t1 - t0 is near 0, or depth = inf) ends the ray as a miss:
 *    the reference loops forever in these cases;
 *  - an optional bisection refinement and absolute normal offset (the
 *    configured render method) next to the reference walk.
 *
 * Pinning: tests/test_oracle.py checks these functions against exact rational
 * hand computations, closed forms on the diagonal, permutation invariants,
 * an mpmath brute force, the analytic one-letter raymarch scene, the
 * closed step recurrence of a constant-lambda ray (jitter, near step), the
 * a/c sums of a hit rebuilt sample by sample in march order
 * (kernel.cu:370-375), the shading's cone, chaos and ambient branches, and
 * the reversing walk's halving. Parity unpinned: RGBA of colour channels
 * outside [0, 1] (undefined in the reference's conversion).
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <omp.h>

#include "lyap_types.h"
#include "lyap.h"

namespace {

constexpr int kModeF64 = 0;
constexpr int kModeF32W = 1;
/* march samples per ray before the ray is given up (the reference has no
 * bound; a zero or negative step would never leave the box) */
constexpr int kMaxSamples = 1 << 20;

/* ---- lyap4d (kernel.cu:108-154) --------------------------------------- */

/* FP64: every quantity double. */
double lyap_f64(double x, double y, double z, double d, unsigned settle,
                unsigned accum, const int32_t *seq)
{
    const double coord[4] = {x, y, z, d};
    unsigned pos = 0;
    double v = 0.5;
    double sum = 0.0;
    for (unsigned n = 0; n < settle; ++n) {
        const double r = coord[seq[pos]];
        ++pos;
        if (seq[pos] == -1) pos = 0;
        v = r * v * (1.0 - v);
    }
    const double off = v - 0.5;
    if (off <= -1e-8 || off >= 1e-8) {
        for (unsigned n = 0; n < accum; ++n) {
            const double r = coord[seq[pos]];
            ++pos;
            if (seq[pos] == -1) pos = 0;
            v = r * v * (1.0 - v);
            double g = r - 2.0 * r * v;   /* derivative with the new iterate */
            if (g < 0) g = -g;
            sum += std::log(g);
            if (!std::isfinite(sum)) return NAN;
        }
    }
    return sum / (double)accum;
}

/* FP32 as written: float storage, double temporaries from the literals. */
float lyap_f32w(float x, float y, float z, float d, unsigned settle,
                unsigned accum, const int32_t *seq)
{
    const float coord[4] = {x, y, z, d};
    unsigned pos = 0;
    float v = 0.5f;
    float sum = 0.0f;
    for (unsigned n = 0; n < settle; ++n) {
        const float r = coord[seq[pos]];
        ++pos;
        if (seq[pos] == -1) pos = 0;
        v = (float)((double)(r * v) * (1.0 - (double)v));
    }
    const double off = (double)v - 0.5;
    if (off <= -1e-8 || off >= 1e-8) {
        for (unsigned n = 0; n < accum; ++n) {
            const float r = coord[seq[pos]];
            ++pos;
            if (seq[pos] == -1) pos = 0;
            v = (float)((double)(r * v) * (1.0 - (double)v));
            float g = (float)((double)r - 2.0 * (double)r * (double)v);
            if (g < 0) g = -g;
            sum += logf(g);
            if (!std::isfinite(sum)) return NAN;
        }
    }
    return sum / (float)accum;
}

/* lambda at a float point in the chosen form, widened to double. */
struct Lambda {
    int mode;
    float d;
    unsigned settle, accum;
    const int32_t *seq;
    double operator()(float x, float y, float z) const {
        if (mode == kModeF64)
            return lyap_f64(x, y, z, d, settle, accum, seq);
        return (double)lyap_f32w(x, y, z, d, settle, accum, seq);
    }
};

/* ---- float Vec helpers (vec3.hpp) -------------------------------------- */

struct V3 { float x, y, z; };
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 divv(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float mag2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
/* normalize (vec3.hpp:141-154): |v|^2 < 1e-12 gives 0, |v|^2 within 1e-12 of
 * 1 is left alone (compared in double, so only exactly 1.0f in float). */
inline V3 normalized(V3 a)
{
    const float m2 = mag2(a);
    if ((double)m2 < 1e-12) return {0.f, 0.f, 0.f};
    if (m2 == 1.0f || ((double)m2 > 1.0 - 1e-12 && (double)m2 < 1.0 + 1e-12))
        return a;
    return divv(a, std::sqrt(m2));
}
inline float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }
inline V3 vec(const LyapVec &v) { return {v.x, v.y, v.z}; }
inline LyapVec lvec(V3 v) { return {v.x, v.y, v.z}; }

struct C4 { float x, y, z, w; };
inline C4 col(const LyapColor &c) { return {c.x, c.y, c.z, c.w}; }
inline C4 cadd(C4 a, C4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline C4 cmul(C4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }

/* ---- shade (kernel.cu:25-106) ------------------------------------------ */

C4 shade(const LyapPoint &pt, const LyapCam &cam, const LyapLight *lights,
         unsigned n_lights)
{
    C4 color = {0, 0, 0, 0};
    if (std::isnan(pt.a)) return color;
    const V3 P = vec(pt.P), N = vec(pt.N);
    const float c = pt.c;
    for (unsigned li = 0; li < n_lights; ++li) {
        const LyapLight &L = lights[li];
        const V3 camV = sub(vec(cam.C), P);
        V3 lightV = sub(vec(L.C), P);
        const float d2 = mag2(lightV);
        lightV = normalized(lightV);
        float i = dot(lightV, N);
        const float j = -dot(lightV, vec(L.V));
        C4 phong;
        if (j > L.lightOuterCone) {
            i = clamp01(i);
            const C4 diffuse = cmul(col(L.diffuseColor), i * L.diffusePower);
            const V3 halfV = normalized(add(camV, lightV));
            float s = clamp01(dot(N, halfV));
            s = powf(s, L.specularHardness);
            const C4 specular = cmul(col(L.specularColor), s * L.specularPower);
            phong = cmul(cadd(specular, diffuse), L.lightRange / d2);
            if (j < L.lightInnerCone)
                phong = cmul(phong, (j - L.lightOuterCone) /
                                    (L.lightInnerCone - L.lightOuterCone));
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

/* to_rgba (color.hpp:169-175): 255*x in double, truncated; saturated here. */
inline unsigned char to_u8(float x)
{
    const double v = 255.0 * (double)x;
    if (!(v > 0.0)) return 0;
    if (v >= 255.0) return 255;
    return (unsigned char)v;
}

/* ---- raymarch (kernel.cu:156-486) -------------------------------------- */

struct RayOut {
    LyapPoint pt;
    float depth;       /* t of the hit, NaN on a miss */
    int hit;
    int hit_index;     /* march sample index of the hit, -1 on a miss */
    uint32_t evals;    /* algorithmic lambda evaluations */
    float margin;      /* smallest |lambda - opaque| over the march decisions */
};

/* Ray entry/exit against the box (kernel.cu:181-285). Returns false on a miss.
 * Plane parameters are formed in double and stored as float, as the
 * reference's double LMIN/LMAX literals make them. */
bool clip_box(V3 C, V3 V, float lo, float hi, float *t0_out, float *t1_out)
{
    const double dlo = lo, dhi = hi;
    float ts[6];
    const float Ca[3] = {C.x, C.y, C.z};
    const float Va[3] = {V.x, V.y, V.z};
    for (int a = 0; a < 3; ++a) {
        ts[2 * a] = (Va[a] != 0.0f) ? (float)((dlo - Ca[a]) / Va[a]) : INFINITY;
        ts[2 * a + 1] = (Va[a] != 0.0f) ? (float)((dhi - Ca[a]) / Va[a]) : INFINITY;
    }
    for (int k = 0; k < 6; ++k) {
        if (ts[k] == INFINITY) continue;
        const V3 P = add(C, mul(V, ts[k]));
        const float Pa[3] = {P.x, P.y, P.z};
        const int ax = k / 2;
        bool out = false;
        for (int b = 0; b < 3; ++b) {
            if (b == ax) continue;
            if ((double)Pa[b] < dlo || (double)Pa[b] > dhi) out = true;
        }
        if (out) ts[k] = NAN;
    }
    float t0 = 3.40282347e+38F, t1 = 0.0f;
    int i0 = -1, i1 = -1;
    for (int k = 0; k < 6; ++k) {
        if (!std::isfinite(ts[k])) continue;
        if (i0 == -1 || ts[k] < t0) { t0 = ts[k]; i0 = k; }
        if (i1 == -1 || ts[k] > t1) { t1 = ts[k]; i1 = k; }
    }
    if (i0 == -1 && i1 == -1) return false;
    if (i1 == -1 || i0 == i1) { t1 = t0; t0 = 0.0f; }
    if (t0 < 0) t0 = 0.0f;
    *t0_out = t0;
    *t1_out = t1;
    return true;
}

V3 ray_direction(const LyapCam &cam, unsigned sx, unsigned sy)
{
    V3 V = add(add(vec(cam.S0), mul(vec(cam.SDX), (float)sx)),
               mul(vec(cam.SDY), (float)sy));
    V = normalized(V);
    return divv(V, cam.M);
}

RayOut raymarch(const LyapCam &cam, const LyapParams &prm, const Lambda &lam,
                const LyapRenderMode &mode, unsigned sx, unsigned sy)
{
    RayOut o;
    std::memset(&o, 0, sizeof(o));
    o.pt.a = NAN;
    o.pt.l = NAN;
    o.depth = NAN;
    o.hit = 0;
    o.hit_index = -1;
    o.evals = 0;
    o.margin = INFINITY;

    const V3 C = vec(cam.C);
    const V3 V = ray_direction(cam, sx, sy);
    float t0, t1;
    if (!clip_box(C, V, prm.lMin, prm.lMax, &t0, &t1)) return o;
    const bool bisect = mode.refine == LYAP_REFINE_BISECT;

    float t = t0;
    V3 P = add(C, mul(V, t));
    float a = 0.0f, c = 0.0f;
    float Fdt;
    if (prm.stepMethod == 1) Fdt = (t1 - t0) / prm.depth;
    else Fdt = std::sqrt(mag2(V)) / prm.depth;
    float dt = Fdt;
    const float Ndt = dt / prm.nearMultiplier;
    bool near = false;
    float last_step = dt;   /* increment that led to the current sample */
    float tb = t0, dtb = dt;  /* bisection method: t = tb + m * dtb */
    uint32_t m = 0;

    auto note = [&](double l) {
        const double m = std::fabs(l - (double)prm.opaqueThreshold);
        if (!(m >= (double)o.margin)) o.margin = (float)m;
    };

    float l = (float)lam(P.x, P.y, P.z);
    o.evals = 1;
    int index = 0;
    note(l);
    while (l > prm.opaqueThreshold) {
        float step = dt;
        if (prm.jitter != 0.0f) {
            float jit = l - truncf(l);
            if (jit < 0) jit = (float)(1.0 - (double)(jit * prm.jitter));
            else jit = (float)(1.0 + (double)(jit * prm.jitter));
            if (std::isfinite(jit)) step = dt * jit;
        }
        /* A step too small to change t (a grazing ray with t1 - t0 near 0,
         * or depth = inf) would never leave the box: the ray ends. The
         * bisection method samples t = tb + m*dtb directly, rebased when the
         * step length changes; the reference walk adds the step (and V*step
         * to P) each time. */
        if (bisect) {
            if (step != dtb) { tb = t; dtb = step; m = 0; }
            const float tn = tb + (float)(m + 1u) * dtb;
            if (!(tn > t)) return o;
            t = tn;
            ++m;
            P = add(C, mul(V, t));
        } else {
            if (!(t + step > t)) return o;
            t += step;
            P = add(P, mul(V, step));
        }
        last_step = step;
        if (t > t1) return o;
        if (index + 1 >= kMaxSamples) return o;
        l = (float)lam(P.x, P.y, P.z);
        ++o.evals;
        ++index;
        note(l);
        if (l > prm.chaosThreshold) c += l;
        else if (l > prm.opaqueThreshold) a += l;
        if (l <= prm.nearThreshold && !near) { near = true; dt = Ndt; }
        else if (l > prm.nearThreshold && near) { near = false; dt = Fdt; }
    }
    if (t > t1) return o;
    o.hit_index = index;

    if (bisect) {
        float lo = t - last_step, hi = t;
        float l_hi = l;
        for (int k = 0; k < mode.bisect_steps; ++k) {
            const float mid = lo + (hi - lo) * 0.5f;
            const V3 Pm = add(C, mul(V, mid));
            const float lm = (float)lam(Pm.x, Pm.y, Pm.z);
            ++o.evals;
            if (!(lm > prm.opaqueThreshold)) { hi = mid; l_hi = lm; }
            else lo = mid;
        }
        t = hi;
        P = add(C, mul(V, t));
        l = l_hi;
    } else {
        /* reversing walk over [t - dt, t] (kernel.cu:402-441) */
        bool sign = false, osign;
        float Qdt = dt * -0.5f;
        V3 QdV = mul(V, Qdt);
        const float Qt1 = t, Qt0 = t - dt;
        const float min_Qdt = dt / prm.refine;
        for (int it = 0; it < 64; ++it) {
            if (!(t <= Qt1 && t >= Qt0 && (Qdt <= -min_Qdt || Qdt >= min_Qdt)))
                break;
            t += Qdt;
            P = add(P, QdV);
            l = (float)lam(P.x, P.y, P.z);
            ++o.evals;
            if (l == prm.opaqueThreshold) break;
            osign = sign;
            sign = (l < prm.opaqueThreshold) ? false : true;
            if (sign != osign) {
                Qdt *= -0.5f;
                QdV = mul(QdV, -0.5f);
            }
        }
    }

    /* normal from six lambda samples (kernel.cu:456-473) */
    const float h = (mode.normal_h > 0.0f) ? mode.normal_h : dt * prm.gradient;
    const float lxm = (float)lam(P.x - h, P.y, P.z);
    const float lxp = (float)lam(P.x + h, P.y, P.z);
    const float lym = (float)lam(P.x, P.y - h, P.z);
    const float lyp = (float)lam(P.x, P.y + h, P.z);
    const float lzm = (float)lam(P.x, P.y, P.z - h);
    const float lzp = (float)lam(P.x, P.y, P.z + h);
    o.evals += 6;
    const V3 N = normalized({lxp - lxm, lyp - lym, lzp - lzm});

    o.pt.P = lvec(P);
    o.pt.N = lvec(N);
    o.pt.a = a;
    o.pt.c = c;
    o.pt.l = l;
    o.depth = t;
    o.hit = 1;
    return o;
}

inline float coord(float lo, float hi, unsigned i, unsigned n)
{
    return lo + ((hi - lo) * (float)i) / (float)n;
}

LyapRenderMode reference_mode()
{
    LyapRenderMode m;
    std::memset(&m, 0, sizeof(m));
    m.refine = LYAP_REFINE_WALK;
    return m;
}

} /* namespace */

extern "C" {

/* lambda at n points (xyz float triples) in the chosen form. */
void oracle_lyap_points(int mode, const float *xyz, size_t n, float d,
                        unsigned settle, unsigned accum, const int32_t *seq,
                        double *out, int threads)
{
    const Lambda lam{mode, d, settle, accum, seq};
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (long i = 0; i < (long)n; ++i)
        out[i] = lam(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

/* FP64 lambda with the accumulated sum split at step `split` (the two
 * halves of the window, each divided by its own length): used to tell
 * converged orbits from ones still in a transient. */
void oracle_lyap_halves(const float *xyz, size_t n, float d, unsigned settle,
                        unsigned accum, unsigned split, const int32_t *seq,
                        double *first, double *second, int threads)
{
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
    for (long i = 0; i < (long)n; ++i) {
        const double coordv[4] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], d};
        unsigned pos = 0;
        double v = 0.5, s1 = 0.0, s2 = 0.0;
        for (unsigned k = 0; k < settle; ++k) {
            const double r = coordv[seq[pos]];
            ++pos;
            if (seq[pos] == -1) pos = 0;
            v = r * v * (1.0 - v);
        }
        for (unsigned k = 0; k < accum; ++k) {
            const double r = coordv[seq[pos]];
            ++pos;
            if (seq[pos] == -1) pos = 0;
            v = r * v * (1.0 - v);
            double g = r - 2.0 * r * v;
            if (g < 0) g = -g;
            if (k < split) s1 += std::log(g);
            else s2 += std::log(g);
        }
        first[i] = split ? s1 / split : NAN;
        second[i] = (accum > split) ? s2 / (accum - split) : NAN;
    }
}

/* Volume over [lo, hi) (kernel.cu:518-532 generalised): slices z in
 * [z_begin, z_end) of the nx*ny*nz grid are written to out[x + (y + (z -
 * z_begin)*ny)*nx]. */
void oracle_volume(int mode, float *out, const LyapParams *prm,
                   const int32_t *seq, const float *lo, const float *hi,
                   unsigned nx, unsigned ny, unsigned nz, unsigned z_begin,
                   unsigned z_end, int threads)
{
    const Lambda lam{mode, prm->d, prm->settle, prm->accum, seq};
    (void)nz;
    const long plane = (long)nx * ny;
    const long total = plane * (long)(z_end - z_begin);
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (long idx = 0; idx < total; ++idx) {
        const unsigned x = (unsigned)(idx % nx);
        const unsigned y = (unsigned)((idx / nx) % ny);
        const unsigned z = z_begin + (unsigned)(idx / plane);
        out[idx] = (float)lam(coord(lo[0], hi[0], x, nx),
                              coord(lo[1], hi[1], y, ny),
                              coord(lo[2], hi[2], z, nz));
    }
}

/* Grid coordinate used by the volume path (for tests). */
float oracle_coord(float lo, float hi, unsigned i, unsigned n)
{
    return coord(lo, hi, i, n);
}

/* Render rows row_begin, row_begin+row_step, ... of a w*h image. Outputs
 * are indexed by pixel x + y*w; any pointer may be NULL. */
void oracle_render(int lam_mode, const LyapRenderMode *mode_in, LyapRGBA *rgba,
                   LyapPoint *points, float *depth, float *lambda,
                   uint32_t *evals, int32_t *hit_index, float *margin,
                   const LyapCam *cam, const LyapParams *prm,
                   const int32_t *seq, const LyapLight *lights,
                   unsigned num_lights, unsigned w, unsigned h,
                   unsigned row_begin, unsigned row_step, int threads)
{
    const LyapRenderMode mode = mode_in ? *mode_in : reference_mode();
    const Lambda lam{lam_mode, prm->d, prm->settle, prm->accum, seq};
    if (row_step == 0) row_step = 1;
    const long rows = (h > row_begin) ? (long)((h - row_begin + row_step - 1) / row_step) : 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (long r = 0; r < rows; ++r) {
        const unsigned y = row_begin + (unsigned)r * row_step;
        for (unsigned x = 0; x < w; ++x) {
            const size_t ind = (size_t)x + (size_t)y * w;
            const RayOut o = raymarch(*cam, *prm, lam, mode, x, y);
            if (points) points[ind] = o.pt;
            if (depth) depth[ind] = o.depth;
            if (lambda) lambda[ind] = o.hit ? o.pt.l : NAN;
            if (evals) evals[ind] = o.evals;
            if (hit_index) hit_index[ind] = o.hit_index;
            if (margin) margin[ind] = o.margin;
            if (rgba) {
                const C4 cc = shade(o.pt, *cam, lights, num_lights);
                rgba[ind].r = to_u8(cc.x);
                rgba[ind].g = to_u8(cc.y);
                rgba[ind].b = to_u8(cc.z);
                rgba[ind].a = to_u8(cc.w);
            }
        }
    }
}

/* shade + to_rgba of given hit records (checks the device's shading). */
void oracle_shade(const LyapPoint *points, size_t n, const LyapCam *cam,
                  const LyapLight *lights, unsigned num_lights, float *color4,
                  LyapRGBA *rgba)
{
    for (size_t i = 0; i < n; ++i) {
        const C4 cc = shade(points[i], *cam, lights, num_lights);
        if (color4) {
            color4[4 * i] = cc.x;
            color4[4 * i + 1] = cc.y;
            color4[4 * i + 2] = cc.z;
            color4[4 * i + 3] = cc.w;
        }
        if (rgba) {
            rgba[i].r = to_u8(cc.x);
            rgba[i].g = to_u8(cc.y);
            rgba[i].b = to_u8(cc.z);
            rgba[i].a = to_u8(cc.w);
        }
    }
}

/* Ray entry/exit of pixel (sx, sy): returns 0 on a miss. */
int oracle_ray_setup(const LyapCam *cam, const LyapParams *prm, unsigned sx,
                     unsigned sy, float *V_out, float *t0, float *t1)
{
    const V3 V = ray_direction(*cam, sx, sy);
    if (V_out) { V_out[0] = V.x; V_out[1] = V.y; V_out[2] = V.z; }
    return clip_box(vec(cam->C), V, prm->lMin, prm->lMax, t0, t1) ? 1 : 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

} /* extern "C" */
