/*
 * lyap_device.hpp - the per-point Lyapunov exponent on gfx950.
 *
 * One lane evaluates one point. The reference routine is lyap4d
 * (kernel.cu:108-154): x0 = 0.5, `settle` steps of x <- r x (1 - x), then
 * `accum` steps summing ln|r_n (1 - 2 x_{n+1})| (the derivative uses the new
 * iterate, kernel.cu:143-144); |x - 0.5| < 1e-8 after the settle returns 0
 * (kernel.cu:138); a non-finite sum returns NaN (kernel.cu:147-149); the
 * sequence position wraps at the -1 terminator and carries over from the
 * settle into the accumulation (kernel.cu:133-134, 141-142).
 *
 * Three arithmetic forms:
 *
 *  plain<L>    FP32 iterate, one hardware log2 per accumulated step. The
 *              reference method in FP32; the baseline kernel.
 *
 *  product<L>  FP32 iterate in u = 1 - 2x, where the map becomes
 *              u' = (1 - r/2) + (r/2) u^2 and the derivative factor is
 *              r_n u_{n+1}. |u| <= 1 while 0 <= r <= 4, so the running product
 *              M of |u| only shrinks between renormalisations: it is carried
 *              as an FP32 mantissa plus an int32 exponent, renormalised every
 *              block of whole periods (up to 32 steps for the volumes, one
 *              16- or 32-slot window for the render), and a block that ends
 *              below the smallest normal float (where gradual underflow could
 *              have cost bits) is recomputed with a renormalisation per step.
 *              Renormalising is exact, so the block length changes no bit. The ln|r_n| terms depend only on the
 *              letter, so they are added once per letter with host-computed
 *              counts. One log2 per evaluation instead of one per step.
 *              Per accumulated step: one multiply (u^2), one FMA, one
 *              multiply into M.
 *
 *  fp64<L>     FP64 map with exactly the reference's operation order
 *              ((r*v)*(1-v), no fused multiply-add; the whole library is
 *              built with -ffp-contract=off), so the orbit is bitwise the one
 *              an IEEE double CPU computes. The derivative r(1 - 2v) only
 *              feeds the log sum and is formed with one fma; the log sum is
 *              carried as a product with exponent renormalisation every 8
 *              steps, which changes lambda only by rounding (~1e-15).
 *
 * Sequence handling: the period L is a template parameter, so the
 * coefficients of each sequence position sit in VGPRs and the step loop is
 * unrolled over one period. The letters themselves are wave-uniform (kernel
 * arguments in SGPRs); a position's coefficient is picked from the lane's
 * (x, y, z, d) once per evaluation.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lyap.h"
#include "sequence.hpp"

namespace lyap {

constexpr float kLn2f = 0.693147180559945309f;
constexpr double kLn2 = 0.693147180559945309417232121458;
/* A product block ending below the smallest normal float may have lost bits
 * to gradual underflow; one ending at or above it did not, since |u| <= 1
 * keeps the product non-increasing within the block. Such a block is
 * recomputed with a renormalisation per step. */
constexpr float kUnderflowGuard = 0x1p-126f;

/* The per-period volume kernels redo a block that ended at infinity (or
 * NaN) as well: the orbit of a point with r > 4 (a volume reaching past 4)
 * can escape [0, 1], and then |u| > 1 grows fast and may overflow between
 * renormalisations where its factors are finite. Redone per step the product
 * stays finite while its factors are, so their adaptive block lengths never
 * change a bit, and lambda is finite exactly when every factor is (as with
 * the reference's sum of logs). The window forms of the persistent launch
 * keep the one compare: their samples lie in the [lMin, lMax] box (at most
 * an ulp past it), and this check costs the render kernels registers. */
__device__ __forceinline__ bool redo_block(float M)
{
    return !(M >= kUnderflowGuard && M <= 0x1.fffffep127f);
}

/* Wave-uniform description of one lambda evaluation: iteration counts,
 * sequence, and host-derived loop bounds. Built by make_eval_desc(). */
struct EvalDesc {
    uint32_t settle, accum;
    float d;
    int32_t L;
    int32_t letters[LYAP_MAX_SEQ];
    /* settle = settle_periods * L + settle_rem */
    uint32_t settle_periods, settle_rem;
    /* accumulation: positions [head_begin, head_end) of the period the settle
     * ended in, then acc_periods full periods, then positions [0, tail). */
    uint32_t head_begin, head_end;
    uint32_t acc_periods, tail;
    /* first part of the window for the convergence check: the head plus
     * half_periods full periods (half_steps steps) */
    uint32_t half_periods, half_steps;
    /* product form: full periods per renormalisation block, and after a
     * block that had to be redone */
    uint32_t block_periods;
    uint32_t block_short;
    /* occurrences of letters A..D in the whole window and in its first part */
    float count[4];
    float count_half[4];
    /* Blocked layout for the period-agnostic kernels: a block is B steps, a
     * whole number of periods that fits the kernel's coefficient window W;
     * position j of a block uses letter wletters[j] = letters[j mod L].
     * settle = set_blocks * B + set_rem; the accumulation finishes that block
     * (positions [head_lo, head_hi)), runs acc_blocks blocks, then positions
     * [0, acc_rem). */
    int32_t wletters[LYAP_MAX_SEQ];
    /* the same as bit masks: bit j of wmask[k] is set when wletters[j] == k,
     * so a kernel builds its coefficients from one 16-byte scalar load */
    uint32_t wmask[4];
    uint32_t W, B;
    uint32_t set_blocks, set_rem;
    uint32_t head_lo, head_hi;
    uint32_t acc_blocks, acc_rem;
};

inline void rebuild_blocks(EvalDesc *e, uint32_t W);

/* Default product blocks of the per-period volume kernels: whole periods
 * making up to kBlockLong steps, and after the first block that had to be
 * redone (see redo_block) the whole periods nearest kBlockShort. The whole wave waits for any
 * lane's redo: where lambda is strongly negative (the stable regions of the
 * [2,4]^3 volumes) 30-step blocks redo about half the time and 10-step ones
 * rarely, while in chaotic regions long blocks redo as rarely and
 * renormalise less often. */
constexpr uint32_t kBlockLong = 32, kBlockShort = 12;

/* Host: derive the loop bounds. Returns false for an unusable sequence. */
inline bool make_eval_desc(EvalDesc *e, const LyapParams &prm,
                           const int32_t *seq, uint32_t renorm)
{
    const int L = sequence_period(seq);
    if (L == 0) return false;
    *e = EvalDesc{};
    e->settle = prm.settle;
    e->accum = prm.accum;
    e->d = prm.d;
    e->L = L;
    for (int j = 0; j < L; ++j) e->letters[j] = seq[j];
    e->settle_periods = prm.settle / L;
    e->settle_rem = prm.settle % L;
    uint32_t left = prm.accum;
    e->head_begin = e->settle_rem;
    e->head_end = e->settle_rem;
    if (e->settle_rem != 0) {
        const uint32_t h = (uint32_t)L - e->settle_rem;
        const uint32_t take = left < h ? left : h;
        e->head_end = e->settle_rem + take;
        left -= take;
    }
    e->acc_periods = left / L;
    e->tail = left % L;
    e->half_periods = e->acc_periods / 2;
    e->half_steps = (e->head_end - e->head_begin) + e->half_periods * (uint32_t)L;
    uint32_t bp = (renorm ? renorm : kBlockLong) / (uint32_t)L;
    e->block_periods = bp ? bp : 1;
    /* an explicit length stays; the default shortens after a redo */
    bp = renorm ? e->block_periods : (kBlockShort + (uint32_t)L / 2u) / (uint32_t)L;
    e->block_short = bp ? bp : 1;
    /* letter counts */
    for (int k = 0; k < 4; ++k) { e->count[k] = 0.f; e->count_half[k] = 0.f; }
    for (uint32_t j = e->head_begin; j < e->head_end; ++j) {
        e->count[seq[j]] += 1.f;
        e->count_half[seq[j]] += 1.f;
    }
    for (int j = 0; j < L; ++j) {
        e->count[seq[j]] += (float)e->acc_periods;
        e->count_half[seq[j]] += (float)e->half_periods;
    }
    for (uint32_t j = 0; j < e->tail; ++j) e->count[seq[j]] += 1.f;
    rebuild_blocks(e, L <= 16 ? 16u : 32u);
    return true;
}

/* Host: blocked layout for a coefficient window of W (16 or 32) positions. */
inline void rebuild_blocks(EvalDesc *e, uint32_t W)
{
    const uint32_t L = (uint32_t)e->L;
    e->W = W;
    /* a 32-slot window with a period of at most 16 takes two 16-slot blocks
     * at a time, so the FP64 form, renormalised at half and end of each
     * block, renormalises after the same steps in either window */
    e->B = (W > 16 && L <= 16) ? 2u * (L * (16u / L)) : L * (W / L);
    for (uint32_t k = 0; k < 4; ++k) e->wmask[k] = 0;
    for (uint32_t j = 0; j < (uint32_t)LYAP_MAX_SEQ; ++j) {
        e->wletters[j] = e->letters[j % L];
        e->wmask[e->wletters[j] & 3] |= 1u << j;
    }
    e->set_blocks = e->settle / e->B;
    e->set_rem = e->settle % e->B;
    uint32_t left = e->accum;
    e->head_lo = e->head_hi = 0;
    if (e->set_rem != 0) {
        const uint32_t h = e->B - e->set_rem;
        const uint32_t take = left < h ? left : h;
        e->head_lo = e->set_rem;
        e->head_hi = e->set_rem + take;
        left -= take;
    }
    e->acc_blocks = left / e->B;
    e->acc_rem = left % e->B;
}

/* Lane's value of sequence letter `s` (0..3 = x, y, z, d); s is uniform. */
__device__ __forceinline__ float pick(int s, float x, float y, float z, float d)
{
    return s == 0 ? x : (s == 1 ? y : (s == 2 ? z : d));
}
__device__ __forceinline__ double pickd(int s, double x, double y, double z,
                                        double d)
{
    return s == 0 ? x : (s == 1 ? y : (s == 2 ? z : d));
}

/* Lane's value of the letter at window position j, from the letter masks
 * (uniform); one scalar load for the whole window. */
template <typename T>
__device__ __forceinline__ T pick_mask(const uint32_t (&m)[4], int j, T x, T y, T z, T d)
{
    return ((m[0] >> j) & 1u) ? x : (((m[1] >> j) & 1u) ? y : (((m[2] >> j) & 1u) ? z : d));
}

/* ---- plain FP32: one log2 per accumulated step ------------------------- */

template <int L>
__device__ __forceinline__ float lambda_plain(float x, float y, float z,
                                              const EvalDesc &e)
{
    float r[L], m2r[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        r[j] = pick(e.letters[j], x, y, z, e.d);
        m2r[j] = -2.0f * r[j];
    }
    float v = 0.5f;
    for (uint32_t p = 0; p < e.settle_periods; ++p) {
#pragma unroll
        for (int j = 0; j < L; ++j) v = r[j] * __builtin_fmaf(-v, v, v);
    }
#pragma unroll
    for (int j = 0; j < L; ++j)
        if ((uint32_t)j < e.settle_rem) v = r[j] * __builtin_fmaf(-v, v, v);

    if (v == 0.5f) return 0.0f / (float)e.accum;   /* kernel.cu:138 in FP32 */

    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j >= e.head_begin && (uint32_t)j < e.head_end) {
            v = r[j] * __builtin_fmaf(-v, v, v);
            s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(m2r[j], v, r[j])));
        }
    }
    for (uint32_t p = 0; p < e.acc_periods; ++p) {
#pragma unroll
        for (int j = 0; j < L; ++j) {
            v = r[j] * __builtin_fmaf(-v, v, v);
            s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(m2r[j], v, r[j])));
        }
    }
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j < e.tail) {
            v = r[j] * __builtin_fmaf(-v, v, v);
            s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(m2r[j], v, r[j])));
        }
    }
    const float lam = s * kLn2f / (float)e.accum;
    return __builtin_isfinite(lam) ? lam : __builtin_nanf("");
}

/* ---- FP32 u-form with a mantissa/exponent product ---------------------- */

struct ProductState {
    float M;     /* mantissa of the running product of |u| */
    int E;       /* its binary exponent */
};

__device__ __forceinline__ void renorm(ProductState &st)
{
    st.E += __builtin_amdgcn_frexp_expf(st.M);
    st.M = __builtin_amdgcn_frexp_mantf(st.M);
}

/* ln of the product plus the letter terms, divided by n; NaN if not finite. */
__device__ __forceinline__ float finish_product(const ProductState &st,
                                                const float lg[4],
                                                const float cnt[4], uint32_t n)
{
    double t = (double)st.E + (double)__builtin_amdgcn_logf(st.M);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (cnt[k] != 0.0f) t += (double)cnt[k] * (double)lg[k];
    const float lam = (float)(t * kLn2 / (double)n);
    return __builtin_isfinite(lam) ? lam : __builtin_nanf("");
}

/* Result of a product-form evaluation. half1/half2 are the lambdas of the
 * two parts of the window (NaN when a part is empty). */
struct ProductResult {
    float lam;
    float half1, half2;
};

/* FP32 maps in u = 1 - 2x with b = r/2. The volume form u' = b u^2 + a with
 * a = 1 - b (a multiply and an FMA) is what every volume evaluator uses (the
 * per-period kernels and the window form of volume blocks in the persistent
 * launch), so they give the same bits for the same voxel; the render's form
 * u' = 1 + b (u^2 - 1) (two FMAs) needs no a and so half the coefficient
 * registers. Renormalising the product is exact, so block lengths change no
 * bit; only the map's rounding does. */
__device__ __forceinline__ float vmap(float u, float a, float b)
{
    return __builtin_fmaf(b, u * u, a);
}

__device__ __forceinline__ float umap(float u, float b)
{
    return __builtin_fmaf(b, __builtin_fmaf(u, u, -1.0f), 1.0f);
}

/* `n` full periods in blocks of `bp` periods, renormalising after each block
 * and recomputing a block whose product redo_block() flags; once a lane of
 * the wave has needed that, the blocks are `bshort` periods long. */
template <int L>
__device__ __forceinline__ void product_periods(float &u, ProductState &st,
                                                const float (&a)[L], const float (&b)[L],
                                                uint32_t n, uint32_t &bp, uint32_t bshort)
{
    for (uint32_t p = 0; p < n;) {
        const uint32_t cnt = n - p < bp ? n - p : bp;
        p += cnt;
        const float u0 = u;
        const ProductState s0 = st;
        for (uint32_t k = 0; k < cnt; ++k) {
#pragma unroll
            for (int j = 0; j < L; ++j) {
                u = vmap(u, a[j], b[j]);
                st.M *= __builtin_fabsf(u);
            }
        }
        const bool rd = redo_block(st.M);
        if (rd) {
            float w = u0;
            st = s0;
            for (uint32_t k = 0; k < cnt; ++k) {
#pragma unroll
                for (int j = 0; j < L; ++j) {
                    w = vmap(w, a[j], b[j]);
                    st.M *= __builtin_fabsf(w);
                    renorm(st);
                }
            }
        }
        if (__any(rd)) bp = bshort;
        renorm(st);
    }
}

template <int L, bool HALVES>
__device__ __forceinline__ ProductResult lambda_product(float x, float y,
                                                        float z,
                                                        const EvalDesc &e)
{
    float a[L], b[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        b[j] = 0.5f * pick(e.letters[j], x, y, z, e.d);
        a[j] = 1.0f - b[j];
    }
    float u = 0.0f;
    for (uint32_t p = 0; p < e.settle_periods; ++p) {
#pragma unroll
        for (int j = 0; j < L; ++j) u = vmap(u, a[j], b[j]);
    }
#pragma unroll
    for (int j = 0; j < L; ++j)
        if ((uint32_t)j < e.settle_rem) u = vmap(u, a[j], b[j]);

    ProductResult res;
    /* kernel.cu:138: |x - 0.5| < 1e-8 is |u| < 2e-8 */
    if (__builtin_fabsf(u) < 2e-8f) {
        res.lam = 0.0f / (float)e.accum;
        res.half1 = res.half2 = res.lam;
        return res;
    }

    ProductState st{1.0f, 0};
    /* head: the rest of the period the settle ended in, renormalised per step */
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j >= e.head_begin && (uint32_t)j < e.head_end) {
            u = vmap(u, a[j], b[j]);
            st.M *= __builtin_fabsf(u);
            renorm(st);
        }
    }
    ProductState st_half{1.0f, 0};
    uint32_t bp = e.block_periods;
    if (HALVES) {
        product_periods<L>(u, st, a, b, e.half_periods, bp, e.block_short);
        st_half = st;
        product_periods<L>(u, st, a, b, e.acc_periods - e.half_periods, bp, e.block_short);
    } else {
        product_periods<L>(u, st, a, b, e.acc_periods, bp, e.block_short);
    }
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j < e.tail) {
            u = vmap(u, a[j], b[j]);
            st.M *= __builtin_fabsf(u);
            renorm(st);
        }
    }

    float lg[4];
    lg[0] = __builtin_amdgcn_logf(__builtin_fabsf(x));
    lg[1] = __builtin_amdgcn_logf(__builtin_fabsf(y));
    lg[2] = __builtin_amdgcn_logf(__builtin_fabsf(z));
    lg[3] = __builtin_amdgcn_logf(__builtin_fabsf(e.d));
    res.lam = finish_product(st, lg, e.count, e.accum);
    res.half1 = res.half2 = res.lam;
    if (HALVES) {
        const uint32_t n1 = e.half_steps, n2 = e.accum - e.half_steps;
        float c2[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) c2[k] = e.count[k] - e.count_half[k];
        ProductState second;
        second.E = st.E - st_half.E;
        second.M = st.M / st_half.M;
        res.half1 = n1 ? finish_product(st_half, lg, e.count_half, n1) : __builtin_nanf("");
        res.half2 = n2 ? finish_product(second, lg, c2, n2) : __builtin_nanf("");
    }
    return res;
}

/* ---- FP64 with the reference's operation order ------------------------- */

struct ProductState64 {
    double M;
    int E;
};

__device__ __forceinline__ void renorm64(ProductState64 &st)
{
    st.E += __builtin_amdgcn_frexp_exp(st.M);
    st.M = __builtin_amdgcn_frexp_mant(st.M);
}

template <int L>
__device__ __forceinline__ double lambda_fp64(float xf, float yf, float zf,
                                              const EvalDesc &e)
{
    const double x = xf, y = yf, z = zf, d = e.d;
    double r[L], m2r[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        r[j] = pickd(e.letters[j], x, y, z, d);
        m2r[j] = -2.0 * r[j];        /* exact */
    }
    double v = 0.5;
    for (uint32_t p = 0; p < e.settle_periods; ++p) {
#pragma unroll
        for (int j = 0; j < L; ++j) v = (r[j] * v) * (1.0 - v);
    }
#pragma unroll
    for (int j = 0; j < L; ++j)
        if ((uint32_t)j < e.settle_rem) v = (r[j] * v) * (1.0 - v);

    const double off = v - 0.5;
    if (!(off <= -1e-8 || off >= 1e-8)) return 0.0 / (double)e.accum;

    ProductState64 st{1.0, 0};
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j >= e.head_begin && (uint32_t)j < e.head_end) {
            v = (r[j] * v) * (1.0 - v);
            st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
            renorm64(st);
        }
    }
    /* factors are at most max|r| * max|1 - 2v|; 8 of them cannot leave the
     * double range for coordinates of float magnitude (or the orbit escapes,
     * which makes lambda NaN either way) */
    for (uint32_t p = 0; p < e.acc_periods; ++p) {
#pragma unroll
        for (int j = 0; j < L; ++j) {
            v = (r[j] * v) * (1.0 - v);
            st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
            if (((j + 1) & 7) == 0) renorm64(st);
        }
        if ((L & 7) != 0) renorm64(st);
    }
#pragma unroll
    for (int j = 0; j < L; ++j) {
        if ((uint32_t)j < e.tail) {
            v = (r[j] * v) * (1.0 - v);
            st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
            renorm64(st);
        }
    }
    const double lam = ((double)st.E * kLn2 + log(st.M)) / (double)e.accum;
    return __builtin_isfinite(lam) ? lam : __builtin_nan("");
}

/* ======================================================================== */
/* Period-agnostic forms: one compiled loop for every period L <= W.         */
/* ======================================================================== */

/* The coefficients of a window of W slots: b_j = r_j / 2 and, for the
 * volume form (VOL), a_j = 1 - b_j; step(j, u) is the map at slot j. */
template <int W, bool VOL>
struct Coefs {
    float b[W];
    float a[VOL ? W : 1];
    __device__ __forceinline__ float step(int j, float u) const
    {
        if constexpr (VOL) return vmap(u, a[j], b[j]);
        else return umap(u, b[j]);
    }
};

/* Steps of a block at positions [lo, hi) (FULL: [0, B), where every block
 * has at least W/2 + 1 steps, so only the upper half is guarded). */
template <int W, bool FULL>
__device__ __forceinline__ bool in_block(int j, uint32_t lo, uint32_t hi)
{
    if (FULL) return j <= W / 2 || (uint32_t)j < hi;
    return (uint32_t)j >= lo && (uint32_t)j < hi;
}

template <int W, bool VOL>
__device__ __forceinline__ void load_half_r(Coefs<W, VOL> &c, float x, float y,
                                            float z, const EvalDesc &e)
{
    const float hx = 0.5f * x, hy = 0.5f * y, hz = 0.5f * z, hd = 0.5f * e.d;
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = e.wmask[k];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        c.b[j] = pick_mask<float>(m, j, hx, hy, hz, hd);
        if constexpr (VOL) c.a[j] = 1.0f - c.b[j];
    }
}

template <int W, bool VOL, bool FULL>
__device__ __forceinline__ void settle_block(float &u, const Coefs<W, VOL> &c,
                                             uint32_t lo, uint32_t hi)
{
#pragma unroll
    for (int j = 0; j < W; ++j)
        if (in_block<W, FULL>(j, lo, hi)) u = c.step(j, u);
}

template <int W, bool VOL, bool FULL>
__device__ __forceinline__ void product_block(float &u, ProductState &st,
                                              const Coefs<W, VOL> &c, uint32_t lo,
                                              uint32_t hi)
{
    const float u0 = u;
    const ProductState s0 = st;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (in_block<W, FULL>(j, lo, hi)) {
            u = c.step(j, u);
            st.M *= __builtin_fabsf(u);
        }
    }
    if (st.M < kUnderflowGuard) {
        /* |u| <= 1 keeps M monotone within the block, so ending above the
         * guard means no step went below it; otherwise redo per step */
        float w = u0;
        st = s0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (in_block<W, FULL>(j, lo, hi)) {
                w = c.step(j, w);
                st.M *= __builtin_fabsf(w);
                renorm(st);
            }
        }
    }
    renorm(st);
}

/* Calls f.template go<B>() with B the block length Bv, one of (W/2, W]. A
 * full block then runs exactly B unguarded steps. Bv comes from the view
 * descriptor and is the same for every lane of the wave, so the choice is a
 * scalar branch. */
template <int W, int B, class F>
__device__ __forceinline__ void with_block_len(uint32_t Bv, F &f)
{
    if constexpr (B > W / 2) {
        if (Bv == (uint32_t)B) f.template go<B>();
        else with_block_len<W, B - 1>(Bv, f);
    }
}

/* n full settle blocks of B steps. */
template <int W, bool VOL>
struct SettleFull {
    float &u;
    const Coefs<W, VOL> &c;
    uint32_t n;
    template <int B>
    __device__ __forceinline__ void go()
    {
        for (uint32_t k = 0; k < n; ++k) {
#pragma unroll
            for (int j = 0; j < B; ++j) u = c.step(j, u);
        }
    }
};

/* n full product blocks of B steps, with the same underflow redo as
 * product_block. */
template <int W, bool VOL>
struct ProductFull {
    float &u;
    ProductState &st;
    const Coefs<W, VOL> &c;
    uint32_t n;
    template <int B>
    __device__ __forceinline__ void go()
    {
        for (uint32_t k = 0; k < n; ++k) {
            const float u0 = u;
            const ProductState s0 = st;
#pragma unroll
            for (int j = 0; j < B; ++j) {
                u = c.step(j, u);
                st.M *= __builtin_fabsf(u);
            }
            if (st.M < kUnderflowGuard) {
                float w = u0;
                st = s0;
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    w = c.step(j, w);
                    st.M *= __builtin_fabsf(w);
                    renorm(st);
                }
            }
            renorm(st);
        }
    }
};

/* Calls f.template go<N>() for n in [1, W); nothing for n = 0. Partial
 * blocks are specialised by their start or length this way, so none of
 * their steps carries a per-step guard. n is wave-uniform. */
template <int W, int N, class F>
__device__ __forceinline__ void with_len(uint32_t n, F &f)
{
    if constexpr (N >= 1) {
        if (n == (uint32_t)N) f.template go<N>();
        else with_len<W, N - 1>(n, f);
    }
}

/* Window positions [LO, HI) of the product form; positions past W/2 are
 * guarded by j < hi when GUARD (they may lie past a block of B < W). */
template <int W, bool VOL, int LO, int HI, bool GUARD>
__device__ __forceinline__ void product_range(float &u, ProductState &st,
                                              const Coefs<W, VOL> &c, uint32_t hi)
{
    const float u0 = u;
    const ProductState s0 = st;
#pragma unroll
    for (int j = LO; j < HI; ++j) {
        if (!GUARD || j <= W / 2 || (uint32_t)j < hi) {
            u = c.step(j, u);
            st.M *= __builtin_fabsf(u);
        }
    }
    if (st.M < kUnderflowGuard) {
        float w = u0;
        st = s0;
#pragma unroll
        for (int j = LO; j < HI; ++j) {
            if (!GUARD || j <= W / 2 || (uint32_t)j < hi) {
                w = c.step(j, w);
                st.M *= __builtin_fabsf(w);
                renorm(st);
            }
        }
    }
    renorm(st);
}

/* Settle steps [0, N) / product steps [0, N) / the rest of a block from N. */
template <int W, bool VOL>
struct SettlePrefix {
    float &u;
    const Coefs<W, VOL> &c;
    template <int N>
    __device__ __forceinline__ void go()
    {
#pragma unroll
        for (int j = 0; j < N; ++j) u = c.step(j, u);
    }
};

template <int W, bool VOL>
struct ProductPrefix {
    float &u;
    ProductState &st;
    const Coefs<W, VOL> &c;
    template <int N>
    __device__ __forceinline__ void go() { product_range<W, VOL, 0, N, false>(u, st, c, 0); }
};

template <int W, bool VOL>
struct ProductHead {
    float &u;
    ProductState &st;
    const Coefs<W, VOL> &c;
    uint32_t hi;
    template <int N>
    __device__ __forceinline__ void go() { product_range<W, VOL, N, W, true>(u, st, c, hi); }
};

/* FP32 product form, any period L <= W, with the render's map. */
template <int W>
__device__ __forceinline__ float lambda_product_w(float x, float y, float z,
                                                  const EvalDesc &e)
{
    constexpr bool VOL = false;
    Coefs<W, VOL> c;
    load_half_r<W, VOL>(c, x, y, z, e);
    float u = 0.0f;
    {
        SettleFull<W, VOL> f{u, c, e.set_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        SettlePrefix<W, VOL> f{u, c};
        with_len<W, W - 1>(e.set_rem, f);
    }
    if (__builtin_fabsf(u) < 2e-8f) return 0.0f / (float)e.accum;   /* kernel.cu:138 */
    ProductState st{1.0f, 0};
    if (e.head_hi == e.B) {
        /* the usual head: the rest of the block the settle ended in */
        ProductHead<W, VOL> f{u, st, c, e.B};
        with_len<W, W - 1>(e.head_lo, f);
    } else {
        product_block<W, VOL, false>(u, st, c, e.head_lo, e.head_hi);
    }
    {
        ProductFull<W, VOL> f{u, st, c, e.acc_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        ProductPrefix<W, VOL> f{u, st, c};
        with_len<W, W - 1>(e.acc_rem, f);
    }
    float lg[4];
    lg[0] = __builtin_amdgcn_logf(__builtin_fabsf(x));
    lg[1] = __builtin_amdgcn_logf(__builtin_fabsf(y));
    lg[2] = __builtin_amdgcn_logf(__builtin_fabsf(z));
    lg[3] = __builtin_amdgcn_logf(__builtin_fabsf(e.d));
    return finish_product(st, lg, e.count, e.accum);
}

/* n product steps from window position p (p < B): the rest of that block,
 * whole blocks, then a prefix of one more block. Every part ends with a
 * renormalisation, which is exact, so where a run is split changes no bit
 * of the product. */
template <int W, bool VOL>
__device__ __forceinline__ void product_run(float &u, ProductState &st, const Coefs<W, VOL> &c,
                                            uint32_t B, uint32_t p, uint32_t n)
{
    if (n == 0) return;
    if (p != 0) {
        const uint32_t hi = p + n < B ? p + n : B;
        if (hi == B) {
            ProductHead<W, VOL> f{u, st, c, B};
            with_len<W, W - 1>(p, f);
        } else {
            product_block<W, VOL, false>(u, st, c, p, hi);
        }
        n -= hi - p;
    }
    {
        ProductFull<W, VOL> f{u, st, c, n / B};
        with_block_len<W, W>(B, f);
    }
    {
        ProductPrefix<W, VOL> f{u, st, c};
        with_len<W, W - 1>(n % B, f);
    }
}

/* FP32 product form with the volume map, any period L <= W, with the
 * lambdas of the two parts of the window split after e.half_steps
 * accumulated steps: the same split, map and products as
 * lambda_product<L, true>, so the same bits (the products differ only in
 * where they were renormalised). */
template <int W>
__device__ __forceinline__ ProductResult lambda_product_halves_w(float x, float y, float z,
                                                                 const EvalDesc &e)
{
    constexpr bool VOL = true;
    Coefs<W, VOL> c;
    load_half_r<W, VOL>(c, x, y, z, e);
    float u = 0.0f;
    {
        SettleFull<W, VOL> f{u, c, e.set_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        SettlePrefix<W, VOL> f{u, c};
        with_len<W, W - 1>(e.set_rem, f);
    }
    ProductResult res;
    if (__builtin_fabsf(u) < 2e-8f) {                 /* kernel.cu:138 */
        res.lam = 0.0f / (float)e.accum;
        res.half1 = res.half2 = res.lam;
        return res;
    }
    ProductState st{1.0f, 0};
    const uint32_t n1 = e.half_steps;
    product_run<W, VOL>(u, st, c, e.B, e.set_rem, n1);
    const ProductState st_half = st;
    product_run<W, VOL>(u, st, c, e.B, (e.set_rem + n1) % e.B, e.accum - n1);
    float lg[4];
    lg[0] = __builtin_amdgcn_logf(__builtin_fabsf(x));
    lg[1] = __builtin_amdgcn_logf(__builtin_fabsf(y));
    lg[2] = __builtin_amdgcn_logf(__builtin_fabsf(z));
    lg[3] = __builtin_amdgcn_logf(__builtin_fabsf(e.d));
    res.lam = finish_product(st, lg, e.count, e.accum);
    const uint32_t n2 = e.accum - n1;
    float c2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c2[k] = e.count[k] - e.count_half[k];
    ProductState second;
    second.E = st.E - st_half.E;
    second.M = st.M / st_half.M;
    res.half1 = n1 ? finish_product(st_half, lg, e.count_half, n1) : __builtin_nanf("");
    res.half2 = n2 ? finish_product(second, lg, c2, n2) : __builtin_nanf("");
    return res;
}

/* FP32 x-form with one log2 per accumulated step (the reference method). */
template <int W, bool FULL>
__device__ __forceinline__ void plain_block(float &v, float &s, const float (&r)[W],
                                            uint32_t lo, uint32_t hi, bool acc)
{
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (in_block<W, FULL>(j, lo, hi)) {
            v = r[j] * __builtin_fmaf(-v, v, v);
            if (acc) s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(-2.0f * r[j], v, r[j])));
        }
    }
}

/* n full blocks of B steps of the x-form, accumulating when ACC. */
template <int W, bool ACC>
struct PlainFull {
    float &v;
    float &s;
    const float (&r)[W];
    uint32_t n;
    template <int B>
    __device__ __forceinline__ void go()
    {
        for (uint32_t k = 0; k < n; ++k) {
#pragma unroll
            for (int j = 0; j < B; ++j) {
                v = r[j] * __builtin_fmaf(-v, v, v);
                if (ACC) s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(-2.0f * r[j], v, r[j])));
            }
        }
    }
};

/* Plain steps at window positions [LO, HI), guarded past W/2 as in
 * product_range; ACC accumulates the log of the derivative. */
template <int W, bool ACC>
struct PlainRange {
    float &v;
    float &s;
    const float (&r)[W];
    uint32_t hi;
    template <int LO, int HI, bool GUARD>
    __device__ __forceinline__ void run()
    {
#pragma unroll
        for (int j = LO; j < HI; ++j) {
            if (!GUARD || j <= W / 2 || (uint32_t)j < hi) {
                v = r[j] * __builtin_fmaf(-v, v, v);
                if (ACC) s += __builtin_amdgcn_logf(__builtin_fabsf(__builtin_fmaf(-2.0f * r[j], v, r[j])));
            }
        }
    }
};

template <int W, bool ACC>
struct PlainPrefix {
    PlainRange<W, ACC> rg;
    template <int N>
    __device__ __forceinline__ void go() { rg.template run<0, N, false>(); }
};

template <int W, bool ACC>
struct PlainHead {
    PlainRange<W, ACC> rg;
    template <int N>
    __device__ __forceinline__ void go() { rg.template run<N, W, true>(); }
};

template <int W>
__device__ __forceinline__ float lambda_plain_w(float x, float y, float z,
                                                const EvalDesc &e)
{
    float r[W];
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = e.wmask[k];
#pragma unroll
    for (int j = 0; j < W; ++j) r[j] = pick_mask<float>(m, j, x, y, z, e.d);
    float v = 0.5f, s = 0.0f;
    {
        PlainFull<W, false> f{v, s, r, e.set_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        PlainPrefix<W, false> f{{v, s, r, 0}};
        with_len<W, W - 1>(e.set_rem, f);
    }
    if (v == 0.5f) return 0.0f / (float)e.accum;
    if (e.head_hi == e.B) {
        PlainHead<W, true> f{{v, s, r, e.B}};
        with_len<W, W - 1>(e.head_lo, f);
    } else {
        plain_block<W, false>(v, s, r, e.head_lo, e.head_hi, true);
    }
    {
        PlainFull<W, true> f{v, s, r, e.acc_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        PlainPrefix<W, true> f{{v, s, r, 0}};
        with_len<W, W - 1>(e.acc_rem, f);
    }
    const float lam = s * kLn2f / (float)e.accum;
    return __builtin_isfinite(lam) ? lam : __builtin_nanf("");
}

/* FP64 with the reference's map ((r*v)*(1-v), separately rounded, so the
 * orbit is the IEEE double one); the derivative r(1 - 2v) only feeds the
 * log sum and is formed with one fma. The product is renormalised at the end
 * of each block and, in a 32-slot window, after its first `mid` = B/2
 * positions: at most 16 factors between renormalisations, each at most 4 in
 * magnitude while the orbit stays in [0, 1], and the same steps in either
 * window (B is then twice the 16-slot block). An orbit that escapes (a
 * sample just outside the box) can overflow between renormalisations; that
 * happens after the same steps wherever the window or launch is. */
template <int W, bool FULL, bool ACC>
__device__ __forceinline__ void fp64_block(double &v, ProductState64 &st,
                                           const double (&r)[W],
                                           const double (&m2r)[W], uint32_t lo,
                                           uint32_t hi, uint32_t mid)
{
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (in_block<W, FULL>(j, lo, hi)) {
            v = (r[j] * v) * (1.0 - v);
            if (ACC) {
                st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
                if (W > 16 && (uint32_t)j + 1u == mid) renorm64(st);
            }
        }
    }
    if (ACC) renorm64(st);
}

/* n full FP64 blocks of B steps. */
template <int W, bool ACC>
struct Fp64Full {
    double &v;
    ProductState64 &st;
    const double (&r)[W];
    const double (&m2r)[W];
    uint32_t n;
    template <int B>
    __device__ __forceinline__ void go()
    {
        for (uint32_t k = 0; k < n; ++k) {
#pragma unroll
            for (int j = 0; j < B; ++j) {
                v = (r[j] * v) * (1.0 - v);
                if (ACC) {
                    st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
                    if (W > 16 && j + 1 == B / 2) renorm64(st);
                }
            }
            if (ACC) renorm64(st);
        }
    }
};

/* FP64 steps at window positions [LO, HI), guarded past W/2 as in
 * product_range; the product is renormalised after position mid - 1 (a
 * 32-slot window) and at the end, as in fp64_block: exact power-of-two
 * scalings, and the final (E, M) with M in [0.5, 1) is the same wherever
 * the earlier ones fell. */
template <int W, bool ACC>
struct Fp64Range {
    double &v;
    ProductState64 &st;
    const double (&r)[W];
    const double (&m2r)[W];
    uint32_t hi, mid;
    template <int LO, int HI, bool GUARD>
    __device__ __forceinline__ void run()
    {
#pragma unroll
        for (int j = LO; j < HI; ++j) {
            if (!GUARD || j <= W / 2 || (uint32_t)j < hi) {
                v = (r[j] * v) * (1.0 - v);
                if (ACC) {
                    st.M *= __builtin_fabs(__builtin_fma(m2r[j], v, r[j]));
                    if (W > 16 && (uint32_t)j + 1u == mid) renorm64(st);
                }
            }
        }
        if (ACC) renorm64(st);
    }
};

template <int W, bool ACC>
struct Fp64Prefix {
    Fp64Range<W, ACC> rg;
    template <int N>
    __device__ __forceinline__ void go() { rg.template run<0, N, false>(); }
};

template <int W, bool ACC>
struct Fp64Head {
    Fp64Range<W, ACC> rg;
    template <int N>
    __device__ __forceinline__ void go() { rg.template run<N, W, true>(); }
};

template <int W>
__device__ __forceinline__ double lambda_fp64_w(float xf, float yf, float zf,
                                                const EvalDesc &e)
{
    double r[W], m2r[W];
    const double x = xf, y = yf, z = zf, d = e.d;
    uint32_t m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = e.wmask[k];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        r[j] = pick_mask<double>(m, j, x, y, z, d);
        m2r[j] = -2.0 * r[j];
    }
    double v = 0.5;
    ProductState64 st{1.0, 0};
    const uint32_t mid = e.B / 2u;
    {
        Fp64Full<W, false> f{v, st, r, m2r, e.set_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        Fp64Prefix<W, false> f{{v, st, r, m2r, 0, mid}};
        with_len<W, W - 1>(e.set_rem, f);
    }
    const double off = v - 0.5;
    if (!(off <= -1e-8 || off >= 1e-8)) return 0.0 / (double)e.accum;
    if (e.head_hi == e.B) {
        Fp64Head<W, true> f{{v, st, r, m2r, e.B, mid}};
        with_len<W, W - 1>(e.head_lo, f);
    } else {
        fp64_block<W, false, true>(v, st, r, m2r, e.head_lo, e.head_hi, mid);
    }
    {
        Fp64Full<W, true> f{v, st, r, m2r, e.acc_blocks};
        with_block_len<W, W>(e.B, f);
    }
    {
        Fp64Prefix<W, true> f{{v, st, r, m2r, 0, mid}};
        with_len<W, W - 1>(e.acc_rem, f);
    }
    const double lam = ((double)st.E * kLn2 + log(st.M)) / (double)e.accum;
    return __builtin_isfinite(lam) ? lam : __builtin_nan("");
}

} /* namespace lyap */
