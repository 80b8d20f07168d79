/*
 * lyap_types.h - plain-old-data types of the Lyapunov hot path.
 *
 * Byte-identical to the FP32 build (Real = float, DOUBLE_REALS unset) of the
 * reference data model, structs.hpp:10-80, so a buffer written by either
 * implementation can be read by the other:
 *   LyapParams  56 B, align 4   (structs.hpp:53-68)
 *   LyapPoint   36 B, align 4   (structs.hpp:70-76)
 *   LyapCam    224 B, align 16  (structs.hpp:26-51; LyapLight is the same struct)
 *   RGBA         4 B            (structs.hpp:78-80)
 * The reference's Vec is a float3 (12 B, align 4) and its Quat and Color derive
 * from float4 (16 B, align 16); the field offsets below follow from that and
 * are checked by static assertions at the end of this header.
 *
 * Plain C: no HIP, CUDA or torch type crosses this boundary.
 */
#ifndef LYAP_TYPES_H
#define LYAP_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__cplusplus)
#define LYAP_ALIGN16 alignas(16)
#else
#define LYAP_ALIGN16 _Alignas(16)
#endif

/* Vec (vec3.hpp:12): three floats, 4-byte aligned like float3. */
typedef struct LyapVec {
    float x, y, z;
} LyapVec;

/* Quat (quat.hpp:17): x, y, z, w with w the scalar part; float4 alignment. */
typedef struct LyapQuat {
    LYAP_ALIGN16 float x;
    float y, z, w;
} LyapQuat;

/* Color (color.hpp:12): r, g, b, a stored as x, y, z, w; float4 alignment. */
typedef struct LyapColor {
    LYAP_ALIGN16 float x;
    float y, z, w;
} LyapColor;

/* LyapCam / LyapLight (structs.hpp:26-51). One struct serves as camera and
 * spotlight: C position, Q orientation, M half-width of the screen plane at
 * distance 1, V view axis, S0 screen corner ray, SDX/SDY per-pixel steps. */
typedef struct LyapCamLight {
    LyapVec C;
    LyapQuat Q;
    float M;
    LyapVec V;
    LyapVec S0;
    LyapVec SDX;
    LyapVec SDY;
    uint32_t textureWidth;
    uint32_t textureHeight;
    uint32_t renderWidth;
    uint32_t renderHeight;
    uint32_t renderDenominator;
    float lightInnerCone, lightOuterCone;
    float lightRange;
    LyapColor ambient;
    LyapColor diffuseColor;
    float diffusePower;
    LyapColor specularColor;
    float specularPower;
    float specularHardness;
    LyapColor chaosColor;
} LyapCamLight;

typedef LyapCamLight LyapCam;
typedef LyapCamLight LyapLight;

/* LyapParams (structs.hpp:53-68); meanings from their use in kernel.cu. */
typedef struct LyapParams {
    float d;               /* value of the 'D' sequence letter (kernel.cu:114) */
    uint32_t settle;       /* iterations without accumulation (kernel.cu:132) */
    uint32_t accum;        /* accumulated iterations (kernel.cu:140) */
    uint32_t stepMethod;   /* 1: (t1-t0)/depth, else |V|/depth (kernel.cu:299-311) */
    float nearThreshold;   /* lambda below which the march uses the near step (kernel.cu:377) */
    float nearMultiplier;  /* near step = step / nearMultiplier (kernel.cu:314) */
    float opaqueThreshold; /* lambda <= this is opaque (kernel.cu:326) */
    float chaosThreshold;  /* lambda above this accumulates into c (kernel.cu:370) */
    float depth;           /* march steps per ray length unit (kernel.cu:303,310) */
    float jitter;          /* step jitter from frac(lambda) (kernel.cu:334-352) */
    float refine;          /* refinement stops below step/refine (kernel.cu:416) */
    float gradient;        /* normal offset = step * gradient (kernel.cu:456) */
    float lMin, lMax;      /* Lyapunov-space box [lMin, lMax]^3 (kernel.cu:12-18) */
} LyapParams;

/* LyapPoint (structs.hpp:70-76): per-pixel hit record. */
typedef struct LyapPoint {
    LyapVec P;  /* hit position */
    LyapVec N;  /* unit normal, pointing toward increasing lambda */
    float a;    /* accumulated lambda of semi-opaque samples; NaN marks a miss */
    float c;    /* accumulated lambda of chaotic samples */
    float l;    /* lambda at P */
} LyapPoint;

/* RGBA (structs.hpp:78-80). */
typedef struct LyapRGBA {
    unsigned char r, g, b, a;
} LyapRGBA;

#ifndef LYAP_NO_REFERENCE_NAMES
typedef LyapRGBA RGBA;
#endif

#ifdef __cplusplus
} /* extern "C" */
#define LYAP_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define LYAP_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

LYAP_STATIC_ASSERT(sizeof(LyapVec) == 12, "Vec is 12 bytes");
LYAP_STATIC_ASSERT(sizeof(LyapQuat) == 16, "Quat is 16 bytes");
LYAP_STATIC_ASSERT(sizeof(LyapParams) == 56, "LyapParams is 56 bytes");
LYAP_STATIC_ASSERT(sizeof(LyapPoint) == 36, "LyapPoint is 36 bytes");
LYAP_STATIC_ASSERT(sizeof(LyapCam) == 224, "LyapCam is 224 bytes");
LYAP_STATIC_ASSERT(sizeof(LyapRGBA) == 4, "RGBA is 4 bytes");
LYAP_STATIC_ASSERT(offsetof(LyapCam, Q) == 16, "Q at 16");
LYAP_STATIC_ASSERT(offsetof(LyapCam, M) == 32, "M at 32");
LYAP_STATIC_ASSERT(offsetof(LyapCam, V) == 36, "V at 36");
LYAP_STATIC_ASSERT(offsetof(LyapCam, S0) == 48, "S0 at 48");
LYAP_STATIC_ASSERT(offsetof(LyapCam, SDX) == 60, "SDX at 60");
LYAP_STATIC_ASSERT(offsetof(LyapCam, SDY) == 72, "SDY at 72");
LYAP_STATIC_ASSERT(offsetof(LyapCam, textureWidth) == 84, "textureWidth at 84");
LYAP_STATIC_ASSERT(offsetof(LyapCam, lightInnerCone) == 104, "lightInnerCone at 104");
LYAP_STATIC_ASSERT(offsetof(LyapCam, ambient) == 128, "ambient at 128");
LYAP_STATIC_ASSERT(offsetof(LyapCam, diffuseColor) == 144, "diffuseColor at 144");
LYAP_STATIC_ASSERT(offsetof(LyapCam, diffusePower) == 160, "diffusePower at 160");
LYAP_STATIC_ASSERT(offsetof(LyapCam, specularColor) == 176, "specularColor at 176");
LYAP_STATIC_ASSERT(offsetof(LyapCam, specularHardness) == 196, "specularHardness at 196");
LYAP_STATIC_ASSERT(offsetof(LyapCam, chaosColor) == 208, "chaosColor at 208");

#endif /* LYAP_TYPES_H */
