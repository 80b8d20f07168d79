/*
 * scene.cpp - host-side scene set-up of the Lyapunov path.
 *
 * Restates params_init (params.cu:21-114), scene_convert_sequence
 * (scene.cu:69-108), scene_cam_recalculate (scene.cu:31-63),
 * scene_lights_recalculate (scene.cu:20-29) and the quaternion helpers they
 * use (quat.hpp:147-165, 194-238, 279-284), in float with the reference's
 * operation order, so the camera vectors (V, S0, SDX, SDY) and the cone
 * cosines are bit for bit the reference's for the same inputs (this file is
 * built with -ffp-contract=off: no multiply-add is fused). Differences from
 * the reference:
 *  - errors are returned (a bad sequence letter exits the reference process);
 *  - a zero renderDenominator is treated as 1 (the reference divides by it).
 * QUAT::normalize is restated with its quirk: outside a 1e-6 band around
 * |q|^2 = 1 it multiplies by |q| (quat.hpp:159-163) instead of dividing;
 * callers keep it a no-op by passing unit quaternions.
 */
#include <cmath>
#include <cstring>

#include "lyap.h"

namespace {

constexpr double kQuatEpsilon = 1e-6;        /* quat.hpp:12 */
constexpr double kVecEpsilon = 1e-12;        /* vec3.hpp:10 */
constexpr double kQuatPi = 3.14159265358979323846264338327950288;

/* QUAT::transform (quat.hpp:279-284): each component is a sum of eight
 * float products, accumulated in the order the reference writes them
 * (left to right), each product formed left to right too. */
LyapVec qtransform(const LyapQuat &q, const LyapVec &v)
{
    const float x = q.x, y = q.y, z = q.z, w = q.w, two = 2.0f;
    float ox = w * w * v.x;
    ox += two * y * w * v.z;
    ox -= two * z * w * v.y;
    ox += x * x * v.x;
    ox += two * y * x * v.y;
    ox += two * z * x * v.z;
    ox -= z * z * v.x;
    ox -= y * y * v.x;
    float oy = two * x * y * v.x;
    oy += y * y * v.y;
    oy += two * z * y * v.z;
    oy += two * w * z * v.x;
    oy -= z * z * v.y;
    oy += w * w * v.y;
    oy -= two * x * w * v.z;
    oy -= x * x * v.y;
    float oz = two * x * z * v.x;
    oz += two * y * z * v.y;
    oz += z * z * v.z;
    oz -= two * w * y * v.x;
    oz -= y * y * v.z;
    oz += two * w * x * v.y;
    oz -= x * x * v.z;
    oz += w * w * v.z;
    return {ox, oy, oz};
}

/* Vec::normalized (vec3.hpp:128-139): float |v|^2, zero below 1e-12, left
 * alone within 1e-12 of 1 (only exactly 1 in float). */
LyapVec vnormalized(LyapVec v)
{
    const float m2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if ((double)m2 < kVecEpsilon) return {0.f, 0.f, 0.f};
    if (m2 == 1.0f || ((double)m2 > 1.0 - kVecEpsilon && (double)m2 < 1.0 + kVecEpsilon)) return v;
    const float m = std::sqrt(m2);
    return {v.x / m, v.y / m, v.z / m};
}

float vdot(const LyapVec &a, const LyapVec &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

/* Vec operator* (vec3.hpp:112-115): the cross product. */
LyapVec vcross(const LyapVec &a, const LyapVec &b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

/* QUAT::normalize (quat.hpp:147-165): float |q|^2; the identity below 1e-6;
 * unchanged within 1e-6 of 1; otherwise each component times |q|. */
void qnormalize(LyapQuat *q)
{
    float m2 = q->x * q->x + q->y * q->y + q->z * q->z + q->w * q->w;
    if ((double)m2 < kQuatEpsilon) {
        q->x = q->y = q->z = 0.0f;
        q->w = 1.0f;
        return;
    }
    if (m2 == 1.0f || ((double)m2 > 1.0 - kQuatEpsilon && (double)m2 < 1.0 + kQuatEpsilon)) return;
    m2 = std::sqrt(m2);
    q->x *= m2;
    q->y *= m2;
    q->z *= m2;
    q->w *= m2;
}

LyapColor color(float r, float g, float b, float a)
{
    LyapColor c;
    c.x = r; c.y = g; c.z = b; c.w = a;
    return c;
}

} /* namespace */

extern "C" {

void lyap_quat_transform(const LyapQuat *q, const LyapVec *v, LyapVec *out)
{
    *out = qtransform(*q, *v);
}

/* QUAT(axis, ang, inDegrees) (quat.hpp:227-238): the half angle formed in
 * double from the double pi and stored as float, float sin/cos. */
void lyap_quat_axis_angle(const LyapVec *axis, float angle, int in_degrees,
                          LyapQuat *out)
{
    const float half = in_degrees ? (float)((double)angle * kQuatPi / (double)360.0f) : angle * 0.5f;
    const float s = std::sin(half);
    out->x = axis->x * s;
    out->y = axis->y * s;
    out->z = axis->z * s;
    out->w = std::cos(half);
    qnormalize(out);
}

/* QUAT(p, q, scale) (quat.hpp:194-220), in float; the reference returns the
 * identity when the vectors are (nearly) perpendicular, |cos| <= 1e-6
 * (quat.hpp:204-208): kept as written. */
void lyap_quat_between(const LyapVec *p, const LyapVec *q, float scale,
                       LyapQuat *out)
{
    float c = vdot(*p, *q);
    if ((double)c < -1.0) c = -1.0f;
    else if ((double)c > 1.0) c = 1.0f;
    if (c == 0.0f || ((double)c >= -kQuatEpsilon && (double)c <= kQuatEpsilon)) {
        out->x = out->y = out->z = 0.0f;
        out->w = 1.0f;
        return;
    }
    const float ang = std::acos(c);
    const LyapVec axis = vcross(*p, *q);
    const float half = (float)((double)ang * 0.5 * (double)scale);
    const float k = std::sin(half) / std::sin(ang);
    out->x = axis.x * k;
    out->y = axis.y * k;
    out->z = axis.z * k;
    out->w = std::cos(half);
}

void lyap_scene_cam_recalculate(LyapCam *cam, unsigned tw, unsigned th,
                                unsigned td)
{
    if (cam->M < 1e-6f) cam->M = 1e-6f;
    qnormalize(&cam->Q);
    if (td > 0) cam->renderDenominator = td;
    if (cam->renderDenominator == 0) cam->renderDenominator = 1;
    if (tw > 0) {
        cam->textureWidth = tw;
        cam->renderWidth = tw / cam->renderDenominator;
    }
    if (th > 0) {
        cam->textureHeight = th;
        cam->renderHeight = th / cam->renderDenominator;
    }
    const float M = cam->M;
    cam->V = vnormalized(qtransform(cam->Q, {0.0f, 0.0f, 1.0f}));
    cam->S0 = qtransform(cam->Q, {-M, -M, 1.0f});
    cam->lightInnerCone = vdot(cam->V, vnormalized(qtransform(cam->Q, {-M, -M, 1.5f})));
    cam->lightOuterCone = vdot(cam->V, vnormalized(qtransform(cam->Q, {-M, -M, 1.0f})));
    /* Vec(2*M / (Real)renderWidth, 0, 0) (scene.cu:57-62) */
    const float sx = 2 * M / (float)(cam->renderWidth ? cam->renderWidth : 1);
    const float sy = 2 * M / (float)(cam->renderHeight ? cam->renderHeight : 1);
    cam->SDX = qtransform(cam->Q, {sx, 0.0f, 0.0f});
    cam->SDY = qtransform(cam->Q, {0.0f, sy, 0.0f});
}

void lyap_scene_lights_recalculate(LyapLight *lights, size_t num_lights)
{
    for (size_t i = 0; i < num_lights; ++i) {
        LyapLight *L = &lights[i];
        const float M = L->M;
        L->V = vnormalized(qtransform(L->Q, {0.0f, 0.0f, 1.0f}));
        L->lightInnerCone = vdot(L->V, vnormalized(qtransform(L->Q, {-M, -M, 1.5f})));
        L->lightOuterCone = vdot(L->V, vnormalized(qtransform(L->Q, {-M, -M, 1.0f})));
    }
}

int lyap_scene_convert_sequence(const char *str, int32_t *out, size_t cap,
                                size_t *out_len)
{
    if (!str || !out_len) return LYAP_ERR_INVALID_ARGUMENT;
    /* The reference inspects the first character even when the string is
     * empty and rejects the terminating NUL as a bad letter. */
    if (*str == '\0') return LYAP_ERR_BAD_SEQUENCE;
    size_t n = 0;
    int32_t last = 1;              /* 'B' precedes the first letter */
    bool overflow = false;
    auto push = [&](int32_t v) {
        if (out && n < cap) out[n] = v;
        else overflow = true;
        ++n;
    };
    for (const char *c = str; *c; ++c) {
        const char ch = *c;
        if (ch >= '1' && ch <= '9') {
            for (int k = 0; k < ch - '0'; ++k) push(last);
        } else if ((ch >= 'A' && ch <= 'D') || (ch >= 'a' && ch <= 'd')) {
            last = (ch >= 'a') ? ch - 'a' : ch - 'A';
            push(last);
        } else {
            return LYAP_ERR_BAD_SEQUENCE;
        }
    }
    push(-1);
    *out_len = n;
    return overflow ? LYAP_ERR_INVALID_ARGUMENT : LYAP_OK;
}

void lyap_params_init(LyapParams *prm, LyapCam *cam, LyapLight *lights,
                      unsigned *num_lights)
{
    if (prm) {
        std::memset(prm, 0, sizeof(*prm));
        prm->d = 2.1f;
        prm->settle = 18;
        prm->accum = 1008;
        prm->stepMethod = 2;
        prm->nearThreshold = -1.0f;
        prm->nearMultiplier = 2.0f;
        prm->opaqueThreshold = -0.75f;
        prm->chaosThreshold = -0.5f;
        prm->depth = 4096.0f;
        prm->jitter = 0.5f;
        prm->refine = 32.0f;
        prm->gradient = 0.01f;
        prm->lMin = 0.0f;
        prm->lMax = 4.0f;
    }
    if (cam) {
        /* The scripted camera of params.cu:42-62 at its end point t = 1: the
         * view axis is (4,4,4) turned 20 degrees about the axis
         * normalize(-4,4,4) x normalize(4,4,4), and the camera sits one unit
         * behind (3.1, 3.1, 3.1) along it. */
        std::memset(cam, 0, sizeof(*cam));
        const LyapVec dir = {4.0f, 4.0f, 4.0f};
        LyapVec side = vnormalized({-4.0f, 4.0f, 4.0f});
        const LyapVec dn = vnormalized(dir);
        LyapVec up = vnormalized(vcross(side, dn));
        /* nlerp(rot0, rot1, 1) is rot1 (quat.hpp:178-179) */
        LyapQuat rot;
        lyap_quat_axis_angle(&up, 20.0f, 1, &rot);
        LyapVec nd = vnormalized(qtransform(rot, dir));
        nd = {-nd.x, -nd.y, -nd.z};
        const float c0 = (float)(4.0 - 0.9);
        cam->C = {c0 - nd.x, c0 - nd.y, c0 - nd.z};
        const LyapVec zhat = {0.0f, 0.0f, 1.0f};
        lyap_quat_between(&zhat, &nd, 1.0f, &cam->Q);
        cam->M = 0.45f;
        cam->renderDenominator = 1;
    }
    if (lights) {
        std::memset(lights, 0, sizeof(LyapLight) * LYAP_MAX_LIGHTS);
        LyapLight *L = &lights[0];
        L->C = {6.0f, 5.0f, 3.0f};
        L->Q.x = 0.710595f; L->Q.y = 0.282082f; L->Q.z = -0.512168f; L->Q.w = 0.391368f;
        L->M = 0.5f;
        L->lightInnerCone = 0.904535f;
        L->lightOuterCone = 0.816497f;
        L->lightRange = 1.0f;
        L->ambient = color(0.1f, 0.0f, 0.0f, 0.0f);
        L->diffuseColor = color(1.0f, 0.25f, 0.125f, 1.0f);
        L->diffusePower = 10.0f;
        L->specularColor = color(1.0f, 1.0f, 1.0f, 1.0f);
        L->specularPower = 10.0f;
        L->specularHardness = 10.0f;
        L->chaosColor = color(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (num_lights) *num_lights = 1;
}

const char *lyap_default_sequence(void) { return "BCABA"; }

void lyap_cam_look_at(LyapCam *cam, const LyapVec *eye, const LyapVec *target,
                      float M, unsigned w, unsigned h)
{
    double fx = (double)target->x - eye->x, fy = (double)target->y - eye->y,
           fz = (double)target->z - eye->z;
    const double fl = std::sqrt(fx * fx + fy * fy + fz * fz);
    LyapQuat q;
    q.x = q.y = q.z = 0.0f;
    q.w = 1.0f;
    if (fl > 0.0) {
        fx /= fl;
        fy /= fl;
        fz /= fl;
        /* shortest-arc rotation of +z onto f: q = (z x f, 1 + z.f), normalised
         * in double, so QUAT::normalize leaves it alone */
        double qx = -fy, qy = fx, qz = 0.0, qw = 1.0 + fz;
        if (qw < 1e-12) { qx = 1.0; qy = 0.0; qz = 0.0; qw = 0.0; }  /* f = -z */
        const double n = std::sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
        q.x = (float)(qx / n);
        q.y = (float)(qy / n);
        q.z = (float)(qz / n);
        q.w = (float)(qw / n);
    }
    cam->C = *eye;
    cam->Q = q;
    cam->M = M;
    cam->renderDenominator = 1;
    lyap_scene_cam_recalculate(cam, w, h, 1);
}

} /* extern "C" */
