/*
 * scene.cpp - host-side scene set-up of the Lyapunov path.
 *
 * Restates params_init (params.cu:21-114), scene_convert_sequence
 * (scene.cu:69-108), scene_cam_recalculate (scene.cu:31-63),
 * scene_lights_recalculate (scene.cu:20-29) and the quaternion helpers they
 * use (quat.hpp:147-165, 194-238, 279-284). Differences from the reference:
 *  - errors are returned (a bad sequence letter exits the reference process);
 *  - quaternion normalisation divides by |q| (the reference multiplies,
 *    quat.hpp:159-163, which is only harmless for already-unit q);
 *  - quaternion rotation is evaluated in double and rounded once, so results
 *    can differ from the reference's float expression in the last bit;
 *  - a zero renderDenominator is treated as 1 (the reference divides by it).
 */
#include <cmath>
#include <cstring>

#include "lyap.h"

namespace {

struct D3 { double x, y, z; };

inline D3 to_d(const LyapVec &v) { return {v.x, v.y, v.z}; }
inline LyapVec to_f(const D3 &v) { return {(float)v.x, (float)v.y, (float)v.z}; }

/* Rotation of v by q = (x, y, z, w): the rotation matrix of q scaled by |q|^2,
 * the value QUAT::transform computes (quat.hpp:279-284). */
D3 rotate(const LyapQuat &q, const D3 &v)
{
    const double x = q.x, y = q.y, z = q.z, w = q.w;
    const double xx = x * x, yy = y * y, zz = z * z, ww = w * w;
    return {
        (ww + xx - yy - zz) * v.x + 2.0 * (x * y - w * z) * v.y + 2.0 * (x * z + w * y) * v.z,
        2.0 * (x * y + w * z) * v.x + (ww - xx + yy - zz) * v.y + 2.0 * (y * z - w * x) * v.z,
        2.0 * (x * z - w * y) * v.x + 2.0 * (y * z + w * x) * v.y + (ww - xx - yy + zz) * v.z,
    };
}

/* Vec::normalize on a float vector (vec3.hpp:141-154). */
LyapVec vnormalized(LyapVec v)
{
    const float m2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if ((double)m2 < 1e-12) return {0.f, 0.f, 0.f};
    if (m2 == 1.0f) return v;
    const float m = std::sqrt(m2);
    return {v.x / m, v.y / m, v.z / m};
}

float vdot(const LyapVec &a, const LyapVec &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

/* QUAT::normalize (quat.hpp:147-165) with the division the code intends. */
void qnormalize(LyapQuat *q)
{
    const double m2 = (double)q->x * q->x + (double)q->y * q->y +
                      (double)q->z * q->z + (double)q->w * q->w;
    if (m2 < 1e-6) {
        q->x = q->y = q->z = 0.0f;
        q->w = 1.0f;
        return;
    }
    if (std::fabs(m2 - 1.0) < 1e-6) return;
    const double m = std::sqrt(m2);
    q->x = (float)(q->x / m);
    q->y = (float)(q->y / m);
    q->z = (float)(q->z / m);
    q->w = (float)(q->w / m);
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
    *out = to_f(rotate(*q, to_d(*v)));
}

void lyap_quat_axis_angle(const LyapVec *axis, float angle, int in_degrees,
                          LyapQuat *out)
{
    const double half = in_degrees ? (double)angle * M_PI / 360.0
                                   : (double)angle * 0.5;
    const double s = std::sin(half);
    out->x = (float)(axis->x * s);
    out->y = (float)(axis->y * s);
    out->z = (float)(axis->z * s);
    out->w = (float)std::cos(half);
    qnormalize(out);
}

void lyap_quat_between(const LyapVec *p, const LyapVec *q, float scale,
                       LyapQuat *out)
{
    double c = (double)p->x * q->x + (double)p->y * q->y + (double)p->z * q->z;
    if (c < -1.0) c = -1.0;
    else if (c > 1.0) c = 1.0;
    /* The reference returns the identity when the vectors are (nearly)
     * perpendicular, |cos| <= 1e-6 (quat.hpp:204-208); kept as written. */
    if (c >= -1e-6 && c <= 1e-6) {
        out->x = out->y = out->z = 0.0f;
        out->w = 1.0f;
        return;
    }
    const double ang = std::acos(c);
    const D3 axis = {(double)p->y * q->z - (double)p->z * q->y,
                     (double)p->z * q->x - (double)p->x * q->z,
                     (double)p->x * q->y - (double)p->y * q->x};
    const double k = std::sin(ang * 0.5 * scale) / std::sin(ang);
    out->x = (float)(axis.x * k);
    out->y = (float)(axis.y * k);
    out->z = (float)(axis.z * k);
    out->w = (float)std::cos(ang * 0.5 * scale);
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
    const double M = cam->M;
    cam->V = vnormalized(to_f(rotate(cam->Q, {0.0, 0.0, 1.0})));
    cam->S0 = to_f(rotate(cam->Q, {-M, -M, 1.0}));
    cam->lightInnerCone = vdot(cam->V, vnormalized(to_f(rotate(cam->Q, {-M, -M, 1.5}))));
    cam->lightOuterCone = vdot(cam->V, vnormalized(to_f(rotate(cam->Q, {-M, -M, 1.0}))));
    const float sx = 2.0f * cam->M / (float)(cam->renderWidth ? cam->renderWidth : 1);
    const float sy = 2.0f * cam->M / (float)(cam->renderHeight ? cam->renderHeight : 1);
    cam->SDX = to_f(rotate(cam->Q, {sx, 0.0, 0.0}));
    cam->SDY = to_f(rotate(cam->Q, {0.0, sy, 0.0}));
}

void lyap_scene_lights_recalculate(LyapLight *lights, size_t num_lights)
{
    for (size_t i = 0; i < num_lights; ++i) {
        LyapLight *L = &lights[i];
        const double M = L->M;
        L->V = vnormalized(to_f(rotate(L->Q, {0.0, 0.0, 1.0})));
        L->lightInnerCone = vdot(L->V, vnormalized(to_f(rotate(L->Q, {-M, -M, 1.5}))));
        L->lightOuterCone = vdot(L->V, vnormalized(to_f(rotate(L->Q, {-M, -M, 1.0}))));
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
        LyapVec up = {side.y * dn.z - side.z * dn.y, side.z * dn.x - side.x * dn.z,
                      side.x * dn.y - side.y * dn.x};
        up = vnormalized(up);
        LyapQuat rot;
        lyap_quat_axis_angle(&up, 20.0f, 1, &rot);
        LyapVec nd = vnormalized(to_f(rotate(rot, to_d(dir))));
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
    D3 f = {(double)target->x - eye->x, (double)target->y - eye->y,
            (double)target->z - eye->z};
    const double fl = std::sqrt(f.x * f.x + f.y * f.y + f.z * f.z);
    LyapQuat q;
    q.x = q.y = q.z = 0.0f;
    q.w = 1.0f;
    if (fl > 0.0) {
        f = {f.x / fl, f.y / fl, f.z / fl};
        /* shortest-arc rotation of +z onto f: q = (z x f, 1 + z.f), normalised */
        double qx = -f.y, qy = f.x, qz = 0.0, qw = 1.0 + f.z;
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
