/*
 * editor.cpp - headless parameter editing and the scripted fly-through.
 *
 * The reference edits its scene from GLUT callbacks: the keyboard
 * (lyap_interactive.cu:311-463) picks a control mode or steps a parameter,
 * and the spaceball's rotation and pan (lyap_interactive.cu:157-309) move the
 * current camera or light, or nudge the parameter of the current mode; the
 * display loop then recalculates what changed (lyap_interactive.cu:118-137).
 * LyapEditor holds that state without a window so a program (or the request
 * server) can drive the same edits and re-render.
 *
 * The fly-through restates scale.pl: frames at t = 0, 0.01, ... 1, eased by
 * easeInOutQuart (scale.pl:5-11), with the camera of scale.pl:33-50 for each
 * eased value. The reference rebuilds and relaunches its viewer for every
 * frame; here a frame is one camera, so a whole fly-through is one batch.
 *
 * Differences from the reference, as in scene.cpp: quaternion normalisation
 * divides by |q| (the reference multiplies, quat.hpp:147-165).
 */
#include <cmath>
#include <cstring>

#include "lyap.h"

namespace {

enum Mode {
    MODE_LOCK = 0, MODE_ROTPAN, MODE_MAG, MODE_NTHRESH, MODE_OTHRESH,
    MODE_CTHRESH, MODE_D, MODE_JITTER, MODE_L_PAN, MODE_L_RANGE, MODE_L_COLOR
};

/* Vec::spaceball_soften (vec3.hpp:272-285): a dead zone of +-lim, then a
 * linear ramp reaching `scale` at +-range. */
float soften1(float v, float lim, float range, float scale)
{
    if (v <= -lim) return scale * (v + lim) / (range - lim);
    if (v >= lim) return scale * (v - lim) / (range - lim);
    return 0.0f;
}

LyapVec soften(float x, float y, float z, float lim, float range, float scale)
{
    return {soften1(x, lim, range, scale), soften1(y, lim, range, scale),
            soften1(z, lim, range, scale)};
}

/* Vec::dominantize (vec3.hpp:173-188): keep the largest component. */
void dominantize(LyapVec &v)
{
    const float ax = std::fabs(v.x), ay = std::fabs(v.y), az = std::fabs(v.z);
    if (ax > ay) {
        if (ax > az) v.y = v.z = 0.0f;
        else v.x = v.y = 0.0f;
    } else {
        if (ay > az) v.x = v.z = 0.0f;
        else v.x = v.y = 0.0f;
    }
}

float clamp01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

/* Hamilton product (quat.hpp:74-80). */
LyapQuat qmul(const LyapQuat &a, const LyapQuat &b)
{
    LyapQuat r;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x;
    r.z = a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}

void qnormalize(LyapQuat &q)
{
    const double m2 = (double)q.x * q.x + (double)q.y * q.y + (double)q.z * q.z + (double)q.w * q.w;
    if (m2 < 1e-6) {
        q.x = q.y = q.z = 0.0f;
        q.w = 1.0f;
        return;
    }
    if (std::fabs(m2 - 1.0) < 1e-6) return;
    const double m = std::sqrt(m2);
    q.x = (float)(q.x / m);
    q.y = (float)(q.y / m);
    q.z = (float)(q.z / m);
    q.w = (float)(q.w / m);
}

/* QUAT::nlerp (quat.hpp:172-192). */
LyapQuat nlerp(const LyapQuat &a, const LyapQuat &b, float t)
{
    if (t == 0.0f || t < 1e-6f) return a;
    if (t == 1.0f || t > 1.0f - 1e-6f) return b;
    const float dot = a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
    const float tA = dot >= 0 ? t : -t;
    const float tI = 1.0f - t;
    LyapQuat r;
    r.x = a.x * tI + b.x * tA;
    r.y = a.y * tI + b.y * tA;
    r.z = a.z * tI + b.z * tA;
    r.w = a.w * tI + b.w * tA;
    qnormalize(r);
    return r;
}

LyapVec vnormalized(LyapVec v)
{
    const float m2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if ((double)m2 < 1e-12) return {0.f, 0.f, 0.f};
    if (m2 == 1.0f) return v;
    const float m = std::sqrt(m2);
    return {v.x / m, v.y / m, v.z / m};
}

LyapCam *current(LyapEditor *ed)
{
    return ed->cur_light >= 0 ? &ed->lights[ed->cur_light] : &ed->cam;
}

} /* namespace */

extern "C" {

int lyap_editor_init(LyapEditor *ed, unsigned window_w, unsigned window_h)
{
    if (!ed) return LYAP_ERR_INVALID_ARGUMENT;
    std::memset(ed, 0, sizeof(*ed));
    lyap_params_init(&ed->prm, &ed->cam, ed->lights, &ed->num_lights);
    ed->home_light = ed->lights[0];
    ed->mode = MODE_ROTPAN;
    ed->cur_light = -1;
    ed->light_index = 0;
    ed->dominant = 0;
    ed->window_w = window_w;
    ed->window_h = window_h;
    ed->render_denominator = 1;
    ed->pending = LYAP_UPDATE_CALC | LYAP_UPDATE_LIGHTS | LYAP_UPDATE_CAM;
    return LYAP_OK;
}

unsigned lyap_editor_key(LyapEditor *ed, int key)
{
    if (!ed) return 0;
    unsigned f = 0;
    LyapParams &p = ed->prm;
    switch (key) {
    case '1': case '2': case '3': case '4': case '5':
    case '6': case '7': case '8': case '9':
        ed->light_index = key - '1';
        ed->cur_light = ed->light_index;
        ed->mode = MODE_ROTPAN;
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_RENDER | LYAP_UPDATE_LIGHTS;
        break;
    case '0':
        ed->cur_light = -1;
        ed->mode = MODE_ROTPAN;
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_CALC;
        break;
    case 'l': case 'L': ed->mode = MODE_L_PAN; f = LYAP_UPDATE_LIGHTS; break;
    case 'r': case 'R': ed->mode = MODE_L_RANGE; f = LYAP_UPDATE_LIGHTS; break;
    case 'o': case 'O': ed->mode = MODE_L_COLOR; f = LYAP_UPDATE_LIGHTS; break;
    case 'd': case 'D': ed->mode = MODE_D; break;
    case 'j': case 'J': ed->mode = MODE_JITTER; break;
    case 't': case 'T': ed->mode = MODE_OTHRESH; break;
    case 'y': case 'Y': ed->mode = MODE_NTHRESH; break;
    case 'c': case 'C': ed->mode = MODE_CTHRESH; break;
    case 'm': case 'M': ed->mode = MODE_MAG; break;
    case '[': case '{': p.stepMethod = 1; f = LYAP_UPDATE_CALC; break;
    case ']': case '}': p.stepMethod = 2; f = LYAP_UPDATE_CALC; break;
    case 'z':
        p.depth -= 1;
        if (p.depth < 1.0f) p.depth = 1.0f;
        f = LYAP_UPDATE_CALC;
        break;
    case 'Z':
        p.depth /= 2;
        if (p.depth < 1.0f) p.depth = 1.0f;
        f = LYAP_UPDATE_CALC;
        break;
    case 'a': p.depth += 1; f = LYAP_UPDATE_CALC; break;
    case 'A': p.depth *= 2; f = LYAP_UPDATE_CALC; break;
    case 'h': case 'H': {
        /* the home light at the camera (lyap_interactive.cu:397-406) */
        LyapLight &L = ed->lights[ed->light_index];
        L = ed->home_light;
        L.C = ed->cam.C;
        L.Q = ed->cam.Q;
        L.M = ed->cam.M;
        lyap_scene_cam_recalculate(&L, ed->window_w, ed->window_h, ed->render_denominator);
        f = LYAP_UPDATE_LIGHTS | LYAP_UPDATE_RENDER;
        break;
    }
    case 'x':
        /* an unsigned count: 0 after the decrement goes back to 1 */
        p.accum--;
        if (p.accum < 1) p.accum = 1;
        f = LYAP_UPDATE_CALC;
        break;
    case 'X':
        p.accum /= 2;
        if (p.accum < 1) p.accum = 1;
        f = LYAP_UPDATE_CALC;
        break;
    case 's': p.accum++; f = LYAP_UPDATE_CALC; break;
    case 'S': p.accum *= 2; f = LYAP_UPDATE_CALC; break;
    case 'q': case 'Q': f = LYAP_UPDATE_QUIT; break;
    case '-': case '_':
        ed->render_denominator++;
        f = LYAP_UPDATE_CALC;
        break;
    case '+': case '=':
        if (ed->render_denominator > 1) {
            ed->render_denominator--;
            f = LYAP_UPDATE_CALC;
        } else {
            ed->render_denominator = 1;
        }
        break;
    case ' ': ed->mode = MODE_ROTPAN; break;
    default: break;
    }
    ed->pending |= f & ~LYAP_UPDATE_QUIT;
    return f;
}

unsigned lyap_editor_spaceball_button(LyapEditor *ed, int button)
{
    if (!ed) return 0;
    if (button == 1) ed->dominant = !ed->dominant;
    return 0;
}

unsigned lyap_editor_spaceball_rot(LyapEditor *ed, int x, int y, int z)
{
    if (!ed) return 0;
    unsigned f = 0;
    LyapParams &p = ed->prm;
    LyapCam *c = current(ed);
    const double fx = (double)(float)x;
    switch (ed->mode) {
    case MODE_ROTPAN: {
        LyapVec axis = soften((float)x, (float)-y, (float)z, 50.0f, 500.0f, 1.0f);
        const float len = std::sqrt(axis.x * axis.x + axis.y * axis.y + axis.z * axis.z);
        if (len < 1e-6f) return 0;
        axis = {axis.x / len, axis.y / len, axis.z / len};
        if (ed->dominant) dominantize(axis);
        LyapQuat rq;
        lyap_quat_axis_angle(&axis, len * 0.05f, 0, &rq);
        c->Q = qmul(c->Q, rq);
        qnormalize(c->Q);
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_CALC;
        break;
    }
    case MODE_MAG:
        c->M = (float)((double)c->M + fx / 2500.0);
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_CALC;
        break;
    case MODE_NTHRESH:
        p.nearThreshold = (float)((double)p.nearThreshold + fx / 10000.0);
        if (p.opaqueThreshold > p.nearThreshold) p.nearThreshold = p.opaqueThreshold;
        f = LYAP_UPDATE_CALC;
        break;
    case MODE_OTHRESH:
        p.opaqueThreshold = (float)((double)p.opaqueThreshold + fx / 10000.0);
        if (p.opaqueThreshold > p.nearThreshold) p.opaqueThreshold = p.nearThreshold;
        f = LYAP_UPDATE_CALC;
        break;
    case MODE_CTHRESH:
        p.chaosThreshold = (float)((double)p.chaosThreshold + fx / 10000.0);
        f = LYAP_UPDATE_CALC;
        break;
    case MODE_D:
        p.d = (float)((double)p.d + fx / 10000.0);
        if (p.d < p.lMin) p.d = p.lMin;
        else if (p.d > p.lMax) p.d = p.lMax;
        f = LYAP_UPDATE_CALC;
        break;
    case MODE_JITTER:
        p.jitter = (float)((double)p.jitter + fx / 10000.0);
        if (p.jitter < 0) p.jitter = 0;
        else if (p.jitter > 1.0f) p.jitter = 1.0f;
        f = LYAP_UPDATE_CALC;
        break;
    default:
        break;   /* MODE_L_RANGE is disabled in the reference */
    }
    ed->pending |= f;
    return f;
}

unsigned lyap_editor_spaceball_pan(LyapEditor *ed, int x, int y, int z)
{
    if (!ed) return 0;
    unsigned f = 0;
    LyapCam *c = current(ed);
    switch (ed->mode) {
    case MODE_ROTPAN: {
        LyapVec d = soften((float)x, (float)-y, (float)z, 50.0f, 500.0f, 0.05f);
        if (ed->dominant) dominantize(d);
        LyapVec w;
        lyap_quat_transform(&c->Q, &d, &w);
        c->C = {c->C.x + w.x, c->C.y + w.y, c->C.z + w.z};
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_LIGHTS | LYAP_UPDATE_CALC;
        break;
    }
    case MODE_L_PAN: {
        if ((unsigned)ed->light_index + 1 > ed->num_lights) return 0;
        LyapVec d = soften((float)x, (float)-y, (float)z, 50.0f, 500.0f, 0.25f);
        if (ed->dominant) dominantize(d);
        LyapVec w;
        lyap_quat_transform(&c->Q, &d, &w);
        LyapLight &L = ed->lights[ed->light_index];
        L.C = {L.C.x + w.x, L.C.y + w.y, L.C.z + w.z};
        f = LYAP_UPDATE_CAM | LYAP_UPDATE_LIGHTS | LYAP_UPDATE_RENDER;
        break;
    }
    case MODE_L_COLOR: {
        if ((unsigned)ed->light_index + 1 > ed->num_lights) return 0;
        LyapVec d = soften((float)x, (float)-y, (float)z, 50.0f, 500.0f, 0.25f);
        if (ed->dominant) dominantize(d);
        LyapLight &L = ed->lights[ed->light_index];
        L.diffuseColor.x = clamp01(L.diffuseColor.x + d.x);
        L.diffuseColor.y = clamp01(L.diffuseColor.y + d.y);
        L.diffuseColor.z = clamp01(L.diffuseColor.z + d.z);
        L.specularColor = L.diffuseColor;
        f = LYAP_UPDATE_LIGHTS | LYAP_UPDATE_RENDER;
        break;
    }
    default:
        break;
    }
    ed->pending |= f;
    return f;
}

unsigned lyap_editor_commit(LyapEditor *ed, unsigned image_w, unsigned image_h)
{
    if (!ed) return 0;
    const unsigned f = ed->pending;
    if (f & LYAP_UPDATE_LIGHTS) lyap_scene_lights_recalculate(ed->lights, ed->num_lights);
    if (f & LYAP_UPDATE_CAM)
        lyap_scene_cam_recalculate(&ed->cam, image_w, image_h, ed->render_denominator);
    ed->pending = 0;
    return f;
}

double lyap_ease_in_out_quart(double t, double b, double c, double d)
{
    t /= d / 2;
    if (t < 1) return c / 2 * t * t + b;
    t--;
    return -c / 2 * (t * (t - 2) - 1) + b;
}

unsigned lyap_flythrough_frames(double step, double *t_out, unsigned cap)
{
    /* for ($t = 0; $t <= 1.00001; $t += step): t accumulates in double */
    unsigned n = 0;
    if (!(step > 0)) return 0;
    for (double t = 0; t <= 1.00001; t += step) {
        if (t_out && n < cap) t_out[n] = t;
        ++n;
    }
    return n;
}

void lyap_flythrough_camera(double eased, LyapCam *cam)
{
    const float i = (float)eased;
    const LyapVec dir = {4.0f, 4.0f, 4.0f};
    const LyapVec side = vnormalized({-4.0f, 4.0f, 4.0f});
    const LyapVec dn = vnormalized(dir);
    LyapVec up = {side.y * dn.z - side.z * dn.y, side.z * dn.x - side.x * dn.z,
                  side.x * dn.y - side.y * dn.x};
    up = vnormalized(up);
    LyapQuat rot0, rot1;
    lyap_quat_axis_angle(&up, -20.0f, 1, &rot0);
    lyap_quat_axis_angle(&up, 20.0f, 1, &rot1);
    const LyapQuat nrot = nlerp(rot0, rot1, i);
    LyapVec nd;
    lyap_quat_transform(&nrot, &dir, &nd);
    nd = vnormalized(nd);
    nd = {-nd.x, -nd.y, -nd.z};
    /* the script pastes the eased value into C++ source as a double literal */
    const float c0 = (float)(4.0 - 0.9 * eased);
    cam->C = {c0 - nd.x, c0 - nd.y, c0 - nd.z};
    const LyapVec zhat = {0.0f, 0.0f, 1.0f};
    lyap_quat_between(&zhat, &nd, 1.0f, &cam->Q);
}

} /* extern "C" */
