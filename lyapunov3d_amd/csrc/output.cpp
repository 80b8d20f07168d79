/*
 * output.cpp - the reference's on-disk formats.
 *
 *  - exps.raw: the lambda volume as raw native-endian FP32, x fastest
 *    (lyap_calculate.cu:79-86);
 *  - Points_*.raw: the LyapPoint array of a frame, raw (lyap_interactive.cu:616-655);
 *  - Render_*.ppm: plain-text P3 of the frame's RGB (lyap_interactive.cu:569-613);
 *  - the file names that carry the frame's parameters
 *    (lyap_interactive.cu:577-593, 601-602).
 * The writers take host buffers; lyap_copy_to_host fetches device outputs.
 */
#include <cstdio>
#include <cstring>

#include "lyap.h"

extern "C" {

int lyap_write_volume_raw(const char *path, const float *exps, size_t n)
{
    if (!path || (n && !exps)) return LYAP_ERR_INVALID_ARGUMENT;
    FILE *f = std::fopen(path, "wb");
    if (!f) return LYAP_ERR_INVALID_ARGUMENT;
    const size_t w = n ? std::fwrite(exps, sizeof(float), n, f) : 0;
    const int rc = std::fclose(f);
    return (w == n && rc == 0) ? LYAP_OK : LYAP_ERR_INVALID_ARGUMENT;
}

int lyap_write_points_raw(const char *path, const LyapPoint *points, size_t n)
{
    if (!path || (n && !points)) return LYAP_ERR_INVALID_ARGUMENT;
    FILE *f = std::fopen(path, "wb");
    if (!f) return LYAP_ERR_INVALID_ARGUMENT;
    const size_t w = n ? std::fwrite(points, sizeof(LyapPoint), n, f) : 0;
    const int rc = std::fclose(f);
    return (w == n && rc == 0) ? LYAP_OK : LYAP_ERR_INVALID_ARGUMENT;
}

int lyap_write_ppm(const char *path, const LyapRGBA *rgba, unsigned w, unsigned h)
{
    if (!path || (w && h && !rgba)) return LYAP_ERR_INVALID_ARGUMENT;
    FILE *f = std::fopen(path, "w");
    if (!f) return LYAP_ERR_INVALID_ARGUMENT;
    bool ok = std::fprintf(f, "P3\n%u %u\n%d\n", w, h, 255) > 0;
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n && ok; ++i)
        ok = std::fprintf(f, "%3d %3d %3d ", rgba[i].r, rgba[i].g, rgba[i].b) > 0;
    ok = ok && std::fprintf(f, "\n") > 0;
    const int rc = std::fclose(f);
    return (ok && rc == 0) ? LYAP_OK : LYAP_ERR_INVALID_ARGUMENT;
}

int lyap_output_leaf(char *out, size_t cap, const char *kind, long stamp, unsigned w,
                     unsigned h, const char *sequence, const LyapCam *cam,
                     const LyapParams *prm)
{
    if (!out || !kind || !sequence || !cam || !prm) return LYAP_ERR_INVALID_ARGUMENT;
    const int n = std::snprintf(
        out, cap, "%s_%ld_%dx%d_%s_cx=%.8g_cy=%.8g_cz=%.8g_step=%d_D=%g_i=%d,%d_d=%d_j=%g_r=%d_ot=%g",
        kind, stamp, (int)w, (int)h, sequence, (double)cam->C.x, (double)cam->C.y,
        (double)cam->C.z, (int)prm->stepMethod, (double)prm->d, (int)prm->settle,
        (int)prm->accum, (int)(unsigned)prm->depth, (double)prm->jitter,
        (int)(unsigned)prm->refine, (double)prm->opaqueThreshold);
    return (n >= 0 && (size_t)n < cap) ? LYAP_OK : LYAP_ERR_INVALID_ARGUMENT;
}

int lyap_output_name(char *out, size_t cap, const char *leaf, unsigned hours,
                     unsigned minutes, unsigned seconds, const char *ext)
{
    if (!out || !leaf || !ext) return LYAP_ERR_INVALID_ARGUMENT;
    const int n = std::snprintf(out, cap, "%s_time=%dh%02dm%02ds.%s", leaf, (int)hours,
                                (int)minutes, (int)seconds, ext);
    return (n >= 0 && (size_t)n < cap) ? LYAP_OK : LYAP_ERR_INVALID_ARGUMENT;
}

} /* extern "C" */
