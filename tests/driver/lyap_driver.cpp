/*
 * lyap_driver.cpp - the reference's two drivers rewritten against lyap.h.
 *
 * A compiled C++ caller of liblyap_amd.so that uses nothing but the C ABI
 * (no HIP header, no torch): the drop-in as a maintainer of the reference
 * would wire it. The GPU tests run it and compare its files with the oracle.
 *
 *   lyap_driver calculate N plain|product|fp64 OUT.raw
 *     The flow of lyap_calculate.cu:58-91: params_init defaults (BCABA,
 *     d = 2.1, settle 18, accum 1008), the sequence converted from its
 *     string, lambda of an N^3 grid over the reference's fixed box [0, 4)^3
 *     (kernel.cu:527-529), copied to the host and written as exps.raw
 *     (raw FP32, x fastest).
 *
 *   lyap_driver interactive W H plain|fp64 POINTS.raw RENDER.ppm
 *     One frame of lyap_interactive.cu: init_scene (:127-137: params_init,
 *     device lights and points, the sequence), update_scene (:118-125:
 *     lights and camera recalculated for a W x H image), render (:696-714:
 *     kernel_calc_render with the reference's method - the refinement walk,
 *     normal offset step * gradient, jitter and near step - then a
 *     synchronise), and the Points/PPM files of :569-655.
 *
 * Exit status 0 on success; on failure the library's status and message
 * are printed and the status is the exit code (negated).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lyap.h"

namespace {

lyap_ctx *g_ctx = nullptr;

int check(int rc, const char *what)
{
    if (rc != LYAP_OK) {
        std::fprintf(stderr, "lyap_driver: %s: %s (%s)\n", what, lyap_strerror(rc),
                     g_ctx ? lyap_last_error(g_ctx) : "no context");
        if (g_ctx) lyap_destroy(g_ctx);
        std::exit(rc < 0 ? -rc : 1);
    }
    return rc;
}

int precision_of(const char *name)
{
    if (!std::strcmp(name, "plain")) return LYAP_PREC_PLAIN;
    if (!std::strcmp(name, "product")) return LYAP_PREC_PRODUCT;
    if (!std::strcmp(name, "fp64")) return LYAP_PREC_FP64;
    std::fprintf(stderr, "lyap_driver: unknown precision %s\n", name);
    std::exit(2);
}

int calculate(unsigned n, int precision, const char *out_path)
{
    LyapParams prm;
    lyap_params_init(&prm, nullptr, nullptr, nullptr);
    int32_t seq[LYAP_MAX_SEQ + 1];
    size_t len = 0;
    check(lyap_scene_convert_sequence(lyap_default_sequence(), seq, LYAP_MAX_SEQ + 1, &len),
          "scene_convert_sequence");
    check(lyap_create(&g_ctx, 0, 0), "lyap_create");
    const size_t count = (size_t)n * n * n;
    const size_t bytes = sizeof(float) * count;
    void *d_exps = nullptr;
    check(lyap_device_alloc(g_ctx, &d_exps, bytes), "device_alloc");
    const float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {4.0f, 4.0f, 4.0f};
    LyapVolumeMode mode;
    std::memset(&mode, 0, sizeof(mode));
    mode.precision = precision;
    check(lyap_calc_volume(g_ctx, (float *)d_exps, &prm, seq, lo, hi, n, n, n, &mode),
          "calc_volume");
    std::vector<float> exps(count);
    check(lyap_copy_to_host(g_ctx, exps.data(), d_exps, bytes), "copy_to_host");
    check(lyap_write_volume_raw(out_path, exps.data(), count), "write_volume_raw");
    std::printf("Points size = %zu\n", bytes);
    check(lyap_device_free(g_ctx, d_exps), "device_free");
    lyap_destroy(g_ctx);
    g_ctx = nullptr;
    return 0;
}

int interactive(unsigned w, unsigned h, int precision, const char *points_path,
                const char *ppm_path)
{
    LyapParams prm;
    LyapCam cam;
    LyapLight lights[LYAP_MAX_LIGHTS];
    unsigned num_lights = 0;
    lyap_params_init(&prm, &cam, lights, &num_lights);
    int32_t seq[LYAP_MAX_SEQ + 1];
    size_t len = 0;
    check(lyap_scene_convert_sequence(lyap_default_sequence(), seq, LYAP_MAX_SEQ + 1, &len),
          "scene_convert_sequence");
    check(lyap_create(&g_ctx, 0, 0), "lyap_create");
    /* init_scene */
    void *d_lights = nullptr, *d_points = nullptr, *d_rgba = nullptr;
    check(lyap_device_alloc(g_ctx, &d_lights, sizeof(LyapLight) * LYAP_MAX_LIGHTS), "alloc lights");
    check(lyap_device_alloc(g_ctx, &d_points, sizeof(LyapPoint) * w * h), "alloc points");
    check(lyap_device_alloc(g_ctx, &d_rgba, sizeof(LyapRGBA) * w * h), "alloc rgba");
    /* update_scene */
    lyap_scene_lights_recalculate(lights, num_lights);
    check(lyap_copy_to_device(g_ctx, d_lights, lights, sizeof(LyapLight) * num_lights), "copy lights");
    lyap_scene_cam_recalculate(&cam, w, h, 1);
    /* render: the reference's method; FP64 is its DOUBLE_REALS build */
    LyapRenderMode mode;
    std::memset(&mode, 0, sizeof(mode));
    mode.refine = LYAP_REFINE_WALK;
    mode.precision = precision;
    check(lyap_calc_render(g_ctx, (LyapRGBA *)d_rgba, (LyapPoint *)d_points, &cam, &prm, seq,
                           (const LyapLight *)d_lights, num_lights, w, h, nullptr, nullptr, nullptr,
                           &mode),
          "calc_render");
    check(lyap_sync(g_ctx), "sync");
    std::vector<LyapPoint> points((size_t)w * h);
    std::vector<LyapRGBA> rgba((size_t)w * h);
    check(lyap_copy_to_host(g_ctx, points.data(), d_points, sizeof(LyapPoint) * w * h), "copy points");
    check(lyap_copy_to_host(g_ctx, rgba.data(), d_rgba, sizeof(LyapRGBA) * w * h), "copy rgba");
    check(lyap_write_points_raw(points_path, points.data(), points.size()), "write_points_raw");
    check(lyap_write_ppm(ppm_path, rgba.data(), w, h), "write_ppm");
    check(lyap_device_free(g_ctx, d_lights), "free");
    check(lyap_device_free(g_ctx, d_points), "free");
    check(lyap_device_free(g_ctx, d_rgba), "free");
    lyap_destroy(g_ctx);
    g_ctx = nullptr;
    return 0;
}

} /* namespace */

int main(int argc, char **argv)
{
    if (argc == 5 && !std::strcmp(argv[1], "calculate"))
        return calculate((unsigned)std::atoi(argv[2]), precision_of(argv[3]), argv[4]);
    if (argc == 7 && !std::strcmp(argv[1], "interactive"))
        return interactive((unsigned)std::atoi(argv[2]), (unsigned)std::atoi(argv[3]),
                           precision_of(argv[4]), argv[5], argv[6]);
    std::fprintf(stderr,
                 "usage: lyap_driver calculate N plain|product|fp64 OUT.raw\n"
                 "       lyap_driver interactive W H plain|fp64 POINTS.raw RENDER.ppm\n");
    return 2;
}
