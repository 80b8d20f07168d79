/*
 * sanitize_main.cpp - runs the library's host code under AddressSanitizer
 * and UndefinedBehaviorSanitizer: the scene helpers, the parameter editor,
 * the file writers, the serving scheduler over the host-memory device, and
 * the CPU oracle the tests check against. Built by this folder's Makefile
 * (target host_sanitize) with g++ only; any sanitizer report ends the
 * program with a non-zero status.
 *
 * usage: host_sanitize <scratch directory>
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fake_server.h"
#include "lyap.h"

extern "C" {
/* the oracle's entry points (oracle/lyap_oracle.cpp) */
void oracle_volume(int mode, float *out, const LyapParams *prm, const int32_t *seq, const float *lo,
                   const float *hi, unsigned nx, unsigned ny, unsigned nz, unsigned z_begin, unsigned z_end,
                   int threads);
void oracle_render(int lam_mode, const LyapRenderMode *mode_in, LyapRGBA *rgba, LyapPoint *points,
                   float *depth, float *lambda, uint32_t *evals, int32_t *hit_index, float *margin,
                   const LyapCam *cam, const LyapParams *prm, const int32_t *seq, const LyapLight *lights,
                   unsigned num_lights, unsigned w, unsigned h, unsigned row_begin, unsigned row_step,
                   int threads);
void oracle_lyap_halves(const float *xyz, size_t n, float d, unsigned settle, unsigned accum,
                        unsigned split, const int32_t *seq, double *first, double *second, int threads);
}

static int g_failures = 0;

#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "check failed: %s (%s:%d)\n", #cond,        \
                         __FILE__, __LINE__);                                \
            ++g_failures;                                                    \
        }                                                                    \
    } while (0)

static void scene_and_editor(LyapParams *prm, LyapCam *cam, LyapLight *lights, unsigned *nl, int32_t *seq)
{
    size_t len = 0;
    CHECK(lyap_scene_convert_sequence("BCABA", seq, LYAP_MAX_SEQ + 1, &len) == LYAP_OK && len == 6);
    int32_t tmp[LYAP_MAX_SEQ + 1];
    CHECK(lyap_scene_convert_sequence("A9B9C9", tmp, LYAP_MAX_SEQ + 1, &len) != LYAP_OK ||
          len <= LYAP_MAX_SEQ + 1);
    CHECK(lyap_scene_convert_sequence("AxB", tmp, LYAP_MAX_SEQ + 1, &len) == LYAP_ERR_BAD_SEQUENCE);
    CHECK(lyap_scene_convert_sequence("3AB2", tmp, 3, &len) != LYAP_OK || len <= 3);
    CHECK(lyap_scene_convert_sequence("", tmp, LYAP_MAX_SEQ + 1, &len) != LYAP_OK || len <= 1);

    lyap_params_init(prm, cam, lights, nl);
    lyap_scene_cam_recalculate(cam, 24, 16, 1);
    lyap_scene_lights_recalculate(lights, *nl);
    LyapVec axis = {0.0f, 1.0f, 0.0f}, v = {1.0f, 2.0f, 3.0f}, o;
    LyapQuat q;
    lyap_quat_axis_angle(&axis, 30.0f, 1, &q);
    lyap_quat_transform(&q, &v, &o);
    LyapVec p1 = {1.0f, 0.0f, 0.0f}, p2 = {0.0f, 0.0f, 1.0f};
    lyap_quat_between(&p1, &p2, 0.5f, &q);
    lyap_quat_between(&p1, &p1, 1.0f, &q);
    LyapVec eye = {0.5f, 0.5f, 0.5f}, target = {4.0f, 4.0f, 4.0f};
    LyapCam c2 = *cam;
    lyap_cam_look_at(&c2, &eye, &target, 1.0f, 20, 12);

    LyapEditor ed;
    CHECK(lyap_editor_init(&ed, 640, 480) == LYAP_OK);
    const char keys[] = "0123456789abcdefghijklmnopqrstuvwxyz+-=_[]{}<>,./?!@#$%^&*() \t\r";
    for (const char *k = keys; *k; ++k) lyap_editor_key(&ed, (unsigned char)*k);
    for (int b = 0; b < 10; ++b) lyap_editor_spaceball_button(&ed, b);
    for (int i = -3; i <= 3; ++i) {
        lyap_editor_spaceball_rot(&ed, 100 * i, -50 * i, 700 * i);
        lyap_editor_spaceball_pan(&ed, -200 * i, 30 * i, 90 * i);
        lyap_editor_key(&ed, 'l');
    }
    lyap_editor_commit(&ed, 32, 24);
    lyap_editor_commit(&ed, 0, 0);

    double ts[64];
    const unsigned nf = lyap_flythrough_frames(0.05, ts, 64);
    CHECK(nf > 0 && nf <= 64);
    for (unsigned i = 0; i < nf; ++i) {
        LyapCam fc = *cam;
        lyap_flythrough_camera(lyap_ease_in_out_quart(ts[i], 0.0, 1.0, 1.0), &fc);
        lyap_scene_cam_recalculate(&fc, 8, 8, 1);
    }
}

static void writers(const char *dir, const LyapParams *prm, const LyapCam *cam)
{
    char leaf[512], name[600];
    CHECK(lyap_output_leaf(leaf, sizeof(leaf), "Render", 1234567L, 24, 16, "BCABA", cam, prm) >= 0);
    CHECK(lyap_output_name(name, sizeof(name), leaf, 1, 2, 3, "ppm") >= 0);
    char small[16];
    lyap_output_leaf(small, sizeof(small), "Points", 1L, 1, 1, "AB", cam, prm);
    lyap_output_name(small, sizeof(small), "x", 100, 59, 59, "raw");
    CHECK(std::strlen(small) < sizeof(small));

    std::vector<float> vol(5 * 4 * 3);
    for (size_t i = 0; i < vol.size(); ++i) vol[i] = (float)i - 7.5f;
    std::vector<LyapRGBA> rgba(7 * 5);
    for (size_t i = 0; i < rgba.size(); ++i) rgba[i] = {(unsigned char)i, 255, 0, 128};
    std::vector<LyapPoint> pts(7 * 5);
    std::memset(pts.data(), 0, pts.size() * sizeof(LyapPoint));
    const std::string d(dir);
    CHECK(lyap_write_volume_raw((d + "/exps.raw").c_str(), vol.data(), vol.size()) == LYAP_OK);
    CHECK(lyap_write_ppm((d + "/r.ppm").c_str(), rgba.data(), 7, 5) == LYAP_OK);
    CHECK(lyap_write_points_raw((d + "/p.raw").c_str(), pts.data(), pts.size()) == LYAP_OK);
    CHECK(lyap_write_ppm((d + "/no/such/dir/r.ppm").c_str(), rgba.data(), 7, 5) != LYAP_OK);
}

static LyapRequest view_request(const LyapCam &cam, const LyapParams &prm, const int32_t *seq, unsigned w,
                                unsigned h, double deadline)
{
    LyapRequest r;
    std::memset(&r, 0, sizeof(r));
    r.kind = LYAP_REQ_VIEW;
    r.view.cam = cam;
    r.view.prm = prm;
    std::memcpy(r.view.seq, seq, sizeof(r.view.seq));
    r.view.width = w;
    r.view.height = h;
    r.deadline_ms = deadline;
    return r;
}

static LyapRequest volume_request(const LyapParams &prm, const int32_t *seq, unsigned nx, unsigned ny,
                                  unsigned nz, float *out, double deadline)
{
    LyapRequest r;
    std::memset(&r, 0, sizeof(r));
    r.kind = LYAP_REQ_VOLUME;
    r.vol_prm = prm;
    std::memcpy(r.vol_seq, seq, sizeof(r.vol_seq));
    for (int k = 0; k < 3; ++k) {
        r.vol_lo[k] = 2.0f;
        r.vol_hi[k] = 4.0f;
    }
    r.vol_n[0] = nx;
    r.vol_n[1] = ny;
    r.vol_n[2] = nz;
    r.vol_out = out;
    r.deadline_ms = deadline;
    return r;
}

static void serving(const LyapParams &prm, const LyapCam &cam, const int32_t *seq)
{
    for (int manual = 0; manual < 2; ++manual) {
        LyapServerConfig cfg;
        std::memset(&cfg, 0, sizeof(cfg));
        cfg.mode.refine = LYAP_REFINE_BISECT;
        cfg.mode.bisect_steps = 8;
        cfg.mode.precision = LYAP_PREC_PRODUCT;
        cfg.vol_mode.precision = LYAP_PREC_PRODUCT;
        cfg.slice_ms = 1e-6;
        cfg.chunk_rays = 64;
        cfg.hbm_budget_bytes = 64 << 10;
        cfg.flags = manual ? LYAP_SERVER_SEPARATE_VOLUMES : 0;
        FakeOptions opt;
        std::memset(&opt, 0, sizeof(opt));
        opt.manual = manual;
        opt.device_bytes = 48 << 10;
        lyap_server *s = nullptr;
        CHECK(fake_server_create(&cfg, &opt, &s) == LYAP_OK);
        if (!s) return;
        std::vector<float> own(6 * 5 * 4);
        std::vector<uint64_t> t;
        uint64_t k = 0;
        const unsigned sizes[][2] = {{17, 9}, {8, 8}, {33, 31}, {1, 1}, {64, 3}};
        for (int i = 0; i < 5; ++i) {
            LyapRequest r = view_request(cam, prm, seq, sizes[i][0], sizes[i][1], 1.0 + i);
            if (lyap_server_submit(s, &r, &k) == LYAP_OK) t.push_back(k);
        }
        LyapRequest v1 = volume_request(prm, seq, 6, 5, 4, own.data(), 2.0);
        LyapRequest v2 = volume_request(prm, seq, 9, 7, 5, nullptr, 0.0);
        CHECK(lyap_server_submit(s, &v1, &k) == LYAP_OK);
        t.push_back(k);
        CHECK(lyap_server_submit(s, &v2, &k) == LYAP_OK);
        t.push_back(k);
        lyap_server_cancel(s, t[1]);
        LyapRequest bad = view_request(cam, prm, seq, 0, 4, 0.0);
        CHECK(lyap_server_submit(s, &bad, &k) == LYAP_ERR_INVALID_ARGUMENT);
        bad = volume_request(prm, seq, 4096, 4096, 4096, nullptr, 0.0);
        CHECK(lyap_server_submit(s, &bad, &k) != LYAP_OK);
        for (int it = 0; it < 20000; ++it) {
            LyapServerStats st;
            lyap_server_stats(s, &st);
            if (st.queued == 0 && st.active == 0) break;
            if (manual) {
                for (int sl = 0; sl < fake_slot_count(s); ++sl)
                    if (fake_running(s, (unsigned)sl))
                        fake_complete(s, (unsigned)sl, (it % 97 == 5) ? LYAP_ERR_DEVICE : LYAP_OK);
            }
            if (it % 13 == 3) fake_fail_next_launch(s, LYAP_ERR_DEVICE);
            lyap_server_progress(s, 0.5);
            uint64_t done[8];
            unsigned n = 0;
            lyap_server_finished(s, done, 8, &n);
            for (unsigned j = 0; j < n; ++j) {
                LyapResult res;
                CHECK(lyap_server_poll(s, done[j], &res) == LYAP_OK);
                CHECK(lyap_server_release(s, done[j]) == LYAP_OK);
            }
        }
        LyapServerStats st;
        lyap_server_stats(s, &st);
        CHECK(st.queued == 0 && st.active == 0);
        CHECK(st.resident_bytes == 0);
        lyap_server_destroy(s);
    }
    CHECK(fake_live_bytes_all() == 0);
}

static void oracle_small(const LyapParams &prm0, const LyapCam &cam0, const LyapLight *lights, unsigned nl,
                         const int32_t *seq)
{
    LyapParams prm = prm0;
    prm.settle = 20;
    prm.accum = 60;
    LyapCam cam = cam0;
    lyap_scene_cam_recalculate(&cam, 12, 8, 1);
    const unsigned w = 12, h = 8;
    std::vector<LyapRGBA> rgba(w * h);
    std::vector<LyapPoint> pts(w * h);
    std::vector<float> depth(w * h), lam(w * h), margin(w * h);
    std::vector<uint32_t> evals(w * h);
    std::vector<int32_t> hit(w * h);
    for (int m = 0; m < 2; ++m)
        oracle_render(m, nullptr, rgba.data(), pts.data(), depth.data(), lam.data(), evals.data(), hit.data(),
                      margin.data(), &cam, &prm, seq, lights, nl, w, h, 0, 1, 2);
    std::vector<float> vol(5 * 4 * 3);
    const float lo[3] = {0.0f, 0.0f, 0.0f}, hi[3] = {4.0f, 4.0f, 4.0f};
    for (int m = 0; m < 2; ++m) oracle_volume(m, vol.data(), &prm, seq, lo, hi, 5, 4, 3, 0, 3, 2);
    const float xyz[6] = {2.5f, 3.5f, 3.0f, 3.9f, 3.9f, 3.9f};
    double a[2], b[2];
    oracle_lyap_halves(xyz, 2, prm.d, prm.settle, prm.accum, 30, seq, a, b, 1);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    LyapParams prm;
    LyapCam cam;
    LyapLight lights[LYAP_MAX_LIGHTS];
    unsigned nl = 0;
    int32_t seq[LYAP_MAX_SEQ + 1];
    scene_and_editor(&prm, &cam, lights, &nl, seq);
    writers(argv[1], &prm, &cam);
    serving(prm, cam, seq);
    oracle_small(prm, cam, lights, nl, seq);
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures);
        return 1;
    }
    std::printf("host code clean under the sanitizers\n");
    return 0;
}
