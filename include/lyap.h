/*
 * lyap.h - C ABI of the MI355X Lyapunov library (liblyap_amd.so).
 *
 * This is the drop-in boundary for the reference's device path. Each entry
 * names the reference interface it replaces. Only plain C types and the PODs
 * of lyap_types.h cross it: no HIP, CUDA or torch type appears here.
 *
 * Conventions
 *  - Every function that can fail returns an int status: LYAP_OK (0) or a
 *    negative LYAP_ERR_* code. Nothing in the library calls exit(); the
 *    reference's helpers print and exit (helper_cuda.h, scene.cu:98-100).
 *    lyap_strerror() names a code and lyap_last_error() gives the message of
 *    the last failure on a context.
 *  - Device buffers are allocated and owned by the caller, as in the
 *    reference (lyap_interactive.cu:115-116, lyap_calculate.cu:64-67).
 *    Parameter structs and the sequence are copied at call time, so the
 *    caller may change them as soon as the call returns (the reference passes
 *    them to the kernel by value).
 *  - Compute calls are asynchronous on the context's stream; lyap_sync()
 *    waits. A context is used from one host thread at a time.
 *  - Sequences are host arrays of letter indices 0..3 (A..D) terminated by
 *    -1, as scene_convert_sequence produces them (scene.cu:69-108). Their
 *    period (entries before the -1) is 1..LYAP_MAX_SEQ.
 *  - Sizes need not be multiples of any block size: every kernel checks its
 *    bounds (the reference requires multiples of 16 or 8).
 */
#ifndef LYAP_H
#define LYAP_H

#include "lyap_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LYAP_ABI_VERSION 2
#define LYAP_MAX_SEQ 32     /* longest sequence period accepted */
#define LYAP_MAX_LIGHTS 16  /* params.hpp:21 */

enum {
    LYAP_OK = 0,
    LYAP_ERR_INVALID_ARGUMENT = -1,
    LYAP_ERR_BAD_SEQUENCE = -2,   /* letter outside A-D/a-d/1-9, empty, too long */
    LYAP_ERR_UNSUPPORTED = -3,
    LYAP_ERR_NO_DEVICE = -4,
    LYAP_ERR_DEVICE = -5,         /* a HIP runtime call failed */
    LYAP_ERR_BUDGET = -6,         /* request does not fit the HBM budget */
    LYAP_ERR_QUEUE_FULL = -7,
    LYAP_ERR_NOT_FOUND = -8,
    LYAP_ERR_CANCELLED = -9,
    LYAP_ERR_PENDING = -10,       /* lyap_poll: request not finished yet */
    LYAP_ERR_NO_LIBRARY = -11     /* reserved for bindings that fail to load the library */
};

const char *lyap_strerror(int status);

/* ------------------------------------------------------------------------ */
/* Host scene helpers (no device needed).                                    */
/* ------------------------------------------------------------------------ */

/* Defaults of params_init() (params.cu:21-114): parameters, the scripted
 * camera and the single default spotlight. Any pointer may be NULL. `lights`
 * must hold LYAP_MAX_LIGHTS entries. The camera's derived vectors are not
 * computed: call lyap_scene_cam_recalculate() for a render size. */
void lyap_params_init(LyapParams *prm, LyapCam *cam, LyapLight *lights,
                      unsigned *num_lights);

/* The default sequence string of params_init, "BCABA" (params.cu:39). */
const char *lyap_default_sequence(void);

/* scene_convert_sequence (scene.cu:69-108): "A".."D" (either case) map to
 * 0..3; a digit n appends n more copies of the previous letter (B before any
 * letter); a -1 terminator is appended. Writes at most `cap` ints and stores
 * the length including the terminator in *out_len. A bad letter returns
 * LYAP_ERR_BAD_SEQUENCE instead of exit(1). */
int lyap_scene_convert_sequence(const char *str, int32_t *out, size_t cap,
                                size_t *out_len);

/* scene_cam_recalculate (scene.cu:31-63): clamps M, normalises Q, sets
 * V, S0, SDX, SDY and the render size. tw/th/td of 0 keep the old value. */
void lyap_scene_cam_recalculate(LyapCam *cam, unsigned tw, unsigned th,
                                unsigned td);

/* scene_lights_recalculate (scene.cu:20-29): spotlight axis and cone cosines. */
void lyap_scene_lights_recalculate(LyapLight *lights, size_t num_lights);

/* QUAT(axis, angle, inDegrees) (quat.hpp:227-238). */
void lyap_quat_axis_angle(const LyapVec *axis, float angle, int in_degrees,
                          LyapQuat *out);
/* QUAT(p, q, scale) (quat.hpp:194-220): rotation taking unit p toward unit q,
 * scaled by `scale` of the angle. */
void lyap_quat_between(const LyapVec *p, const LyapVec *q, float scale,
                       LyapQuat *out);
/* QUAT::transform (quat.hpp:279-284). */
void lyap_quat_transform(const LyapQuat *q, const LyapVec *v, LyapVec *out);

/* Camera at `eye` looking at `target` (new; used by the multi-view
 * workloads): Q is the rotation of +z onto the view direction, M the screen
 * half-width, then lyap_scene_cam_recalculate(w, h, 1). */
void lyap_cam_look_at(LyapCam *cam, const LyapVec *eye, const LyapVec *target,
                      float M, unsigned w, unsigned h);

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct lyap_ctx lyap_ctx;

/* Opens `device`, creates a non-blocking stream at default priority.
 * hbm_budget_bytes caps the library's own device allocations (scratch,
 * queues, server state); 0 means no cap. */
int lyap_create(lyap_ctx **out, int device, size_t hbm_budget_bytes);
void lyap_destroy(lyap_ctx *ctx);
int lyap_sync(lyap_ctx *ctx);
/* The hipStream_t the context launches on, as an opaque pointer. */
void *lyap_get_stream(lyap_ctx *ctx);
const char *lyap_last_error(lyap_ctx *ctx);
/* Bytes of device memory the library currently holds for this context. */
size_t lyap_bytes_in_use(lyap_ctx *ctx);

/* Convenience device memory helpers for callers without a HIP toolchain
 * (the reference's drivers call cudaMalloc/cudaMemcpy directly). */
int lyap_device_alloc(lyap_ctx *ctx, void **ptr, size_t bytes);
int lyap_device_free(lyap_ctx *ctx, void *ptr);
int lyap_copy_to_device(lyap_ctx *ctx, void *dst, const void *src, size_t bytes);
int lyap_copy_to_host(lyap_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------------------ */
/* Arithmetic forms of the lambda iteration                                  */
/* ------------------------------------------------------------------------ */

enum {
    /* FP32 iterate and one log per accumulated step: the reference's method
     * (kernel.cu:140-150). */
    LYAP_PREC_PLAIN = 0,
    /* FP32 iterate in u = 1 - 2x, derivative product carried as an FP32
     * mantissa with an integer exponent, renormalised every `renorm` steps;
     * one log per evaluation. */
    LYAP_PREC_PRODUCT = 1,
    /* FP64 iterate with the reference's operation order (DOUBLE_REALS build). */
    LYAP_PREC_FP64 = 2
};

/* Volume options. NULL selects {LYAP_PREC_PLAIN, no re-evaluation}. */
typedef struct LyapVolumeMode {
    int32_t precision;      /* LYAP_PREC_* */
    uint32_t renorm;        /* PRODUCT: steps between renormalisations; 0 = whole
                               periods up to 32 steps, about 12 once a wave has
                               had to redo a block (results do not depend on it) */
    float fp64_margin;      /* re-evaluate in FP64 where |lambda| < margin; 0 = off */
    float converge_tau;     /* also re-evaluate where the two halves of the window */
    float converge_band;    /*   differ by > tau and the smaller half-lambda < band */
    uint32_t reserved[3];
} LyapVolumeMode;

/* kernel_calc_volume (kernel.hpp:19, kernel.cu:518-532; launched at
 * lyap_calculate.cu:72): lambda of every voxel of an nx*ny*nz grid over the
 * box [lo, hi) with coordinate lo + (hi-lo)*i/n per axis and index
 * x + (y + z*ny)*nx. The reference's box is lo = 0, hi = 4. */
int lyap_calc_volume(lyap_ctx *ctx, float *d_exps, const LyapParams *prm,
                     const int32_t *seq, const float lo[3], const float hi[3],
                     unsigned nx, unsigned ny, unsigned nz,
                     const LyapVolumeMode *mode);

/* Allocation-free variant of lyap_calc_volume. It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results as lyap_calc_volume, but
 * the FP64 re-evaluation list and its counter live in a fixed per-device
 * array of 16 MiB that the library holds until the process ends, and the
 * grid is processed in z slabs that the list holds entirely. Limit: one z
 * slice (nx*ny voxels) of at most LYAP_NOALLOC_MAX_SLICE, else
 * LYAP_ERR_UNSUPPORTED. Nothing on this path calls hipMalloc or hipFree.
 * Every call on a device shares that device's array: calls that overlap on
 * one device (two streams, or two host threads) race and must not be made.
 * The call does not update lyap_last_reevaluated. */
#define LYAP_NOALLOC_MAX_SLICE 4194304u
int lyap_calc_volume_noalloc(lyap_ctx *ctx, float *d_exps, const LyapParams *prm,
                             const int32_t *seq, const float lo[3], const float hi[3],
                             unsigned nx, unsigned ny, unsigned nz,
                             const LyapVolumeMode *mode);

/* lambda at n arbitrary points (x, y, z float triples in device memory):
 * lyap4d (kernel.hpp:13, kernel.cu:108-154) as a batch entry. */
int lyap_calc_points(lyap_ctx *ctx, float *d_out, const float *d_xyz, size_t n,
                     const LyapParams *prm, const int32_t *seq,
                     const LyapVolumeMode *mode);

/* The PRODUCT form's lambda and the lambdas of the two halves of its
 * accumulation window for each point (a convergence diagnostic: a point
 * whose halves disagree has not settled on its attractor). No reference
 * counterpart. */
int lyap_calc_points_halves(lyap_ctx *ctx, float *d_lambda, float *d_half1, float *d_half2,
                            const float *d_xyz, size_t n, const LyapParams *prm,
                            const int32_t *seq, unsigned renorm);

/* Statistics of the last volume/points call on this context: voxels whose
 * lambda was re-evaluated in FP64. Waits for that call to finish. */
int lyap_last_reevaluated(lyap_ctx *ctx, uint64_t *count);

/* ------------------------------------------------------------------------ */
/* Render                                                                    */
/* ------------------------------------------------------------------------ */

enum {
    LYAP_REFINE_WALK = 0,   /* reversing half-step walk (kernel.cu:402-441) */
    LYAP_REFINE_BISECT = 1  /* fixed-count bisection of the last march step */
};

/* Render options. NULL selects the reference's method: refinement walk,
 * normal offset step*gradient, incremental positions, LYAP_PREC_PLAIN. */
typedef struct LyapRenderMode {
    int32_t refine;         /* LYAP_REFINE_* */
    int32_t bisect_steps;   /* BISECT: number of halvings */
    float normal_h;         /* > 0: absolute normal offset; 0: step * gradient */
    int32_t precision;      /* LYAP_PREC_PLAIN or LYAP_PREC_PRODUCT */
    float fp64_margin;      /* BISECT: samples with |lambda - opaque| < margin
                               are re-evaluated in FP64; 0 = off */
    uint32_t lanes_per_ray; /* 0 = chosen by the library */
    int32_t schedule;       /* LYAP_SCHEDULE_* */
    uint32_t reserved;
} LyapRenderMode;

/* How rays are spread over lanes. BATCHED: the persistent work queue with
 * lane groups per ray, refill of finished lanes and continuation launches;
 * with bisection, rays whose march found its opaque sample are refined in a
 * launch of their own, two lanes per ray. A small launch of up to 16 views
 * refined by bisection (fewer rays than about two per resident lane) marches in passes
 * instead: a set-up launch lists the rays that meet the box, then launches
 * of one round each give every listed ray K lanes (4, 8, 16, 32, then 64
 * samples at once), the last running its rays to the end of their march.
 * LEVELS: BATCHED without the passes (the persistent launches only).
 * DIRECT: the reference's own launch shape (kernel.cu:500-516): one lane
 * per ray for 64 consecutive rays (an 8x8 pixel block), each lane marching
 * its ray to the end and a wave taking new rays only when all 64 are done.
 * All give the same bytes; DIRECT is the plain per-request baseline. */
enum { LYAP_SCHEDULE_BATCHED = 0, LYAP_SCHEDULE_DIRECT = 1, LYAP_SCHEDULE_LEVELS = 2 };

/* kernel_calc_render (kernel.hpp:17, kernel.cu:500-516; launched at
 * lyap_interactive.cu:711): one primary ray per pixel of a w*h image through
 * the box [lMin, lMax]^3, hit record into d_points[x + y*w] and shaded
 * colour into d_rgba. A miss writes P = N = 0, a = l = NaN, c = 0 and colour
 * (0,0,0,0) (the reference leaves the record unwritten). d_lights is a device
 * array of num_lights <= 16 lights (lyap_interactive.cu:115,130). d_depth and
 * d_lambda are optional per-pixel outputs (ray parameter t and lambda of the
 * hit, NaN on a miss); d_evals optionally receives the number of lambda
 * evaluations each ray's method needed. Any output pointer except d_rgba may
 * be NULL. */
int lyap_calc_render(lyap_ctx *ctx, LyapRGBA *d_rgba, LyapPoint *d_points,
                     const LyapCam *cam, const LyapParams *prm,
                     const int32_t *seq, const LyapLight *d_lights,
                     unsigned num_lights, unsigned w, unsigned h,
                     float *d_depth, float *d_lambda, uint32_t *d_evals,
                     const LyapRenderMode *mode);

/* Re-shades a rendered frame from its LyapPoint buffer (d_points as
 * lyap_calc_render wrote it) for new lights, without marching: with the
 * frame's own camera the RGBA equals what lyap_calc_render gives with those
 * lights (kernel.cu:511-514). For the viewer's light edits, which raise
 * update_lights/update_render but not update_calc and leave the camera's C
 * and Q alone (lyap_interactive.cu:267-307, 340-356); the reference
 * re-renders the whole frame for them (:124-142, 711). */
int lyap_shade_points(lyap_ctx *ctx, LyapRGBA *d_rgba, const LyapPoint *d_points,
                      const LyapCam *cam, const LyapLight *d_lights,
                      unsigned num_lights, unsigned w, unsigned h);

/* Diagnostic: a device buffer of 16 uint64 per wave slot, 8192 slots per
 * launch level, 4 levels (4 MiB), that later render launches (in the
 * persistent levels, which frames use while the buffer is set) fill with
 * {rounds, rays started, start clock, end clock, clocks evaluating, clocks
 * refilling, clocks combining, clocks of those refilling spent claiming
 * tiles, clocks of those combining spent handing rays over, clocks of
 * those refilling spent starting rays, start and end on the constant-rate
 * wall clock (100 MHz, common to the chip's XCDs, which the shader clock
 * is not), 0 x 4}; NULL switches it off. No reference counterpart. */
int lyap_debug_wave_log(lyap_ctx *ctx, void *d_log);

/* Kernel timing: while on, every render kernel launch on the context's
 * stream is bracketed by a pair of HIP events; switching it off pauses
 * (launches recorded so far stay), so a caller may time a sample of its
 * frames. lyap_timing_collect waits for the stream, returns the summed
 * kernel milliseconds and the number of launches recorded since the last
 * collection, and starts over. No reference counterpart (measurement
 * only). */
int lyap_timing_enable(lyap_ctx *ctx, int on);
int lyap_timing_collect(lyap_ctx *ctx, double *kernel_ms, uint64_t *launches);

/* Total algorithmic lambda evaluations of the last render/batch call
 * (march samples through the first opaque one, refinement, normal). Waits
 * for that call to finish. */
int lyap_last_evals(lyap_ctx *ctx, uint64_t *count);

/* ------------------------------------------------------------------------ */
/* Batched views: many cameras in one persistent launch                      */
/* ------------------------------------------------------------------------ */

typedef struct LyapView {
    LyapCam cam;                       /* derived vectors already computed */
    LyapParams prm;
    int32_t seq[LYAP_MAX_SEQ + 1];     /* -1 terminated */
    uint32_t width, height;
    LyapRGBA *rgba;                    /* device outputs, as in lyap_calc_render */
    LyapPoint *points;
    float *depth;
    float *lambda;
    uint32_t *evals;
    const LyapLight *lights;           /* device array */
    uint32_t num_lights;
    uint32_t reserved;
} LyapView;

/* Launch shape of the persistent kernels (the reference's fixed 16x16 and
 * 8x8x8 blocks, lyap_interactive.cu:64-66 and lyap_calculate.cu:19-24, are
 * what it replaces). Zero fields are filled from the autotuner's choice
 * (lyap_autotune, lyap_autotune_batch) or built-in defaults. */
typedef struct LyapTuning {
    uint32_t tile_rays;        /* rays handed out per work-queue claim */
    uint32_t waves_per_cu;     /* resident waves requested per CU */
    uint32_t lanes_per_ray;    /* lanes sharing one ray's march */
    uint32_t slots;            /* samples each lane evaluates per round, 1..4 */
    uint32_t level_budget[3];  /* march samples a ray may take in launch levels 0, 1
                                  and 2 before it is handed to the next level (0: 24
                                  at four or more lanes per ray, else 64 / 64 / 128) */
    uint32_t voxel_tile;       /* voxels per work-queue claim of a volume block,
                                  rounded up to a multiple of 64 (0: 256) */
} LyapTuning;

/* Renders n_views views in one persistent launch fed by a device work queue
 * of ray tiles. Views may differ in size, camera, parameters and sequence. */
int lyap_render_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                      const LyapRenderMode *mode, const LyapTuning *tuning);

/* A block of a volume for lyap_run_batch: z slices [z_begin, z_end) of the
 * n[0] x n[1] x n[2] grid over [lo, hi) that lyap_calc_volume computes,
 * lambda written to `out` (device memory; slice z_begin first, x fastest). */
typedef struct LyapVolumeBlock {
    LyapParams prm;
    int32_t seq[LYAP_MAX_SEQ + 1];     /* -1 terminated */
    float lo[3], hi[3];
    uint32_t n[3];
    uint32_t z_begin, z_end;           /* z_end 0: n[2] */
    float *out;
} LyapVolumeBlock;

/* Views and volume blocks of different sizes in one persistent launch
 * sequence (kernel_calc_render and kernel_calc_volume, kernel.hpp:17,19,
 * in one): the first launch's device work queue deals the views' ray tiles
 * and then the blocks' voxel tiles (one lane per voxel), so waves whose
 * rays have ended take voxels; voxels the volume mode's rule lists are
 * re-evaluated in FP64 by the next launch. Each view gets the bytes
 * lyap_calc_render gives it and each block the bytes of lyap_calc_volume
 * over the same grid (the same arithmetic). Blocks are evaluated in the
 * views' precision: when vol_mode's precision differs they run in the
 * volume kernels after the views' launches (same results, not fused). */
int lyap_run_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                   const LyapVolumeBlock *blocks, unsigned n_blocks,
                   const LyapRenderMode *mode, const LyapVolumeMode *vol_mode,
                   const LyapTuning *tuning);

/* Times candidate launch shapes on a representative batch and stores the
 * fastest in the context (used when a LyapTuning field is 0): lanes per
 * ray x samples per lane (one lane per ray included) x tile length, then
 * each level's march-sample budget, then (with volume blocks) the voxel
 * tile. The results do not depend on the shape. */
int lyap_autotune(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                  const LyapRenderMode *mode, LyapTuning *best);
int lyap_autotune_batch(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                        const LyapVolumeBlock *blocks, unsigned n_blocks,
                        const LyapRenderMode *mode, const LyapVolumeMode *vol_mode,
                        LyapTuning *best);
/* Sets the context's tuned shape (what lyap_autotune stores), or clears it
 * with NULL; the server's launches use it too. */
int lyap_set_tuning(lyap_ctx *ctx, const LyapTuning *tuning);
int lyap_get_tuning(lyap_ctx *ctx, LyapTuning *out);

/* A captured frame: the batch upload and the frame's launches (the march
 * passes or levels, and the refinement) recorded once as a HIP graph and
 * replayed. The views' device buffers must stay valid. */
typedef struct lyap_frame lyap_frame;
int lyap_frame_capture(lyap_ctx *ctx, const LyapView *views, unsigned n_views,
                       const LyapRenderMode *mode, const LyapTuning *tuning,
                       lyap_frame **out);
int lyap_frame_replay(lyap_ctx *ctx, lyap_frame *frame);
void lyap_frame_destroy(lyap_frame *frame);

/* ------------------------------------------------------------------------ */
/* Headless parameter editing (the viewer's keyboard and spaceball)          */
/* ------------------------------------------------------------------------ */

/* Control modes of lyap_interactive.cu:43-55. */
enum {
    LYAP_EDIT_LOCK = 0, LYAP_EDIT_ROTPAN, LYAP_EDIT_MAG, LYAP_EDIT_NTHRESH,
    LYAP_EDIT_OTHRESH, LYAP_EDIT_CTHRESH, LYAP_EDIT_D, LYAP_EDIT_JITTER,
    LYAP_EDIT_L_PAN, LYAP_EDIT_L_RANGE, LYAP_EDIT_L_COLOR
};
/* What an edit invalidated (update_cam/lights/calc/render,
 * lyap_interactive.cu:57-60), and the quit key. */
enum {
    LYAP_UPDATE_CAM = 1, LYAP_UPDATE_LIGHTS = 2, LYAP_UPDATE_CALC = 4,
    LYAP_UPDATE_RENDER = 8, LYAP_UPDATE_QUIT = 16
};

/* The viewer's editable state (lyap_interactive.cu:36-71, params.cu:9). */
typedef struct LyapEditor {
    LyapParams prm;
    LyapCam cam;
    LyapLight lights[LYAP_MAX_LIGHTS];
    LyapLight home_light;        /* L0: the first light as initialised */
    uint32_t num_lights;
    int32_t mode;                /* LYAP_EDIT_* */
    int32_t cur_light;           /* -1: the camera is current (curC), else a light */
    int32_t light_index;         /* curL */
    int32_t dominant;            /* spaceball axis locking */
    uint32_t window_w, window_h, render_denominator;
    uint32_t pending;            /* LYAP_UPDATE_* since the last commit */
} LyapEditor;

/* params_init plus the viewer's initial control state. */
int lyap_editor_init(LyapEditor *ed, unsigned window_w, unsigned window_h);
/* glut_keyboard (lyap_interactive.cu:311-463); returns LYAP_UPDATE_* flags. */
unsigned lyap_editor_key(LyapEditor *ed, int key);
/* glut_spaceball_rot / _pan / _but (lyap_interactive.cu:144-309). */
unsigned lyap_editor_spaceball_rot(LyapEditor *ed, int x, int y, int z);
unsigned lyap_editor_spaceball_pan(LyapEditor *ed, int x, int y, int z);
unsigned lyap_editor_spaceball_button(LyapEditor *ed, int button);
/* update_scene (lyap_interactive.cu:118-137): recalculates the lights and
 * the camera for an image_w x image_h frame as the pending flags ask; returns
 * and clears them. A caller then renders (or submits) the new frame. */
unsigned lyap_editor_commit(LyapEditor *ed, unsigned image_w, unsigned image_h);

/* ------------------------------------------------------------------------ */
/* Scripted fly-through (scale.pl)                                           */
/* ------------------------------------------------------------------------ */

/* easeInOutQuart(t, b, c, d) (scale.pl:5-11). */
double lyap_ease_in_out_quart(double t, double b, double c, double d);
/* The frame times of `for ($t = 0; $t <= 1.00001; $t += step)`
 * (scale.pl:15): their number, and up to `cap` of them in t_out. */
unsigned lyap_flythrough_frames(double step, double *t_out, unsigned cap);
/* The camera of scale.pl:33-50 for an eased value i: the view axis (4,4,4)
 * turned by nlerp(-20, +20 degrees about `up`, i), the camera one unit
 * behind (4 - 0.9 i) along each axis. Sets C and Q; the caller then calls
 * lyap_scene_cam_recalculate. */
void lyap_flythrough_camera(double eased, LyapCam *cam);

/* ------------------------------------------------------------------------ */
/* On-disk formats                                                           */
/* ------------------------------------------------------------------------ */

/* exps.raw: raw FP32, x fastest (lyap_calculate.cu:79-86). Host buffers. */
int lyap_write_volume_raw(const char *path, const float *exps, size_t n);
/* Points_*.raw: raw LyapPoint[w*h] (lyap_interactive.cu:616-655). */
int lyap_write_points_raw(const char *path, const LyapPoint *points, size_t n);
/* Render_*.ppm: "P3\n<w> <h>\n255\n" then "%3d %3d %3d " per pixel in
 * buffer order and a newline (lyap_interactive.cu:569-613). */
int lyap_write_ppm(const char *path, const LyapRGBA *rgba, unsigned w, unsigned h);
/* The parameter-encoding leaf name "<kind>_<stamp>_<w>x<h>_<seq>_cx=..._ot=..."
 * (lyap_interactive.cu:577-591), kind "Render" or "Points". */
int lyap_output_leaf(char *out, size_t cap, const char *kind, long stamp, unsigned w,
                     unsigned h, const char *sequence, const LyapCam *cam,
                     const LyapParams *prm);
/* "<leaf>_time=<h>h<mm>m<ss>s.<ext>" (lyap_interactive.cu:594, 607). */
int lyap_output_name(char *out, size_t cap, const char *leaf, unsigned hours,
                     unsigned minutes, unsigned seconds, const char *ext);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching of render and volume requests               */
/* ------------------------------------------------------------------------ */

enum { LYAP_REQ_VIEW = 0, LYAP_REQ_VOLUME = 1 };

typedef struct LyapRequest {
    int32_t kind;                      /* LYAP_REQ_* */
    uint32_t priority;                 /* reserved, 0 */
    LyapView view;                     /* VIEW: outputs may be NULL: the server
                                          then keeps RGBA and lambda resident */
    LyapParams vol_prm;                /* VOLUME */
    int32_t vol_seq[LYAP_MAX_SEQ + 1];
    float vol_lo[3], vol_hi[3];
    uint32_t vol_n[3];
    float *vol_out;                    /* VOLUME: device output or NULL */
    double deadline_ms;                /* relative deadline used for ordering */
} LyapRequest;

/* LyapServerConfig.flags */
enum { LYAP_SERVER_SEPARATE_VOLUMES = 1 };

typedef struct LyapServerConfig {
    size_t hbm_budget_bytes;           /* cap on resident outputs (0: none) */
    uint32_t render_slots;             /* launches in flight, one stream each:
                                          render_slots + volume_slots in all (0: 2) */
    uint32_t volume_slots;             /* (0: 1) */
    double slice_ms;                   /* device time aimed at per launch (0: 12) */
    uint32_t chunk_rays;               /* smallest piece of a view a launch
                                          takes, a multiple of 64 (0: 4096) */
    uint32_t flags;                    /* LYAP_SERVER_SEPARATE_VOLUMES: render slots
                                          take views only and volume slots run slabs
                                          in the volume kernels; otherwise every
                                          launch may carry both kinds */
    LyapRenderMode mode;
    LyapVolumeMode vol_mode;
    uint32_t max_queued;               /* requests waiting for admission; a submit
                                          beyond it returns LYAP_ERR_QUEUE_FULL
                                          (0: no cap) */
} LyapServerConfig;

typedef struct LyapResult {
    int32_t status;                    /* LYAP_OK, LYAP_ERR_PENDING, ... */
    uint32_t kind;
    double submit_ms, start_ms, finish_ms;   /* server clock */
    uint64_t rays, voxels;
    uint64_t evals;                    /* volumes: lambda evaluations; views: 0
                                          (a launch mixes views, so evaluations
                                          are counted per server, not per view) */
    void *rgba;                        /* view RGBA (device) */
    float *lambda;                     /* view lambda, or the volume (device) */
} LyapResult;

typedef struct LyapServerStats {
    uint64_t submitted, completed, cancelled, failed;
    uint64_t rays, voxels, evals;      /* work of finished launches */
    uint64_t launches;
    size_t resident_bytes, peak_resident_bytes;
    uint32_t queued, active;           /* requests waiting / admitted */
    double device_ms;                  /* sum of launch durations */
} LyapServerStats;

/* Continuous batching over one device. A view is cut into 64-ray pieces
 * and a volume into z slabs; each launch takes the pieces of the requests
 * with the earliest deadlines, up to about `slice_ms` of device time, so a
 * request that arrives while others run joins the next launch. A launch
 * carries views and volume slabs together (one persistent launch, as
 * lyap_run_batch), each slot on its own stream, so the launches' work
 * overlaps on the device. A request's outputs are admitted against
 * `hbm_budget_bytes` when its first piece is issued and stay resident until
 * lyap_server_release. Requests are admitted in deadline order: one that
 * does not fit yet holds back the admission of those with later deadlines
 * (requests already admitted keep running), so a large request is not
 * passed over indefinitely. A request without a deadline orders after every
 * request that has one. When the device has no room for a request's outputs,
 * the server first frees its pooled buffers; the request then waits while
 * other requests hold resident outputs, and ends with LYAP_ERR_DEVICE when
 * none do. A launch that fails ends the requests it carried with its error
 * once their other pieces are done. One host thread drives a server. */
typedef struct lyap_server lyap_server;
int lyap_server_create(lyap_ctx *ctx, const LyapServerConfig *cfg,
                       lyap_server **out);
void lyap_server_destroy(lyap_server *srv);
/* Queues a request. A VIEW's lights and any caller-given outputs must stay
 * valid until the request finishes. LYAP_ERR_QUEUE_FULL when `max_queued`
 * requests already wait for admission. */
int lyap_server_submit(lyap_server *srv, const LyapRequest *req,
                       uint64_t *ticket);
/* Advances the server: retires finished launches, admits requests, and
 * issues new launches. Returns once something changed, or after `wait_ms`. */
int lyap_server_progress(lyap_server *srv, double wait_ms);
/* LYAP_OK with out->status LYAP_ERR_PENDING while the request runs. */
int lyap_server_poll(lyap_server *srv, uint64_t ticket, LyapResult *out);
/* Frees a finished request's resident outputs and forgets its ticket. */
int lyap_server_release(lyap_server *srv, uint64_t ticket);
/* Stops issuing a request's pieces; it ends with LYAP_ERR_CANCELLED once
 * the pieces already on the device are done. */
int lyap_server_cancel(lyap_server *srv, uint64_t ticket);
double lyap_server_clock_ms(lyap_server *srv);
int lyap_server_stats(lyap_server *srv, LyapServerStats *out);
/* Tickets of requests that finished (any status) since the last call, in
 * the order they finished: up to `cap` of them, their number in *n. */
int lyap_server_finished(lyap_server *srv, uint64_t *tickets, unsigned cap, unsigned *n);
/* Blocks until every request has finished (or `timeout_ms` passes). */
int lyap_server_drain(lyap_server *srv, double timeout_ms);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* LYAP_H */
