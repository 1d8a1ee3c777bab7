/*
 * eqfluid.h -- C ABI of libeqfluid.so, the MI355X (gfx950) Stam frame step.
 *
 * Drop-in boundary for the reference's simulation core, `Fluid`
 * (src/simulation/fluid.rs).  The reference has no FFI layer: its boundary is
 * the Rust API of `Fluid` that `CurrentSimulation::simulate`
 * (src/simulation/renderer_helpers.rs:52-72) drives once per frame.  Each
 * entry point below names the reference item it replaces.  Fields stay
 * resident on the GPU; host buffers are only borrowed for the duration of a
 * call.  Every function returns an eqf_status; on failure eqf_last_error()
 * gives a message (thread-local).  No function aborts or unwinds.
 *
 * Layout of every host array: row-major `idx(x, y) = x + y*size`
 * (fluid.rs:31-35), size*size elements, ring included.
 */
#ifndef EQFLUID_H
#define EQFLUID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EQF_ABI_VERSION 1

typedef enum eqf_status {
    EQF_OK = 0,
    EQF_EINVAL = 1,   /* bad argument (reference: panics, e.g. obstacle.rs:67) */
    EQF_ENOMEM = 2,   /* device or host allocation failed / budget exceeded */
    EQF_EDEVICE = 3,  /* HIP runtime error */
    EQF_EBUSY = 4,    /* object in use (scene owned by a batch or server) */
    EQF_ENOTFOUND = 5 /* unknown ticket */
} eqf_status;

/* The six fields of `Fluid` (fluid.rs:59-77). */
typedef enum eqf_field {
    EQF_DENSITY = 0,    /* pub density */
    EQF_SCRATCH = 1,    /* scratch_space (private in the reference) */
    EQF_VX = 2,         /* pub velocities_x */
    EQF_VY = 3,         /* pub velocities_y */
    EQF_VX0 = 4,        /* velocities_x0 (private) */
    EQF_VY0 = 5         /* velocities_y0 (private) */
} eqf_field;

/* `ContainerWall` (fluid.rs:12-17), one byte per cell, declaration order. */
enum { EQF_DEFAULT_WALL = 0, EQF_NO_WALL = 1 };

/* Gauss-Seidel ordering of every lin_solve (fluid.rs:301-325). */
typedef enum eqf_gs_mode {
    EQF_GS_WAVEFRONT = 0, /* lexicographic order of the reference: bit-exact */
    EQF_GS_REDBLACK = 1   /* red-black order: parallel, checked against FP64 */
} eqf_gs_mode;

/* `SimulationConfigs` (configs.rs:5-12) + `FluidConfigs` (configs.rs:37-48)
 * minus the colours, plus the GPU knobs. */
typedef struct eqf_params {
    uint32_t size;          /* SimulationConfigs.size, ring included, >= 3 */
    float delta_t;          /* SimulationConfigs.delta_t */
    int64_t frames;         /* SimulationConfigs.frames */
    float diffusion;        /* FluidConfigs.diffusion */
    float viscosity;        /* FluidConfigs.viscousity */
    int64_t gs_iters;       /* sweeps per diffuse lin_solve; < 0: `frames`, as the reference */
    int64_t pressure_iters; /* sweeps per project lin_solve; < 0: same as diffuse */
    int32_t gs_mode;        /* eqf_gs_mode */
    int32_t init_times;     /* 1: state of Fluid::new (fluid.rs:93); 2: Fluid::default (fluid.rs:83) */
} eqf_params;

typedef struct eqf_ctx eqf_ctx;
typedef struct eqf_scene eqf_scene;
typedef struct eqf_batch eqf_batch;
typedef struct eqf_server eqf_server;

/* Defaults of SimulationConfigs::default (configs.rs:15-21) and
 * FluidConfigs::default (configs.rs:51-59); init_times = 1. */
void eqf_params_default(eqf_params *p);

/* Message of the last failed call on this thread ("" if none). */
const char *eqf_last_error(void);

/* One context per GPU: owns the device selection. */
int eqf_ctx_create(int device, eqf_ctx **out);
int eqf_ctx_destroy(eqf_ctx *ctx);
/* No counterpart in the reference (host plumbing for tuning caches): writes
 * "<device name>/<gfx arch>/<CU count>CU" of the context's GPU, NUL-terminated,
 * into buf ("gpu" when the driver reports no name); EQF_EINVAL if it does not
 * fit in len bytes. */
int eqf_ctx_device_key(const eqf_ctx *ctx, char *buf, size_t len);

/* Fluid::new / Fluid::default (fluid.rs:83-110): allocates the six fields and
 * the mask on the device and writes the reference's initial state. */
int eqf_scene_create(eqf_ctx *ctx, const eqf_params *p, eqf_scene **out);
/* #[derive(Clone)] on Fluid (fluid.rs:51): device-to-device deep copy. */
int eqf_scene_clone(const eqf_scene *src, eqf_scene **out);
int eqf_scene_destroy(eqf_scene *s);
int eqf_scene_get_params(const eqf_scene *s, eqf_params *out);
/* Changes delta_t, diffusion, viscosity, iteration counts; size and mode are fixed. */
int eqf_scene_set_params(eqf_scene *s, const eqf_params *p);

/* Rectangle::are_all_points_valid (obstacle.rs:74-87); EQF_EINVAL with
 * "Invalid input for Rectangle" (obstacle.rs:67) when invalid. */
int eqf_rect_validate(int64_t x0, int64_t y0, int64_t x1, int64_t y1, uint32_t size);
/* Fluid::fill_obstacle (fluid.rs:610-619): walls on [x0,x1) x [y0,y1), clamped. */
int eqf_fill_rect(eqf_scene *s, int64_t x0, int64_t y0, int64_t x1, int64_t y1);
/* pub cells_type (fluid.rs:80): read or replace the whole mask.  The ring
 * must stay DefaultWall (init_walls, fluid.rs:552-570). */
int eqf_set_mask(eqf_scene *s, const uint8_t *mask);
int eqf_get_mask(const eqf_scene *s, uint8_t *mask);

/* add_density / add_velocity (fluid.rs:120-131), made public: one cell,
 * clamped index; density is added to the density and scratch fields. */
int eqf_add_density(eqf_scene *s, uint32_t x, uint32_t y, float amount);
int eqf_add_velocity(eqf_scene *s, uint32_t x, uint32_t y, float ax, float ay);
/* One cell of one field, with eqf_add_source's field meaning: EQF_DENSITY is
 * add_density(x, y, amount); EQF_VX / EQF_VY add to that velocity component
 * only (add_velocity with the other component untouched). */
int eqf_add_cell(eqf_scene *s, int field, uint32_t x, uint32_t y, float amount);
/* Fluid::add_noise (fluid.rs:575-599) with its random inputs supplied by the
 * caller: the line from the centre cell (size/2, size/2) to (rand_x, rand_y)
 * is rotated by `angle_degrees` about the centre, and twice the rotated end
 * point is added to the centre cell's velocity.  The reference draws the
 * angle from noise-0.7 Perlin at [dt, dt] and the point from an unseeded
 * thread_rng; neither is reproduced (its parity is unpinned). */
int eqf_add_noise(eqf_scene *s, float angle_degrees, uint32_t rand_x, uint32_t rand_y);
/* add_density(x, y, a[idx]) for every cell (field EQF_DENSITY), or
 * add_velocity with one component (EQF_VX / EQF_VY). */
int eqf_add_source(eqf_scene *s, int field, const float *amounts);
/* Same as eqf_add_source, but applied on the device before every frame of
 * eqf_step / batch / server (the frame loop of renderer_helpers.rs:54-59
 * calls add_noise before step).  NULL clears it. */
int eqf_set_frame_source(eqf_scene *s, int field, const float *amounts);

/* Read / write a whole field (the pub Vec<f32> fields, fluid.rs:61-77). */
int eqf_upload(eqf_scene *s, int field, const float *host);
int eqf_download(eqf_scene *s, int field, float *host);

/* Fluid::step (fluid.rs:437-524), `nframes` times, each preceded by the
 * frame sources.  Asynchronous on the scene's stream. */
int eqf_step(eqf_scene *s, int64_t nframes);
/* Waits for the scene's queued work. */
int eqf_sync(eqf_scene *s);

/* Red-black mode tiling (no counterpart in the reference): band shape index
 * (0: 8 waves x 16 rows, 1: 8 waves x 32 rows, 2: 16 waves x 16 rows, 3: 8
 * waves x 18 rows, 4: 8 waves x 14 rows, 5: 8 waves x 10 rows; regions 128
 * columns wide) and Gauss-Seidel sweeps per launch (1, 2, 4, 5, 6 or 8); 5
 * and 5 until set.
 * Results do not depend on it.  Chosen by the autotuner
 * (equilibrium_amd/tune.py); the server's red-black groups use the default. */
int eqf_scene_set_tuning(eqf_scene *s, int tile, int sweeps_per_block);

/* Batches: many scenes of one mode advanced together; K frames are captured
 * once as a HIP graph and replayed.  Wavefront mode (sizes up to 1026): one
 * fused launch per frame runs every scene's whole frame.  Red-black mode (any
 * sizes, one iteration count): every kernel of the frame covers all scenes
 * (band passes over all scenes' tiles, then the grid-wide phases).  Scenes
 * stay owned by the caller.  While in a batch a scene is not stepped alone
 * and its mask, parameters and frame sources are fixed (eqf_step,
 * eqf_fill_rect, eqf_set_mask, eqf_scene_set_params, eqf_set_frame_source
 * and eqf_scene_destroy return EQF_EBUSY).  The other scene calls (adds,
 * uploads, downloads, clone, render, snapshots, eqf_sync) stay available and
 * are ordered with the batch: each runs after the batch steps queued before
 * it, and the next eqf_batch_step runs after it -- the reference's frame loop
 * (add sources, then step) works on batched scenes.  A scene appears in a
 * batch at most once (EQF_EINVAL). */
int eqf_batch_create(eqf_ctx *ctx, eqf_scene *const *scenes, int n, eqf_batch **out);
int eqf_batch_step(eqf_batch *b, int64_t nframes);
int eqf_batch_sync(eqf_batch *b);
int eqf_batch_destroy(eqf_batch *b);
/* Wall time of the last eqf_batch_step in milliseconds (HIP events). */
int eqf_batch_last_ms(const eqf_batch *b, double *ms);
/* Launch shape of a batch (no counterpart in the reference): waves per
 * workgroup at most (1..16, at least the largest scene's strip count of
 * ceil(N/64); fewer waves means fewer sweep streams per strip) and the
 * scene-to-workgroup order.  Results do not depend on it.  Chosen by the
 * autotuner (equilibrium_amd/tune.py). */
enum { EQF_BATCH_LARGEST_FIRST = 0, EQF_BATCH_GIVEN_ORDER = 1 };
int eqf_batch_set_tuning(eqf_batch *b, int max_waves, int order);
/* Wavefront schedule of a wavefront-mode batch (no counterpart in the
 * reference; results do not depend on it): EQF_SCHED_WIDE, a workgroup
 * barrier every 16 steps, every sweep stream reading a lin_solve's
 * right-hand side from device memory; EQF_SCHED_RING, a barrier every 8
 * steps, the right-hand side handed from sweep stream to sweep stream in
 * LDS and up to P/48 streams per strip instead of P/64; EQF_SCHED_FINE, a
 * barrier every 4 steps, 4 steps of prefetch, the LDS hand-off and up to
 * P/32 streams per strip (small scenes); EQF_SCHED_AUTO (the default): FINE
 * when it gives the batch's largest scene more sweep streams than RING (N
 * 49..192 with max_waves 16), else RING when it gives more than WIDE or two
 * or more streams and at least 8 waves (N 193..512), else WIDE.  Single
 * scenes (eqf_step) and server ticks use the AUTO rule.  Red-black batches:
 * EQF_EINVAL. */
enum { EQF_SCHED_AUTO = 0, EQF_SCHED_WIDE = 1, EQF_SCHED_RING = 2, EQF_SCHED_FINE = 3 };
int eqf_batch_set_schedule(eqf_batch *b, int sched);
/* Band shape and sweeps per launch of a red-black batch (as
 * eqf_scene_set_tuning); chosen by the autotuner. */
int eqf_batch_set_band(eqf_batch *b, int tile, int sweeps_per_block);

/* Serving: continuous batching of scene requests (no counterpart in the
 * reference, whose frame loop runs one scene; renderer_helpers.rs:52-72).
 * A server thread admits queued requests while their device memory fits the
 * budget, keeps admitted scenes resident, and advances every resident scene
 * by one frame per tick: wavefront-mode scenes (size <= 1026) with one fused
 * launch, red-black scenes (any size) with their frame graphs on a second
 * stream beside it; scenes join and leave at tick boundaries.  Admission
 * uploads overlap the running tick. */
typedef struct eqf_server_config {
    uint64_t hbm_budget_bytes; /* device memory the resident scenes may use */
    int32_t max_active;        /* resident scenes at most (0: no limit) */
    int32_t reserved;
} eqf_server_config;

typedef struct eqf_request {
    eqf_params params;         /* size, delta_t, diffusion, viscosity, gs_iters, gs_mode */
    int64_t nframes;           /* frames to advance */
    int32_t n_rects;           /* obstacles: n_rects x (x0, y0, x1, y1), fill_obstacle semantics */
    int32_t keep_velocity;     /* 1: also keep vx, vy for eqf_server_result */
    int32_t borrow_inputs;     /* 1: the arrays below stay valid until the request leaves the
                                  queue (state != EQF_REQ_QUEUED) and are read in place;
                                  0: eqf_server_submit copies them */
    int32_t reserved;
    const int64_t *rects;
    const float *src_density;  /* per-frame sources (size*size, row-major) or NULL */
    const float *src_vx;
    const float *src_vy;
    const float *init_vx;      /* initial velocity replacing Fluid::new's, or NULL */
    const float *init_vy;
} eqf_request;

enum { EQF_REQ_QUEUED = 0, EQF_REQ_ACTIVE = 1, EQF_REQ_DONE = 2, EQF_REQ_FAILED = 3, EQF_REQ_CANCELLED = 4 };

typedef struct eqf_request_status {
    int32_t state;
    int32_t reserved;
    int64_t frames_done;
    double submit_ms, admit_ms, done_ms; /* since server start */
} eqf_request_status;

typedef struct eqf_server_stats {
    int64_t ticks;
    int64_t scene_frames;
    double cell_updates;      /* sum over resident scenes and ticks of N^2 */
    double busy_ms;           /* wall time with at least one resident scene */
    double p50_frame_ms;      /* per scene-frame latency percentiles (tick durations) */
    double p99_frame_ms;
    double p50_queue_ms;      /* submit -> admit */
    double p99_queue_ms;
    int64_t max_resident;
    uint64_t peak_bytes;
} eqf_server_stats;

int eqf_server_create(eqf_ctx *ctx, const eqf_server_config *cfg, eqf_server **out);
int eqf_server_submit(eqf_server *sv, const eqf_request *req, uint64_t *ticket);
int eqf_server_poll(eqf_server *sv, uint64_t ticket, eqf_request_status *st);
/* Copies a finished request's field (EQF_DENSITY, or EQF_VX/EQF_VY when
 * keep_velocity was set) into `out` (size*size floats). */
int eqf_server_result(eqf_server *sv, uint64_t ticket, int field, float *out);
/* Forgets a ticket: a finished request's record is dropped at once; a queued
 * or resident one is cancelled and dropped when the server retires it (poll
 * then reports EQF_ENOTFOUND). */
int eqf_server_release(eqf_server *sv, uint64_t ticket);
/* Waits until no request is queued or resident (timeout_ms < 0: forever). */
int eqf_server_wait_idle(eqf_server *sv, double timeout_ms);
int eqf_server_stats_get(eqf_server *sv, eqf_server_stats *st);
/* The same statistics restricted to requests of size min_size..max_size
 * (frame and queue latency percentiles, scene_frames, cell_updates); ticks,
 * busy_ms, max_resident and peak_bytes stay server-wide. */
int eqf_server_stats_sizes(eqf_server *sv, uint32_t min_size, uint32_t max_size, eqf_server_stats *st);
int eqf_server_destroy(eqf_server *sv);

/* ---- the renderer's side: RGBA image and per-frame snapshots ---------- */

/* Renderer colours as RGBA bytes. */
typedef struct eqf_colors {
  uint8_t fluid[4];    /* FluidConfigs::fluid_color (configs.rs:56) */
  uint8_t world[4];    /* FluidConfigs::world_color (configs.rs:57) */
  uint8_t obstacle[4]; /* RenderingListener::obstacles_color (renderer_helpers.rs:101) */
} eqf_colors;
/* The reference defaults: (208,88,157,220), (94,146,162,128), (255,0,0,255). */
void eqf_colors_default(eqf_colors *c);

/* RenderingListener::render_image's pixels (renderer_helpers.rs:143-165),
 * computed on the device from the scene's density and mask: size*size RGBA8
 * pixels, row-major (x + y*size), into `rgba` (4*size*size bytes).  NULL
 * colours = eqf_colors_default.  Synchronous. */
int eqf_render_rgba(eqf_scene *s, const eqf_colors *c, uint8_t *rgba);

/* Allocation-free variant of eqf_render_rgba (no counterpart in the
 * reference; for callers that cannot allocate device memory: render loops
 * that must not call hipMalloc/hipFree, which also synchronise the device,
 * and harnesses that forbid it).  It calls no hipMalloc, hipFree or other
 * allocating routine: the density and image are staged in a fixed device
 * array of EQF_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it
 * runs on until the process ends.  Limits: size <= EQF_NOALLOC_RENDER_MAX_SIZE
 * (EQF_EINVAL beyond), and the scene's device mask must be current -- after
 * eqf_fill_rect / eqf_set_mask, step the scene (or call eqf_render_rgba) once
 * before using this call (EQF_EINVAL otherwise).  Calls on one GPU share that
 * array; the library serialises them with a lock (each call is synchronous),
 * so concurrent calls from several threads wait for each other rather than
 * race.  Same output as eqf_render_rgba. */
#define EQF_NOALLOC_SCRATCH_BYTES (36u << 20)
#define EQF_NOALLOC_RENDER_MAX_SIZE 2172
int eqf_render_rgba_noalloc(eqf_scene *s, const eqf_colors *c, uint8_t *rgba);

/* Per-frame snapshots: CurrentSimulation::simulate sends the renderer a
 * clone of the Fluid after every step (renderer_helpers.rs:59-65).  A ring of
 * `depth` (1..8) pinned host slots; a capture is queued behind the scene's
 * steps (density, row-major f32, and, with `with_rgba`, the RGBA image) and
 * its copy to host runs on a separate stream, overlapping later steps. */
typedef struct eqf_snapshot eqf_snapshot;
int eqf_snapshot_create(eqf_scene *s, int depth, int with_rgba, const eqf_colors *c, eqf_snapshot **out);
/* Queues a capture of the scene's state after its queued steps, tagged with
 * `frame`.  EQF_EBUSY when every slot holds an unreleased capture. */
int eqf_snapshot_capture(eqf_snapshot *sn, int64_t frame);
/* Waits for the oldest capture; the pointers (size*size floats, 4*size*size
 * bytes or NULL) stay valid until eqf_snapshot_release. */
int eqf_snapshot_acquire(eqf_snapshot *sn, int64_t *frame, const float **density, const uint8_t **rgba);
/* Frees the oldest capture's slot. */
int eqf_snapshot_release(eqf_snapshot *sn);
/* Waits for outstanding copies and frees the ring; call before destroying
 * the scene.  A ring and its scene are used from one thread at a time. */
int eqf_snapshot_destroy(eqf_snapshot *sn);

#ifdef __cplusplus
}
#endif
#endif
