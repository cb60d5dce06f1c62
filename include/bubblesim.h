/*
 * bubblesim.h -- C ABI of the MI355X-native lattice solver for a first-order
 * cosmological phase transition (scalar field phi/pi coupled by friction eta to
 * a relativistic bag-EOS fluid, periodic cubic lattice, SSP-RK3).
 *
 * The reference (AgenttiX/bubble-simulation) has no FFI: its boundary is the
 * Rust `Simulation` struct API plus a handful of #[repr(C)] byte layouts that
 * mirror WGSL uniforms.  Every declaration below names the reference item it
 * replaces (file:line, relative to the reference repository root).
 *
 * Conventions
 *   - Every entry point that can fail returns an int status (BS_OK == 0) and
 *     leaves a human-readable message in a thread-local string read with
 *     bs_last_error().  No entry point aborts the process.
 *   - Handles (bs_sim, bs_batch, bs_server) are opaque and owned by the
 *     library; destroy them with the matching *_destroy call.
 *   - The library owns all device memory.  Lattice state only crosses the
 *     boundary through the explicit download/upload debug calls.
 *   - One handle per host thread: calls on one bs_sim/bs_batch are not
 *     thread-safe.  bs_server_submit/poll/wait are thread-safe.
 *   - State layout (download/upload): two arrays per lattice, x fastest,
 *     linear index i = (z*ny + y)*nx + x  (common.wgsl:19-23):
 *       field[i] = (phi, pi)                 2 reals
 *       fluid[i] = (E, Zx, Zy, Zz)           4 reals
 *     where a real is float for BS_PRECISION_F32 and double for BS_PRECISION_F64.
 */
#ifndef BUBBLESIM_H
#define BUBBLESIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/*  Status codes (the reference uses anyhow::Result; sim.rs:764-781,        */
/*  physics.rs:191-217, headless.rs:63-67)                                  */
/* ------------------------------------------------------------------------ */
enum {
    BS_OK = 0,
    BS_ERR_INVALID = 1,   /* invalid parameters (Model::validate, Config::validate) */
    BS_ERR_OOM = 2,       /* device memory exhausted or over budget              */
    BS_ERR_DEVICE = 3,    /* HIP runtime error                                   */
    BS_ERR_NONFINITE = 4, /* non-finite diagnostics (headless.rs:63-67)          */
    BS_ERR_NOT_READY = 5, /* asynchronous result not available yet              */
    BS_ERR_NOT_FOUND = 6  /* unknown job id / box index                          */
};

/* Lattice precision of a simulation.  F32 is the reference's arithmetic
 * (physics.rs:21 `Scalar = f32`); F64 is the same scheme in double precision,
 * used as the accuracy oracle (docs/PRECISION.md:154-177). */
enum { BS_PRECISION_F32 = 0, BS_PRECISION_F64 = 1 };

/* Arithmetic policy of the stage kernels.
 *   BS_MATH_FAST : hardware reciprocal/square root (<= 2 ulp, inside WGSL's
 *                  2.5 ulp bound for `/` and `sqrt`) and FMA contraction.
 *   BS_MATH_EXACT: IEEE correctly rounded division/square root, no FMA
 *                  contraction, reference operation order; bit-identical to a
 *                  sequential CPU evaluation of the same operation sequence.
 * F64 lattices always use EXACT. */
enum { BS_MATH_FAST = 0, BS_MATH_EXACT = 1 };

/* Stage kernel implementation.
 *   BS_KERNEL_FUSED  : one fused launch per RK stage (primitive recovery,
 *                      MUSCL/HLL fluxes computed once per face, friction,
 *                      field update and stage combination) with an LDS plane
 *                      ring marching along z.
 *   BS_KERNEL_TWO_PASS: primitive pass + stage pass per RK stage, the
 *                      reference's dispatch structure (sim.rs:881-916). */
enum { BS_KERNEL_FUSED = 0, BS_KERNEL_TWO_PASS = 1 };

/* ------------------------------------------------------------------------ */
/*  Physics model (physics.rs:23-83)                                        */
/* ------------------------------------------------------------------------ */

/* SeedSize (physics.rs:29-42) */
enum { BS_SEED_CRITICAL = 0, BS_SEED_CELLS = 1 };

/* NucleationMode (physics.rs:224-234).  UNIFORM_STEPS is an extension: times
 * are whole steps drawn uniformly in [0, uniform_max_step] from a generator
 * seeded by `time_seed`, separate from the position stream, so positions stay
 * identical to the reference's for the same `seed`. */
enum { BS_NUC_SIMULTANEOUS = 0, BS_NUC_EXPONENTIAL = 1, BS_NUC_UNIFORM_STEPS = 2 };

/* Equation of state behind the swap point of docs/EOS.md:10-34:
 *   BS_EOS_BAG      : e = 3p, closed-form inversion (eos.wgsl, the default);
 *   BS_EOS_GAMMA_LAW: p = (eos_gamma - 1) e, c_s^2 = eos_gamma - 1, inverted
 *                     by a fixed count of Newton steps (docs/EOS.md:86-128);
 *                     eos_gamma = 4/3 is the bag model.
 * dof_ratio (docs/EOS.md:65-84) interpolates the radiation constant of the
 * temperature across the wall, a_broken / a_symmetric; it changes only the
 * temperature channel of the visualisation.  A batch or a server may mix
 * both models (one stage launch per model present). */
enum { BS_EOS_BAG = 0, BS_EOS_GAMMA_LAW = 1 };

/* Model (physics.rs:44-68).  seed_kind/seed_value encode SeedSize. */
typedef struct bs_model {
    float lambda;
    float eps_ratio;
    float alpha;
    float eta;
    float phi_b;
    float dx;
    float cfl;
    int32_t seed_kind;   /* BS_SEED_CRITICAL: factor of R_c; BS_SEED_CELLS: cells */
    float seed_value;
} bs_model;

/* Derived quantities (physics.rs:90-189), all evaluated in f32 in the
 * reference's operation order. */
typedef struct bs_derived {
    float eps;                     /* physics.rs:90-92   */
    float m2;                      /* physics.rs:95-97   */
    float delta;                   /* physics.rs:100-102 */
    float wall_width;              /* physics.rs:109-111 */
    float surface_tension;         /* physics.rs:114-116 */
    float critical_radius;         /* physics.rs:120-122 */
    float e_ref;                   /* physics.rs:125-127 */
    float t_ref;                   /* physics.rs:129-131 */
    float a_rad;                   /* physics.rs:134-136 */
    float dt;                      /* physics.rs:138-140 */
    float seed_radius;             /* physics.rs:154-160 */
    float critical_to_wall_ratio;  /* physics.rs:173-175 */
    float estimated_wall_velocity; /* physics.rs:180-184 */
    float sound_speed;             /* physics.rs:187-189 */
} bs_derived;

/* BubbleEvent (physics.rs:236-241) */
typedef struct bs_event {
    float time;
    float pos[3];
    float radius;
} bs_event;

/* ------------------------------------------------------------------------ */
/*  Byte layouts mirrored from the WGSL uniforms (sim.rs:66-131)            */
/* ------------------------------------------------------------------------ */

/* SimParams (common.wgsl:25-58, SimParamsGpu sim.rs:66-98): 96 bytes. */
typedef struct bs_sim_params {
    uint32_t nx, ny, nz, pad_u0;
    float dx, inv_dx, dt, eta;
    float lambda, eps, phi_b, m2;
    float delta, a_rad, wall_width, p_floor;
    float e_floor, z_max, e_ref, t_ref;
    float time, vis_gain, pad0, pad1;
} bs_sim_params;

/* StageParams (common.wgsl:63-68, StageParamsGpu sim.rs:100-107): 16 bytes. */
typedef struct bs_stage_params {
    float a0, a1, dt_stage, pad;
} bs_stage_params;

/* Bubble (nucleate.wgsl:29-35, BubbleGpu sim.rs:109-116): 16 bytes. */
typedef struct bs_bubble {
    float cx, cy, cz, r0;
} bs_bubble;

/* NucleationBatch (nucleate.wgsl:40-45, sim.rs:118-123): 16 bytes. */
typedef struct bs_nucleation_batch {
    uint32_t count, pad[3];
} bs_nucleation_batch;

/* Diagnostics readback block (reduce.wgsl:32-38, DiagnosticsGpu
 * sim.rs:125-131): 32 bytes.  Filled by bs_sim_diagnostics_raw with the
 * FP64-accumulated sums rounded to f32. */
typedef struct bs_diagnostics_gpu {
    float sums[4];
    uint32_t max_v_bits;
    uint32_t pad[3];
} bs_diagnostics_gpu;

/* Host diagnostics (Diagnostics, sim.rs:141-167).  Means are the FP64
 * compensated sums divided by the cell count in FP64; the reference sums in
 * f32 and divides in f32 (reduce.wgsl:81-140, sim.rs:991-1010). */
typedef struct bs_diagnostics {
    double mean_energy;        /* <1/2 pi^2 + 1/2 |grad phi|^2 + V + max(E,e_floor)> */
    double broken_fraction;    /* <[phi > phi_b/2]>                                   */
    double mean_kinetic;       /* <(max(E,e_floor) + p) v^2>                          */
    double mean_fluid_energy;  /* <max(E,e_floor)>                                    */
    double max_velocity;       /* max |v|                                             */
    double kinetic_fraction;   /* Diagnostics::kinetic_fraction, sim.rs:160-166       */
    uint64_t cells;
    uint32_t nonfinite;        /* non-zero if any summand was NaN/Inf                 */
    uint32_t pad;
} bs_diagnostics;

/* Everything Simulation::new needs (SimulationSpec sim.rs:246-255, built by
 * Config::sim_spec config.rs:138-149), plus this library's extensions. */
typedef struct bs_spec {
    bs_model model;
    uint32_t grid[3];
    int32_t nucleation;          /* BS_NUC_*                                   */
    uint32_t bubbles;            /* config.rs:65-67                            */
    float nucleation_duration;   /* config.rs:73-75                            */
    uint64_t seed;               /* config.rs:77-79                            */
    /* extensions */
    int32_t precision;           /* BS_PRECISION_*                             */
    int32_t math;                /* BS_MATH_*                                  */
    int32_t kernel;              /* BS_KERNEL_*                                */
    int32_t bubble_at_centre;    /* 1: single bubble at (nx/2, ny/2, nz/2)     */
    uint32_t uniform_max_step;   /* BS_NUC_UNIFORM_STEPS: times in [0, this]   */
    int32_t eos;                 /* BS_EOS_*                                   */
    uint64_t time_seed;          /* BS_NUC_UNIFORM_STEPS time generator seed   */
    float eos_gamma;             /* BS_EOS_GAMMA_LAW: 1 < eos_gamma < 2        */
    float dof_ratio;             /* temperature dof ratio across the wall (> 0) */
} bs_spec;

/* ------------------------------------------------------------------------ */
/*  Library / error surface                                                 */
/* ------------------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char *bs_last_error(void);
/* Library version string. */
const char *bs_version(void);
/* sizeof of each ABI struct, so bindings can check their mirrors
 * (parameter_blocks_have_the_expected_size, sim.rs:1070-1078).
 * which: 0 sim_params, 1 stage_params, 2 bubble, 3 nucleation_batch,
 * 4 diagnostics_gpu, 5 diagnostics, 6 spec, 7 model, 8 event, 9 derived,
 * 10 server_config, 11 job, 12 job_result, 13 server_stats. */
size_t bs_struct_size(int which);

/* ------------------------------------------------------------------------ */
/*  Host physics (physics.rs, config.rs, sim.rs host helpers)               */
/* ------------------------------------------------------------------------ */

/* Model::default (physics.rs:70-83). */
void bs_model_default(bs_model *out);
/* Model::validate (physics.rs:191-217).  BS_OK or BS_ERR_INVALID. */
int bs_model_validate(const bs_model *m);
/* Model derivations (physics.rs:90-189). */
int bs_model_derive(const bs_model *m, bs_derived *out);
/* make_schedule (physics.rs:248-296): Pcg64Mcg positions with 2.5 R0
 * rejection (<= 64 attempts), simultaneous/exponential times, stable sort by
 * time.  `out` holds `count` events.  Bit-exact with the reference for
 * BS_NUC_SIMULTANEOUS and BS_NUC_EXPONENTIAL.  For BS_NUC_UNIFORM_STEPS,
 * `duration` is unused and times are k*dt for k drawn in [0, uniform_max_step]
 * from `time_seed`; `dt` must then be the model's time step. */
int bs_make_schedule(int32_t mode, uint32_t count, float radius, const uint32_t grid[3],
                     float duration, uint64_t seed, uint32_t uniform_max_step,
                     uint64_t time_seed, float dt, bs_event *out);
/* Raw Pcg64Mcg stream (rand_pcg 0.9 Mcg128Xsl64, seeded by rand_core 0.9
 * SeedableRng::seed_from_u64) -- the generator make_schedule draws from
 * (physics.rs:256, 278-280).  Writes n next_u64 outputs. */
void bs_pcg64mcg_u64(uint64_t seed, uint32_t n, uint64_t *out);
/* rand 0.9 `random::<f32>()` on the same stream: n floats in [0, 1). */
void bs_pcg64mcg_f32(uint64_t seed, uint32_t n, float *out);
/* Mcg128Xsl64::new(state) then n next_u64 (state given as lo/hi halves). */
void bs_mcg128_from_state(uint64_t state_lo, uint64_t state_hi, uint32_t n, uint64_t *out);

/* gpu_params (sim.rs:613-644): the 96-byte parameter block for a model and
 * grid at time `time`. */
int bs_sim_params_fill(const bs_model *m, const uint32_t grid[3], float time,
                       bs_sim_params *out);
/* RK3 (sim.rs:135) as (a0, a1) x 3 -> out[6]; upload_stage_params
 * (sim.rs:650-656) uses dt_stage = dt for every stage. */
void bs_rk3_coefficients(float out[6]);
/* stage_slots (sim.rs:1046-1051): out[2*s] = input slot, out[2*s+1] = output
 * slot of stage s, for U^n in slot `cur`. */
void bs_stage_slots(uint32_t cur, uint32_t out[6]);
/* Step index at which each event is stamped: the first step s whose f32
 * accumulated time (sum of s dt's, sim.rs:913-915) is >= event time, as
 * record_nucleation pops events with time <= self.time before each step
 * (sim.rs:843-879, headless.rs:126-131). */
int bs_event_steps(const bs_event *events, uint32_t n, float dt, uint64_t *steps_out);

/* Config::validate (config.rs:151-178) applied to a spec. */
int bs_spec_validate(const bs_spec *spec);
/* The spec Config::parse_from(["bubble-simulation"]) produces (config.rs
 * defaults), with precision F32, fast math and the fused kernel. */
void bs_spec_default(bs_spec *out);
/* Parse a --grid value: "192" or "256,128,64" (parse_grid config.rs:194-203). */
int bs_parse_grid(const char *s, uint32_t grid_out[3]);

/* Device bytes per lattice cell held by a simulation of this precision:
 * three RK slots of (field 2 + fluid 4 reals), plus the primitive cache of
 * the two-pass kernel (4 reals) when `kernel` is BS_KERNEL_TWO_PASS.  The
 * reference's figure is 96 B/cell (Simulation::lattice_bytes sim.rs:705-716). */
uint64_t bs_bytes_per_cell(int32_t precision, int32_t kernel);
/* cells * bs_bytes_per_cell (lattice_bytes sim.rs:712-716). */
uint64_t bs_lattice_bytes(const uint32_t grid[3], int32_t precision, int32_t kernel);
/* max_cubic_size (sim.rs:718-734) with a memory budget only (HIP has no
 * storage-binding ceiling).  side is clamped to [8, 2048]. */
int bs_max_cubic_size(uint64_t budget_bytes, int32_t precision, int32_t kernel,
                      uint32_t *side_out, int32_t *limited_by_memory_out);
/* lattice_budget (vram.rs:58,67-69): (available + current) * 0.85. */
uint64_t bs_lattice_budget(uint64_t available, uint64_t current_lattice);

/* ------------------------------------------------------------------------ */
/*  Device                                                                   */
/* ------------------------------------------------------------------------ */

/* Number of visible HIP devices (0 if none). */
int bs_device_count(void);
/* Free/total device memory of the current device (hipMemGetInfo). */
int bs_device_memory(uint64_t *free_out, uint64_t *total_out);
/* Device name into buf. */
int bs_device_name(char *buf, size_t len);

/* Achievable HBM bandwidth of the current device, in GB/s (1e9 B/s):
 * streaming read-only, write-only and copy kernels over two buffers of
 * `bytes` each (>= 1 MiB; far above the caches, e.g. 2 GiB), `repeats`
 * launches per measurement, best over grids of 8..64 workgroups per CU.
 * No counterpart in the reference: it gives the practical ceiling that the
 * stage kernel's measured HBM traffic is compared against, beside the
 * published peak.  Allocates and frees its buffers. */
typedef struct bs_hbm_rates {
    double read_gbps;
    double write_gbps;
    double copy_gbps; /* bytes read + bytes written, per second */
} bs_hbm_rates;
int bs_hbm_probe(uint64_t bytes, uint32_t repeats, bs_hbm_rates *out);

/* ------------------------------------------------------------------------ */
/*  Simulation (Simulation, sim.rs:180-1026)                                */
/* ------------------------------------------------------------------------ */
typedef struct bs_sim bs_sim;

/* Simulation::new (sim.rs:257-607): validate, allocate, build the schedule
 * (make_schedule), init the lattice (init.wgsl) at time 0. */
int bs_sim_create(const bs_spec *spec, bs_sim **out);
void bs_sim_destroy(bs_sim *sim);
/* Simulation::reset (sim.rs:790-821).  schedule == NULL keeps the current
 * schedule; otherwise it is replaced by the n given events. */
int bs_sim_reset(bs_sim *sim, const bs_event *schedule, uint32_t n);
/* Simulation::set_model (sim.rs:663-670).  An invalid model is ignored and
 * BS_ERR_INVALID returned; the simulation is untouched. */
int bs_sim_set_model(bs_sim *sim, const bs_model *m);
/* Simulation::record_steps (sim.rs:829-841): stamp due events once, then n
 * steps.  Asynchronous on the simulation's stream. */
int bs_sim_record_steps(bs_sim *sim, uint32_t n);
/* headless::advance's stepping loop (headless.rs:126-131): n x
 * record_steps(1), i.e. due events are stamped before every step. */
int bs_sim_advance(bs_sim *sim, uint64_t n);
/* record_diagnostics + after_submit + poll_diagnostics (sim.rs:951-1010),
 * blocking.  Primitives are recomputed from the current state inside the
 * reduction, so no separate refresh is needed (headless.rs:135-137). */
int bs_sim_diagnostics(bs_sim *sim, bs_diagnostics *out);
/* The same reduction as the reference's 32-byte block. */
int bs_sim_diagnostics_raw(bs_sim *sim, bs_diagnostics_gpu *out);
/* Non-blocking form: request enqueues the reduction, poll returns
 * BS_ERR_NOT_READY until it has completed (after_submit/poll_diagnostics). */
int bs_sim_request_diagnostics(bs_sim *sim);
int bs_sim_poll_diagnostics(bs_sim *sim, bs_diagnostics *out);
/* Block until all queued work on the simulation's stream has finished. */
int bs_sim_synchronize(bs_sim *sim);
/* Simulation::resize (sim.rs:736-762): validates first; a rejected size
 * leaves the simulation unchanged.  Re-initialises the lattice.  If the new
 * lattice then fails to allocate, the simulation is left at the minimum 8^3
 * lattice with its other parameters, as the reference's is (sim.rs:756-760),
 * and the error is returned. */
int bs_sim_resize(bs_sim *sim, const uint32_t grid[3]);
/* Simulation::nucleate_now (sim.rs:1012-1025): queue a bubble at `pos`
 * (lattice cells) stamped before the next step. */
int bs_sim_nucleate_now(bs_sim *sim, const float pos[3]);
/* Public fields / getters (sim.rs:180-242, 686-703). */
uint64_t bs_sim_cell_count(const bs_sim *sim);
uint32_t bs_sim_remaining_bubbles(const bs_sim *sim);
float bs_sim_time(const bs_sim *sim);
uint64_t bs_sim_steps(const bs_sim *sim);
int32_t bs_sim_nucleated_this_frame(const bs_sim *sim);
uint64_t bs_sim_device_memory_bytes(const bs_sim *sim);
int bs_sim_grid(const bs_sim *sim, uint32_t grid_out[3]);
int bs_sim_model(const bs_sim *sim, bs_model *out);
/* Copy of the schedule (remaining + already stamped), n_out = total. */
int bs_sim_schedule(const bs_sim *sim, bs_event *out, uint32_t cap, uint32_t *n_out);
/* Simulation::record_visualisation (sim.rs:920-942; vis.wgsl:32-55): pack
 * slot `cur` into one rgba16f texel per cell, x fastest, channels
 * (phi/phi_b, |v|, T/T_ref - 1, (E+p)|v|^2/e_ref) as IEEE half bits, round
 * to nearest even.  The texel buffer (8 B per cell) is allocated on the
 * first call and counted by bs_sim_device_memory_bytes; it replaces the
 * reference's vis_tex storage texture.  Enqueued on the simulation's stream. */
int bs_sim_record_visualisation(bs_sim *sim);
/* Device address and size of the texel buffer, for a renderer on the same
 * device (BS_ERR_NOT_READY before the first record_visualisation). */
int bs_sim_visualisation_buffer(bs_sim *sim, void **device_ptr, uint64_t *bytes);
/* Blocking copy of the texels: 4 halves per cell (n_halfs >= 4 * cells). */
int bs_sim_download_visualisation(bs_sim *sim, uint16_t *out, uint64_t n_halfs);

/* ---- offscreen volume rendering (--screenshot) ----------------------------
 * The reference draws the texels with a ray-marching fragment shader
 * (volume.wgsl, gpu/render.rs) and writes one offscreen frame to a PNG
 * (capture.rs:19-117).  Here the same ray march is a HIP kernel, one thread
 * per pixel, with the trilinear, clamp-to-edge texel filter written out. */

/* RenderSettings (render.rs:73-126); field_mode: 0 phi, 1 speed,
 * 2 temperature, 3 kinetic (FieldMode); colormap: 0 viridis, 1 inferno,
 * 2 signed; clip_axis -1 disables the cutaway. */
typedef struct bs_render_settings {
    int32_t field_mode;
    int32_t colormap;
    float iso_level;
    float wall_opacity;
    float samples_per_cell;
    float exposure;
    float field_gain;
    float absorption;
    int32_t show_iso;
    int32_t show_volume;
    int32_t show_box;
    float background;
    int32_t clip_axis;
    float clip_pos;
} bs_render_settings;

/* The view uniform of volume.wgsl:22-31 / ViewUniform (render.rs:128-139):
 * inverse view-projection (column-major), eye, box half-extents with the
 * cell size in w, grid, and the packed settings; 176 bytes. */
typedef struct bs_render_view {
    float inv_view_proj[16];
    float cam_pos[4];
    float box_half[4];
    float grid[4];
    float render0[4]; /* iso_level, wall_opacity, samples_per_cell, exposure */
    float render1[4]; /* field_mode, clip_axis, clip_pos, colormap */
    float render2[4]; /* field_gain, absorption, show_iso, show_volume */
    float render3[4]; /* frame jitter, show_box, background, unused */
} bs_render_view;

/* RenderSettings::with_field (render.rs:96-105): the defaults with the
 * field's gain and colour map (FieldMode::default_gain/default_colormap). */
int bs_render_settings_default(int32_t field_mode, bs_render_settings *out);
/* The uniform VolumeRenderer::update writes (render.rs:288-334) for the
 * default orbit camera framed on the box (camera.rs:24-34, frame_box
 * camera.rs:98-101; box_dimensions render.rs:363-371), a width x height
 * target, and frame number `frame` (the dither offset is (frame % 64) *
 * 0.618034; capture.rs renders frame 1). */
int bs_render_view_make(const uint32_t grid[3], const bs_render_settings *settings, uint32_t width, uint32_t height,
                        uint32_t frame, bs_render_view *out);
/* volume.wgsl fs_main over the texels of the last
 * bs_sim_record_visualisation: writes width x height RGBA8 pixels, row 0 at
 * the top, to host memory (blocking).  BS_ERR_NOT_READY if nothing was
 * recorded yet. */
int bs_sim_render(bs_sim *sim, const bs_render_view *view, uint32_t width, uint32_t height, uint8_t *rgba_out);
/* capture::render_to_png (capture.rs:19-117): records the visualisation of
 * the current state, renders frame 1 with `settings` at width x height and
 * writes an 8-bit RGBA PNG to `path`. */
int bs_sim_screenshot(bs_sim *sim, const bs_render_settings *settings, uint32_t width, uint32_t height,
                      const char *path);
/* The PNG writer of bs_sim_screenshot: 8-bit RGBA, rows top to bottom,
 * zlib-compressed (no GPU involved). */
int bs_png_write(const char *path, const uint8_t *rgba, uint32_t width, uint32_t height);
/* Debug dump / load of U^n (slot `cur`), layout in the header comment. */
int bs_sim_download(bs_sim *sim, void *field, void *fluid);
int bs_sim_upload(bs_sim *sim, const void *field, const void *fluid);
/* Run n steps (advance semantics) and report the summed duration of the
 * stage-kernel launches, measured with HIP events on the simulation's stream
 * around each launch, and the wall time of the whole call. */
int bs_sim_advance_timed(bs_sim *sim, uint64_t n, double *stage_kernel_ms, double *wall_ms);

/* Result of bs_reduce_noalloc, written to device memory: the FP64
 * compensated sums of reduce.wgsl's per-cell terms (e_total, broken,
 * kinetic, E), max |v|, and a non-zero `nonfinite` if any term was NaN/Inf.
 * bs_diagnostics' means are these sums divided by the cell count. */
typedef struct bs_reduce_result {
    double sums[4];
    double max_velocity;
    double nonfinite;
} bs_reduce_result;

/* No counterpart in the reference: for callers that cannot allocate (code
 * under hipGraph capture, tight serving loops), the diagnostics reduction
 * of bs_sim_diagnostics over a lattice the caller holds in device memory,
 * allocating nothing.  field/fluid: device pointers to one state (layout as
 * bs_sim_download, in `precision`; the bag equation of state); params: its
 * SimParams block (bs_sim_params_fill); out_device: device memory for one
 * bs_reduce_result; stream: the hipStream_t to enqueue on (NULL: the null
 * stream).  Enqueues the reduction and returns.  The reduction's 512
 * partial results live in a fixed 40 KB device array of the library, one
 * per device, held until the process ends, whatever the lattice's size;
 * lattices bs_spec_validate would reject (an axis below 8 cells, an x-y
 * plane of 2^28 cells or more in f32, 2^27 in f64) are rejected with
 * BS_ERR_INVALID.  Calls that overlap on one device share that array and
 * race: serialise them (one stream, or events).  The sums equal
 * bs_sim_diagnostics' bit for bit on the same state. */
int bs_reduce_noalloc(const bs_sim_params *params, int32_t precision, const void *field, const void *fluid,
                      void *out_device, void *stream);

/* ------------------------------------------------------------------------ */
/*  Batched sweep: many independent boxes advanced together (no reference   */
/*  counterpart; each box follows Simulation semantics exactly)            */
/* ------------------------------------------------------------------------ */
typedef struct bs_batch bs_batch;

/* All specs must share precision and math.  Boxes may differ in grid, model
 * and schedule. */
int bs_batch_create(const bs_spec *specs, uint32_t n, bs_batch **out);
void bs_batch_destroy(bs_batch *b);
/* Advance every box by n steps (advance semantics per box).  use_graph != 0
 * captures steps into a hipGraph and replays it.  The steps are enqueued on
 * the batch's streams and the call returns without waiting for them
 * (bs_batch_synchronize waits; the other calls on a box order themselves
 * after them). */
int bs_batch_advance(bs_batch *b, uint64_t n, int32_t use_graph);
int bs_batch_synchronize(bs_batch *b);
uint32_t bs_batch_size(const bs_batch *b);
uint64_t bs_batch_total_cells(const bs_batch *b);
int bs_batch_diagnostics(bs_batch *b, uint32_t box, bs_diagnostics *out);
int bs_batch_download(bs_batch *b, uint32_t box, void *field, void *fluid);
/* Load U^n of one box (layout as bs_sim_upload).  Boxes never share a task,
 * so a box's state reaches no other box: this is how a host forces a failing
 * (e.g. non-finite) box to check that its neighbours are untouched. */
int bs_batch_upload(bs_batch *b, uint32_t box, const void *field, const void *fluid);
/* Steps of one box, counting those enqueued by bs_batch_advance (at most the
 * box's step target), whether or not the device has finished them. */
int bs_batch_steps(const bs_batch *b, uint32_t box, uint64_t *steps_out);
/* Tile shape chosen by the autotuner for RK stages 2 and 3 of this batch
 * (x, y). */
int bs_batch_tile(const bs_batch *b, uint32_t *tx, uint32_t *ty);
/* The autotuner's choice for both classes of stage launch:
 * out = (tx, ty, zc) of RK stage 1, then (tx, ty, zc) of stages 2 and 3.
 * It times every built tile shape with z chunks of 16, 24, 32 and 64 planes on
 * the batch's boxes with their bubbles stamped, before the first step. */
int bs_batch_plan(const bs_batch *b, uint32_t out[6]);
/* Steps advanced so far through replay of a captured hipGraph (25 steps per
 * graph; bs_batch_advance steps a remainder below that eagerly). */
uint64_t bs_batch_graph_steps(const bs_batch *b);
/* Steps advanced so far without the nucleation stamp launch: a step (or a
 * 25-step graph) in which no box of a stream group has an event due leaves
 * the stamp kernel out, as it would write nothing (least over the groups).
 * Simulation::record_nucleation (sim.rs:843-860) likewise records no
 * nucleation pass when its batch of due events is empty. */
uint64_t bs_batch_stamps_skipped(const bs_batch *b);
/* Streams the batch's boxes are spread over: groups of about equal cell
 * count, each stepped by its own launches (and graph) on its own stream, so
 * that the groups' launches overlap on the device. */
uint32_t bs_batch_streams(const bs_batch *b);

/* ------------------------------------------------------------------------ */
/*  Serving: continuous batching of sweep jobs under an HBM budget          */
/* ------------------------------------------------------------------------ */
typedef struct bs_server bs_server;

/* bs_server_config.flags */
enum {
    /* step every tick eagerly instead of replaying a captured hipGraph when
     * the active set is unchanged (the results are the same bits); for
     * tools that do not follow graph launches made from a second thread */
    BS_SERVER_NO_GRAPHS = 1
};

typedef struct bs_server_config {
    uint64_t hbm_budget_bytes;   /* 0: use (total device memory - reserve)     */
    uint64_t reserve_bytes;      /* kept free                                  */
    uint32_t max_active;         /* 0: unlimited                               */
    uint32_t chunk_steps;        /* steps per scheduling tick (>= 1) of jobs of
                                  * 128^3 cells and more; smaller jobs step
                                  * in a class of their own, 1 to 8 chunks
                                  * per tick                                   */
    int32_t precision;           /* BS_PRECISION_* for every job               */
    int32_t math;                /* BS_MATH_*                                  */
    uint32_t flags;              /* BS_SERVER_*                                */
    uint32_t pad;
} bs_server_config;

typedef struct bs_job {
    bs_spec spec;
    uint64_t steps;              /* steps to advance                            */
    int32_t keep_state;          /* 1: keep the final state for bs_server_fetch */
    uint32_t checksum;           /* 1: fill state_checksum (reads the state back) */
    /* Optional initial U^n (both or neither; layout as bs_sim_upload, in the
     * server's precision), copied at submission and loaded over the
     * initialised lattice when the job is admitted: a job restarted from a
     * saved state.  Its schedule still runs from step 0.  NULL: the
     * reference's initial state (init.wgsl:18-26). */
    const void *init_field;
    const void *init_fluid;
} bs_job;

typedef struct bs_job_result {
    uint64_t job_id;
    uint64_t steps;
    bs_diagnostics final_diag;
    /* diagnostics at the first tick boundary after the job's last bubble was
     * stamped, the headless run's energy baseline E0 (headless.rs:54-78):
     * energy drift = final_diag.mean_energy / baseline_diag.mean_energy - 1 */
    bs_diagnostics baseline_diag;
    uint64_t baseline_step;      /* steps done when the baseline was taken     */
    double submit_s;             /* seconds since server start                 */
    double admit_s;
    double finish_s;
    double latency_s;            /* finish - submit                            */
    uint64_t state_checksum;     /* FNV-1a of the final field+fluid bytes, when
                                    requested (keep_state or checksum), else 0  */
    int32_t status;              /* BS_OK; BS_ERR_NONFINITE when the state went
                                    NaN/infinite (checked every tick; the job is
                                    evicted at that tick, as the headless run
                                    aborts, headless.rs:63-67); BS_ERR_OOM /
                                    BS_ERR_DEVICE when admission or a launch
                                    failed                                      */
    uint32_t pad;
} bs_job_result;

typedef struct bs_server_stats {
    uint64_t submitted, admitted, completed, rejected;
    uint64_t site_updates;       /* sum over completed jobs of cells * steps   */
    uint64_t ticks;
    uint64_t bytes_in_use;
    uint64_t budget_bytes;
    uint32_t active, queued;
    double busy_s;               /* wall time spent advancing                  */
    double elapsed_s;            /* since start                                */
    uint64_t failed;             /* jobs finished with a status other than BS_OK */
    uint64_t nonfinite;          /* of which evicted as non-finite             */
} bs_server_stats;

/* Starts two host threads: a worker that admits jobs and steps the active
 * set on one stream, and a retire thread that, on a second stream, reads
 * back each finished job's diagnostics (and state, when asked for) and
 * returns its memory while the worker keeps the other jobs stepping. */
int bs_server_create(const bs_server_config *cfg, bs_server **out);
/* Thread-safe.  Rejects jobs that can never fit the budget (BS_ERR_OOM) and
 * invalid specs (BS_ERR_INVALID). */
int bs_server_submit(bs_server *s, const bs_job *job, uint64_t *job_id_out);
/* BS_OK with the result once the job has completed, BS_ERR_NOT_READY before. */
int bs_server_poll(bs_server *s, uint64_t job_id, bs_job_result *out);
/* Block until the job completes or timeout_s elapses (BS_ERR_NOT_READY). */
int bs_server_wait(bs_server *s, uint64_t job_id, double timeout_s, bs_job_result *out);
/* Final state of a job submitted with keep_state (layout as bs_sim_download). */
int bs_server_fetch_state(bs_server *s, uint64_t job_id, void *field, void *fluid);
/* Counters so far.  Returns the error of the most recent tick when that
 * tick failed (every job active in it then finished with that error), once;
 * a later tick that goes through clears it as well. */
int bs_server_stats_get(bs_server *s, bs_server_stats *out);
/* Stop accepting work, finish nothing further, free device memory. */
void bs_server_destroy(bs_server *s);

#ifdef __cplusplus
}
#endif

#endif /* BUBBLESIM_H */
