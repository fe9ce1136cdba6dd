/*
 * wq3d.h - C ABI of the MI355X WaveQ3D wavefront-propagation library
 * (libwq3d.so, built from underseamodelinglibrary_amd/csrc).
 *
 * The library replaces the hot path of usml::waveq3d::wave_queue
 * (waveq3d/wave_queue.h:78-831 of the reference) with FP64 HIP kernels for
 * gfx950.  Everything crossing this boundary is plain C: fixed-layout structs,
 * raw host pointers and sizes.  No exceptions cross it; every entry point
 * returns a wq3d_status and wq3d_last_error() holds the message of the last
 * failure on the calling thread.
 *
 * Mapping to the reference (file:line of the interface each entry replaces):
 *   wq3d_fan_create    wave_queue::wave_queue      waveq3d/wave_queue.cc:30-98
 *   wq3d_step          wave_queue::step            waveq3d/wave_queue.cc:155-192
 *   wq3d_drain_eigenrays   eigenray_notifier::notify_eigenray_listeners
 *                          (eigenrays/eigenray_notifier.cc:26-33), replayed
 *                          by the caller in the reference's listener order
 *   wq3d_drain_eigenverbs  eigenverb_notifier::notify_eigenverb_listeners
 *                          (eigenverbs/eigenverb_notifier.cc:13-33)
 *   wq3d_drain_reflections reflection_notifier::notify_reflection_listeners
 *                          (waveq3d/reflection_notifier.cc:15-38)
 *   wq3d_snapshot      wave_queue::curr()/next()/prev()/past()
 *                          (waveq3d/wave_queue.h:181-216)
 *   wq3d_env_create    ocean_model construction (ocean/ocean_model.h:62-70)
 *                      lowered to the supported analytic / gridded models
 *   wq3d_sum_eigenrays eigenray_collection::add_eigenray + sum_eigenrays
 *                          (eigenrays/eigenray_collection.cc:90-191)
 */
#ifndef WQ3D_H
#define WQ3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum wq3d_status {
    WQ3D_OK = 0,
    WQ3D_ERR_ARGUMENT = 1,     /* reference: std::invalid_argument          */
    WQ3D_ERR_UNSUPPORTED = 2,  /* model outside the supported set           */
    WQ3D_ERR_DEVICE = 3,       /* HIP runtime failure                       */
    WQ3D_ERR_CAPACITY = 4,     /* caller buffer too small / limit exceeded  */
    WQ3D_ERR_STATE = 5         /* call out of order                         */
} wq3d_status;

/* ------------------------------------------------------------- sequences */
/* find_index() rule of the seq_vector subclass (types/seq_linear.h:89-94,
 * types/seq_log.h:84, types/seq_data.cc:15-36). */
typedef enum wq3d_seq_search {
    WQ3D_SEQ_LINEAR = 0,
    WQ3D_SEQ_LOG = 1,
    WQ3D_SEQ_DATA = 2
} wq3d_seq_search;

/* One usml::types::seq_vector: values and increments as the reference
 * stores them (types/seq_vector.h:99-143).  increment[i] is value[i+1] -
 * value[i] as accumulated by the subclass; the last entry repeats. */
typedef struct wq3d_seq {
    int64_t size;
    const double* value;
    const double* increment;
    int32_t search; /* wq3d_seq_search */
    int32_t pad_;
} wq3d_seq;

/* ------------------------------------------------------------------ grids */
typedef enum wq3d_interp {
    WQ3D_INTERP_NEAREST = 0,
    WQ3D_INTERP_LINEAR = 1,
    WQ3D_INTERP_PCHIP = 2
} wq3d_interp;

/* gen_grid<N> (types/gen_grid.h:32-560), N = ndims <= 3.  data is row-major
 * with axis 0 slowest (types/data_grid.h:47-61). */
typedef struct wq3d_grid {
    int32_t ndims;
    int32_t interp[3];     /* wq3d_interp per axis     */
    int32_t edge_limit[3]; /* 0/1 per axis             */
    int32_t pad_;
    wq3d_seq axis[3];
    const double* data;
} wq3d_grid;

/* --------------------------------------------------------- environment */
typedef enum wq3d_profile_kind {
    WQ3D_PROFILE_LINEAR = 0, /* ocean/profile_linear.cc:14-40  p = {c0, g0, z1, g1}          */
    WQ3D_PROFILE_MUNK = 1,   /* ocean/profile_munk.cc:19-32    p = {axis_depth, scale, c1, eps} */
    WQ3D_PROFILE_GRID = 2    /* ocean/profile_grid.h:68-130    grid 1-D (rho) or 3-D (rho,theta,phi) */
} wq3d_profile_kind;

typedef enum wq3d_atten_kind {
    WQ3D_ATTEN_THORP = 0,    /* ocean/attenuation_thorp.cc:14-36    */
    WQ3D_ATTEN_CONSTANT = 1  /* ocean/attenuation_constant.cc:50-62 */
} wq3d_atten_kind;

typedef struct wq3d_profile_desc {
    int32_t kind;       /* wq3d_profile_kind */
    int32_t flat_earth; /* profile_model::flat_earth (ocean/profile_model.cc:14-26) */
    double p[4];
    wq3d_grid grid;
    int32_t atten_kind; /* wq3d_atten_kind */
    int32_t pad_;
    double atten_coeff; /* attenuation_constant coefficient (dB/m/Hz) */
} wq3d_profile_desc;

typedef enum wq3d_boundary_kind {
    WQ3D_BOUNDARY_FLAT = 0, /* ocean/boundary_flat.cc:19-63                     */
    WQ3D_BOUNDARY_GRID = 1  /* ocean/boundary_grid.h:88-196, 1-D or 2-D rho grid */
} wq3d_boundary_kind;

typedef enum wq3d_loss_kind {
    WQ3D_LOSS_CONSTANT = 0, /* ocean/reflect_loss_constant.h   loss = {amplitude, phase} */
    WQ3D_LOSS_RAYLEIGH = 1  /* ocean/reflect_loss_rayleigh.cc:68-144
                               loss = {density, speed, att_bottom, speed_shear, att_shear}
                               as ratios/dB-per-wavelength given to the constructor */
} wq3d_loss_kind;

typedef struct wq3d_boundary_desc {
    int32_t kind;      /* wq3d_boundary_kind */
    int32_t loss_kind; /* wq3d_loss_kind */
    double rho;        /* FLAT: earth_radius - |depth| at construction */
    double normal_rho; /* FLAT: -1 for the ocean surface, +1 otherwise   */
    double loss[5];
    wq3d_grid grid;    /* GRID: rho = f(theta) or f(theta, phi) */
} wq3d_boundary_desc;

/* volume_flat (ocean/volume_flat.h): layer at rho = earth_radius - |depth|. */
typedef struct wq3d_volume_desc {
    double rho;
    double thickness;
} wq3d_volume_desc;

typedef struct wq3d_env_desc {
    wq3d_profile_desc profile;
    wq3d_boundary_desc surface;
    wq3d_boundary_desc bottom;
    int32_t num_volume;
    int32_t pad_;
    const wq3d_volume_desc* volume;
} wq3d_env_desc;

/* ------------------------------------------------------------------ fans */
typedef enum wq3d_spreading {
    WQ3D_SPREADING_CLASSIC_RAY = 0,
    WQ3D_SPREADING_HYBRID_GAUSSIAN = 1
} wq3d_spreading;

/* wave_thresholds (waveq3d/wave_thresholds.h:39-200) in stored form:
 * intensity in positive dB, eigenverb as linear power. */
typedef struct wq3d_thresholds {
    double intensity;  /* default 300      */
    double eigenverb;  /* default 1e-30    */
    int32_t max_bottom, max_surface, max_caustic, max_upper, max_lower; /* default 999 */
    int32_t pad_;
} wq3d_thresholds;

typedef struct wq3d_fan_desc {
    wq3d_seq freq;
    wq3d_seq de;          /* degrees */
    wq3d_seq az;          /* degrees */
    double src_rho, src_theta, src_phi; /* wposition1 before the water-column clamp */
    double time_step;
    double earth_radius;  /* wposition::earth_radius in effect while propagating */
    int64_t num_t1, num_t2;             /* target grid (0 x 0 = no targets) */
    const double* target_rho;           /* num_t1*num_t2, row-major */
    const double* target_theta;
    const double* target_phi;
    int32_t spreading;    /* wq3d_spreading */
    int32_t want_eigenverbs;   /* an eigenverb listener is attached */
    int32_t want_reflections;  /* a reflection listener is attached */
    int32_t pad_;
    wq3d_thresholds thresholds;
} wq3d_fan_desc;

/* ---------------------------------------------------------------- records */
/* One eigenray_model (eigenrays/eigenray_model.h:24-102) plus the keys the
 * reference leaves implicit.  intensity/phase (num_freq each) are returned in
 * separate arrays.  Drained records are ordered (step, t1, t2, de, az), the
 * reference's listener call order within detect_eigenrays(). */
typedef struct wq3d_eigenray {
    int64_t step;
    int32_t t1, t2, de, az;
    double travel_time, source_de, source_az, target_de, target_az;
    int32_t surface, bottom, caustic, upper, lower;
    int32_t pad_;
} wq3d_eigenray;

/* One eigenverb_model (eigenverbs/eigenverb_model.h:34-160); power in a
 * separate array (num_freq each).  Ordered (step, de, az, seq) where seq
 * counts the emissions of one ray within one step. */
typedef struct wq3d_eigenverb {
    int64_t step;
    int32_t de, az, seq, type;
    double travel_time, length, width, direction, grazing, sound_speed;
    double pos_rho, pos_theta, pos_phi;
    double source_de, source_az;
    int32_t surface, bottom, caustic, upper, lower;
    int32_t pad_;
} wq3d_eigenverb;

/* Arguments of reflection_listener::reflect (waveq3d/reflection_listener.h). */
typedef struct wq3d_reflection {
    int64_t step;
    int32_t de, az, seq, type;
    double time, dt, grazing, speed;
    double pos_rho, pos_theta, pos_phi;
    double ndir_rho, ndir_theta, ndir_phi;
} wq3d_reflection;

/* Host copy of one wave_front (waveq3d/wave_front.h:152-258).  Every array
 * pointer may be NULL (not copied); otherwise it holds num_de*num_az entries
 * (de-major), attenuation/phase num_de*num_az*num_freq ([ray][freq]). */
typedef struct wq3d_front_view {
    double *pos_rho, *pos_theta, *pos_phi;
    double *pgrad_rho, *pgrad_theta, *pgrad_phi;
    double *ndir_rho, *ndir_theta, *ndir_phi;
    double *dgrad_rho, *dgrad_theta, *dgrad_phi;
    double *sound_speed, *sgrad_rho, *sgrad_theta, *sgrad_phi;
    double *distance, *path_length;
    double *attenuation, *phase;
    int32_t *surface, *bottom, *caustic, *upper, *lower;
    uint8_t* on_edge;
} wq3d_front_view;

typedef enum wq3d_slot { WQ3D_PAST = 0, WQ3D_PREV = 1, WQ3D_CURR = 2, WQ3D_NEXT = 3 } wq3d_slot;

typedef struct wq3d_env wq3d_env;
typedef struct wq3d_fan wq3d_fan;

/* ------------------------------------------------------------ entry points */
/* Message of the last failing call on this thread ("" if none). */
const char* wq3d_last_error(void);

/* Number of visible HIP devices (0 when none); never fails. */
int32_t wq3d_device_count(void);

/* Upload an environment to the current HIP device.  Grids are copied; the
 * descriptor may be freed on return. */
wq3d_status wq3d_env_create(const wq3d_env_desc* desc, wq3d_env** env);
wq3d_status wq3d_env_destroy(wq3d_env* env);

/* Build a ray fan on one environment: source clamp, init_wave, RK3 start-up
 * and first AB3 step (wave_queue.cc:30-150).  The env must outlive the fan. */
wq3d_status wq3d_fan_create(wq3d_env* env, const wq3d_fan_desc* desc, wq3d_fan** fan);
wq3d_status wq3d_fan_destroy(wq3d_fan* fan);

/* Advance n steps of wave_queue::step().  Asynchronous on the fan's stream. */
wq3d_status wq3d_step(wq3d_fan* fan, int64_t n);

/* ---------------------------------------------------------------- batches */
/* Many fans stepped together.  Replaces the reference's pool of
 * wavefront_generator tasks, each stepping its own wave_queue
 * (wavegen/wavefront_generator.cc:52-127 on threads/thread_pool.cc:21-75):
 * wq3d_batch_step(batch, n) is n calls of step() on every fan of the batch,
 * with one fused launch per kernel of a step for all fans (each fan owns a
 * range of the launch's blocks), 64-step loops replayed as HIP graphs, and
 * the records of each fan kept and drained per fan as wq3d_step keeps them
 * (wq3d_drain_* on the fan).  The fans may differ in size, frequencies,
 * targets and environment; they must be on one device and at the same step
 * (fresh fans are), and are stepped only through the batch while it exists.
 * wq3d_batch_step returns when the n steps are done. */
typedef struct wq3d_batch wq3d_batch;

/* Launch shape of a batch's fused launches (results never depend on it). */
typedef struct wq3d_batch_shape {
    int32_t ray_threads;   /* threads per block of the per-ray launches: 64, 128 or 256 */
    int32_t build_blocks;  /* blocks of the eigenray-construction launch */
    int32_t use_graph;     /* 1: 64-step loops replayed as a captured HIP graph */
    int32_t pad_;
} wq3d_batch_shape;

wq3d_status wq3d_batch_create(wq3d_fan* const* fans, int32_t num_fans, wq3d_batch** batch);
wq3d_status wq3d_batch_destroy(wq3d_batch* batch);
wq3d_status wq3d_batch_step(wq3d_batch* batch, int64_t n);
wq3d_status wq3d_batch_get_shape(const wq3d_batch* batch, wq3d_batch_shape* shape);
wq3d_status wq3d_batch_set_shape(wq3d_batch* batch, const wq3d_batch_shape* shape);
/* HIP events around each step's search stage (the fused launches of the
 * tiles' super-block masks and of the eigenray search filter) on the
 * batch's stream; while on, steps are launched one by one (no graph).
 * wq3d_batch_stage_times returns the mean stage time per step since the
 * last call and the number of steps it covers, then resets them. */
wq3d_status wq3d_batch_set_timing(wq3d_batch* batch, int32_t enable);
wq3d_status wq3d_batch_stage_times(wq3d_batch* batch, double* ms_per_step, int64_t* steps);
/* The autotuner: times `trial` real steps of the batch for each candidate
 * shape (threads per block, then construction blocks), keeps the fastest and
 * returns its time per step.  Advances the batch by the steps it times plus
 * one per candidate; steps_taken returns that count. */
wq3d_status wq3d_batch_tune(wq3d_batch* batch, int32_t trial, double* best_ms_per_step, int64_t* steps_taken);

/* wave_thresholds setters (waveq3d/wave_thresholds.h:54-192) and listener
 * presence (eigenverb_notifier::has_eigenverb_listeners,
 * reflection_notifier::has_reflection_listeners); both take effect at the
 * next wq3d_step. */
wq3d_status wq3d_set_thresholds(wq3d_fan* fan, const wq3d_thresholds* thresholds);
wq3d_status wq3d_set_listeners(wq3d_fan* fan, int32_t want_eigenverbs, int32_t want_reflections);

/* Wait for queued work. */
wq3d_status wq3d_sync(wq3d_fan* fan);

/* Current wave_queue::time() (host mirror, exact accumulation). */
wq3d_status wq3d_time(const wq3d_fan* fan, double* time);

/* Source position after the water-column clamp (wave_queue.cc:58-68). */
wq3d_status wq3d_source(const wq3d_fan* fan, double* rho, double* theta, double* phi);

/* Count / copy out the records produced since the last drain.  A call with
 * out == NULL only reports the pending count in *n. */
wq3d_status wq3d_drain_eigenrays(wq3d_fan* fan, wq3d_eigenray* out, double* intensity,
                                 double* phase, int64_t capacity, int64_t* n);
wq3d_status wq3d_drain_eigenverbs(wq3d_fan* fan, wq3d_eigenverb* out, double* power,
                                  int64_t capacity, int64_t* n);
wq3d_status wq3d_drain_reflections(wq3d_fan* fan, wq3d_reflection* out, int64_t capacity,
                                   int64_t* n);

/* ------------------------------------------------------ eigenray totals */
/* eigenray_collection's per-target total (eigenrays/eigenray_collection.cc:
 * 105-191): the eigenray_model of the sum, whose intensity and phase
 * (num_freq each) are returned in separate arrays, plus what add_eigenray
 * keeps per target (:90-101).  Targets with no eigenray get the reference's
 * values for an empty list: 300 dB, phase 0, NaN weighted averages, source
 * and target AZ 90 degrees, counters -1, initial time 0. */
typedef struct wq3d_eigenray_total {
    double travel_time, source_de, source_az, target_de, target_az;
    double initial_time;  /* eigenray_collection::initial_time(t1, t2) */
    int32_t surface, bottom, caustic, upper, lower;
    int32_t count;        /* eigenray_collection::eigenrays(t1, t2).size() */
} wq3d_eigenray_total;

/* eigenray_collection::sum_eigenrays on the GPU (the current device) for
 * eigenrays given in the order the collection received them, which is the
 * order wq3d_drain_eigenrays returns them in.  rays[k].t1 / .t2 index a
 * num_t1 x num_t2 target grid; intensity / phase hold num_freq values per
 * ray.  coherent: phasor sum with phase 2 pi f t + phi (the reference's
 * default), else a sum of pressure amplitudes.  Outputs are num_t1 x num_t2,
 * row-major: totals, and total_intensity / total_phase with num_freq values
 * per target.  WQ3D_ERR_ARGUMENT for a ray outside the grid. */
wq3d_status wq3d_sum_eigenrays(const wq3d_eigenray* rays, const double* intensity, const double* phase,
                               int64_t num_rays, const double* frequencies, int32_t num_freq,
                               int64_t num_t1, int64_t num_t2, int32_t coherent,
                               wq3d_eigenray_total* totals, double* total_intensity, double* total_phase);

/* wq3d_sum_eigenrays without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device or host memory
 * (tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * results and errors as wq3d_sum_eigenrays; the device memory of the call
 * (the uploaded records, the per-target lists and the totals) lives in a
 * 128 MiB array the library holds on each device for the life of the
 * process, and a call whose records need more than that returns
 * WQ3D_ERR_CAPACITY before any device work.  Calls on the same device share
 * that array: calls that overlap on one device race. */
wq3d_status wq3d_sum_eigenrays_noalloc(const wq3d_eigenray* rays, const double* intensity, const double* phase,
                                       int64_t num_rays, const double* frequencies, int32_t num_freq,
                                       int64_t num_t1, int64_t num_t2, int32_t coherent,
                                       wq3d_eigenray_total* totals, double* total_intensity,
                                       double* total_phase);

/* Copy one queue slot to the host (synchronises the fan's stream). */
wq3d_status wq3d_snapshot(wq3d_fan* fan, int32_t slot, wq3d_front_view* view);

/* Average duration (ms) per launch of each kernel of the last step() calls,
 * measured with HIP events on the fan's stream when timing is enabled.
 * Kernel ids: 0 init, 1 (not launched), 2 advance, 3 the eigenray search
 * filter (its two launches, the tiles' super-block masks and the filter
 * itself, timed together), 4 build, 5 (not launched), 6 bounce (caustics,
 * vertices and reflections).
 * enable: 0 off, 1 every kernel, 2 the eigenray search (detect) only, which
 * keeps the events' own cost out of the other launches. */
wq3d_status wq3d_set_timing(wq3d_fan* fan, int32_t enable);
wq3d_status wq3d_kernel_times(wq3d_fan* fan, double* ms_per_launch, int64_t* launches,
                              int32_t num_kernels);

/* ------------------------------------------------ adjoint (differentiation)
 * The discrete adjoint of eigenray travel time and transmission loss (dB)
 * with respect to the node speeds of the fan's sound speed grid.  No
 * counterpart in the reference, which has no derivatives: it is for
 * callers that fit or invert a sound speed profile (gradient descent over
 * profiles, sensitivity kernels).  Supported: a 1-D profile grid with linear
 * interpolation between flat boundaries, hybrid Gaussian spreading
 * (WQ3D_ERR_UNSUPPORTED otherwise).
 *
 * wq3d_adjoint_checkpoint keeps a device copy of the fan's whole state; the
 * first must be taken before the first step, later ones every K steps of
 * the forward run (memory grows by one copy of the state per call).
 * wq3d_adjoint_gradient takes eigenrays of this fan as drained
 * (wq3d_drain_eigenrays: records, intensities [num_rays][num_freq]), the
 * gradient of a loss with respect to each one's travel time and
 * intensities, and writes d loss / d node speed into grad_nodes (num_nodes,
 * the grid's size).  It replays each segment between checkpoints on a
 * copy of the state (the fan itself is untouched and may step on) and
 * checks that the replay rebuilds every given eigenray bit for bit
 * (WQ3D_ERR_STATE if not: records of another fan or run).  The backward
 * workspace holds a window of the segment's steps (416 bytes per ray and
 * step) plus 1 GiB of tapes and entries; wq3d_adjoint_release frees it and
 * the checkpoints. */
wq3d_status wq3d_adjoint_checkpoint(wq3d_fan* fan);
wq3d_status wq3d_adjoint_gradient(wq3d_fan* fan, const wq3d_eigenray* rays, const double* intensity,
                                  const double* grad_time, const double* grad_intensity, int64_t num_rays,
                                  double* grad_nodes, int64_t num_nodes);
wq3d_status wq3d_adjoint_release(wq3d_fan* fan);

/* Diagnostics with no counterpart in the reference: the eigenray search's
 * pair counts of the last step (pairs the first-stage filter kept - the
 * pre-test window and the neighbour forms - and pairs that passed the exact
 * pre-test and neighbour comparison, the eigenray candidates).
 * Synchronises the fan's stream. */
wq3d_status wq3d_search_stats(wq3d_fan* fan, int64_t* window_hits, int64_t* candidates);

/* Diagnostic with no counterpart in the reference: the work of the eigenray
 * search's filter since the fan was created or since the last call, which
 * resets the counts: the (target, ray) pairs put to the window test (pairs
 * in target blocks the whole-block bounds did not rule out) and the
 * (ray, box of targets) whole-block bounds tested.  Synchronises the fan's
 * stream. */
wq3d_status wq3d_search_totals(wq3d_fan* fan, int64_t* pairs_tested, int64_t* box_tests);

#ifdef __cplusplus
}
#endif
#endif /* WQ3D_H */
