/*
 * wema_dev.h - C ABI of the MI355X (gfx950) WEMA per-frame amplification pipeline.
 *
 * This is the drop-in boundary for the path that curioustorvald/wema drives through
 * wema_init (and the _color/_iir variants), wema_process_frame and wema_batch_process (src/wema.h:253-283). One handle is one
 * stream session: its DWT pyramids, IIR state and sliding-window ring live in HBM for the
 * life of the handle. Frames cross the boundary as RGB24 bytes (or float RGB in [0,1] for
 * the reference's Frame API); every entry point takes plain pointers and sizes.
 *
 * Return values follow the reference: 0 on success, -1 on error (wema.c:437-439); the text
 * of the last error is available from wema_dev_last_error().
 *
 * Threading: a handle is used by one host thread at a time (as WemaContext, wema.h:102).
 * Work is enqueued on the HIP stream given by the caller (NULL = the null stream) and is
 * stream-ordered; consecutive calls on one handle must be ordered (same stream, or
 * synchronised by the caller). wema_dev_sync waits for the last call's stream.
 */
#ifndef WEMA_DEV_H
#define WEMA_DEV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wema_dev wema_dev_t;

/* Options of one session. Mirrors WemaConfig (src/wema.h:192-212) and the fields main.c
 * writes after wema_init (src/main.c:288-290). */
/* Largest frame, in pixels (16384 x 16384). Frame offsets stay below 2^31 bytes per plane
 * and the exact Q32 threshold sum cannot overflow its 64 bits. The reference indexes a
 * frame's floats with an int (src/wema.c:44-50), which limits it to 715,827,882 pixels. */
#define WEMA_DEV_MAX_PIXELS 268435456L

typedef struct {
    int   width, height;  /* 1 <= width * height <= WEMA_DEV_MAX_PIXELS */
    float fps;
    float amp;            /* amplification factor (-a, main.c:96-105) */
    float f_low, f_high;  /* band edges in Hz (--fl/--fh); IIR mode */
    int   window;         /* sliding window size (--temporal-window); window mode */
    int   levels;         /* DWT levels; 0 = the reference's automatic rule (dwt.c:588-596) */
    int   use_iir;        /* 1 = IIR band-pass (default), 0 = sliding window ("--haar") */
    int   color;          /* --color: filter Y, Cb, Cr */
    int   edge_aware;     /* guided filter (default) or 5x5 binomial (--no-edge-aware) */
    int   bilateral;      /* --bilateral-filter (window mode only) */
    int   batch_path;     /* 1: wema_batch_process semantics (guided filter always on,
                             wema_batch.c:452) */
    int   max_batch;      /* largest number of frames per process call (1..256) */
    int   device;         /* HIP device ordinal */
} wema_dev_params;

/* Fills p with the reference CLI defaults (main.c:63-75) for a width x height, fps clip. */
void wema_dev_params_default(wema_dev_params *p, int width, int height, float fps);

/* Host-only: the band-pass of iir_design_bandpass (iir_filter.c:70-101, order 2: one
 * high-pass then one low-pass Butterworth biquad), c10 = HP b0 b1 b2 a1 a2, LP b0 b1 b2 a1
 * a2, and the warm-up count of iir_temporal_init (iir_filter.c:150-151). Same rejections
 * (f_low <= 0, f_low >= f_high, f_high >= fs / 2) and the same float arithmetic. */
int wema_dev_design_bandpass(float f_low, float f_high, float fs, float *c10, int *warmup);

/* The reference's automatic level count (dwt_compute_levels, dwt.c:588-596: halvings of
 * the smaller side while it is >= 16, at most 4, at least 1). */
int wema_dev_auto_levels(int width, int height);

/* Device bytes a session with these options holds (HBM admission; no allocation). */
int wema_dev_query_bytes(const wema_dev_params *p, size_t *bytes);

/* Replaces wema_init + wema_init_color + wema_init_iir + wema_init_iir_color
 * (wema.c:359-589): validates the options, designs the band-pass exactly as
 * iir_design_bandpass does (iir_filter.c:70-101), allocates all device state. */
int  wema_dev_open(wema_dev_t **out, const wema_dev_params *p);
void wema_dev_close(wema_dev_t *s);                      /* wema_free (wema.c:591) */
const char *wema_dev_last_error(void);

/* Number of DWT levels, unique coefficient positions (3 per subband position) and the
 * reference's position count N6 (dwt_num_positions, dwt.c:769-778). */
int wema_dev_geometry(const wema_dev_t *s, int *levels, size_t *n_unique, size_t *n6);
/* IIR warm-up iterations (iir_filter.c:150-151) and the 10 section coefficients
 * (HP b0 b1 b2 a1 a2, LP b0 b1 b2 a1 a2). */
int wema_dev_iir_coeffs(const wema_dev_t *s, float *c10, int *warmup);
/* Frames processed so far by this session. */
long wema_dev_frames(const wema_dev_t *s);

/* Process n consecutive frames of the session (n <= max_batch), device buffers, RGB24
 * HWC. Output frame i corresponds to input frame i, as wema_process_frame (wema.c:746)
 * called n times, or wema_batch_process (wema_batch.c:352) when batch_path is set. */
int wema_dev_process_u8(wema_dev_t *s, const uint8_t *d_in, uint8_t *d_out, int n,
                        void *hip_stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * wema_dev_process_u8 with the kernels' per-call buffers (coefficient planes, LL pyramid,
 * reconstructions, guided-filter a and b, threshold accumulators) in one static device
 * array of WEMA_DEV_SCRATCH_BYTES per GPU instead of the session's own memory; the session
 * keeps its temporal state. It calls no hipMalloc or hipFree. Calls that need more scratch
 * than that (about one 1080p colour frame, or three 1080p luma frames) return -1 with an
 * error that says so. Every call on one GPU shares that GPU's array: calls that overlap in
 * time on one GPU (on different streams, or from different threads) race and must be
 * ordered by the caller. The array's address is looked up when the first session is opened
 * on a device. */
#define WEMA_DEV_SCRATCH_BYTES (64L << 20)
int wema_dev_process_u8_noalloc(wema_dev_t *s, const uint8_t *d_in, uint8_t *d_out, int n,
                                void *hip_stream);
/* Same on float RGB in [0,1] (the reference's Frame data, wema.h:28-33). d_out_u8 and
 * d_out_f32 may each be NULL. */
int wema_dev_process_f32(wema_dev_t *s, const float *d_in, uint8_t *d_out_u8,
                         float *d_out_f32, int n, void *hip_stream);
/* Host-buffer variant (one frame): copies in, processes, copies out, waits. */
int wema_dev_process_host_u8(wema_dev_t *s, const uint8_t *h_in, uint8_t *h_out);
/* Host-buffer variant on the reference's Frame data (float RGB in [0,1], HWC), n frames
 * (n <= max_batch) given as arrays of frame pointers: what wema_process_frame
 * (wema.c:746) and wema_batch_process (wema_batch.c:352) receive and return. Copies in,
 * processes, copies out (clamped float RGB), waits. Pinned frames (wema_dev_host_alloc)
 * copy at full link speed. */
int wema_dev_process_host_f32(wema_dev_t *s, const float *const *h_in, float *const *h_out, int n);
int wema_dev_sync(wema_dev_t *s);

/* Asynchronous host-buffer pipeline, the GPU counterpart of the reader/writer slots of
 * async_io.c (async_io.c:183-272): n RGB24 frames (n <= max_batch) are uploaded on a copy
 * stream, processed on the session's stream and downloaded on a second copy stream; the
 * call returns at once with a ticket. The session has two device slots, so the upload of
 * one batch overlaps the processing and download of the other; further submits queue on
 * the device behind the slot they reuse. h_in and h_out must stay valid
 * until wema_dev_wait_ticket returns for the ticket; with pinned memory
 * (wema_dev_host_alloc) the copies run at full link speed. */
int wema_dev_submit_host_u8(wema_dev_t *s, const uint8_t *h_in, uint8_t *h_out, int n, long *ticket);
int wema_dev_wait_ticket(wema_dev_t *s, long ticket);

/* The same pipeline one frame at a time, with presentation timestamps (the shape of
 * async_io.c's slot queue, src/async_io.c:183-272, where frames travel with their pts):
 * wema_dev_submit queues one RGB24 frame and the buffer its output goes to; a full batch
 * (max_batch frames) is dispatched at once. wema_dev_collect waits for the oldest queued
 * frame (dispatching a partial batch if that frame is still in one) and returns its pts and
 * output buffer: frames come back in submission order. Both buffers must stay valid until
 * the frame is collected; pinned memory (wema_dev_host_alloc) copies at full link speed.
 * Collecting with nothing queued is an error, and so is wema_dev_submit_host_u8 while
 * frames submitted here wait to be dispatched. wema_dev_close drops frames not collected
 * (those already dispatched finish first). */
int wema_dev_submit(wema_dev_t *s, const uint8_t *rgb24_in, uint8_t *rgb24_out, int64_t pts);
int wema_dev_collect(wema_dev_t *s, int64_t *pts, uint8_t **rgb24_out);
/* The same on the reference's Frame data (float RGB in [0,1], HWC; clamped float RGB out),
 * through float device slots allocated on first use. RGB24 and float frames may be queued
 * in any order (a change of kind starts a new batch); each is collected with the call of
 * its kind, in submission order. */
int wema_dev_submit_f32(wema_dev_t *s, const float *rgb_in, float *rgb_out, int64_t pts);
int wema_dev_collect_f32(wema_dev_t *s, int64_t *pts, float **rgb_out);

/* Temporal state of a session (no counterpart in the reference, whose IIR state lives in
 * IIRTemporalFilter, iir_filter.h:45-61, and window ring in TemporalBuffer, wema.h:67-81):
 * the IIR sections' state, or the last window - 1 frames of the window ring, with the
 * number of frames processed. Exported from one session and imported into another of the
 * same size, levels, colour mode and temporal filter (IIR band and frame rate, or window
 * size; max_batch, device and amplification may differ), the second continues with the
 * outputs the first would have given. Both calls wait for the session's work; import
 * rejects the state of another geometry or filter. The buffer is plain host memory:
 * a versioned header, then the state, in the host's byte order. */
int wema_dev_state_bytes(const wema_dev_t *s, size_t *bytes);
int wema_dev_export_state(wema_dev_t *s, void *host, size_t bytes);
int wema_dev_import_state(wema_dev_t *s, const void *host, size_t bytes);

/* Tile shapes (no counterpart in the reference). The guided filter runs as one kernel that
 * streams strips of 240 output columns down segments of rows, each segment recomputing the
 * 16 rows above it. A plain session whose output stage is the guided filter takes the
 * segment height an autotuner picked for its shape: the first session of a frame size and
 * frames per call on a device times the strips split into 1, 2, 4, .. 64 equal segments on
 * its own buffers when it is opened and keeps the fastest for the process. Pooled sessions,
 * and sessions opened with that tuner off, leave the split to each launch (one segment
 * height for all its frames, from a cost model of its rows and workgroups). The DWT levels
 * run as strips 4, 8, 16 or 32 pair rows tall, and the levels below full resolution also as
 * LDS tiles of 16 x 16 output pairs or with one output pair per thread: the first session
 * of a shape (level size, channels, frames per call) on a device times the candidates on
 * its own buffers when it is opened and keeps the fastest for each level
 * (wema_dev_strip_rows: a level's forward or inverse height, 1 = LDS tiles, 2 = one pair
 * per thread, 0 = the fixed rule). Results never depend on either shape.
 * wema_dev_autotune(mask) selects the tuners for sessions opened later: WEMA_TUNE_GF,
 * WEMA_TUNE_DWT (both by default; 0: the launch's split and the strips' fixed rule).
 * wema_dev_tile_rows: the session's guided-filter segment height (0: the launch's split);
 * wema_dev_set_tile_rows forces one (a positive multiple of 4, or 0) for its next calls;
 * wema_dev_set_strip_rows one strip height for every level (0 = the rule; 1 = LDS tiles
 * or 2 = one pair per thread below full resolution, level 0 then keeps the rule). */
#define WEMA_TUNE_GF 1
#define WEMA_TUNE_DWT 2
void wema_dev_autotune(int mask);
int  wema_dev_strip_rows(const wema_dev_t *s, int level, int inverse);
int  wema_dev_set_strip_rows(wema_dev_t *s, int rows);
int  wema_dev_tile_rows(const wema_dev_t *s);
int  wema_dev_set_tile_rows(wema_dev_t *s, int rows);

/* Pinned host memory for frame slots (hipHostMalloc). */
void *wema_dev_host_alloc(size_t bytes);
void  wema_dev_host_free(void *p);

/* ---- multi-stream pool (no counterpart in the reference: it processes one stream) --------
 * Sessions of many streams share one GPU. wema_pool_add admits a session only if the
 * pool's device bytes stay within hbm_budget (-1 and "HBM budget exceeded" otherwise).
 * wema_pool_tick processes the listed frames with one launch per pipeline stage covering
 * all of them: each entry of ids is one frame of that session, and a session listed k
 * times in a row (k <= its max_batch) has its next k frames processed in order, as one
 * call of k frames would; a session may appear in one such run per tick. Each session's
 * results are identical to running it alone (no computation crosses sessions). With graphs
 * enabled, the launch sequence of each tick shape (sessions, their frame counts and output
 * stage) is captured once in a hipGraph and replayed. NULL stream = the pool's own stream
 * (graphs run only there). */
typedef struct wema_pool wema_pool_t;
int    wema_pool_create(wema_pool_t **out, int device, size_t hbm_budget);
void   wema_pool_destroy(wema_pool_t *p);
int    wema_pool_add(wema_pool_t *p, const wema_dev_params *prm, int *id);
int    wema_pool_remove(wema_pool_t *p, int id);
int    wema_pool_session(wema_pool_t *p, int id, wema_dev_t **s);          /* for queries */
size_t wema_pool_bytes_in_use(const wema_pool_t *p);
int    wema_pool_graphs(wema_pool_t *p, int enable);
int    wema_pool_tick(wema_pool_t *p, int n, const int *ids, const uint8_t *const *d_in,
                      uint8_t *const *d_out, void *hip_stream);
int    wema_pool_sync(wema_pool_t *p);
void  *wema_pool_stream(wema_pool_t *p);
/* Ticks run so far, and how many of them were replayed from a captured hipGraph. */
int    wema_pool_stats(const wema_pool_t *p, long *ticks, long *graph_ticks);
int    wema_pool_timing(wema_pool_t *p, int enable);
int    wema_pool_timing_report(wema_pool_t *p, int kind, char *name, int name_len, long *launches,
                               double *total_ms, double *bytes);

/* ---- serving loop over a pool (no counterpart in the reference) -----------------------
 * Sessions arrive as a Poisson process (arrival_rate per second until duration_s; each
 * lives life_min_s..life_max_s), with a resolution from the mix 8:8:6:2 of 640x480,
 * 1280x720, 1920x1080, 3840x2160, a frame rate from that mix (fps_from_mix) or 30/60 at
 * random, a band in 0.3..5 Hz and amplification in 20..150. They are admitted first come
 * first served while the pool's bytes stay within hbm_budget (and max_sessions, if > 0),
 * and deliver frames at their frame rate. batched: all due frames run as one pool tick, up
 * to 2 frames of a session that has fallen behind (as one call of 2 frames; sessions are
 * opened with max_batch 2); otherwise one frame of one session at a time (the plain path,
 * max_batch 1). saturate: `sessions` sessions admitted
 * at once, taking the mix in its exact 8:8:6:2 proportions, with a frame always due, for
 * duration_s (capacity). warm_start: the Poisson run starts in steady state (the sessions
 * that arrived in the life_max_s before the start and are still alive are admitted before
 * the clock starts, each at a random frame phase). frames_per_session > 0: a session leaves
 * after that many measured frames (a saturate run then ends when all have left).
 * warmup_rounds: the first rounds (iterations that ran frames) are run but not measured.
 * host_io: every frame is uploaded from pinned host memory and its output downloaded into
 * pinned host memory, on copy streams that overlap the ticks (the path of async_io.c's
 * slots, src/async_io.c:183-272); otherwise frames stay in HBM. Latency = frame due time to
 * output complete (in host memory with host_io).
 * Input frames: session k of the plan (wema_serve_plan, in arrival order) of resolution class
 * c (0..3 = 640x480, 1280x720, 1920x1080, 3840x2160) gets as its frame t the synthetic frame
 * t % WEMA_SERVE_CLASS_FRAMES of preset 4 at that class's size and mix frame rate (30, 30, 60,
 * 30) with seed offset c (wema_synth_preset). on_frame, if set, is called from the serving
 * thread for every output frame once it is complete, with the session's plan index, the
 * frame index and the output (host memory with host_io, else device memory; valid during
 * the call). */
#define WEMA_SERVE_CLASS_FRAMES 4
typedef void (*wema_serve_frame_cb)(void *user, int session, long frame, const uint8_t *rgb24,
                                    int width, int height);
typedef struct {
    int      device;
    double   duration_s, drain_s;
    double   arrival_rate;
    double   life_min_s, life_max_s;
    size_t   hbm_budget;
    uint64_t seed;
    int      saturate, sessions;
    int      batched, use_graphs;
    int      max_sessions;
    int      fps_from_mix;
    int      warm_start;
    long     frames_per_session;
    int      warmup_rounds;
    int      host_io;
    wema_serve_frame_cb on_frame;
    void    *cb_user;
} wema_serve_config;
typedef struct {
    long   frames;
    double mpix, wall_s, mpix_per_s;
    double lat_p50_ms, lat_p99_ms, lat_max_ms;
    int    sessions_arrived, sessions_admitted, admission_waits, peak_sessions;
    size_t peak_bytes;
    long   ticks;
    double mean_frames_per_tick;
    double host_ms_per_tick;    /* host time to build and launch one tick (mean, measured ticks) */
    double graph_tick_frac;     /* fraction of measured ticks replayed from a hipGraph */
    double gpu_busy_frac;       /* GPU time of the measured ticks / measured wall time */
    double link_gbs;            /* host_io: bytes copied over the link, both ways, per second */
} wema_serve_stats;
int wema_serve_run(const wema_serve_config *cfg, wema_serve_stats *out);
/* The sessions a run with this configuration admits, in arrival order (deterministic for a
 * seed): their parameters, arrival times (s) and resolution classes, up to max of them;
 * returns the number of sessions planned, or -1. */
int wema_serve_plan(const wema_serve_config *cfg, int max, wema_dev_params *params, double *arrive_s,
                    int *cls);

/* ---- measurement (no counterpart in the reference) --------------------------------
 * While enabled, HIP events are recorded on the call's stream between the launches of
 * every process call. wema_dev_timing_report sums them per kernel kind (0 .. kinds-1; it
 * returns -1 past the last kind): launches, total milliseconds and algorithmic HBM bytes
 * (compulsory traffic of the launches: inputs read once, outputs written once). */
int wema_dev_timing(wema_dev_t *s, int enable);
int wema_dev_timing_report(wema_dev_t *s, int kind, char *name, int name_len, long *launches,
                           double *total_ms, double *bytes);

/* ---- test hooks (no counterpart in the reference) --------------------------------
 * Thresholds used by the last processed frames: thr[f * 3 + ch] (NaN for unused
 * channels / passthrough frames). Call after wema_dev_sync. */
int wema_dev_debug_thresholds(wema_dev_t *s, float *thr, int n);
/* Keep original coefficients and the temporal filter output of every frame for
 * wema_dev_debug_read (call before the first frame). */
int wema_dev_debug_capture(wema_dev_t *s, int enable);
/* Override the thresholds of the next process call: thr[f * 3 + ch]; NULL clears. */
int wema_dev_debug_set_thresholds(wema_dev_t *s, const float *thr, int n);
/* Copy a device buffer of frame f (of the last call) to host:
 * which 0: original coefficients (unique planes, n_unique floats),
 *       1: temporal filter output (IIR y / window delta), 2: amplified coefficients,
 *       3: IIR state (4 * n_unique floats, per position HP w1, HP w2, LP w1, LP w2). */
int wema_dev_debug_read(wema_dev_t *s, int which, int frame, int ch, float *host, size_t count);
/* Injected failures: the next `after`-th occurrence of a site fails as if its HIP call had
 * returned an error, and the call returns -1 with the error text. site 1: the upload of a
 * call's job tables (before anything of the call is enqueued: no frame is consumed and the
 * session stays usable); 2: a stage launch (part of the call already enqueued: the session
 * is marked failed, every later call on it returns -1 and it can only be closed, or removed
 * from its pool); 3: the upload of a batch of host frames (wema_dev_submit*, before anything
 * of the batch is enqueued). A real HIP error at the same points behaves the same way.
 * site 0 disarms. */
int wema_dev_debug_inject_error(int site, int after);
/* The guided filter's covariance division q[i] = a[i] / b[i] as the kernels compute it, and
 * the hardware reciprocal estimate rcp[i] of b[i] it starts from; n elements, device
 * buffers, enqueued on the stream. */
int wema_dev_debug_div_mid(const float *d_a, const float *d_b, float *d_q, float *d_rcp, long n,
                           void *hip_stream);

/* ---- synthetic video (deterministic, integer-only; identical on host and device) ---- */
typedef struct {
    int      width, height;
    float    fps;
    uint32_t seed;
    int      kind;               /* 0 texture + discs, 1 checkerboard + shear */
    int      color;              /* independent noise per channel + disc tint */
    int      ndiscs;             /* <= 4 */
    float    disc_cx[4], disc_cy[4];     /* centre, fraction of width / height */
    float    disc_amp[4], disc_freq[4];  /* px, Hz */
    float    disc_radius;        /* px */
    float    disc_contrast;      /* levels */
    float    tint_r, tint_b;     /* extra disc levels on R and B when color */
    float    noise_sigma;        /* levels */
    float    shear_amp, shear_freq;       /* px, Hz */
    int      cell;               /* checkerboard cell size (px) */
    float    checker_contrast;   /* levels */
} wema_synth_params;

/* Presets for the workloads: 1 = 320x240@30 disc, 2 = 1920x1080@30 three discs (colour
 * tint), 3 = 3840x2160@60 checkerboard shear, 4 = disc scene at any size (multi-stream).
 * seed_offset is added to the base seed 0x57454D41. */
int  wema_synth_preset(wema_synth_params *p, int preset, int width, int height, float fps,
                       uint32_t seed_offset);
/* Static texture of the scene (Q8 levels, width*height uint16). */
int  wema_synth_texture_host(const wema_synth_params *p, uint16_t *tex);
/* Frame `frame` on the host (RGB24). */
int  wema_synth_frame_host(const wema_synth_params *p, const uint16_t *tex, int frame,
                           uint8_t *rgb);
/* Frames frame0..frame0+n-1 on the device (RGB24, consecutive), from a device texture. */
int  wema_synth_frames_device(const wema_synth_params *p, const uint16_t *d_tex, int frame0,
                              int n, uint8_t *d_rgb, void *hip_stream);

#ifdef __cplusplus
}
#endif
#endif
