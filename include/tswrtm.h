/*
 * tswrtm.h - C ABI of the MI355X (gfx950) reverse-time-migration engine.
 *
 * Drop-in for the native boundary of tsw_rtm (ll23002/DAS).  Loaded by
 * ctypes from das_amd/_native.py, the way tsw_rtm/wave_solver.py:14-71 loads
 * libwave2d.so.  Plain C types only; arrays are row-major (nz, nx) float32,
 * flat index iz*nx + ix, exactly as in the reference.
 *
 * Two groups of entry points:
 *   1. The four reference kernels with identical signatures and host-pointer,
 *      synchronous semantics (tsw_rtm/c_engine/wave2d.c).  Each copies its
 *      arrays to HBM, runs the HIP kernel and copies the result back.  They
 *      return void like the reference; tsw_last_error() is non-empty after a
 *      failed call.
 *   2. Coarse entry points that keep the time loop, snapshots and images in
 *      HBM.  They return 0 on success and a negative status on failure
 *      (-1 bad argument, -2 HIP error, -3 out of device memory / budget,
 *      -4 wrong state, -5 injected failure, -6 diverged: a finished image or
 *      illumination holds NaN/Inf); tsw_last_error() describes it.
 *      Host pointers are borrowed for the duration of the call.  The
 *      *_destroy calls accept a null handle and then do nothing (return 0).
 */
#ifndef TSWRTM_H
#define TSWRTM_H

#ifdef __cplusplus
extern "C" {
#endif

#define TSW_ABI_VERSION 1

typedef struct tsw_grid_s* tsw_grid_t;
typedef struct tsw_batch_s* tsw_batch_t;
typedef struct tsw_server_s* tsw_server_t;
typedef struct tsw_fwi_s* tsw_fwi_t;

/* One shot of a differentiable acoustic modelling problem.  src (nt) and
 * d_obs (nt, nrec) are host arrays in the problem's precision. */
typedef struct tsw_fwi_shot {
  int nt;
  const void* src;
  int six, siz;
  const int* rix;
  int nrec, riz;
  const void* d_obs;
} tsw_fwi_shot;

/* One virtual shot of a batch, on its own (nz, nx) grid.  The shot's grid is
 * the window [z0, z0+nz) x [x0, x0+nx) of a larger stacking grid. */
typedef struct tsw_shot_desc {
  int nx, nz, nt;          /* grid and time steps (source and data both nt long) */
  const float* vel_fwd;    /* (nz, nx) forward (sediment-flood) model */
  const float* vel_bwd;    /* (nz, nx) back-propagation (rock-flood) model */
  const float* src;        /* (nt) source wavelet */
  int six, siz;            /* source cell */
  const float* data;       /* (nt, nrec) receiver gather to back-inject */
  const int* rix;          /* (nrec) receiver columns on row riz */
  int nrec, riz;
  int x0, z0;              /* window offset in the stacking grid */
  int kind;                /* TSW_SHOT_ACOUSTIC or TSW_SHOT_ELASTIC */
  const float* ela_fwd[4]; /* elastic shots: forward lambda, mu, buoyancy_x, buoyancy_z */
  const float* ela_bwd[4]; /* elastic shots: back-propagation model */
} tsw_shot_desc;

#define TSW_SHOT_ACOUSTIC 0
#define TSW_SHOT_ELASTIC 1

#define TSW_RUN_BATCHED 0  /* all active shots in one launch per step (graphs) */
#define TSW_RUN_PLAIN 1    /* one shot at a time, one launch per step */

/* ------------------------------------------------------------------ 1. */

/* Replaces acoustic_2d_step, tsw_rtm/c_engine/wave2d.c:19-31.  With
 * nz < 3 and a free surface the reference's row 1 = -row 2 (wave2d.c:79)
 * reads past the grid; here p_next is then left unchanged. */
void acoustic_2d_step(const float* p_curr, const float* p_prev, float* p_next, const float* vel,
                      const float* sponge, int nx, int nz, float dx, float dz, float dt,
                      int free_surface_top);

/* Replaces elastic_2d_step, tsw_rtm/c_engine/wave2d.c:95-112.  The row above
 * the grid that the reference reads at iz = 1 with a free surface
 * (wave2d.c:135, :170; out of bounds there) reads as 0 here. */
void elastic_2d_step(float* vx, float* vz, float* sxx, float* szz, float* sxz, const float* lambda,
                     const float* mu, const float* buo_x, const float* buo_z, const float* sponge,
                     int nx, int nz, float dx, float dz, float dt, int free_surface_top);

/* Replaces correlate_imaging_condition, tsw_rtm/c_engine/wave2d.c:218-225. */
void correlate_imaging_condition(const float* src_wavefield, const float* rec_wavefield,
                                 float* image, int nx, int nz, float dt);

/* Replaces build_cerjan_sponge, tsw_rtm/c_engine/wave2d.c:237-244. */
void build_cerjan_sponge(float* sponge, int nx, int nz, int n_sponge, float alpha,
                         int free_surface_top);

/* ------------------------------------------------------------------ 2. */

/* Last error message of the calling thread ("" if none). No reference
 * counterpart: the reference has no error channel (wave2d.c returns void). */
const char* tsw_last_error(void);
/* ABI version (TSW_ABI_VERSION). */
int tsw_version(void);
/* Binds the engine to HIP device `device` (-1: keep the current one). */
int tsw_init(int device);
/* Free and total device memory of the engine's device, in bytes. */
int tsw_mem_info(long long* free_bytes, long long* total_bytes);
/* "<gcnArchName>:<compute units>" of the engine's device, for tuning caches. */
int tsw_device_name(char* out, int len);
/* Test hook: the next `count` coarse calls fail with -5 before doing work. */
void tsw_debug_fail_next(int count);

/* WaveSolver2D.__init__ (tsw_rtm/wave_solver.py:89-117): grid constants and
 * the Cerjan sponge, built on the device. */
int tsw_grid_create(int nx, int nz, double dx, double dz, double dt, int n_sponge,
                    double sponge_alpha, int free_surface_top, tsw_grid_t* out);
int tsw_grid_destroy(tsw_grid_t grid);
/* Copies the grid's sponge (nz*nx floats) to host `out` (WaveSolver2D.sponge). */
int tsw_grid_sponge(tsw_grid_t grid, float* out);
/* Replaces the grid's sponge by host `sponge` (nz*nx floats).  The reference
 * reads WaveSolver2D.sponge on every step (wave_solver.py:167), so the Python
 * layer re-uploads it when a caller has modified that array. */
int tsw_grid_set_sponge(tsw_grid_t grid, const float* sponge);

/* WaveSolver2D.forward_propagate_acoustic (tsw_rtm/wave_solver.py:128-186).
 * rec_out: (nt, nrec); snaps_out: NULL or ((nt-1)/k + 1, nz, nx). */
int tsw_forward_acoustic(tsw_grid_t grid, const float* vel, const float* src, int nt, int six,
                         int siz, const int* rix, int nrec, int riz, int k, float* rec_out,
                         float* snaps_out);

/* WaveSolver2D.backward_propagate_and_correlate (tsw_rtm/wave_solver.py:188-242).
 * data: (nt, nrec); snaps: (nsnap, nz, nx); image_inout accumulated in place. */
int tsw_backward_correlate(tsw_grid_t grid, const float* vel, const float* data, int nt,
                           const int* rix, int nrec, int riz, const float* snaps, long long nsnap,
                           float* image_inout, int k);

/* TSWRTMEngine.migrate_shot (tsw_rtm/rtm_engine.py:49-104): forward in
 * vel_fwd with snapshots kept in HBM, illumination, back-propagation in
 * vel_bwd.  image_out and illum_out: (nz, nx). */
int tsw_migrate_shot(tsw_grid_t grid, const float* vel_fwd, const float* vel_bwd,
                     const float* src, int nt_src, int six, int siz, const float* data, int nt,
                     const int* rix, int nrec, int riz, int k, float* image_out,
                     float* illum_out);

/* TSWRTMEngine.migrate_multiple_shots (tsw_rtm/rtm_engine.py:106-171):
 * every shot migrated as tsw_migrate_shot, images and illuminations stacked
 * in FP32 in shot order, optional sponge mute, illumination normalisation and
 * Laplacian post-filter.  srcs[i]: nt_srcs[i] samples; datas[i]: (nts[i], nrec).
 * Outputs (nz, nx): raw (stacked) image, Laplacian image, normalised image. */
int tsw_migrate_shots(tsw_grid_t grid, const float* vel_fwd, const float* vel_bwd, int nshot,
                      const float* const* srcs, const int* nt_srcs, const int* sixs,
                      const int* sizs, const float* const* datas, const int* nts, const int* rix,
                      int nrec, int riz, int k, int mute, float* raw_out, float* lap_out,
                      float* norm_out);

/* apply_laplacian_filter (tsw_rtm/rtm_engine.py:27-38): 5-point Laplacian with
 * symmetric boundary, bit-identical to scipy.signal.convolve2d(.., 'same',
 * 'symm') with the reference's float32 kernel. */
int tsw_laplacian(const float* image, float* out, int nx, int nz, double dx, double dz);

/* Elastic P-SV driver around elastic_2d_step (wave2d.c:95-212), which has no
 * caller in the reference (wave_solver.py:36-53 only declares it).  Build-
 * defined conventions: iteration it adds the vertical force (dt*bz)*s[it] to
 * vz at the source, snapshots / images the post-injection vz, steps, then
 * records the vertical strain rate (4th-order staggered d(vz)/dz, rows outside
 * the grid read 0) at the receivers on row riz.  Back-propagation injects the
 * data like the source and images vz against the forward vz snapshots with the
 * acoustic loop's snapshot cursor.  Models: lambda, mu, buoyancy_x,
 * buoyancy_z, each (nz, nx).  rec_out: (nt, nrec); snaps_out: NULL or
 * ((nt-1)/k+1, nz, nx) of vz; illum_out: NULL or sum of squared snapshots. */
int tsw_forward_elastic(tsw_grid_t grid, const float* lam, const float* mu, const float* bx,
                        const float* bz, const float* src, int nt, int six, int siz,
                        const int* rix, int nrec, int riz, int k, float* rec_out,
                        float* snaps_out, float* illum_out);
int tsw_backward_correlate_elastic(tsw_grid_t grid, const float* lam, const float* mu,
                                   const float* bx, const float* bz, const float* data, int nt,
                                   const int* rix, int nrec, int riz, const float* snaps,
                                   long long nsnap, float* image_inout, int k);
/* Elastic RTM of one shot: model_fwd / model_bwd are 4 pointers each
 * (lambda, mu, buoyancy_x, buoyancy_z); illumination = sum of squared vz
 * snapshots * dt*k. */
int tsw_migrate_shot_elastic(tsw_grid_t grid, const float* const* model_fwd,
                             const float* const* model_bwd, const float* src, int six, int siz,
                             const float* data, int nt, const int* rix, int nrec, int riz, int k,
                             float* image_out, float* illum_out);

/* Batched RTM (no reference counterpart; it batches what
 * TSWRTMEngine.migrate_multiple_shots, rtm_engine.py:125-143, runs one shot at
 * a time).  Acoustic shots form the acoustic lane and elastic shots
 * (kind TSW_SHOT_ELASTIC, vel_* unused) the elastic lane; the two lanes run on
 * their own streams and overlap.  Uploads every shot's models, source and data, sizes the HBM state
 * and captures the time loop into hipGraphs of `graph_steps` time steps
 * (0: no graphs; every graph of a blocked kernel ends with a join of its
 * launch groups, so fewer, larger graphs run faster).  Shots are packed into `slots` concurrent lanes (<= 0: all
 * shots at once).  kernel 0: one launch per time step over the tiles of every
 * active shot; kernel c in [1, tsw_kernel_configs()]: temporally blocked
 * launches of configuration c (T steps per launch).  The elastic lane runs
 * one launch per time step over every active elastic shot, its velocity and
 * stress passes fused (each shot holds two sets of its fields, read and
 * written alternately).  Per-shot results equal tsw_migrate_shot's (elastic:
 * tsw_migrate_shot_elastic's) bit for bit. */
int tsw_batch_create(int nshot, const tsw_shot_desc* shots, double dx, double dz, double dt,
                     int n_sponge, double sponge_alpha, int free_surface_top, int k, int slots,
                     int graph_steps, int kernel, tsw_batch_t* out);
/* Number of temporally blocked kernel configurations, and configuration
 * `id`'s steps per launch, core tile (W columns x H rows) and threads. */
int tsw_kernel_configs(void);
int tsw_kernel_config(int id, int* T, int* W, int* H, int* threads);
/* Configuration `id`'s kernel with its template arguments (for example
 * "k_tc<5,6>"), NUL-terminated into `buf` of `len` bytes. */
int tsw_kernel_name(int id, char* buf, int len);
/* Migrates every shot (mode TSW_RUN_BATCHED or TSW_RUN_PLAIN); synchronous. */
int tsw_batch_run(tsw_batch_t batch, int mode);
/* Runs one lane of the batch (0 acoustic, 1 elastic) once and returns the
 * kernel time and the number of launches.  The temporally blocked kernels
 * replay the batch's hipGraphs inside one HIP event pair (without graphs:
 * one event pair around each run of 32 launches); a launch there is one block
 * of T steps of every active shot, issued as one kernel per launch group (up
 * to 8 groups of slots, each on its own stream, forked from and joined back
 * into the batch's stream).  The per-step kernel is timed without graphs, one
 * event pair around each run of 32 back-to-back launches. */
int tsw_batch_time_kernels(tsw_batch_t batch, int lane, double* kernel_ms, int* launches);
/* Per-shot image and illumination (nz, nx) of the last run; either may be NULL. */
int tsw_batch_shot_result(tsw_batch_t batch, int shot, float* image_out, float* illum_out);
/* FP32 stack, in shot order, of every shot's image and illumination into a
 * (gnz, gnx) grid at the shots' window offsets. */
int tsw_batch_stack(tsw_batch_t batch, int gnx, int gnz, float* image_out, float* illum_out);
/* The same stack into caller-owned device arrays (gnz * gnx floats each),
 * enqueued on `stream` (a hipStream_t; NULL: the library's stream) without
 * waiting for it.  Replaces the stacking loop of rtm_engine.py:142-143. */
int tsw_batch_stack_device(tsw_batch_t batch, int gnx, int gnz, float* d_image, float* d_illum,
                           void* stream);
/* Cell-updates per run, device bytes held, global steps, captured graphs
 * (both lanes), and the elastic lane's share of the cell-updates. */
int tsw_batch_info(tsw_batch_t batch, long long* cell_updates, long long* device_bytes,
                   int* total_steps, int* n_graphs, long long* elastic_cell_updates);
int tsw_batch_destroy(tsw_batch_t batch);

/* Serving loop (no reference counterpart; the reference migrates a fixed
 * list of shots in a Python loop, rtm_engine.py:125-143).  A worker thread
 * admits submitted requests (acoustic or elastic shots, tsw_shot_desc) in
 * arrival order while fewer than `max_active` run and their HBM footprint fits
 * `budget_bytes`, and advances all running requests together `tick_steps`
 * steps at a time (blocked acoustic kernel configuration `kernel` on one
 * stream, elastic steps on another).  Results are bit-identical to migrating
 * each request alone.  Times are seconds on a monotonic clock since creation.
 * Request states: 0 queued, 1 running, 2 done, 3 failed, 4 cancelled. */
int tsw_server_create(double dx, double dz, double dt, int n_sponge, double sponge_alpha,
                      int free_surface_top, int k, long long budget_bytes, int max_active,
                      int kernel, int tick_steps, tsw_server_t* out);
/* Copies the request's host arrays and queues it; -3 if it can never fit. */
/* Models shared by requests, resident in HBM outside the budget: acoustic
 * forward/backward velocities and/or elastic forward/backward (lambda, mu,
 * buoyancy_x, buoyancy_z), each (gnz, gnx).  A request whose model pointers
 * are all null migrates in the window [z0, z0+nz) x [x0, x0+nx) of these,
 * cropped on the device at admission.  Only while the server is idle. */
int tsw_server_set_models(tsw_server_t srv, int gnx, int gnz, const float* vel_fwd,
                          const float* vel_bwd, const float* const* ela_fwd,
                          const float* const* ela_bwd);
int tsw_server_submit(tsw_server_t server, const tsw_shot_desc* req, long long* id_out);
int tsw_server_poll(tsw_server_t server, long long id, int* state, double* t_submit,
                    double* t_admit, double* t_done);
/* Blocks until request `id` is done, failed or cancelled (or the timeout). */
int tsw_server_wait(tsw_server_t server, long long id, double timeout_s);
/* Copies a done request's image and illumination and forgets the request.
 * For a failed or cancelled request it returns kErrState with the reason in
 * tsw_last_error() and forgets the request as well (a cancelled request that
 * was running is forgotten once its slot has been released at the next tick
 * boundary; until then the call reports "not done"). */
int tsw_server_result(tsw_server_t server, long long id, float* image_out, float* illum_out);
/* Cancels a queued request (it leaves the queue and frees its host copies at
 * once), or a running one at the next tick boundary. */
int tsw_server_cancel(tsw_server_t server, long long id);
int tsw_server_stats(tsw_server_t server, long long* completed, long long* failed,
                     long long* cancelled, long long* queued, long long* running,
                     long long* mem_used, long long* cell_updates, double* busy_s);
/* Cancels everything outstanding, stops the worker and frees the server. */
int tsw_server_destroy(tsw_server_t server);

/* Waveform misfit and its exact discrete adjoint (no reference counterpart;
 * the reference's back-propagation, wave_solver.py:188-242, images with the
 * forward operator and is not an adjoint).  J(v) = 1/2 sum |R P - d_obs|^2 over
 * the shots, with P from the reference's forward loop (wave_solver.py:157-183,
 * wave2d.c:19-82) in precision 4 (FP32) or 8 (FP64).  The forward levels are
 * checkpointed every `ckpt_interval` steps (0: ceil(sqrt(nt)); -1: every level
 * stored, no recompute) and recomputed during the adjoint sweep.  Device
 * pointers (vel, grad: nz*nx in the working precision) and `stream` (a
 * hipStream_t, e.g. torch's current stream) are the caller's. */
int tsw_fwi_create(int nx, int nz, double dx, double dz, double dt, int n_sponge,
                   double sponge_alpha, int free_surface_top, int precision, int nshot,
                   const tsw_fwi_shot* shots, int ckpt_interval, tsw_fwi_t* out);
/* Forward modelling of every shot in velocity d_vel; returns J. */
int tsw_fwi_forward(tsw_fwi_t fwi, const void* d_vel, void* stream, double* loss_out);
/* dJ/dv into d_grad (adjoint sweep with recompute); needs tsw_fwi_forward first. */
int tsw_fwi_backward(tsw_fwi_t fwi, const void* d_vel, void* d_grad, void* stream);
/* FP32 forward and recompute: 0 = one launch per step, c = 1..9 = temporally
 * blocked configuration c (the default is 2).  FP64 runs one launch per step. */
int tsw_fwi_set_blocking(tsw_fwi_t fwi, int config);
/* FP32 adjoint: 0 = the per-step kernels (see tsw_fwi_set_adjoint_steps),
 * c = 1..9 = temporally blocked configuration c, the shapes of
 * tsw_fwi_set_blocking (the default is 2).  The same bits either way.  FP64
 * runs the per-step kernels. */
int tsw_fwi_set_adjoint_blocking(tsw_fwi_t fwi, int config);
/* Per-step adjoint: steps per launch, 1 (default) or 2; the same bits either way. */
int tsw_fwi_set_adjoint_steps(tsw_fwi_t fwi, int steps);
/* Predicted data of a shot from the last forward, (nt, nrec), to the host. */
int tsw_fwi_predicted(tsw_fwi_t fwi, int shot, void* host_out);
/* Overwrites a shot's residual (nt, nrec) before tsw_fwi_backward (adjoint tests). */
int tsw_fwi_set_residual(tsw_fwi_t fwi, int shot, const void* host_res);
/* Adjoint field at the shot's source cell for every iteration of the last
 * backward sweep, (nt), to the host (adjoint tests). */
int tsw_fwi_probe(tsw_fwi_t fwi, int shot, void* host_out);
int tsw_fwi_destroy(tsw_fwi_t fwi);

/* Virtual-shot gather construction (the input of back-injection).
 *
 * Cross-correlations of source channels with receiver channels of one block
 * of continuous data; replaces compute_cross_correlation
 * (tsw_rtm/interferometry.py:26-79) for several sources at once and the
 * stacking over files (run_das_pipeline.py:202-213, accumulate != 0 adds in
 * FP32).  data: (nt, nch) float32 row-major; out: (nsrc, 2*nlag+1, nrec), lag
 * -nlag..nlag, out[s][j][r] = sum_t data[t][src_ch[s]] * data[t+j-nlag][rec_ch[r]]
 * (exact products, FP64 sums, one rounding to FP32). */
int tsw_xcorr(const float* data, int nt, int nch, const int* src_ch, int nsrc, const int* rec_ch,
              int nrec, int nlag, int accumulate, float* out);
/* Same on device pointers (d_data, d_out) and a caller stream; channel lists on the host. */
int tsw_xcorr_device(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                     const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                     void* stream);
/* Allocation-free variant of tsw_xcorr_device, with no counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and results, bit
 * for bit; the channel gathers, FP64 source rows and pair-major lag sums live
 * in a fixed 64 MiB device array per GPU (held from the library's load until
 * the process ends) instead of library-allocated buffers, and the call makes
 * no hipMalloc/hipFree.  Returns kErrArg (-1) when the call would need more
 * than that (tsw_xcorr_noalloc_scratch_bytes tells the need).  Every call on
 * one GPU uses that GPU's array: calls that overlap on one GPU (from several
 * host threads without the library's lock, or on several streams) race; the
 * call returns after its stream has drained, so calls in sequence are safe. */
int tsw_xcorr_device_noalloc(const float* d_data, int nt, int nch, const int* src_ch, int nsrc,
                             const int* rec_ch, int nrec, int nlag, int accumulate, float* d_out,
                             void* stream);
/* Scratch bytes tsw_xcorr_device_noalloc needs for these sizes (-1 for bad
 * sizes), and the size of its fixed array. */
long long tsw_xcorr_noalloc_scratch_bytes(int nt, int nsrc, int nrec, int nlag);
long long tsw_xcorr_noalloc_scratch_max(void);
/* time_derivative_greens (interferometry.py:82-104) of (nl, ncol) float32:
 * scheme 0 "central" (np.gradient, edge order 1), 1 spectral (irfft of
 * i*omega*rfft over the next power of two). */
int tsw_time_derivative(const float* in, int nl, int ncol, double dt, int scheme, float* out);
/* Branch of build_virtual_shot_gather (interferometry.py:140-155): 0 causal,
 * 1 anticausal (reversed), 2 symmetric (average); *nout must equal the
 * branch's row count. */
int tsw_green_branch(const float* in, int nl, int ncol, int zero_idx, int branch, float* out,
                     int* nout);
/* resample_gather_to_sim_dt (run_das_pipeline.py:48-55): linear
 * interpolation of (nt_orig, ncol) from step dt_orig to target_nt samples of
 * step dt_sim, 0 outside the original time range. */
int tsw_resample_linear(const float* in, int nt_orig, int ncol, double dt_orig, double dt_sim,
                        int target_nt, float* out);

/* One flood model of generate_velocity_flood_models
 * (tsw_rtm/velocity_flood.py:24-94): out[iz][ix] = base[iz][colmap[ix]]
 * (0 where colmap is -1), times `factor` when scale != 0, computed in the
 * base's precision (4: float32 with factor rounded to float32, 8: float64)
 * and stored as float32. */
int tsw_flood_columns(const void* base, int precision, int nx, int nz, const int* colmap,
                      int scale, double factor, float* out);

/* save_matrix_csv (tsw_rtm/visualizer.py:19-39), byte for byte: header
 * "row_name/col_name," + column axis (%.6g), then per row its axis value
 * (%.6g) and the values (%.6e).  matrix: (nr, nc) float32 (precision 4) or
 * float64 (8); axes float64.  Host only. */
int tsw_save_matrix_csv(const char* path, const void* matrix, int precision, int nr, int nc,
                        const double* row_axis, const double* col_axis, const char* row_name,
                        const char* col_name);

/* Preprocessing of continuous data (run_das_pipeline.py:165-168, :280; the
 * reference's preprocessing module is absent, the steps follow README.md:27-52).
 * These calls keep their FFT plans (per shape) and device scratch buffers (the
 * largest call's size) until the process ends, so that block after block of
 * one recording pays for neither again.
 *
 * Zero-phase filtering of x (nt, nch) along time with scipy.signal.sosfiltfilt
 * semantics: `sos` (nsec, 6) normalised sections, `zi` (nsec, 2) =
 * sosfilt_zi(sos), odd extension of `padlen` samples; FP64 arithmetic in
 * scipy's order.  Every `step`-th output sample is kept (decimation; out is
 * (ceil(nt/step), nch)), as its sign when sign != 0 (1-bit). */
int tsw_sosfiltfilt(const float* x, int nt, int nch, const double* sos, int nsec,
                    const double* zi, int padlen, int step, int sign, float* out);
/* np.sign(x).astype(float32) of n values. */
int tsw_sign(const float* x, long long n, float* out);
/* Spectral whitening along time of x (nt, nch): rfft, X / (|X| + eps) for
 * fmin <= f <= fmax and 0 elsewhere, irfft (FP32 transforms). */
int tsw_spectral_whitening(const float* x, int nt, int nch, double fs, double fmin, double fmax,
                           double eps, float* out);
/* The chain of one block on the device (run_das_pipeline.py:160-200, the
 * steps of README.md:27-52): x (nt, nch) on the host is copied once, low-pass
 * filtered and decimated by q (lp_*: as tsw_sosfiltfilt; skipped when q <= 1,
 * then lp_* may be null), band-pass filtered with the sign of each output
 * sample (bp_*), whitened at fs_proc between fmin and fmax, and left in d_out
 * (ceil(nt/q), nch) float32 on the device, for tsw_xcorr_device.  The results
 * equal tsw_sosfiltfilt (step q), tsw_sosfiltfilt (sign), then
 * tsw_spectral_whitening, bit for bit.  Runs on `stream` and returns once it
 * has drained. */
int tsw_preprocess_device(const float* x, int nt, int nch, const double* lp_sos, int lp_nsec,
                          const double* lp_zi, int lp_padlen, int q, const double* bp_sos,
                          int bp_nsec, const double* bp_zi, int bp_padlen, double fs_proc,
                          double fmin, double fmax, double eps, float* d_out, void* stream);
/* The same for nblk blocks of one shape at once (x[b] on the host, d_out[b]
 * on the device): the zero-phase filters run the channels of every block
 * side by side in one launch (they are sequential in time, one thread per
 * channel, so a single block leaves most of the GPU idle); each block's
 * result equals tsw_preprocess_device's bit for bit. */
int tsw_preprocess_device_blocks(int nblk, const float* const* x, int nt, int nch,
                                 const double* lp_sos, int lp_nsec, const double* lp_zi,
                                 int lp_padlen, int q, const double* bp_sos, int bp_nsec,
                                 const double* bp_zi, int bp_padlen, double fs_proc, double fmin,
                                 double fmax, double eps, float* const* d_out, void* stream);
/* F-K fan filter of a gather g (nt, nx): 2-D FFT, keep wavenumbers k != 0 with
 * vmin <= |f| / |k| <= vmax, inverse FFT, real part (FP32 transforms). */
int tsw_fk_filter(const float* g, int nt, int nx, double dt, double dx, double vmin, double vmax,
                  float* out);

#ifdef __cplusplus
}
#endif
#endif /* TSWRTM_H */
