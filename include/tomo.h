/*
 * tomo.h -- C ABI of the MI355X wave-tomography engine (libtomo_hip.so).
 *
 * The engine runs the nonlinear wave-based SIRT / ART loop of
 * crogers719/Tomographic-Image-Reconstruction-Algorithms on one AMD GPU:
 * 2-D variable-velocity acoustic FDTD forward propagation with an absorbing
 * edge, sensor-trace record, residual, time-reversed adjoint propagation
 * with residual injection, forward x adjoint correlation (the gradient
 * image) and the uniform SIRT, non-uniform weighted SIRT and sequential ART
 * field updates.
 *
 * The reference has no library API: its "boundary" is a set of CUDA
 * kernels launched in a fixed order by one host function per executable,
 * plus the RAII buffer types of cuda_ptr.cuh.  Every entry point below names
 * the reference code it replaces (file:line, paths relative to the
 * reference repository root).
 *
 * Conventions
 *  - Arrays are row-major with i (x) fastest: a 2-D field is a(i, j) =
 *    a[i + nx * j], exactly the host_ptr indexing of cuda_ptr.cuh:54.
 *  - Traces are float[shots][nt][nr]: receiver r of shot s at time level k
 *    is traces[(s * nt + k) * nr + r].  tomo_traces_to_sides() converts to
 *    the reference's four per-side arrays (generate-data.cu:187-223).
 *  - Every function returns an int status: TOMO_OK (0) or a negative
 *    TOMO_ERR_* code; tomo_last_error() gives the message of the last
 *    failure on a context.  The caller owns every host buffer passed in;
 *    the context owns all device memory.  One context = one HIP stream on
 *    one device; a context is not thread-safe.
 */
#ifndef TOMO_H
#define TOMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TOMO_ABI_VERSION 1

enum tomo_status {
  TOMO_OK = 0,
  TOMO_ERR_ARG = -1,    /* bad argument (null pointer, size out of range) */
  TOMO_ERR_HIP = -2,    /* HIP runtime error */
  TOMO_ERR_ALLOC = -3,  /* device or host allocation failed */
  TOMO_ERR_STATE = -4,  /* call made before its prerequisites (no field, no data ...) */
  TOMO_ERR_IO = -5,     /* file missing / malformed */
  TOMO_ERR_NODEV = -6   /* no usable GPU */
};

/* source wavelets */
enum tomo_wavelet {
  /* reference pulse cos(wc t) exp(-t^2 / (2 tau^2)), t = k dt - t_shift,
   * added while k + 1 <= src_last_level (sirt-mimo.cu:553-562) */
  TOMO_WAVELET_REF_GAUSS_COS = 0,
  /* Ricker (1 - 2 pi^2 fp^2 t^2) exp(-pi^2 fp^2 t^2), t = k dt - t0 */
  TOMO_WAVELET_RICKER = 1
};

/* receiver layouts */
enum tomo_receivers {
  /* four sensor lines: rows/cols line_lo and line_hi, positions
   * line_first..line_last; residual injected by overwrite from the inward
   * neighbour (sirt-mimo.cu:617-649, 712-740) */
  TOMO_RECV_LINES = 0,
  /* point receivers set by tomo_set_receivers(); residual added at the cell */
  TOMO_RECV_POINTS = 1
};

/* uniform SIRT update variants */
enum tomo_uniform_mode {
  /* f += scale * mask * sum_g G_g / (1 + f)          (sirt-mimo.cu:858-877)   */
  TOMO_UPDATE_SUM = 0,
  /* f += scale * mask * (sum_g G_g) / (1 + f) / Ng   (sirt-mimo-io.cu:774-796) */
  TOMO_UPDATE_MEAN = 1
};

/* relaxation schedules (host scalars) */
enum tomo_schedule_kind {
  /* omega * (alpha + beta / (k + 1)^gamma), k = 0-based iteration
   * (sirt-mimo-io.cu:355); params = {omega, alpha, beta, gamma} */
  TOMO_SCHED_POWER = 0,
  /* k 1-based: odd k -> a1 + b1 / k^n1, even k -> a2 - b2 / k^n2
   * (sirt-mimo-norm-alt.cu:472-477); params = {a1, b1, n1, a2, b2, n2} */
  TOMO_SCHED_ALTERNATING = 1
};

typedef struct tomo_config {
  int32_t nx, ny;          /* grid cells (i fastest) */
  int32_t nt;              /* time levels 0..nt-1; levels 0 and 1 are zero */
  int32_t wavelet;         /* enum tomo_wavelet */
  double h;                /* cell size */
  double dt;               /* time step */
  double c0;               /* background speed, v = c0 sqrt(1 + f) */
  double wl_omega;         /* gauss-cos: angular frequency (OMEGAC) */
  double wl_tau;           /* gauss-cos: envelope width (TAO) */
  double wl_shift;         /* gauss-cos: time shift (TT) */
  double ricker_fp;        /* Ricker peak frequency */
  double ricker_t0;        /* Ricker delay */
  int32_t src_last_level;  /* source active while k + 1 <= this; <= 0: no window */
  int32_t sponge_width;    /* absorbing sponge cells inside the edge; 0 = edge rule only */
  double sponge_gamma;     /* peak sponge damping (per-step gamma at the edge) */
  int32_t recv_mode;       /* enum tomo_receivers */
  int32_t record_first;    /* first recorded time level (reference: 2) */
  int32_t line_lo, line_hi;       /* TOMO_RECV_LINES: the two line rows/cols */
  int32_t line_first, line_last;  /* TOMO_RECV_LINES: inclusive span along a line */
  double recv_gain;        /* adjoint injection factor (reference: H * 1000) */
  int32_t mask_x0, mask_x1, mask_y0, mask_y1;  /* update mask [x0,x1) x [y0,y1) */
  int32_t checkpoint_every;  /* 0: keep the whole forward history; C > 0: keep
                                every C-th level pair and recompute segments */
  int32_t shots_per_pass;    /* 0: all shots per pass; P: P shots per pass
                                (memory-scalable SIRT, sirt-mimo-norm-mem.cu:429) */
} tomo_config;

typedef struct tomo_ctx tomo_ctx;

/* ---------------------------------------------------------------- config */

/* Reference-parity preset: 192 x 192, NT = 401, H = 1e-3, DT = 3.3333e-7,
 * c0 = 1500, gauss-cos pulse cut at k + 1 <= 24, edge rule only, four sensor
 * lines at 20 / 180 spanning 21..179, record from level 2, gain H * 1000,
 * mask [21,180)^2, full history.  Replaces the #defines of
 * sirt-mimo.cu:17-35 and the literals of sirt-mimo.cu:525, 629, 724, 870. */
void tomo_config_reference(tomo_config *cfg);

/* Unit-spacing preset for an n x n grid and nt levels: h = 1, c0 = 1,
 * dt = 0.4 (CFL 0.4), Ricker fp = 0.1 (10 cells per wavelength), t0 = 1.5/fp,
 * 20-cell sponge, point receivers, mask = central n/2 square.  No
 * counterpart in the reference (it hard-codes the preset above). */
void tomo_config_unit(tomo_config *cfg, int n, int nt);

/* ------------------------------------------------------------- lifecycle */

/* Replaces the device_ptr allocations of Ultrasonic_Tomography()
 * (sirt-mimo-io.cu:97-231).  device < 0 picks the current device. */
int tomo_create(const tomo_config *cfg, int device, tomo_ctx **out);
void tomo_destroy(tomo_ctx *ctx);
const char *tomo_last_error(const tomo_ctx *ctx);
const char *tomo_strerror(int status);
int tomo_abi_version(void);

/* ------------------------------------------------------------- geometry */

/* Reference transducer ring: 640 positions on the square [21,181]^2, 160 per
 * side, bottom -> right -> top -> left, 1-based (Position_Transducers,
 * sirt-mimo.cu:438-466).  ns <= 640. */
int tomo_transducers_reference(int *ii, int *jj, int ns);

/* count transducers evenly on a circle of radius `radius` cells centred on
 * the grid centre, 0-based cell indices.  No reference counterpart. */
int tomo_transducers_ring(int nx, int ny, int count, double radius, int *ii, int *jj);

/* Set the transducer positions (one_based: 1 for the reference convention). */
int tomo_set_transducers(tomo_ctx *ctx, const int *ii, const int *jj, int ns, int one_based);

/* Consecutive source groups of `group_size` transducers; group g fires the
 * filled bounding box of transducers g*G and g*G+G-1 (sirt-mimo.cu:157 and
 * 543-554).  Returns the number of shots (the groups, once per phantom) or
 * a negative status. */
int tomo_set_groups(tomo_ctx *ctx, int group_size);

/* Stack `count` phantoms of this context's geometry (grid, transducers,
 * groups, receivers, steps) along the shot axis: phantom p owns shots
 * [p Ng, (p + 1) Ng) and its own field, truth and coefficient image, so one
 * stencil launch per level advances every shot of every phantom.  Fields
 * (tomo_set_field / tomo_get_field / tomo_set_truth) become
 * float[count][ny][nx], coverage maps float[count][ny][nx], tomo_error
 * writes one epsilon per phantom, and tomo_update_uniform /
 * tomo_update_weighted update each phantom from its own shots' gradients
 * (weights_out: double[count][Ng]).  ART, the intermediate-field average,
 * the line search and the learned updates take one phantom
 * (TOMO_ERR_STATE).  Resets the fields to 0 and drops the truth.  No
 * reference counterpart: the reference runs one phantom per process
 * (sirt-mimo-norm-mem.cu:429-458 splits one phantom's shots into passes). */
int tomo_set_phantoms(tomo_ctx *ctx, int count);
int tomo_num_phantoms(const tomo_ctx *ctx);

/* TOMO_RECV_POINTS only: receiver cells (one_based as above). */
int tomo_set_receivers(tomo_ctx *ctx, const int *ri, const int *rj, int nr, int one_based);

int tomo_num_shots(const tomo_ctx *ctx);
int tomo_num_receivers(const tomo_ctx *ctx);
/* receiver cells as 0-based (i, j) pairs, for trace layout conversions */
int tomo_get_receivers(const tomo_ctx *ctx, int *ri, int *rj);

/* ---------------------------------------------------------------- fields */

/* Image field f = c^2 / c0^2 - 1, float[ny][nx]. */
int tomo_set_field(tomo_ctx *ctx, const float *f);
int tomo_get_field(tomo_ctx *ctx, float *f);
/* ground truth fo for the error epsilon (sirt-mimo-io.cu:167-168) */
int tomo_set_truth(tomo_ctx *ctx, const float *fo);

/* Reference phantom: disk r = 15 mm at 0.03 plus three r = 5 mm disks at
 * 0.05, in float arithmetic (field_setup, sirt-mimo.cu:470-509;
 * generate-field.cu:117-172).  x = -0.1 + i * h. */
int tomo_phantom_reference(float *fo, int nx, int ny, double h);

/* Seeded circular inclusions inside the region [x0,x1) x [y0,y1): count
 * disks, radius uniform in [rmin, rmax] cells, velocity contrast uniform in
 * +-[cmin, cmax]; every cell of a disk is set to f = (1 + contrast)^2 - 1,
 * other cells keep their value.  No reference counterpart. */
int tomo_phantom_inclusions(float *fo, int nx, int ny, uint64_t seed, int count,
                            double rmin, double rmax, double cmin, double cmax,
                            int x0, int x1, int y0, int y1);

/* Seeded smooth random velocity: white noise filtered by a Gaussian of
 * width sigma cells, scaled to a peak contrast of +-amp and tapered to zero
 * over sigma cells at the edge of [x0,x1) x [y0,y1); sets every cell of fo
 * to f = (1 + c)^2 - 1 (0 outside the region).  sigma must lie in
 * (0, max(nx, ny)].  No reference counterpart. */
int tomo_phantom_smooth(float *fo, int nx, int ny, uint64_t seed, double sigma, double amp,
                        int x0, int x1, int y0, int y1);

/* ------------------------------------------------------ forward / data */

/* Forward-propagate shots [g0, g0+ng) on the current field and return their
 * traces float[ng][nt][nr] (propagation + propagation_at_corners +
 * initial_signal, generate-data.cu:157-195).  fp64 != 0 runs the FP64
 * solver (used to make measured data). */
int tomo_forward(tomo_ctx *ctx, int g0, int ng, int fp64, float *traces);

/* Measured traces float[nshots][nt][nr] (sirt-mimo-io.cu:170-180). */
int tomo_set_measured(tomo_ctx *ctx, const float *traces);

/* Add seeded Gaussian noise of standard deviation rel * RMS(trace) to each
 * receiver trace of float[nshots][nt][nr] in place.  No reference
 * counterpart (the reference's noisy generators are unseeded). */
int tomo_add_noise(float *traces, int nshots, int nt, int nr, double rel, uint64_t seed);

/* Reference file layout: four arrays float[nshots][nt][nx] (bottom, right,
 * top, left) with zeros off the sensor span, from/to float[nshots][nt][nr]
 * (TOMO_RECV_LINES contexts only; generate-data.cu:187-223). */
int tomo_traces_to_sides(const tomo_ctx *ctx, const float *traces, float *bottom,
                         float *right, float *top, float *left);
int tomo_sides_to_traces(const tomo_ctx *ctx, const float *bottom, const float *right,
                         const float *top, const float *left, float *traces);

/* -------------------------------------------------------------- gradient */

/* Data misfit sum_s sum_k sum_r (meas - u)^2 of all shots at the current
 * field: the forward half of tomo_gradient (sirt-mimo.cu:617-684 record and
 * difference), without the adjoint; for step-length searches. */
int tomo_misfit(tomo_ctx *ctx, double *misfit);

/* For shots [g0, g0+ng): forward on the current field, residual
 * measured - simulated at the receivers, time-reversed adjoint with residual
 * injection, and the correlation G_g(i,j) = sum_k z(i,j,k) * Lu(i,j,k)
 * (sirt-mimo-io.cu:333-353).  G_g is kept on the device per shot; misfit
 * (may be NULL) receives sum of squared residuals of those shots (FP64). */
int tomo_gradient(tomo_ctx *ctx, int g0, int ng, double *misfit);

/* raw per-shot gradient G_g, float[ny][nx] */
int tomo_get_gradient(tomo_ctx *ctx, int g, float *out);

/* Sensor-coverage map: forward-wavefield energy sum_g sum_k u(k)^2 over the
 * adjoint sweep (k = 1..nt-2), accumulated by tomo_gradient() while enabled
 * (off by default; it adds one read-modify-write per cell update).
 * float[ny][nx].  No reference counterpart. */
int tomo_enable_coverage(tomo_ctx *ctx, int on);
int tomo_get_coverage(tomo_ctx *ctx, float *out);

/* ---------------------------------------------------------------- updates */

/* Ground-truth line search over a scale grid s_k = s_lo + k s_step, k =
 * 0 .. count-1 (sirt-test-weight.cu:320-400, sirt-norm-test-weight.cu:
 * 407-500: scales 0, 100, .., 1e6): the scale whose update (uniform MEAN, or
 * weighted when `weighted`) gives the smallest relative error against the
 * truth, the first one on ties.  The update is linear in s, so eps(s)^2 is a
 * quadratic whose three FP64 coefficients locate the grid minimum exactly;
 * the field is not changed.  Needs the gradients of all shots. */
int tomo_line_search(tomo_ctx *ctx, int weighted, double s_lo, double s_step, int count, double *best_scale,
                     double *best_eps);

/* Per-pixel illumination weights of the field updates (C2's "per-pixel
 * weights from sensor-coverage normalisation"; no reference counterpart):
 * mode 1 multiplies each pixel's update by w(x) = cbar / (c(x) + eps cmax),
 * c the coverage (tomo_enable_coverage) summed over the shots of the update
 * (the single shot for tomo_update_group), cbar its mean over the mask and
 * cmax its maximum; mode 0 (default) leaves the updates as they are. */
int tomo_set_pixel_weights(tomo_ctx *ctx, int mode, double eps);

/* f += scale * mask * (sum_g G_g) / (1 + f) [/ Ng]  (enum tomo_uniform_mode). */
int tomo_update_uniform(tomo_ctx *ctx, float scale, int mode);
/* the same with phantom p of a stack (tomo_set_phantoms) at relaxation
 * scales[p], float[count]: each phantom keeps the relaxation a run of its
 * own would use.  No reference counterpart (one phantom per process). */
int tomo_update_uniform_each(tomo_ctx *ctx, const float *scales, int mode);

/* Non-uniform weighted SIRT (sirt-mimo-norm.cu:56-118, 873-892):
 * df_m = G_m / (1 + f); d_m = || df_m - mean_m df_m ||_2 over the full grid;
 * f += scale * mask * sum_m (d_m^-1 / sum d^-1) df_m.  All reductions FP64.
 * If some d_m is zero the weights fall back to uniform 1/M.
 * weights_out (may be NULL) receives the normalised weights, double[Ng]. */
int tomo_update_weighted(tomo_ctx *ctx, float scale, double *weights_out);
/* the same with phantom p of a stack at relaxation scales[p], float[count].
 * No reference counterpart. */
int tomo_update_weighted_each(tomo_ctx *ctx, const float *scales, double *weights_out);

/* The update of tomo_update_weighted on caller-owned DEVICE arrays, with no
 * device allocation: no counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops).
 * d_f float[ny][nx] is updated in place from d_grad float[nshots][ny][nx]
 * (raw per-shot gradients G_m); mask4 (host) = {x0, x1, y0, y1} half-open;
 * d_pixw (device, may be NULL) multiplies each pixel's step, as the
 * coverage weights of tomo_set_pixel_weights do; d_weights (device, may be
 * NULL) receives the normalised weights, double[nshots].  Same kernels,
 * order and bits as tomo_update_weighted.  The FP64 scratch (mean image,
 * partial distances, weights) lives in one static device array of
 * TOMO_NOALLOC_SCRATCH_DOUBLES per GPU, held until the process ends: calls
 * needing nx*ny + nshots*(p + 1) + 1 doubles beyond it (p = partial blocks,
 * 64 at 512 x 512; e.g. 512 x 512 with up to 4032 shots fits) return
 * TOMO_ERR_ARG, with the message in tomo_last_error(NULL).  Work is queued
 * on `stream` (a hipStream_t; NULL = the null stream) and not waited for.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (different streams or host threads) race.  The first call on a device
 * looks up the array's address (hipGetSymbolAddress); make it once before
 * any graph capture. */
#define TOMO_NOALLOC_SCRATCH_DOUBLES (1 << 19) /* 4 MiB */
int tomo_update_weighted_noalloc(float *d_f, const float *d_grad, int nshots, int nx, int ny, const int *mask4,
                                 double scale, const float *d_pixw, double *d_weights, void *stream);

/* Weighted average of intermediate fields (sirt-mimo-avg-weight.cu:80-203,
 * 530-544): fg_g = f + scale * mask * G_g / (1 + f) per group;
 * w_g = || fg_g - mean_g fg_g ||_2 over the full grid; f += sum_g fg_g / w_g
 * (the whole intermediate fields, f included, as the reference adds them).
 * All sums FP64.  A zero w_g (one group, or identical groups) leaves a
 * non-finite field, as the reference's division does.  weights_out (may be
 * NULL) receives double[Ng + 1]: w_1 .. w_Ng and their sum. */
int tomo_update_avg_weight(tomo_ctx *ctx, float scale, double *weights_out);

/* ART step for one shot (art-mimo-io.cu:381-400, 684):
 * f += scale * mask * G_g / (1 + f). */
int tomo_update_group(tomo_ctx *ctx, int g, float scale);

/* epsilon = 100 * ||f - fo|| / ||fo|| over the full grid, FP64
 * (sirt-mimo.cu:342-344, 427-436); one value per phantom (eps[count] after
 * tomo_set_phantoms). */
int tomo_error(tomo_ctx *ctx, double *eps);

/* ------------------------------------------------------------ iterations */

/* One SIRT iteration over all shots (gradient of every shot in passes of
 * shots_per_pass, then the update): weighted != 0 -> tomo_update_weighted,
 * else tomo_update_uniform(mode).  misfit / eps may be NULL (eps: one value
 * per phantom). */
int tomo_sirt_iteration(tomo_ctx *ctx, float scale, int weighted, int mode,
                        double *misfit, double *eps);

/* One ART iteration: shots in order, gradient then update per shot
 * (art-mimo-io.cu:637-685). */
int tomo_art_iteration(tomo_ctx *ctx, float scale, double *misfit, double *eps);

/* relaxation scale for 0-based iteration k (enum tomo_schedule_kind). */
double tomo_schedule(int kind, int k, const double *params);

/* -------------------------------------------------------------- text I/O */

/* mimo-io.cuh:22-45 layout: for each l, k, j: values i = 0..nx-1 each
 * followed by " ", then "\n"; "\n" after each k block and each l block.
 * precision <= 0 -> max_digits10 (9).  trailing_blank_lines = 0 gives the
 * IO_Files layout of sirt-mimo.cu:389-425 (one 2-D block, no blank lines). */
int tomo_write_text(const char *path, const float *a, int nx, int ny, int nz, int nw,
                    int precision, int trailing_blank_lines);
/* whitespace-separated read of nx*ny*nz*nw values (mimo-io.cuh:6-20) */
int tomo_read_text(const char *path, float *a, int nx, int ny, int nz, int nw);

/* ----------------------------------------------------- instrumentation */

/* Kernel timing gathered with HIP events on the context stream around the
 * forward-step and adjoint-step loops since the last reset:
 * out[0] = forward-step ms, out[1] = forward-step launches,
 * out[2] = adjoint-step ms, out[3] = adjoint-step launches,
 * out[4] = forward cell-updates, out[5] = adjoint cell-updates. */
int tomo_timing(tomo_ctx *ctx, double *out6);
int tomo_timing_reset(tomo_ctx *ctx);

/* Shape of the stencil kernels: vec = cells per lane along x (1 scalar, 2 or
 * 4 vectorised; FP32 solver only, nx % vec == 0); rows per lane along y for
 * the forward step (vec 1: 4/8/16; vec 2/4: 4/8/16/32) and the adjoint step
 * (vec 1: 4/8; vec 2/4: 4/8/16/32, and 64, which only the three-step adjoint
 * takes: the other adjoint kernels then run 32); prefetch depth of the vectorised kernels
 * (rows whose loads are in flight ahead of the row being computed: 1, 2, 4).
 * 0 in a field picks the default.  Every shape gives the same results up to
 * FP32 rounding.  No reference counterpart (its launch shapes are fixed,
 * sirt-mimo.cu:165-299). */
int tomo_set_shape(tomo_ctx *ctx, int vec, int fwd_rows, int adj_rows, int fwd_depth, int adj_depth);

/* Stencil steps per launch in the forward passes: 2 runs two leapfrog
 * steps in one pass over each strip (the intermediate level kept in
 * registers: 8 instead of 12 bytes of traffic per forward cell-update) where
 * it applies -- vectorised kernels, nx of 128, 192, 256, 384 or 512 (64
 * lanes of 2, 3, 4, 6 or 8 cells), checkpoint interval >= 4, with or
 * without the sponge -- and 1 otherwise.  3 (default) runs three
 * forward steps per launch where, in addition, the strips tile the rows
 * exactly and the checkpoint interval is >= 5 (a recording pass then writes
 * only the last two levels of each launch: 16 bytes per three
 * cell-updates), else two or one.  Where the forward pass runs two or three
 * steps, the adjoint sweep runs two (16 instead of 24 bytes per
 * cell-update) with the sponge preset, when the receivers are points with
 * additive injection and no
 * 64th of a row holds two of them; the coverage map (tomo_enable_coverage)
 * is then summed per checkpoint segment by a pass of its own, in the
 * one-step sweep's order.  Results are the same bit for bit.  tomo_get_time_blocking returns the forward steps per
 * launch in effect, tomo_get_adjoint_blocking the adjoint's.  No reference
 * counterpart (one launch per step, sirt-mimo.cu:291-299). */
int tomo_set_time_blocking(tomo_ctx *ctx, int steps);
int tomo_get_time_blocking(const tomo_ctx *ctx);
int tomo_get_adjoint_blocking(const tomo_ctx *ctx);

/* Most adjoint steps per launch (1, 2 or 3; default 3).  Where the two-step
 * adjoint applies (above) and the grid is 256 or 512 cells wide, 3 runs
 * three adjoint steps per launch: z(k) and z(k-1) stay in registers and
 * only z(k-1), z(k-2) are stored, 12 instead of 16 bytes per cell-update.
 * Results are the same bit for bit.  No reference counterpart
 * (backpropagation1/2 run one step per launch, sirt-mimo.cu:316-327). */
int tomo_set_adjoint_blocking(tomo_ctx *ctx, int steps);

/* Autotuner: times the vectorised forward and adjoint kernels of this
 * context (its grid, shots per pass, sponge and coverage) over the shapes
 * vec in {4, 2} x rows in {8, 16, 32} x prefetch depth in {1, 2, 4}
 * (adjoint depth <= 2) on the context's own buffers, picks per lane width
 * the fastest forward and adjoint settings, and keeps the lane width with
 * the lowest gradient-pass cost (forward time x forward launches per
 * adjoint launch + adjoint time).  Where two steps per launch apply
 * (tomo_set_time_blocking), the strip heights are then re-timed on the
 * two-step kernels (the forward's on the three-step kernel where that
 * applies), which run all but the tails.  The result is set as
 * the context's shape and cached per size class for the process (force = 1
 * re-times).  out5 (may be NULL) = {vec, fwd_rows, adj_rows, fwd_depth,
 * adj_depth}; ms2 (may be NULL) = {forward, adjoint} ms per launch of one
 * step (per level for the two-step kernels).  Invalidates stored
 * gradients.  No reference counterpart (the reference launches one thread
 * per cell in fixed 2-D blocks, sirt-mimo.cu:290-300). */
int tomo_autotune(tomo_ctx *ctx, int force, int *out5, double *ms2);

/* Graph replay of gradient passes (off by default).  When on, each pass of
 * tomo_gradient (all its fills, forward steps, misfit, segment recomputes
 * and adjoint steps: about 3 nt launches) is captured into a HIP graph on
 * first use and replayed afterwards; a pass is captured again when any
 * buffer, geometry, shape or coverage setting it baked in has changed.
 * The results are bitwise those of the eager launches.  Replaces the
 * reference's per-step launch sequence (sirt-mimo.cu:326-366) with one
 * launch per pass. */
int tomo_set_graphs(tomo_ctx *ctx, int on);
/* Replayed graphs: out3 = {ms of the replays, replays, cell updates of
 * replays whose loops were not timed inside the graph}.  A captured pass
 * records its stencil loops' event pairs as event-record nodes of the graph,
 * so on replay they are timed and counted in tomo_timing like eager loops;
 * where the runtime refused such nodes at capture, the replay's cell updates
 * are counted here instead. */
int tomo_timing_graphs(tomo_ctx *ctx, double *out3);
/* out5 = {vec, fwd_rows, adj_rows, fwd_depth, adj_depth} in effect */
int tomo_get_shape(const tomo_ctx *ctx, int *out5);

/* -------------------------------------------------------------- workspace
 * Wavefield storage (checkpoint level pairs, the segment being replayed and
 * the adjoint ring: the bulk of a context's device memory) shared by
 * contexts that are never used at the same time, e.g. contexts stepped one
 * after another by one host thread.  The storage holds nothing between
 * calls; it grows to the largest need among its contexts, so a set of
 * problems of different sizes can run every shot in one pass
 * (shots_per_pass = 0) in the memory of its largest member.  No reference
 * counterpart (the reference allocates per problem,
 * sirt-mimo-norm-mem.cu:262-282). */
typedef struct tomo_workspace tomo_workspace;

int tomo_workspace_create(int device, tomo_workspace **out);
/* TOMO_ERR_STATE while contexts still use it (destroy or detach them first) */
int tomo_workspace_destroy(tomo_workspace *ws);
/* attach ctx to ws (NULL: back to storage of its own); calls on contexts
 * that share a workspace must not overlap; batched contexts cannot share one */
int tomo_set_workspace(tomo_ctx *ctx, tomo_workspace *ws);

/* ------------------------------------------------------------------ batch
 * Many independent problems stepped together (uneven batches of phantoms
 * of different grid, shot and step counts).  No reference counterpart: the
 * reference runs one problem per process.  Every time level of every shot
 * pass is one launch per stencil shape over the tiles of all problems
 * active at that level (problems aligned on the absolute level k; a common
 * checkpoint interval; shot passes in lockstep).  The contexts stay owned by
 * the caller and must outlive the batch; each needs its shots, receivers
 * and measured traces set, the sponge preset and the vectorised kernels
 * (nx a multiple of 2 or 4, nx >= 8), one checkpoint interval for all.  The
 * launch program is built at creation for the contexts' buffers and shapes
 * as they are then (set the shapes, e.g. with tomo_autotune, before);
 * tomo_batch_gradient returns TOMO_ERR_STATE once any context has changed
 * them (shots, receivers, coverage, shape ...).  Programs of more than
 * 16384 launches per gradient call are refused at creation (TOMO_ERR_ARG):
 * split such sets into several batches. */
typedef struct tomo_batch tomo_batch;

int tomo_batch_create(tomo_ctx *const *ctxs, int n, tomo_batch **out);
void tomo_batch_destroy(tomo_batch *batch);
const char *tomo_batch_last_error(const tomo_batch *batch);
/* replay the whole program as one captured HIP graph (off by default) */
int tomo_batch_set_graph(tomo_batch *batch, int on);
/* the gradients of all shots of every context (as tomo_gradient on each),
 * misfits[n] (may be NULL) per context */
int tomo_batch_gradient(tomo_batch *batch, double *misfits);
/* out4 = {ms, cell updates, launches, gradient calls} since creation */
int tomo_batch_timing(const tomo_batch *batch, double *out4);

/* ------------------------------------------------------------------ learned
 * Learned preconditioner.  No counterpart in the reference (which has no
 * learned component): a 5-layer 3x3 convolutional network, channels
 * 3 -> 32 -> 32 -> 32 -> 32 -> 1 with ReLU after layers 1-4, bfloat16
 * weights and activations on the matrix cores with FP32 accumulation.  Its
 * input at each pixel is (field / 0.1, plain SIRT step / 0.1, summed coverage
 * / its maximum), the plain SIRT step being scale * mask * (sum_s G_s) /
 * (1 + f) / S for the relaxation `scale` of the plain iteration; its output,
 * times 0.1 and the imaging mask, is a field update.  Trained with Adam on
 * samples taken from plain SIRT trajectories, target (truth - field) / 0.1
 * over the mask.
 *
 * Parameters (FP32, 28929 values): per layer W[o][c][ky][kx] then b[o],
 * y[p, o] = b[o] + sum W[o][c][ky][kx] x[p + (ky - 1, kx - 1), c], zero
 * padding.  Image width must be a multiple of 32.
 */
typedef struct tomo_cnn tomo_cnn;

int tomo_cnn_num_params(void);
/* network on `device` (-1: current), initialised with seed 1 */
int tomo_cnn_create(int device, tomo_cnn **out);
void tomo_cnn_destroy(tomo_cnn *net);
const char *tomo_cnn_last_error(const tomo_cnn *net);
/* He-normal weights (output layer scaled by 0.1), zero biases; resets Adam */
int tomo_cnn_init(tomo_cnn *net, uint64_t seed);
int tomo_cnn_set_params(tomo_cnn *net, const float *params);  /* resets Adam */
int tomo_cnn_get_params(tomo_cnn *net, float *params);
/* y[B][H][W] from x[B][H][W][3] (host arrays) */
int tomo_cnn_forward(tomo_cnn *net, int batch, int h, int w, const float *x, float *y);
/* grad[P] = d(sum dy * y)/d(params) for y = net(x) (host arrays) */
int tomo_cnn_backward(tomo_cnn *net, int batch, int h, int w, const float *x, const float *dy, float *grad);
/* masked mean-square loss sum_{mask} (y - t)^2 / (2 B |mask|), mask4 =
 * {x0, x1, y0, y1} half-open; grad may be NULL */
int tomo_cnn_loss(tomo_cnn *net, int batch, int h, int w, const float *x, const float *target, const int *mask4,
                  double *loss, float *grad);
/* training samples from an engine context: its current field, the plain
 * SIRT step with relaxation `scale` from the gradients of all shots
 * (tomo_gradient) and the coverage, target from its truth; all samples share
 * one grid size and mask */
int tomo_cnn_add_sample(tomo_cnn *net, tomo_ctx *ctx, float scale);
/* the same for phantom p of a context of stacked phantoms
 * (tomo_set_phantoms): its field, the plain step from its own shots'
 * gradients, its coverage and its truth; one call per phantom collects the
 * samples of many phantoms from one batched gradient.  No reference
 * counterpart. */
int tomo_cnn_add_sample_phantom(tomo_cnn *net, tomo_ctx *ctx, int p, float scale);
/* storage for `count` samples of ny x nx cells, allocated once (samples
 * already held are kept), so that collecting many samples does not regrow
 * it step by step */
int tomo_cnn_reserve_samples(tomo_cnn *net, int count, int ny, int nx);
int tomo_cnn_num_samples(const tomo_cnn *net);
/* new targets (field - f_s) / 0.1 for samples [first, first + count), f_s the
 * field each was taken at (e.g. field = a later iterate of the trajectory) */
int tomo_cnn_set_targets(tomo_cnn *net, int first, int count, const float *field);
int tomo_cnn_clear_samples(tomo_cnn *net);
/* `steps` Adam steps (beta 0.9 / 0.999) on random batches; losses[steps] optional */
int tomo_cnn_train(tomo_cnn *net, int steps, int batch, double lr, uint64_t seed, double *losses);
/* f += mask * (alpha * 0.1 * net(input) + max(1 - alpha, 0) * plain step),
 * the input from the context's state with the plain step of relaxation
 * `scale` (alpha = 1: the learned step alone; alpha -> 0: plain SIRT);
 * needs the gradients of all shots at the current field */
int tomo_learned_update(tomo_ctx *ctx, tomo_cnn *net, float alpha, float scale);
/* One learned iteration with the plain fallback: the gradient of all shots
 * at the current field f (misfit m0); the plain SIRT step from f (uniform
 * MEAN, or weighted when `weighted`, with relaxation `scale`) is kept as the
 * fallback; the learned candidate of tomo_learned_update with factor a is
 * formed and its data misfit m1 computed (forward only).  If m1 < m0 the
 * candidate is the new field (*accepted = 1, a grows by 1.5x up to `alpha`),
 * else the plain step is (*accepted = 0, a halves).  a starts at `alpha` on
 * a new context.  *misfit = m0; *eps = relative error (%) of the new field
 * when the context has a truth.  Cost: one gradient and one forward pass. */
int tomo_learned_iteration(tomo_ctx *ctx, tomo_cnn *net, float alpha, float scale, int weighted, double *misfit,
                           double *eps, int *accepted);
/* forget a pending learned step (start of a new run) */
int tomo_learned_reset(tomo_cnn *net);
/* while on, each tomo_learned_iteration also adds a training sample (as
 * tomo_cnn_add_sample) at the state its gradient was taken at */
int tomo_cnn_set_sampling(tomo_cnn *net, int on);

#ifdef __cplusplus
}
#endif

#endif /* TOMO_H */
