// tomo_kernels.h -- device kernel arguments and launchers of the gfx950
// wave-tomography engine (internal; the public ABI is include/tomo.h).
//
// Data layout in HBM (all i-fastest, rows of nx cells):
//   wavefield level   float[P][ny][nx]     (P = shots of the current pass)
//   coefficients pw   T[ny][nx]            shared by every shot
//   receiver map      int[ny][nx]          receiver index or -1, shared
//   traces/residuals  float[S][nt][nr]
//   gradient images   float[S][ny][nx]
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tomo {

// Stencil thread block: 64 lanes along x (one wavefront per row of the tile)
// x 4 wavefronts along y; each lane owns a vertical strip of RY cells.
constexpr int kBlockX = 64;
constexpr int kBlockY = 4;

struct Sponge {
  int width;           // 0: no sponge
  const void *a;       // T[width]  1/(1+g(d))
  const void *ab;      // T[width]  (1-g(d))/(1+g(d))
};

template <typename T>
struct FwdArgs {
  const T *um;          // level k-1 of local shot 0
  const T *u0;          // level k
  T *up;                // level k+1
  long long sstride;    // elements between consecutive shots in a level
  const T *pw;          // a * (v dt / h)^2 per cell
  const float *f;       // field (edge cells recompute r from it)
  const int *cellmap;   // receiver index per cell or -1
  const int4 *bbox;     // per global shot: x0, x1, y0, y1 (inclusive)
  const float *meas;    // measured traces [S][nt][nr] or null
  float *rec;           // traces / residuals [S][nt][nr] or null
  int nx, ny, nt, nr;
  int tiles_x, tiles_per_shot;
  int shot0;            // global index of local shot 0
  int level;            // k + 1
  int record;           // record this level
  T amp_h2;             // pulse(k) * h^2, 0 when no source this step
  T c0dt_h;             // c0 dt / h
  int sponge_w;
  const T *sp_a, *sp_ab;
  const int2 *ring;     // cells of the edge/receiver pass {flat cell, receiver or -1}
  int nring;
  // stacked phantoms: shot gs reads the coefficient (and field) image of
  // phantom gs / spp, pstride elements after phantom 0's; spp = 0: one image
  int spp;
  long long pstride;
};

struct AdjArgs {
  const float *z1;      // adjoint level k+1 of local shot 0
  const float *z2;      // level k+2
  float *z0;            // level k (written)
  const float *u;       // forward level k
  long long sstride;
  const float *pw;
  const int *cellmap;
  const int *anchor;    // per receiver: anchor cell (flat index)
  const float *rr;      // residual [S][nt][nr]
  float *acc;           // gradient accumulators [S][n2] (global shot index)
  float *cov;           // coverage accumulators [S][n2] or null
  int nx, ny, nt, nr;
  int tiles_x, tiles_per_shot;
  int shot0;
  int level;            // k
  float gain;           // residual injection factor
  float inv_h2;
  const unsigned char *row_has_recv;  // per row: 1 if the row holds a receiver
  int sponge_w;
  const float *sp_a, *sp_ab;
  const int2 *ring;     // cells of the edge/receiver pass
  int nring;
  // stacked phantoms: shot gs reads the coefficient (and field) image of
  // phantom gs / spp, pstride elements after phantom 0's; spp = 0: one image
  int spp;
  long long pstride;
};

// Kernel parameter of the stencil kernels: one problem (a), or a batch of
// problems whose descriptors live in device memory, item p owning the
// blocks [start[p], start[p + 1]) of the launch.
template <typename Args>
struct Launch {
  Args a;
  const Args *items;
  const int *start;
  int n;
};

// Stencil kernel shape: vec = cells per lane along x (1: scalar kernels,
// 2 or 4: vectorised kernels, FP32 only, nx % vec == 0), rows = cells per
// lane along y (scalar: forward 4/8/16, adjoint 4/8; vectorised: 4/8/16/32
// for both).
struct StencilShape {
  int vec;
  int fwd_rows;
  int adj_rows;
  int fwd_depth;  // vectorised kernels: rows whose loads are in flight ahead (1, 2, 4)
  int adj_depth;
  int rows;   // rows of the launch at hand (set by the caller from the fields above)
  int depth;  // prefetch depth of the launch at hand
  // vectorised kernels: cells done by the edge/receiver pass, {flat cell,
  // receiver index or -1} (edges and receivers of one shot)
  const int2 *ring;
  int nring;
};

// Stencil launchers (grid = tiles x P shots).
template <typename T>
hipError_t launch_forward_step(const FwdArgs<T> &a, int nshots, const StencilShape &sh, hipStream_t st);
hipError_t launch_adjoint_step(const AdjArgs &a, int nshots, const StencilShape &sh, hipStream_t st);

// Two forward steps k -> k+1 -> k+2 in one launch (sponge preset, waves of
// full rows: nx == 64 vec, vec in {2, 4, 8}), every row of both levels: the
// first and last strips form rows 0 and ny-1 of level k+1 from the inputs
// and of level k+2 from the level-k+1 rows in their registers (one-way edge
// rule).  With recv_seg set, the receiver records of levels k+1 (rec1) and
// k+2 (rec2) are written in the kernel too; otherwise they are left to
// launch_forward2_edges over the receivers.
struct Fwd2Args {
  const float *um, *u0;  // levels k-1, k of local shot 0
  float *up1, *up2;      // levels k+1, k+2
  long long sstride;
  const float *pw;
  const float *f;
  const int4 *bbox;      // per global shot
  int nx, ny;
  int shot0;
  int blocks_per_shot;
  float amp1, amp2;      // pulse * h^2 of steps k and k+1 (0: no source)
  float c0dt_h;
  int sponge_w;
  const float *sp_a, *sp_ab;
  // receiver records in the kernel (recv_seg null: none here)
  const int *recv_seg;   // [ny][nx / vec]: receiver * vec + cell of the lane's segment, or -1
  const unsigned char *row_has_recv;
  const float *meas;     // measured traces [S][nt][nr] (null: record u itself)
  float *rec;            // traces / residuals [S][nt][nr]
  int nt, nr;
  int level1;            // k+1
  int rec1, rec2;        // record levels k+1, k+2
  // stacked phantoms: shot gs reads the coefficient (and field) image of
  // phantom gs / spp, pstride elements after phantom 0's; spp = 0: one image
  int spp;
  long long pstride;
};
bool forward2_supported(int nx, int sponge_w);
// depth: rows whose loads are in flight ahead of the row being formed (1, 2)
hipError_t launch_forward2(Fwd2Args a, int nshots, int rows, int depth, hipStream_t st);

// Three forward steps k -> k+1 -> k+2 -> k+3 in one launch (same conditions
// as the two-step kernel, and ny a multiple of the strip height): u(k+1) and
// u(k+2) are kept in registers, u(k+3) and u(k+2) are always stored, u(k+1)
// only when store1 is set (a recording pass needs it only at checkpoints).
// Edge rows and, with recv_seg set, the receiver records of all three levels
// are formed in the kernel.
struct Fwd3Args {
  Fwd2Args b;            // levels k-1, k in, k+1 (up1), k+2 (up2) out; records of k+1, k+2
  float *up3;            // level k+3
  float amp3;            // pulse * h^2 of step k+2
  int rec3;              // record level k+3
  int store1;            // store level k+1
};
bool forward3_supported(int nx, int ny, int sponge_w, int rows);
hipError_t launch_forward3(Fwd3Args a, int nshots, int rows, hipStream_t st);

// The receiver records of a pair over ring[0 .. nring) when the main kernel
// does not write them (ring cells with a receiver index: records of levels
// k+1 (rec1) and k+2 (rec2); ring cells without one, the top and bottom
// rows: rows 0 and ny-1 of level k+2, and of level k+3 into up3 when set).
struct Fwd2EdgeArgs {
  const float *u0;        // level k of local shot 0
  float *up1, *up2, *up3; // levels k+1, k+2, k+3 (up3 may be null)
  long long sstride;
  const float *f;
  const int2 *ring;
  int nring;
  int nx, ny, nt, nr;
  int shot0;
  int level1;             // k+1
  int rec1, rec2;
  const float *meas;      // measured traces (null: record u itself)
  float *rec;
  float c0dt_h;
  // stacked phantoms: shot gs reads the coefficient (and field) image of
  // phantom gs / spp, pstride elements after phantom 0's; spp = 0: one image
  int spp;
  long long pstride;
};
hipError_t launch_forward2_edges(const Fwd2EdgeArgs &a, int nshots, hipStream_t st);

// Two adjoint steps k+1, k+2 -> k -> k-1 in one pass over a strip (sponge
// preset, point receivers with additive injection, no coverage; a wave
// holds full rows, nx == 64 VEC).  z(k) is kept in registers one row ahead
// of z(k-1), the receivers are injected in registers, and the gradient row
// is read and written once for both levels.  All rows of both levels: rows
// 0 and ny-1 are the copies of rows 1 and ny-2, with their correlation.
struct Adj2Args {
  const float *z1, *z2;  // adjoint levels k+1, k+2 of local shot 0
  float *z0, *zm;        // levels k, k-1 (written)
  const float *u0, *um;  // forward levels k, k-1
  long long sstride;
  const float *pw;
  const int *recv_seg;   // [ny][nx / VEC]: receiver * VEC + cell of the lane's segment, or -1
  const float *rr;       // residual [S][nt][nr]
  float *acc;            // gradient accumulators [S][n2] (global shot index)
  int nx, ny, nt, nr;
  int shot0;
  int blocks_per_shot;
  int level;             // k
  float gain, inv_h2;
  const unsigned char *row_has_recv;
  int sponge_w;
  const float *sp_a, *sp_ab;
  // stacked phantoms: shot gs reads the coefficient (and field) image of
  // phantom gs / spp, pstride elements after phantom 0's; spp = 0: one image
  int spp;
  long long pstride;
};
// Three adjoint steps k+1, k+2 -> k -> k-1 -> k-2 in one pass over a strip
// (same conditions as the two-step adjoint).  z(k) and z(k-1) are kept in
// registers; only z(k-1) and z(k-2), the levels the next launch starts
// from, are stored, and the gradient row is read and written once for all
// three levels (12 instead of 16 bytes per cell-update).  b holds levels
// k+1, k+2 (z1, z2), k-1 (zm, written) and the forward levels k (u0), k-1
// (um); b.z0 is not used.
struct Adj3Args {
  Adj2Args b;
  float *zm2;            // adjoint level k-2 (written)
  const float *um2;      // forward level k-2
};
bool adjoint3_supported(int nx, int sponge_w);
// depth: rows whose loads are in flight ahead of the row being formed (1, 2)
hipError_t launch_adjoint3(Adj3Args a, int nshots, int rows, int depth, hipStream_t st);

// Coverage of one checkpoint segment's forward levels k = khi .. klo (newest
// first) into cov[shot0 + s] for the nshots local shots; level k is
// lev01[k - k0] for k - k0 < 2, else seg + (k - k0) * lstride; shots are
// sstride apart.  Needs nx % 4 == 0.
struct CovSegArgs {
  const float *lev01[2];
  const float *seg;
  long long lstride, sstride;
  float *cov;
  int nx, ny, n2, nshots, shot0;
  int k0, klo, khi;
};
hipError_t launch_coverage_segment(const CovSegArgs &a, hipStream_t st);
// nx for which the two-step adjoint exists (vec = nx / 64)
bool adjoint2_supported(int nx, int sponge_w);
hipError_t launch_adjoint2(Adj2Args a, int nshots, int rows, hipStream_t st);

// Batched stencil launches (vectorised kernels, sponge preset): the main
// kernel over blocks_main blocks (item p: start_main[p] ..), then / before it
// the edge/receiver kernel over blocks_ring blocks (item p: start_ring[p] ..,
// shots x ceil(nring / 256) blocks).  Item descriptors carry tiles_x and
// tiles_per_shot for the shape (vec, rows).
hipError_t launch_forward_batch(const FwdArgs<float> *items, const int *start_main, const int *start_ring, int n,
                                int blocks_main, int blocks_ring, int vec, int rows, int depth, hipStream_t st);
hipError_t launch_adjoint_batch(const AdjArgs *items, const int *start_main, const int *start_ring, int n,
                                int blocks_main, int blocks_ring, bool cov, int vec, int rows, int depth,
                                hipStream_t st);
// tiles of one shot for a vectorised shape
void vec_tiles(int nx, int ny, int vec, int rows, int *tiles_x, int *tiles_per_shot);

// pw = a(d) * (c0 dt / h)^2 * (1 + f)
template <typename T>
hipError_t launch_coefficients(const float *f, T *pw, int nx, int ny, double c0dt_h, int sponge_w,
                               const T *sp_a, hipStream_t st);

hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t st);
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t st);

// per-shot sum of squares of rr[s][nt][nr] in FP64 -> out[s]
hipError_t launch_misfit(const float *rr, long long per_shot, int nshots, double *out, hipStream_t st);

// Mask rectangle for the update kernels.
struct Mask {
  int x0, x1, y0, y1;
};

// f += scale * mask * (sum_g G_g) / (1 + f) / div  [* pixw per pixel when given]
hipError_t launch_update_uniform(float *f, const float *G, int nshots, int nx, int ny, Mask m,
                                 double scale, double div, hipStream_t st, const float *pixw = nullptr);
// f += scale * mask * G_g / (1 + f)  [* pixw]
hipError_t launch_update_group(float *f, const float *Gg, int nx, int ny, Mask m, double scale,
                               hipStream_t st, const float *pixw = nullptr);
// per-pixel illumination weights w = cbar / (c + eps cmax) from the sum c of
// S coverage images (stats: 2 doubles of scratch)
hipError_t launch_pixel_weights(const float *cov, int S, int nx, int ny, Mask m, double eps, double *stats,
                                float *w, hipStream_t st);
// weighted SIRT: mean image (FP64), per-shot squared distance partials,
// weights (normalised, degenerate fallback), update.
hipError_t launch_weighted(float *f, const float *G, int M, int nx, int ny, Mask m, double scale,
                           double *mean_scratch, double *part_scratch, int part_blocks,
                           double *w_out, hipStream_t st, const float *pixw = nullptr);
// the weights of weighted SIRT only (no update)
hipError_t launch_weights_only(const float *f, const float *G, int M, int nx, int ny, double *mean, double *part,
                               int part_blocks, double *w, hipStream_t st);
// d = update direction of unit relaxation (uniform when w is null, else
// weighted), then out4 = {sum (f-fo)^2, sum (f-fo) d, sum d^2, sum fo^2}
hipError_t launch_error_quadratic(const float *f, const float *fo, const float *G, int M, int nx, int ny, Mask m,
                                  double div, const double *w, const float *pixw, double *d, double *out4,
                                  hipStream_t st);
// sums {sum (f - fo)^2, sum fo^2} in FP64 into out[2]
hipError_t launch_error(const float *f, const float *fo, int n, double *out, hipStream_t st);

// weighted average of intermediate fields (sirt-mimo-avg-weight.cu:80-203):
// mean[n2], part[M * part_blocks], w[M + 1] (w_g and their sum) are scratch
// the caller provides
hipError_t launch_avg_weight(float *f, const float *G, int M, int nx, int ny, Mask m, double scale, double *mean,
                             double *part, int part_blocks, double *w, hipStream_t st);
// number of partial blocks used by launch_weighted for n2 cells
int weighted_part_blocks(int n2);

}  // namespace tomo
