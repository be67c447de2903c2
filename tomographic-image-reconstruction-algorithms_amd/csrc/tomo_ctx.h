// tomo_ctx.h -- the engine context and the helpers that the engine and the
// learned-preconditioner sources share (internal; the public ABI is
// include/tomo.h).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tomo.h"
#include "tomo_kernels.h"

// adjoint levels held per pass: three-step launches write z(k-1), z(k-2)
// while other strips still read z(k+1), z(k+2) on their halo rows
constexpr int kAdjRing = 5;

struct TimedSpan {
  hipEvent_t a = nullptr, b = nullptr;
  int kind = 0;          // 0 forward, 1 adjoint, 2 graph replay
  double launches = 0;
  double cells = 0;
  bool owned = false;    // events belong to a captured graph (not recycled)
};

// One captured gradient pass (shots [p0, p0 + np) in passes of P) and the
// pointers and parameters its launches baked in.  spans: the stencil loops'
// event pairs, recorded by event-record nodes inside the graph on every
// replay (empty when the runtime refused such nodes at capture; launches and
// cells then count for the replay as a whole).
struct GraphEntry {
  int p0 = 0, np = 0, P = 0;
  std::vector<uintptr_t> sig;
  hipGraphExec_t exec = nullptr;
  double launches[2] = {0, 0}, cells[2] = {0, 0};
  std::vector<TimedSpan> spans;
};

// Wavefield storage shared by contexts that are never used at the same time
// (tomo_workspace_create, tomo_set_workspace): it grows to the largest need
// among them.
struct tomo_workspace {
  int device = 0;
  float *d_ckpt = nullptr, *d_seg = nullptr, *d_z = nullptr;
  size_t ckpt_n = 0, seg_n = 0, z_n = 0;
  int users = 0;
};

struct tomo_ctx {
  tomo_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nx = 0, ny = 0, nt = 0;
  long long n2 = 0;
  // geometry (host)
  std::vector<int> tx, ty;  // transducers, 0-based
  std::vector<int4> bbox;   // per shot
  std::vector<int> rcell, ranchor;
  int nshots = 0, nr = 0;
  // phantoms stacked along the shot axis (tomo_set_phantoms): shots
  // [p groups, (p + 1) groups) belong to phantom p, whose field, truth and
  // coefficient images are n2 elements apart; nshots = nph * groups
  int nph = 1, groups = 0;
  // persistent device buffers
  float *d_f = nullptr, *d_fo = nullptr, *d_pw = nullptr;
  double *d_pw64 = nullptr;
  bool pw64_valid = false;
  int *d_cellmap = nullptr, *d_anchor = nullptr;
  int2 *d_ring = nullptr;                 // edge cells + receiver cells for the vectorised kernels
  int nring = 0;
  unsigned char *d_rowflag = nullptr;     // per row: holds a receiver
  // receiver of each 64th of a row (the lane segments of the two-step
  // adjoint, vec = nx / 64): receiver * vec + cell, or -1; recvseg_ok is
  // false when a segment holds two receivers (the two-step adjoint is then off)
  int *d_recvseg = nullptr;
  bool recvseg_ok = false;
  int4 *d_bbox = nullptr;
  float *d_spa = nullptr, *d_spab = nullptr;
  double *d_spa64 = nullptr, *d_spab64 = nullptr;
  float *d_meas = nullptr, *d_rr = nullptr, *d_grad = nullptr, *d_cov = nullptr;
  double *d_misfit = nullptr;
  double *d_scratch = nullptr;
  size_t scratch_n = 0;
  // wavefield pool: the context's own, or views of a shared workspace (ws)
  float *d_ckpt = nullptr, *d_seg = nullptr, *d_z = nullptr;
  size_t ckpt_n = 0, seg_n = 0, z_n = 0;
  tomo_workspace *ws = nullptr;
  double *d_ring64 = nullptr;
  size_t ring64_n = 0;
  // state
  bool have_truth = false, have_meas = false, coverage = false;
  std::vector<char> grad_valid;
  std::vector<float> amp32;   // pulse per step, float arithmetic
  std::vector<double> amp64;  // pulse per step, double arithmetic
  tomo::StencilShape shape_req{};  // requested fields (0 = automatic)
  // timing
  std::vector<hipEvent_t> ev_free;
  std::vector<TimedSpan> spans;
  double t_ms[3] = {0, 0, 0}, t_launch[3] = {0, 0, 0}, t_cells[3] = {0, 0, 0};
  // most forward steps per launch (1, 2 or 3; fewer where the larger
  // blockings do not apply)
  int time_block = 3;
  int adj_block = 3;  // most adjoint steps per launch (tomo_set_adjoint_blocking)
  int num_cus = 256;  // compute units of the device
  // per-pixel illumination weights of the updates (0 off, 1 coverage)
  int pixw_mode = 0;
  double pixw_eps = 0.05;
  float *d_pixw = nullptr;
  // graph replay of gradient passes
  bool use_graphs = false;
  bool capturing = false;
  double cap_launches[2] = {0, 0}, cap_cells[2] = {0, 0};
  std::vector<TimedSpan> cap_spans;  // event pairs recorded into the graph being captured
  bool cap_events_ok = true;         // every event-record node of this capture was accepted
  std::vector<GraphEntry> graphs;
};

namespace tomo_internal {

inline int fail(tomo_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail((ctx), TOMO_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
inline int dalloc(tomo_ctx *c, T **p, size_t n) {
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) return TOMO_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(c, TOMO_ERR_ALLOC, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  return TOMO_OK;
}

template <typename T>
inline void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int check_ctx(tomo_ctx *c);
int need_shots(tomo_ctx *c);
int need_all_gradients(tomo_ctx *c);
int upload_coefficients(tomo_ctx *c);  // per-cell coefficients from the current field
int ensure_scratch(tomo_ctx *c, size_t n);
// per-pixel update weights from the coverage of shots [g0, g0 + ng) into
// d_pixw (*w = d_pixw), or *w = nullptr when the weights are off
int pixel_weights(tomo_ctx *c, int g0, int ng, const float **w);
inline tomo::Mask ctx_mask(const tomo_ctx *c) {
  return tomo::Mask{c->cfg.mask_x0, c->cfg.mask_x1, c->cfg.mask_y0, c->cfg.mask_y1};
}

}  // namespace tomo_internal
