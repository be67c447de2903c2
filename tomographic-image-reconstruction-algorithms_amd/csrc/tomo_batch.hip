// tomo_batch.hip -- many independent problems (contexts of different grid
// sizes, shot counts and step counts) stepped together: every time level is
// one launch per stencil shape covering the tiles of every problem active at
// that level.  No reference counterpart: the reference runs one problem per
// process.
//
// Alignment: problems share the absolute time level k.  The recording
// forward pass runs k = 1 .. max(nt) - 2, each problem while k < nt - 1;
// with a common checkpoint interval C the backward sweep walks the segments
// m = max .. 0 and, inside a segment, recomputes levels mC + 1 .. and then
// steps the adjoint down from the segment's top, each problem while the
// level lies inside its own run.  Shot passes (shots_per_pass) are processed
// in lockstep: pass j of every problem that has one.
//
// The whole launch program (per launch: the problems' argument descriptors
// and block prefix tables) is built once into device memory; a gradient call
// replays it, optionally as one captured HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "tomo.h"
#include "tomo_ctx.h"
#include "tomo_host.h"
#include "tomo_kernels.h"

using namespace tomo;
using namespace tomo_internal;

// engine functions shared with this file (tomo_engine.hip)
namespace tomo_internal {
StencilShape context_shape(const tomo_ctx *c);
float *level_pass1(tomo_ctx *c, int P, int k);
float *level_segment(tomo_ctx *c, int P, int m, int k);
FwdArgs<float> forward_args(tomo_ctx *c, int shot0);
AdjArgs adjoint_args(tomo_ctx *c, int shot0);
float pulse_h2_f32(const tomo_ctx *c, int k);
int wavefields(tomo_ctx *c, int P);
int shots_per_pass(const tomo_ctx *c);
std::vector<uintptr_t> context_signature(const tomo_ctx *c);
}  // namespace tomo_internal

namespace {

// Largest launch program accepted (kernel launches plus fills per gradient
// call).  The largest programs run so far are the engine's captured gradient
// passes of about 3 nt launches (12k at nt = 4000); the one much larger batch
// program attempted (about 260k launches in one graph, 64 contexts of up to
// 512 x 512 and 4096 steps) ended in a host-side crash whose cause is not
// established, so programs beyond this size are refused.
constexpr double kMaxProgramLaunches = 16384;

struct Rec {
  int kind;            // 0 forward step, 1 adjoint step
  int vec, rows, depth;
  bool cov;
  size_t item0;        // first descriptor of this launch
  size_t start0;       // first entry of its main / ring prefix tables (n + 1 each)
  int n;
  int blocks_main, blocks_ring;
  double cells;
};

struct PassItem {
  int ctx;
  int p0, np;
};

struct Pass {
  std::vector<PassItem> items;
  size_t fwd_begin, fwd_end;  // records of the recording forward pass
  size_t seg_begin, seg_end;  // records of the backward sweep
};

}  // namespace

struct tomo_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<tomo_ctx *> ctx;
  std::vector<std::vector<uintptr_t>> sig;  // context_signature of each context at build
  std::vector<int> P;
  std::vector<Pass> passes;
  std::vector<Rec> recs;
  std::vector<FwdArgs<float>> hfwd;
  std::vector<AdjArgs> hadj;
  std::vector<int> hstart;
  FwdArgs<float> *d_fwd = nullptr;
  AdjArgs *d_adj = nullptr;
  int *d_start = nullptr;
  bool use_graph = false;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ms = 0.0, cells = 0.0, launches = 0.0, calls = 0.0;
  double cells_per_call = 0.0, launches_per_call = 0.0;
};

namespace {

int bfail(tomo_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}

#define BCHK(b, call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return bfail((b), TOMO_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// one forward or adjoint launch over the (item, level) pairs given
template <typename Args>
void add_rec(tomo_batch *b, int kind, const std::vector<std::pair<Args, int>> &items, std::vector<Args> &store,
             int vec, int rows, int depth, bool cov) {
  Rec r{};
  r.kind = kind;
  r.vec = vec;
  r.rows = rows;
  r.depth = depth;
  r.cov = cov;
  r.item0 = store.size();
  r.start0 = b->hstart.size();
  r.n = (int)items.size();
  int main = 0, ring = 0;
  std::vector<int> sm, sr;
  for (const auto &it : items) {
    Args a = it.first;
    const int np = it.second;
    vec_tiles(a.nx, a.ny, vec, rows, &a.tiles_x, &a.tiles_per_shot);
    store.push_back(a);
    sm.push_back(main);
    sr.push_back(ring);
    main += a.tiles_per_shot * np;
    ring += np * ((a.nring + 255) / 256);
    r.cells += (double)np * a.nx * a.ny;
  }
  sm.push_back(main);
  sr.push_back(ring);
  b->hstart.insert(b->hstart.end(), sm.begin(), sm.end());
  b->hstart.insert(b->hstart.end(), sr.begin(), sr.end());
  r.blocks_main = main;
  r.blocks_ring = ring;
  b->recs.push_back(r);
}

// group the (args, shots) pairs of one level by stencil shape and add a
// launch per group
template <typename Args>
void add_level(tomo_batch *b, int kind, const std::vector<std::tuple<Args, int, StencilShape>> &items,
               std::vector<Args> &store) {
  std::map<std::tuple<int, int, int, int>, std::vector<std::pair<Args, int>>> groups;
  for (const auto &t : items) {
    const StencilShape &s = std::get<2>(t);
    const Args &a = std::get<0>(t);
    const int rows = kind == 0 ? s.fwd_rows : s.adj_rows;
    const int depth = kind == 0 ? s.fwd_depth : s.adj_depth;
    int cov = 0;
    if constexpr (std::is_same<Args, AdjArgs>::value) cov = a.cov != nullptr;
    groups[std::make_tuple(s.vec, rows, depth, cov)].push_back(std::make_pair(a, std::get<1>(t)));
  }
  for (const auto &g : groups)
    add_rec(b, kind, g.second, store, std::get<0>(g.first), std::get<1>(g.first), std::get<2>(g.first),
            std::get<3>(g.first) != 0);
}

int build(tomo_batch *b) {
  const int K = (int)b->ctx.size();
  int C = -1;
  for (int i = 0; i < K; ++i) {
    tomo_ctx *c = b->ctx[i];
    const int ci = (c->cfg.checkpoint_every <= 0 || c->cfg.checkpoint_every >= c->nt) ? 0 : c->cfg.checkpoint_every;
    if (C < 0) C = ci;
    if (ci != C) return bfail(b, TOMO_ERR_ARG, "contexts of a batch need one checkpoint interval");
  }
  const bool full = C == 0;
  int max_pass = 0;
  for (int i = 0; i < K; ++i) {
    tomo_ctx *c = b->ctx[i];
    b->P.push_back(shots_per_pass(c));
    max_pass = std::max(max_pass, (c->nshots + b->P[i] - 1) / b->P[i]);
    int rc = wavefields(c, b->P[i]);
    if (rc) return bfail(b, rc, c->err);
  }
  for (int j = 0; j < max_pass; ++j) {
    Pass pass{};
    int max_nt = 0;
    for (int i = 0; i < K; ++i) {
      const int p0 = j * b->P[i];
      if (p0 >= b->ctx[i]->nshots) continue;
      pass.items.push_back(PassItem{i, p0, std::min(b->P[i], b->ctx[i]->nshots - p0)});
      max_nt = std::max(max_nt, b->ctx[i]->nt);
    }
    // recording forward pass
    pass.fwd_begin = b->recs.size();
    for (int k = 1; k < max_nt - 1; ++k) {
      std::vector<std::tuple<FwdArgs<float>, int, StencilShape>> lv;
      for (const auto &it : pass.items) {
        tomo_ctx *c = b->ctx[it.ctx];
        if (k >= c->nt - 1) continue;
        const int P = b->P[it.ctx];
        FwdArgs<float> a = forward_args(c, it.p0);
        a.um = level_pass1(c, P, k - 1);
        a.u0 = level_pass1(c, P, k);
        a.up = level_pass1(c, P, k + 1);
        a.level = k + 1;
        a.record = (k + 1 >= c->cfg.record_first) && c->nr > 0;
        a.meas = c->d_meas;
        a.rec = c->d_rr;
        a.amp_h2 = pulse_h2_f32(c, k);
        lv.emplace_back(a, it.np, context_shape(c));
      }
      add_level(b, 0, lv, b->hfwd);
    }
    pass.fwd_end = b->recs.size();
    // backward sweep, newest segment first
    pass.seg_begin = b->recs.size();
    const int Cc = full ? max_nt : C;
    const int nseg = (max_nt + Cc - 1) / Cc;
    for (int m = nseg - 1; m >= 0; --m) {
      const int k0 = m * Cc;
      if (!full) {
        for (int q = 2; q < Cc; ++q) {
          std::vector<std::tuple<FwdArgs<float>, int, StencilShape>> lv;
          for (const auto &it : pass.items) {
            tomo_ctx *c = b->ctx[it.ctx];
            const int cnt = std::min(Cc, c->nt - k0);
            if (cnt <= 2 || q >= cnt) continue;
            const int P = b->P[it.ctx];
            const int k = k0 + q - 1;
            FwdArgs<float> a = forward_args(c, it.p0);
            a.um = level_segment(c, P, m, k - 1);
            a.u0 = level_segment(c, P, m, k);
            a.up = level_segment(c, P, m, k + 1);
            a.level = k + 1;
            a.record = 0;
            a.amp_h2 = pulse_h2_f32(c, k);
            lv.emplace_back(a, it.np, context_shape(c));
          }
          add_level(b, 0, lv, b->hfwd);
        }
      }
      for (int q = Cc - 1; q >= 0; --q) {
        const int k = k0 + q;
        std::vector<std::tuple<AdjArgs, int, StencilShape>> lv;
        for (const auto &it : pass.items) {
          tomo_ctx *c = b->ctx[it.ctx];
          const int cnt = std::min(Cc, c->nt - k0);
          if (cnt <= 0 || q >= cnt || k < 1 || k > c->nt - 2) continue;
          const int P = b->P[it.ctx];
          const size_t lvl = (size_t)P * c->n2;
          AdjArgs a = adjoint_args(c, it.p0);
          a.z1 = c->d_z + (size_t)((k + 1) % 3) * lvl;
          a.z2 = c->d_z + (size_t)((k + 2) % 3) * lvl;
          a.z0 = c->d_z + (size_t)(k % 3) * lvl;
          a.u = level_segment(c, P, m, k);
          a.level = k;
          lv.emplace_back(a, it.np, context_shape(c));
        }
        add_level(b, 1, lv, b->hadj);
      }
    }
    pass.seg_end = b->recs.size();
    b->passes.push_back(pass);
  }
  // size check before anything is allocated for the program
  double launches = (double)b->recs.size();
  for (const auto &pass : b->passes)
    for (const auto &it : pass.items) launches += 8 + (b->ctx[it.ctx]->coverage ? 1 : 0);
  if (launches > kMaxProgramLaunches)
    return bfail(b, TOMO_ERR_ARG,
                 "launch program of " + std::to_string((long long)launches) + " launches per gradient is above the " +
                     std::to_string((long long)kMaxProgramLaunches) + " accepted; split the contexts into smaller batches");
  for (tomo_ctx *c : b->ctx) b->sig.push_back(context_signature(c));
  // device tables
  BCHK(b, hipMalloc((void **)&b->d_fwd, sizeof(FwdArgs<float>) * std::max<size_t>(b->hfwd.size(), 1)));
  BCHK(b, hipMalloc((void **)&b->d_adj, sizeof(AdjArgs) * std::max<size_t>(b->hadj.size(), 1)));
  BCHK(b, hipMalloc((void **)&b->d_start, sizeof(int) * std::max<size_t>(b->hstart.size(), 1)));
  BCHK(b, hipMemcpy(b->d_fwd, b->hfwd.data(), sizeof(FwdArgs<float>) * b->hfwd.size(), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_adj, b->hadj.data(), sizeof(AdjArgs) * b->hadj.size(), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->d_start, b->hstart.data(), sizeof(int) * b->hstart.size(), hipMemcpyHostToDevice));
  for (const auto &r : b->recs) {
    b->cells_per_call += r.cells;
    b->launches_per_call += 1;
  }
  return TOMO_OK;
}

int run_rec(tomo_batch *b, const Rec &r) {
  const int *sm = b->d_start + r.start0;
  const int *sr = sm + (r.n + 1);
  if (r.kind == 0) {
    BCHK(b, launch_forward_batch(b->d_fwd + r.item0, sm, sr, r.n, r.blocks_main, r.blocks_ring, r.vec, r.rows,
                                 r.depth, b->stream));
  } else {
    BCHK(b, launch_adjoint_batch(b->d_adj + r.item0, sm, sr, r.n, r.blocks_main, r.blocks_ring, r.cov, r.vec,
                                 r.rows, r.depth, b->stream));
  }
  return TOMO_OK;
}

// the whole gradient of every problem on the batch stream
int run_all(tomo_batch *b) {
  for (const auto &pass : b->passes) {
    for (const auto &it : pass.items) {
      tomo_ctx *c = b->ctx[it.ctx];
      const int P = b->P[it.ctx];
      const size_t tr = (size_t)c->nt * c->nr;
      BCHK(b, launch_fill_f32(c->d_rr + (size_t)it.p0 * tr, (size_t)it.np * tr, 0.f, b->stream));
      BCHK(b, launch_fill_f32(level_pass1(c, P, 0), (size_t)it.np * c->n2, 0.f, b->stream));
      BCHK(b, launch_fill_f32(level_pass1(c, P, 1), (size_t)it.np * c->n2, 0.f, b->stream));
    }
    for (size_t r = pass.fwd_begin; r < pass.fwd_end; ++r) {
      int rc = run_rec(b, b->recs[r]);
      if (rc) return rc;
    }
    for (const auto &it : pass.items) {
      tomo_ctx *c = b->ctx[it.ctx];
      const int P = b->P[it.ctx];
      const size_t tr = (size_t)c->nt * c->nr;
      const size_t lvl = (size_t)P * c->n2;
      BCHK(b, launch_misfit(c->d_rr + (size_t)it.p0 * tr, (long long)tr, it.np, c->d_misfit + it.p0, b->stream));
      BCHK(b, launch_fill_f32(c->d_grad + (size_t)it.p0 * c->n2, (size_t)it.np * c->n2, 0.f, b->stream));
      if (c->coverage)
        BCHK(b, launch_fill_f32(c->d_cov + (size_t)it.p0 * c->n2, (size_t)it.np * c->n2, 0.f, b->stream));
      BCHK(b, launch_fill_f32(c->d_z + (size_t)((c->nt - 1) % 3) * lvl, (size_t)it.np * c->n2, 0.f, b->stream));
      BCHK(b, launch_fill_f32(c->d_z + (size_t)(c->nt % 3) * lvl, (size_t)it.np * c->n2, 0.f, b->stream));
    }
    for (size_t r = pass.seg_begin; r < pass.seg_end; ++r) {
      int rc = run_rec(b, b->recs[r]);
      if (rc) return rc;
    }
  }
  return TOMO_OK;
}

}  // namespace

extern "C" {

int tomo_batch_create(tomo_ctx *const *ctxs, int n, tomo_batch **out) {
  if (!ctxs || n <= 0 || !out) return TOMO_ERR_ARG;
  *out = nullptr;
  tomo_batch *b = new tomo_batch;
  b->device = ctxs[0]->device;
  int rc = TOMO_OK;
  for (int i = 0; i < n && !rc; ++i) {
    tomo_ctx *c = ctxs[i];
    if (!c) rc = bfail(b, TOMO_ERR_ARG, "null context");
    else if (c->device != b->device) rc = bfail(b, TOMO_ERR_ARG, "contexts of a batch share one device");
    else if (c->nshots <= 0 || c->nr <= 0) rc = bfail(b, TOMO_ERR_STATE, "context without shots or receivers");
    else if (!c->have_meas) rc = bfail(b, TOMO_ERR_STATE, "context without measured traces");
    else if (c->cfg.sponge_width <= 0) rc = bfail(b, TOMO_ERR_ARG, "batched contexts use the sponge preset");
    else if (c->ws) rc = bfail(b, TOMO_ERR_ARG, "batched contexts run together: they cannot share a workspace");
    else if (context_shape(c).vec < 2) rc = bfail(b, TOMO_ERR_ARG, "grid too small for the vectorised kernels");
    else b->ctx.push_back(c);
  }
  if (!rc && hipSetDevice(b->device) != hipSuccess) rc = bfail(b, TOMO_ERR_HIP, "hipSetDevice");
  if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    rc = bfail(b, TOMO_ERR_HIP, "hipStreamCreate");
  if (!rc && (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess))
    rc = bfail(b, TOMO_ERR_HIP, "hipEventCreate");
  if (!rc) {
    for (tomo_ctx *c : b->ctx) (void)hipStreamSynchronize(c->stream);
    rc = build(b);
  }
  if (rc) {
    tomo_host::set_create_error(b->err);
    tomo_batch_destroy(b);
    return rc;
  }
  *out = b;
  return TOMO_OK;
}

void tomo_batch_destroy(tomo_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->d_fwd) (void)hipFree(b->d_fwd);
  if (b->d_adj) (void)hipFree(b->d_adj);
  if (b->d_start) (void)hipFree(b->d_start);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *tomo_batch_last_error(const tomo_batch *b) { return b ? b->err.c_str() : ""; }

int tomo_batch_set_graph(tomo_batch *b, int on) {
  if (!b) return TOMO_ERR_ARG;
  b->use_graph = on != 0;
  return TOMO_OK;
}

int tomo_batch_gradient(tomo_batch *b, double *misfits) {
  if (!b) return TOMO_ERR_ARG;
  // the program holds the contexts' device pointers and shapes as they were
  // at creation: refuse to run once any of them changed
  for (size_t i = 0; i < b->ctx.size(); ++i)
    if (context_signature(b->ctx[i]) != b->sig[i])
      return bfail(b, TOMO_ERR_STATE, "context " + std::to_string(i) +
                                          " changed its buffers, geometry or shape since the batch was created");
  BCHK(b, hipSetDevice(b->device));
  for (tomo_ctx *c : b->ctx) BCHK(b, hipStreamSynchronize(c->stream));
  BCHK(b, hipEventRecord(b->ev0, b->stream));
  if (b->use_graph) {
    if (!b->exec) {
      hipGraph_t g = nullptr;
      BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = run_all(b);
      const hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : bfail(b, TOMO_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      }
      const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) return bfail(b, TOMO_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
  } else {
    int rc = run_all(b);
    if (rc) return rc;
  }
  BCHK(b, hipEventRecord(b->ev1, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float ms = 0.f;
  BCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->ms += ms;
  b->cells += b->cells_per_call;
  b->launches += b->launches_per_call;
  b->calls += 1;
  for (size_t i = 0; i < b->ctx.size(); ++i) {
    tomo_ctx *c = b->ctx[i];
    std::vector<double> mis(c->nshots);
    BCHK(b, hipMemcpy(mis.data(), c->d_misfit, sizeof(double) * c->nshots, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double v : mis) s += v;
    if (misfits) misfits[i] = s;
    std::fill(c->grad_valid.begin(), c->grad_valid.end(), 1);
  }
  return TOMO_OK;
}

int tomo_batch_timing(const tomo_batch *b, double *out4) {
  if (!b || !out4) return TOMO_ERR_ARG;
  out4[0] = b->ms;
  out4[1] = b->cells;
  out4[2] = b->launches;
  out4[3] = b->calls;
  return TOMO_OK;
}

}  // extern "C"
