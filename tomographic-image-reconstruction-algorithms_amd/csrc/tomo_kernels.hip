// tomo_kernels.hip -- gfx950 kernels of the wave-tomography engine.
//
// Stencil kernels: a 256-thread block covers a 64 x (4*RY) tile of one shot;
// lane x of a wavefront reads row-contiguous, 256-byte coalesced cells, and
// each lane walks a vertical strip of RY cells keeping the rows above /
// below in registers, so a cell's y-neighbours are never re-read.  Blocks
// of all shots of a pass are launched together (grid.y = shot).  Cells on
// the outer edge, and receiver cells in the adjoint, take a slow path with
// direct loads; they are O(n) of the O(n^2) cells.
//
// Reductions are FP64 with a fixed summation order (deterministic).
#include "tomo_kernels.h"

#include <hip/hip_runtime.h>

namespace tomo {

namespace {

__device__ __forceinline__ int edge_dist(int x, int y, int nx, int ny) {
  return min(min(x, y), min(nx - 1 - x, ny - 1 - y));
}

// The wavefront's row of the block.  Blocks are kBlockX = 64 lanes wide, one
// wavefront per threadIdx.y, so the value is the same in every lane; reading
// it through readfirstlane lets the compiler keep the strip bounds and the
// row loop in scalar registers.  The two-step adjoint uses it (it frees the
// VGPRs of a spill there); in the forward kernels it measured slower.
__device__ __forceinline__ int wave_row() { return __builtin_amdgcn_readfirstlane((int)threadIdx.y); }
static_assert(kBlockX == 64, "one wavefront per block row");

// ------------------------------------------------------- cell arithmetic
// Device code is compiled without automatic FMA contraction; every cell rule
// below is one fixed chain of fused multiply-adds, shared by all kernels that
// apply it (scalar, vectorised, two-step, ring, batched), so each of them
// rounds a cell exactly as the others do and their fields agree bit for bit.

__device__ __forceinline__ float fused(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fused(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Leapfrog interior update (propagation, sirt-mimo.cu:540-563, written with
// the per-cell weight pw = c0^2 dt^2 (1 + f) / h^2 and the sponge factors a,
// ab): u(k+1) = pw * sum4 + (2 a - 4 pw) u(k) - ab u(k-1), where sum4 is the
// sum of the four neighbours of u(k)
template <typename T>
__device__ __forceinline__ T leapfrog(T pw, T sum4, T u, T a, T ab, T m) {
  return fused(-ab, m, fused(fused(T(-4), pw, T(2) * a), u, pw * sum4));
}

// Adjoint interior update (backpropagation1, sirt-mimo.cu:686-710): the
// weighted neighbour sum sum_n pw_n z_n + (2 a - 4 pw) z(k+1) - ab z(k+2)
__device__ __forceinline__ float adjoint_rule(float pn, float zn, float ps, float zs, float pl, float zl, float pr,
                                              float zr, float pc, float zc, float a, float ab, float z2) {
  const float s = fused(pr, zr, fused(pl, zl, fused(ps, zs, pn * zn)));
  return fused(-ab, z2, fused(fused(-4.f, pc, 2.f * a), zc, s));
}

// 5-point Laplacian (laplace, sirt-mimo.cu:776-804)
__device__ __forceinline__ float lap5(float un, float us, float ul, float ur, float uc, float inv_h2) {
  return fused(-4.f, uc, ((un + us) + ul) + ur) * inv_h2;
}

// ----------------------------------------------------------------- forward

// One-way edge rule of propagation (sirt-mimo.cu:566-587): corners use the
// diagonal neighbour.
// u(k+1) at an edge cell from u(k) at the cell (u), one (ua) and two (ub)
// cells inward, and u(k-1) at the cell (m) and one cell inward (ma):
// (2 - 2r - r^2) u + 2r (1 + r) ua - r^2 ub + (2r - 1) m - 2r ma
template <typename T>
__device__ __forceinline__ T mur_edge(T r, T u, T ua, T ub, T m, T ma) {
  const T r2 = T(2) * r, rr = r * r;
  T v = (T(2) - r2 - rr) * u;
  v = fused(r2 * (T(1) + r), ua, v);
  v = fused(-rr, ub, v);
  v = fused(r2 - T(1), m, v);
  return fused(-r2, ma, v);
}

// the cells one (ida) and two (idb) steps inward of edge cell (x, y)
__device__ __forceinline__ void edge_inward(int nx, int ny, int x, int y, int &ida, int &idb) {
  const bool top = y == 0, bottom = y == ny - 1, left = x == 0, right = x == nx - 1;
  const int ja = top ? y + 1 : bottom ? y - 1 : y;
  const int jb = top ? y + 2 : bottom ? y - 2 : y;
  const int ia = left ? x + 1 : right ? x - 1 : x;
  const int ib = left ? x + 2 : right ? x - 2 : x;
  ida = ja * nx + ia;
  idb = jb * nx + ib;
}

// Offset of the coefficient / field image of global shot gs in a context that
// stacks several phantoms of one geometry along the shot axis (spp shots per
// phantom, images pstride elements apart); 0 for a single phantom (spp = 0).
__device__ __forceinline__ long long phantom_offset(int spp, long long pstride, int gs) {
  return spp > 0 ? (long long)(gs / spp) * pstride : 0;
}

template <typename T>
__device__ T forward_edge(const FwdArgs<T> &A, const float *__restrict__ f, const T *__restrict__ u0, const T *__restrict__ um,
                          int x, int y) {
  const int id = y * A.nx + x;
  const T r = A.c0dt_h * sqrt(T(1) + T(f[id]));
  int ida, idb;
  edge_inward(A.nx, A.ny, x, y, ida, idb);
  return mur_edge<T>(r, u0[id], u0[ida], u0[idb], um[id], um[ida]);
}

// Leapfrog step k -> k+1 for every cell of one tile of one shot
// (propagation, sirt-mimo.cu:511-593, and the receiver record of
// initial_signal / difference_signal, sirt-mimo.cu:617-684, fused).
// Each lane first issues every load of its strip (RY + 2 rows of u(k), the
// x-neighbours, u(k-1), the coefficients, the receiver map), then computes
// and stores: the loads of all rows are in flight together.
template <typename T, bool SPONGE, int RY>
__global__ __launch_bounds__(kBlockX *kBlockY) void forward_step_kernel(FwdArgs<T> A) {
  const int nx = A.nx, ny = A.ny;
  const int tile = blockIdx.x;
  const int x = (tile % A.tiles_x) * kBlockX + threadIdx.x;
  const int y0 = ((tile / A.tiles_x) * kBlockY + threadIdx.y) * RY;
  if (x >= nx || y0 >= ny) return;
  const int gs = A.shot0 + blockIdx.y;
  const long long so = (long long)blockIdx.y * A.sstride;
  const T *__restrict__ u0 = A.u0 + so;
  const T *__restrict__ um = A.um + so;
  T *__restrict__ up = A.up + so;
  const long long po = phantom_offset(A.spp, A.pstride, gs);
  const T *__restrict__ pw = A.pw + po;
  const bool xin = x > 0 && x < nx - 1;

  T c[RY + 2], l[RY], r[RY], p[RY], m[RY];
  int cm[RY];
#pragma unroll
  for (int i = 0; i < RY + 2; ++i) {
    const int y = y0 - 1 + i;
    c[i] = (y >= 0 && y < ny) ? u0[y * nx + x] : T(0);
  }
#pragma unroll
  for (int i = 0; i < RY; ++i) {
    const int y = y0 + i;
    const bool ok = y < ny;
    const int id = y * nx + x;
    l[i] = (ok && xin) ? u0[id - 1] : T(0);
    r[i] = (ok && xin) ? u0[id + 1] : T(0);
    p[i] = ok ? pw[id] : T(0);
    m[i] = ok ? um[id] : T(0);
    cm[i] = (ok && A.record) ? A.cellmap[id] : -1;
  }

  const int4 bb = A.bbox[gs];
  const bool src_col = (A.amp_h2 != T(0)) && x >= bb.x && x <= bb.y;
#pragma unroll
  for (int i = 0; i < RY; ++i) {
    const int y = y0 + i;
    if (y >= ny) break;
    const int id = y * nx + x;
    T val;
    if (xin && y > 0 && y < ny - 1) {
      const T sum = ((r[i] + l[i]) + c[i]) + c[i + 2];  // u(i+1) + u(i-1) + u(j-1) + u(j+1)
      T a = T(1), ab = T(1);
      if (SPONGE) {
        const int d = edge_dist(x, y, nx, ny);
        if (d < A.sponge_w) {
          a = A.sp_a[d];
          ab = A.sp_ab[d];
        }
      }
      val = leapfrog<T>(p[i], sum, c[i + 1], a, ab, m[i]);
      if (src_col && y >= bb.z && y <= bb.w) val = fused(p[i], T(A.amp_h2), val);
    } else {
      val = forward_edge(A, A.f + po, u0, um, x, y);
    }
    up[id] = val;
    if (cm[i] >= 0) {
      const long long o = ((long long)gs * A.nt + A.level) * A.nr + cm[i];
      A.rec[o] = A.meas ? A.meas[o] - float(val) : float(val);
    }
  }
}

// ----------------------------------------------------------------- adjoint

template <bool SPONGE>
__device__ __forceinline__ void sponge_factors(const AdjArgs &A, int x, int y, float &a, float &ab) {
  a = 1.f;
  ab = 1.f;
  if (SPONGE) {
    const int d = edge_dist(x, y, A.nx, A.ny);
    if (d < A.sponge_w) {
      a = A.sp_a[d];
      ab = A.sp_ab[d];
    }
  }
}

// Interior adjoint stencil at (x, y) (backpropagation1, sirt-mimo.cu:686-710,
// written with the per-cell weight pw = c0^2 dt^2 (1 + f) / h^2).
template <bool SPONGE>
__device__ float adjoint_interior(const AdjArgs &A, const float *__restrict__ pw, const float *__restrict__ z1,
                                  const float *__restrict__ z2, int x, int y) {
  const int nx = A.nx;
  const int id = y * nx + x;
  float a, ab;
  sponge_factors<SPONGE>(A, x, y, a, ab);
  return adjoint_rule(pw[id - nx], z1[id - nx], pw[id + nx], z1[id + nx], pw[id - 1], z1[id - 1], pw[id + 1],
                      z1[id + 1], pw[id], z1[id], a, ab, z2[id]);
}

// forward level with the corner cells replaced by the mean of their two edge
// neighbours (propagation_at_corners, sirt-mimo.cu:595-615)
__device__ __forceinline__ float u_corner_mean(const float *__restrict__ u, int nx, int ny, int x, int y) {
  const bool cx = (x == 0 || x == nx - 1), cy = (y == 0 || y == ny - 1);
  if (cx && cy) {
    const int xi = x == 0 ? 1 : nx - 2, yi = y == 0 ? 1 : ny - 2;
    return 0.5f * (u[yi * nx + x] + u[y * nx + xi]);
  }
  return u[y * nx + x];
}

// clamped 5-point Laplacian (laplace, sirt-mimo.cu:776-804) for a non-corner cell
__device__ float lap_clamped(const float *__restrict__ u, int nx, int ny, int x, int y, float inv_h2) {
  const int ya = y > 0 ? y - 1 : y, yb = y < ny - 1 ? y + 1 : y;
  const int xa = x > 0 ? x - 1 : x, xb = x < nx - 1 ? x + 1 : x;
  return lap5(u_corner_mean(u, nx, ny, x, ya), u_corner_mean(u, nx, ny, x, yb), u_corner_mean(u, nx, ny, xa, y),
              u_corner_mean(u, nx, ny, xb, y), u_corner_mean(u, nx, ny, x, y), inv_h2);
}

// Lu at any cell with clamped indices; a corner cell takes the mean of its
// two edge neighbours' values (laplace_corners, sirt-mimo.cu:806-831)
__device__ float edge_lu(const float *__restrict__ u, int nx, int ny, int x, int y, float inv_h2) {
  const bool ex = (x == 0 || x == nx - 1), ey = (y == 0 || y == ny - 1);
  if (ex && ey) {
    const int xi = x == 0 ? 1 : nx - 2, yi = y == 0 ? 1 : ny - 2;
    return (lap_clamped(u, nx, ny, xi, y, inv_h2) + lap_clamped(u, nx, ny, x, yi, inv_h2)) / 2.f;
  }
  return lap_clamped(u, nx, ny, x, y, inv_h2);
}

// Slow path: edge, corner and receiver cells.  Final z(k) follows
// backpropagation2 (sirt-mimo.cu:712-774): receiver = inward anchor + gain *
// residual, edge = copy of the inward cell, corner = mean of the two edge
// copies (= the diagonal interior cell); Lu corners are the mean of the edge
// neighbours' Lu (laplace_corners, sirt-mimo.cu:806-831).
template <bool SPONGE>
__device__ void adjoint_slow(const AdjArgs &A, const float *__restrict__ z1, const float *__restrict__ z2,
                             const float *__restrict__ u, int gs, int x, int y, int ri, float &zval,
                             float &lu, float &uval) {
  const int nx = A.nx, ny = A.ny;
  const float *__restrict__ pw = A.pw + phantom_offset(A.spp, A.pstride, gs);
  const bool ex = (x == 0 || x == nx - 1), ey = (y == 0 || y == ny - 1);
  const int xi = x == 0 ? 1 : (x == nx - 1 ? nx - 2 : x);
  const int yi = y == 0 ? 1 : (y == ny - 1 ? ny - 2 : y);
  if (ex || ey) {
    zval = adjoint_interior<SPONGE>(A, pw, z1, z2, xi, yi);
  } else {
    const int an = A.anchor[ri];
    zval = fused(A.rr[((long long)gs * A.nt + A.level) * A.nr + ri], A.gain,
                 adjoint_interior<SPONGE>(A, pw, z1, z2, an % nx, an / nx));
  }
  lu = edge_lu(u, nx, ny, x, y, A.inv_h2);
  uval = u_corner_mean(u, nx, ny, x, y);
}

// Adjoint step k+1, k+2 -> k with the correlation acc += z(k) * Lu(k)
// fused (backpropagation1/2 + laplace + update_differential,
// sirt-mimo-io.cu:342-353): the adjoint field is consumed while it is in
// registers, so z and Lu histories are never stored.  Same load-then-compute
// structure as the forward step.
template <bool SPONGE, bool COV, int RY>
__global__ __launch_bounds__(kBlockX *kBlockY) void adjoint_step_kernel(AdjArgs A) {
  const int nx = A.nx, ny = A.ny;
  const int tile = blockIdx.x;
  const int x = (tile % A.tiles_x) * kBlockX + threadIdx.x;
  const int y0 = ((tile / A.tiles_x) * kBlockY + threadIdx.y) * RY;
  if (x >= nx || y0 >= ny) return;
  const int gs = A.shot0 + blockIdx.y;
  const long long so = (long long)blockIdx.y * A.sstride;
  const long long n2 = (long long)nx * ny;
  const float *__restrict__ z1 = A.z1 + so;
  const float *__restrict__ z2 = A.z2 + so;
  const float *__restrict__ u = A.u + so;
  float *__restrict__ z0 = A.z0 + so;
  float *__restrict__ acc = A.acc + gs * n2;
  float *__restrict__ cov = COV ? A.cov + gs * n2 : nullptr;
  const float *__restrict__ pw = A.pw + phantom_offset(A.spp, A.pstride, gs);
  const bool xin = x > 0 && x < nx - 1;

  float zc[RY + 2], pc[RY + 2], uc[RY + 2];
  float zl[RY], zr[RY], pl[RY], pr[RY], ul[RY], ur[RY], zz[RY], ac[RY], cv[RY];
  int cm[RY];
#pragma unroll
  for (int i = 0; i < RY + 2; ++i) {
    const int y = y0 - 1 + i;
    const bool ok = y >= 0 && y < ny;
    const int id = y * nx + x;
    zc[i] = ok ? z1[id] : 0.f;
    pc[i] = ok ? pw[id] : 0.f;
    uc[i] = ok ? u[id] : 0.f;
  }
#pragma unroll
  for (int i = 0; i < RY; ++i) {
    const int y = y0 + i;
    const bool ok = y < ny;
    const bool in = ok && xin;
    const int id = y * nx + x;
    zl[i] = in ? z1[id - 1] : 0.f;
    zr[i] = in ? z1[id + 1] : 0.f;
    pl[i] = in ? pw[id - 1] : 0.f;
    pr[i] = in ? pw[id + 1] : 0.f;
    ul[i] = in ? u[id - 1] : 0.f;
    ur[i] = in ? u[id + 1] : 0.f;
    zz[i] = ok ? z2[id] : 0.f;
    ac[i] = ok ? acc[id] : 0.f;
    cv[i] = (ok && COV) ? cov[id] : 0.f;
    cm[i] = (in && y > 0 && y < ny - 1) ? A.cellmap[id] : -1;
  }

#pragma unroll
  for (int i = 0; i < RY; ++i) {
    const int y = y0 + i;
    if (y >= ny) break;
    const int id = y * nx + x;
    const bool interior = xin && y > 0 && y < ny - 1;
    float zval, lu, uval;
    if (interior && cm[i] < 0) {
      float a, ab;
      sponge_factors<SPONGE>(A, x, y, a, ab);
      zval = adjoint_rule(pc[i], zc[i], pc[i + 2], zc[i + 2], pl[i], zl[i], pr[i], zr[i], pc[i + 1], zc[i + 1], a,
                          ab, zz[i]);
      lu = lap5(uc[i], uc[i + 2], ul[i], ur[i], uc[i + 1], A.inv_h2);
      uval = uc[i + 1];
    } else {
      adjoint_slow<SPONGE>(A, z1, z2, u, gs, x, y, interior ? cm[i] : -1, zval, lu, uval);
    }
    z0[id] = zval;
    acc[id] = fused(zval, lu, ac[i]);
    if (COV) cov[id] = fused(uval, uval, cv[i]);
  }
}

// ------------------------------------------------------ vectorised stencils
// Same rules as the kernels above, with each lane owning VEC consecutive
// cells of a row (one VEC-wide load per row and array instead of three
// scalar loads): a wavefront covers 64 * VEC columns, the x-neighbours
// crossing a lane boundary move between lanes with DPP whole-wave shifts,
// and only the two tile-boundary lanes load their outer neighbour.
// Requires nx % VEC == 0.  All lanes stay active through the DPP shifts.

// value held by lane - 1 (lane 0 receives 0)
__device__ __forceinline__ float from_prev_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
// value held by lane + 1 (lane 63 receives 0)
__device__ __forceinline__ float from_next_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map the physical block id so that every XCD owns one contiguous range
// of logical blocks: neighbouring tiles of a shot, which share halo rows,
// then run on the same XCD.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_logical_block() {
  const int b = blockIdx.x, nb = gridDim.x;
  const int per = nb / kXcds, rem = nb % kXcds;
  const int xcd = b % kXcds, loc = b / kXcds;
  return xcd * per + min(xcd, rem) + loc;
}

// the problem of this block in a batched launch; lb becomes the block index
// within that problem
template <bool BATCH, typename Args>
__device__ __forceinline__ Args pick_item(const Launch<Args> &L, int &lb) {
  if (!BATCH) return L.a;
  int lo = 0, hi = L.n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (L.start[mid] <= lb) lo = mid;
    else hi = mid;
  }
  lb -= L.start[lo];
  return L.items[lo];
}

// A row segment of VEC cells; hl / hr hold the cells just left / right of
// the wave's segment for the tile-edge lanes (load_row_halo).
template <int VEC>
struct Row {
  float e[VEC];
  float hl, hr;
};

// Sponge factors of a lane's VEC cells.  The factor depends on the cell's
// distance to the nearest edge, min(side distance, row distance): the side
// part is fixed per lane, so its factors sit in registers for the whole
// strip, and only rows within the sponge width of the top or bottom edge
// (one branch per row, the same for the whole wave) look the table up.
template <int VEC>
struct LaneSponge {
  float a[VEC], ab[VEC];
  int x0, nx;  // the distance of cell e to the side edges is recomputed where needed
};

template <int VEC>
__device__ __forceinline__ LaneSponge<VEC> lane_sponge(int x0, int nx, int w, const float *__restrict__ sp_a,
                                                       const float *__restrict__ sp_ab, bool ok) {
  LaneSponge<VEC> s;
  s.x0 = x0;
  s.nx = nx;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int x = x0 + e;
    const int dx = min(x, nx - 1 - x);
    s.a[e] = 1.f;
    s.ab[e] = 1.f;
    if (ok && dx < w) {
      s.a[e] = sp_a[dx];
      s.ab[e] = sp_ab[dx];
    }
  }
  return s;
}

template <int VEC>
__device__ __forceinline__ void sponge_at(const LaneSponge<VEC> &s, int e, int y, int ny, int w, const float *sp_a,
                                          const float *sp_ab, float &a, float &ab) {
  const int dy = min(y, ny - 1 - y);
  if (dy >= w) {
    a = s.a[e];
    ab = s.ab[e];
  } else {
    const int x = s.x0 + e;
    const int d = min(min(x, s.nx - 1 - x), dy);
    a = sp_a[d];
    ab = sp_ab[d];
  }
}

// Widest access that divides a lane segment of VEC cells and keeps its
// alignment: the whole segment for VEC = 2, 4, 8 (16- or 32-byte aligned), 2
// cells for VEC = 6 (segments start at 24-byte multiples), 1 cell for VEC = 3.
template <int VEC>
struct RowAccess {
  static constexpr bool kPow2 = (VEC & (VEC - 1)) == 0;
  static constexpr int kChunk = kPow2 ? VEC : (VEC % 2 == 0 ? 2 : 1);
};

template <int VEC>
__device__ __forceinline__ Row<VEC> load_row(const float *__restrict__ p, bool ok) {
  constexpr int C = RowAccess<VEC>::kChunk;
  typedef float vt __attribute__((ext_vector_type(C)));
  Row<VEC> r;
  r.hl = r.hr = 0.f;
  if (ok) {
#pragma unroll
    for (int c = 0; c < VEC / C; ++c) {
      if constexpr (C == 1) {
        r.e[c] = p[c];
      } else {
        const vt v = *reinterpret_cast<const vt *>(p + c * C);
#pragma unroll
        for (int e = 0; e < C; ++e) r.e[c * C + e] = v[e];
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < VEC; ++e) r.e[e] = 0.f;
  }
  return r;
}

// Row y (lane segment at x0) with the tile-edge lanes' outer neighbours, all
// issued together so that no load is waited for when the row is consumed.
template <int VEC>
__device__ __forceinline__ Row<VEC> load_row_halo(const float *__restrict__ base, int y, int x0, int nx, bool ok) {
  Row<VEC> r = load_row<VEC>(base + (ok ? y * nx + x0 : 0), ok);
  if (ok && threadIdx.x == 0 && x0 > 0) r.hl = base[y * nx + x0 - 1];
  if (ok && threadIdx.x == kBlockX - 1 && x0 + VEC < nx) r.hr = base[y * nx + x0 + VEC];
  return r;
}

// left / right neighbours of a lane's segment: DPP within the wave, the
// prefetched halo at the tile edges (0 outside the grid)
template <int VEC>
__device__ __forceinline__ void halo_ends(const Row<VEC> &r, float &lft, float &rgt) {
  const float l = from_prev_lane(r.e[VEC - 1]);
  const float g = from_next_lane(r.e[0]);
  lft = threadIdx.x == 0 ? r.hl : l;
  rgt = threadIdx.x == kBlockX - 1 ? r.hr : g;
}

// f at the side-edge cells of row y for the lanes holding them (else 0)
template <int VEC>
__device__ __forceinline__ float edge_f(const float *__restrict__ f, int y, int x0, int nx, bool ok) {
  if (!ok) return 0.f;
  if (x0 == 0) return f[y * nx];
  if (x0 + VEC == nx) return f[y * nx + nx - 1];
  return 0.f;
}

// Sponge tables in LDS (filled by the whole block before any wave returns)
constexpr int kSpongeLds = 256;

template <int VEC>
__device__ __forceinline__ void store_row(float *__restrict__ p, const Row<VEC> &r) {
  constexpr int C = RowAccess<VEC>::kChunk;
  typedef float vt __attribute__((ext_vector_type(C)));
#pragma unroll
  for (int c = 0; c < VEC / C; ++c) {
    if constexpr (C == 1) {
      p[c] = r.e[c];
    } else {
      vt v;
#pragma unroll
      for (int e = 0; e < C; ++e) v[e] = r.e[c * C + e];
      *reinterpret_cast<vt *>(p + c * C) = v;
    }
  }
}

template <int VEC>
__device__ __forceinline__ void load_row_int(const int *__restrict__ p, bool ok, int *out) {
  typedef int vt __attribute__((ext_vector_type(VEC)));
  if (ok) {
    const vt v = *reinterpret_cast<const vt *>(p);
#pragma unroll
    for (int e = 0; e < VEC; ++e) out[e] = v[e];
  } else {
#pragma unroll
    for (int e = 0; e < VEC; ++e) out[e] = -1;
  }
}

// x-neighbours of a lane's VEC cells: left of cell 0 and right of cell VEC-1
// (the rest are in the lane's own row).  Every lane must execute this.
template <int VEC>
__device__ __forceinline__ void row_ends(const Row<VEC> &r, const float *__restrict__ row_base, int x0, int nx,
                                         bool ok, float &lft, float &rgt) {
  lft = from_prev_lane(r.e[VEC - 1]);
  rgt = from_next_lane(r.e[0]);
  if (threadIdx.x == 0) lft = (ok && x0 > 0) ? row_base[x0 - 1] : 0.f;
  if (threadIdx.x == kBlockX - 1) rgt = (ok && x0 + VEC < nx) ? row_base[x0 + VEC] : 0.f;
}

// Forward step, vectorised: each lane walks a strip of RY rows as a sliding
// window (rows above / at / below in registers) and issues the loads of the
// next row before computing the current one, so a wavefront streams every
// input row once (halo overhead (RY + 2) / RY) with few live registers.
// Rows 1..ny-2 including their two side-edge cells (whose one-way rule needs
// only cells of the same row); the top and bottom rows and the receiver
// record are done by forward_ring_kernel right after, so the hot loop has no
// cell map.  Needs nx >= 4 VEC.

template <bool SPONGE, int VEC, int RY, int PD, bool BATCH>
__global__ __launch_bounds__(kBlockX *kBlockY) void forward_vec_kernel(Launch<FwdArgs<float>> L) {
  int lb = xcd_logical_block();
  const FwdArgs<float> A = pick_item<BATCH>(L, lb);
  __shared__ float s_spa[kSpongeLds], s_spab[kSpongeLds];
  if (SPONGE) {
    for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY) {
      s_spa[i] = A.sp_a[i];
      s_spab[i] = A.sp_ab[i];
    }
    __syncthreads();
  }
  const int nx = A.nx, ny = A.ny;
  const int tile = lb % A.tiles_per_shot, shot = lb / A.tiles_per_shot;
  const int x0 = (tile % A.tiles_x) * (kBlockX * VEC) + threadIdx.x * VEC;
  const int y0 = ((tile / A.tiles_x) * kBlockY + threadIdx.y) * RY;
  if (y0 >= ny) return;  // whole wavefront
  const bool xok = x0 < nx;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const float *__restrict__ u0 = A.u0 + so;
  const float *__restrict__ um = A.um + so;
  float *__restrict__ up = A.up + so;
  const long long po = phantom_offset(A.spp, A.pstride, gs);
  const float *__restrict__ pw = A.pw + po;
  const float *__restrict__ fld = A.f + po;
  const int4 bb = A.bbox[gs];
  const bool src = A.amp_h2 != 0.f;
  const int yend = min(y0 + RY, ny);  // rows of this strip: [y0, yend)
  const LaneSponge<VEC> ls = lane_sponge<VEC>(x0, nx, SPONGE ? A.sponge_w : 0, s_spa, s_spab, xok);

  // u(k) rows y0-1 .. yend are needed; u(k-1) and the coefficients rows y0 .. yend-1
  auto uok = [&](int y) { return xok && y >= y0 - 1 && y <= yend && y >= 0 && y < ny; };
  auto sok = [&](int y) { return xok && y >= y0 && y < yend; };
  auto at = [&](int y, bool ok) { return ok ? y * nx + x0 : 0; };

  // queue of the next PD rows: qu[j] = u(k) row y+1+j (with the tile-edge
  // halo), qm/qp = u(k-1) / pw row y+j, qf = f at the row's side edges
  Row<VEC> cu = load_row<VEC>(u0 + at(y0 - 1, uok(y0 - 1)), uok(y0 - 1));
  Row<VEC> cc = load_row_halo<VEC>(u0, y0, x0, nx, uok(y0));
  Row<VEC> qu[PD], qm[PD], qp[PD];
  float qf[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) {
    qu[j] = load_row_halo<VEC>(u0, y0 + 1 + j, x0, nx, uok(y0 + 1 + j));
    qm[j] = load_row<VEC>(um + at(y0 + j, sok(y0 + j)), sok(y0 + j));
    qp[j] = load_row<VEC>(pw + at(y0 + j, sok(y0 + j)), sok(y0 + j));
    qf[j] = edge_f<VEC>(fld, y0 + j, x0, nx, sok(y0 + j));
  }

#pragma unroll 1
  for (int base = y0; base < yend; base += PD) {
    // loads of the next PD rows are in flight while these PD rows compute
    Row<VEC> nu[PD], nm[PD], np[PD];
    float nf[PD];
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int yu = base + PD + 1 + j, ys = base + PD + j;
      nu[j] = load_row_halo<VEC>(u0, yu, x0, nx, uok(yu));
      nm[j] = load_row<VEC>(um + at(ys, sok(ys)), sok(ys));
      np[j] = load_row<VEC>(pw + at(ys, sok(ys)), sok(ys));
      nf[j] = edge_f<VEC>(fld, ys, x0, nx, sok(ys));
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int y = base + j;
      if (y < yend) {  // wavefront-uniform
        float lft, rgt;
        halo_ends<VEC>(cc, lft, rgt);
        if (xok && y > 0 && y < ny - 1) {
          Row<VEC> out;
          const bool src_row = src && y >= bb.z && y <= bb.w;
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const int x = x0 + e;
            const float l = e == 0 ? lft : cc.e[e - 1];
            const float r = e == VEC - 1 ? rgt : cc.e[e + 1];
            const float sum = ((r + l) + cu.e[e]) + qu[j].e[e];
            const float pe = qp[j].e[e];
            float a = 1.f, ab = 1.f;
            if (SPONGE) sponge_at<VEC>(ls, e, y, ny, A.sponge_w, s_spa, s_spab, a, ab);
            float val = leapfrog<float>(pe, sum, cc.e[e], a, ab, qm[j].e[e]);
            if (src_row && x >= bb.x && x <= bb.y) val = fused(pe, A.amp_h2, val);
            out.e[e] = val;
          }
          // side edges of this row: every operand is in this lane's registers
          if (x0 == 0) {
            const float r = A.c0dt_h * sqrtf(1.f + qf[j]);
            out.e[0] = mur_edge<float>(r, cc.e[0], cc.e[1], VEC == 2 ? rgt : cc.e[2 % VEC], qm[j].e[0],
                                       qm[j].e[1]);
          }
          if (x0 + VEC == nx) {
            constexpr int E = VEC - 1;
            const float r = A.c0dt_h * sqrtf(1.f + qf[j]);
            out.e[E] = mur_edge<float>(r, cc.e[E], cc.e[E - 1], VEC == 2 ? lft : cc.e[(E + VEC - 2) % VEC],
                                       qm[j].e[E], qm[j].e[E - 1]);
          }
          store_row<VEC>(up + y * nx + x0, out);
        }
      }
      cu = cc;
      cc = qu[j];
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      qu[j] = nu[j];
      qm[j] = nm[j];
      qp[j] = np[j];
      qf[j] = nf[j];
    }
  }
}

// Two leapfrog steps k -> k+1 -> k+2 in one pass over a strip of RY rows
// (sponge preset; a wave holds full rows, nx == 64 VEC, so x-neighbours move
// by DPP only).  Each lane walks rows r = y0-1 .. yend: it forms u(k+1) at
// row r from the u(k) window (rows r-1, r, r+1), stores it on its own rows,
// and forms u(k+2) at row r-1 from the u(k+1) window (rows r-2, r-1, r) and
// u(k) row r-1.  u(k) and u(k-1) are read once per strip (plus two halo rows
// of u(k)), u(k+1) is kept in registers: 8 B of traffic per cell-update
// instead of 12.  Rows 0 and ny-1 of both levels are formed by the first and
// last strips (edge_row_step), and in recording passes the receiver records
// are written as the rows are formed.  The per-cell arithmetic is that of
// forward_vec_kernel and forward_ring_kernel.
template <int VEC>
__device__ __forceinline__ Row<VEC> step_row(const Fwd2Args &A, const LaneSponge<VEC> &ls, const float *s_spa,
                                             const float *s_spab, const Row<VEC> &up, const Row<VEC> &at,
                                             const Row<VEC> &dn, const Row<VEC> &prev, const Row<VEC> &p,
                                             float fedge, int y, int x0, float amp, const int4 &bb) {
  const int nx = A.nx, ny = A.ny;
  const float lft = from_prev_lane(at.e[VEC - 1]);
  const float rgt = from_next_lane(at.e[0]);
  const bool src_row = amp != 0.f && y >= bb.z && y <= bb.w;
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int x = x0 + e;
    const float l = e == 0 ? lft : at.e[e - 1];
    const float r = e == VEC - 1 ? rgt : at.e[e + 1];
    const float sum = ((r + l) + up.e[e]) + dn.e[e];
    const float pe = p.e[e];
    float a, ab;
    sponge_at<VEC>(ls, e, y, ny, A.sponge_w, s_spa, s_spab, a, ab);
    float val = leapfrog<float>(pe, sum, at.e[e], a, ab, prev.e[e]);
    if (src_row && x >= bb.x && x <= bb.y) val = fused(pe, amp, val);
    out.e[e] = val;
  }
  if (x0 == 0) {
    const float r = A.c0dt_h * sqrtf(1.f + fedge);
    out.e[0] = mur_edge<float>(r, at.e[0], at.e[1], VEC == 2 ? rgt : at.e[2 % VEC], prev.e[0], prev.e[1]);
  }
  if (x0 + VEC == nx) {
    constexpr int E = VEC - 1;
    const float r = A.c0dt_h * sqrtf(1.f + fedge);
    out.e[E] = mur_edge<float>(r, at.e[E], at.e[E - 1], VEC == 2 ? lft : at.e[(E + VEC - 2) % VEC], prev.e[E],
                               prev.e[E - 1]);
  }
  return out;
}

// One edge row (0 or ny-1) of the next level by the one-way edge rule
// (forward_edge, sirt-mimo.cu:566-587): E, I, I2 = the current level's edge
// row and the rows one and two inward, Em, Im = the previous level's edge row
// and the row one inward, fr = f on the edge row.  The corner cells use the
// diagonal neighbours (for 2-cell lanes one of them sits in the next or
// previous lane).
template <int VEC>
__device__ __forceinline__ Row<VEC> edge_row_step(const Row<VEC> &E, const Row<VEC> &I, const Row<VEC> &I2,
                                                  const Row<VEC> &Em, const Row<VEC> &Im, const Row<VEC> &fr,
                                                  float c0dt_h, int x0, int nx) {
  const float i2n = from_next_lane(I2.e[0]), i2p = from_prev_lane(I2.e[VEC - 1]);
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const float r = c0dt_h * sqrtf(1.f + fr.e[e]);
    out.e[e] = mur_edge<float>(r, E.e[e], I.e[e], I2.e[e], Em.e[e], Im.e[e]);
  }
  if (x0 == 0) {
    const float r = c0dt_h * sqrtf(1.f + fr.e[0]);
    out.e[0] = mur_edge<float>(r, E.e[0], I.e[1], VEC > 2 ? I2.e[2 % VEC] : i2n, Em.e[0], Im.e[1]);
  }
  if (x0 + VEC == nx) {
    constexpr int L = VEC - 1;
    const float r = c0dt_h * sqrtf(1.f + fr.e[L]);
    out.e[L] = mur_edge<float>(r, E.e[L], I.e[L - 1], VEC > 2 ? I2.e[(L + VEC - 2) % VEC] : i2p, Em.e[L],
                               Im.e[L - 1]);
  }
  return out;
}

// record of the lane's receiver (descriptor d = receiver * VEC + cell, or -1)
// at one level: the residual meas - u, or u without measured traces
template <int VEC>
__device__ __forceinline__ void record_lane(const Row<VEC> &u, int d, float meas_v, bool has_meas,
                                            float *__restrict__ rec_level) {
#pragma unroll
  for (int e = 0; e < VEC; ++e)  // one predicated store per cell: no dynamic register index
    if (d >= 0 && d % VEC == e) rec_level[d / VEC] = has_meas ? meas_v - u.e[e] : u.e[e];
}

template <int VEC, int RY, int PD, bool REC>
__global__ __launch_bounds__(kBlockX *kBlockY) void forward2_kernel(Fwd2Args A) {
  const int nx = A.nx, ny = A.ny;
  __shared__ float s_spa[kSpongeLds], s_spab[kSpongeLds];
  for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY) {
    s_spa[i] = A.sp_a[i];
    s_spab[i] = A.sp_ab[i];
  }
  __syncthreads();
  const int lb = xcd_logical_block();
  const int shot = lb / A.blocks_per_shot;
  const int y0 = ((lb % A.blocks_per_shot) * kBlockY + threadIdx.y) * RY;
  if (y0 >= ny) return;  // whole wavefront
  const int yend = min(y0 + RY, ny);
  const int x0 = threadIdx.x * VEC;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const float *__restrict__ u0 = A.u0 + so;
  const float *__restrict__ um = A.um + so;
  float *__restrict__ up1 = A.up1 + so;
  float *__restrict__ up2 = A.up2 + so;
  const long long po = phantom_offset(A.spp, A.pstride, gs);
  const float *__restrict__ pw = A.pw + po;
  const float *__restrict__ fld = A.f + po;
  const int4 bb = A.bbox[gs];
  const LaneSponge<VEC> ls = lane_sponge<VEC>(x0, nx, A.sponge_w, s_spa, s_spab, true);
  auto ld = [&](const float *__restrict__ base, int y) {
    const bool ok = y >= 0 && y < ny;
    return load_row<VEC>(base + (ok ? y * nx + x0 : 0), ok);
  };
  const int r0 = max(y0 - 1, 0), r1 = min(yend, ny - 1);
  auto own = [&](int y) { return y >= y0 && y < yend; };
  // receiver records: rows r0 .. r1 holding receivers, the lane's receiver
  // descriptor read two rows ahead and its measured values one row ahead
  unsigned long long rflags = 0;
  if (REC) {
#pragma unroll
    for (int i = 0; i < RY + 2; ++i)
      if (r0 + i <= r1 && A.row_has_recv[r0 + i]) rflags |= 1ull << i;
  }
  auto desc = [&](int y) {
    const bool on = REC && y >= r0 && y <= r1 && ((rflags >> (y - r0)) & 1ull);
    return on ? A.recv_seg[y * kBlockX + threadIdx.x] : -1;
  };
  const long long o1 = ((long long)gs * A.nt + A.level1) * A.nr;  // level k+1 of this shot; k+2 at + nr
  const bool has_meas = A.meas != nullptr;
  auto mv = [&](int d, long long o) { return (REC && d >= 0 && has_meas) ? A.meas[o + d / VEC] : 0.f; };
  int dr = desc(r0), dn = desc(r0 + 1), dq = -1;
  float m1r = mv(dr, o1), m2r = mv(dr, o1 + A.nr), m2q = 0.f;

  // queues of the next PD rows: qn[j] = u(k) row base+1+j; qm / qp / qf =
  // u(k-1), coefficients, side-edge f of row base+j
  Row<VEC> cu = ld(u0, r0 - 1), cc = ld(u0, r0);
  Row<VEC> qn[PD], qm[PD], qp[PD];
  float qf[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) {
    const int y = r0 + j;
    qn[j] = ld(u0, y <= r1 ? y + 1 : -1);
    qm[j] = ld(um, y <= r1 ? y : -1);
    qp[j] = ld(pw, y <= r1 ? y : -1);
    qf[j] = edge_f<VEC>(fld, y, x0, nx, y <= r1);
  }
  Row<VEC> pp = cc, v2 = cc, v1 = cc;
  float fp = 0.f;
#pragma unroll 1
  for (int base = r0; base <= r1; base += PD) {
    // the next PD rows' loads are in flight while these PD rows compute
    Row<VEC> nn[PD], nm[PD], np[PD];
    float nf[PD];
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      // rows read: u(k) r0-1 .. r1+1, u(k-1) and coefficients r0 .. r1
      const int y = base + PD + j;
      nn[j] = ld(u0, y <= r1 ? y + 1 : -1);
      nm[j] = ld(um, y <= r1 ? y : -1);
      np[j] = ld(pw, y <= r1 ? y : -1);
      nf[j] = edge_f<VEC>(fld, y, x0, nx, y <= r1);
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int r = base + j;
      if (r <= r1) {  // wavefront-uniform
        const int d2 = desc(r + 2);
        const Row<VEC> &cn = qn[j];
        Row<VEC> v0;
        // rows 0 and ny-1 of u(k+1), formed by the strips that use them from
        // the inputs (loaded here: once per kernel, in the first and last
        // strips only); the strip owning the row stores it
        if (r == 0) {
          v0 = edge_row_step<VEC>(ld(u0, 0), ld(u0, 1), ld(u0, 2), ld(um, 0), ld(um, 1), ld(fld, 0), A.c0dt_h, x0,
                                  nx);
          if (own(0)) store_row<VEC>(up1 + x0, v0);
        } else if (r == ny - 1) {
          v0 = edge_row_step<VEC>(ld(u0, ny - 1), ld(u0, ny - 2), ld(u0, ny - 3), ld(um, ny - 1), ld(um, ny - 2),
                                  ld(fld, ny - 1), A.c0dt_h, x0, nx);
          if (own(ny - 1)) store_row<VEC>(up1 + (ny - 1) * nx + x0, v0);
        } else {
          v0 = step_row<VEC>(A, ls, s_spa, s_spab, cu, cc, cn, qm[j], qp[j], qf[j], r, x0, A.amp1, bb);
          if (own(r)) {
            store_row<VEC>(up1 + r * nx + x0, v0);
            if (REC && A.rec1) record_lane<VEC>(v0, dr, m1r, has_meas, A.rec + o1);
          }
        }
        const int q = r - 1;
        if (q >= y0 && q >= 1 && q <= ny - 2) {
          const Row<VEC> w = step_row<VEC>(A, ls, s_spa, s_spab, v2, v1, v0, cu, pp, fp, q, x0, A.amp2, bb);
          store_row<VEC>(up2 + q * nx + x0, w);
          if (REC && A.rec2) record_lane<VEC>(w, dq, m2q, has_meas, A.rec + o1 + A.nr);
        }
        // rows 0 and ny-1 of u(k+2), by the strips owning rows 0 and ny-2
        if (r == 2 && own(0))
          store_row<VEC>(up2 + x0, edge_row_step<VEC>(v2, v1, v0, ld(u0, 0), cu, ld(fld, 0), A.c0dt_h, x0, nx));
        if (r == ny - 1 && own(ny - 2))
          store_row<VEC>(up2 + (ny - 1) * nx + x0,
                         edge_row_step<VEC>(v0, v1, v2, cc, cu, ld(fld, ny - 1), A.c0dt_h, x0, nx));
        // measured values of row r+1 (its descriptor was read a row ago)
        const float m1n = mv(dn, o1), m2n = mv(dn, o1 + A.nr);
        v2 = v1;
        v1 = v0;
        pp = qp[j];
        fp = qf[j];
        cu = cc;
        cc = cn;
        dq = dr;
        dr = dn;
        dn = d2;
        m2q = m2r;
        m1r = m1n;
        m2r = m2n;
      }
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      qn[j] = nn[j];
      qm[j] = nm[j];
      qp[j] = np[j];
      qf[j] = nf[j];
    }
  }
}

// The edge pass after forward2_kernel: rows 0 and ny-1 of level k+2 and,
// when the next launch is also a pair, of level k+3 (its one-way rule reads
// the cell's own level-k+2 value, formed here by the same thread, and
// interior cells already stored), then the receiver records of levels k+1
// and k+2.  Replaces three edge launches per pair with one.
__global__ __launch_bounds__(256) void forward2_edge_kernel(Fwd2EdgeArgs A) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= A.nring) return;
  const int shot = blockIdx.y, gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const int2 c = A.ring[i];
  const int nx = A.nx;
  if (c.y < 0) {
    const int x = c.x % nx, y = c.x / nx;
    const float r = A.c0dt_h * sqrtf(1.f + A.f[phantom_offset(A.spp, A.pstride, gs) + c.x]);
    int ida, idb;
    edge_inward(nx, A.ny, x, y, ida, idb);
    const float *__restrict__ u0 = A.u0 + so;
    const float *__restrict__ u1 = A.up1 + so;
    const float v2 = mur_edge<float>(r, u1[c.x], u1[ida], u1[idb], u0[c.x], u0[ida]);
    A.up2[so + c.x] = v2;
    if (A.up3) {
      const float *__restrict__ u2 = A.up2 + so;
      A.up3[so + c.x] = mur_edge<float>(r, v2, u2[ida], u2[idb], u1[c.x], u1[ida]);
    }
  } else {
    const long long o1 = ((long long)gs * A.nt + A.level1) * A.nr + c.y, o2 = o1 + A.nr;
    if (A.rec1) {
      const float v = A.up1[so + c.x];
      A.rec[o1] = A.meas ? A.meas[o1] - v : v;
    }
    if (A.rec2) {
      const float v = A.up2[so + c.x];
      A.rec[o2] = A.meas ? A.meas[o2] - v : v;
    }
  }
}

// Three leapfrog steps k -> k+1 -> k+2 -> k+3 in one pass over a strip:
// each lane walks rows r of its strip and forms level 1 (u(k+1)) at row r,
// level 2 at row r-1 and level 3 at row r-2, each from the three-row window
// of the level below kept in registers (level 0 = u(k), streamed).  Level 1
// is formed on two halo rows above and below the strip, level 2 on one.
// Edge rows: rows 0 and ny-1 of level 1 from the inputs; of level 2 one row
// later than its window position (the top row is patched into the window),
// of level 3 at the last iteration that still holds its inputs.  Per three
// cell-updates a recording pass reads u(k), u(k-1) and writes u(k+2),
// u(k+3): 16 B instead of 24; a recompute pass also writes u(k+1).  Needs
// ny a multiple of RY (a strip never starts one or two rows above the
// bottom edge, whose rows it could not form).  The per-cell arithmetic is
// that of forward_vec_kernel and forward_ring_kernel.
template <int VEC, int RY, bool REC>
__global__ __launch_bounds__(kBlockX *kBlockY) __attribute__((amdgpu_waves_per_eu(2))) void forward3_kernel(
    Fwd3Args A3) {
  const Fwd2Args &A = A3.b;
  const int nx = A.nx, ny = A.ny;
  __shared__ float s_spa[kSpongeLds], s_spab[kSpongeLds];
  // row ny-1 of level 2, formed one iteration before it is used: a lane-private
  // LDS slot instead of VEC registers held through the loop
  __shared__ float s_e2[kBlockY][VEC][kBlockX];
  for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY) {
    s_spa[i] = A.sp_a[i];
    s_spab[i] = A.sp_ab[i];
  }
  __syncthreads();
  const int lb = xcd_logical_block();
  const int shot = lb / A.blocks_per_shot;
  const int y0 = ((lb % A.blocks_per_shot) * kBlockY + threadIdx.y) * RY;
  if (y0 >= ny) return;  // whole wavefront
  const int yend = min(y0 + RY, ny);
  const int x0 = threadIdx.x * VEC;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const float *__restrict__ u0 = A.u0 + so;
  const float *__restrict__ um = A.um + so;
  float *__restrict__ up1 = A.up1 + so;
  float *__restrict__ up2 = A.up2 + so;
  float *__restrict__ up3 = A3.up3 + so;
  const long long po = phantom_offset(A.spp, A.pstride, gs);
  const float *__restrict__ pw = A.pw + po;
  const float *__restrict__ fld = A.f + po;
  const int4 bb = A.bbox[gs];
  const LaneSponge<VEC> ls = lane_sponge<VEC>(x0, nx, A.sponge_w, s_spa, s_spab, true);
  auto ld = [&](const float *__restrict__ base, int y) {
    const bool ok = y >= 0 && y < ny;
    return load_row<VEC>(base + (ok ? y * nx + x0 : 0), ok);
  };
  auto own = [&](int y) { return y >= y0 && y < yend; };
  const int q1lo = max(y0 - 2, 0), q1hi = min(yend + 1, ny - 1);  // rows of level 1
  const int q2lo = max(y0 - 1, 0), q2hi = min(yend, ny - 1);      // rows of level 2
  const int rmax = min(yend + 1, ny);                              // last iteration
  // receiver records: descriptors two rows ahead, measured values one ahead
  unsigned long long rflags = 0;
  if (REC) {
#pragma unroll
    for (int i = 0; i < RY + 4; ++i)
      if (q1lo + i <= q1hi && A.row_has_recv[q1lo + i]) rflags |= 1ull << i;
  }
  auto desc = [&](int y) {
    const bool on = REC && y >= q1lo && y <= q1hi && ((rflags >> (y - q1lo)) & 1ull);
    return on ? A.recv_seg[y * kBlockX + threadIdx.x] : -1;
  };
  const long long o1 = ((long long)gs * A.nt + A.level1) * A.nr;  // level k+1; k+2 at + nr, k+3 at + 2 nr
  const bool has_meas = A.meas != nullptr;
  auto mv = [&](int d, long long o) { return (REC && d >= 0 && has_meas) ? A.meas[o + d / VEC] : 0.f; };
  int d0 = desc(q1lo), dn = desc(q1lo + 1), d1 = -1, d2r = -1;  // rows r, r+1, r-1, r-2
  float m1 = mv(d0, o1), m2 = 0.f, m3 = 0.f;                    // level 1 at r, 2 at r-1, 3 at r-2

  // streams: u(k) rows r-1, r, r+1; u(k-1) row r; coefficients and side-edge
  // f of rows r, r-1, r-2; windows of levels 1 (rows r-2, r-1) and 2 (rows
  // r-3, r-2) before row r is formed
  const Row<VEC> zero = ld(u0, -1);
  Row<VEC> cu = ld(u0, q1lo - 1), cc = ld(u0, q1lo), cn = ld(u0, q1lo + 1), qm = ld(um, q1lo);
  Row<VEC> p0 = ld(pw, q1lo), p1 = zero, p2 = zero;
  float f0 = edge_f<VEC>(fld, q1lo, x0, nx, true), f1 = 0.f, f2 = 0.f;
  Row<VEC> w1a = zero, w1b = zero, w2a = zero, w2b = zero;
#pragma unroll 1
  for (int r = q1lo; r <= rmax; ++r) {
    // the next row's inputs are in flight while this row computes
    const Row<VEC> nn = ld(u0, r + 2 <= q1hi + 1 ? r + 2 : -1);
    const Row<VEC> nm = ld(um, r + 1 <= q1hi ? r + 1 : -1);
    const Row<VEC> np = ld(pw, r + 1 <= q1hi ? r + 1 : -1);
    const float nf = edge_f<VEC>(fld, r + 1, x0, nx, r + 1 <= q1hi);
    const int d2 = desc(r + 2);
    // level 1 at row r
    Row<VEC> v1 = zero;
    if (r <= q1hi) {
      if (r == 0) {
        v1 = edge_row_step<VEC>(ld(u0, 0), ld(u0, 1), ld(u0, 2), ld(um, 0), ld(um, 1), ld(fld, 0), A.c0dt_h, x0, nx);
      } else if (r == ny - 1) {
        v1 = edge_row_step<VEC>(ld(u0, ny - 1), ld(u0, ny - 2), ld(u0, ny - 3), ld(um, ny - 1), ld(um, ny - 2),
                                ld(fld, ny - 1), A.c0dt_h, x0, nx);
      } else {
        v1 = step_row<VEC>(A, ls, s_spa, s_spab, cu, cc, cn, qm, p0, f0, r, x0, A.amp1, bb);
        if (REC && A.rec1 && own(r)) record_lane<VEC>(v1, d0, m1, has_meas, A.rec + o1);
      }
      if (A3.store1 && own(r)) store_row<VEC>(up1 + r * nx + x0, v1);
    }
    // level 2 at row q = r-1
    const int q = r - 1;
    Row<VEC> v2 = zero;
    if (q >= q2lo && q <= q2hi) {
      if (q == ny - 1) {
#pragma unroll
        for (int e = 0; e < VEC; ++e) v2.e[e] = s_e2[threadIdx.y][e][threadIdx.x];
      } else if (q > 0) {
        v2 = step_row<VEC>(A, ls, s_spa, s_spab, w1a, w1b, v1, cu, p1, f1, q, x0, A.amp2, bb);
        if (own(q)) {
          store_row<VEC>(up2 + q * nx + x0, v2);
          if (REC && A.rec2) record_lane<VEC>(v2, d1, m2, has_meas, A.rec + o1 + A.nr);
        }
      }  // q == 0: formed at r = 2 (below) and patched into the window
    }
    // level 3 at row s = r-2
    const int s3 = r - 2;
    if (own(s3) && s3 > 0 && s3 < ny - 1) {
      const Row<VEC> v3 = step_row<VEC>(A, ls, s_spa, s_spab, w2a, w2b, v2, w1a, p2, f2, s3, x0, A3.amp3, bb);
      store_row<VEC>(up3 + s3 * nx + x0, v3);
      if (REC && A3.rec3) record_lane<VEC>(v3, d2r, m3, has_meas, A.rec + o1 + 2 * A.nr);
    }
    // edge rows of levels 2 and 3
    if (r == 2 && own(0)) {  // row 0 of level 2: level-1 rows 0, 1, 2 are w1a, w1b, v1
      const Row<VEC> e = edge_row_step<VEC>(w1a, w1b, v1, ld(u0, 0), cu, ld(fld, 0), A.c0dt_h, x0, nx);
      store_row<VEC>(up2 + x0, e);
      w2b = e;  // the window's row 0 (row r-2 after this iteration's shift)
    }
    if (r == 3 && own(0)) {  // row 0 of level 3: level-2 rows 0, 1, 2 are w2a, w2b, v2; level-1 row 1 is w1a
      const Row<VEC> l1top =
          edge_row_step<VEC>(ld(u0, 0), ld(u0, 1), ld(u0, 2), ld(um, 0), ld(um, 1), ld(fld, 0), A.c0dt_h, x0, nx);
      store_row<VEC>(up3 + x0, edge_row_step<VEC>(w2a, w2b, v2, l1top, w1a, ld(fld, 0), A.c0dt_h, x0, nx));
    }
    if (r == ny - 1 && q2hi == ny - 1) {  // row ny-1 of level 2: level-1 rows ny-1, ny-2, ny-3 are v1, w1b, w1a
      const Row<VEC> e = edge_row_step<VEC>(v1, w1b, w1a, cc, cu, ld(fld, ny - 1), A.c0dt_h, x0, nx);
#pragma unroll
      for (int i = 0; i < VEC; ++i) s_e2[threadIdx.y][i][threadIdx.x] = e.e[i];
      if (own(ny - 1)) store_row<VEC>(up2 + (ny - 1) * nx + x0, e);
    }
    if (r == ny && own(ny - 1)) {  // row ny-1 of level 3: level-2 rows ny-1, ny-2, ny-3 are v2, w2b, w2a
      store_row<VEC>(up3 + (ny - 1) * nx + x0,
                     edge_row_step<VEC>(v2, w2b, w2a, w1b, w1a, ld(fld, ny - 1), A.c0dt_h, x0, nx));
    }
    // measured values for the next iteration: level 1 at row r+1, level 2
    // at row r, level 3 at row r-1
    const float m1n = mv(dn, o1), m2n = mv(d0, o1 + A.nr), m3n = mv(d1, o1 + 2 * A.nr);
    w1a = w1b;
    w1b = v1;
    w2a = w2b;
    w2b = v2;
    cu = cc;
    cc = cn;
    cn = nn;
    qm = nm;
    p2 = p1;
    p1 = p0;
    p0 = np;
    f2 = f1;
    f1 = f0;
    f0 = nf;
    d2r = d1;
    d1 = d0;
    d0 = dn;
    dn = d2;
    m1 = m1n;
    m2 = m2n;
    m3 = m3n;
  }
}

// Edge cells and receiver record of one forward step, for the vectorised
// kernels (one thread per listed cell and shot; runs after forward_vec_kernel
// on the same stream).  cells[i] = {flat cell, receiver index or -1}.
template <bool SPONGE, bool BATCH>
__global__ __launch_bounds__(256) void forward_ring_kernel(Launch<FwdArgs<float>> L) {
  int lb = blockIdx.x, shot, i;
  const FwdArgs<float> A = pick_item<BATCH>(L, lb);
  if (BATCH) {
    const int chunks = (A.nring + 255) / 256;
    shot = lb / chunks;
    i = (lb % chunks) * 256 + threadIdx.x;
  } else {
    shot = blockIdx.y;
    i = blockIdx.x * blockDim.x + threadIdx.x;
  }
  if (i >= A.nring) return;
  const int2 c = A.ring[i];
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const int x = c.x % A.nx, y = c.x / A.nx;
  float val;
  if (c.y < 0) {
    val = forward_edge(A, A.f + phantom_offset(A.spp, A.pstride, gs), A.u0 + so, A.um + so, x, y);
    A.up[so + c.x] = val;
  } else {
    val = A.up[so + c.x];
  }
  if (A.record && c.y >= 0) {
    const long long o = ((long long)gs * A.nt + A.level) * A.nr + c.y;
    A.rec[o] = A.meas ? A.meas[o] - val : val;
  }
}

// Side-edge cells of an adjoint row y in 1..ny-2 for the lanes holding them
// (backpropagation2): z = the inward cell's value; Lu with clamped indices,
// the corner cells of rows 0 / ny-1 read as the mean of their edge
// neighbours (laplace, laplace_corners).  uu / uc / ud = u rows y-1, y, y+1.
template <int VEC>
__device__ __forceinline__ void adjoint_side_edges(const Row<VEC> &uu, const Row<VEC> &uc, const Row<VEC> &ud,
                                                   const Row<VEC> &acc, Row<VEC> &zo, Row<VEC> &ao, int x0, int nx,
                                                   int y, int ny, float inv_h2) {
  if (x0 == 0) {
    const float upv = y == 1 ? 0.5f * (uc.e[0] + uu.e[1]) : uu.e[0];
    const float dnv = y == ny - 2 ? 0.5f * (uc.e[0] + ud.e[1]) : ud.e[0];
    const float lu = lap5(upv, dnv, uc.e[0], uc.e[1], uc.e[0], inv_h2);
    zo.e[0] = zo.e[1];
    ao.e[0] = fused(zo.e[0], lu, acc.e[0]);
  }
  if (x0 + VEC == nx) {
    constexpr int E = VEC - 1;
    const float upv = y == 1 ? 0.5f * (uc.e[E] + uu.e[E - 1]) : uu.e[E];
    const float dnv = y == ny - 2 ? 0.5f * (uc.e[E] + ud.e[E - 1]) : ud.e[E];
    const float lu = lap5(upv, dnv, uc.e[E - 1], uc.e[E], uc.e[E], inv_h2);
    zo.e[E] = zo.e[E - 1];
    ao.e[E] = fused(zo.e[E], lu, acc.e[E]);
  }
}

// Adjoint step, vectorised, same sliding-window structure: z(k+1), the
// coefficients and u(k) stream through a three-row window; z(k+2) and the
// accumulators are read for the current row only.  Rows 1..ny-2 including
// their side-edge cells, receivers excepted (rows holding receivers are
// flagged per row; their receiver cells are skipped): adjoint_ring_kernel
// does the receivers and the top and bottom rows.
template <bool SPONGE, bool COV, int VEC, int RY, int PD, bool BATCH>
__global__ __launch_bounds__(kBlockX *kBlockY) void adjoint_vec_kernel(Launch<AdjArgs> L) {
  int lb = xcd_logical_block();
  const AdjArgs A = pick_item<BATCH>(L, lb);
  __shared__ float s_spa[kSpongeLds], s_spab[kSpongeLds];
  if (SPONGE) {
    for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY) {
      s_spa[i] = A.sp_a[i];
      s_spab[i] = A.sp_ab[i];
    }
    __syncthreads();
  }
  const int nx = A.nx, ny = A.ny;
  const int tile = lb % A.tiles_per_shot, shot = lb / A.tiles_per_shot;
  const int x0 = (tile % A.tiles_x) * (kBlockX * VEC) + threadIdx.x * VEC;
  const int y0 = ((tile / A.tiles_x) * kBlockY + threadIdx.y) * RY;
  if (y0 >= ny) return;  // whole wavefront
  const bool xok = x0 < nx;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const long long n2 = (long long)nx * ny;
  const float *__restrict__ z1 = A.z1 + so;
  const float *__restrict__ z2 = A.z2 + so;
  const float *__restrict__ u = A.u + so;
  float *__restrict__ z0 = A.z0 + so;
  float *__restrict__ acc = A.acc + gs * n2;
  float *__restrict__ cov = COV ? A.cov + gs * n2 : nullptr;
  const float *__restrict__ pw = A.pw + phantom_offset(A.spp, A.pstride, gs);
  const int yend = min(y0 + RY, ny);
  const LaneSponge<VEC> ls = lane_sponge<VEC>(x0, nx, SPONGE ? A.sponge_w : 0, s_spa, s_spab, xok);

  auto wok = [&](int y) { return xok && y >= y0 - 1 && y <= yend && y >= 0 && y < ny; };  // window rows
  auto sok = [&](int y) { return xok && y >= y0 && y < yend; };                            // strip rows
  auto at = [&](int y, bool ok) { return ok ? y * nx + x0 : 0; };

  // window rows (z(k+1), pw, u(k)) above / at the current row, queue of the
  // next PD rows below, and the strip-only rows (z(k+2), accumulators)
  Row<VEC> zu = load_row<VEC>(z1 + at(y0 - 1, wok(y0 - 1)), wok(y0 - 1));
  Row<VEC> pu = load_row<VEC>(pw + at(y0 - 1, wok(y0 - 1)), wok(y0 - 1));
  Row<VEC> uu = load_row<VEC>(u + at(y0 - 1, wok(y0 - 1)), wok(y0 - 1));
  Row<VEC> zc = load_row_halo<VEC>(z1, y0, x0, nx, wok(y0));
  Row<VEC> pc = load_row_halo<VEC>(pw, y0, x0, nx, wok(y0));
  Row<VEC> uc = load_row_halo<VEC>(u, y0, x0, nx, wok(y0));
  // receivers of a row (rare): the row's flag, then its cell map, read ahead
  // with the row so that nothing is waited for when the row is consumed
  unsigned rflags = 0;  // strip rows holding receivers, read once
#pragma unroll
  for (int i = 0; i < RY; ++i)
    if (y0 + i < ny && A.row_has_recv[y0 + i]) rflags |= 1u << i;
  auto recv_bits = [&](int y) {
    int bits = 0;
    if (sok(y) && ((rflags >> (y - y0)) & 1u)) {
      int cm[VEC];
      load_row_int<VEC>(A.cellmap + y * nx + x0, true, cm);
#pragma unroll
      for (int e = 0; e < VEC; ++e) bits |= (cm[e] >= 0) << e;
    }
    return bits;
  };
  Row<VEC> qz[PD], qp[PD], qu[PD], q2[PD], qa[PD], qc[PD];
  int qr[PD];
#pragma unroll
  for (int j = 0; j < PD; ++j) {
    const int yw = y0 + 1 + j, ys = y0 + j;
    qz[j] = load_row_halo<VEC>(z1, yw, x0, nx, wok(yw));
    qp[j] = load_row_halo<VEC>(pw, yw, x0, nx, wok(yw));
    qu[j] = load_row_halo<VEC>(u, yw, x0, nx, wok(yw));
    q2[j] = load_row<VEC>(z2 + at(ys, sok(ys)), sok(ys));
    qa[j] = load_row<VEC>(acc + at(ys, sok(ys)), sok(ys));
    if (COV) qc[j] = load_row<VEC>(cov + at(ys, sok(ys)), sok(ys));
    qr[j] = recv_bits(ys);
  }

#pragma unroll 1
  for (int base = y0; base < yend; base += PD) {
    Row<VEC> nz[PD], np[PD], nu[PD], n2r[PD], na[PD], nc[PD];
    int nr[PD];
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int yw = base + PD + 1 + j, ys = base + PD + j;
      nz[j] = load_row_halo<VEC>(z1, yw, x0, nx, wok(yw));
      np[j] = load_row_halo<VEC>(pw, yw, x0, nx, wok(yw));
      nu[j] = load_row_halo<VEC>(u, yw, x0, nx, wok(yw));
      n2r[j] = load_row<VEC>(z2 + at(ys, sok(ys)), sok(ys));
      na[j] = load_row<VEC>(acc + at(ys, sok(ys)), sok(ys));
      if (COV) nc[j] = load_row<VEC>(cov + at(ys, sok(ys)), sok(ys));
      nr[j] = recv_bits(ys);
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      const int y = base + j;
      if (y < yend) {  // wavefront-uniform
        float zl, zr, pl, pr, ul, ur;
        halo_ends<VEC>(zc, zl, zr);
        halo_ends<VEC>(pc, pl, pr);
        halo_ends<VEC>(uc, ul, ur);
        if (xok && y > 0 && y < ny - 1) {
          const int skip = qr[j];
          Row<VEC> zo, ao, co;
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const int x = x0 + e;
            float a = 1.f, ab = 1.f;
            if (SPONGE) sponge_at<VEC>(ls, e, y, ny, A.sponge_w, s_spa, s_spab, a, ab);
            const float zle = e == 0 ? zl : zc.e[e - 1];
            const float zre = e == VEC - 1 ? zr : zc.e[e + 1];
            const float ple = e == 0 ? pl : pc.e[e - 1];
            const float pre = e == VEC - 1 ? pr : pc.e[e + 1];
            const float ule = e == 0 ? ul : uc.e[e - 1];
            const float ure = e == VEC - 1 ? ur : uc.e[e + 1];
            const float zval = adjoint_rule(pu.e[e], zu.e[e], qp[j].e[e], qz[j].e[e], ple, zle, pre, zre, pc.e[e],
                                            zc.e[e], a, ab, q2[j].e[e]);
            const float lu = lap5(uu.e[e], qu[j].e[e], ule, ure, uc.e[e], A.inv_h2);
            zo.e[e] = zval;
            ao.e[e] = fused(zval, lu, qa[j].e[e]);
            if (COV) co.e[e] = fused(uc.e[e], uc.e[e], qc[j].e[e]);
          }
          adjoint_side_edges<VEC>(uu, uc, qu[j], qa[j], zo, ao, x0, nx, y, ny, A.inv_h2);
          const int id = y * nx + x0;
          if (!skip) {
            store_row<VEC>(z0 + id, zo);
            store_row<VEC>(acc + id, ao);
            if (COV) store_row<VEC>(cov + id, co);
          } else {
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              if (skip & (1 << e)) continue;
              z0[id + e] = zo.e[e];
              acc[id + e] = ao.e[e];
              if (COV) cov[id + e] = co.e[e];
            }
          }
        }
      }
      zu = zc; zc = qz[j];
      pu = pc; pc = qp[j];
      uu = uc; uc = qu[j];
    }
#pragma unroll
    for (int j = 0; j < PD; ++j) {
      qz[j] = nz[j];
      qp[j] = np[j];
      qu[j] = nu[j];
      q2[j] = n2r[j];
      qa[j] = na[j];
      if (COV) qc[j] = nc[j];
      qr[j] = nr[j];
    }
  }
}

// Edge, corner and receiver cells of one adjoint step for the vectorised
// kernels (the cells adjoint_vec_kernel skips); inputs are the previous
// levels only, so the order against adjoint_vec_kernel does not matter.
template <bool SPONGE, bool COV, bool BATCH>
__global__ __launch_bounds__(256) void adjoint_ring_kernel(Launch<AdjArgs> L) {
  int lb = blockIdx.x, shot, i;
  const AdjArgs A = pick_item<BATCH>(L, lb);
  if (BATCH) {
    const int chunks = (A.nring + 255) / 256;
    shot = lb / chunks;
    i = (lb % chunks) * 256 + threadIdx.x;
  } else {
    shot = blockIdx.y;
    i = blockIdx.x * blockDim.x + threadIdx.x;
  }
  if (i >= A.nring) return;
  const int2 c = A.ring[i];
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const long long n2 = (long long)A.nx * A.ny;
  const int x = c.x % A.nx, y = c.x / A.nx;
  float zval, lu, uval;
  adjoint_slow<SPONGE>(A, A.z1 + so, A.z2 + so, A.u + so, gs, x, y, c.y, zval, lu, uval);
  A.z0[so + c.x] = zval;
  A.acc[gs * n2 + c.x] = fused(zval, lu, A.acc[gs * n2 + c.x]);
  if (COV) A.cov[gs * n2 + c.x] = fused(uval, uval, A.cov[gs * n2 + c.x]);
}

// Sponge factors of cell (x, y) from an interleaved LDS table {a, ab}: the
// same table entries as lane_sponge / sponge_at, looked up per cell instead
// of held in registers (for kernels short of registers).
__device__ __forceinline__ void sponge_cell(int x, int nx, int y, int ny, int w, const float2 *tab, float &a,
                                            float &ab) {
  const int d = min(min(x, nx - 1 - x), min(y, ny - 1 - y));
  const float2 v = d < w ? tab[d] : make_float2(1.f, 1.f);
  a = v.x;
  ab = v.y;
}

// One adjoint row by the interior rule, for full-row waves (the x-neighbours
// of every lane come by DPP; the side-edge cells get the rule's value and are
// overwritten by the caller).  pn / zn, ps / zs = coefficient and z(k+1)
// rows above and below, pc / zc the row itself, z2 = z(k+2) at the row.
template <int VEC>
__device__ __forceinline__ Row<VEC> adjoint_row(const float2 *s_sp, int x0, int nx, const Row<VEC> &pn,
                                                const Row<VEC> &zn, const Row<VEC> &ps, const Row<VEC> &zs,
                                                const Row<VEC> &pc, const Row<VEC> &zc, const Row<VEC> &z2, int y,
                                                int ny, int w) {
  const float zl = from_prev_lane(zc.e[VEC - 1]), zr = from_next_lane(zc.e[0]);
  const float pl = from_prev_lane(pc.e[VEC - 1]), pr = from_next_lane(pc.e[0]);
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    float a, ab;
    sponge_cell(x0 + e, nx, y, ny, w, s_sp, a, ab);
    const float zle = e == 0 ? zl : zc.e[e - 1];
    const float zre = e == VEC - 1 ? zr : zc.e[e + 1];
    const float ple = e == 0 ? pl : pc.e[e - 1];
    const float pre = e == VEC - 1 ? pr : pc.e[e + 1];
    out.e[e] = adjoint_rule(pn.e[e], zn.e[e], ps.e[e], zs.e[e], ple, zle, pre, zre, pc.e[e], zc.e[e], a, ab, z2.e[e]);
  }
  return out;
}

// Correlation of one row: acc + z * Lu for every cell (interior Laplacian of
// the u rows above / at / below), then the side-edge cells' copies and
// clamped Laplacians.
template <int VEC>
__device__ __forceinline__ Row<VEC> correlate_row(const Row<VEC> &uu, const Row<VEC> &uc, const Row<VEC> &ud,
                                                  const Row<VEC> &acc, Row<VEC> &z, int x0, int nx, int y, int ny,
                                                  float inv_h2) {
  const float ul = from_prev_lane(uc.e[VEC - 1]), ur = from_next_lane(uc.e[0]);
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const float ule = e == 0 ? ul : uc.e[e - 1];
    const float ure = e == VEC - 1 ? ur : uc.e[e + 1];
    out.e[e] = fused(z.e[e], lap5(uu.e[e], ud.e[e], ule, ure, uc.e[e], inv_h2), acc.e[e]);
  }
  adjoint_side_edges<VEC>(uu, uc, ud, acc, z, out, x0, nx, y, ny, inv_h2);
  return out;
}

// Lu of the top (TOP) or bottom edge row from registers, for full-row waves:
// E = the edge row, I = the row one inward, I2 = two inward.  The same
// operations as edge_lu: clamped 5-point Laplacian whose corner neighbours
// read as the mean of their two edge neighbours, and at a corner cell the
// mean of the Laplacians at its two edge neighbours (laplace,
// laplace_corners, sirt-mimo.cu:776-831).
template <int VEC, bool TOP>
__device__ __forceinline__ Row<VEC> edge_row_lu(const Row<VEC> &E, const Row<VEC> &I, const Row<VEC> &I2, int x0,
                                                int nx, float inv_h2) {
  const float el = from_prev_lane(E.e[VEC - 1]), er = from_next_lane(E.e[0]);
  const float cm0 = 0.5f * (I.e[0] + E.e[1]);                  // corner mean at x = 0 (lane 0)
  const float cmr = 0.5f * (I.e[VEC - 1] + E.e[VEC - 2]);      // at x = nx-1 (last lane)
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) {
    const int x = x0 + e;
    float w = e == 0 ? el : E.e[e - 1];
    float ea = e == VEC - 1 ? er : E.e[e + 1];
    w = x - 1 == 0 ? cm0 : w;
    ea = x + 1 == nx - 1 ? cmr : ea;
    const float n = TOP ? E.e[e] : I.e[e], s = TOP ? I.e[e] : E.e[e];
    out.e[e] = lap5(n, s, w, ea, E.e[e], inv_h2);
  }
  if (x0 == 0) {
    const float b = TOP ? lap5(cm0, I2.e[0], I.e[0], I.e[1], I.e[0], inv_h2)
                        : lap5(I2.e[0], cm0, I.e[0], I.e[1], I.e[0], inv_h2);
    out.e[0] = (out.e[1] + b) / 2.f;
  }
  if (x0 + VEC == nx) {
    constexpr int L = VEC - 1;
    const float b = TOP ? lap5(cmr, I2.e[L], I.e[L - 1], I.e[L], I.e[L], inv_h2)
                        : lap5(I2.e[L], cmr, I.e[L - 1], I.e[L], I.e[L], inv_h2);
    out.e[L] = (out.e[L - 1] + b) / 2.f;
  }
  return out;
}

// Row 0 (TOP) or ny-1 of one adjoint level, from row 1 / ny-2 already formed
// (z, with its side copies) and the u rows E, I, I2 (edge row first): stores
// the copy and adds its correlation to the edge row of the gradient.  Run by
// the one strip that owns row 1 / ny-2; level k before level k-1, the order
// of two one-step launches.
template <int VEC, bool TOP>
__device__ __forceinline__ void adjoint_edge_corr(const Row<VEC> &z, const Row<VEC> &E, const Row<VEC> &I,
                                                  const Row<VEC> &I2, float *__restrict__ arow, int x0, int nx,
                                                  float inv_h2) {
  const Row<VEC> lu = edge_row_lu<VEC, TOP>(E, I, I2, x0, nx, inv_h2);
  const Row<VEC> a = load_row<VEC>(arow + x0, true);
  Row<VEC> out;
#pragma unroll
  for (int e = 0; e < VEC; ++e) out.e[e] = fused(z.e[e], lu.e[e], a.e[e]);
  store_row<VEC>(arow + x0, out);
}

template <int VEC, bool TOP>
__device__ __forceinline__ void adjoint_edge_row(const Row<VEC> &z, const Row<VEC> &E, const Row<VEC> &I,
                                                 const Row<VEC> &I2, float *__restrict__ zrow,
                                                 float *__restrict__ arow, int x0, int nx, float inv_h2) {
  store_row<VEC>(zrow + x0, z);
  adjoint_edge_corr<VEC, TOP>(z, E, I, I2, arow, x0, nx, inv_h2);
}

// additive injection at the lane's receiver cell (descriptor d = receiver *
// VEC + cell, or -1), as adjoint_slow does for a point receiver
template <int VEC>
__device__ __forceinline__ void inject(Row<VEC> &z, int d, float res, float gain) {
#pragma unroll
  for (int e = 0; e < VEC; ++e) z.e[e] = (d >= 0 && d % VEC == e) ? fused(res, gain, z.e[e]) : z.e[e];
}

template <int VEC>
__device__ __forceinline__ void side_copies(Row<VEC> &z, int x0, int nx) {
  if (x0 == 0) z.e[0] = z.e[1];
  if (x0 + VEC == nx) z.e[VEC - 1] = z.e[VEC - 2];
}

// One strip of a two-step adjoint launch.  Each lane walks the rows of its
// strip in direction D (+1: down, r = rs .. qb+1; -1: up, r = re .. qa-1):
// it forms z(k) at row r from the z(k+1) window (rows r-1, r, r+1) and
// z(k+2) row r, injects the residual at receivers, stores it and correlates
// it with Lu(k) on the strip's own rows; then it forms z(k-1) at row q = r-D
// from the z(k) window held in registers (rows q-1, q, q+1) and z(k+1) row
// q, and correlates it with Lu(k-1) into the same gradient row, read once
// and written once.  z(k) is formed on one halo row above and below the
// strip.  Row 0 of each level is the copy of row 1 and row ny-1 the copy of
// row ny-2 (backpropagation2), so the window needs nothing from outside the
// strip's rows and halo; the strip that owns row 1 (ny-2) also writes row 0
// (ny-1) and its correlation, level k before level k-1 in either direction.
// The windows hold rows behind (r-D) and ahead (r+D); every cell rule takes
// them as north / south rows in the same operand order whichever way the
// strip walks, so both directions give the same bits.  Per two cell-updates:
// reads z(k+1), z(k+2), u(k), u(k-1), the gradient, writes z(k), z(k-1), the
// gradient (16 B per cell-update instead of 24).  The arithmetic of every
// cell is that of adjoint_vec_kernel and adjoint_ring_kernel, so the fields
// and gradients agree with two one-step launches bit for bit.
template <int VEC, int RY, int D>
__device__ __forceinline__ void adjoint2_strip(const Adj2Args &A, const float2 *s_sp,
                                               float (&stash)[VEC][kBlockX], int shot, int y0) {
  const int nx = A.nx, ny = A.ny;
  const int yend = min(y0 + RY, ny);
  const int qa = max(y0, 1), qb = min(yend, ny - 1) - 1;  // own rows of the interior rule
  const int x0 = threadIdx.x * VEC;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const long long n2 = (long long)nx * ny;
  const float *__restrict__ z1 = A.z1 + so;
  const float *__restrict__ z2 = A.z2 + so;
  const float *__restrict__ u0 = A.u0 + so;
  const float *__restrict__ um = A.um + so;
  float *__restrict__ z0 = A.z0 + so;
  float *__restrict__ zm = A.zm + so;
  float *__restrict__ acc = A.acc + gs * n2;
  const float *__restrict__ pw = A.pw + phantom_offset(A.spp, A.pstride, gs);
  const float *__restrict__ rrk = A.rr + ((long long)gs * A.nt + A.level) * A.nr;  // residuals at level k
  const float *__restrict__ rrm = rrk - A.nr;                                        // and at level k-1
  auto ld = [&](const float *__restrict__ base, int y, bool ok) {
    ok = ok && y >= 0 && y < ny;
    return load_row<VEC>(base + (ok ? y * nx + x0 : 0), ok);
  };
  // z(k) by the rule on rows [rs, re]
  const int rs = max(qa - 1, 1), re = min(qb + 1, ny - 2);
  auto rule = [&](int y) { return y >= rs && y <= re; };
  auto own = [&](int y) { return y >= qa && y <= qb; };
  auto lrow = [&](int y) { return y >= qa - 1 && y <= qb + 1; };  // u rows of the Laplacians
  unsigned long long rflags = 0;                                     // rows rs .. re holding receivers
#pragma unroll
  for (int i = 0; i < RY + 2; ++i)
    if (rs + i <= re && A.row_has_recv[rs + i]) rflags |= 1ull << i;
  auto desc = [&](int y) {
    const bool on = rule(y) && ((rflags >> (y - rs)) & 1ull);
    return on ? A.recv_seg[y * kBlockX + threadIdx.x] : -1;
  };
  auto resid = [&](int d, const float *__restrict__ r) { return d >= 0 ? r[d / VEC] : 0.f; };
  // north / south of a row from the windows' rows behind (r-D) and ahead (r+D)
  auto north = [](const Row<VEC> &behind, const Row<VEC> &ahead) -> const Row<VEC> & {
    return D > 0 ? behind : ahead;
  };
  auto south = [](const Row<VEC> &behind, const Row<VEC> &ahead) -> const Row<VEC> & {
    return D > 0 ? ahead : behind;
  };

  const int r0 = D > 0 ? rs : re, rlast = D > 0 ? qb + 1 : qa - 1;
  // windows: z(k+1) rows r-D, r, r+D; coefficients r-2D .. r+D; u(k) rows
  // r-D .. r+D; u(k-1) rows r-2D .. r; z(k+2) and the gradient at row r;
  // z(k) rows r-2D, r-D before row r is formed
  Row<VEC> zb = ld(z1, r0 - D, true), zc = ld(z1, r0, true), za = ld(z1, r0 + D, true);
  Row<VEC> pp = ld(pw, r0 - 2 * D, false), pb = ld(pw, r0 - D, true), pc = ld(pw, r0, true),
           pa = ld(pw, r0 + D, true);
  Row<VEC> ub = ld(u0, r0 - D, lrow(r0 - D)), uc = ld(u0, r0, lrow(r0)), ua = ld(u0, r0 + D, lrow(r0 + D));
  Row<VEC> m2 = ld(um, r0 - 2 * D, false), m1 = ld(um, r0 - D, lrow(r0 - D)), m0 = ld(um, r0, lrow(r0));
  Row<VEC> zz = ld(z2, r0, true);
  Row<VEC> ac = ld(acc, r0, own(r0));
  Row<VEC> v2 = ld(acc, 0, false), v1 = v2;
  int dr = desc(r0), dn = desc(r0 + D), dq = -1;
  float rk = resid(dr, rrk), rkm = resid(dr, rrm), rq = 0.f;

#pragma unroll 1
  for (int r = r0; D > 0 ? r <= rlast : r >= rlast; r += D) {
    // the next row's inputs are in flight while this row computes
    const Row<VEC> za2 = ld(z1, r + 2 * D, rule(r + D));
    const Row<VEC> pa2 = ld(pw, r + 2 * D, rule(r + D));
    const Row<VEC> ua2 = ld(u0, r + 2 * D, lrow(r + 2 * D));
    const Row<VEC> mn = ld(um, r + D, lrow(r + D));
    const Row<VEC> zzn = ld(z2, r + D, rule(r + D));
    const Row<VEC> an = ld(acc, r + D, own(r + D));
    const int d2 = desc(r + 2 * D);
    // z(k) at row r
    Row<VEC> v0;
    if (rule(r)) {
      v0 = adjoint_row<VEC>(s_sp, x0, nx, north(pb, pa), north(zb, za), south(pb, pa), south(zb, za), pc, zc, zz, r,
                            ny, A.sponge_w);
      inject<VEC>(v0, dr, rk, A.gain);
      if (own(r)) {
        const Row<VEC> &un = north(ub, ua), &us = south(ub, ua);
        ac = correlate_row<VEC>(un, uc, us, ac, v0, x0, nx, r, ny, A.inv_h2);
        store_row<VEC>(z0 + r * nx + x0, v0);
        // rows 0 and ny-1 of z(k): copies of rows 1 and ny-2
        if (r == 1) adjoint_edge_row<VEC, true>(v0, un, uc, us, z0, acc, x0, nx, A.inv_h2);
        if (r == ny - 2)
          adjoint_edge_row<VEC, false>(v0, us, uc, un, z0 + (ny - 1) * nx, acc + (ny - 1) * nx, x0, nx, A.inv_h2);
      } else {
        side_copies<VEC>(v0, x0, nx);
      }
    } else {
      v0 = v1;  // row ny-1 (0): the copy of row ny-2 (1)
    }
    if (r == (D > 0 ? 1 : ny - 2)) v1 = v0;  // row 0 (ny-1) before the walk reaches it
    // z(k-1) at row q = r-D
    const int q = r - D;
    if (own(q)) {
      Row<VEC> aq;
#pragma unroll
      for (int e = 0; e < VEC; ++e) aq.e[e] = stash[e][threadIdx.x];
      Row<VEC> w = adjoint_row<VEC>(s_sp, x0, nx, north(pp, pc), north(v2, v0), south(pp, pc), south(v2, v0), pb,
                                    v1, zb, q, ny, A.sponge_w);
      inject<VEC>(w, dq, rq, A.gain);
      const Row<VEC> &mn_ = north(m2, m0), &ms_ = south(m2, m0);
      const Row<VEC> ao = correlate_row<VEC>(mn_, m1, ms_, aq, w, x0, nx, q, ny, A.inv_h2);
      store_row<VEC>(zm + q * nx + x0, w);
      store_row<VEC>(acc + q * nx + x0, ao);
      // rows 0 and ny-1 of z(k-1)
      if (q == 1) adjoint_edge_row<VEC, true>(w, mn_, m1, ms_, zm, acc, x0, nx, A.inv_h2);
      if (q == ny - 2)
        adjoint_edge_row<VEC, false>(w, ms_, m1, mn_, zm + (ny - 1) * nx, acc + (ny - 1) * nx, x0, nx, A.inv_h2);
    }
    // residuals of row r+D (its descriptor was read a row ago)
    const float rkn = resid(dn, rrk), rkmn = resid(dn, rrm);
    if (own(r)) {
#pragma unroll
      for (int e = 0; e < VEC; ++e) stash[e][threadIdx.x] = ac.e[e];
    }
    ac = an;
    v2 = v1;
    v1 = v0;
    pp = pb;
    pb = pc;
    pc = pa;
    pa = pa2;
    zb = zc;
    zc = za;
    za = za2;
    ub = uc;
    uc = ua;
    ua = ua2;
    m2 = m1;
    m1 = m0;
    m0 = mn;
    zz = zzn;
    dq = dr;
    dr = dn;
    dn = d2;
    rq = rkm;
    rk = rkn;
    rkm = rkmn;
  }
}

// Two adjoint steps per launch over strips of RY rows.  Even strips walk
// down, odd strips up, so two neighbouring strips reach their shared halo
// rows at the same time (both at the start or both at the end of their
// walks) and the second read of a row is served by L2.
template <int VEC, int RY>
__global__ __launch_bounds__(kBlockX *kBlockY) __attribute__((amdgpu_waves_per_eu(2))) void adjoint2_kernel(
    Adj2Args A) {
  const int ny = A.ny;
  __shared__ float2 s_sp[kSpongeLds];
  // the gradient row after level k, held from one row to the next for level
  // k-1 in a lane-private LDS slot instead of registers ([wave][cell][lane]:
  // conflict-free)
  __shared__ float s_acc[kBlockY][VEC][kBlockX];
  for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY)
    s_sp[i] = make_float2(A.sp_a[i], A.sp_ab[i]);
  __syncthreads();
  const int lb = xcd_logical_block();
  const int shot = lb / A.blocks_per_shot;
  const int strip = (lb % A.blocks_per_shot) * kBlockY + wave_row();
  const int y0 = strip * RY;
  const int yend = min(y0 + RY, ny);
  if (y0 >= ny || max(y0, 1) > min(yend, ny - 1) - 1) return;  // whole wavefront
  if (strip & 1)
    adjoint2_strip<VEC, RY, -1>(A, s_sp, s_acc[threadIdx.y], shot, y0);
  else
    adjoint2_strip<VEC, RY, 1>(A, s_sp, s_acc[threadIdx.y], shot, y0);
}

// One strip of a three-step adjoint launch, the two-step strip extended by a
// level.  Each lane walks the rows of its strip in direction D and, at row r,
// forms z(k) at row r (from the z(k+1) window and z(k+2) row r), z(k-1) at
// row r-D (from the z(k) window and z(k+1) row r-D) and z(k-2) at row r-2D
// (from the z(k-1) window and z(k) row r-2D), injecting each level's
// residual at receivers.  z(k) is formed on two halo rows above and below
// the strip, z(k-1) on one; z(k-1) and z(k-2) are stored on the strip's own
// rows (the next launch starts from them), z(k) is not stored.  Each level
// is correlated with its Lu on the own rows: the gradient row, read once,
// waits in a lane-private LDS slot (one of three, by row) between the levels
// and is written once after level k-2.  Rows 0 and ny-1 of every level are
// copies of rows 1 and ny-2 with their correlations, added level k first.
// Per three cell-updates: reads z(k+1), z(k+2), u(k), u(k-1), u(k-2), the
// gradient, writes z(k-1), z(k-2), the gradient (12 B per cell-update).
// Every cell is the same fixed FMA chain as in the one-step kernels, so the
// fields and gradients agree with three one-step launches bit for bit.
template <int VEC, int RY, int D, int PD>
__device__ __forceinline__ void adjoint3_strip(const Adj3Args &A3, const float2 *s_sp,
                                               float (&stash)[3][VEC][kBlockX], int shot, int y0) {
  const Adj2Args &A = A3.b;
  const int nx = A.nx, ny = A.ny;
  const int yend = min(y0 + RY, ny);
  const int qa = max(y0, 1), qb = min(yend, ny - 1) - 1;  // own rows of the interior rule
  const int x0 = threadIdx.x * VEC;
  const int gs = A.shot0 + shot;
  const long long so = (long long)shot * A.sstride;
  const long long n2 = (long long)nx * ny;
  const float *__restrict__ z1 = A.z1 + so;
  const float *__restrict__ z2 = A.z2 + so;
  const float *__restrict__ u0 = A.u0 + so;
  const float *__restrict__ um = A.um + so;
  const float *__restrict__ un2 = A3.um2 + so;
  float *__restrict__ zm = A.zm + so;
  float *__restrict__ zm2 = A3.zm2 + so;
  float *__restrict__ acc = A.acc + gs * n2;
  const float *__restrict__ pw = A.pw + phantom_offset(A.spp, A.pstride, gs);
  const float *__restrict__ rrk = A.rr + ((long long)gs * A.nt + A.level) * A.nr;  // residuals at level k
  const float *__restrict__ rrm = rrk - A.nr;                                        // level k-1
  const float *__restrict__ rrm2 = rrm - A.nr;                                       // level k-2
  auto ld = [&](const float *__restrict__ base, int y, bool ok) {
    ok = ok && y >= 0 && y < ny;
    return load_row<VEC>(base + (ok ? y * nx + x0 : 0), ok);
  };
  // rule rows: z(k) on [rs0, re0], z(k-1) on [rs1, re1], z(k-2) on the own rows
  const int rs0 = max(qa - 2, 1), re0 = min(qb + 2, ny - 2);
  const int rs1 = max(qa - 1, 1), re1 = min(qb + 1, ny - 2);
  auto rule0 = [&](int y) { return y >= rs0 && y <= re0; };
  auto rule1 = [&](int y) { return y >= rs1 && y <= re1; };
  auto own = [&](int y) { return y >= qa && y <= qb; };
  auto lrow = [&](int y) { return y >= qa - 1 && y <= qb + 1; };  // u rows of the Laplacians
  // rows rs0 .. re0 holding receivers, one bit each (RY + 4 rows: two words
  // at RY = 64)
  unsigned long long rflags0 = 0, rflags1 = 0;
#pragma unroll
  for (int i = 0; i < RY + 4; ++i)
    if (rs0 + i <= re0 && A.row_has_recv[rs0 + i]) {
      if (i < 64) rflags0 |= 1ull << i;
      else rflags1 |= 1ull << (i - 64);
    }
  auto desc = [&](int y) {
    const int i = y - rs0;
    const bool on = rule0(y) && (((i < 64 ? rflags0 >> i : rflags1 >> (i - 64))) & 1ull);
    return on ? A.recv_seg[y * kBlockX + threadIdx.x] : -1;
  };
  auto resid = [&](int d, const float *__restrict__ r) { return d >= 0 ? r[d / VEC] : 0.f; };
  auto north = [](const Row<VEC> &behind, const Row<VEC> &ahead) -> const Row<VEC> & {
    return D > 0 ? behind : ahead;
  };
  auto south = [](const Row<VEC> &behind, const Row<VEC> &ahead) -> const Row<VEC> & {
    return D > 0 ? ahead : behind;
  };
  auto slot = [](int y) { return ((y % 3) + 3) % 3; };

  const int r0 = D > 0 ? rs0 : re0, rlast = D > 0 ? qb + 2 : qa - 2;
  // windows: z(k+1) rows r-D, r, r+D; coefficients r-3D .. r+D; u(k) rows
  // r-D .. r+D; u(k-1) rows r-2D .. r; u(k-2) rows r-3D .. r-D; z(k+2) and
  // the gradient at row r; z(k) rows r-2D, r-D and z(k-1) rows r-3D, r-2D
  // before this row's levels are formed
  Row<VEC> zb = ld(z1, r0 - D, true), zc = ld(z1, r0, true), za = ld(z1, r0 + D, true);
  Row<VEC> p3 = ld(pw, r0 - 3 * D, false), pp = ld(pw, r0 - 2 * D, false), pb = ld(pw, r0 - D, true),
           pc = ld(pw, r0, true), pa = ld(pw, r0 + D, true);
  Row<VEC> ub = ld(u0, r0 - D, lrow(r0 - D)), uc = ld(u0, r0, lrow(r0)), ua = ld(u0, r0 + D, lrow(r0 + D));
  Row<VEC> m2 = ld(um, r0 - 2 * D, false), m1 = ld(um, r0 - D, lrow(r0 - D)), m0 = ld(um, r0, lrow(r0));
  Row<VEC> t3 = ld(un2, r0 - 3 * D, false), t2 = ld(un2, r0 - 2 * D, lrow(r0 - 2 * D)),
           t1 = ld(un2, r0 - D, lrow(r0 - D));
  Row<VEC> zz = ld(z2, r0, true);
  Row<VEC> ac = ld(acc, r0, own(r0));
  Row<VEC> v2 = ld(acc, 0, false), v1 = v2, w2 = v2, w1 = v2;
  int dr = desc(r0), dn = desc(r0 + D), dq1 = -1, dq2 = -1;
  float rk = resid(dr, rrk), rkm = resid(dr, rrm), rkm2 = resid(dr, rrm2);
  float rq1 = 0.f, rq1b = 0.f, rq2 = 0.f;

  // read-ahead of PD rows: with PD = 2 the next row's inputs were issued a
  // row earlier (the *_q registers) and those of the row after are issued
  // while this row computes
  Row<VEC> za_q, pa_q, ua_q, mn_q, tn_q, zzn_q, an_q;
  if constexpr (PD == 2) {
    za_q = ld(z1, r0 + 2 * D, rule0(r0 + D));
    pa_q = ld(pw, r0 + 2 * D, rule0(r0 + D));
    ua_q = ld(u0, r0 + 2 * D, lrow(r0 + 2 * D));
    mn_q = ld(um, r0 + D, lrow(r0 + D));
    tn_q = ld(un2, r0, lrow(r0));
    zzn_q = ld(z2, r0 + D, rule0(r0 + D));
    an_q = ld(acc, r0 + D, own(r0 + D));
  }

#pragma unroll 1
  for (int r = r0; D > 0 ? r <= rlast : r >= rlast; r += D) {
    // the next row's inputs are in flight while this row computes
    Row<VEC> za2, pa2, ua2, mn, tn, zzn, an;
    if constexpr (PD == 2) {
      za2 = za_q;
      pa2 = pa_q;
      ua2 = ua_q;
      mn = mn_q;
      tn = tn_q;
      zzn = zzn_q;
      an = an_q;
      za_q = ld(z1, r + 3 * D, rule0(r + 2 * D));
      pa_q = ld(pw, r + 3 * D, rule0(r + 2 * D));
      ua_q = ld(u0, r + 3 * D, lrow(r + 3 * D));
      mn_q = ld(um, r + 2 * D, lrow(r + 2 * D));
      tn_q = ld(un2, r + D, lrow(r + D));
      zzn_q = ld(z2, r + 2 * D, rule0(r + 2 * D));
      an_q = ld(acc, r + 2 * D, own(r + 2 * D));
    } else {
      za2 = ld(z1, r + 2 * D, rule0(r + D));
      pa2 = ld(pw, r + 2 * D, rule0(r + D));
      ua2 = ld(u0, r + 2 * D, lrow(r + 2 * D));
      mn = ld(um, r + D, lrow(r + D));
      tn = ld(un2, r, lrow(r));
      zzn = ld(z2, r + D, rule0(r + D));
      an = ld(acc, r + D, own(r + D));
    }
    const int d2 = desc(r + 2 * D);
    // z(k) at row r
    Row<VEC> v0;
    if (rule0(r)) {
      v0 = adjoint_row<VEC>(s_sp, x0, nx, north(pb, pa), north(zb, za), south(pb, pa), south(zb, za), pc, zc, zz, r,
                            ny, A.sponge_w);
      inject<VEC>(v0, dr, rk, A.gain);
      if (own(r)) {
        const Row<VEC> &un = north(ub, ua), &us = south(ub, ua);
        const Row<VEC> ao = correlate_row<VEC>(un, uc, us, ac, v0, x0, nx, r, ny, A.inv_h2);
        if (r == 1) adjoint_edge_corr<VEC, true>(v0, un, uc, us, acc, x0, nx, A.inv_h2);
        if (r == ny - 2) adjoint_edge_corr<VEC, false>(v0, us, uc, un, acc + (ny - 1) * nx, x0, nx, A.inv_h2);
        const int sl = slot(r);
#pragma unroll
        for (int e = 0; e < VEC; ++e) stash[sl][e][threadIdx.x] = ao.e[e];
      } else {
        side_copies<VEC>(v0, x0, nx);
      }
    } else {
      v0 = v1;  // row ny-1 (0): the copy of row ny-2 (1)
    }
    if (r == (D > 0 ? 1 : ny - 2)) v1 = v0;  // row 0 (ny-1) before the walk reaches it
    // z(k-1) at row q1 = r-D
    const int q1 = r - D;
    Row<VEC> w0;
    if (rule1(q1)) {
      w0 = adjoint_row<VEC>(s_sp, x0, nx, north(pp, pc), north(v2, v0), south(pp, pc), south(v2, v0), pb, v1, zb, q1,
                            ny, A.sponge_w);
      inject<VEC>(w0, dq1, rq1, A.gain);
      if (own(q1)) {
        const int sl = slot(q1);
        Row<VEC> aq;
#pragma unroll
        for (int e = 0; e < VEC; ++e) aq.e[e] = stash[sl][e][threadIdx.x];
        const Row<VEC> &mn_ = north(m2, m0), &ms_ = south(m2, m0);
        const Row<VEC> ao = correlate_row<VEC>(mn_, m1, ms_, aq, w0, x0, nx, q1, ny, A.inv_h2);
        store_row<VEC>(zm + q1 * nx + x0, w0);
        if (q1 == 1) adjoint_edge_row<VEC, true>(w0, mn_, m1, ms_, zm, acc, x0, nx, A.inv_h2);
        if (q1 == ny - 2)
          adjoint_edge_row<VEC, false>(w0, ms_, m1, mn_, zm + (ny - 1) * nx, acc + (ny - 1) * nx, x0, nx, A.inv_h2);
#pragma unroll
        for (int e = 0; e < VEC; ++e) stash[sl][e][threadIdx.x] = ao.e[e];
      } else {
        side_copies<VEC>(w0, x0, nx);
      }
    } else {
      w0 = w1;  // row ny-1 (0) of z(k-1): the copy of row ny-2 (1)
    }
    if (q1 == (D > 0 ? 1 : ny - 2)) w1 = w0;  // row 0 (ny-1) of z(k-1) before the walk reaches it
    // z(k-2) at row q2 = r-2D
    const int q2 = r - 2 * D;
    if (own(q2)) {
      const int sl = slot(q2);
      Row<VEC> aq;
#pragma unroll
      for (int e = 0; e < VEC; ++e) aq.e[e] = stash[sl][e][threadIdx.x];
      Row<VEC> x = adjoint_row<VEC>(s_sp, x0, nx, north(p3, pb), north(w2, w0), south(p3, pb), south(w2, w0), pp, w1,
                                    v2, q2, ny, A.sponge_w);
      inject<VEC>(x, dq2, rq2, A.gain);
      const Row<VEC> &tn_ = north(t3, t1), &ts_ = south(t3, t1);
      const Row<VEC> ao = correlate_row<VEC>(tn_, t2, ts_, aq, x, x0, nx, q2, ny, A.inv_h2);
      store_row<VEC>(zm2 + q2 * nx + x0, x);
      store_row<VEC>(acc + q2 * nx + x0, ao);
      if (q2 == 1) adjoint_edge_row<VEC, true>(x, tn_, t2, ts_, zm2, acc, x0, nx, A.inv_h2);
      if (q2 == ny - 2)
        adjoint_edge_row<VEC, false>(x, ts_, t2, tn_, zm2 + (ny - 1) * nx, acc + (ny - 1) * nx, x0, nx, A.inv_h2);
    }
    // residuals of row r+D (its descriptor was read a row ago)
    const float rkn = resid(dn, rrk), rkmn = resid(dn, rrm), rkm2n = resid(dn, rrm2);
    ac = an;
    v2 = v1;
    v1 = v0;
    w2 = w1;
    w1 = w0;
    p3 = pp;
    pp = pb;
    pb = pc;
    pc = pa;
    pa = pa2;
    zb = zc;
    zc = za;
    za = za2;
    ub = uc;
    uc = ua;
    ua = ua2;
    m2 = m1;
    m1 = m0;
    m0 = mn;
    t3 = t2;
    t2 = t1;
    t1 = tn;
    zz = zzn;
    dq2 = dq1;
    dq1 = dr;
    dr = dn;
    dn = d2;
    rq2 = rq1b;
    rq1b = rkm2;
    rq1 = rkm;
    rk = rkn;
    rkm = rkmn;
    rkm2 = rkm2n;
  }
}

// Three adjoint steps per launch over strips of RY rows; even strips walk
// down, odd strips up (as the two-step kernel).
template <int VEC, int RY, int PD>
__global__ __launch_bounds__(kBlockX *kBlockY) __attribute__((amdgpu_waves_per_eu(VEC > 4 ? 1 : 2))) void adjoint3_kernel(
    Adj3Args A3) {
  const Adj2Args &A = A3.b;
  const int ny = A.ny;
  __shared__ float2 s_sp[kSpongeLds];
  // the gradient rows between their levels, in lane-private LDS slots
  // ([wave][slot][cell][lane]: conflict-free)
  __shared__ float s_acc[kBlockY][3][VEC][kBlockX];
  for (int i = threadIdx.y * kBlockX + threadIdx.x; i < A.sponge_w; i += kBlockX * kBlockY)
    s_sp[i] = make_float2(A.sp_a[i], A.sp_ab[i]);
  __syncthreads();
  const int lb = xcd_logical_block();
  const int shot = lb / A.blocks_per_shot;
  const int strip = (lb % A.blocks_per_shot) * kBlockY + wave_row();
  const int y0 = strip * RY;
  const int yend = min(y0 + RY, ny);
  if (y0 >= ny || max(y0, 1) > min(yend, ny - 1) - 1) return;  // whole wavefront
  if (strip & 1)
    adjoint3_strip<VEC, RY, -1, PD>(A3, s_sp, s_acc[threadIdx.y], shot, y0);
  else
    adjoint3_strip<VEC, RY, 1, PD>(A3, s_sp, s_acc[threadIdx.y], shot, y0);
}


// Coverage of the forward levels of one checkpoint segment, for passes whose
// adjoint runs two steps per launch (which carry no coverage):
// cov(x) = fma(v, v, cov(x)) over the segment's levels k = khi .. klo, newest
// first, with v = u(k) at the cell and, at the four corner cells, the mean
// of their two edge neighbours -- the order and the values of the one-step
// adjoint sweep (adjoint_vec_kernel, adjoint_ring_kernel), so the sums are
// the same bits.  Level k sits at lev01[k - k0] (k - k0 < 2, the checkpoint
// pair) or seg + (k - k0) * lstride; each thread owns 4 consecutive cells of
// one shot and keeps their sums in registers across the levels.
__global__ __launch_bounds__(256) void coverage_segment_kernel(CovSegArgs A) {
  const long long n4 = (long long)A.n2 / 4;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n4 * A.nshots) return;
  const int shot = (int)(t / n4);
  const int id = (int)(t % n4) * 4;
  const int nx = A.nx, ny = A.ny;
  const long long so = (long long)shot * A.sstride;
  float *__restrict__ cov = A.cov + (long long)(A.shot0 + shot) * A.n2 + id;
  float4 c = *reinterpret_cast<const float4 *>(cov);
  // corner cells among these four: x = 0 or nx-1 on rows 0 / ny-1
  const int y = id / nx, x = id % nx;
  const bool edge_row = y == 0 || y == ny - 1;
  const int yi = y == 0 ? 1 : ny - 2;
#pragma unroll 1
  for (int k = A.khi; k >= A.klo; --k) {
    const int q = k - A.k0;
    const float *__restrict__ u = (q < 2 ? A.lev01[q] : A.seg + (long long)q * A.lstride) + so;
    float4 v = *reinterpret_cast<const float4 *>(u + id);
    if (edge_row) {
      if (x == 0) v.x = 0.5f * (u[yi * nx] + u[y * nx + 1]);
      if (x + 3 == nx - 1) v.w = 0.5f * (u[yi * nx + nx - 1] + u[y * nx + nx - 2]);
    }
    c.x = fused(v.x, v.x, c.x);
    c.y = fused(v.y, v.y, c.y);
    c.z = fused(v.z, v.z, c.z);
    c.w = fused(v.w, v.w, c.w);
  }
  *reinterpret_cast<float4 *>(cov) = c;
}

// ------------------------------------------------------------- elementwise

template <typename T>
__global__ void coefficients_kernel(const float *__restrict__ f, T *__restrict__ pw, int nx, int ny,
                                    T c0dt_h, int sponge_w, const T *__restrict__ sp_a) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny) return;
  const T r = c0dt_h * sqrt(T(1) + T(f[id]));
  T a = T(1);
  if (sponge_w > 0) {
    const int d = edge_dist(id % nx, id / nx, nx, ny);
    if (d < sponge_w) a = sp_a[d];
  }
  pw[id] = a * (r * r);
}

template <typename T>
__global__ void fill_kernel(T *__restrict__ p, size_t n, T v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Fixed-order block sum of one double per thread (blockDim.x = 256).
__device__ double block_sum_256(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void misfit_kernel(const float *__restrict__ rr, long long per_shot,
                                                     double *__restrict__ out) {
  __shared__ double sh[256];
  const float *p = rr + blockIdx.x * per_shot;
  double acc = 0.0;
  for (long long i = threadIdx.x; i < per_shot; i += 256) {
    const double v = p[i];
    acc += v * v;
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

__device__ __forceinline__ bool in_mask(const Mask &m, int x, int y) {
  return x >= m.x0 && x < m.x1 && y >= m.y0 && y < m.y1;
}

// uniform SIRT (sirt-mimo.cu:858-877; sirt-mimo-io.cu:774-796)
__global__ void update_uniform_kernel(float *__restrict__ f, const float *__restrict__ G, int S, int nx,
                                      int ny, Mask m, double scale, double div, const float *__restrict__ pixw) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const long long n2 = (long long)nx * ny;
  if (id >= n2 || !in_mask(m, id % nx, id / nx)) return;
  double s = 0.0;
  for (int g = 0; g < S; ++g) s += (double)G[g * n2 + id];
  const double fv = f[id];
  const double d = scale * (s / (1.0 + fv)) / div;
  f[id] = (float)(fv + (pixw ? d * (double)pixw[id] : d));
}

// ART step (art-mimo-io.cu:381-400)
__global__ void update_group_kernel(float *__restrict__ f, const float *__restrict__ Gg, int nx, int ny,
                                    Mask m, double scale, const float *__restrict__ pixw) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nx * ny || !in_mask(m, id % nx, id / nx)) return;
  const double fv = f[id];
  const double d = scale * ((double)Gg[id] / (1.0 + fv));
  f[id] = (float)(fv + (pixw ? d * (double)pixw[id] : d));
}

// weighted SIRT, pass 1: mean increment image (weights_differential1's
// df_avg / Ng, sirt-mimo-norm.cu:56-74)
__global__ void weighted_mean_kernel(const float *__restrict__ f, const float *__restrict__ G, int M,
                                     long long n2, double *__restrict__ mean) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n2) return;
  const double q = 1.0 + (double)f[id];
  double s = 0.0;
  for (int m = 0; m < M; ++m) s += (double)G[m * n2 + id] / q;
  mean[id] = s / M;
}

// pass 2: partial squared distances of each shot's increment to the mean
__global__ __launch_bounds__(256) void weighted_norm_kernel(const float *__restrict__ f,
                                                            const float *__restrict__ G,
                                                            const double *__restrict__ mean, long long n2,
                                                            int chunk, double *__restrict__ part) {
  __shared__ double sh[256];
  const int m = blockIdx.y;
  const long long beg = (long long)blockIdx.x * chunk;
  const long long end = min(beg + chunk, n2);
  double acc = 0.0;
  for (long long id = beg + threadIdx.x; id < end; id += 256) {
    const double v = (double)G[m * n2 + id] / (1.0 + (double)f[id]) - mean[id];
    acc += v * v;
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) part[(long long)m * gridDim.x + blockIdx.x] = s;
}

// pass 3: d_m = sqrt(sum parts); w_m = d_m^-1 / sum d^-1, uniform 1/M when a
// d_m is zero (weights_differential2, sirt-mimo-norm.cu:76-97)
__global__ __launch_bounds__(256) void weighted_weights_kernel(const double *__restrict__ part, int M,
                                                               int nparts, double *__restrict__ w) {
  __shared__ int degenerate;
  if (threadIdx.x == 0) degenerate = 0;
  __syncthreads();
  for (int m = threadIdx.x; m < M; m += 256) {
    double d2 = 0.0;
    for (int b = 0; b < nparts; ++b) d2 += part[(long long)m * nparts + b];
    if (!(d2 > 0.0)) atomicOr(&degenerate, 1);
    w[m] = 1.0 / sqrt(d2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double W = 0.0;
    for (int m = 0; m < M; ++m) W += w[m];
    if (degenerate || !isfinite(W)) degenerate = 1;
    else degenerate = 0;
    w[M] = W;  // scratch slot
  }
  __syncthreads();
  const double W = w[M];
  const int deg = degenerate;
  __syncthreads();
  for (int m = threadIdx.x; m < M; m += 256) w[m] = deg ? 1.0 / M : w[m] / W;
}

// pass 4: f += scale * mask * sum_m w_m df_m (average_differential + update_field,
// sirt-mimo-norm.cu:100-118, 873-892)
__global__ void weighted_update_kernel(float *__restrict__ f, const float *__restrict__ G, int M, int nx,
                                       int ny, Mask msk, double scale, const double *__restrict__ w,
                                       const float *__restrict__ pixw) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const long long n2 = (long long)nx * ny;
  if (id >= n2 || !in_mask(msk, id % nx, id / nx)) return;
  const double q = 1.0 + (double)f[id];
  double s = 0.0;
  for (int m = 0; m < M; ++m) s += w[m] * ((double)G[m * n2 + id] / q);
  const double d = scale * s;
  f[id] = (float)((double)f[id] + (pixw ? d * (double)pixw[id] : d));
}

// Weighted average of intermediate fields (sirt-mimo-avg-weight.cu:80-203,
// 530-544) in FP64 sums of fixed order: fg_g = f + scale mask G_g / (1 + f)
// per group, formed where it is needed and never stored.
__device__ __forceinline__ double intermediate_field(const float *__restrict__ f, const float *__restrict__ G,
                                                     long long n2, int g, long long id, bool inside,
                                                     double scale) {
  const double fv = (double)f[id];
  return inside ? fv + scale * ((double)G[g * n2 + id] / (1.0 + fv)) : fv;
}

// pass 1: the mean of the intermediate fields over the groups
__global__ void avgw_mean_kernel(const float *__restrict__ f, const float *__restrict__ G, int M, int nx, int ny,
                                 Mask msk, double scale, double *__restrict__ mean) {
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n2 = (long long)nx * ny;
  if (id >= n2) return;
  const bool inside = in_mask(msk, (int)(id % nx), (int)(id / nx));
  double s = 0.0;
  for (int g = 0; g < M; ++g) s += intermediate_field(f, G, n2, g, id, inside, scale);
  mean[id] = s / M;
}

// pass 2: partial squared distances of each intermediate field to the mean,
// over the full grid
__global__ __launch_bounds__(256) void avgw_norm_kernel(const float *__restrict__ f, const float *__restrict__ G,
                                                        const double *__restrict__ mean, int nx, int ny, Mask msk,
                                                        double scale, int chunk, double *__restrict__ part) {
  __shared__ double sh[256];
  const int g = blockIdx.y;
  const long long n2 = (long long)nx * ny;
  const long long beg = (long long)blockIdx.x * chunk;
  const long long end = min(beg + chunk, n2);
  double acc = 0.0;
  for (long long id = beg + threadIdx.x; id < end; id += 256) {
    const bool inside = in_mask(msk, (int)(id % nx), (int)(id / nx));
    const double v = intermediate_field(f, G, n2, g, id, inside, scale) - mean[id];
    acc += v * v;
  }
  const double s = block_sum_256(acc, sh);
  if (threadIdx.x == 0) part[(long long)g * gridDim.x + blockIdx.x] = s;
}

// pass 3: w_g = sqrt(sum of its parts); w[M] = sum_g w_g (get_normalized_weights,
// get_total_weight)
__global__ void avgw_weights_kernel(const double *__restrict__ part, int M, int nparts, double *__restrict__ w) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double total = 0.0;
  for (int g = 0; g < M; ++g) {
    double d2 = 0.0;
    for (int b = 0; b < nparts; ++b) d2 += part[(long long)g * nparts + b];
    w[g] = sqrt(d2);
    total += w[g];
  }
  w[M] = total;
}

// pass 4: f += sum_g fg_g / w_g (update_weighted_average_field, update_field);
// every cell, the whole intermediate fields included
__global__ void avgw_update_kernel(float *__restrict__ f, const float *__restrict__ G, int M, int nx, int ny,
                                   Mask msk, double scale, const double *__restrict__ w) {
  const long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long n2 = (long long)nx * ny;
  if (id >= n2) return;
  const bool inside = in_mask(msk, (int)(id % nx), (int)(id / nx));
  double s = 0.0;
  for (int g = 0; g < M; ++g) s += intermediate_field(f, G, n2, g, id, inside, scale) / w[g];
  f[id] = (float)((double)f[id] + s);
}

// Update direction of one unit of relaxation, d = mask * update(s = 1), in
// FP64: uniform (w == nullptr: sum_g G_g / (1 + f) / div) or weighted
// (sum_m w_m G_m / (1 + f)); pixw optional
__global__ void direction_kernel(const float *__restrict__ f, const float *__restrict__ G, int M, int nx, int ny,
                                 Mask msk, double div, const double *__restrict__ w, const float *__restrict__ pixw,
                                 double *__restrict__ d) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const long long n2 = (long long)nx * ny;
  if (id >= n2) return;
  if (!in_mask(msk, id % nx, id / nx)) {
    d[id] = 0.0;
    return;
  }
  const double q = 1.0 + (double)f[id];
  double s = 0.0;
  if (w) {
    for (int m = 0; m < M; ++m) s += w[m] * ((double)G[m * n2 + id] / q);
  } else {
    for (int m = 0; m < M; ++m) s += (double)G[m * n2 + id];
    s = s / q / div;
  }
  d[id] = pixw ? s * (double)pixw[id] : s;
}

// out[0..3] = {sum (f - fo)^2, sum (f - fo) d, sum d^2, sum fo^2}
__global__ __launch_bounds__(256) void quadratic_kernel(const float *__restrict__ f, const float *__restrict__ fo,
                                                        const double *__restrict__ d, long long n,
                                                        double *__restrict__ out) {
  __shared__ double sh[256];
  double a = 0.0, b = 0.0, c = 0.0, e = 0.0;
  for (long long i = threadIdx.x; i < n; i += 256) {
    const double r = (double)f[i] - (double)fo[i];
    a += r * r;
    b += r * d[i];
    c += d[i] * d[i];
    e += (double)fo[i] * (double)fo[i];
  }
  const double sa = block_sum_256(a, sh);
  const double sb = block_sum_256(b, sh);
  const double sc = block_sum_256(c, sh);
  const double se = block_sum_256(e, sh);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb;
    out[2] = sc;
    out[3] = se;
  }
}

// Illumination (coverage) normalisation: stats = {mean of c over the mask,
// max of c}, c = sum of the S coverage images
__global__ __launch_bounds__(256) void coverage_stats_kernel(const float *__restrict__ cov, int S, int nx, int ny,
                                                             Mask m, double *__restrict__ stats) {
  __shared__ double sh[256];
  const long long n2 = (long long)nx * ny;
  double sum = 0.0, mx = 0.0;
  for (long long i = threadIdx.x; i < n2; i += 256) {
    double c = 0.0;
    for (int g = 0; g < S; ++g) c += (double)cov[g * n2 + i];
    if (in_mask(m, (int)(i % nx), (int)(i / nx))) sum += c;
    mx = fmax(mx, c);
  }
  const double ssum = block_sum_256(sum, sh);
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    stats[0] = ssum / fmax((double)(m.x1 - m.x0) * (double)(m.y1 - m.y0), 1.0);
    stats[1] = sh[0];
  }
}

// w(x) = cbar / (c(x) + eps cmax)
__global__ void pixel_weight_kernel(const float *__restrict__ cov, int S, long long n2, double eps,
                                    const double *__restrict__ stats, float *__restrict__ w) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double c = 0.0;
  for (int g = 0; g < S; ++g) c += (double)cov[g * n2 + i];
  const double den = c + eps * stats[1];
  w[i] = den > 0.0 ? (float)(stats[0] / den) : 0.f;
}

__global__ __launch_bounds__(256) void error_kernel(const float *__restrict__ f, const float *__restrict__ fo,
                                                    int n, double *__restrict__ out) {
  __shared__ double sh[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double d = (double)f[i] - (double)fo[i];
    a += d * d;
    b += (double)fo[i] * (double)fo[i];
  }
  const double sa = block_sum_256(a, sh);
  const double sb = block_sum_256(b, sh);
  if (threadIdx.x == 0) {
    out[0] = sa;
    out[1] = sb;
  }
}

inline dim3 stencil_grid(int nx, int ny, int ry, int nshots, int *tiles_x) {
  *tiles_x = (nx + kBlockX - 1) / kBlockX;
  const int tiles_y = (ny + kBlockY * ry - 1) / (kBlockY * ry);
  return dim3(*tiles_x * tiles_y, nshots, 1);
}

template <typename T, bool SPONGE>
hipError_t fwd_dispatch(FwdArgs<T> a, int nshots, int ry, hipStream_t st) {
  const dim3 block(kBlockX, kBlockY, 1);
  const dim3 grid = stencil_grid(a.nx, a.ny, ry, nshots, &a.tiles_x);
  switch (ry) {
    case 4: forward_step_kernel<T, SPONGE, 4><<<grid, block, 0, st>>>(a); break;
    case 16: forward_step_kernel<T, SPONGE, 16><<<grid, block, 0, st>>>(a); break;
    default: forward_step_kernel<T, SPONGE, 8><<<grid, block, 0, st>>>(a); break;
  }
  return hipGetLastError();
}

// 1-D grid of tiles x shots (the kernels remap it per XCD)
inline dim3 vec_grid(int nx, int ny, int vec, int ry, int nshots, int *tiles_x, int *tiles_per_shot) {
  *tiles_x = (nx + kBlockX * vec - 1) / (kBlockX * vec);
  const int tiles_y = (ny + kBlockY * ry - 1) / (kBlockY * ry);
  *tiles_per_shot = *tiles_x * tiles_y;
  return dim3(*tiles_per_shot * nshots, 1, 1);
}

template <bool SPONGE, int VEC, int PD, bool BATCH>
void fwd_vec_rows(const Launch<FwdArgs<float>> &L, dim3 grid, dim3 block, int ry, hipStream_t st) {
  switch (ry) {
    case 4: forward_vec_kernel<SPONGE, VEC, 4, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    case 8: forward_vec_kernel<SPONGE, VEC, 8, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    case 32: forward_vec_kernel<SPONGE, VEC, 32, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    default: forward_vec_kernel<SPONGE, VEC, 16, PD, BATCH><<<grid, block, 0, st>>>(L); break;
  }
}

template <bool SPONGE, int VEC, bool BATCH>
void fwd_vec_depth(const Launch<FwdArgs<float>> &L, dim3 grid, int ry, int pd, hipStream_t st) {
  const dim3 block(kBlockX, kBlockY, 1);
  switch (pd) {
    case 1: fwd_vec_rows<SPONGE, VEC, 1, BATCH>(L, grid, block, ry, st); break;
    case 4: fwd_vec_rows<SPONGE, VEC, 4, BATCH>(L, grid, block, ry, st); break;
    default: fwd_vec_rows<SPONGE, VEC, 2, BATCH>(L, grid, block, ry, st); break;
  }
}

template <bool SPONGE, int VEC>
hipError_t fwd_vec_dispatch(FwdArgs<float> a, int nshots, int ry, int pd, const int2 *ring, int nring,
                            hipStream_t st) {
  const dim3 grid = vec_grid(a.nx, a.ny, VEC, ry, nshots, &a.tiles_x, &a.tiles_per_shot);
  a.ring = ring;
  a.nring = nring;
  const Launch<FwdArgs<float>> L{a, nullptr, nullptr, 0};
  fwd_vec_depth<SPONGE, VEC, false>(L, grid, ry, pd, st);
  // edge cells and receiver record read this step's interior result: after
  if (nring > 0) forward_ring_kernel<SPONGE, false><<<dim3((nring + 255) / 256, nshots, 1), 256, 0, st>>>(L);
  return hipGetLastError();
}

template <bool SPONGE, bool COV>
hipError_t adj_dispatch(AdjArgs a, int nshots, int ry, hipStream_t st) {
  const dim3 block(kBlockX, kBlockY, 1);
  const dim3 grid = stencil_grid(a.nx, a.ny, ry, nshots, &a.tiles_x);
  switch (ry) {
    case 4: adjoint_step_kernel<SPONGE, COV, 4><<<grid, block, 0, st>>>(a); break;
    default: adjoint_step_kernel<SPONGE, COV, 8><<<grid, block, 0, st>>>(a); break;
  }
  return hipGetLastError();
}

template <bool SPONGE, bool COV, int VEC, int PD, bool BATCH>
void adj_vec_rows(const Launch<AdjArgs> &L, dim3 grid, dim3 block, int ry, hipStream_t st) {
  switch (ry) {
    case 4: adjoint_vec_kernel<SPONGE, COV, VEC, 4, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    case 8: adjoint_vec_kernel<SPONGE, COV, VEC, 8, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    case 32: adjoint_vec_kernel<SPONGE, COV, VEC, 32, PD, BATCH><<<grid, block, 0, st>>>(L); break;
    default: adjoint_vec_kernel<SPONGE, COV, VEC, 16, PD, BATCH><<<grid, block, 0, st>>>(L); break;
  }
}

template <bool SPONGE, bool COV, int VEC, bool BATCH>
void adj_vec_depth(const Launch<AdjArgs> &L, dim3 grid, int ry, int pd, hipStream_t st) {
  const dim3 block(kBlockX, kBlockY, 1);
  switch (pd) {
    case 1: adj_vec_rows<SPONGE, COV, VEC, 1, BATCH>(L, grid, block, ry, st); break;
    case 4: adj_vec_rows<SPONGE, COV, VEC, 4, BATCH>(L, grid, block, ry, st); break;
    default: adj_vec_rows<SPONGE, COV, VEC, 2, BATCH>(L, grid, block, ry, st); break;
  }
}

template <bool SPONGE, bool COV, int VEC>
hipError_t adj_vec_dispatch(AdjArgs a, int nshots, int ry, int pd, const int2 *ring, int nring, hipStream_t st) {
  const dim3 grid = vec_grid(a.nx, a.ny, VEC, ry, nshots, &a.tiles_x, &a.tiles_per_shot);
  a.ring = ring;
  a.nring = nring;
  const Launch<AdjArgs> L{a, nullptr, nullptr, 0};
  if (nring > 0)
    adjoint_ring_kernel<SPONGE, COV, false><<<dim3((nring + 255) / 256, nshots, 1), 256, 0, st>>>(L);
  adj_vec_depth<SPONGE, COV, VEC, false>(L, grid, ry, pd, st);
  return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------- launchers

template <typename T>
hipError_t launch_forward_step(const FwdArgs<T> &a, int nshots, const StencilShape &sh, hipStream_t st) {
  int ry = sh.rows;
  if (ry != 4 && ry != 8 && ry != 16) ry = 8;
  return a.sponge_w > 0 ? fwd_dispatch<T, true>(a, nshots, ry, st) : fwd_dispatch<T, false>(a, nshots, ry, st);
}

template <>
hipError_t launch_forward_step<float>(const FwdArgs<float> &a, int nshots, const StencilShape &sh, hipStream_t st) {
  const int vec = sh.vec;
  if (vec == 2 || vec == 4) {
    if (a.nx % vec != 0) return hipErrorInvalidValue;
    const int ry = (sh.rows == 4 || sh.rows == 8 || sh.rows == 32) ? sh.rows : 16;
    const bool sp = a.sponge_w > 0;
    if (vec == 4)
      return sp ? fwd_vec_dispatch<true, 4>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st)
                : fwd_vec_dispatch<false, 4>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st);
    return sp ? fwd_vec_dispatch<true, 2>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st)
              : fwd_vec_dispatch<false, 2>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st);
  }
  int ry = sh.rows;
  if (ry != 4 && ry != 8 && ry != 16) ry = 8;
  return a.sponge_w > 0 ? fwd_dispatch<float, true>(a, nshots, ry, st) : fwd_dispatch<float, false>(a, nshots, ry, st);
}
template hipError_t launch_forward_step<double>(const FwdArgs<double> &, int, const StencilShape &, hipStream_t);

hipError_t launch_adjoint_step(const AdjArgs &a, int nshots, const StencilShape &sh, hipStream_t st) {
  const bool sp = a.sponge_w > 0, cov = a.cov != nullptr;
  const int vec = sh.vec;
  if (vec == 2 || vec == 4) {
    if (a.nx % vec != 0) return hipErrorInvalidValue;
    const int ry = (sh.rows == 4 || sh.rows == 8 || sh.rows == 32) ? sh.rows : 16;
#define TOMO_ADJ_VEC(V)                                                                                   \
  if (sp)                                                                                                 \
    return cov ? adj_vec_dispatch<true, true, V>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st)          \
               : adj_vec_dispatch<true, false, V>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st);        \
  return cov ? adj_vec_dispatch<false, true, V>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st)           \
             : adj_vec_dispatch<false, false, V>(a, nshots, ry, sh.depth, sh.ring, sh.nring, st);
    if (vec == 4) {
      TOMO_ADJ_VEC(4)
    }
    TOMO_ADJ_VEC(2)
#undef TOMO_ADJ_VEC
  }
  const int ry = sh.rows <= 4 ? 4 : 8;  // the adjoint strip holds ~14 values per row in registers
  if (sp) return cov ? adj_dispatch<true, true>(a, nshots, ry, st) : adj_dispatch<true, false>(a, nshots, ry, st);
  return cov ? adj_dispatch<false, true>(a, nshots, ry, st) : adj_dispatch<false, false>(a, nshots, ry, st);
}

template <typename T>
hipError_t launch_coefficients(const float *f, T *pw, int nx, int ny, double c0dt_h, int sponge_w,
                               const T *sp_a, hipStream_t st) {
  const int n = nx * ny;
  coefficients_kernel<T><<<(n + 255) / 256, 256, 0, st>>>(f, pw, nx, ny, T(c0dt_h), sponge_w, sp_a);
  return hipGetLastError();
}
template hipError_t launch_coefficients<float>(const float *, float *, int, int, double, int, const float *,
                                               hipStream_t);
template hipError_t launch_coefficients<double>(const float *, double *, int, int, double, int, const double *,
                                                hipStream_t);

hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  fill_kernel<float><<<blocks, 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  fill_kernel<double><<<blocks, 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_misfit(const float *rr, long long per_shot, int nshots, double *out, hipStream_t st) {
  if (nshots <= 0) return hipSuccess;
  misfit_kernel<<<nshots, 256, 0, st>>>(rr, per_shot, out);
  return hipGetLastError();
}

hipError_t launch_update_uniform(float *f, const float *G, int nshots, int nx, int ny, Mask m, double scale,
                                 double div, hipStream_t st, const float *pixw) {
  const int n = nx * ny;
  update_uniform_kernel<<<(n + 255) / 256, 256, 0, st>>>(f, G, nshots, nx, ny, m, scale, div, pixw);
  return hipGetLastError();
}

hipError_t launch_update_group(float *f, const float *Gg, int nx, int ny, Mask m, double scale,
                               hipStream_t st, const float *pixw) {
  const int n = nx * ny;
  update_group_kernel<<<(n + 255) / 256, 256, 0, st>>>(f, Gg, nx, ny, m, scale, pixw);
  return hipGetLastError();
}

hipError_t launch_pixel_weights(const float *cov, int S, int nx, int ny, Mask m, double eps, double *stats,
                                float *w, hipStream_t st) {
  const long long n2 = (long long)nx * ny;
  coverage_stats_kernel<<<1, 256, 0, st>>>(cov, S, nx, ny, m, stats);
  pixel_weight_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, st>>>(cov, S, n2, eps, stats, w);
  return hipGetLastError();
}

int weighted_part_blocks(int n2) { return std::max(1, std::min(256, (n2 + 4095) / 4096)); }

hipError_t launch_weighted(float *f, const float *G, int M, int nx, int ny, Mask m, double scale,
                           double *mean, double *part, int part_blocks, double *w, hipStream_t st,
                           const float *pixw) {
  const long long n2 = (long long)nx * ny;
  const int nb = (int)((n2 + 255) / 256);
  weighted_mean_kernel<<<nb, 256, 0, st>>>(f, G, M, n2, mean);
  const int chunk = (int)((n2 + part_blocks - 1) / part_blocks);
  weighted_norm_kernel<<<dim3(part_blocks, M), 256, 0, st>>>(f, G, mean, n2, chunk, part);
  weighted_weights_kernel<<<1, 256, 0, st>>>(part, M, part_blocks, w);
  weighted_update_kernel<<<nb, 256, 0, st>>>(f, G, M, nx, ny, m, scale, w, pixw);
  return hipGetLastError();
}

hipError_t launch_avg_weight(float *f, const float *G, int M, int nx, int ny, Mask m, double scale, double *mean,
                             double *part, int part_blocks, double *w, hipStream_t st) {
  const long long n2 = (long long)nx * ny;
  const int nb = (int)((n2 + 255) / 256);
  avgw_mean_kernel<<<nb, 256, 0, st>>>(f, G, M, nx, ny, m, scale, mean);
  const int chunk = (int)((n2 + part_blocks - 1) / part_blocks);
  avgw_norm_kernel<<<dim3(part_blocks, M), 256, 0, st>>>(f, G, mean, nx, ny, m, scale, chunk, part);
  avgw_weights_kernel<<<1, 64, 0, st>>>(part, M, part_blocks, w);
  avgw_update_kernel<<<nb, 256, 0, st>>>(f, G, M, nx, ny, m, scale, w);
  return hipGetLastError();
}

hipError_t launch_coverage_segment(const CovSegArgs &a, hipStream_t st) {
  if (a.n2 % 4 != 0 || a.nx % 4 != 0) return hipErrorInvalidValue;
  if (a.khi < a.klo) return hipSuccess;
  const long long threads = (long long)a.n2 / 4 * a.nshots;
  coverage_segment_kernel<<<dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_error(const float *f, const float *fo, int n, double *out, hipStream_t st) {
  error_kernel<<<1, 256, 0, st>>>(f, fo, n, out);
  return hipGetLastError();
}

void vec_tiles(int nx, int ny, int vec, int rows, int *tiles_x, int *tiles_per_shot) {
  (void)vec_grid(nx, ny, vec, rows, 1, tiles_x, tiles_per_shot);
}

static int batch_rows(int rows) { return (rows == 4 || rows == 8 || rows == 32) ? rows : 16; }

hipError_t launch_forward_batch(const FwdArgs<float> *items, const int *start_main, const int *start_ring, int n,
                                int blocks_main, int blocks_ring, int vec, int rows, int depth, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int ry = batch_rows(rows);
  const Launch<FwdArgs<float>> Lm{FwdArgs<float>{}, items, start_main, n};
  if (vec == 4) fwd_vec_depth<true, 4, true>(Lm, dim3(blocks_main, 1, 1), ry, depth, st);
  else if (vec == 2) fwd_vec_depth<true, 2, true>(Lm, dim3(blocks_main, 1, 1), ry, depth, st);
  else return hipErrorInvalidValue;
  if (blocks_ring > 0) {
    const Launch<FwdArgs<float>> Lr{FwdArgs<float>{}, items, start_ring, n};
    forward_ring_kernel<true, true><<<dim3(blocks_ring, 1, 1), 256, 0, st>>>(Lr);
  }
  return hipGetLastError();
}

hipError_t launch_adjoint_batch(const AdjArgs *items, const int *start_main, const int *start_ring, int n,
                                int blocks_main, int blocks_ring, bool cov, int vec, int rows, int depth,
                                hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int ry = batch_rows(rows);
  const Launch<AdjArgs> Lr{AdjArgs{}, items, start_ring, n};
  const Launch<AdjArgs> Lm{AdjArgs{}, items, start_main, n};
  if (blocks_ring > 0) {
    if (cov) adjoint_ring_kernel<true, true, true><<<dim3(blocks_ring, 1, 1), 256, 0, st>>>(Lr);
    else adjoint_ring_kernel<true, false, true><<<dim3(blocks_ring, 1, 1), 256, 0, st>>>(Lr);
  }
  const dim3 grid(blocks_main, 1, 1);
  if (vec == 4) {
    if (cov) adj_vec_depth<true, true, 4, true>(Lm, grid, ry, depth, st);
    else adj_vec_depth<true, false, 4, true>(Lm, grid, ry, depth, st);
  } else if (vec == 2) {
    if (cov) adj_vec_depth<true, true, 2, true>(Lm, grid, ry, depth, st);
    else adj_vec_depth<true, false, 2, true>(Lm, grid, ry, depth, st);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_error_quadratic(const float *f, const float *fo, const float *G, int M, int nx, int ny, Mask m,
                                  double div, const double *w, const float *pixw, double *d, double *out4,
                                  hipStream_t st) {
  const long long n2 = (long long)nx * ny;
  direction_kernel<<<(unsigned)((n2 + 255) / 256), 256, 0, st>>>(f, G, M, nx, ny, m, div, w, pixw, d);
  quadratic_kernel<<<1, 256, 0, st>>>(f, fo, d, n2, out4);
  return hipGetLastError();
}

hipError_t launch_weights_only(const float *f, const float *G, int M, int nx, int ny, double *mean, double *part,
                               int part_blocks, double *w, hipStream_t st) {
  const long long n2 = (long long)nx * ny;
  const int nb = (int)((n2 + 255) / 256);
  weighted_mean_kernel<<<nb, 256, 0, st>>>(f, G, M, n2, mean);
  const int chunk = (int)((n2 + part_blocks - 1) / part_blocks);
  weighted_norm_kernel<<<dim3(part_blocks, M), 256, 0, st>>>(f, G, mean, n2, chunk, part);
  weighted_weights_kernel<<<1, 256, 0, st>>>(part, M, part_blocks, w);
  return hipGetLastError();
}

// the blocked forward kernels: full-row waves (nx = 64 VEC); the sponge may
// be absent (the reference preset: a = ab = 1 everywhere, the same FMA chain
// as the one-step kernels without sponge)
bool forward2_supported(int nx, int sponge_w) {
  return sponge_w >= 0 && sponge_w <= kSpongeLds && (nx == 128 || nx == 192 || nx == 256 || nx == 384 || nx == 512);
}

hipError_t launch_forward2(Fwd2Args a, int nshots, int rows, int depth, hipStream_t st) {
  const int vec = a.nx / kBlockX;
  // records in the kernel: one row of read-ahead (the record pipeline keeps
  // per-row state)
  const bool rec = a.recv_seg != nullptr && (a.rec1 || a.rec2);
  const int ry = (rows == 8 || rows == 32) ? rows : 16;
  const int strips = (a.ny + ry - 1) / ry;
  a.blocks_per_shot = (strips + kBlockY - 1) / kBlockY;
  const dim3 grid(a.blocks_per_shot * nshots, 1, 1), block(kBlockX, kBlockY, 1);
#define TOMO_FWD2_PD(V, R)                                                            \
  if (rec) forward2_kernel<V, R, 1, true><<<grid, block, 0, st>>>(a);                 \
  else if (depth >= 2) forward2_kernel<V, R, 2, false><<<grid, block, 0, st>>>(a);    \
  else forward2_kernel<V, R, 1, false><<<grid, block, 0, st>>>(a);
#define TOMO_FWD2(V)                                                                  \
  switch (ry) {                                                                       \
    case 8: TOMO_FWD2_PD(V, 8) break;                                                 \
    case 32: TOMO_FWD2_PD(V, 32) break;                                               \
    default: TOMO_FWD2_PD(V, 16) break;                                               \
  }
  if (vec == 8) {
    TOMO_FWD2(8)
  } else if (vec == 4) {
    TOMO_FWD2(4)
  } else if (vec == 2) {
    TOMO_FWD2(2)
  } else if (vec == 6) {
    TOMO_FWD2(6)
  } else if (vec == 3) {
    TOMO_FWD2(3)
  } else {
    return hipErrorInvalidValue;
  }
#undef TOMO_FWD2
#undef TOMO_FWD2_PD
  return hipGetLastError();
}

bool forward3_supported(int nx, int ny, int sponge_w, int rows) {
  const int ry = (rows == 8 || rows == 32) ? rows : 16;
  return forward2_supported(nx, sponge_w) && ny % ry == 0 && ny >= 2 * ry;
}

hipError_t launch_forward3(Fwd3Args a, int nshots, int rows, hipStream_t st) {
  const int vec = a.b.nx / kBlockX;
  const int ry = (rows == 8 || rows == 32) ? rows : 16;
  const int strips = (a.b.ny + ry - 1) / ry;
  a.b.blocks_per_shot = (strips + kBlockY - 1) / kBlockY;
  const bool rec = a.b.recv_seg != nullptr && (a.b.rec1 || a.b.rec2 || a.rec3);
  const dim3 grid(a.b.blocks_per_shot * nshots, 1, 1), block(kBlockX, kBlockY, 1);
#define TOMO_FWD3_R(V, R)                                                             \
  if (rec) forward3_kernel<V, R, true><<<grid, block, 0, st>>>(a);                    \
  else forward3_kernel<V, R, false><<<grid, block, 0, st>>>(a);
#define TOMO_FWD3(V)                                                                  \
  switch (ry) {                                                                       \
    case 8: TOMO_FWD3_R(V, 8) break;                                                  \
    case 32: TOMO_FWD3_R(V, 32) break;                                                \
    default: TOMO_FWD3_R(V, 16) break;                                                \
  }
  if (vec == 8) {
    TOMO_FWD3(8)
  } else if (vec == 4) {
    TOMO_FWD3(4)
  } else if (vec == 2) {
    TOMO_FWD3(2)
  } else if (vec == 6) {
    TOMO_FWD3(6)
  } else if (vec == 3) {
    TOMO_FWD3(3)
  } else {
    return hipErrorInvalidValue;
  }
#undef TOMO_FWD3
#undef TOMO_FWD3_R
  return hipGetLastError();
}

hipError_t launch_forward2_edges(const Fwd2EdgeArgs &a, int nshots, hipStream_t st) {
  if (a.nring <= 0) return hipSuccess;
  forward2_edge_kernel<<<dim3((a.nring + 255) / 256, nshots, 1), 256, 0, st>>>(a);
  return hipGetLastError();
}

// the blocked adjoint kernels: the sponge preset only (point receivers,
// additive injection)
bool adjoint2_supported(int nx, int sponge_w) { return sponge_w > 0 && forward2_supported(nx, sponge_w); }

hipError_t launch_adjoint2(Adj2Args a, int nshots, int rows, hipStream_t st) {
  const int vec = a.nx / kBlockX;
  const int ry = (rows == 8 || rows == 32) ? rows : 16;
  const int strips = (a.ny + ry - 1) / ry;
  a.blocks_per_shot = (strips + kBlockY - 1) / kBlockY;
  const dim3 grid(a.blocks_per_shot * nshots, 1, 1), block(kBlockX, kBlockY, 1);
#define TOMO_ADJ2(V)                                                                  \
  switch (ry) {                                                                       \
    case 8: adjoint2_kernel<V, 8><<<grid, block, 0, st>>>(a); break;                  \
    case 32: adjoint2_kernel<V, 32><<<grid, block, 0, st>>>(a); break;                \
    default: adjoint2_kernel<V, 16><<<grid, block, 0, st>>>(a); break;                \
  }
  if (vec == 8) {
    TOMO_ADJ2(8)
  } else if (vec == 4) {
    TOMO_ADJ2(4)
  } else if (vec == 2) {
    TOMO_ADJ2(2)
  } else if (vec == 6) {
    TOMO_ADJ2(6)
  } else if (vec == 3) {
    TOMO_ADJ2(3)
  } else {
    return hipErrorInvalidValue;
  }
#undef TOMO_ADJ2
  return hipGetLastError();
}

bool adjoint3_supported(int nx, int sponge_w) {
  return sponge_w > 0 && sponge_w <= kSpongeLds && (nx == 256 || nx == 512);
}

// rows: 16, 32 or 64 (64-row strips run with one row of read-ahead)
hipError_t launch_adjoint3(Adj3Args a, int nshots, int rows, int depth, hipStream_t st) {
  Adj2Args &b = a.b;
  if (!adjoint3_supported(b.nx, b.sponge_w)) return hipErrorInvalidValue;
  const int vec = b.nx / kBlockX;
  const int ry = rows == 16 ? 16 : rows == 64 ? 64 : 32;
  const int strips = (b.ny + ry - 1) / ry;
  b.blocks_per_shot = (strips + kBlockY - 1) / kBlockY;
  const dim3 grid(b.blocks_per_shot * nshots, 1, 1), block(kBlockX, kBlockY, 1);
#define TOMO_ADJ3(V)                                                                  \
  if (ry == 16 && depth == 2) adjoint3_kernel<V, 16, 2><<<grid, block, 0, st>>>(a);   \
  else if (ry == 16) adjoint3_kernel<V, 16, 1><<<grid, block, 0, st>>>(a);            \
  else if (ry == 64) adjoint3_kernel<V, 64, 1><<<grid, block, 0, st>>>(a);            \
  else if (depth == 2) adjoint3_kernel<V, 32, 2><<<grid, block, 0, st>>>(a);          \
  else adjoint3_kernel<V, 32, 1><<<grid, block, 0, st>>>(a);
  if (vec == 8) {
    TOMO_ADJ3(8)
  } else {
    TOMO_ADJ3(4)
  }
#undef TOMO_ADJ3
  return hipGetLastError();
}

}  // namespace tomo
