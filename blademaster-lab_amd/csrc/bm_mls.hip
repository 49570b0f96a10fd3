// MLS-MPM cutting baseline (blademaster_warp/mlsmpm.py:42-229) on gfx950.
//
// Updated-Lagrangian MLS-MPM with quadratic B-splines and APIC transfers on a dense grid,
// fixed-corotated stress from the SVD of F (shear scaled for damaged particles), sticky
// walls one node inside the domain, G2P with singular-value clamp, optional Hencky von Mises
// return map with softening, and the J-based damage of TopoCut.  The knife/board contact
// impulses come from the TLMPM contact kernel (bm_rigid_contacts) run on the MLS particle
// state; this file holds the grid transfers.
//
// The grid is dense: node (i, j, k) is (i * dims_y + j) * dims_z + k, momentum and velocity
// are stored as three planes.  P2G sorts each workgroup's particles by stencil cell in LDS and
// gathers per node before adding to the grid (see k_mls_p2g).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

#include "bm_device.h"
#include "bm_hip.h"

using namespace bm;

namespace {

constexpr int kBlock = 256;
constexpr float kMinSingularValue = 0.05f;

inline int blocks_for(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return int(b < 1 ? 1 : b);
}

struct Base {
  int x, y, z;
};

// cell_base and inside_domain (mlsmpm.py:54-72); false when the stencil leaves the grid.
__device__ __forceinline__ bool stencil_base(const bm_mls_params& M, f3 x, Base& b, f3& fx) {
  f3 local = f3{(x.x - float(M.origin[0])) / float(M.dx), (x.y - float(M.origin[1])) / float(M.dx),
                (x.z - float(M.origin[2])) / float(M.dx)};
  float fbx = floorf(local.x - 0.5f), fby = floorf(local.y - 0.5f), fbz = floorf(local.z - 0.5f);
  // Compare in float before converting: a diverged position stays outside.
  if (!(fbx >= 0.0f && fby >= 0.0f && fbz >= 0.0f && fbx < float(M.dims[0] - 2) && fby < float(M.dims[1] - 2) &&
        fbz < float(M.dims[2] - 2)))
    return false;
  b = Base{int(fbx), int(fby), int(fbz)};
  fx = f3{local.x - fbx, local.y - fby, local.z - fbz};
  return true;
}

// Quadratic B-spline weights of node offsets 0, 1, 2 per axis (mlsmpm.py:42-51).
__device__ __forceinline__ void spline(float f, float w[3]) {
  w[0] = 0.5f * (1.5f - f) * (1.5f - f);
  w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
  w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

__device__ __forceinline__ bool finite3(f3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// p2g (mlsmpm.py:75-119).
//
// Instead of 27 x 4 float atomics per particle, a workgroup takes the particles of one brick
// of 3 x 3 x 3 stencil base cells (`perm`, grouped on the host from the reference positions;
// particles that have moved are still handled, since the cells are recomputed here) and:
//   1. computes each particle's APIC terms into LDS, in a form where the contribution to node
//      offset (i, j, k) is
//        w_i w_j w_k (q + a_0 i + a_1 j + a_2 k),  q = p - a f,  a = dx (scale P F^T + m C)
//      (f: the particle's offset from its base node, in cells; w from f);
//   2. finds the box of base cells its particles occupy (wavefront min / max: LDS atomics
//      on six shared words cost 126 us on the 1M block) and counting-sorts the particles
//      by base cell, with one scan inside each wavefront;
//   3. sums, for every occupied cell and node offset, the contributions of the cell's
//      particles: the lane of (cell, j, k) walks the cell's particles, which every lane of
//      the cell reads at the same LDS address, and keeps the 3 offsets i in registers;
//   4. lets each node of the box gather the <= 27 cell sums that reach it and add each
//      component to the grid with one global atomic.
// Cell sums are made kMaxSumCells occupied cells at a time (more only when particles have
// moved across cells: lattice particles that sit on base-cell boundaries split a brick over
// up to 4 x 4 x 4 cells).  A workgroup whose cell box exceeds kMaxCells adds every
// contribution with global atomics instead.  Sums agree with the reference's up to float
// order.  The LDS budget keeps four workgroups per CU: a 21-float record with the weights,
// or 64 cells of sums, left three and cost 40-50 us more.
constexpr int kMaxCells = 512;     // box of base cells (2 per thread in the scan)
constexpr int kMaxSumCells = 52;   // occupied cells whose offset sums fit in LDS at once
constexpr int kRec = 15;           // floats per particle record: f 3, q 3, a 9

__device__ __forceinline__ float spline_at(float f, int i) {
  return i == 0 ? 0.5f * (1.5f - f) * (1.5f - f) : (i == 1 ? 0.75f - (f - 1.0f) * (f - 1.0f) : 0.5f * (f - 0.5f) * (f - 0.5f));
}

__global__ void __launch_bounds__(kBlock) k_mls_p2g(const bm_mls_params M, bm_mls_state S) {
  // 40.5 KB in all, so that four workgroups fit in a CU's LDS.
  __shared__ float rec[kBlock * kRec];
  __shared__ __attribute__((aligned(16))) float cell_sum[kMaxSumCells * 27 * 4];
  // The per-cell counts live in cell_sum's space: they are read by the scan before any sum
  // is written.
  int* cell_count = reinterpret_cast<int*>(cell_sum);
  __shared__ unsigned short cell_start[kMaxCells + 1];
  __shared__ short occ_of[kMaxCells];  // occupied-cell index, -1 for an empty cell
  __shared__ int wave_total[kBlock / 64];
  __shared__ unsigned short sorted[kBlock];
  __shared__ int lo[3], hi[3];
  __shared__ int box_part[kBlock / 64][6];
  const int t = threadIdx.x;
  int64_t p = -1;
  if (S.perm != nullptr) {
    p = S.perm[int64_t(blockIdx.x) * blockDim.x + t];
  } else {
    p = int64_t(blockIdx.x) * blockDim.x + t;
    if (p >= M.n) p = -1;
  }
  Base b = Base{0, 0, 0};
  f3 fx = f3{0.f, 0.f, 0.f};
  bool live = p >= 0;
  f3 x = live ? ld3(S.x, p) : f3{0.f, 0.f, 0.f};
  live = live && finite3(x) && stencil_base(M, x, b, fx);
  const float dx = float(M.dx), mass = float(M.mass);
  f3 mom = f3{0.f, 0.f, 0.f};
  m3 affine = mzero();
  if (live) {
    float mu = float(M.mu);
    if (S.damage[p] > 0.0f) mu = float(M.mu) * float(M.damaged_shear);
    m3 F = ld9(S.F, p);
    m3 U, V;
    f3 sg;
    svd3(F, U, sg, V);
    m3 I = mzero();
    I.a[0] = I.a[4] = I.a[8] = 1.0f;
    m3 P = corotated(U, sg, V, I, mu, float(M.lam));
    float scale = -float(M.dt) * float(M.volume) * 4.0f / (dx * dx);
    m3 stress = mmul_t(P, F);  // P F^T
    m3 C = ld9(S.C, p);
#pragma unroll
    for (int k = 0; k < 9; ++k) affine.a[k] = scale * stress.a[k] + mass * C.a[k];
    mom = mass * ld3(S.v, p) + ld3(S.impulse, p);
  }
  // Box of the base cells: a min / max over each wavefront, then over the wavefronts.
  {
    int v[6] = {live ? b.x : INT_MAX, live ? b.y : INT_MAX, live ? b.z : INT_MAX,
                live ? b.x : INT_MIN, live ? b.y : INT_MIN, live ? b.z : INT_MIN};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
      for (int a = 0; a < 6; ++a) {
        int o = __shfl_xor(v[a], off, 64);
        v[a] = a < 3 ? min(v[a], o) : max(v[a], o);
      }
    if ((t & 63) == 0)
#pragma unroll
      for (int a = 0; a < 6; ++a) box_part[t >> 6][a] = v[a];
    __syncthreads();
    if (t < 6) {
      int r = box_part[0][t];
      for (int w = 1; w < int(blockDim.x >> 6); ++w) r = t < 3 ? min(r, box_part[w][t]) : max(r, box_part[w][t]);
      if (t < 3) lo[t] = r;
      else hi[t - 3] = r;
    }
    __syncthreads();
  }
  const bool any = hi[0] >= lo[0];
  if (!any) return;
  const int cx = hi[0] - lo[0] + 1, cy = hi[1] - lo[1] + 1, cz = hi[2] - lo[2] + 1;
  const int64_t cells64 = int64_t(cx) * cy * cz;
  const int64_t dz = M.dims[2];
  const int64_t plane = M.dims[0] * M.dims[1] * M.dims[2];
  // Direct atomics: this workgroup's particles (when its box of cells is too large) or,
  // below, after counting its occupied cells.
  auto scatter = [&]() {
    if (!live) return;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float weight = spline_at(fx.x, i) * spline_at(fx.y, j) * spline_at(fx.z, k);
          f3 dpos = f3{(float(i) - fx.x) * dx, (float(j) - fx.y) * dx, (float(k) - fx.z) * dx};
          f3 c = weight * (mom + mvec(affine, dpos));
          int64_t node = (int64_t(b.x + i) * M.dims[1] + (b.y + j)) * dz + (b.z + k);
          atomicAdd(S.grid_p + node, c.x);
          atomicAdd(S.grid_p + plane + node, c.y);
          atomicAdd(S.grid_p + 2 * plane + node, c.z);
          atomicAdd(S.grid_m + node, weight * mass);
        }
  };
  if (cells64 > kMaxCells) {
    scatter();
    return;
  }
  const int cells = int(cells64);
  // 1. particle records; counts per cell
  for (int c = t; c < kMaxCells; c += blockDim.x) cell_count[c] = 0;
  int my_cell = -1, rank = 0;
  if (live) {
    float* r = rec + t * kRec;
    r[0] = fx.x;
    r[1] = fx.y;
    r[2] = fx.z;
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      float a0 = dx * affine.a[3 * row], a1 = dx * affine.a[3 * row + 1], a2 = dx * affine.a[3 * row + 2];
      r[6 + 3 * row] = a0;
      r[7 + 3 * row] = a1;
      r[8 + 3 * row] = a2;
      r[3 + row] = comp(mom, row) - (a0 * fx.x + a1 * fx.y + a2 * fx.z);
    }
    my_cell = ((b.x - lo[0]) * cy + (b.y - lo[1])) * cz + (b.z - lo[2]);
  }
  __syncthreads();
  if (live) rank = atomicAdd(&cell_count[my_cell], 1);
  __syncthreads();
  // 2. exclusive scan of the counts and of the occupied flags (packed: count in the low 16
  //    bits, flag in the high bits), thread t holding cells 2t and 2t + 1: a scan inside the
  //    wavefront, then the wavefronts' totals.
  {
    const int c0 = 2 * t, c1 = 2 * t + 1;
    const int a0 = c0 < cells ? cell_count[c0] : 0, a1 = c1 < cells ? cell_count[c1] : 0;
    const int mine = (a0 + a1) | ((a0 > 0 ? 1 : 0) + (a1 > 0 ? 1 : 0)) << 16;
    int incl = mine;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int v = __shfl_up(incl, off, 64);
      if ((t & 63) >= off) incl += v;
    }
    if ((t & 63) == 63) wave_total[t >> 6] = incl;
    __syncthreads();
    int before = incl - mine;
    for (int w = 0; w < (t >> 6); ++w) before += wave_total[w];
    const int start = before & 0xFFFF, occ = before >> 16;
    if (c0 < cells) {
      cell_start[c0] = (unsigned short)start;
      occ_of[c0] = short(a0 > 0 ? occ : -1);
    }
    if (c1 < cells) {
      cell_start[c1] = (unsigned short)(start + a0);
      occ_of[c1] = short(a1 > 0 ? occ + (a0 > 0 ? 1 : 0) : -1);
    }
    if (t == int(blockDim.x) - 1) cell_start[cells] = (unsigned short)((before + mine) & 0xFFFF);
  }
  int occupied = 0;
  for (int w = 0; w < int(blockDim.x >> 6); ++w) occupied += wave_total[w] >> 16;
  __syncthreads();
  if (live) sorted[int(cell_start[my_cell]) + rank] = (unsigned short)t;
  __syncthreads();
  const int ex = cx + 2, ey = cy + 2, ez = cz + 2;
  const int nodes = ex * ey * ez;
  // Occupied cells kMaxSumCells at a time (almost always one pass).
  for (int first = 0; first < occupied; first += kMaxSumCells) {
    const int last = min(occupied, first + kMaxSumCells);
    // 3. cell sums: lane (occupied cell, j, k) over the cell's particles, offsets i in
    //    registers.  Every lane of a cell reads the same particle record.
    for (int item = t; item < cells * 9; item += blockDim.x) {
      const int c = item / 9, jk = item - 9 * c;
      const int occ = occ_of[c];
      if (occ < first || occ >= last) continue;
      const int start = cell_start[c], end = cell_start[c + 1];
      const int j = jk / 3, k = jk - 3 * j;
      const float fj = float(j), fk = float(k);
      float acc[3][4];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[i][q] = 0.0f;
      for (int e = start; e < end; ++e) {
        const float* r = rec + int(sorted[e]) * kRec;
        const float wjk = spline_at(r[1], j) * spline_at(r[2], k);
        float base[3];
#pragma unroll
        for (int row = 0; row < 3; ++row) base[row] = r[3 + row] + r[7 + 3 * row] * fj + r[8 + 3 * row] * fk;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const float w = spline_at(r[0], i) * wjk;
#pragma unroll
          for (int row = 0; row < 3; ++row) acc[i][row] += w * (base[row] + r[6 + 3 * row] * float(i));
          acc[i][3] += w * mass;
        }
      }
      float* dst = cell_sum + (occ - first) * 27 * 4;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[((i * 3 + j) * 3 + k) * 4 + q] = acc[i][q];
    }
    __syncthreads();
    // 4. per node of the node box: the cell sums reaching it (the 4 components of an offset
    //    are one 16-byte read), then one atomic per component.
    const float4* sums = reinterpret_cast<const float4*>(cell_sum);
    for (int n = t; n < nodes; n += blockDim.x) {
      const int nk = n % ez, nj = (n / ez) % ey, ni = n / (ez * ey);
      float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int i = 0; i < 3; ++i) {
        const int bi = ni - i;
        if (bi < 0 || bi >= cx) continue;
        for (int j = 0; j < 3; ++j) {
          const int bj = nj - j;
          if (bj < 0 || bj >= cy) continue;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const int bk = nk - k;
            if (bk < 0 || bk >= cz) continue;
            const int occ = occ_of[(bi * cy + bj) * cz + bk];
            if (occ >= first && occ < last) {
              const float4 v = sums[(occ - first) * 27 + (i * 3 + j) * 3 + k];
              sum.x += v.x;
              sum.y += v.y;
              sum.z += v.z;
              sum.w += v.w;
            }
          }
        }
      }
      const int64_t node = (int64_t(lo[0] + ni) * M.dims[1] + (lo[1] + nj)) * dz + (lo[2] + nk);
      if (sum.x != 0.0f) atomicAdd(S.grid_p + node, sum.x);
      if (sum.y != 0.0f) atomicAdd(S.grid_p + plane + node, sum.y);
      if (sum.z != 0.0f) atomicAdd(S.grid_p + 2 * plane + node, sum.z);
      if (sum.w != 0.0f) atomicAdd(S.grid_m + node, sum.w);
    }
    __syncthreads();
  }
}

// grid_update (mlsmpm.py:122-137): velocity, sticky walls one node inside the boundary.
__global__ void __launch_bounds__(kBlock) k_mls_grid(const bm_mls_params M, bm_mls_state S) {
  int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t plane = M.dims[0] * M.dims[1] * M.dims[2];
  if (node >= plane) return;
  int64_t k = node % M.dims[2], j = (node / M.dims[2]) % M.dims[1], i = node / (M.dims[1] * M.dims[2]);
  float m = S.grid_m[node];
  f3 v = f3{0.f, 0.f, 0.f};
  if (m > 0.0f) v = f3{S.grid_p[node], S.grid_p[plane + node], S.grid_p[2 * plane + node]} / m;
  if (i < 1 || j < 1 || k < 1 || i >= M.dims[0] - 1 || j >= M.dims[1] - 1 || k >= M.dims[2] - 1) v = f3{0.f, 0.f, 0.f};
  S.grid_v[node] = v.x;
  S.grid_v[plane + node] = v.y;
  S.grid_v[2 * plane + node] = v.z;
  // The momentum and mass planes are consumed: leave them zero for the next P2G.
  S.grid_m[node] = 0.0f;
  S.grid_p[node] = 0.0f;
  S.grid_p[plane + node] = 0.0f;
  S.grid_p[2 * plane + node] = 0.0f;
}

// g2p (mlsmpm.py:157-229)
__global__ void __launch_bounds__(kBlock) k_mls_g2p(const bm_mls_params M, bm_mls_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= M.n) return;
  f3 x = ld3(S.x, p);
  if (!finite3(x)) return;
  Base b;
  f3 fx;
  if (!stencil_base(M, x, b, fx)) return;
  float wx[3], wy[3], wz[3];
  spline(fx.x, wx);
  spline(fx.y, wy);
  spline(fx.z, wz);
  float dx = float(M.dx), dt = float(M.dt);
  float inv_dx2 = 4.0f / (dx * dx);
  int64_t dz = M.dims[2];
  int64_t plane = M.dims[0] * M.dims[1] * M.dims[2];
  f3 nv = f3{0.f, 0.f, 0.f};
  m3 nC = mzero();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float weight = wx[i] * wy[j] * wz[k];
        f3 dpos = f3{(float(i) - fx.x) * dx, (float(j) - fx.y) * dx, (float(k) - fx.z) * dx};
        int64_t node = (int64_t(b.x + i) * M.dims[1] + (b.y + j)) * dz + (b.z + k);
        f3 gv = f3{S.grid_v[node], S.grid_v[plane + node], S.grid_v[2 * plane + node]};
        nv = nv + weight * gv;
        float s = inv_dx2 * weight;
        nC.a[0] += s * gv.x * dpos.x; nC.a[1] += s * gv.x * dpos.y; nC.a[2] += s * gv.x * dpos.z;
        nC.a[3] += s * gv.y * dpos.x; nC.a[4] += s * gv.y * dpos.y; nC.a[5] += s * gv.y * dpos.z;
        nC.a[6] += s * gv.z * dpos.x; nC.a[7] += s * gv.z * dpos.y; nC.a[8] += s * gv.z * dpos.z;
      }
  st3(S.v, p, nv);
  st9(S.C, p, nC);
  st3(S.x, p, x + dt * nv);
  m3 A;
#pragma unroll
  for (int k = 0; k < 9; ++k) A.a[k] = dt * nC.a[k];
  A.a[0] += 1.0f;
  A.a[4] += 1.0f;
  A.a[8] += 1.0f;
  m3 f = mmul(A, ld9(S.F, p));
  m3 U, V;
  f3 s;
  svd3(f, U, s, V);
  float smin = fminf(fminf(s.x, s.y), s.z);
  if (smin <= 0.0f) atomicAdd(S.inversions, 1);
  auto rebuild = [&](f3 sv) {
    m3 US;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      US.a[3 * r] = U.a[3 * r] * sv.x;
      US.a[3 * r + 1] = U.a[3 * r + 1] * sv.y;
      US.a[3 * r + 2] = U.a[3 * r + 2] * sv.z;
    }
    return mmul_t(US, V);
  };
  if (smin < kMinSingularValue) {
    s = f3{fmaxf(s.x, kMinSingularValue), fmaxf(s.y, kMinSingularValue), fmaxf(s.z, kMinSingularValue)};
    f = rebuild(s);
  }
  float reference_J = S.damage_reference[p];
  if (M.plastic) {
    f3 e = f3{logf(s.x), logf(s.y), logf(s.z)};
    float mean = (e.x + e.y + e.z) / 3.0f;
    f3 dev = e - f3{mean, mean, mean};
    float norm = len(dev);
    float limit = S.yield_stress[p] / (2.0f * float(M.mu)) * sqrtf(2.0f / 3.0f);
    if (norm > limit) {
      e = e - (1.0f - limit / norm) * dev;
      s = f3{expf(e.x), expf(e.y), expf(e.z)};
      f = rebuild(s);
      if (M.softening > 0.0 && reference_J > 0.0f) {
        float ys = fmaxf(0.0f, S.yield_stress[p] - float(M.softening) * (norm - limit));
        S.yield_stress[p] = ys;
        if (ys <= 0.0f) S.damage[p] = 1.0f;
      }
    }
  }
  st9(S.F, p, f);
  if (reference_J > 0.0f) {
    float J = s.x * s.y * s.z / reference_J;
    if (J <= float(M.compressed) || J >= float(M.stretched)) S.damage[p] = 1.0f;
  }
}

// accumulate_reaction of the MLS substep (mlsmpm.py:140-154).
__global__ void k_mls_accumulate(bm_mls_state S, bm_state T) {
  for (int i = 0; i < 3; ++i) {
    S.interval_impulse[i] += T.reaction[i];
    S.interval_angular[i] += T.torque[i];
    T.reaction[i] = 0.0f;
    T.torque[i] = 0.0f;
    T.board[i] = 0.0f;
  }
  T.speed[0] = 0.0f;
}

inline int status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

bool mls_ok(const bm_mls_params* m, const bm_mls_state* s) {
  return m && s && s->x && s->v && s->C && s->F && s->grid_m && s->grid_p && s->grid_v && m->n >= 0 &&
         m->dims[0] >= 3 && m->dims[1] >= 3 && m->dims[2] >= 3;
}

}  // namespace

extern "C" {

int bm_mls_p2g(const bm_mls_params* m, const bm_mls_state* s, void* stream) {
  if (!mls_ok(m, s)) return -1;
  if (m->n == 0) return 0;
  // One workgroup per group of `perm` (kBlock slots each), or per kBlock particles in index order.
  const int groups = s->perm != nullptr ? int(m->groups) : blocks_for(m->n);
  if (s->perm != nullptr && m->groups < 1) return -1;
  hipLaunchKernelGGL(k_mls_p2g, dim3(groups), dim3(kBlock), 0, (hipStream_t)stream, *m, *s);
  return status();
}

int bm_mls_grid_update(const bm_mls_params* m, const bm_mls_state* s, void* stream) {
  if (!mls_ok(m, s)) return -1;
  int64_t nodes = m->dims[0] * m->dims[1] * m->dims[2];
  hipLaunchKernelGGL(k_mls_grid, dim3(blocks_for(nodes)), dim3(kBlock), 0, (hipStream_t)stream, *m, *s);
  return status();
}

int bm_mls_g2p(const bm_mls_params* m, const bm_mls_state* s, void* stream) {
  if (!mls_ok(m, s)) return -1;
  if (m->n == 0) return 0;
  hipLaunchKernelGGL(k_mls_g2p, dim3(blocks_for(m->n)), dim3(kBlock), 0, (hipStream_t)stream, *m, *s);
  return status();
}

int bm_mls_substep(const bm_params* p, const bm_state* t, const bm_mls_params* m, const bm_mls_state* s,
                   void* stream) {
  if (!mls_ok(m, s) || !p || !t) return -1;
  int r = bm_rigid_contacts(p, t, 0, stream);
  if (r) return r;
  hipLaunchKernelGGL(k_mls_accumulate, dim3(1), dim3(1), 0, (hipStream_t)stream, *s, *t);
  if ((r = bm_mls_p2g(m, s, stream))) return r;
  if ((r = bm_mls_grid_update(m, s, stream))) return r;
  return bm_mls_g2p(m, s, stream);
}

}  // extern "C"
