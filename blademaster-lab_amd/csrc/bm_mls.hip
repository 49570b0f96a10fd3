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
// The workgroup's particles are consecutive in index order, which the sampling makes
// neighbours in space (z fastest): their 27-node stencils overlap heavily.  Instead of 27 x 4
// float atomics per particle, the workgroup
//   1. computes each particle's stencil base cell, momentum and APIC affine term into LDS,
//   2. sorts its particles by base cell inside the box of base cells they occupy (a counting
//      sort: one LDS atomic per particle and a scan),
//   3. lets each node of the box gather the contributions of the particles in the 27 base
//      cells whose stencils reach it, in registers, and adds the node's 4 sums to the grid
//      with one global atomic each.
// A workgroup whose particles spread over more than kMaxCells base cells (large motion) adds
// every contribution with a global atomic instead.  Sums agree up to float order.
constexpr int kMaxCells = 1024;

// Particle record of the gather.  15 floats: the odd stride spreads the lanes' scattered reads
// over the LDS banks (a padded 16-float record with 16-byte reads measured 1.5x slower).
struct P2gParticle {
  float fx, fy, fz, mx, my, mz;  // offset from the base node (cells), momentum
  float a[9];                    // affine term: scale * P F^T + m C
};

__device__ __forceinline__ float spline_at(float f, int i) {
  return i == 0 ? 0.5f * (1.5f - f) * (1.5f - f) : (i == 1 ? 0.75f - (f - 1.0f) * (f - 1.0f) : 0.5f * (f - 0.5f) * (f - 0.5f));
}

__global__ void __launch_bounds__(kBlock) k_mls_p2g(const bm_mls_params M, bm_mls_state S) {
  __shared__ P2gParticle part[kBlock];
  __shared__ int cell_start[kMaxCells + 1];
  __shared__ unsigned short sorted[kBlock];
  __shared__ int lo[3], hi[3];
  const int t = threadIdx.x;
  int64_t p = int64_t(blockIdx.x) * blockDim.x + t;
  Base b = Base{0, 0, 0};
  f3 fx = f3{0.f, 0.f, 0.f};
  bool live = p < M.n;
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
  if (t < 3) {
    lo[t] = INT_MAX;
    hi[t] = INT_MIN;
  }
  __syncthreads();
  if (live) {
    atomicMin(&lo[0], b.x);
    atomicMin(&lo[1], b.y);
    atomicMin(&lo[2], b.z);
    atomicMax(&hi[0], b.x);
    atomicMax(&hi[1], b.y);
    atomicMax(&hi[2], b.z);
  }
  __syncthreads();
  const bool any = hi[0] >= lo[0];
  const int cx = hi[0] - lo[0] + 1, cy = hi[1] - lo[1] + 1, cz = hi[2] - lo[2] + 1;
  const int64_t cells64 = any ? int64_t(cx) * cy * cz : 0;
  const int64_t dz = M.dims[2];
  const int64_t plane = M.dims[0] * M.dims[1] * M.dims[2];
  if (!any) return;
  if (cells64 > kMaxCells) {
    // Scattered particles: every contribution goes to the grid directly.
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
    return;
  }
  const int cells = int(cells64);
  // 1. particle data and counting sort by base cell
  for (int c = t; c <= cells; c += blockDim.x) cell_start[c] = 0;
  if (live) {
    P2gParticle q;
    q.fx = fx.x;
    q.fy = fx.y;
    q.fz = fx.z;
    q.mx = mom.x;
    q.my = mom.y;
    q.mz = mom.z;
#pragma unroll
    for (int k = 0; k < 9; ++k) q.a[k] = affine.a[k];
    part[t] = q;
  }
  __syncthreads();
  int my_cell = -1, rank = 0;
  if (live) {
    my_cell = ((b.x - lo[0]) * cy + (b.y - lo[1])) * cz + (b.z - lo[2]);
    rank = atomicAdd(&cell_start[my_cell + 1], 1);
  }
  __syncthreads();
  // Inclusive scan of the counts in cell_start[1..cells] (one workgroup, cells <= 1024).
  for (int off = 1; off < cells; off <<= 1) {
    int v[kMaxCells / kBlock];
#pragma unroll
    for (int r = 0; r < kMaxCells / kBlock; ++r) {
      int c = 1 + t + r * kBlock;
      v[r] = (c <= cells && c - off >= 1) ? cell_start[c - off] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kMaxCells / kBlock; ++r) {
      int c = 1 + t + r * kBlock;
      if (c <= cells) cell_start[c] += v[r];
    }
    __syncthreads();
  }
  if (live) sorted[cell_start[my_cell] + rank] = (unsigned short)t;
  __syncthreads();
  // 2. gather per node of the node box (cells + 2 per axis)
  const int ex = cx + 2, ey = cy + 2, ez = cz + 2;
  const int nodes = ex * ey * ez;
  for (int n = t; n < nodes; n += blockDim.x) {
    int nk = n % ez, nj = (n / ez) % ey, ni = n / (ez * ey);
    float sx = 0.f, sy = 0.f, sz = 0.f, sm = 0.f;
    for (int i = 0; i < 3; ++i) {
      int bi = ni - i;
      if (bi < 0 || bi >= cx) continue;
      for (int j = 0; j < 3; ++j) {
        int bj = nj - j;
        if (bj < 0 || bj >= cy) continue;
        for (int k = 0; k < 3; ++k) {
          int bk = nk - k;
          if (bk < 0 || bk >= cz) continue;
          int c = (bi * cy + bj) * cz + bk;
          for (int e = cell_start[c]; e < cell_start[c + 1]; ++e) {
            const P2gParticle& q = part[sorted[e]];
            float weight = spline_at(q.fx, i) * spline_at(q.fy, j) * spline_at(q.fz, k);
            float px = (float(i) - q.fx) * dx, py = (float(j) - q.fy) * dx, pz = (float(k) - q.fz) * dx;
            sx += weight * (q.mx + q.a[0] * px + q.a[1] * py + q.a[2] * pz);
            sy += weight * (q.my + q.a[3] * px + q.a[4] * py + q.a[5] * pz);
            sz += weight * (q.mz + q.a[6] * px + q.a[7] * py + q.a[8] * pz);
            sm += weight * mass;
          }
        }
      }
    }
    if (sm == 0.0f && sx == 0.0f && sy == 0.0f && sz == 0.0f) continue;
    int64_t node = (int64_t(lo[0] + ni) * M.dims[1] + (lo[1] + nj)) * dz + (lo[2] + nk);
    atomicAdd(S.grid_p + node, sx);
    atomicAdd(S.grid_p + plane + node, sy);
    atomicAdd(S.grid_p + 2 * plane + node, sz);
    atomicAdd(S.grid_m + node, sm);
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
  hipLaunchKernelGGL(k_mls_p2g, dim3(blocks_for(m->n)), dim3(kBlock), 0, (hipStream_t)stream, *m, *s);
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
