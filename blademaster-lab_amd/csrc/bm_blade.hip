// Held-blade grid constraint (blade_constraint.py:13-324) on gfx950.
//
// Particles beside the broad faces of an externally held blade get a face contact with the
// side they came from; the solver finds nonnegative face impulses lambda with
// W lambda >= rhs (W = N^T M^-1 C N through the compatible-field exchange C) by projected
// gradient with a diagonal row-sum bound and k/(k+3) extrapolation, 48 iterations per batch,
// up to 8 batches until the largest inward residual is <= 1e-6 m/s.  The grid correction
// is added to new_v and the reaction to the tool.
//
// The active set (particles with a face contact, and the nodes they touch) is compacted in
// the prepare pass, so the iterations touch only it: one workgroup of 1024 threads runs all
// batches with workgroup barriers between the scatter, exchange and projection phases, and
// leaves the marks and the scatter grid zero for the next substep.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm_device.h"
#include "bm_hip.h"

using namespace bm;

namespace {

constexpr int kBlock = 256;
constexpr int kSolve = 1024;
constexpr int kIterations = 48;
constexpr int kMaxBatches = 8;

inline int blocks_for(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return int(b < 1 ? 1 : b);
}

__device__ __forceinline__ int field_at(const uint8_t* field_idx, int64_t p, int a) { return field_idx[8 * p + a]; }
__device__ __forceinline__ int64_t corner_node_of(const bm_params& P, int32_t base, int a) {
  return int64_t(base) + corner_offset(a, int(P.dims[1] * P.dims[2]), int(P.dims[2]));
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// prepare (blade_constraint.py:13-70)
__global__ void __launch_bounds__(kBlock) k_bc_prepare(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  quat q = ldq(S.tool_rot);
  f3 pos = ld3(S.tool_pos, 0);
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  float bevel = float(P.bevel), radius = float(P.blade_offset), dt = float(P.dt);
  f3 local = rotate_inv(q, ld3(S.x, p) - pos);
  st3(S.bc_normal, p, f3{0.f, 0.f, 0.f});
  S.bc_impulse[p] = 0.0f;
  S.bc_extrapolated[p] = 0.0f;
  // Going around a finite edge releases the previous entry side.
  if (fabsf(local.y) >= hf.y || local.z <= -hf.z + bevel || local.z >= hf.z) {
    S.bc_side[p] = 0.0f;
    return;
  }
  float w[8];
  stencil_weights(ld3(S.local, p), w);
  int32_t base = S.base_node[p];
  int64_t N = P.node_count;
  f3 grid_v = f3{0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int64_t slot = int64_t(field_at(S.field_idx, p, a)) * N + corner_node_of(P, base, a);
    grid_v = grid_v + w[a] * ld3(S.new_v, slot);
  }
  float s = S.bc_side[p];
  if (s == 0.0f || fabsf(local.x) > hf.x + radius) s = local.x < 0.0f ? -1.0f : 1.0f;
  f3 n = rotate(q, f3{s, 0.f, 0.f});
  float gap = s * local.x - hf.x;
  // Fast arrivals from beyond the proxy radius count with the grid velocity.
  if (gap > radius && gap + dt * dot(grid_v, n) > 0.0f) {
    S.bc_side[p] = 0.0f;
    return;
  }
  S.bc_side[p] = s;
  st3(S.bc_normal, p, n);
  st3(S.bc_lever, p, rotate(q, f3{s * hf.x, local.y, local.z}));
  float rhs = -fmaxf(gap, 0.0f) / dt - dot(grid_v, n);
  S.bc_rhs[p] = rhs;
  S.bc_active[atomicAdd(S.bc_counts + 0, 1)] = int32_t(p);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int64_t node = corner_node_of(P, base, a);
    if (atomicExch(S.bc_node_mark + node, 1) == 0) S.bc_nodes[atomicAdd(S.bc_counts + 1, 1)] = int32_t(node);
    int64_t slot = int64_t(field_at(S.field_idx, p, a)) * N + node;
    atomicAdd(S.bc_grid + 3 * slot, w[a]);
  }
  if (rhs > 1.0e-7f) S.bc_counts[2] = 1;
}

struct Solve {
  const bm_params& P;
  const bm_state& S;
  int na, nn;
  int64_t N;

  // couple (material.py:155-179) over the fields of the active nodes: dst = src + exchange.
  __device__ void couple(const float* src, float* dst) const {
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      int64_t node = S.bc_nodes[i];
      int nf = S.node_nf[node];
      float total = 0.0f;
      for (int f = 0; f < nf; ++f) total += S.grid_mass[int64_t(f) * N + node];
      for (int a = 0; a < nf; ++a) {
        int64_t sa = int64_t(a) * N + node;
        float ma = S.grid_mass[sa];
        if (ma <= 0.0f) continue;
        int64_t ca = S.field_code[sa];
        f3 xa = ld3(src, sa);
        f3 ex = f3{0.f, 0.f, 0.f};
        for (int b = 0; b < nf; ++b) {
          if (b == a) continue;
          int64_t sb = int64_t(b) * N + node;
          if (!compatible(ca, S.field_code[sb])) continue;
          ex = ex + (ma * ld3(src, sb) - S.grid_mass[sb] * xa);
        }
        st3(dst, sa, xa + ex / total);
      }
    }
  }

  // sum_a w_a coupled[s_a] / m[s_a] over the particle's stencil (response_at, :114-127)
  __device__ f3 response(int64_t p) const {
    float w[8];
    stencil_weights(ld3(S.local, p), w);
    int32_t base = S.base_node[p];
    f3 r = f3{0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      int64_t slot = int64_t(field_at(S.field_idx, p, a)) * N + corner_node_of(P, base, a);
      float m = S.grid_mass[slot];
      if (m > 0.0f) r = r + (w[a] / m) * ld3(S.bc_coupled, slot);
    }
    return r;
  }

  // scatter (:93-111) of values[p] * normal[p] onto the grid, its exchange, and the totals.
  __device__ void scatter(const float* values, float* sums) const {
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      int64_t node = S.bc_nodes[i];
      int nf = S.node_nf[node];
      for (int f = 0; f < nf; ++f) st3(S.bc_grid, int64_t(f) * N + node, f3{0.f, 0.f, 0.f});
    }
    if (threadIdx.x < 6) sums[threadIdx.x] = 0.0f;
    __syncthreads();
    f3 lin = f3{0.f, 0.f, 0.f}, ang = lin;
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
      int64_t p = S.bc_active[i];
      f3 j = values[p] * ld3(S.bc_normal, p);
      float w[8];
      stencil_weights(ld3(S.local, p), w);
      int32_t base = S.base_node[p];
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        int64_t slot = int64_t(field_at(S.field_idx, p, a)) * N + corner_node_of(P, base, a);
        f3 c = w[a] * j;
        atomicAdd(S.bc_grid + 3 * slot, c.x);
        atomicAdd(S.bc_grid + 3 * slot + 1, c.y);
        atomicAdd(S.bc_grid + 3 * slot + 2, c.z);
      }
      lin = lin + j;
      ang = ang + cross(ld3(S.bc_lever, p), j);
    }
    float v[6] = {lin.x, lin.y, lin.z, ang.x, ang.y, ang.z};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      float t = wave_sum(v[k]);
      if ((threadIdx.x & 63) == 0 && t != 0.0f) atomicAdd(sums + k, t);
    }
    __syncthreads();
    couple(S.bc_grid, S.bc_coupled);
    __syncthreads();
  }
};

__global__ void __launch_bounds__(kSolve) k_bc_solve(const bm_params P, bm_state S) {
  __shared__ float sums[6];
  __shared__ float red[kSolve / 64];
  __shared__ int refine;
  Solve sv{P, S, S.bc_counts[0], S.bc_counts[1], P.node_count};
  int na = sv.na, nn = sv.nn;
  int64_t N = P.node_count;
  if (S.bc_counts[2] != 0) {
    // coupled active weights and the row-sum bound (row_bound, :73-90)
    sv.couple(S.bc_grid, S.bc_coupled);
    __syncthreads();
    for (int i = threadIdx.x; i < na; i += blockDim.x) {
      int64_t p = S.bc_active[i];
      float w[8];
      stencil_weights(ld3(S.local, p), w);
      int32_t base = S.base_node[p];
      float total = 0.0f;
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        int64_t slot = int64_t(field_at(S.field_idx, p, a)) * N + corner_node_of(P, base, a);
        float m = S.grid_mass[slot];
        if (m > 0.0f) total += w[a] * S.bc_coupled[3 * slot] / m;
      }
      S.bc_bound[p] = total;
    }
    __syncthreads();
    int batches = 0;
    do {
      for (int i = threadIdx.x; i < na; i += blockDim.x) {
        int64_t p = S.bc_active[i];
        S.bc_extrapolated[p] = S.bc_impulse[p];
      }
      __syncthreads();
      for (int it = 0; it < kIterations; ++it) {
        sv.scatter(S.bc_extrapolated, sums);
        float accel = float(it) / float(it + 3);
        // project (:130-152)
        for (int i = threadIdx.x; i < na; i += blockDim.x) {
          int64_t p = S.bc_active[i];
          f3 r = sv.response(p);
          float previous = S.bc_impulse[p];
          float current = fmaxf(0.0f, S.bc_extrapolated[p] +
                                          (S.bc_rhs[p] - dot(ld3(S.bc_normal, p), r)) / fmaxf(S.bc_bound[p], 1.0e-12f));
          S.bc_impulse[p] = current;
          S.bc_extrapolated[p] = current + accel * (current - previous);
        }
        __syncthreads();
      }
      sv.scatter(S.bc_impulse, sums);
      // record_residual (:155-168) and finish_refinement (:171-183)
      float m = 0.0f;
      for (int i = threadIdx.x; i < na; i += blockDim.x) {
        int64_t p = S.bc_active[i];
        f3 r = sv.response(p);
        m = fmaxf(m, fmaxf(0.0f, S.bc_rhs[p] - dot(ld3(S.bc_normal, p), r)));
      }
      m = wave_max(m);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
      __syncthreads();
      if (threadIdx.x == 0) {
        float residual = 0.0f;
        for (int w = 0; w < kSolve / 64; ++w) residual = fmaxf(residual, red[w]);
        batches += 1;
        refine = residual > 1.0e-6f && batches < kMaxBatches;
        if (!refine) S.bc_sums[6] = fmaxf(S.bc_sums[6], residual);
        S.bc_sums[7] = residual;
        S.bc_counts[3] = batches;
      }
      __syncthreads();
      batches = S.bc_counts[3];
    } while (refine);
    // apply_grid (:186-194) on the fields of the active nodes, apply_reaction (:197-205)
    for (int i = threadIdx.x; i < nn; i += blockDim.x) {
      int64_t node = S.bc_nodes[i];
      int nf = S.node_nf[node];
      for (int f = 0; f < nf; ++f) {
        int64_t s = int64_t(f) * N + node;
        float mass = S.grid_mass[s];
        if (mass > 0.0f) st3(S.new_v, s, ld3(S.new_v, s) + ld3(S.bc_coupled, s) / mass);
      }
    }
    if (threadIdx.x < 3) {
      S.reaction[threadIdx.x] -= sums[threadIdx.x];
      S.torque[threadIdx.x] -= sums[3 + threadIdx.x];
    }
    __syncthreads();
  }
  // Leave the marks, the scatter grid and the counters clean for the next substep.
  for (int i = threadIdx.x; i < nn; i += blockDim.x) {
    int64_t node = S.bc_nodes[i];
    S.bc_node_mark[node] = 0;
    for (int f = 0; f < int(P.fields); ++f) st3(S.bc_grid, int64_t(f) * N + node, f3{0.f, 0.f, 0.f});
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    S.bc_counts[0] = 0;
    S.bc_counts[1] = 0;
    S.bc_counts[2] = 0;
  }
}

__global__ void k_store_residual(bm_state S, float* result, int size) { result[size - 1] = S.bc_sums[6]; }

}  // namespace

extern "C" {

int bm_blade_constraint(const bm_params* p, const bm_state* s, void* stream) {
  if (!p || !s || !s->bc_side || !s->bc_active || !s->bc_nodes || !s->bc_node_mark || !s->bc_grid ||
      !s->bc_coupled || !s->bc_counts || !s->bc_sums || !s->new_v || !s->field_idx)
    return -1;
  if (p->n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bc_prepare, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s);
  hipLaunchKernelGGL(k_bc_solve, dim3(1), dim3(kSolve), 0, st, *p, *s);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

int bm_store_residual(const bm_params* p, const bm_state* s, float* result, int32_t size, void* stream) {
  if (!p || !s || !s->bc_sums || !result || size < 1) return -1;
  hipLaunchKernelGGL(k_store_residual, dim3(1), dim3(1), 0, (hipStream_t)stream, *s, result, int(size));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

}  // extern "C"
