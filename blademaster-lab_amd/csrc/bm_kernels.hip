// BladeMaster TLMPM cutting substep for MI355X (gfx950): HIP kernels and the C-ABI.
//
// Design (differences from the reference Warp kernels, same physics):
//  * Grid fields live in a node-major table of F slots per node (field_code[f][node]); the
//    field of particle p at stencil corner a is the slot holding its transfer code at that
//    node.  It replaces the reference's int64 open-addressing key table, compaction and
//    per-node linked lists (material.py:37-75, 621-653): the compatible-field exchange
//    walks at most F slots of one node instead of a list.
//  * The stencil is rebuilt in registers from the base node and three local fractions (16 B
//    per particle) instead of loading 8 node ids, 8 weights and 24 gradient terms (128 B).
//  * The contact sums of a step (speed, reaction, torque, board, work) are reduced per
//    workgroup and stored as partials that one workgroup adds in order: no global atomics.
//    Batched instances add per-workgroup sums into 16 slots per instance.
//  * Particles are binned into bricks of cells; P2G and the MUSL re-scatter sum each cell's
//    particles with DPP moves and write per-brick window partials that the node kernels
//    gather in a fixed order, so the transfers need no global float atomics.
//  * Pair candidates come from bucket chains whose heads carry the build count, so a stale
//    head reads as an empty bucket and no clearing pass is needed; after relabelling only,
//    just the lists around the relabelled particles are rebuilt.  Batched instances keep one
//    list each.
//  * After relabelling only, registration empties and refills just the nodes around the
//    relabelled particles.
//  * Inside a recorded interval the F update of a substep runs in the next substep's P2G,
//    which uses the new stress at once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm_device.h"
#include "bm_hip.h"

using namespace bm;

namespace {

constexpr int kBlock = 256;

inline int blocks_for(int64_t n, int block = kBlock) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return int(b);
}

// ------------------------------------------------------------------------ wave helpers
// DPP lane move within rows of 16 lanes (all lanes of the wavefront active).
template <int Ctrl>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
// Wavefront sum / max in a fixed order, result in every lane: DPP butterflies inside each row
// of 16 lanes (swap neighbours, swap pairs, mirror halves, mirror rows), then the four row
// results.  All lanes of the wavefront must be active.
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dpp_move<0xB1>(v));
  v = fmaxf(v, dpp_move<0x4E>(v));
  v = fmaxf(v, dpp_move<0x141>(v));
  v = fmaxf(v, dpp_move<0x140>(v));
  return fmaxf(fmaxf(lane_value(v, 0), lane_value(v, 16)), fmaxf(lane_value(v, 32), lane_value(v, 48)));
}
__device__ __forceinline__ bool lane0() { return (threadIdx.x & 63) == 0; }
// Instance of particle p (0 for a single solver).
__device__ __forceinline__ int instance_of(const bm_state& S, int64_t p) { return S.instance ? S.instance[p] : 0; }
// Bricks holding particles and nodes touched by particles (static in TLMPM): the transfer and
// grid kernels run over these lists; without lists, over every brick and node.
__device__ __forceinline__ int active_brick(const bm_state& S, int k) { return S.brick_list ? S.brick_list[k] : k; }
__device__ __forceinline__ int64_t active_node_count(const bm_params& P) {
  return P.active_nodes > 0 ? P.active_nodes : P.node_count;
}
__device__ __forceinline__ int64_t active_node(const bm_state& S, int64_t i) {
  return S.node_list ? int64_t(S.node_list[i]) : i;
}

// ------------------------------------------------------------------------ contact laws
struct Contact {
  float dt, sn, mus, mud, gk, gc;
};

// friction (contact.py:21-43): trial tangential impulse from the history spring and
// damper, Coulomb cap; returns impulse and the new history.
__device__ __forceinline__ void friction(const Contact& c, f3 n, f3 relative, float jn, f3 history, float me,
                                         f3& jt, f3& ht) {
  f3 vt = relative - dot(relative, n) * n;
  f3 htan = history - dot(history, n) * n;
  float k = c.gk * me / (c.dt * c.dt);
  float damping = 2.0f * c.gc * sqrtf(me * k);
  f3 trial = ((-k) * htan - damping * vt) * c.dt;
  float magnitude = len(trial);
  if (magnitude <= c.mus * jn) {
    jt = trial;
    ht = htan + c.dt * vt;
    return;
  }
  f3 applied = trial * (c.mud * jn / magnitude);
  jt = applied;
  ht = -((applied / c.dt + damping * vt) / k);
}

// penalty_impulse (contact.py:123-147)
__device__ __forceinline__ void penalty(const Contact& c, float gap, f3 relative, f3 n, float me, f3 history,
                                        float sn, f3& j, f3& h) {
  float predicted = gap + c.dt * fminf(dot(relative, n), 0.0f);
  if (predicted >= 0.0f) {
    j = f3{0.f, 0.f, 0.f};
    h = f3{0.f, 0.f, 0.f};
    return;
  }
  float jn = sn * me / c.dt * (-predicted);
  f3 jt, ht;
  friction(c, n, relative, jn, history, me, jt, ht);
  j = jn * n + jt;
  h = ht;
}

__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// blade_query (contact.py:46-120): closest point, outward normal and signed distance of a
// point in blade coordinates to the (optionally beveled) box.
__device__ inline void blade_query(f3 local, f3 hf, float bevel, f3& closest, f3& normal, float& distance) {
  closest = f3{clampf(local.x, -hf.x, hf.x), clampf(local.y, -hf.y, hf.y), clampf(local.z, -hf.z, hf.z)};
  f3 delta = local - closest;
  distance = len(delta);
  normal = f3{0.f, 0.f, 0.f};
  if (bevel == 0.0f) {
    if (distance > 1.0e-12f) {
      normal = delta / distance;
    } else {
      f3 margin = f3{hf.x - fabsf(local.x), hf.y - fabsf(local.y), hf.z - fabsf(local.z)};
      int axis = 0;
      float m = margin.x;
      if (margin.y < m) {
        axis = 1;
        m = margin.y;
      }
      if (margin.z < m) {
        axis = 2;
        m = margin.z;
      }
      float sgn = comp(local, axis) < 0.0f ? -1.0f : 1.0f;
      if (axis == 0) {
        normal.x = sgn;
        closest.x = sgn * hf.x;
      } else if (axis == 1) {
        normal.y = sgn;
        closest.y = sgn * hf.y;
      } else {
        normal.z = sgn;
        closest.z = sgn * hf.z;
      }
      distance = -m;
    }
    return;
  }
  float qx = fabsf(local.x), qz = local.z;
  float sign = local.x < 0.0f ? -1.0f : 1.0f;
  float edge_len_sq = hf.x * hf.x + bevel * bevel;
  float t = clampf((qx * hf.x + (qz + hf.z) * bevel) / edge_len_sq, 0.0f, 1.0f);
  float nx = t * hf.x, nz = -hf.z + t * bevel;
  float best_sq = (qx - nx) * (qx - nx) + (qz - nz) * (qz - nz);
  float il = sqrtf(edge_len_sq);
  float ox = bevel / il, oz = -hf.x / il;
  float cx = hf.x, cz = clampf(qz, -hf.z + bevel, hf.z);
  float dsq = (qx - cx) * (qx - cx) + (qz - cz) * (qz - cz);
  if (dsq < best_sq) {
    best_sq = dsq;
    nx = cx;
    nz = cz;
    ox = 1.0f;
    oz = 0.0f;
  }
  cx = clampf(qx, 0.0f, hf.x);
  cz = hf.z;
  dsq = (qx - cx) * (qx - cx) + (qz - cz) * (qz - cz);
  if (dsq < best_sq) {
    best_sq = dsq;
    nx = cx;
    nz = cz;
    ox = 0.0f;
    oz = 1.0f;
  }
  bool inside = qx <= hf.x && qz <= hf.z && hf.x * (qz + hf.z) >= bevel * qx;
  float best = sqrtf(best_sq);
  nx = sign * nx;
  ox = sign * ox;
  closest = f3{nx, clampf(local.y, -hf.y, hf.y), nz};
  if (inside && fabsf(local.y) <= hf.y) {
    normal = f3{ox, 0.0f, oz};
    distance = -best;
    if (hf.y - fabsf(local.y) < best) {
      float sy = local.y < 0.0f ? -1.0f : 1.0f;
      closest = f3{local.x, sy * hf.y, local.z};
      normal = f3{0.0f, sy, 0.0f};
      distance = -(hf.y - fabsf(local.y));
    }
  } else {
    if (inside) {
      closest.x = local.x;
      closest.z = local.z;
    }
    delta = local - closest;
    distance = len(delta);
    normal = f3{ox, 0.0f, oz};
    if (distance > 1.0e-12f) normal = delta / distance;
  }
}

struct RigidArgs {
  bm_params p;
  bm_state s;
  int cut_bits;
};

// blade_effective_mass (contact.py:150-178)
__device__ __forceinline__ float effective_mass(const bm_params& P, const float* mobility, f3 lever, f3 n, quat q) {
  float inv = 1.0f / float(P.blade_mass);
  if (P.tool_model < 0) {
    f3 rn = cross(lever, n);
    float u[6] = {n.x, n.y, n.z, rn.x, rn.y, rn.z};
    inv = 0.0f;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int k = 0; k < 6; ++k) inv += u[r] * mobility[6 * r + k] * u[k];
  } else if (P.tool_model > 0) {
    f3 rn = rotate_inv(q, cross(lever, n));
    f3 ii = f3{float(P.inertia_inv[0]), float(P.inertia_inv[1]), float(P.inertia_inv[2])};
    inv += dot(rn, f3{ii.x * rn.x, ii.y * rn.y, ii.z * rn.z});
  }
  return 1.0f / (1.0f / float(P.mass) + inv);
}

// Modes of the contact kernel.
enum RigidMode : int {
  kAccumulate = 0,        // add the step sums to reaction/torque/board/work (standalone launch)
  kOverwrite = 1,         // the sums replace them (clear_step + accumulate of one substep)
  kToolStep = 2,          // then tool.step of a driven or free knife (last block)
  kFinishExternal = 4,    // then finish_external_step (last block)
  kAccumulateInterval = 8,// accumulate_reaction before the tool step
  kPairPrep = 16,         // clear stale pair slots and bin particles for a pending list rebuild
  kTopologyPrep = 32,     // clear the field table for a pending registration
};

struct ToolTarget {
  const float* targets;
  const float* rotations;
  int index;
};

__device__ void tool_step_dev(const bm_params& P, const bm_state& S, const float* targets, const float* rotations,
                              int index, int k = 0);
__device__ void finish_external_dev(const bm_params& P, const bm_state& S);
__device__ void accumulate_reaction_dev(const bm_state& S);

// Tool and interval state read by finish_external_step, loaded in one batch.
struct ExternalTool {
  f3 impulse, angular, pos, vel, omega;
  quat rot;
  float mobility[36];
};
__device__ __forceinline__ ExternalTool load_external_tool(const bm_params& P, const bm_state& S, int i = 0) {
  ExternalTool e;
  e.impulse = ld3(S.interval_impulse, i);
  e.angular = ld3(S.interval_angular, i);
  e.pos = ld3(S.tool_pos, i);
  e.vel = ld3(S.tool_vel, i);
  e.omega = ld3(S.tool_omega, i);
  e.rot = ldq(S.tool_rot + 4 * i);
  for (int k = 0; k < 36; ++k) e.mobility[k] = P.predict ? S.mobility[36 * i + k] : 0.0f;
  return e;
}
__device__ void finish_external_regs(const bm_params& P, const bm_state& S, const ExternalTool& e, f3 j, f3 tau,
                                     int i = 0);

// The pair list of one solver, or of one batched instance: the first bucket and bucket mask of
// its cell list, the tag of its current build and whether it is rebuilt this step.  Batched
// instances keep separate lists, each rebuilt when its own particles have moved too far (as
// the separate runs of a sweep are), in a bucket region of its own.
struct ListRegion {
  int base, mask;
  int tag;   // build tag of the cell lists this step's pair kernel reads (now + bin)
  int now;   // build tag of the cell lists before this step's contact kernel bins
  bool build;  // every list of the region is rebuilt this step
  bool bin;    // the contact kernel bins the particles this step (a rebuild)
};
constexpr int kPairList = 0, kPairHash = 1, kPairAll = 2;  // bm_params.pair_mode
// The search lists of the particles of instance inst; `build` when they are rebuilt this step
// (always in the hash mode, which queries the cell buckets every step).  The build count is
// advanced once the step's pair kernel is done, so a list rebuilt this step carries the tag
// count + 1 from its binning on.
__device__ __forceinline__ ListRegion list_region(const bm_params& P, const bm_state& S, int inst) {
  const bool every_step = P.pair_mode == kPairHash;
  if (S.instance == nullptr) {
    const bool build = every_step || S.pair_dirty[0] != 0;
    const int now = S.pair_builds[0];
    return ListRegion{0, int(P.list_buckets - 1), now + (build ? 1 : 0), now, build, build};
  }
  const bool build = every_step || S.pair_dirty[0] != 0 || S.inst_pair_dirty[inst] != 0;
  const int now = S.inst_pair_builds[inst];
  return ListRegion{S.inst_buckets[2 * inst], S.inst_buckets[2 * inst + 1] - 1, now + (build ? 1 : 0), now, build,
                    build};
}
__device__ void add_relabelled_pairs(const bm_params& P, const bm_state& S, int32_t d, int lane, const ListRegion& L,
                                     bool refreshing);
// Pair-list refresh (one solver, list mode): a particle that moved more than half the skin
// from its build position (a mover) takes its current position as its build position and gets
// a new partner list in the next pair kernel, while every other list keeps its build.
// refresh_mark holds each particle's refresh time, 2 E + 3 for a refresh at pair epoch E
// (times only grow), and refresh_state[3] the time of the last full build: a particle
// refreshed since then is displaced (the cell lists hold it at a stale position, so searches
// take it from the displaced list instead).  A pair belongs to the list of the particle
// refreshed last, the smaller index on a tie; rank: its refresh time, or the build time when
// it is not displaced.  A refresh then never edits another particle's list: the lists that
// hold a mover skip it from then on.
__device__ __forceinline__ int refresh_time(const bm_state& S) { return 2 * S.pair_epoch[0] + 3; }
__device__ __forceinline__ int refresh_rank(const bm_state& S, int32_t q, int built) {
  return max(S.refresh_mark[q], built);
}
// Whether the list of p (rank rp) holds the pair (p, q) (rank rq).
__device__ __forceinline__ bool owns_pair(int rp, int32_t p, int rq, int32_t q) {
  return rq < rp || (rq == rp && q > p);
}


__device__ __forceinline__ void bin_particle(const bm_params& P, const bm_state& S, int64_t p, f3 xp,
                                             const ListRegion& L);

__device__ __forceinline__ int corner_node(int32_t base, int a, int dyz, int dz) { return base + corner_offset(a, dyz, dz); }

constexpr int kRigidSums = 12;  // board 3, reaction 3, torque 3, work, speed, pad
// topology_dirty bits: register every particle (host writes, a full record of relabelled
// particles), or only around the particles relabelled last step (`changed`).
constexpr int kTopoFull = 1;
constexpr int kTopoChanged = 2;
constexpr int kInstSlots = 16;  // slots per instance of the batched contact sums (inst_sums)

// resist_cut (contact.py:182-232): cutting resistance of a particle (position xp, velocity vp)
// just ahead of the knife edge (knife origin, rotation q, half sizes hf, velocity tvel, angular
// velocity tom).  Returns the blade impulse j, less its push along the cutting direction once
// the particle is fully damaged.  While it is intact and the edge advances on it, the push
// does work that adds to its damage; the new damage and the work done come back in `damage`
// and `work` (unchanged when the particle is not ahead of the edge).  radius: the blade
// contact offset; threshold: the work that damages one particle fully.
__device__ inline f3 resist_cut(f3 j, f3 xp, f3 vp, f3 origin, f3 tvel, quat q, f3 tom, f3 hf, float radius,
                                float spacing, float threshold, int cut_bits, float& damage, float& work) {
  f3 edge = origin + rotate(q, f3{0.f, 0.f, -hf.z});
  f3 down = rotate(q, f3{0.f, 0.f, -1.f});
  f3 edge_velocity = tvel + cross(tom, edge - origin);
  float front = dot(xp - edge, down);
  float advance = (cut_bits & BM_CUT_ADVANCE) ? dot(edge_velocity, down) : dot(edge_velocity - vp, down);
  f3 le = rotate_inv(q, xp - edge);
  bool ahead = front >= 0.0f && front <= 2.0f * radius && fabsf(le.x) <= 6.0f * spacing && fabsf(le.y) <= hf.y;
  if (!ahead) return j;
  float push = fmaxf(dot(j, down), 0.0f);
  if (damage >= 1.0f) return j - push * down;
  if (advance > 0.0f) {
    float delta = push * advance;
    work = delta;
    damage = fminf(1.0f, damage + delta / threshold);
  }
  return j;
}

__device__ void refresh_list_of(const bm_params& P, const bm_state& S, int32_t a, int time, int built, float reach,
                                int* group_lds);
constexpr int kRefreshLanes = 16;  // list_lanes at most this with refresh (LDS of a refresh)
constexpr int kRefreshBlocks = 16;
__device__ void refresh_movers(const bm_params& P, const bm_state& S, int block, int blocks);
// Particles relabelled this step (assign) add their new candidate pairs to the partner lists
// that the next step's pair kernel uses (add_relabelled_pairs), unless those lists are rebuilt
// anyway; one wave per relabelled particle: wave `wave` of `waves`.  Beside the refresh of the
// movers' lists (refreshing), a mover's pairs are left to its refresh.
__device__ void mark_relabelled(const bm_params& P, const bm_state& S, int wave, int waves, bool refreshing) {
  if (S.pair_dirty[0] != 0 || P.pair_mode != kPairList) return;
  const int lane = int(threadIdx.x) & 63;
  const int relabels = min(S.changed_count[0], int(P.changed_capacity));
  for (int i = wave; i < relabels; i += waves) {
    const int32_t d = S.changed[i];
    ListRegion L = list_region(P, S, instance_of(S, d));
    // A particle relabelled by two knives in one step is listed twice: one walk is enough.
    int seen = 0;
    if (lane == 0) seen = atomicExch(S.pair_mark + d, S.pair_epoch[0] + 1) == S.pair_epoch[0] + 1;
    seen = __shfl(seen, 0, 64);
    if (!L.build && !seen) add_relabelled_pairs(P, S, d, lane, L, refreshing);  // a list rebuilt in full needs nothing
  }
}

// rigid_and_speed (contact.py:332-497): gravity, board and blade contact of every particle.
// Per-block sums go to rigid_partial (reduced by the pair kernel's first workgroups and the
// node kernel's last).  In a substep (kPairPrep) the same pass also empties the pair slots of
// the new table and resets its insertion count, bins particles for a pair-list rebuild and
// clears the field table for a re-registration.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) k_rigid(const bm_params P, const bm_state S, int cut_bits, int mode,
                                                  int parity_new, int refresh, int marks) {
  // The first `refresh` workgroups refresh the movers' pair lists (refresh_movers), the next
  // `marks` add the last G2P's relabelled particles' new pairs to the lists (mark_relabelled).
  if (int(blockIdx.x) < refresh) {
    refresh_movers(P, S, int(blockIdx.x), refresh);
    return;
  }
  if (int(blockIdx.x) < refresh + marks) {
    mark_relabelled(P, S, (int(blockIdx.x) - refresh) * (kBlock / 64) + int(threadIdx.x >> 6), marks * (kBlock / 64),
                    refresh > 0);
    return;
  }
  const int bid = int(blockIdx.x) - refresh - marks;
  int64_t p = int64_t(bid) * blockDim.x + threadIdx.x;
  bool live = p < P.n;
  // The particle's own loads go first, ahead of the step's rare preparation loops below
  // (whose stores the compiler cannot move them past).
  // (A lane past the last particle loads particle 0 and is ignored.)
  const f3 xp = ld3(S.x, live ? p : 0), vp = ld3(S.v, live ? p : 0);
  // Friction histories that may be nonzero (bit p & 63 of the wavefront's word per row, one
  // scalar load per row): a zero history is written only over one of them.
  const int64_t mask_words = (P.n + 63) >> 6;
  uint64_t may0 = ~0ull, may1 = ~0ull;
  if (S.history_mask != nullptr) {
    const int64_t word = ((int64_t(bid) * blockDim.x) >> 6) + (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);
    if (word < mask_words) {
      may0 = uniform_load(S.history_mask + word);
      may1 = uniform_load(S.history_mask + mask_words + word);
    }
  }
  const float dt = float(P.dt);
  Contact c{dt, float(P.sn), float(P.mus), float(P.mud), float(P.gk), float(P.gc)};
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  // The knife of this particle's instance (for a single solver, knife 0 by scalar loads).
  int inst = 0;
  quat q;
  f3 origin, tvel, tom;
  if (S.instance == nullptr) {
    q = quat{uniform_load(S.tool_rot), uniform_load(S.tool_rot + 1), uniform_load(S.tool_rot + 2),
             uniform_load(S.tool_rot + 3)};
    origin = uniform_ld3(S.tool_pos, 0);
    tvel = uniform_ld3(S.tool_vel, 0);
    tom = uniform_ld3(S.tool_omega, 0);
  } else {
    inst = live ? S.instance[p] : 0;
    q = ldq(S.tool_rot + 4 * inst);
    origin = ld3(S.tool_pos, inst);
    tvel = ld3(S.tool_vel, inst);
    tom = ld3(S.tool_omega, inst);
  }
  // A new substep: empty the node kernels' multi-field node list.
  if (p == 0 && S.multi_count != nullptr) {
    S.multi_count[0] = 0;
    S.multi_count[1] += 1;
  }
  // The pair kernel inserts into the new table from an empty count.
  if (p == 0 && (mode & kPairPrep) != 0) S.counters[1 + parity_new] = 0u;
  f3 board_j = f3{0.f, 0.f, 0.f}, react = board_j, torq = board_j;
  float work = 0.0f, spd = 0.0f;
  int64_t stride = int64_t(int(gridDim.x) - refresh - marks) * blockDim.x;
  if ((mode & kPairPrep) != 0) {
    // clear_pair_slots (contact.py:643-652): empty the slots of the table used two steps ago.
    const int count = uniform_load(S.pair_count + parity_new);
    int64_t* keys = S.pair_keys + int64_t(parity_new) * P.pair_capacity;
    const int32_t* used = S.pair_used + int64_t(parity_new) * P.pair_capacity;
    for (int64_t a = p; a < count; a += stride) keys[used[a]] = -1;
  }
  const int topo = (mode & kTopologyPrep) != 0 ? uniform_load(S.topology_dirty) : 0;
  if (topo & kTopoFull) {
    // register, first half (solver.py:392-393): empty field table before the inserts.
    int64_t nodes = active_node_count(P);
    for (int64_t i = p; i < nodes; i += stride) {
      int64_t node = active_node(S, i);
      for (int f = 0; f < int(P.fields); ++f) S.field_code[int64_t(f) * P.node_count + node] = -1;
      S.node_nf[node] = 0;
    }
  } else if (topo & kTopoChanged) {
    // Only the stencil nodes of the particles relabelled last step can change their codes:
    // empty those (the particles around them register again in the pair kernel).
    int64_t count = min(uniform_load(S.changed_count), int(P.changed_capacity));
    int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
    for (int64_t a = p; a < count * 8; a += stride) {
      int64_t node = corner_node(S.base_node[S.changed[a >> 3]], int(a & 7), dyz, dz);
      for (int f = 0; f < int(P.fields); ++f) S.field_code[int64_t(f) * P.node_count + node] = -1;
      S.node_nf[node] = 0;
    }
  }
  f3 total = f3{0.f, 0.f, 0.f};
  const bool nonzero0 = (may0 >> (p & 63)) & 1ull, nonzero1 = (may1 >> (p & 63)) & 1ull;
  bool board_contact = false, blade_contact = false;
  if (live) {
    spd = len(vp);
    if ((mode & kPairPrep) != 0) {
      ListRegion L = list_region(P, S, inst);
      if (L.bin) bin_particle(P, S, p, xp, L);
    }
    total = f3{0.f, 0.f, -9.81f * float(P.mass) * dt};
    // board: plane z = 0 inside its rectangle; penalty_impulse is zero (history zero) unless
    // the predicted gap closes (contact.py:123-147)
    float gap = xp.z - float(P.radius);
    if (P.board_half[0] > 0.0 && fabsf(xp.x) > float(P.board_half[0])) gap = 1.0e30f;
    if (P.board_half[1] > 0.0 && fabsf(xp.y) > float(P.board_half[1])) gap = 1.0e30f;
    if (gap + dt * fminf(vp.z, 0.0f) < 0.0f) {
      board_contact = true;
      f3 jb, hb;
      penalty(c, gap, vp, f3{0.f, 0.f, 1.f}, float(P.mass), ld3(S.histories, p), float(P.board_sn), jb, hb);
      st3(S.histories, p, hb);
      if (dot(jb, jb) > 0.0f) {
        total = total + jb;
        board_j = jb;
      }
    } else if (nonzero0) {
      st3(S.histories, p, f3{0.f, 0.f, 0.f});
    }
    // blade (blade_at, contact.py:235-329)
    float blade_radius = len(hf);
    float max_closing = spd + len(tvel) + len(tom) * blade_radius;
    float offset = float(P.blade_offset);
    int64_t hrow = P.n + p;
    f3 local = f3{0.f, 0.f, 0.f};
    bool near = len(xp - origin) - blade_radius - offset <= dt * max_closing;
    if (near) {
      // The distance to the box without the bevel bounds the blade's from below (the bevel
      // only removes material), and |relative . n| <= max_closing: a particle whose box
      // distance stays beyond the offset over the step cannot touch the blade.  The column
      // under the edge of a held cut keeps the exact query (its distance is the edge's).
      local = rotate_inv(q, xp - origin);
      f3 out = f3{fmaxf(fabsf(local.x) - hf.x, 0.f), fmaxf(fabsf(local.y) - hf.y, 0.f),
                  fmaxf(fabsf(local.z) - hf.z, 0.f)};
      float box = len(out);
      bool hold_column = (cut_bits & BM_CUT_HOLD) && fabsf(local.x) <= hf.x + offset && fabsf(local.y) <= hf.y &&
                         -hf.z - local.z >= 0.0f;
      if (!hold_column && box - offset > dt * max_closing + 1.0e-6f * (box + blade_radius)) near = false;
    }
    if (!near) {
      if (nonzero1) st3(S.histories, hrow, f3{0.f, 0.f, 0.f});
    } else {
      f3 closest, normal;
      float distance;
      blade_query(local, hf, float(P.bevel), closest, normal, distance);
      float front = -hf.z - local.z;
      float dmg = S.damage[p];
      if ((cut_bits & BM_CUT_HOLD) && dmg < 1.0f && fabsf(local.x) <= hf.x + offset && fabsf(local.y) <= hf.y &&
          front >= 0.0f) {
        closest = f3{local.x, local.y, -hf.z};
        normal = f3{0.f, 0.f, -1.f};
        distance = front;
      }
      f3 lever = rotate(q, closest);
      f3 n = rotate(q, normal);
      f3 relative = vp - tvel - cross(tom, lever);
      if (distance - offset + dt * fminf(dot(relative, n), 0.0f) >= 0.0f) {
        if (nonzero1) st3(S.histories, hrow, f3{0.f, 0.f, 0.f});
      } else {
        blade_contact = true;
        float me = effective_mass(P, S.mobility + 36 * inst, lever, n, q);
        f3 j, h;
        penalty(c, distance - offset, relative, n, me, ld3(S.histories, hrow), float(P.sn), j, h);
        st3(S.histories, hrow, h);
        if (dot(j, j) > 0.0f) {
          if (cut_bits != 0) {
            float d = dmg;
            j = resist_cut(j, xp, vp, origin, tvel, q, tom, hf, offset, float(P.spacing), float(P.threshold), cut_bits,
                           d, work);
            if (d != dmg) S.damage[p] = d;
          }
          total = total + j;
          react = -j;
          torq = -cross(lever, j);
        }
      }
    }
  }
  // This step's contacts are the histories that may be nonzero next step.
  const uint64_t now0 = __ballot(board_contact), now1 = __ballot(blade_contact);
  if (S.history_mask != nullptr && lane0() && p < P.n) {
    S.history_mask[p >> 6] = now0;
    S.history_mask[mask_words + (p >> 6)] = now1;
  }
  if (live) st3(S.impulses, p, total);
  // Batched instances: the step sums go to each instance's own slots.  The block's first
  // particle names the instance of the block reduction below; a wave with a live particle of
  // another instance (at most the waves where instances meet, since they are stored one after
  // the other) adds its lanes' sums directly instead.
  const int block_inst = S.instance != nullptr ? S.instance[int64_t(bid) * blockDim.x] : 0;
  if (S.instance != nullptr && __ballot(live && inst != block_inst) != 0) {
    float v[10] = {board_j.x, board_j.y, board_j.z, react.x, react.y, react.z, torq.x, torq.y, torq.z, work};
    float* dst = S.inst_sums + (int64_t(inst) * kInstSlots + (bid & (kInstSlots - 1))) * 10;
#pragma unroll
    for (int k = 0; k < 10; ++k)
      if (v[k] != 0.0f) atomicAdd(dst + k, v[k]);
    board_j = react = torq = f3{0.f, 0.f, 0.f};
    work = 0.0f;
  }
  // Block sums: wave reduction, then the waves of the block in order.
  __shared__ float wsum[kBlock / 64][kRigidSums];
  float v[kRigidSums] = {board_j.x, board_j.y, board_j.z, react.x, react.y, react.z,
                         torq.x,    torq.y,    torq.z,    work,    spd,     0.0f};
  // Most waves have no particle at the blade (and many none on the board): their reaction,
  // torque and work sums are exact zeros and skip the reduction.
  bool board_any = __ballot(board_j.x != 0.0f || board_j.y != 0.0f || board_j.z != 0.0f) != 0;
  bool blade_any = __ballot(react.x != 0.0f || react.y != 0.0f || react.z != 0.0f || work != 0.0f) != 0;
  if (board_any) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = wave_sum(v[k]);
  }
  if (blade_any) {
#pragma unroll
    for (int k = 3; k < 10; ++k) v[k] = wave_sum(v[k]);
  }
  v[10] = wave_max(v[10]);
  int wave = threadIdx.x >> 6;
  if (lane0()) {
#pragma unroll
    for (int k = 0; k < kRigidSums; ++k) wsum[wave][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < kRigidSums) {
    int k = threadIdx.x;
    float acc = wsum[0][k];
    for (int w = 1; w < kBlock / 64; ++w) acc = (k == 10) ? fmaxf(acc, wsum[w][k]) : acc + wsum[w][k];
    if (S.instance != nullptr && k < 10) {
      // One add per sum and block into a slot of the block's instance (the step's finish reads
      // only the speed of rigid_partial for batched instances).
      if (acc != 0.0f) atomicAdd(S.inst_sums + (int64_t(block_inst) * kInstSlots + (bid & (kInstSlots - 1))) * 10 + k, acc);
      acc = 0.0f;
    }
    S.rigid_partial[int64_t(bid) * kRigidSums + k] = acc;
  }
}

// ------------------------------------------------------------------ extra knives
constexpr int kExtraSums = 8;  // reaction 3, torque 3, work, pad

// blade_at (contact.py:235-329) of knife k >= 1 on particle p (extra_blade, contact.py:500-566):
// the friction history row and the damage are knife k's own; the reaction, torque and work go
// to the caller's sums.  Returns the impulse on the particle.
__device__ inline f3 extra_knife_contact(const bm_params& P, const bm_state& S, const Contact& c, int64_t p, f3 xp,
                                         f3 vp, int k, int cut_bits, f3& react, f3& torq, float& work) {
  const float dt = c.dt;
  const f3 zero = f3{0.f, 0.f, 0.f};
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  quat q = ldq(S.tool_rot + 4 * k);
  f3 origin = ld3(S.tool_pos, k), tvel = ld3(S.tool_vel, k), tom = ld3(S.tool_omega, k);
  float* hist = S.extra_histories + 3 * int64_t(k - 1) * P.n;
  float* dmg_row = S.extra_damage + int64_t(k - 1) * P.n;
  float blade_radius = len(hf);
  float max_closing = len(vp) + len(tvel) + len(tom) * blade_radius;
  float offset = float(P.blade_offset);
  if (len(xp - origin) - blade_radius - offset > dt * max_closing) {
    st3(hist, p, zero);
    return zero;
  }
  f3 local = rotate_inv(q, xp - origin);
  f3 closest, normal;
  float distance;
  blade_query(local, hf, float(P.bevel), closest, normal, distance);
  float front = -hf.z - local.z;
  float dmg = dmg_row[p];
  if ((cut_bits & BM_CUT_HOLD) && dmg < 1.0f && fabsf(local.x) <= hf.x + offset && fabsf(local.y) <= hf.y &&
      front >= 0.0f) {
    closest = f3{local.x, local.y, -hf.z};
    normal = f3{0.f, 0.f, -1.f};
    distance = front;
  }
  f3 lever = rotate(q, closest);
  f3 n = rotate(q, normal);
  f3 relative = vp - tvel - cross(tom, lever);
  if (distance - offset + dt * fminf(dot(relative, n), 0.0f) >= 0.0f) {
    st3(hist, p, zero);
    return zero;
  }
  float me = effective_mass(P, S.mobility, lever, n, q);
  f3 j, h;
  penalty(c, distance - offset, relative, n, me, ld3(hist, p), float(P.sn), j, h);
  st3(hist, p, h);
  if (!(dot(j, j) > 0.0f)) return zero;
  if (cut_bits != 0) {
    float d = dmg, w = 0.0f;
    j = resist_cut(j, xp, vp, origin, tvel, q, tom, hf, offset, float(P.spacing), float(P.threshold), cut_bits, d, w);
    if (d != dmg) dmg_row[p] = d;
    work += w;
  }
  react = react - j;
  torq = torq - cross(lever, j);
  return j;
}

// extra_blade (contact.py:500-566) for knives 1 .. tools-1, after k_rigid: each knife's impulse is
// added to the particle's in knife order, and each knife's block sums go to extra_partial.
__global__ void __launch_bounds__(kBlock) k_extra_blades(const bm_params P, bm_state S, int cut_bits) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = p < P.n;
  Contact c{float(P.dt), float(P.sn), float(P.mus), float(P.mud), float(P.gk), float(P.gc)};
  f3 xp = f3{0.f, 0.f, 0.f}, vp = xp, imp = xp;
  if (live) {
    xp = ld3(S.x, p);
    vp = ld3(S.v, p);
    imp = ld3(S.impulses, p);
  }
  __shared__ float wsum[kBlock / 64][kExtraSums];
  const int wave = int(threadIdx.x >> 6);
  for (int k = 1; k < int(P.tools); ++k) {
    f3 react = f3{0.f, 0.f, 0.f}, torq = react;
    float work = 0.0f;
    if (live) imp = imp + extra_knife_contact(P, S, c, p, xp, vp, k, cut_bits, react, torq, work);
    float v[kExtraSums] = {react.x, react.y, react.z, torq.x, torq.y, torq.z, work, 0.0f};
#pragma unroll
    for (int i = 0; i < kExtraSums - 1; ++i) v[i] = wave_sum(v[i]);
    if (lane0()) {
#pragma unroll
      for (int i = 0; i < kExtraSums; ++i) wsum[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < kExtraSums) {
      float acc = wsum[0][threadIdx.x];
      for (int w = 1; w < kBlock / 64; ++w) acc += wsum[w][threadIdx.x];
      S.extra_partial[(int64_t(k - 1) * gridDim.x + blockIdx.x) * kExtraSums + threadIdx.x] = acc;
    }
    __syncthreads();
  }
  if (live) st3(S.impulses, p, imp);
}

// register (material.py:50-65): the field slot of code `code` at node `node`, claiming the
// first empty slot when the code is new there.  Used slots stay a prefix of the node's F
// slots; a node needing more than F fields counts errors[0] and takes slot 0.
__device__ __forceinline__ int claim_slot(const bm_params& P, const bm_state& S, int64_t node, int64_t code) {
  const int64_t N = P.node_count;
  const int F = int(P.fields);
  for (int f = 0; f < F; ++f) {
    int64_t* slot = S.field_code + int64_t(f) * N + node;
    int64_t cur = *slot;
    if (cur == -1) {
      unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)(-1LL),
                                          (unsigned long long)code);
      cur = int64_t(prev) == -1 ? code : int64_t(prev);
      if (int64_t(prev) == -1) atomicMax(S.node_nf + node, f + 1);
    }
    if (cur == code) return f;
  }
  atomicAdd(S.errors + 0, 1);
  return 0;
}

// While the particle's F update waits for the next P2G (update_pending 2), that update needs
// the slots of the grid velocities it was computed on: the first registration of the step
// keeps them in field_idx_prev and marks the particle.  Runs before any of the particle's
// slots of this registration are written.
__device__ __forceinline__ void keep_previous_slots(const bm_state& S, int64_t p) {
  const uint2 before = reinterpret_cast<const uint2*>(S.field_idx)[p];
  if (S.field_idx_prev != nullptr && S.update_pending != nullptr && S.update_pending[0] == 2 &&
      atomicExch(S.update_done + p, 1) == 0)
    reinterpret_cast<uint2*>(S.field_idx_prev)[p] = before;
}

// The field slot of the particle's transfer code at each of its 8 nodes, in field_idx.
__device__ void insert_particle(const bm_params& P, const bm_state& S, int64_t p) {
  keep_previous_slots(S, p);
  int64_t code = S.transfer_codes[p];
  int32_t base = S.base_node[p];
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  uint32_t packed[2] = {0u, 0u};
  for (int a = 0; a < 8; ++a) {
    const int found = claim_slot(P, S, corner_node(base, a, dyz, dz), code);
    packed[a >> 2] |= uint32_t(found & 0xff) << (8 * (a & 3));
  }
  reinterpret_cast<uint2*>(S.field_idx)[p] = make_uint2(packed[0], packed[1]);
}


// Sum (the speed, k = 10: maximum) of `count` rows of kRigidSums contact sums, in row order
// within each thread's stride and in wave order across the block, so that the result does not
// depend on timing.  Every thread of the block calls it; the sums end in part[0].
__device__ void reduce_rigid_rows(const float* rows, int count, float (*part)[kRigidSums]) {
  float acc[kRigidSums];
#pragma unroll
  for (int k = 0; k < kRigidSums; ++k) acc[k] = 0.0f;
#pragma unroll 4
  for (int b = threadIdx.x; b < count; b += blockDim.x) {
    const float4* src = reinterpret_cast<const float4*>(rows + int64_t(b) * kRigidSums);
    float4 q0 = src[0], q1 = src[1], q2 = src[2];
    float v[kRigidSums] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y, q2.z, q2.w};
#pragma unroll
    for (int k = 0; k < kRigidSums; ++k) acc[k] = (k == 10) ? fmaxf(acc[k], v[k]) : acc[k] + v[k];
  }
#pragma unroll
  for (int k = 0; k < kRigidSums; ++k) acc[k] = (k == 10) ? wave_max(acc[k]) : wave_sum(acc[k]);
  if (lane0()) {
#pragma unroll
    for (int k = 0; k < kRigidSums; ++k) part[threadIdx.x >> 6][k] = acc[k];
  }
  __syncthreads();
  // One thread per sum adds the waves' partials in wave order.
  if (threadIdx.x < kRigidSums) {
    int k = int(threadIdx.x);
    float a = part[0][k];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) a = (k == 10) ? fmaxf(a, part[w][k]) : a + part[w][k];
    part[0][k] = a;
  }
  __syncthreads();
}

// The step's contact sums t (board 3, reaction 3, torque 3, work, speed) applied: stored or
// added (kOverwrite), then finish_external_step (kFinishExternal) or the interval sums and
// tool.step (kAccumulateInterval, kToolStep).  Batched instances: each thread finishes some
// instances from their contact-sum slots, which it clears for the next step.  Every thread of
// the workgroup calls it.
__device__ void finish_rigid_step(const bm_params& P, const bm_state& S, const float* t, int mode,
                                  const ToolTarget& tt) {
  if (S.instance != nullptr) {
    for (int i = int(threadIdx.x); i < int(P.instances); i += int(blockDim.x)) {
      // This step's sums of instance i: its slots, which are cleared for the next step.
      float sum[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float* slots = S.inst_sums + int64_t(i) * kInstSlots * 10;
      for (int b = 0; b < kInstSlots; ++b)
#pragma unroll
        for (int k = 0; k < 10; ++k) sum[k] += slots[b * 10 + k];
      for (int b = 0; b < kInstSlots; ++b)
#pragma unroll
        for (int k = 0; k < 10; ++k) slots[b * 10 + k] = 0.0f;
      f3 board = ld3(S.board, i) + f3{sum[0], sum[1], sum[2]};
      f3 j = ld3(S.reaction, i) + f3{sum[3], sum[4], sum[5]};
      f3 tau = ld3(S.torque, i) + f3{sum[6], sum[7], sum[8]};
      S.work[i] += sum[9];
      if (mode & kFinishExternal) {
        ExternalTool e = load_external_tool(P, S, i);
        finish_external_regs(P, S, e, j, tau, i);
      } else {
        st3(S.board, i, board);
        st3(S.reaction, i, j);
        st3(S.torque, i, tau);
      }
    }
    if (threadIdx.x == 0) S.speed[0] = t[10];
    return;
  }
  if (threadIdx.x != 0) return;
  // A load's wait also waits for this thread's earlier stores, so everything the serial tail
  // reads is loaded before anything is written.
  f3 board = f3{t[0], t[1], t[2]}, react = f3{t[3], t[4], t[5]}, torq = f3{t[6], t[7], t[8]};
  float speed = t[10];
  if (!(mode & kOverwrite)) {
    board = board + ld3(S.board, 0);
    react = react + ld3(S.reaction, 0);
    torq = torq + ld3(S.torque, 0);
    speed = fmaxf(speed, S.speed[0]);
  }
  float work = S.work[0] + t[9];
  bool finish = (mode & kFinishExternal) != 0;
  ExternalTool ext;
  if (finish) ext = load_external_tool(P, S);
  // stores
  S.work[0] = work;
  if (finish) {
    // finish_external_step: the step sums are consumed and reset here.
    finish_external_regs(P, S, ext, react, torq);
    return;
  }
  st3(S.board, 0, board);
  st3(S.reaction, 0, react);
  st3(S.torque, 0, torq);
  S.speed[0] = speed;
  if (mode & kAccumulateInterval) accumulate_reaction_dev(S);
  if (mode & kToolStep) tool_step_dev(P, S, tt.targets, tt.rotations, tt.index);
}

// Registration of the field slots when the labels changed (topology_dirty), by workgroup
// `block` of `blocks`, after the contact kernel emptied the affected nodes.
__device__ void register_changed(const bm_params& P, const bm_state& S, int block, int blocks) {
  const int topo = S.topology_dirty[0];
  if (topo == 0) return;
  const int64_t stride = int64_t(blocks) * blockDim.x;
  const int64_t first = int64_t(block) * blockDim.x + threadIdx.x;
  if ((topo & kTopoFull) == 0) {
    // After relabelling only: every particle of the 27 cells around a relabelled particle's
    // cell touches a node emptied in k_rigid and registers again; the other particles' nodes
    // and slots are unchanged.  A particle met twice registers twice, finding its slots.
    // One wavefront per (relabelled particle, neighbouring cell), a lane per (particle slot,
    // corner): the 8 lanes of a particle claim its 8 nodes side by side.
    const int count = min(S.changed_count[0], int(P.changed_capacity));
    const int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
    const int BX = int(P.brick[0]), BY = int(P.brick[1]), BZ = int(P.brick[2]);
    const int cpb = BX * BY * BZ;
    const int corner = int(threadIdx.x) & 7, slot = (int(threadIdx.x) >> 3) & 7;
    const int waves = blocks * int(blockDim.x >> 6);
    for (int w = block * int(blockDim.x >> 6) + int(threadIdx.x >> 6); w < count * 27; w += waves) {
      const int base = S.base_node[S.changed[w / 27]];
      const int k = w % 27;
      const int cx = base / dyz + k / 9 - 1, cy = (base % dyz) / dz + (k / 3) % 3 - 1, cz = base % dz + k % 3 - 1;
      if (cx < 0 || cy < 0 || cz < 0 || cx >= int(P.bricks[0]) * BX || cy >= int(P.bricks[1]) * BY ||
          cz >= int(P.bricks[2]) * BZ)
        continue;
      const int64_t brick = (int64_t(cx / BX) * P.bricks[1] + cy / BY) * P.bricks[2] + cz / BZ;
      const int64_t cell = brick * cpb + ((cx % BX) * BY + cy % BY) * BZ + cz % BZ;
      const int32_t q = S.perm[8 * cell + slot];
      if (q < 0) continue;
      // The particle's first lane keeps its earlier slots before any lane writes a new one.
      if (corner == 0) keep_previous_slots(S, q);
      const int f = claim_slot(P, S, corner_node(S.base_node[q], corner, dyz, dz), S.transfer_codes[q]);
      S.field_idx[8 * int64_t(q) + corner] = uint8_t(f);
    }
    return;
  }
  // Every particle registers in the pair kernel's particle threads.  These workgroups, which
  // the launch dispatches first, let one anchor particle per cell with even coordinates in its
  // brick (its first particle) claim its code at its 8 nodes.  With even brick sides every
  // node is a corner of exactly one such cell, so the claims meet no contention, and most
  // particles find their code in place, reading it instead of competing for the same slots.
  // A particle that registers twice finds its own slots.
  const int ax = int((P.brick[0] + 1) / 2), ay = int((P.brick[1] + 1) / 2), az = int((P.brick[2] + 1) / 2);
  const int per = ax * ay * az;
  const int64_t nbricks = P.active_bricks > 0 ? P.active_bricks : P.bricks[0] * P.bricks[1] * P.bricks[2];
  const int64_t cpb = P.brick[0] * P.brick[1] * P.brick[2];
  for (int64_t a = first; a < nbricks * per; a += stride) {
    int b = active_brick(S, int(a / per));
    int k = int(a % per);
    int cz = 2 * (k % az), cy = 2 * ((k / az) % ay), cx = 2 * (k / (az * ay));
    int64_t cell = int64_t(b) * cpb + (cx * P.brick[1] + cy) * P.brick[2] + cz;
    const int32_t* slots = S.perm + 8 * cell;
    for (int s = 0; s < 8; ++s) {
      int32_t q = slots[s];
      if (q >= 0) {
        insert_particle(P, S, q);
        break;
      }
    }
  }
}

// The contact kernel's block sums reduced in block order and applied (finish_rigid_step), for
// a standalone rigid-contact call.  One workgroup.
constexpr int kPrepBlock = 1024;  // 16 waves reduce the rows

__global__ void __launch_bounds__(kPrepBlock) k_prep(const bm_params P, bm_state S, int nblocks, int mode, ToolTarget tt) {
  __shared__ float part[kPrepBlock / 64][kRigidSums];
  reduce_rigid_rows(S.rigid_partial, nblocks, part);
  finish_rigid_step(P, S, part[0], mode, tt);
}

// The step sums of knives 1 .. tools-1 from extra_partial (nblocks blocks each, in block order):
// stored as their reaction and torque (added with kAccumulate), their work added to the shared
// work; with kToolStep each of them then takes its tool.step (tool.py:146-259).  One workgroup.
__global__ void __launch_bounds__(kPrepBlock) k_extra_tools(const bm_params P, bm_state S, int nblocks, int mode,
                                                            ToolTarget tt) {
  __shared__ float part[kPrepBlock / 64][kExtraSums];
  for (int k = 1; k < int(P.tools); ++k) {
    float acc[kExtraSums - 1];
#pragma unroll
    for (int i = 0; i < kExtraSums - 1; ++i) acc[i] = 0.0f;
    for (int b = int(threadIdx.x); b < nblocks; b += int(blockDim.x)) {
      const float* src = S.extra_partial + (int64_t(k - 1) * nblocks + b) * kExtraSums;
#pragma unroll
      for (int i = 0; i < kExtraSums - 1; ++i) acc[i] += src[i];
    }
#pragma unroll
    for (int i = 0; i < kExtraSums - 1; ++i) acc[i] = wave_sum(acc[i]);
    if (lane0()) {
#pragma unroll
      for (int i = 0; i < kExtraSums - 1; ++i) part[threadIdx.x >> 6][i] = acc[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float t[kExtraSums - 1];
      for (int i = 0; i < kExtraSums - 1; ++i) {
        t[i] = part[0][i];
        for (int w = 1; w < int(blockDim.x >> 6); ++w) t[i] += part[w][i];
      }
      f3 react = f3{t[0], t[1], t[2]}, torq = f3{t[3], t[4], t[5]};
      if (!(mode & kOverwrite)) {
        react = react + ld3(S.reaction, k);
        torq = torq + ld3(S.torque, k);
      }
      st3(S.reaction, k, react);
      st3(S.torque, k, torq);
      S.work[0] += t[6];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && (mode & kToolStep))
    for (int k = 1; k < int(P.tools); ++k) tool_step_dev(P, S, tt.targets, tt.rotations, tt.index, k);
}

__global__ void k_clear_step(const bm_params P, bm_state S) {
  for (int i = 0; i < 3 * max(int(P.tools), 1); ++i) {
    S.reaction[i] = 0.0f;
    S.torque[i] = 0.0f;
  }
  for (int i = 0; i < 3; ++i) S.board[i] = 0.0f;
  S.speed[0] = 0.0f;
}

// ------------------------------------------------------------------------ pair contacts
__device__ __forceinline__ uint64_t mix_key(int64_t key) { return uint64_t(key) * 2654435761ull; }

// Insert into / find in an open-addressing int64 table with linear probing
// (material.py:37-47, contact.py:569-580).
__device__ __forceinline__ int64_t table_insert(int64_t* keys, int64_t cap, int64_t key) {
  int64_t slot = int64_t(mix_key(key) & uint64_t(cap - 1));
  for (int64_t probe = 0; probe < cap; ++probe) {
    unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(keys + slot),
                                        (unsigned long long)(-1LL), (unsigned long long)key);
    if (int64_t(prev) == -1 || int64_t(prev) == key) return slot;
    slot = (slot + 1) & (cap - 1);
  }
  return -1;
}
__device__ __forceinline__ int64_t table_find(const int64_t* keys, int64_t cap, int64_t key) {
  int64_t slot = int64_t(mix_key(key) & uint64_t(cap - 1));
  for (int64_t probe = 0; probe < cap; ++probe) {
    int64_t k = keys[slot];
    if (k == -1) return -1;
    if (k == key) return slot;
    slot = (slot + 1) & (cap - 1);
  }
  return -1;
}

struct Cell {
  int x, y, z;
};
// Cell of a point: floor(x / size), taken as floor(x * (1 / size)) so the candidate checks of
// a search cost multiplies instead of divisions.  Binning and every search use this one
// function with the same reciprocal, so they agree on every particle's cell.
__device__ __forceinline__ Cell cell_of(f3 x, float inv) {
  return Cell{int(floorf(x.x * inv)), int(floorf(x.y * inv)), int(floorf(x.z * inv))};
}
// Bucket of a cell: its linear index in a 2^13 x 2^7 x 2^6-cell frame (z fastest), wrapped
// into the table.  Neighbouring cells fall into neighbouring buckets, so the bucket heads a
// search reads are a few runs of adjacent words, shared by the particles next to it in memory
// (a mixing hash scattered every head read of a rebuild: 610 -> 340 us for a full rebuild at 1M
// particles).  Cells that wrap onto the same bucket share its chain; the searches check every
// particle's cell.
__device__ __forceinline__ int bucket_of(Cell c, int mask) {
  return int((uint32_t(c.x) << 13) + (uint32_t(c.y) << 6) + uint32_t(c.z)) & mask;
}
__device__ __forceinline__ int bucket_of_point(f3 x, float inv, int mask) { return bucket_of(cell_of(x, inv), mask); }
// The cell list bins particles in cells of the list's base reach 2 r + skin; a rebuild
// whose reach is larger (fast particles) searches ceil(reach / cell) rings of cells.
__device__ __forceinline__ float list_cell_size(const bm_params& P) { return 2.0f * float(P.radius) + float(P.pair_skin); }
__device__ __forceinline__ float list_cell_inv(const bm_params& P) { return 1.0f / list_cell_size(P); }
// set_pair_reach (contact.py:749-761): the search reach of a list built at the largest
// particle speed `speed`.
__device__ __forceinline__ float list_reach_of(const bm_params& P, float speed) {
  const float r = float(P.radius), skin = float(P.pair_skin), dt = float(P.dt);
  float value = 2.0f * r + skin;
  if (dt * speed > 0.5f * skin) value += 2.0f * dt * speed;
  return value;
}
// At most 4 rings: only a diverged run (non-finite or absurd speeds) asks for more, and its
// contacts are void anyway (the run fails its state checks).
__device__ __forceinline__ int list_rings(const bm_params& P, float reach) {
  float r = reach / list_cell_size(P);
  if (!(r <= 4.0f)) r = 4.0f;
  return max(1, int(ceilf(r)));
}
// Distance along one axis from coordinate x (in cell `home`) to the cell home + d.
__device__ __forceinline__ float axis_gap(float x, int home, int d, float cell) {
  float g = d > 0 ? float(home + d) * cell - x : (d < 0 ? x - float(home + d + 1) * cell : 0.0f);
  return fmaxf(g, 0.0f);
}
// A cell of the search cube around x can hold a particle within `reach` of x only if its box
// is nearer than reach; the others are skipped.  The bound carries a margin of 1e-3 cells
// over the rounding of the box faces, so no particle within reach is skipped.
struct CellCull {
  f3 x;
  Cell home;
  float cell, limit2;
  __device__ __forceinline__ CellCull(f3 xp, Cell h, float c, float reach) : x(xp), home(h), cell(c) {
    float r = reach + 1.0e-3f * c;
    limit2 = r * r;
  }
  __device__ __forceinline__ bool near(int dx, int dy, int dz) const {
    float gx = axis_gap(x.x, home.x, dx, cell), gy = axis_gap(x.y, home.y, dy, cell), gz = axis_gap(x.z, home.z, dz, cell);
    return gx * gx + gy * gy + gz * gz <= limit2;
  }
};

__global__ void k_clear_pair_slots(bm_state S, int64_t P, int parity_new) {
  int count = S.pair_count[parity_new];
  int64_t* keys = S.pair_keys + int64_t(parity_new) * P;
  const int32_t* used = S.pair_used + int64_t(parity_new) * P;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < count; a += gridDim.x * blockDim.x) keys[used[a]] = -1;
}

// bin_particles (contact.py:764-775) into bucket lists without a clearing pass: a head
// written by an earlier build carries an older tag and reads as an empty bucket.
__device__ __forceinline__ void bin_particle(const bm_params& P, const bm_state& S, int64_t p, f3 xp,
                                             const ListRegion& L) {
  int tag = L.tag;
  int b = L.base + bucket_of_point(xp, list_cell_inv(P), L.mask);
  unsigned long long entry = (uint64_t(uint32_t(tag)) << 32) | uint64_t(uint32_t(p));
  unsigned long long old = atomicExch(reinterpret_cast<unsigned long long*>(S.list_heads + b), entry);
  S.list_links[p] = (uint32_t(old >> 32) == uint32_t(tag)) ? int32_t(uint32_t(old)) : -1;
}
__device__ __forceinline__ int32_t bucket_head(const bm_state& S, int b, int tag) {
  uint64_t e = uint64_t(S.list_heads[b]);
  return (uint32_t(e >> 32) == uint32_t(tag)) ? int32_t(uint32_t(e)) : -1;
}

// Standalone set_pair_reach + binning (contact.py:749-775) for bm_pair_contacts.
__global__ void k_list_bin(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  ListRegion L = list_region(P, S, instance_of(S, p));
  if (L.bin) bin_particle(P, S, p, ld3(S.x, p), L);
}
__global__ void k_list_reach(const bm_params P, bm_state S, int parity_new) {
  // A standalone pair call takes the reach of the speed of the last rigid-contact call.
  const float speed = fmaxf(S.speed[0], 0.0f);
  S.list_reach[0] = list_reach_of(P, speed);
  S.counters[0] = __float_as_uint(speed);
  S.counters[1 + parity_new] = 0u;
}

// The step's pair-list bookkeeping once its pair kernel is done (every reader of these flags
// and counts has finished): the build counts advance for the lists built this step, whose
// reach is kept; the insertion counts become the tables' counts; the rebuild requests, the
// relabelled list and the step's speed are reset.  Every thread of one workgroup calls it.
__device__ void finish_pair_step(const bm_params& P, const bm_state& S) {
  const bool all = P.pair_mode == kPairHash || S.pair_dirty[0] != 0;
  int built = all ? 1 : 0;
  if (S.instance != nullptr) {
    for (int i = int(threadIdx.x); i < int(P.instances); i += int(blockDim.x)) {
      if (all || S.inst_pair_dirty[i] != 0) {
        S.inst_pair_builds[i] += 1;
        built = 1;
      }
      S.inst_pair_dirty[i] = 0;
    }
    built = __syncthreads_or(built);
  }
  if (threadIdx.x != 0) return;
  if (built) {
    S.pair_builds[0] += 1;
    S.list_reach[0] = list_reach_of(P, __uint_as_float(S.counters[0]));
  }
  if (P.mover_capacity > 0) {
    // A full build bins every particle afresh: none is displaced any more (this step's marks
    // included).  Otherwise the step's movers were refreshed (refresh_state[2] counts such
    // steps).
    if (all) {
      S.refresh_state[1] = 0;
      S.refresh_state[3] = refresh_time(S);
    } else if (S.refresh_state[0] > 0) {
      S.refresh_state[2] += 1;
    }
    S.refresh_state[0] = 0;
  }
  S.pair_count[0] = int32_t(S.counters[1]);
  S.pair_count[1] = int32_t(S.counters[2]);
  S.counters[0] = 0u;
  S.pair_dirty[0] = 0;
  S.changed_count[0] = 0;
  S.pair_epoch[0] += 1;
}

struct PairCtx {
  const int64_t* old_keys;
  const float* old_hist;
  int64_t* new_keys;
  float* new_hist;
  int32_t* new_used;
  int32_t* new_count;
};

// pair_at (contact.py:584-640)
__device__ void pair_at(const bm_params& P, const bm_state& S, const PairCtx& T, int64_t p, int64_t q) {
  float dt = float(P.dt), r = float(P.radius);
  f3 xp = ld3(S.x, p), xq = ld3(S.x, q);
  f3 vp = ld3(S.v, p), vq = ld3(S.v, q);
  f3 delta = xp - xq;
  float dist = len(delta);
  if (dist >= 2.0f * r + dt * len(vp - vq)) return;
  if (compatible(S.codes[p], S.codes[q]) && len(ld3(S.ref, p) - ld3(S.ref, q)) < 2.0f * float(P.spacing)) return;
  if (dist < 1.0e-12f) {
    atomicAdd(S.errors + 2, 1);
    return;
  }
  f3 normal = delta / dist;
  f3 relative = vp - vq;
  float predicted = 2.0f * r - dist - dt * fminf(dot(relative, normal), 0.0f);
  if (predicted <= 0.0f) return;
  int64_t key = p * P.n + q;
  int64_t old = table_find(T.old_keys, P.pair_capacity, key);
  f3 history = f3{0.f, 0.f, 0.f};
  if (old >= 0) history = ld3(T.old_hist, old);
  int64_t slot = table_insert(T.new_keys, P.pair_capacity, key);
  if (slot < 0) {
    atomicAdd(S.errors + 0, 1);
    return;
  }
  float half_mass = float(P.mass) / 2.0f;
  float jn = float(P.sn) * half_mass / dt * predicted;
  Contact c{dt, float(P.sn), float(P.mus), float(P.mud), float(P.gk), float(P.gc)};
  f3 jt, ht;
  friction(c, normal, relative, jn, history, half_mass, jt, ht);
  st3(T.new_hist, slot, ht);
  T.new_used[atomicAdd(T.new_count, 1)] = int32_t(slot);
  f3 impulse = jn * normal + jt;
  atomic_add3(S.impulses, p, impulse);
  atomic_add3(S.impulses, q, -impulse);
}

// Particle-pair contacts of particle p with partners q > p.  When the list is stale, p first
// rebuilds its own partner list (build_pair_list, contact.py:778-822) from the bucket lists
// binned by the contact kernel: partners within reach that pair_at does not skip as
// compatible reference neighbours, at most `lanes`, else -1 and a bucket search at every
// step (pairs_listed, contact.py:825-925).
// In a substep the pair kernel's last workgroups do other work of the step, which needs the
// contact kernel's results and must end before the transfers: after the `particles`
// workgroups of the pair contacts, `reduce` workgroups sum the contact kernel's `rows` block
// sums into rows `rows` .. `rows + reduce - 1` of rigid_partial (for the step's finish in the
// node kernel), then `insert` workgroups register the field slots after a label change.
struct PairExtras {
  int particles, reduce, insert, rows;
};
constexpr int kReduceBlocks = 16;

// pair_at for the pair of a and q in index order (the order of the keys and the impulse).
__device__ __forceinline__ void pair_ordered(const bm_params& P, const bm_state& S, const PairCtx& T, int32_t a,
                                             int32_t q) {
  pair_at(P, S, T, min(a, q), max(a, q));
}

__global__ void __launch_bounds__(kBlock) k_pairs(const bm_params P, bm_state S, int parity_old, int topology,
                                                  PairExtras X) {
  if (int(blockIdx.x) >= X.particles) {
    const int block = int(blockIdx.x) - X.particles;
    if (block < X.reduce) {
      __shared__ float part[kBlock / 64][kRigidSums];
      const int chunk = (X.rows + X.reduce - 1) / X.reduce;
      const int first = block * chunk;
      reduce_rigid_rows(S.rigid_partial + int64_t(first) * kRigidSums, max(0, min(chunk, X.rows - first)), part);
      if (threadIdx.x < kRigidSums)
        S.rigid_partial[int64_t(X.rows + block) * kRigidSums + threadIdx.x] = part[0][threadIdx.x];
    } else {
      register_changed(P, S, block - X.reduce, X.insert);
    }
    return;
  }
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool live = p < P.n;
  const int32_t listed = live ? S.partner_count[p] : 0;
  if (P.pair_mode == kPairList) {
    // The relabelled particles' build codes (their new pairs are in the lists already).
    const int count = min(S.changed_count[0], int(P.changed_capacity));
    for (int64_t i = p; i < count; i += int64_t(X.particles) * blockDim.x) {
      const int32_t d = S.changed[i];
      S.built_codes[d] = S.codes[d];
    }
  }
  // This step's largest particle speed, for lists rebuilt or refreshed this step: the maximum
  // of the contact kernel's block maxima (rows of rigid_partial), taken by each workgroup that
  // rebuilds, which keeps it in counters[0] for the step's bookkeeping.
  // A standalone pair call (no rows) finds it in counters[0].
  float build_speed = 0.0f;
  if (P.pair_mode != kPairAll) {
    // Uniform over a single solver's particles; batched instances rebuild their own lists.
    const int build = S.instance == nullptr ? int(list_region(P, S, 0).bin)
                                            : __syncthreads_or(live && list_region(P, S, instance_of(S, p)).build);
    if (build && X.rows > 0) {
      __shared__ float wave_top[kBlock / 64];
      float top = 0.0f;
      for (int b = int(threadIdx.x); b < X.rows; b += int(blockDim.x))
        top = fmaxf(top, S.rigid_partial[int64_t(b) * kRigidSums + 10]);
      top = wave_max(top);
      if (lane0()) wave_top[threadIdx.x >> 6] = top;
      __syncthreads();
      build_speed = wave_top[0];
      for (int w = 1; w < int(blockDim.x >> 6); ++w) build_speed = fmaxf(build_speed, wave_top[w]);
      if (threadIdx.x == 0) S.counters[0] = __float_as_uint(build_speed);  // the same value in every such block
    } else if (build) {
      build_speed = __uint_as_float(S.counters[0]);
    }
  }
  if (!live) return;
  // Registration of every particle (after the anchors of the first workgroups).
  if (topology && (S.topology_dirty[0] & kTopoFull) != 0) insert_particle(P, S, p);
  // The new table's insertion count (copied to pair_count once the kernel is done).
  int32_t* const filled = reinterpret_cast<int32_t*>(S.counters + 2 - parity_old);
  if (P.pair_mode == kPairAll) {
    // pairs_all (contact.py:929-993): every pair q > p, no lists.
    int parity_new = 1 - parity_old;
    int64_t cap = P.pair_capacity;
    PairCtx T{S.pair_keys + parity_old * cap, S.pair_history + parity_old * cap * 3, S.pair_keys + parity_new * cap,
              S.pair_history + parity_new * cap * 3, S.pair_used + parity_new * cap, filled};
    for (int64_t q = p + 1; q < P.n; ++q) pair_at(P, S, T, p, q);
    return;
  }
  ListRegion L = list_region(P, S, instance_of(S, p));
  int tag = L.tag;
  // A list rebuilt this step takes the reach of this step's largest speed (set_pair_reach,
  // contact.py:749-761); the others keep the reach of their build.
  float reach = L.build ? list_reach_of(P, build_speed) : S.list_reach[0];
  const float reach2 = reach * reach;
  // Displaced particles (refreshed lists): lists skip the pairs that another list owns, and
  // searches take displaced particles from their list.
  const bool displaced_lists =
      S.instance == nullptr && P.mover_capacity > 0 && !L.build && S.refresh_state[1] > 0;
  const int built = displaced_lists ? S.refresh_state[3] : 0;
  float cell = list_cell_size(P), inv = list_cell_inv(P);
  int rings = list_rings(P, reach);
  int base = L.base, mask = L.mask;
  int lanes = int(P.list_lanes);
  int count;
  bool dirty = L.build;
  // A rebuild uses the current positions.  Relabelled particles do not rebuild: the contact
  // kernel's first workgroups added their new pairs to the lists (add_relabelled_pairs); their
  // build codes are brought up to date at the start of this kernel.
  const bool rebuild = dirty;
  const float* bx = dirty ? S.x : S.built_x;
  if (rebuild) {
    f3 xp = ld3(bx, p);
    if (dirty) st3(S.built_x, p, xp);
    int64_t cp = S.codes[p];
    S.built_codes[p] = cp;
    f3 rp = ld3(S.ref, p);
    float two_h = 2.0f * float(P.spacing);
    Cell home = cell_of(xp, inv);
    count = 0;
    // The bucket chains of up to 9 cells are walked in lockstep, so each step of the walk has
    // up to 9 independent loads in flight instead of one.  A lane skips the cells of the
    // search cube too far away to hold a partner.
    constexpr int kChains = 9;
    CellCull cull(xp, home, cell, reach);
    int side = 2 * rings + 1, nyz = side * side;
    // With at least 2^17 buckets, two cells share a bucket only when they are at least 16
    // cells apart along some axis (bucket_of is the linear index x 2^13 + y 2^6 + z, wrapped).
    // The cells of the search cube (at most 9 per axis) then fall into distinct buckets, so
    // each particle met is in one searched cell's chain only, and a particle of another cell
    // whose bucket aliases a searched one lies beyond the reach: the cell of each particle
    // met is not checked.
    const bool distinct = mask >= (1 << 17) - 1;
    for (int dx = -rings; dx <= rings; ++dx) {
      for (int g0 = 0; g0 < nyz; g0 += kChains) {
        int32_t cur[kChains];
        Cell cc[kChains];
#pragma unroll
        for (int j = 0; j < kChains; ++j) {
          int k = g0 + j;
          int dy = k / side - rings, dz = k % side - rings;
          cc[j] = Cell{home.x + dx, home.y + dy, home.z + dz};
          cur[j] = (k < nyz && cull.near(dx, dy, dz)) ? bucket_head(S, base + bucket_of(cc[j], mask), tag) : -1;
        }
        // A chain holds each particle at most once, so every walk ends within n steps; the
        // bound keeps a damaged table from looping.
        bool more = true;
        for (int64_t step = 0; more && step < P.n; ++step) {
          more = false;
#pragma unroll
          for (int j = 0; j < kChains; ++j) {
            int32_t q = cur[j];
            if (q < 0) continue;
            int32_t next = S.list_links[q];
            cur[j] = next;
            more = more || next >= 0;
            if (q <= p || count < 0) continue;
            f3 xq = ld3(bx, q);
            if (!distinct) {
              Cell cq = cell_of(xq, inv);
              if (cq.x != cc[j].x || cq.y != cc[j].y || cq.z != cc[j].z) continue;
            }
            f3 d = xq - xp;
            if (!(dot(d, d) < reach2)) continue;
            if (compatible(cp, S.codes[q]) && len(rp - ld3(S.ref, q)) < two_h) continue;
            if (count == lanes) {
              count = -1;
            } else {
              S.partners[p * lanes + count] = q;
              count += 1;
            }
          }
        }
      }
    }
    S.partner_count[p] = count;
  } else {
    count = listed;
  }
  if (count == 0) return;
  const int rp = displaced_lists ? refresh_rank(S, int32_t(p), built) : 0;
  int parity_new = 1 - parity_old;
  int64_t cap = P.pair_capacity;
  PairCtx T{S.pair_keys + parity_old * cap, S.pair_history + parity_old * cap * 3, S.pair_keys + parity_new * cap,
            S.pair_history + parity_new * cap * 3, S.pair_used + parity_new * cap, filled};
  if (count > 0) {
    for (int k = 0; k < count; ++k) {
      const int32_t q = S.partners[p * lanes + k];
      if (!displaced_lists || owns_pair(rp, int32_t(p), refresh_rank(S, q, built), q))
        pair_ordered(P, S, T, int32_t(p), q);
    }
    return;
  }
  // Overflowing particle: search the buckets of the last build around its build position.
  // In a full build step built_x is being written by this kernel; the build positions are x.
  // A displaced particle owns its pairs with the particles that are not displaced (any index)
  // and some of the displaced ones (from their list); a particle that is not displaced owns
  // none with displaced ones.
  const bool displaced_p = rp > built;
  f3 xb = ld3(bx, p);
  Cell home = cell_of(xb, inv);
  CellCull cull(xb, home, cell, reach);
  for (int dx = -rings; dx <= rings; ++dx)
    for (int dy = -rings; dy <= rings; ++dy)
      for (int dz = -rings; dz <= rings; ++dz) {
        if (!cull.near(dx, dy, dz)) continue;
        Cell c = Cell{home.x + dx, home.y + dy, home.z + dz};
        int64_t step = 0;
        for (int32_t q = bucket_head(S, base + bucket_of(c, mask), tag); q >= 0 && step < P.n;
             q = S.list_links[q], ++step) {
          if (displaced_p ? q == p : q <= p) continue;
          if (displaced_lists && S.refresh_mark[q] > built) continue;
          Cell cq = cell_of(ld3(bx, q), inv);
          if (cq.x != c.x || cq.y != c.y || cq.z != c.z) continue;
          pair_ordered(P, S, T, int32_t(p), q);
        }
      }
  if (displaced_p) {
    const int nd = min(S.refresh_state[1], int(P.mover_capacity));
    for (int i = 0; i < nd; ++i) {
      const int32_t q = S.displaced[i];
      if (q != p && owns_pair(rp, int32_t(p), refresh_rank(S, q, built), q)) pair_ordered(P, S, T, int32_t(p), q);
    }
  }
}
__global__ void k_pairs_done(const bm_params P, bm_state S) { finish_pair_step(P, S); }

// Refresh of mover a's partner list around its current position (its new build position), by
// one workgroup, after the G2P that found it (time: its refresh time, built: the time of the
// last full build).  The list holds the partners q whose pairs a owns once refreshed (every
// particle refreshed before it, and the other movers q > a) that lie within reach + e_q of a
// and are not compatible reference neighbours, where e_q is q's distance from its build
// position (0 for a mover); more than list_lanes: -1 (a search at every step).  A pair left
// out then has build positions at least `reach` apart and, with each particle within half the
// skin of its build position and dt |v| + that distance <= skin / 2 (the staleness check),
// stays beyond 2 r + dt |v_p - v_q| until either is refreshed or rebuilt again (the Verlet
// guarantee of a full build at reach >= 2 r + skin).  The candidates come from the cell lists
// of the last full build, which hold a particle that is not displaced at its build position,
// within half the skin of its current one (search reach + skin, a thread per cell; at least
// 2^17 buckets, so the cells searched have distinct buckets), and from the displaced list (the
// movers included).  group_lds: 1 + list_lanes ints of LDS.
__device__ void refresh_list_of(const bm_params& P, const bm_state& S, int32_t a, int time, int built, float reach,
                                int* group_lds) {
  const ListRegion L = list_region(P, S, 0);
  const int tid = int(threadIdx.x), nt = int(blockDim.x);
  const float two_h = 2.0f * float(P.spacing);
  const f3 xa = ld3(S.x, a), ra = ld3(S.ref, a);
  const int64_t ca = S.codes[a];
  const int lanes = int(P.list_lanes);
  const float search = reach + float(P.pair_skin);
  const float cell = list_cell_size(P), inv = list_cell_inv(P);
  const int rings = list_rings(P, search);
  const Cell home = cell_of(xa, inv);
  CellCull cull(xa, home, cell, search);
  const int side = 2 * rings + 1;
  const int nd = min(S.refresh_state[1], int(P.mover_capacity));
  int* found = group_lds;
  int* list = group_lds + 1;
  if (tid == 0) *found = 0;
  __syncthreads();
  auto consider = [&](int32_t q, int mq, f3 xq) {
    const float eq = mq == time ? 0.0f : len(xq - ld3(S.built_x, q));
    if (!(len(xq - xa) < reach + eq)) return;
    if (compatible(ca, S.codes[q]) && len(ra - ld3(S.ref, q)) < two_h) return;
    const int k = atomicAdd(found, 1);
    if (k < lanes) list[k] = q;
  };
  for (int k = tid; k < side * side * side; k += nt) {
    const int dx = k / (side * side) - rings, dy = (k / side) % side - rings, dz = k % side - rings;
    if (!cull.near(dx, dy, dz)) continue;
    const Cell c = Cell{home.x + dx, home.y + dy, home.z + dz};
    int64_t step = 0;  // a chain holds each particle at most once
    for (int32_t q = bucket_head(S, L.base + bucket_of(c, L.mask), L.now); q >= 0 && step < P.n; ++step) {
      // The next link, the mark and the position of q in one round of loads.
      const int32_t next = S.list_links[q];
      const int mq = S.refresh_mark[q];
      const f3 xq = ld3(S.x, q);
      if (mq <= built) consider(q, mq, xq);  // displaced particles (a included) come from their list
      q = next;
    }
  }
  // The displaced list, four entries per thread at a time, their loads issued together.
  constexpr int kAhead = 4;
  for (int i0 = tid; i0 < nd; i0 += kAhead * nt) {
    int32_t qs[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) qs[u] = i0 + u * nt < nd ? S.displaced[i0 + u * nt] : -1;
    int ms[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) ms[u] = qs[u] >= 0 ? S.refresh_mark[qs[u]] : 0;
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
      if (qs[u] >= 0 && qs[u] != a && owns_pair(time, a, max(ms[u], built), qs[u])) consider(qs[u], ms[u], ld3(S.x, qs[u]));
  }
  __syncthreads();
  const int n = *found;
  const int count = n > lanes ? -1 : n;
  for (int i = tid; i < count; i += nt) S.partners[int64_t(a) * lanes + i] = list[i];
  if (tid == 0) {
    S.partner_count[a] = count;
    S.built_codes[a] = ca;
    st3(S.built_x, a, xa);
  }
  __syncthreads();  // the LDS is taken again for the workgroup's next mover
}

// The movers' list refreshes, one workgroup per mover (workgroup `block` of `blocks`), in the
// first workgroups of the contact kernel after the G2P that found them: they run beside the
// contacts, and the pair kernel that follows uses the new lists.
__device__ void refresh_movers(const bm_params& P, const bm_state& S, int block, int blocks) {
  if (S.instance != nullptr || P.mover_capacity <= 0 || S.pair_dirty[0] != 0 || S.refresh_state[0] == 0) return;
  __shared__ int group_lds[1 + kRefreshLanes];
  const int n = min(S.refresh_state[0], int(P.mover_capacity));
  // A refreshed list keeps the reach of the last build, at least 2 r + skin, which suffices
  // while every mover has dt |v| <= skin / 2 (a faster one rebuilds every list).
  const float reach = S.list_reach[0];
  const int time = refresh_time(S), built = S.refresh_state[3];
  for (int i = block; i < n; i += blocks) refresh_list_of(P, S, S.movers[i], time, built, reach, group_lds);
}

// After particle d's code changed, adds the pairs (d, q) that pair_at no longer skips to the
// partner lists, so the lists need no rebuild: q within the list reach of d at the build
// positions, the pair now a candidate (codes incompatible or references apart).  A pair goes
// to the list of its lower index unless it is there already; when q was relabelled too (its
// code differs from its build code), the walk of the lower of the two adds it.  A list that
// fills up turns into a bucket search (-1).  Pairs that become compatible stay listed and
// pair_at skips them, as with any listed pair.  Candidates only grow by a relabel of the
// knives (a code gains a digit), and the lists keep the Verlet guarantee of the last build.
// One wave per relabelled particle: the lanes take the cells of the search cube (those near
// enough to matter) and walk their bucket chains side by side.
__device__ void add_relabelled_pairs(const bm_params& P, const bm_state& S, int32_t d, int lane, const ListRegion& L,
                                     bool refreshing) {
  const int tag = L.now;  // the cell lists binned before this walk (none is binned meanwhile)
  const float reach = S.list_reach[0];
  const float reach2 = reach * reach, two_h = 2.0f * float(P.spacing);
  const float cell = list_cell_size(P), inv = list_cell_inv(P);
  const int rings = list_rings(P, reach);
  const int lanes = int(P.list_lanes);
  const f3 xd = ld3(S.built_x, d), rd = ld3(S.ref, d);
  const int64_t cd = S.codes[d];
  const Cell home = cell_of(xd, inv);
  CellCull cull(xd, home, cell, reach);
  const int side = 2 * rings + 1;
  // Displaced particles (refresh) sit at stale positions in the cell lists: they are taken
  // from the displaced list (after the cells, lanes over its entries).  A pair goes to the
  // list of its owner.
  const bool displaced_lists = S.instance == nullptr && P.mover_capacity > 0;
  const int nd = displaced_lists ? min(S.refresh_state[1], int(P.mover_capacity)) : 0;
  const int built = displaced_lists ? S.refresh_state[3] : 0;
  const int rank_d = displaced_lists ? refresh_rank(S, d, built) : 0;
  // A mover's list is refreshed beside this walk, with its new pairs.
  const int mover = refreshing && displaced_lists ? refresh_time(S) : -1;
  if (rank_d == mover) return;
  const int cube = side * side * side;
  // Distinct buckets for the cells of a search cube (at least 2^17 buckets, as in the rebuild):
  // a particle met is in one searched chain only, and needs no cell check.
  const bool distinct = L.mask >= (1 << 17) - 1;
  for (int k = lane; k < cube + nd; k += 64) {
    int32_t first = -1;
    Cell c = Cell{0, 0, 0};
    if (k < cube) {
      const int dx = k / (side * side) - rings, dy = (k / side) % side - rings, dz = k % side - rings;
      if (!cull.near(dx, dy, dz)) continue;
      c = Cell{home.x + dx, home.y + dy, home.z + dz};
      first = bucket_head(S, L.base + bucket_of(c, L.mask), tag);
    } else {
      first = S.displaced[k - cube];
    }
    int64_t step = 0;  // a chain holds each particle at most once
    for (int32_t q = first; q >= 0 && step < P.n; q = k < cube ? S.list_links[q] : -1, ++step) {
      if (q == d) continue;
      const int rq = displaced_lists ? refresh_rank(S, q, built) : 0;
      if ((k < cube && rq > built) || rq == mover) continue;
      const f3 xq = ld3(S.built_x, q);
      if (k < cube && !distinct) {
        const Cell cq = cell_of(xq, inv);
        if (cq.x != c.x || cq.y != c.y || cq.z != c.z) continue;
      }
      const f3 dd = xq - xd;
      if (!(dot(dd, dd) < reach2)) continue;
      const int64_t cq_code = S.codes[q];
      if (compatible(cd, cq_code) && len(rd - ld3(S.ref, q)) < two_h) continue;
      if (q < d && cq_code != S.built_codes[q]) continue;  // q's own walk adds the pair
      const bool mine = owns_pair(rank_d, d, rq, q);
      const int32_t owner = mine ? d : q, other = mine ? q : d;
      int32_t* count = S.partner_count + owner;
      const int32_t* list = S.partners + int64_t(owner) * lanes;
      int32_t n = *count;
      bool listed = false;
      for (int i = 0; i < n && i < lanes; ++i) listed = listed || list[i] == other;
      // Append (the other walks append other partners to the same list concurrently).
      while (!listed && n >= 0) {
        const int32_t next = n < lanes ? n + 1 : -1;
        const int32_t seen = atomicCAS(count, n, next);
        if (seen == n) {
          if (next > 0) S.partners[int64_t(owner) * lanes + n] = other;
          break;
        }
        n = seen;
      }
    }
  }
}

// check_pair_list (contact.py:716-735) as its own pass for a standalone pair-contact call.
__global__ void k_check_pair_list(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  float dt = float(P.dt);
  f3 xp = ld3(S.x, p);
  if (len(xp - ld3(S.built_x, p)) + dt * len(ld3(S.v, p)) > 0.5f * float(P.pair_skin) ||
      S.codes[p] != S.built_codes[p]) {
    if (S.instance != nullptr)
      S.inst_pair_dirty[S.instance[p]] = 1;
    else
      S.pair_dirty[0] = 1;
  }
}

// ------------------------------------------------------------------------ topology

__global__ void k_topo_clear(const bm_params P, bm_state S) {
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < active_node_count(P); i += stride) {
    int64_t node = active_node(S, i);
    for (int f = 0; f < int(P.fields); ++f) S.field_code[int64_t(f) * P.node_count + node] = -1;
    S.node_nf[node] = 0;
  }
}
__global__ void k_topo_insert(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < P.n) insert_particle(P, S, p);
}
__global__ void k_topo_done(bm_state S) { S.topology_dirty[0] = 0; }

__device__ __forceinline__ int field_of(uint2 idx, int a) {
  uint32_t w = a < 4 ? idx.x : idx.y;
  return int((w >> (8 * (a & 3))) & 0xffu);
}

// ------------------------------------------------------------------------ material
// Particles are binned into bricks of brick[0] x brick[1] x brick[2] cells (perm lists the
// particles of each brick).  One workgroup per brick accumulates its particles' stencil
// contributions in LDS over the brick's node window (brick + 1 nodes per axis) and writes
// the window to a per-brick partial buffer with plain stores.  The grid kernels gather, for
// every node, the windows of the (at most 8) bricks that contain it in a fixed order: P2G and
// the MUSL re-scatter need no global atomics and their sums are deterministic.
// Division by launch constants without the integer-divide expansion (about 15 VALU per
// quotient on gfx950): floor(n / d) = mulhi(n, m) with m = ceil(2^32 / d) is exact for
// n < 2^16 (Div32), and with m = ceil(2^64 / d) and a 64-bit high product for n < 2^32,
// d < 2^32 (Div64); d = 1 is stored as m = 0.  The host fills them per launch (divisors_of).
struct Div32 {
  uint32_t m, d;
};
struct Div64 {
  uint64_t m;
  uint32_t d;
};
__device__ __forceinline__ int quot(const Div32& v, int n) { return v.m ? int(__umulhi(uint32_t(n), v.m)) : n; }
__device__ __forceinline__ int64_t quot(const Div64& v, int64_t n) {
  return v.m ? int64_t(__umul64hi(uint64_t(n), v.m)) : n;
}
struct Divisors {
  Div32 bx, by, bz;   // brick cells per axis (node coordinate -> brick)
  Div32 bybz;         // cells per brick column: cell index -> x within the brick
  Div32 wz, wywz;     // window nodes along z, per x plane: window index -> (x, y, z)
  Div64 nbz, nbynbz;  // bricks along z, per x plane: brick index -> (x, y, z)
  Div64 dz, dyz;      // grid nodes along z, per x plane: node -> (x, y, z)
};

struct Window {
  int x0, y0, z0;
  int wy, wz, W;
};
__device__ __forceinline__ Window window_of(const bm_params& P, const Divisors& D, int b) {
  int nby = int(P.bricks[1]), nbz = int(P.bricks[2]);
  int qz = int(quot(D.nbz, b)), bx = int(quot(D.nbynbz, b));
  int bz = b - qz * nbz, by = qz - bx * nby;
  Window w;
  w.x0 = bx * int(P.brick[0]);
  w.y0 = by * int(P.brick[1]);
  w.z0 = bz * int(P.brick[2]);
  w.wy = int(P.brick[1]) + 1;
  w.wz = int(P.brick[2]) + 1;
  w.W = (int(P.brick[0]) + 1) * w.wy * w.wz;
  return w;
}
__device__ __forceinline__ int window_size(const bm_params& P) {
  return int((P.brick[0] + 1) * (P.brick[1] + 1) * (P.brick[2] + 1));
}
// Window index of corner 0 of the particle's base cell.
__device__ __forceinline__ int window_base(const bm_params& P, const Window& w, int32_t base) {
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  int cx = base / dyz, rem = base - cx * dyz;
  int cy = rem / dz, cz = rem - cy * dz;
  return ((cx - w.x0) * w.wy + (cy - w.y0)) * w.wz + (cz - w.z0);
}
__device__ __forceinline__ int window_corner(const Window& w, int a) {
  return ((a >> 2) & 1) * w.wy * w.wz + ((a >> 1) & 1) * w.wz + (a & 1);
}
// (x, y, z) of window index i.
__device__ __forceinline__ void window_coords(const Window& w, const Divisors& D, int i, int& ix, int& iy, int& iz) {
  int qz = quot(D.wz, i);
  ix = quot(D.wywz, i);
  iz = i - qz * w.wz;
  iy = qz - ix * w.wy;
}
// Global node of window index i, or -1 outside the grid.
__device__ __forceinline__ int64_t window_node(const bm_params& P, const Window& w, const Divisors& D, int i) {
  int ix, iy, iz;
  window_coords(w, D, i, ix, iy, iz);
  int nx = w.x0 + ix, ny = w.y0 + iy, nz = w.z0 + iz;
  if (nx >= P.dims[0] || ny >= P.dims[1] || nz >= P.dims[2]) return -1;
  return (int64_t(nx) * P.dims[1] + ny) * P.dims[2] + nz;
}
// Largest used field count over the window's nodes: each wave stores the maximum of its
// lanes' counts in wave_nf[wave]; after a barrier, window_max reads the block's maximum.
__device__ __forceinline__ void wave_max_store(int v, int* wave_nf) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) wave_nf[threadIdx.x >> 6] = v;
}
__device__ __forceinline__ int window_max(const int* wave_nf) {
  int m = 0;
  for (int k = 0; k < int(blockDim.x >> 6); ++k) m = max(m, wave_nf[k]);
  return m;
}

// Brick of workgroup blk.  Workgroups go to the 8 XCDs round-robin (blk % 8), each XCD with
// its own L2: give every XCD a contiguous run of bricks, so bricks that share window nodes and
// particle cache lines are processed under the same L2.
__device__ __forceinline__ int brick_of_block(int blk, int nb) {
  int q = nb >> 3, r = nb & 7;
  int xcd = blk & 7, k = blk >> 3;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

// Window index of corner 0 of the c-th cell of a brick (cells z-fastest).
__device__ __forceinline__ int cell_window_base(const bm_params& P, const Window& w, const Divisors& D, int c) {
  int bz = int(P.brick[2]), by = int(P.brick[1]);
  int qz = quot(D.bz, c), lx = quot(D.bybz, c);
  int lz = c - qz * bz, ly = qz - lx * by;
  return (lx * w.wy + ly) * w.wz + lz;
}
__device__ __forceinline__ int cells_per_brick(const bm_params& P) { return int(P.brick[0] * P.brick[1] * P.brick[2]); }

// Lanes of a brick workgroup and the particle slots of its cells.  Slot s of a cell (0..7)
// is 4 bx + 2 by + bz; on a regular lattice (bm_params.regular) the particle in slot s sits
// at local cell coordinates (bx, by, bz) / 2 + 1/4.  Two cells share each row of 16 lanes so
// that the slot's position bits z, y, x are lane bits 0, 1, 3: lane l holds slot
// (l & 3) | (l >> 3 & 1) << 2 of cell 2 (l >> 4) + (l >> 2 & 1), and lanes differing in one
// position bit are one DPP move apart (xor 1 and 2 by quad_perm, xor 8 by row_ror:8).
__device__ __forceinline__ int lane_cell(int l) { return ((l >> 4) << 1) | ((l >> 2) & 1); }
__device__ __forceinline__ int lane_slot(int l) { return (l & 3) | (((l >> 3) & 1) << 2); }

// The lane's value from its partner across slot position bit Axis (0 x, 1 y, 2 z).  Call
// with the whole wavefront active.
template <int Axis>
__device__ __forceinline__ float partner(float v) {
  if constexpr (Axis == 0) return dpp_move<0x128>(v);  // row_ror:8
  else if constexpr (Axis == 1) return dpp_move<0x4E>(v);  // quad_perm [2,3,0,1]
  else return dpp_move<0xB1>(v);  // quad_perm [1,0,3,2]
}

// Sum over the 8 lanes of a cell: every lane of the cell ends with the total.
__device__ __forceinline__ float cell_sum(float v) {
  v += partner<2>(v);
  v += partner<1>(v);
  v += partner<0>(v);
  return v;
}

// Separable trilinear stencil of a regular lattice.  Along one axis the weights of corner bit
// b at the particles of position bit p are phi_b(1/4 + p/2): 3/4 when b = p, 1/4 otherwise,
// and their derivatives are (2 b - 1) / h whatever p.  So the 8 x 8 (corner, particle) weight
// matrix of a cell is a tensor product of 2 x 2 factors, applied one axis at a time by lane
// pairs:
//   phi:  4 x the weighted pair   3 q + q'      (weights; symmetric, so scatter = gather)
//   both: the pair sum            q + q'        (scatter with the derivative: x (2b-1)/h)
//   diff: the pair difference     q - q'        (gather with the derivative:  x (2p-1)/h)
// where q' is the partner's value.  Scattering particle values with phi along all three
// axes leaves 64 x the cell's corner sum of w q in the lane of that corner; gathering corner
// values the same way leaves 64 x sum_a w_a q_a in the lane of each particle.
template <int Axis>
__device__ __forceinline__ float phi(float q) { return 3.0f * q + partner<Axis>(q); }
template <int Axis>
__device__ __forceinline__ float both(float q) { return q + partner<Axis>(q); }
template <int Axis>
__device__ __forceinline__ float diff(float q) { return q - partner<Axis>(q); }
__device__ __forceinline__ float phi3(float q) { return phi<0>(phi<1>(phi<2>(q))); }
// Signs of the derivative at the lane's position bits, over 16 h: the scale of a stage
// sequence with one both/diff and two phi stages.
__device__ __forceinline__ f3 gradient_signs(int lane, float inv_h) {
  float g = inv_h * (1.0f / 16.0f);
  return f3{(lane >> 3) & 1 ? g : -g, (lane >> 1) & 1 ? g : -g, lane & 1 ? g : -g};
}
// Any live particle in the lane's cell, from a wavefront ballot of live lanes.
__device__ __forceinline__ bool cell_live(uint64_t live_mask, int lane) {
  return ((live_mask >> (lane & 48)) & (0x0F0Full << (4 * ((lane >> 2) & 1)))) != 0;
}
// Grid node of corner s of the c-th cell of a brick (the cell must hold a particle, so the
// node exists).
__device__ __forceinline__ int64_t cell_corner_node(const bm_params& P, const Window& w, const Divisors& D, int c,
                                                   int s) {
  int bz = int(P.brick[2]), by = int(P.brick[1]);
  int qz = quot(D.bz, c), lx = quot(D.bybz, c);
  int lz = c - qz * bz, ly = qz - lx * by;
  int64_t dz = P.dims[2], dyz = P.dims[1] * P.dims[2];
  return int64_t(w.x0 + lx) * dyz + int64_t(w.y0 + ly) * dz + (w.z0 + lz) + corner_offset(s, int(dyz), int(dz));
}
// Velocity gradient sum_a v_a grad w_a^T at the lane's particle from the lane's corner
// velocity v (corner = the lane's slot), by the separable stencil; sg from gradient_signs.
__device__ __forceinline__ m3 separable_rate(f3 v, f3 sg) {
  m3 r;
  float vi[3] = {v.x, v.y, v.z};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float a = phi<2>(vi[i]);
    r.a[3 * i] = sg.x * diff<0>(phi<1>(a));
    r.a[3 * i + 1] = sg.y * phi<0>(diff<1>(a));
    r.a[3 * i + 2] = sg.z * phi<0>(phi<1>(diff<2>(vi[i])));
  }
  return r;
}

// Sum over the (up to 8) cells of the brick around window node wi of their corner entries
// cc[k * CS + 8 c + a], in fixed corner order.
template <int K>
__device__ __forceinline__ void gather_cells(const bm_params& P, const Window& w, const Divisors& D, const float* cc,
                                             int CS, int wi, float* out) {
  int bx = int(P.brick[0]), by = int(P.brick[1]), bz = int(P.brick[2]);
  int ix, iy, iz;
  window_coords(w, D, wi, ix, iy, iz);
  // Entry of corner 0 of cell (ix, iy, iz); corner a's cell is offset by a uniform stride.
  int e0 = 8 * __mul24(__mul24(ix, by) + iy, bz) + 8 * iz;
  int sy = 8 * bz, sx = sy * by;
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = 0.0f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int dx = (a >> 2) & 1, dy = (a >> 1) & 1, dz = a & 1;
    int cx = ix - dx, cy = iy - dy, cz = iz - dz;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= bx || cy >= by || cz >= bz) continue;
    int e = e0 - (dx * sx + dy * sy + dz * 8) + a;
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] += cc[k * CS + e];
  }
}

// p2g (material.py:127-152) for one brick: mass, momentum and stress + contact impulse.
// One lane per particle slot, 8 slots per cell (at most 1024 slots per brick; lane_cell and
// lane_slot give a lane's cell and slot).  The lane of slot a of a cell ends with the cell's
// sums at corner a of the 7 values of its particles on field slot 0, and stores them in a
// per-(cell, corner) LDS table with plain stores; a particle on another field slot (a cut
// through the node) adds its values to an LDS node window with atomics.  Each window node
// then gathers its cells' corner sums in fixed order, adds the window, and the brick's
// partial goes out with plain stores.  The cell sums come from one of two paths, chosen per
// wavefront:
//  - separable: on a regular lattice, when none of the wavefront's particles has a field
//    slot other than 0, the corner sums are the separable stencil's lane-pair stages
//    (phi/both), about 100 VALU per lane;
//  - per corner: otherwise each lane forms its 7 values per stencil corner and the cell's
//    lanes sum them (cell_sum).
// When the last substep's F update is pending (update_pending), each lane first updates its
// particle's F and stress (update_F_plastic, material.py:445-475) and uses that stress; on
// the separable path the corner velocities of that update are one gather per lane (the lane's
// corner) and the stencil's gather stages.
// kMx: the Maxwell branch is on (a separate instantiation, so that the elastic kernels keep
// their registers).
template <bool kMx>
__device__ inline m3 finish_update(const bm_params& P, const bm_state& S, int64_t p, const m3& rate,
                                   bool store_stress);
template <bool kMx>
__device__ inline m3 update_particle(const bm_params& P, const bm_state& S, const Divisors& D, int64_t p, f3 lp,
                                     uint2 fidx, bool store_stress);

// Appends a node with several fields to the node kernels' list, once per epoch (many window
// threads of neighbouring bricks see the same node).
__device__ __forceinline__ void note_multi_node(const bm_state& S, int64_t node) {
  const int epoch = S.multi_count[1];
  if (atomicExch(S.multi_stamp + node, epoch) != epoch) S.multi_nodes[atomicAdd(S.multi_count, 1)] = int32_t(node);
}

template <bool kMx>
// Elastic: six wavefronts per SIMD (at most 80 registers; the few spills are cheaper than the
// occupancy they buy), 33.4 -> 32.6 us on the 1M block; the Maxwell branch keeps its registers.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(kMx ? 1 : 6, 8))) k_p2g(const bm_params P, bm_state S, int flags, const Divisors D) {
  extern __shared__ float lds[];
  __shared__ int wave_nf[16];
  int b = active_brick(S, brick_of_block(int(blockIdx.x), int(gridDim.x)));
  // Registration (bit 0) and the pair kernel (bit 1) of this step are complete.
  if ((flags & 1) && blockIdx.x == 0 && threadIdx.x == 0) S.topology_dirty[0] = 0;
  if ((flags & 2) && blockIdx.x == 0) finish_pair_step(P, S);
  Window w = window_of(P, D, b);
  int F = int(P.fields);
  int FW = F * w.W;
  int cpb = cells_per_brick(P);
  int CS = 8 * cpb;
  float* win = lds;          // [7][F][W]
  float* cc = lds + 7 * FW;  // [7][CS]
  int lane = int(threadIdx.x & 63);
  int c = lane_cell(int(threadIdx.x)), a_mine = lane_slot(int(threadIdx.x));
  int64_t p = c < cpb ? int64_t(S.perm[int64_t(b) * CS + 8 * c + a_mine]) : -1;
  bool live = p >= 0;
  int64_t my_node = int(threadIdx.x) < w.W ? window_node(P, w, D, int(threadIdx.x)) : -1;
  int my_nf = my_node >= 0 ? S.node_nf[my_node] : 0;
  if (my_nf >= 2 && S.multi_nodes != nullptr) note_multi_node(S, my_node);
  if (F > 1) {
    for (int k = 0; k < 7; ++k)
      for (int i = threadIdx.x; i < FW; i += blockDim.x) win[k * FW + i] = 0.0f;
    __syncthreads();
  }
  float mass = float(P.mass), sv = float(P.dt) * float(P.volume), inv_h = 1.0f / float(P.spacing);
  float mine[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  // The last substep's F update, when it was deferred to this pass (bm_substep_deferred).
  const int pending = S.update_pending != nullptr ? S.update_pending[0] : 0;
  const uint64_t live_mask = __ballot(live);
  if (live_mask != 0) {
    f3 vp = f3{0.f, 0.f, 0.f}, imp = vp, lp = vp;
    m3 Pk = mzero();
    uint2 fidx = make_uint2(0u, 0u), ufidx = fidx;
    bool update = false;
    if (live) {
      vp = ld3(S.v, p);
      imp = ld3(S.impulses, p);
      fidx = reinterpret_cast<const uint2*>(S.field_idx)[p];
      update = pending != 0;
      // pending 2: a particle marked in update_done registered again this step; its update
      // reads the grid velocities of its earlier slots.  The marks are cleared here.
      ufidx = fidx;
      if (pending == 2 && S.update_done[p] != 0) {
        ufidx = reinterpret_cast<const uint2*>(S.field_idx_prev)[p];
        S.update_done[p] = 0;
      }
    }
    // The separable path needs no local coordinates: the slot gives them.
    const bool separable = P.regular && __ballot(live && (fidx.x | fidx.y | ufidx.x | ufidx.y) != 0u) == 0;
    if (live && !separable) lp = ld3(S.local, p);
    const f3 sg = gradient_signs(lane, inv_h);
    if (separable) {
      if (__ballot(update) != 0) {
        // Corner velocity of the lane's slot in its cell; every lane of an occupied cell
        // gathers, whether or not its own slot holds a particle to update.
        f3 vc = f3{0.f, 0.f, 0.f};
        if (cell_live(live_mask, lane)) vc = ld3(S.grid_vel, cell_corner_node(P, w, D, c, a_mine));
        m3 rate = separable_rate(vc, sg);
        if (update) Pk = finish_update<kMx>(P, S, p, rate, false);
      }
    } else if (update) {
      Pk = update_particle<kMx>(P, S, D, p, lp, ufidx, false);
    }
    if (live) {
      if (!update) Pk = ld9(S.stress, p);
#pragma unroll
      for (int k = 0; k < 9; ++k) Pk.a[k] *= sv;
    }
    if (separable) {
      // mass, momentum: 64 x the corner sums of m w and m w v; impulse: 64 x sum w j minus
      // sum_j (2 b_j - 1) / (16 h) x the stages with the derivative along j of the stress.
      float m64 = mass * (1.0f / 64.0f);
      mine[0] = m64 * phi3(live ? 1.0f : 0.0f);
      mine[1] = m64 * phi3(vp.x);
      mine[2] = m64 * phi3(vp.y);
      mine[3] = m64 * phi3(vp.z);
      float ji[3] = {imp.x, imp.y, imp.z};
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        float gx = both<0>(phi<1>(phi<2>(Pk.a[3 * i])));
        float gy = phi<0>(both<1>(phi<2>(Pk.a[3 * i + 1])));
        float gz = phi<0>(phi<1>(both<2>(Pk.a[3 * i + 2])));
        mine[4 + i] = phi3(ji[i]) * (1.0f / 64.0f) - (sg.x * gx + sg.y * gy + sg.z * gz);
      }
    } else {
      int wb = cell_window_base(P, w, D, c < cpb ? c : 0);
#pragma unroll 2
      for (int a = 0; a < 8; ++a) {
        int f = field_of(fidx, a);
        float wt;
        f3 ga;
        stencil_corner(lp, inv_h, a, wt, ga);
        if (!live) wt = 0.0f;
        float mw = mass * wt;
        f3 jm = wt * imp - mvec(Pk, ga);
        float cv[7] = {mw, mw * vp.x, mw * vp.y, mw * vp.z, jm.x, jm.y, jm.z};
        bool own = live && f != 0;
        if (own) {
          int idx = __mul24(f, w.W) + wb + window_corner(w, a);
#pragma unroll
          for (int k = 0; k < 7; ++k) atomicAdd(win + k * FW + idx, cv[k]);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
          float t = cell_sum(own ? 0.0f : cv[k]);
          if (a_mine == a) mine[k] = t;
        }
      }
    }
  }
  if (c < cpb) {
#pragma unroll
    for (int k = 0; k < 7; ++k) cc[k * CS + 8 * c + a_mine] = mine[k];
  }
  wave_max_store(my_nf, wave_nf);
  __syncthreads();
  int Fb = window_max(wave_nf);
  float* out = S.p2g_partial + int64_t(b) * FW * 7;
  for (int i = threadIdx.x; i < Fb * w.W; i += blockDim.x) {
    float sum[7];
    if (i < w.W) {
      gather_cells<7>(P, w, D, cc, CS, i, sum);
    } else {
#pragma unroll
      for (int k = 0; k < 7; ++k) sum[k] = 0.0f;
    }
    if (F > 1) {
#pragma unroll
      for (int k = 0; k < 7; ++k) sum[k] += win[k * FW + i];
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) out[int64_t(i) * 7 + k] = sum[k];
  }
}
// Bricks along one axis whose window holds node coordinate c: c / B, and c / B - 1 when c
// is on a brick boundary.  The stores use fixed indices (out[0] is c / B when that is a
// brick, else c / B - 1) so that the caller's arrays stay in registers.
__device__ __forceinline__ int bricks_at(int c, const Div32& B, int nb, int* out) {
  int b0 = quot(B, c);
  bool inside = b0 < nb, below = c == b0 * int(B.d) && b0 > 0;
  out[0] = inside ? b0 : b0 - 1;
  out[1] = b0 - 1;
  return int(inside) + int(below);
}

// Sum of the window entries of node (nx, ny, nz), field f, components [0, C) over its (at
// most 2 x 2 x 2) bricks, in brick order.  All loads are issued before the first add.
template <int C>
__device__ __forceinline__ void gather_node(const bm_params& P, const Divisors& D, const float* partial, int nx, int ny,
                                            int nz, int f, float* out) {
  int bx[2], by[2], bz[2];
  int BX = int(P.brick[0]), BY = int(P.brick[1]), BZ = int(P.brick[2]);
  int kx = bricks_at(nx, D.bx, int(P.bricks[0]), bx);
  int ky = bricks_at(ny, D.by, int(P.bricks[1]), by);
  int kz = bricks_at(nz, D.bz, int(P.bricks[2]), bz);
  int wy = BY + 1, wz = BZ + 1, W = (BX + 1) * wy * wz;
  int F = int(P.fields);
  float v[8][C];
  bool ok[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    int i = t >> 2, j = (t >> 1) & 1, l = t & 1;
    ok[t] = i < kx && j < ky && l < kz;
    int64_t b = ok[t] ? (int64_t(bx[i]) * P.bricks[1] + by[j]) * P.bricks[2] + bz[l] : 0;
    int widx = ok[t] ? ((nx - bx[i] * BX) * wy + (ny - by[j] * BY)) * wz + (nz - bz[l] * BZ) : 0;
    const float* src = partial + ((b * F + f) * W + widx) * C;
    if (ok[t]) {
#pragma unroll
      for (int c = 0; c < C; ++c) v[t][c] = src[c];
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) v[t][c] = 0.0f;
    }
  }
#pragma unroll
  for (int c = 0; c < C; ++c) out[c] = 0.0f;
#pragma unroll
  for (int t = 0; t < 8; ++t)
    if (ok[t]) {
#pragma unroll
      for (int c = 0; c < C; ++c) out[c] += v[t][c];
    }
}

__device__ __forceinline__ void node_coords(const bm_params& P, const Divisors& D, int64_t node, int& nx, int& ny,
                                            int& nz) {
  int64_t dz = P.dims[2], dyz = P.dims[1] * P.dims[2];
  nx = int(quot(D.dyz, node));
  int64_t rem = node - int64_t(nx) * dyz;
  ny = int(quot(D.dz, rem));
  nz = int(rem - int64_t(ny) * dz);
}

// Nodes of 2 .. kMultiFields fields are exchanged by the first multi_blocks(P) workgroups of
// each node kernel: four lanes per listed node, one per field.  The lanes load the list entry
// (with the count), then their field's window sums, code and mass in one round, and trade
// them by cross-lane reads within the four.  These workgroups are dispatched first and run
// beside the node threads, which skip such nodes; a node thread that met one would otherwise
// wait for one more round of loads per field, and on the few multi-field nodes of a cut that
// chain, not bandwidth, sets the kernel's duration.
constexpr int kMultiBlocks = 8;
constexpr int kMultiFields = 4;
// Batched solvers, whose many cuts give far more multi-field nodes, take eight times as many
// of these workgroups (eight had cost the batched sweep a quarter of its throughput).
__host__ __device__ __forceinline__ int multi_blocks(const bm_params& P) {
  return P.instances > 1 ? 8 * kMultiBlocks : kMultiBlocks;
}

__device__ __forceinline__ bool multi_listed(const bm_state& S, int nf) {
  return S.multi_nodes != nullptr && nf >= 2 && nf <= kMultiFields;
}
__device__ __forceinline__ int64_t quad_read64(int64_t v, int src) {
  const int lo = __shfl(int(uint32_t(uint64_t(v))), src, 64), hi = __shfl(int(uint64_t(v) >> 32), src, 64);
  return int64_t((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo)));
}

// The exchange of couple_and_velocity (below) for the listed nodes; field f of a node is lane
// f of its four, and the sums run over the fields in the same order as a node thread's.
__device__ void multi_velocity(const bm_params& P, const bm_state& S, const Divisors& D) {
  if (S.multi_nodes == nullptr) return;
  const int f = int(threadIdx.x) & 3, base = int(threadIdx.x) & 60;
  const int64_t N = P.node_count;
  const int stride = multi_blocks(P) * kBlock / 4;
  for (int q = (int(blockIdx.x) * kBlock + int(threadIdx.x)) >> 2;; q += stride) {
    const int count = S.multi_count[0];
    const int64_t node = q < N ? S.multi_nodes[q] : 0;  // in bounds whatever the count
    if (q >= count) return;
    int nx, ny, nz;
    node_coords(P, D, node, nx, ny, nz);
    float g[7];
    int64_t code = 0;
    if (f < int(P.fields)) {
      gather_node<7>(P, D, S.p2g_partial, nx, ny, nz, f, g);
      code = S.field_code[int64_t(f) * N + node];
    } else {
#pragma unroll
      for (int c = 0; c < 7; ++c) g[c] = 0.0f;
    }
    const int nf = S.node_nf[node];
    if (nf < 2 || nf > kMultiFields) continue;  // the same for the four lanes
    float mb[kMultiFields], pb[kMultiFields][3], jb[kMultiFields][3];
    int64_t cb[kMultiFields];
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b) {
      mb[b] = __shfl(g[0], base + b, 64);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        pb[b][k] = __shfl(g[1 + k], base + b, 64);
        jb[b][k] = __shfl(g[4 + k], base + b, 64);
      }
      cb[b] = quad_read64(code, base + b);
    }
    float total = 0.0f;
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b)
      if (b < nf) total += mb[b];
    if (f >= nf) continue;
    const int64_t sa = int64_t(f) * N + node;
    const float ma = g[0];
    S.grid_mass[sa] = ma;
    if (!(ma > 0.0f)) continue;
    f3 pa = f3{g[1], g[2], g[3]}, ja = f3{g[4], g[5], g[6]};
    f3 dm = f3{0.f, 0.f, 0.f}, di = dm;
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b) {
      if (b == f || b >= nf || !compatible(code, cb[b])) continue;
      dm = dm + (ma * f3{pb[b][0], pb[b][1], pb[b][2]} - mb[b] * pa);
      di = di + (ma * f3{jb[b][0], jb[b][1], jb[b][2]} - mb[b] * ja);
    }
    f3 pm = pa + dm / total, jm = ja + di / total;
    st3(S.old_v, sa, pm / ma);
    st3(S.new_v, sa, (pm + jm) / ma);
  }
}

// The exchange of couple_remapped (below) for the listed nodes.
__device__ void multi_remapped(const bm_params& P, const bm_state& S, const Divisors& D) {
  if (S.multi_nodes == nullptr) return;
  const int f = int(threadIdx.x) & 3, base = int(threadIdx.x) & 60;
  const int64_t N = P.node_count;
  const int stride = multi_blocks(P) * kBlock / 4;
  for (int q = (int(blockIdx.x) * kBlock + int(threadIdx.x)) >> 2;; q += stride) {
    const int count = S.multi_count[0];
    const int64_t node = q < N ? S.multi_nodes[q] : 0;
    if (q >= count) return;
    int nx, ny, nz;
    node_coords(P, D, node, nx, ny, nz);
    float g[3], m = 0.0f;
    int64_t code = 0;
    if (f < int(P.fields)) {
      gather_node<3>(P, D, S.remap_partial, nx, ny, nz, f, g);
      m = S.grid_mass[int64_t(f) * N + node];
      code = S.field_code[int64_t(f) * N + node];
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) g[c] = 0.0f;
    }
    const int nf = S.node_nf[node];
    if (nf < 2 || nf > kMultiFields) continue;
    float mb[kMultiFields], xb[kMultiFields][3];
    int64_t cb[kMultiFields];
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b) {
      mb[b] = __shfl(m, base + b, 64);
#pragma unroll
      for (int k = 0; k < 3; ++k) xb[b][k] = __shfl(g[k], base + b, 64);
      cb[b] = quad_read64(code, base + b);
    }
    float total = 0.0f;
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b)
      if (b < nf) total += mb[b];
    if (f >= nf) continue;
    const int64_t sa = int64_t(f) * N + node;
    if (m <= 0.0f) {
      st3(S.grid_vel, sa, f3{0.f, 0.f, 0.f});
      continue;
    }
    f3 xa = f3{g[0], g[1], g[2]};
    f3 ex = f3{0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < kMultiFields; ++b) {
      if (b == f || b >= nf || !compatible(code, cb[b])) continue;
      ex = ex + (m * f3{xb[b][0], xb[b][1], xb[b][2]} - mb[b] * xa);
    }
    st3(S.grid_vel, sa, (xa + ex / total) / m);
  }
}

// Compatible-field exchange of one node (Eq. 8, material.py:252-285) on the gathered P2G
// sums: x_a += sum_{b compatible, b != a} (m_a x_b - m_b x_a) / sum_all m_b; then
// old_v = p / m, new_v = (p + j) / m.

// The step's finish (finish_rigid_step) in a substep, by the node kernel's last workgroup:
// `mode` of the contact sums (-1: none), whose block sums the pair kernel reduced to rows
// `rows` .. `rows + kReduceBlocks - 1` of rigid_partial; tt: the knife's targets.  Nothing
// between the contact kernel and here reads what it writes, and the knife's new pose is
// first read by G2P's labelling.
struct StepFinish {
  int mode, rows;
  ToolTarget tt;
};

__global__ void __launch_bounds__(kBlock) k_couple_velocity(const bm_params P, bm_state S, const Divisors D,
                                                            StepFinish SF) {
  if (SF.mode >= 0 && blockIdx.x == gridDim.x - 1) {
    __shared__ float part[kBlock / 64][kRigidSums];
    reduce_rigid_rows(S.rigid_partial + int64_t(SF.rows) * kRigidSums, kReduceBlocks, part);
    finish_rigid_step(P, S, part[0], SF.mode, SF.tt);
    return;
  }
  if (int(blockIdx.x) < multi_blocks(P)) {
    multi_velocity(P, S, D);
    return;
  }
  int64_t i = int64_t(int(blockIdx.x) - multi_blocks(P)) * blockDim.x + threadIdx.x;
  // Every P2G block has read update_pending by now.
  if (i == 0 && S.update_pending != nullptr) S.update_pending[0] = 0;
  if (i >= active_node_count(P)) return;
  int64_t node = active_node(S, i);
  int64_t N = P.node_count;
  int nx, ny, nz;
  node_coords(P, D, node, nx, ny, nz);
  // Field 0 is gathered in the same round of loads as the node's field count: most nodes
  // have one field and need nothing else.
  float g0[7];
  gather_node<7>(P, D, S.p2g_partial, nx, ny, nz, 0, g0);
  int nf = S.node_nf[node];
  if (nf == 0) return;
  constexpr int kMaxLocal = 4;
  if (nf == 1) {
    const float* g = g0;
    S.grid_mass[node] = g[0];
    if (g[0] > 0.0f) {
      f3 pm = f3{g[1], g[2], g[3]}, j = f3{g[4], g[5], g[6]};
      st3(S.old_v, node, pm / g[0]);
      st3(S.new_v, node, (pm + j) / g[0]);
    }
    return;
  }
  if (multi_listed(S, nf)) return;  // exchanged by the multi-field workgroups
  // Several fields: gather all, then exchange.  Up to kMaxLocal fields stay in registers;
  // more go through the grid arrays.
  if (nf <= kMaxLocal) {
    float g[kMaxLocal][7];
    int64_t code[kMaxLocal];
    float total = 0.0f;
#pragma unroll
    for (int f = 0; f < kMaxLocal; ++f) {
      if (f < nf) {
        if (f == 0) {
#pragma unroll
          for (int c = 0; c < 7; ++c) g[0][c] = g0[c];
        } else {
          gather_node<7>(P, D, S.p2g_partial, nx, ny, nz, f, g[f]);
        }
        code[f] = S.field_code[int64_t(f) * N + node];
        S.grid_mass[int64_t(f) * N + node] = g[f][0];
        total += g[f][0];
      }
    }
#pragma unroll
    for (int a = 0; a < kMaxLocal; ++a) {
      if (a >= nf || !(g[a][0] > 0.0f)) continue;
      float ma = g[a][0];
      f3 pa = f3{g[a][1], g[a][2], g[a][3]}, ja = f3{g[a][4], g[a][5], g[a][6]};
      f3 dm = f3{0.f, 0.f, 0.f}, di = dm;
#pragma unroll
      for (int b = 0; b < kMaxLocal; ++b) {
        if (b == a || b >= nf || !compatible(code[a], code[b])) continue;
        float mb = g[b][0];
        dm = dm + (ma * f3{g[b][1], g[b][2], g[b][3]} - mb * pa);
        di = di + (ma * f3{g[b][4], g[b][5], g[b][6]} - mb * ja);
      }
      f3 pm = pa + dm / total, jm = ja + di / total;
      int64_t sa = int64_t(a) * N + node;
      st3(S.old_v, sa, pm / ma);
      st3(S.new_v, sa, (pm + jm) / ma);
    }
    return;
  }
  float total = 0.0f;
  for (int f = 0; f < nf; ++f) {
    float g[7];
    gather_node<7>(P, D, S.p2g_partial, nx, ny, nz, f, g);
    int64_t s = int64_t(f) * N + node;
    S.grid_mass[s] = g[0];
    st3(S.momentum, s, f3{g[1], g[2], g[3]});
    st3(S.grid_impulse, s, f3{g[4], g[5], g[6]});
    total += g[0];
  }
  for (int a = 0; a < nf; ++a) {
    int64_t sa = int64_t(a) * N + node;
    float ma = S.grid_mass[sa];
    if (ma <= 0.0f) continue;
    int64_t ca = S.field_code[sa];
    f3 pa = ld3(S.momentum, sa), ja = ld3(S.grid_impulse, sa);
    f3 dm = f3{0.f, 0.f, 0.f}, di = dm;
    for (int b = 0; b < nf; ++b) {
      if (b == a) continue;
      int64_t sb = int64_t(b) * N + node;
      if (!compatible(ca, S.field_code[sb])) continue;
      float mb = S.grid_mass[sb];
      dm = dm + (ma * ld3(S.momentum, sb) - mb * pa);
      di = di + (ma * ld3(S.grid_impulse, sb) - mb * ja);
    }
    f3 pm = pa + dm / total, jm = ja + di / total;
    st3(S.old_v, sa, pm / ma);
    st3(S.new_v, sa, (pm + jm) / ma);
  }
}

// viscous_impulse (material.py:321-340) of one particle: with the coupled old velocities of
// its 8 corners, L = (sum_a old_v_a (x) grad w_a) F^-1 and the first Piola stress of a
// Newtonian fluid eta J (L + L^T) F^-T; -dt V0 stress grad w_a goes to each corner's field.
__global__ void __launch_bounds__(kBlock) k_viscous_impulse(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  int32_t base = S.base_node[p];
  int64_t N = P.node_count;
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  float inv_h = 1.0f / float(P.spacing);
  uint2 fidx = reinterpret_cast<const uint2*>(S.field_idx)[p];
  f3 lp = ld3(S.local, p);
  int64_t slot[8];
  f3 g[8];
  m3 rate = mzero();
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    slot[a] = int64_t(field_of(fidx, a)) * N + (base + corner_offset(a, dyz, dz));
    float wt;
    stencil_corner(lp, inv_h, a, wt, g[a]);
    f3 v = ld3(S.old_v, slot[a]);
    rate.a[0] += v.x * g[a].x; rate.a[1] += v.x * g[a].y; rate.a[2] += v.x * g[a].z;
    rate.a[3] += v.y * g[a].x; rate.a[4] += v.y * g[a].y; rate.a[5] += v.y * g[a].z;
    rate.a[6] += v.z * g[a].x; rate.a[7] += v.z * g[a].y; rate.a[8] += v.z * g[a].z;
  }
  m3 F = ld9(S.F, p);
  m3 inv = minv(F);
  m3 L = mmul(rate, inv);
  m3 sym;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) sym.a[3 * i + j] = L.a[3 * i + j] + L.a[3 * j + i];
  m3 stress = mmul_t(sym, inv);
  // -dt V0 eta J folded into one factor.
  float c = -float(P.dt) * float(P.volume) * (float(P.viscosity) * mdet(F));
#pragma unroll
  for (int a = 0; a < 8; ++a) atomic_add3(S.viscous_impulse, slot[a], c * mvec(stress, g[a]));
}

// Couples the viscous impulses of each node's fields (Eq. 8, as the node kernel does for the
// P2G impulse; the exchange is linear, so coupling the two impulses apart equals coupling
// their sum) and adds them to new_v: new_v_a += couple(j_visc)_a / m_a.  Clears the impulses
// for the next substep.
__global__ void __launch_bounds__(kBlock) k_viscous_couple(const bm_params P, bm_state S) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= active_node_count(P)) return;
  int64_t node = active_node(S, i), N = P.node_count;
  int nf = S.node_nf[node];
  float total = 0.0f;
  for (int f = 0; f < nf; ++f) total += S.grid_mass[int64_t(f) * N + node];
  for (int a = 0; a < nf; ++a) {
    int64_t sa = int64_t(a) * N + node;
    float ma = S.grid_mass[sa];
    if (!(ma > 0.0f)) continue;
    int64_t ca = S.field_code[sa];
    f3 ja = ld3(S.viscous_impulse, sa), di = f3{0.f, 0.f, 0.f};
    for (int b = 0; b < nf; ++b) {
      if (b == a) continue;
      int64_t sb = int64_t(b) * N + node;
      if (!compatible(ca, S.field_code[sb])) continue;
      di = di + (ma * ld3(S.viscous_impulse, sb) - S.grid_mass[sb] * ja);
    }
    st3(S.new_v, sa, ld3(S.new_v, sa) + (ja + di / total) / ma);
  }
  for (int f = 0; f < nf; ++f) st3(S.viscous_impulse, int64_t(f) * N + node, f3{0.f, 0.f, 0.f});
}

__device__ void assign_dev(const bm_params& P, const bm_state& S, int64_t p, int64_t digit, f3 xp, f3 vp,
                           int knife = 0);

// g2p_remap (material.py:384-401) for one brick: each particle lane gathers the old and new
// velocities of its 8 corners, updates v and x, and re-scatters m w v through the same
// cell-corner table as k_p2g into remap_partial.  As in k_p2g, a wavefront whose particles
// are all on field slot 0 of a regular lattice gathers one corner per lane and applies the
// separable stencil's stages (phi) for the gather and the re-scatter; otherwise each lane
// reads its 8 corners and the re-scatter sums per corner (cell_sum).  In a substep the updated
// particle is also labelled (assign, contact.py:996-1039) and checked against the pair list's
// skin (check_pair_list, contact.py:716-735) while x and v are in registers.
// Eight wavefronts per SIMD: 56 registers, no spills (the 1024-thread bound alone let it take
// 89): 23.6 -> 22.0 us on the 1M block.
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8, 8))) k_g2p(const bm_params P, bm_state S, int64_t digit, int flags,
                                              const Divisors D) {
  extern __shared__ float lds[];
  __shared__ int wave_nf[16];
  int b = active_brick(S, brick_of_block(int(blockIdx.x), int(gridDim.x)));
  Window w = window_of(P, D, b);
  int F = int(P.fields);
  int FW = F * w.W;
  int cpb = cells_per_brick(P);
  int CS = 8 * cpb;
  float* win = lds;             // [3][F][W]
  float* cc = lds + 3 * FW;     // [3][CS]
  int64_t N = P.node_count;
  int lane = int(threadIdx.x & 63);
  int c = lane_cell(int(threadIdx.x)), a_mine = lane_slot(int(threadIdx.x));
  int64_t p = c < cpb ? int64_t(S.perm[int64_t(b) * CS + 8 * c + a_mine]) : -1;
  bool live = p >= 0;
  f3 lp = f3{0.f, 0.f, 0.f}, xp = lp, vold = lp;
  uint2 fidx = make_uint2(0u, 0u);
  int32_t base = 0;
  if (live) {
    fidx = reinterpret_cast<const uint2*>(S.field_idx)[p];
    xp = ld3(S.x, p);
    vold = ld3(S.v, p);
  }
  const uint64_t live_mask = __ballot(live);
  const bool separable = P.regular && __ballot(live && (fidx.x | fidx.y) != 0u) == 0;
  if (live && !separable) {
    lp = ld3(S.local, p);
    base = S.base_node[p];
  }
  // The separable path's corner velocities: the lane's corner of its cell, field slot 0.
  f3 o_c = f3{0.f, 0.f, 0.f}, n_c = o_c;
  if (separable && cell_live(live_mask, lane)) {
    int64_t node = cell_corner_node(P, w, D, c, a_mine);
    o_c = ld3(S.old_v, node);
    n_c = ld3(S.new_v, node);
  }
  // The window's field count and the cleared cross-field accumulator; the barrier that
  // orders the clearing before the accumulation comes after the velocity gather below, so
  // the waves wait on their global loads and not on each other.
  int my_nf = 0;
  for (int i = threadIdx.x; i < w.W; i += blockDim.x) {
    int64_t node = window_node(P, w, D, i);
    my_nf = max(my_nf, node >= 0 ? S.node_nf[node] : 0);
    for (int f = 0; f < F; ++f) {
      int j = f * w.W + i;
      win[j] = 0.0f;
      win[FW + j] = 0.0f;
      win[2 * FW + j] = 0.0f;
    }
  }
  wave_max_store(my_nf, wave_nf);
  float flip = float(P.flip), dt = float(P.dt), mass = float(P.mass);
  float mine[3] = {0.f, 0.f, 0.f};
  f3 vp = f3{0.f, 0.f, 0.f};
  f3 pic = vp, dv = vp;
  if (separable) {
    if (live_mask != 0) {
      const float s64 = 1.0f / 64.0f;
      pic = s64 * f3{phi3(n_c.x), phi3(n_c.y), phi3(n_c.z)};
      f3 d = n_c - o_c;
      dv = s64 * f3{phi3(d.x), phi3(d.y), phi3(d.z)};
    }
  } else if (live) {
    // The corner velocities come straight from old_v / new_v (node arrays small enough to
    // stay in L2).
    int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
    float wt[8];
    stencil_weights(lp, wt);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      int64_t s = int64_t(field_of(fidx, a)) * N + (base + corner_offset(a, dyz, dz));
      f3 o = ld3(S.old_v, s), nv = ld3(S.new_v, s);
      pic = pic + wt[a] * nv;
      dv = dv + wt[a] * (nv - o);
    }
  }
  if (live) {
    vp = (1.0f - flip) * pic + flip * (vold + dv);
    xp = xp + dt * pic;
    st3(S.v, p, vp);
    st3(S.x, p, xp);
    if (flags & 1) assign_dev(P, S, p, digit, xp, vp);
    if ((flags & 2) && len(xp - ld3(S.built_x, p)) + dt * len(vp) > 0.5f * float(P.pair_skin)) {
      if (S.instance != nullptr) {
        S.inst_pair_dirty[S.instance[p]] = 1;
      } else if (P.mover_capacity > 0) {
        // A mover: its list is refreshed next step (a mover with dt |v| > skin / 2, or too many
        // movers or displaced particles, rebuild every list instead).  It takes the next
        // refresh time and, unless displaced already, joins the displaced list.
        const bool fast = dt * len(vp) > 0.5f * float(P.pair_skin);
        const int k = fast ? int(P.mover_capacity) : atomicAdd(S.refresh_state, 1);
        if (k < P.mover_capacity) {
          S.movers[k] = int32_t(p);
          if (atomicExch(S.refresh_mark + p, refresh_time(S)) <= S.refresh_state[3]) {
            const int j = atomicAdd(S.refresh_state + 1, 1);
            if (j < P.mover_capacity)
              S.displaced[j] = int32_t(p);
            else
              S.pair_dirty[0] = 1;
          }
        } else {
          S.pair_dirty[0] = 1;
        }
      } else {
        S.pair_dirty[0] = 1;
      }
    }
  }
  __syncthreads();
  if (live_mask != 0) {
    if (separable) {
      float m64 = mass * (1.0f / 64.0f);
      mine[0] = m64 * phi3(vp.x);
      mine[1] = m64 * phi3(vp.y);
      mine[2] = m64 * phi3(vp.z);
    } else {
      int wb = cell_window_base(P, w, D, c < cpb ? c : 0);
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        int f = field_of(fidx, a);
        float wt;
        f3 ga;
        stencil_corner(lp, 0.0f, a, wt, ga);
        if (!live) wt = 0.0f;
        f3 r = (mass * wt) * vp;
        bool own = live && f != 0;
        if (own) {
          int idx = __mul24(f, w.W) + wb + window_corner(w, a);
          atomicAdd(win + idx, r.x);
          atomicAdd(win + FW + idx, r.y);
          atomicAdd(win + 2 * FW + idx, r.z);
        }
        float tx = cell_sum(own ? 0.0f : r.x), ty = cell_sum(own ? 0.0f : r.y), tz = cell_sum(own ? 0.0f : r.z);
        if (a_mine == a) {
          mine[0] = tx;
          mine[1] = ty;
          mine[2] = tz;
        }
      }
    }
  }
  if (c < cpb) {
    cc[8 * c + a_mine] = mine[0];
    cc[CS + 8 * c + a_mine] = mine[1];
    cc[2 * CS + 8 * c + a_mine] = mine[2];
  }
  __syncthreads();
  int Fb = window_max(wave_nf);
  float* out = S.remap_partial + int64_t(b) * FW * 3;
  for (int i = threadIdx.x; i < Fb * w.W; i += blockDim.x) {
    float sum[3] = {0.f, 0.f, 0.f};
    if (i < w.W) gather_cells<3>(P, w, D, cc, CS, i, sum);
    sum[0] += win[i];
    sum[1] += win[FW + i];
    sum[2] += win[2 * FW + i];
    out[int64_t(i) * 3] = sum[0];
    out[int64_t(i) * 3 + 1] = sum[1];
    out[int64_t(i) * 3 + 2] = sum[2];
  }
}
__device__ void defer_update(const bm_state& S, int block);

// couple (material.py:155-179) of the gathered re-scattered momentum, then the field
// velocity p / m read by the F update (0 on massless fields, which the update skips).
// With update_blocks > 0 the last update_blocks workgroups instead defer this substep's F
// update to the next P2G (defer_update), so no separate F-update launch follows.
__global__ void __launch_bounds__(kBlock) k_couple_remapped(const bm_params P, bm_state S, const Divisors D,
                                                            int update_blocks) {
  if (int(blockIdx.x) < multi_blocks(P)) {
    multi_remapped(P, S, D);
    return;
  }
  const int node_end = int(gridDim.x) - update_blocks;
  if (int(blockIdx.x) >= node_end) {
    defer_update(S, int(blockIdx.x) - node_end);
    return;
  }
  int64_t i = int64_t(int(blockIdx.x) - multi_blocks(P)) * blockDim.x + threadIdx.x;
  if (i >= active_node_count(P)) return;
  int64_t node = active_node(S, i);
  int64_t N = P.node_count;
  int nx, ny, nz;
  node_coords(P, D, node, nx, ny, nz);
  // Field 0 and its mass are loaded in the same round as the node's field count.
  float g0[3];
  gather_node<3>(P, D, S.remap_partial, nx, ny, nz, 0, g0);
  float m0 = S.grid_mass[node];
  int nf = S.node_nf[node];
  if (nf == 0) return;
  if (nf == 1) {
    const float* g = g0;
    float m = m0;
    f3 vel = f3{0.f, 0.f, 0.f};
    if (m > 0.0f) vel = f3{g[0], g[1], g[2]} / m;
    st3(S.grid_vel, node, vel);
    return;
  }
  if (multi_listed(S, nf)) return;  // exchanged by the multi-field workgroups
  constexpr int kMaxLocal = 4;
  if (nf <= kMaxLocal) {
    float g[kMaxLocal][3], m[kMaxLocal];
    int64_t code[kMaxLocal];
    float total = 0.0f;
#pragma unroll
    for (int f = 0; f < kMaxLocal; ++f) {
      if (f < nf) {
        if (f == 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) g[0][c] = g0[c];
        } else {
          gather_node<3>(P, D, S.remap_partial, nx, ny, nz, f, g[f]);
        }
        m[f] = f == 0 ? m0 : S.grid_mass[int64_t(f) * N + node];
        code[f] = S.field_code[int64_t(f) * N + node];
        total += m[f];
      }
    }
#pragma unroll
    for (int a = 0; a < kMaxLocal; ++a) {
      if (a >= nf) continue;
      int64_t sa = int64_t(a) * N + node;
      float ma = m[a];
      if (ma <= 0.0f) {
        st3(S.grid_vel, sa, f3{0.f, 0.f, 0.f});
        continue;
      }
      f3 xa = f3{g[a][0], g[a][1], g[a][2]};
      f3 ex = f3{0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < kMaxLocal; ++b) {
        if (b == a || b >= nf || !compatible(code[a], code[b])) continue;
        ex = ex + (ma * f3{g[b][0], g[b][1], g[b][2]} - m[b] * xa);
      }
      st3(S.grid_vel, sa, (xa + ex / total) / ma);
    }
    return;
  }
  float total = 0.0f;
  for (int f = 0; f < nf; ++f) {
    float g[3];
    gather_node<3>(P, D, S.remap_partial, nx, ny, nz, f, g);
    int64_t s = int64_t(f) * N + node;
    st3(S.remapped, s, f3{g[0], g[1], g[2]});
    total += S.grid_mass[s];
  }
  for (int a = 0; a < nf; ++a) {
    int64_t sa = int64_t(a) * N + node;
    float ma = S.grid_mass[sa];
    if (ma <= 0.0f) {
      st3(S.grid_vel, sa, f3{0.f, 0.f, 0.f});
      continue;
    }
    int64_t ca = S.field_code[sa];
    f3 xa = ld3(S.remapped, sa);
    f3 ex = f3{0.f, 0.f, 0.f};
    for (int b = 0; b < nf; ++b) {
      if (b == a) continue;
      int64_t sb = int64_t(b) * N + node;
      if (!compatible(ca, S.field_code[sb])) continue;
      ex = ex + (ma * ld3(S.remapped, sb) - S.grid_mass[sb] * xa);
    }
    st3(S.grid_vel, sa, (xa + ex / total) / ma);
  }
}

// Lame constants of particle p: its instance's for batched instances, else the solver's.
__device__ __forceinline__ void lame_of(const bm_params& P, const bm_state& S, int64_t p, float& mu, float& lam) {
  if (S.inst_mu != nullptr) {
    int i = instance_of(S, p);
    mu = S.inst_mu[i];
    lam = S.inst_lam[i];
  } else {
    mu = float(P.mu);
    lam = float(P.lam);
  }
}

// Hencky von Mises return map (material.py:404-431).  Returns false when inverted.
__device__ inline bool plastic_at(const bm_params& P, const bm_state& S, int64_t p, const m3& F, m3& Fp, m3& U,
                                  f3& sg, m3& V, float mu) {
  svd3(mmul(F, minv(Fp)), U, sg, V);
  if (fminf(fminf(sg.x, sg.y), sg.z) <= 0.0f) {
    atomicAdd(S.errors + 1, 1);
    return false;
  }
  f3 e = f3{logf(sg.x), logf(sg.y), logf(sg.z)};
  float mean = (e.x + e.y + e.z) / 3.0f;
  f3 dev = e - f3{mean, mean, mean};
  float norm = len(dev);
  float limit = float(P.yield_stress) / (2.0f * mu) * sqrtf(2.0f / 3.0f);
  if (norm > limit) {
    f3 ep = e - (1.0f - limit / norm) * dev;
    sg = f3{expf(ep.x), expf(ep.y), expf(ep.z)};
    m3 US;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      US.a[3 * i] = U.a[3 * i] * sg.x;
      US.a[3 * i + 1] = U.a[3 * i + 1] * sg.y;
      US.a[3 * i + 2] = U.a[3 * i + 2] * sg.z;
    }
    m3 next_f = mmul_t(US, V);
    Fp = mmul(minv(next_f), F);
    st9(S.Fp, p, Fp);
    if (S.fp_identity != nullptr && !is_identity(Fp)) S.fp_identity[0] = 0;
  }
  return true;
}

// stress_at (material.py:94-112)
__device__ inline m3 elastic_stress(const bm_params& P, const bm_state& S, const m3& F, const m3& Fp, float mu,
                                   float lam) {
  // Fp = I (no plastic flow yet, and always without plasticity): f = F, Fp^-T = I exactly.
  bool plain = is_identity(Fp);
  m3 pinv = plain ? Fp : minv(Fp);
  m3 f = plain ? F : mmul(F, pinv);
  float J = mdet(f);
  if (!(J > 0.0f) || !isfinite(J)) {
    atomicAdd(S.errors + 1, 1);
    return mzero();
  }
  m3 Pk;
  if (corotated_polar(f, J, mu, lam, Pk)) return plain ? Pk : mmul_t(Pk, pinv);
  m3 U, V;
  f3 sg;
  svd3(f, U, sg, V);
  return corotated(U, sg, V, pinv, mu, lam);
}

// stress_at with Fp = I (no plastic flow): the corotated stress of F itself.
__device__ inline m3 elastic_stress_plain(const bm_state& S, const m3& F, float mu, float lam) {
  float J = mdet(F);
  if (!(J > 0.0f) || !isfinite(J)) {
    atomicAdd(S.errors + 1, 1);
    return mzero();
  }
  m3 Pk;
  if (corotated_polar(F, J, mu, lam, Pk)) return Pk;
  m3 U, V, I3 = mzero();
  I3.a[0] = I3.a[4] = I3.a[8] = 1.0f;
  f3 sg;
  svd3(F, U, sg, V);
  return corotated(U, sg, V, I3, mu, lam);
}

__device__ inline m3 stress_of(const bm_params& P, const bm_state& S, int64_t p, const m3& F, m3 Fp) {
  float mu, lam;
  lame_of(P, S, p, mu, lam);
  if (P.plastic) {
    m3 U, V;
    f3 sg;
    if (!plastic_at(P, S, p, F, Fp, U, sg, V, mu)) return mzero();
    return corotated(U, sg, V, minv(Fp), mu, lam);
  }
  return elastic_stress(P, S, F, Fp, mu, lam);
}

// Generalized standard linear solid (maxwell_at, material.py:485-535).  The elastic part
// f = U diag(s) V^T = F Fp^-1 loads two branches in parallel:
//   equilibrium: fixed corotated with mu - mu_m and lam + 2/3 mu_m (so that the instantaneous
//     small-strain shear and bulk moduli are those of mu and lam);
//   Maxwell: f_m = f Fv^-1, Fv = I + viscous[p], Kirchhoff stress 2 mu_m dev(log V_m), with
//     f_m = U_m diag(s_m) V_m^T.
// relax > 0 moves a fraction relax of the deviatoric Hencky strain of f_m into Fv:
// f_m <- f_m W with W = V_m diag(exp(-relax dev)) V_m^T, i.e. Fv <- W^-1 Fv, stored as
// Fv - I += delta + delta (Fv - I), delta = V_m diag(expm1(relax dev)) V_m^T.  The per-step
// change of Fv is far below float32 resolution for long relaxation times, so Fv - I is kept
// in float64.  The first Piola stress is equilibrium + tau_m F^-T.
__device__ inline m3 maxwell_stress(const bm_params& P, const bm_state& S, int64_t p, const m3& F, const m3& Fp,
                                    const m3& U, f3 sg, const m3& V,
                                                       float mu, float lam, float relax) {
  double* Dp = S.viscous + 9 * p;
  double D[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) D[k] = Dp[k];
  m3 Fv, US;
#pragma unroll
  for (int k = 0; k < 9; ++k) Fv.a[k] = float(D[k]) + ((k % 4 == 0) ? 1.0f : 0.0f);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    US.a[3 * i] = U.a[3 * i] * sg.x;
    US.a[3 * i + 1] = U.a[3 * i + 1] * sg.y;
    US.a[3 * i + 2] = U.a[3 * i + 2] * sg.z;
  }
  m3 fm = mmul(mmul_t(US, V), minv(Fv));
  m3 Um, Vm;
  f3 sm;
  svd3(fm, Um, sm, Vm);
  if (fminf(fminf(sm.x, sm.y), sm.z) <= 0.0f) {
    atomicAdd(S.errors + 1, 1);
    return mzero();
  }
  f3 e = f3{logf(sm.x), logf(sm.y), logf(sm.z)};
  float mean = (e.x + e.y + e.z) / 3.0f;
  f3 dev = e - f3{mean, mean, mean};
  if (relax > 0.0f) {
    f3 step = relax * dev;
    dev = dev - step;
    f3 growth = f3{expm1f(step.x), expm1f(step.y), expm1f(step.z)};
    m3 VG;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      VG.a[3 * i] = Vm.a[3 * i] * growth.x;
      VG.a[3 * i + 1] = Vm.a[3 * i + 1] * growth.y;
      VG.a[3 * i + 2] = Vm.a[3 * i + 2] * growth.z;
    }
    m3 delta = mmul_t(VG, Vm);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double dd = double(delta.a[3 * i]) * D[j] + double(delta.a[3 * i + 1]) * D[3 + j] +
                    double(delta.a[3 * i + 2]) * D[6 + j];
        Dp[3 * i + j] = D[3 * i + j] + double(delta.a[3 * i + j]) + dd;
      }
  }
  float mu_m = float(P.mu_maxwell);
  f3 k2 = (2.0f * mu_m) * dev;
  m3 UK;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    UK.a[3 * i] = Um.a[3 * i] * k2.x;
    UK.a[3 * i + 1] = Um.a[3 * i + 1] * k2.y;
    UK.a[3 * i + 2] = Um.a[3 * i + 2] * k2.z;
  }
  m3 kirchhoff = mmul_t(UK, Um);
  m3 eq = corotated(U, sg, V, minv(Fp), mu - mu_m, lam + (2.0f / 3.0f) * mu_m);
  m3 tf = mmul_t(kirchhoff, minv(F));
#pragma unroll
  for (int k = 0; k < 9; ++k) eq.a[k] += tf.a[k];
  return eq;
}

// Stress of particle p with the Maxwell branch (maxwell_stress_at, material.py:538-566):
// with `project` and plasticity, the return map at the instantaneous shear modulus first
// (it stores Fp on flow); relax as in maxwell_stress (0 only evaluates the stress, as
// compute_stress_maxwell, material.py:569-587).
__device__ inline m3 maxwell_update(const bm_params& P, const bm_state& S, int64_t p,
                                                       const m3& F, m3 Fp, bool project, float relax) {
  float mu, lam;
  lame_of(P, S, p, mu, lam);
  m3 U, V;
  f3 sg;
  if (project && P.plastic) {
    if (!plastic_at(P, S, p, F, Fp, U, sg, V, mu)) return mzero();
  } else {
    m3 f = mmul(F, minv(Fp));
    float J = mdet(f);
    if (!(J > 0.0f) || !isfinite(J)) {
      atomicAdd(S.errors + 1, 1);
      return mzero();
    }
    svd3(f, U, sg, V);
  }
  return maxwell_stress(P, S, p, F, Fp, U, sg, V, mu, lam, relax);
}

// F += dt rate for particle p, then the stress; stores F (and Fp after plastic flow) and,
// with store_stress, the stress, and returns the stress.  P2G uses the stress of the particles
// it updates at once and does not store it: the stored stress is read only by a P2G with no
// update pending, and every path to such a P2G (the F-update kernel, early deferred
// updates, refresh_stress) stores it.
template <bool kMx>
__device__ inline m3 finish_update(const bm_params& P, const bm_state& S, int64_t p, const m3& rate,
                                   bool store_stress) {
  // While no plastic flow has happened and the host has not written Fp, every Fp is I.
  bool fp_identity = S.fp_identity != nullptr && S.fp_identity[0] != 0;
  m3 Fm = ld9(S.F, p), Fp;
  if (fp_identity) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Fp.a[k] = (k % 4 == 0) ? 1.0f : 0.0f;
  } else {
    Fp = ld9(S.Fp, p);
  }
  float dt = float(P.dt);
#pragma unroll
  for (int k = 0; k < 9; ++k) Fm.a[k] += dt * rate.a[k];
  st9(S.F, p, Fm);
  m3 stress;
  if constexpr (kMx) {
    stress = maxwell_update(P, S, p, Fm, Fp, true, float(P.maxwell_relax));
  } else if (fp_identity && !P.plastic) {
    float mu, lam;
    lame_of(P, S, p, mu, lam);
    stress = elastic_stress_plain(S, Fm, mu, lam);
  } else {
    stress = stress_of(P, S, p, Fm, Fp);
  }
  if (store_stress) st9(S.stress, p, stress);
  return stress;
}

// update_F_plastic (material.py:445-475) for one particle: F += dt Σ_a v_a ⊗ ∇w_a over its
// 8 corners, then the stress.  The corner velocities of the field slots `fidx` come straight
// from grid_vel (a node array small enough to stay in L2).  Stores F (and Fp after plastic
// flow) and, with store_stress, the stress, and returns the stress.
template <bool kMx>
__device__ inline m3 update_particle(const bm_params& P, const bm_state& S, const Divisors& D, int64_t p, f3 lp,
                                     uint2 fidx, bool store_stress) {
  (void)D;
  int32_t base = S.base_node[p];
  int64_t N = P.node_count;
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  float inv_h = 1.0f / float(P.spacing);
  m3 rate = mzero();
  f3 vel[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) vel[a] = ld3(S.grid_vel, int64_t(field_of(fidx, a)) * N + (base + corner_offset(a, dyz, dz)));
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    float wt;
    f3 g;
    stencil_corner(lp, inv_h, a, wt, g);
    rate.a[0] += vel[a].x * g.x; rate.a[1] += vel[a].x * g.y; rate.a[2] += vel[a].x * g.z;
    rate.a[3] += vel[a].y * g.x; rate.a[4] += vel[a].y * g.y; rate.a[5] += vel[a].y * g.z;
    rate.a[6] += vel[a].z * g.x; rate.a[7] += vel[a].z * g.y; rate.a[8] += vel[a].z * g.z;
  }
  return finish_update<kMx>(P, S, p, rate, store_stress);
}

// The deferred F update of a substep (workgroup `block` of the extra workgroups of
// k_couple_remapped, or of a k_update_F launch after it): the update of every particle is
// left to the next P2G, which updates each particle and uses its stress at once.
// update_pending tells that P2G; it is 2 when this step relabelled particles, so that the next
// registration keeps the earlier field slots of the particles it registers again
// (insert_particle) for their updates.
__device__ void defer_update(const bm_state& S, int block) {
  const int topo = S.topology_dirty[0];
  if (block == 0 && threadIdx.x == 0) S.update_pending[0] = (topo & (kTopoChanged | kTopoFull)) ? 2 : 1;
}

// The F update of a substep, one lane per particle slot of a brick; with `defer`, the
// deferred update above.  The workgroups stride over the bricks (at most kUpdateThreads
// threads in all, in a multiple of 8 workgroups, so a workgroup keeps the bricks of its XCD).
constexpr int kUpdateThreads = 1 << 20;
constexpr int kDeferredUpdateBlocks = 1;
template <bool kMx>
__global__ void __launch_bounds__(1024) k_update_F(const bm_params P, bm_state S, const Divisors D, int defer,
                                                   int nbricks) {
  if (defer) {
    defer_update(S, int(blockIdx.x));
    return;
  }
  int cpb = cells_per_brick(P), CS = 8 * cpb;
  int lane = int(threadIdx.x & 63);
  int c = lane_cell(int(threadIdx.x)), a_mine = lane_slot(int(threadIdx.x));
  const f3 sg = gradient_signs(lane, 1.0f / float(P.spacing));
  // Every lane stays for the whole loop: the separable path's lane-pair stages need the
  // whole wavefront.
  for (int k = int(blockIdx.x); k < nbricks; k += int(gridDim.x)) {
    int b = active_brick(S, brick_of_block(k, nbricks));
    int64_t p = c < cpb ? int64_t(S.perm[int64_t(b) * CS + 8 * c + a_mine]) : -1;
    bool live = p >= 0;
    uint2 fidx = live ? reinterpret_cast<const uint2*>(S.field_idx)[p] : make_uint2(0u, 0u);
    const uint64_t live_mask = __ballot(live);
    if (live_mask == 0) continue;
    if (P.regular && __ballot(live && (fidx.x | fidx.y) != 0u) == 0) {
      f3 vc = f3{0.f, 0.f, 0.f};
      if (cell_live(live_mask, lane)) vc = ld3(S.grid_vel, cell_corner_node(P, window_of(P, D, b), D, c, a_mine));
      m3 rate = separable_rate(vc, sg);
      if (live) finish_update<kMx>(P, S, p, rate, true);
    } else if (live) {
      update_particle<kMx>(P, S, D, p, ld3(S.local, p), fidx, true);
    }
  }
}

__global__ void k_refresh_stress(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  if (P.maxwell) {
    st9(S.stress, p, maxwell_update(P, S, p, ld9(S.F, p), ld9(S.Fp, p), false, 0.0f));
    return;
  }
  float mu, lam;
  lame_of(P, S, p, mu, lam);
  st9(S.stress, p, elastic_stress(P, S, ld9(S.F, p), ld9(S.Fp, p), mu, lam));
}

__global__ void k_plastic_project(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  m3 Fp = ld9(S.Fp, p), U, V;
  f3 sg;
  float mu, lam;
  lame_of(P, S, p, mu, lam);
  plastic_at(P, S, p, ld9(S.F, p), Fp, U, sg, V, mu);
}

// ------------------------------------------------------------------------ tool
__device__ __forceinline__ f3 inv_inertia(quat q, f3 ii, f3 t) {
  f3 l = rotate_inv(q, t);
  return rotate(q, f3{ii.x * l.x, ii.y * l.y, ii.z * l.z});
}
__device__ __forceinline__ f3 blade_vertex(int i, f3 hf, float bevel) {
  int j = i % 5;
  float vx = 0.0f, vz = -hf.z;
  if (j == 1) {
    vx = hf.x;
    vz += bevel;
  } else if (j == 2) {
    vx = hf.x;
    vz = hf.z;
  } else if (j == 3) {
    vx = -hf.x;
    vz = hf.z;
  } else if (j == 4) {
    vx = -hf.x;
    vz += bevel;
  }
  return f3{vx, i >= 5 ? hf.y : -hf.y, vz};
}

// tool.step (tool.py:146-259): drive, gyroscopic term, 6 sweeps x 10 vertices of the
// knife/board velocity constraint with Coulomb friction, integration.
__device__ void tool_step_dev(const bm_params& P, const bm_state& S, const float* targets, const float* rotations,
                              int index, int k) {
  float dt = float(P.dt), mass = float(P.blade_mass), bevel = float(P.bevel);
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  f3 x = ld3(S.tool_pos, k);
  quat q = qnormalize(ldq(S.tool_rot + 4 * k));
  f3 v = ld3(S.tool_vel, k), w = ld3(S.tool_omega, k);
  f3 force = f3{0.f, 0.f, -9.81f * mass};
  f3 servo = f3{0.f, 0.f, 0.f};
  if (P.free_blade == 0) {
    quat tq = ldq(rotations + 4 * index);
    f3 target_center = ld3(targets, index) + rotate(tq, f3{0.f, 0.f, hf.z});
    force = float(P.drive_k) * (target_center - x) - float(P.drive_c) * v;
    float norm = len(force);
    if (norm > float(P.drive_limit)) force = force * (float(P.drive_limit) / norm);
    quat dq = qnormalize(qmul(tq, quat{-q.x, -q.y, -q.z, q.w}));
    if (dq.w < 0.0f) dq = quat{-dq.x, -dq.y, -dq.z, -dq.w};
    f3 vec = qvec(dq);
    float s = len(vec);
    f3 error = 2.0f * vec;
    if (s > 1.0e-8f) error = vec * (2.0f * atan2f(s, dq.w) / s);
    servo = float(P.rot_k) * error - float(P.rot_c) * w;
  }
  v = v + (force * dt + ld3(S.reaction, k)) / mass;
  f3 ii = f3{float(P.inertia_inv[0]), float(P.inertia_inv[1]), float(P.inertia_inv[2])};
  if (P.lock_rotation) {
    ii = f3{0.f, 0.f, 0.f};
    w = f3{0.f, 0.f, 0.f};
  } else {
    f3 lw = rotate_inv(q, w);
    f3 inertia = f3{1.0f / ii.x, 1.0f / ii.y, 1.0f / ii.z};
    f3 gyro = rotate(q, cross(lw, f3{inertia.x * lw.x, inertia.y * lw.y, inertia.z * lw.z}));
    w = w + inv_inertia(q, ii, ld3(S.torque, k) + dt * (servo - gyro));
  }
  f3 normal = f3{0.f, 0.f, 1.f};
  f3 total = f3{0.f, 0.f, 0.f};
  for (int sweep = 0; sweep < 6; ++sweep)
    for (int i = 0; i < 10; ++i) {
      f3 r = rotate(q, blade_vertex(i, hf, bevel));
      f3 corner = x + r;
      if ((P.board_half[0] > 0.0 && fabsf(corner.x) > float(P.board_half[0])) ||
          (P.board_half[1] > 0.0 && fabsf(corner.y) > float(P.board_half[1])))
        continue;
      float gap = corner.z;
      float vn = dot(v + cross(w, r), normal);
      float desired = -gap / dt;
      if (gap < 0.0f) desired *= 0.2f;
      if (vn < desired) {
        f3 rn = cross(r, normal);
        float effective = 1.0f / mass + dot(rn, inv_inertia(q, ii, rn));
        float jn = (desired - vn) / effective;
        f3 impulse = jn * normal;
        v = v + impulse / mass;
        w = w + inv_inertia(q, ii, cross(r, impulse));
        f3 vt = v + cross(w, r);
        vt = vt - dot(vt, normal) * normal;
        float sp = len(vt);
        if (sp > 1.0e-10f) {
          f3 direction = -vt / sp;
          f3 rt = cross(r, direction);
          float wt = 1.0f / mass + dot(rt, inv_inertia(q, ii, rt));
          float jt = fminf(sp / wt, float(P.mud) * jn);
          f3 tangent = jt * direction;
          impulse = impulse + tangent;
          v = v + tangent / mass;
          w = w + inv_inertia(q, ii, cross(r, tangent));
        }
        total = total + impulse;
      }
    }
  x = x + dt * v;
  float sp = len(w);
  if (sp > 1.0e-10f) q = qnormalize(qmul(axis_angle(w / sp, dt * sp), q));
  st3(S.tool_pos, k, x);
  st3(S.tool_vel, k, v);
  S.tool_rot[4 * k] = q.x;
  S.tool_rot[4 * k + 1] = q.y;
  S.tool_rot[4 * k + 2] = q.z;
  S.tool_rot[4 * k + 3] = q.w;
  st3(S.tool_omega, k, w);
  st3(S.table, k, total);
}
// Every knife steps with the same target (free knives ignore it).
__global__ void k_tool_step(const bm_params P, bm_state S, const float* targets, const float* rotations, int index) {
  for (int k = 0; k < max(int(P.tools), 1); ++k) tool_step_dev(P, S, targets, rotations, index, k);
}

// finish_external_step (tool.py:70-115)
// finish_external_step (tool.py:70-115) on values already in registers: interval impulse and
// angular impulse about the world origin; with predict, twist += Mobility [j; tau] and the
// pose is integrated; the step sums are reset.
__device__ void finish_external_regs(const bm_params& P, const bm_state& S, const ExternalTool& e, f3 j, f3 tau,
                                     int i) {
  f3 impulse = e.impulse + j;
  f3 angular = e.angular + (tau + cross(e.pos, j));
  f3 vel = e.vel, w = e.omega, pos = e.pos;
  quat rot = e.rot;
  bool turn = false;
  if (P.predict) {
    float u[6] = {j.x, j.y, j.z, tau.x, tau.y, tau.z};
    float dv[3], dw[3];
    for (int r = 0; r < 3; ++r) {
      float a = 0.0f, b = 0.0f;
      for (int k = 0; k < 6; ++k) {
        a += e.mobility[6 * r + k] * u[k];
        b += e.mobility[6 * (r + 3) + k] * u[k];
      }
      dv[r] = a;
      dw[r] = b;
    }
    vel = vel + f3{dv[0], dv[1], dv[2]};
    w = w + f3{dw[0], dw[1], dw[2]};
    pos = pos + float(P.dt) * vel;
    float rate = len(w);
    if (rate > 1.0e-12f) {
      rot = qnormalize(qmul(axis_angle(w / rate, float(P.dt) * rate), rot));
      turn = true;
    }
  }
  st3(S.interval_impulse, i, impulse);
  st3(S.interval_angular, i, angular);
  if (P.predict) {
    st3(S.tool_vel, i, vel);
    st3(S.tool_omega, i, w);
    st3(S.tool_pos, i, pos);
    if (turn) {
      S.tool_rot[4 * i] = rot.x;
      S.tool_rot[4 * i + 1] = rot.y;
      S.tool_rot[4 * i + 2] = rot.z;
      S.tool_rot[4 * i + 3] = rot.w;
    }
  }
  st3(S.reaction, i, f3{0.f, 0.f, 0.f});
  st3(S.torque, i, f3{0.f, 0.f, 0.f});
  st3(S.board, i, f3{0.f, 0.f, 0.f});
  if (S.instance == nullptr) S.speed[0] = 0.0f;
}

__device__ void finish_external_dev(const bm_params& P, const bm_state& S) {
  f3 j = ld3(S.reaction, 0), tau = ld3(S.torque, 0);
  f3 pos = ld3(S.tool_pos, 0);
  st3(S.interval_impulse, 0, ld3(S.interval_impulse, 0) + j);
  st3(S.interval_angular, 0, ld3(S.interval_angular, 0) + (tau + cross(pos, j)));
  if (P.predict) {
    float u[6] = {j.x, j.y, j.z, tau.x, tau.y, tau.z};
    float dv[3], dw[3];
    for (int r = 0; r < 3; ++r) {
      float a = 0.0f, b = 0.0f;
      for (int k = 0; k < 6; ++k) {
        a += S.mobility[6 * r + k] * u[k];
        b += S.mobility[6 * (r + 3) + k] * u[k];
      }
      dv[r] = a;
      dw[r] = b;
    }
    f3 vel = ld3(S.tool_vel, 0) + f3{dv[0], dv[1], dv[2]};
    st3(S.tool_vel, 0, vel);
    f3 w = ld3(S.tool_omega, 0) + f3{dw[0], dw[1], dw[2]};
    st3(S.tool_omega, 0, w);
    st3(S.tool_pos, 0, pos + float(P.dt) * vel);
    float rate = len(w);
    if (rate > 1.0e-12f) {
      quat nq = qnormalize(qmul(axis_angle(w / rate, float(P.dt) * rate), ldq(S.tool_rot)));
      S.tool_rot[0] = nq.x;
      S.tool_rot[1] = nq.y;
      S.tool_rot[2] = nq.z;
      S.tool_rot[3] = nq.w;
    }
  }
  for (int i = 0; i < 3; ++i) {
    S.reaction[i] = 0.0f;
    S.torque[i] = 0.0f;
    S.board[i] = 0.0f;
  }
  S.speed[0] = 0.0f;
}
__global__ void k_finish_external(const bm_params P, bm_state S) { finish_external_dev(P, S); }

__device__ void accumulate_reaction_dev(const bm_state& S) {
  f3 j = ld3(S.reaction, 0);
  st3(S.interval_impulse, 0, ld3(S.interval_impulse, 0) + j);
  st3(S.interval_angular, 0, ld3(S.interval_angular, 0) + (ld3(S.torque, 0) + cross(ld3(S.tool_pos, 0), j)));
}
__global__ void k_accumulate_reaction(bm_state S) { accumulate_reaction_dev(S); }

// assign (contact.py:996-1039): side digit for a particle behind the edge band or pinched
// on the board under the edge of knife `knife` (batched instances: the knife of its instance).
__device__ void assign_dev(const bm_params& P, const bm_state& S, int64_t p, int64_t digit, f3 xp, f3 vp,
                           int knife) {
  // The band test below needs only the particle's x and v and the knife, so it comes first;
  // the code (an int64 load and division) is read only for particles inside the band.
  int inst = instance_of(S, p) + knife;
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  quat q = ldq(S.tool_rot + 4 * inst);
  f3 pos = ld3(S.tool_pos, inst);
  f3 edge = pos + rotate(q, f3{0.f, 0.f, -hf.z});
  float spacing = float(P.spacing);
  // The band is a slab |local.x| <= 6 h about the knife plane: its x coordinate alone, from
  // the knife's x axis, sends most particles away.  The margin leaves every particle near the
  // slab faces to the exact test below.
  if (fabsf(dot(xp - edge, rotate(q, f3{1.f, 0.f, 0.f}))) > 6.0f * spacing * 1.0001f + 1.0e-6f) return;
  f3 local = rotate_inv(q, xp - edge);
  float front = -local.z;
  f3 edge_v = ld3(S.tool_vel, inst) + cross(ld3(S.tool_omega, inst), edge - pos);
  f3 d = rotate(q, f3{0.f, 0.f, -1.f});
  float advance = dot(edge_v - vp, d);
  float r = float(P.radius), dt = float(P.dt);
  float back = 2.0f * spacing + fmaxf(advance * dt, 0.0f);
  bool behind = front >= -back && front <= 0.0f;
  bool on_board = (P.board_half[0] == 0.0 || fabsf(xp.x) <= float(P.board_half[0])) &&
                  (P.board_half[1] == 0.0 || fabsf(xp.y) <= float(P.board_half[1]));
  f3 edge_here = edge + rotate(q, f3{0.f, local.y, 0.f});
  bool pinched = on_board && edge_here.z <= r && xp.z <= edge_here.z + r;
  if (fabsf(local.x) <= 6.0f * spacing && fabsf(local.y) <= hf.y && (behind || pinched)) {
    int64_t code = S.codes[p];
    if ((code / digit) % 3 != 0) return;
    int64_t side = local.x < 0.0f ? 1 : 2;
    S.codes[p] = code + digit * side;
    // A new side code changes the grid fields and the pair candidates: the fields and lists
    // around the relabelled particle are rebuilt next step (all of them if the record
    // overflows).
    int k = atomicAdd(S.changed_count, 1);
    if (k < P.changed_capacity) {
      S.changed[k] = int32_t(p);
      atomicOr(S.topology_dirty, kTopoChanged);
    } else {
      S.pair_dirty[0] = 1;
      atomicOr(S.topology_dirty, kTopoFull);
    }
  }
}

// assign_sides (solver.py:866-890) for knives first .. tools-1: knife k labels digit * 3^k.
__global__ void __launch_bounds__(kBlock) k_mark_relabelled(const bm_params P, bm_state S) {
  mark_relabelled(P, S, int(blockIdx.x) * (kBlock / 64) + int(threadIdx.x >> 6), int(gridDim.x) * (kBlock / 64), false);
}
__global__ void __launch_bounds__(kBlock) k_assign(const bm_params P, bm_state S, int64_t digit, int cutting,
                                                   int first) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n || cutting == 0) return;
  f3 xp = ld3(S.x, p), vp = ld3(S.v, p);
  for (int k = 0; k < max(int(P.tools), 1); ++k, digit *= 3)
    if (k >= first) assign_dev(P, S, p, digit, xp, vp, k);
}

// ------------------------------------------------------------------------ port kernels
__global__ void k_load_external(bm_state S, const float* packed) {
  for (int i = 0; i < 3; ++i) {
    S.tool_pos[i] = packed[i];
    S.tool_vel[i] = packed[7 + i];
    S.tool_omega[i] = packed[10 + i];
    S.interval_impulse[i] = 0.0f;
    S.interval_angular[i] = 0.0f;
  }
  for (int i = 0; i < 4; ++i) S.tool_rot[i] = packed[3 + i];
  for (int i = 0; i < 36; ++i) S.mobility[i] = packed[13 + i];
}
__global__ void k_record_pose(bm_state S, int index, float* out) {
  int offset = 9 + 7 * index;
  for (int i = 0; i < 3; ++i) out[offset + i] = S.tool_pos[i];
  for (int i = 0; i < 4; ++i) out[offset + 3 + i] = S.tool_rot[i];
}
__global__ void k_store_interval(bm_state S, float* out) {
  for (int i = 0; i < 3; ++i) {
    out[i] = S.interval_impulse[i];
    out[3 + i] = S.interval_angular[i];
    out[6 + i] = float(S.errors[i]);
  }
}
// Batched instances: one thread per instance.
__global__ void k_load_external_batch(const bm_params P, bm_state S, const float* packed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.instances) return;
  const float* r = packed + 49 * i;
  st3(S.tool_pos, i, f3{r[0], r[1], r[2]});
  for (int k = 0; k < 4; ++k) S.tool_rot[4 * i + k] = r[3 + k];
  st3(S.tool_vel, i, f3{r[7], r[8], r[9]});
  st3(S.tool_omega, i, f3{r[10], r[11], r[12]});
  for (int k = 0; k < 36; ++k) S.mobility[36 * i + k] = r[13 + k];
  st3(S.interval_impulse, i, f3{0.f, 0.f, 0.f});
  st3(S.interval_angular, i, f3{0.f, 0.f, 0.f});
}
__global__ void k_store_interval_batch(const bm_params P, bm_state S, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.instances) return;
  st3(out, 2 * i, ld3(S.interval_impulse, i));
  st3(out, 2 * i + 1, ld3(S.interval_angular, i));
}

// validate_interval (material.py:6-20)
__global__ void k_validate(const bm_params P, bm_state S) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  float d = mdet(ld9(S.F, p));
  int bad = (!(d > 0.0f) || !isfinite(d)) ? 1 : 0;
  f3 x = ld3(S.x, p), v = ld3(S.v, p);
  bad += !isfinite(x.x) + !isfinite(x.y) + !isfinite(x.z) + !isfinite(v.x) + !isfinite(v.y) + !isfinite(v.z);
  if (bad) atomicAdd(S.errors + 1, bad);
}

__global__ void k_count_embedded(const bm_params P, bm_state S, int32_t* count) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  f3 hf = f3{float(P.half[0]), float(P.half[1]), float(P.half[2])};
  f3 local = rotate_inv(ldq(S.tool_rot), ld3(S.x, p) - ld3(S.tool_pos, 0));
  if (fabsf(local.x) <= 6.0f * float(P.spacing) && fabsf(local.y) <= hf.y && fabsf(local.z) <= hf.z)
    atomicAdd(count, 1);
}

// metrics (diagnostics.py:7-35)
__global__ void __launch_bounds__(kBlock) k_metrics(const bm_params P, bm_state S, float* out) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float kin = 0.f, grav = 0.f, el = 0.f, vmax = 0.f, zmax = -3.0e38f, detmin = 3.0e38f;
  if (p < P.n) {
    m3 F = ld9(S.F, p);
    m3 fe = mmul(F, minv(ld9(S.Fp, p)));
    m3 U, V;
    f3 sg;
    svd3(fe, U, sg, V);
    m3 R = mmul_t(U, V);
    float e = 0.f;
    float mu, lam;
    lame_of(P, S, p, mu, lam);
    for (int k = 0; k < 9; ++k) {
      float d = fe.a[k] - R.a[k];
      e += mu * d * d;
    }
    float J = mdet(fe);
    e += 0.5f * lam * (J - 1.0f) * (J - 1.0f);
    f3 v = ld3(S.v, p), x = ld3(S.x, p);
    kin = 0.5f * float(P.mass) * dot(v, v);
    grav = float(P.mass) * 9.81f * x.z;
    el = float(P.volume) * e;
    vmax = len(v);
    zmax = x.z;
    detmin = mdet(F);
  }
  kin = wave_sum(kin);
  grav = wave_sum(grav);
  el = wave_sum(el);
  vmax = wave_max(vmax);
  zmax = wave_max(zmax);
  float dm = detmin;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dm = fminf(dm, __shfl_xor(dm, off, 64));
  if (lane0()) {
    atomicAdd(out + 0, kin);
    atomicAdd(out + 1, grav);
    atomicAdd(out + 2, el);
    atomicMax(out + 3, vmax);
    atomicMax(out + 4, zmax);
    atomicMin(out + 5, dm);
  }
}

// ------------------------------------------------------------------------ probes
__global__ void k_probe_blade_query(const float* local, f3 hf, float bevel, float* out, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  f3 c, nrm;
  float d;
  blade_query(ld3(local, i), hf, bevel, c, nrm, d);
  out[7 * i] = c.x; out[7 * i + 1] = c.y; out[7 * i + 2] = c.z;
  out[7 * i + 3] = nrm.x; out[7 * i + 4] = nrm.y; out[7 * i + 5] = nrm.z;
  out[7 * i + 6] = d;
}
__global__ void k_probe_friction(const float* in, Contact c, float* out, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float* a = in + 13 * i;
  f3 jt, ht;
  friction(c, f3{a[0], a[1], a[2]}, f3{a[3], a[4], a[5]}, a[6], f3{a[7], a[8], a[9]}, a[10], jt, ht);
  st3(out, 2 * i, jt);
  st3(out, 2 * i + 1, ht);
}
// Corotated stress of each F (elastic, Fp = I) as the F update computes it; svd_path[i] is 0
// for the polar iteration, 1 when it fell back to the SVD, -1 for det F <= 0.
__global__ void k_probe_stress(const float* Fin, float mu, float lam, float* out, int32_t* svd_path, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  m3 F = ld9(Fin, i), P = mzero();
  float J = mdet(F);
  int path = -1;
  if (J > 0.0f && isfinite(J)) {
    path = corotated_polar(F, J, mu, lam, P) ? 0 : 1;
    if (path == 1) {
      m3 U, V, I = mzero();
      f3 sg;
      I.a[0] = I.a[4] = I.a[8] = 1.0f;
      svd3(F, U, sg, V);
      P = corotated(U, sg, V, I, mu, lam);
    }
  }
  st9(out, i, P);
  svd_path[i] = path;
}

__global__ void k_probe_svd(const float* A, float* U, float* s, float* V, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  m3 u, v;
  f3 sg;
  svd3(ld9(A, i), u, sg, v);
  st9(U, i, u);
  st9(V, i, v);
  st3(s, i, sg);
}

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

bool state_ok(const bm_params* p, const bm_state* s) {
  if (p == nullptr || s == nullptr || p->n < 0 || s->x == nullptr || s->v == nullptr) return false;
  // Several knives: free knives of a single solver, with their own histories, damage and sums.
  if (p->tools > 1 && (s->extra_histories == nullptr || s->extra_damage == nullptr || s->extra_partial == nullptr ||
                       s->instance != nullptr || p->tool_model < 0))
    return false;
  // The Maxwell branch needs its Fv - I; viscosity its grid impulses.
  if ((p->maxwell && s->viscous == nullptr) || (p->viscosity > 0.0 && s->viscous_impulse == nullptr)) return false;
  // Batched instances carry their Lame constants and the slots of their contact sums (and
  // take neither the Maxwell branch nor viscosity).
  if (s->instance != nullptr && (p->maxwell || p->viscosity > 0.0)) return false;
  return s->instance == nullptr || (s->inst_mu != nullptr && s->inst_lam != nullptr && s->inst_sums != nullptr &&
                                    s->inst_buckets != nullptr && s->inst_pair_builds != nullptr &&
                                    s->inst_pair_dirty != nullptr);
}

}  // namespace

// ======================================================================== C-ABI
// Test probe of resist_cut: lane i is particle (x, v, damage)[i] under the knife edge moving
// at tool_velocity[i]; c holds the impulse j (3), knife origin (3), rotation xyzw (4),
// angular velocity (3), half sizes (3), contact radius, grid spacing and the damage work.
struct CutProbe {
  float a[19];
};
__global__ void k_probe_resist_cut(const float* x, const float* v, const float* tool_velocity, float* damage,
                                   CutProbe c, int cut_bits, float* out, float* work, int count) {
  int i = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= count) return;
  const float* a = c.a;
  float d = damage[i], w = 0.0f;
  f3 r = resist_cut(f3{a[0], a[1], a[2]}, ld3(x, i), ld3(v, i), f3{a[3], a[4], a[5]}, ld3(tool_velocity, i),
                    quat{a[6], a[7], a[8], a[9]}, f3{a[10], a[11], a[12]}, f3{a[13], a[14], a[15]}, a[16], a[17],
                    a[18], cut_bits, d, w);
  damage[i] = d;
  st3(out, i, r);
  if (w != 0.0f) atomicAdd(work, w);
}

// couple (material.py:155-209) on caller tables: field a has key code * node_count + node
// (-1 unused), mass and a state vector; heads[node] starts the node's linked list of fields
// (links[field], -1 ends it).  result[a] = state[a] + sum over compatible b != a of
// (m_a state_b - m_b state_a) / (sum of every field's mass at the node); fields with a negative
// key or no mass are left as they are in result.  A list longer than count is cut there.
__global__ void k_couple_keys(const int64_t* keys, const int32_t* heads, const int32_t* links, const float* mass,
                              const float* state, float* result, int64_t count, int64_t node_count) {
  int64_t a = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= count) return;
  const int64_t key = keys[a];
  const float ma = mass[a];
  if (key < 0 || !(ma > 0.0f)) return;
  const int64_t node = key % node_count, code = key / node_count;
  const f3 sa = ld3(state, a);
  float total = 0.0f;
  f3 exchange = f3{0.f, 0.f, 0.f};
  int64_t walked = 0;
  for (int32_t b = heads[node]; b >= 0 && b < count && walked < count; b = links[b], ++walked) {
    const float mb = mass[b];
    total += mb;
    if (b != a && compatible(code, keys[b] / node_count)) exchange = exchange + (ma * ld3(state, b) - mb * sa);
  }
  st3(result, a, sa + exchange / total);
}

// g2p (material.py:343-384) on caller tables: particle p reads its 8 corners' old and new
// velocities at slots[8 p + a] with weights[8 p + a]; v = (1 - flip) pic + flip (v + sum w
// (new - old)) and x += dt pic (TLMPM Eq. 13: positions move with the grid velocity).  A slot
// outside [0, grid_count) contributes nothing.
__global__ void k_g2p_slots(float* x, float* v, const int32_t* slots, const float* weights, const float* old_v,
                            const float* new_v, float dt, float flip, int64_t n, int64_t grid_count) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  f3 pic = f3{0.f, 0.f, 0.f}, dv = pic;
  for (int a = 0; a < 8; ++a) {
    int32_t s = slots[8 * p + a];
    if (s < 0 || s >= grid_count) continue;
    float w = weights[8 * p + a];
    f3 nv = ld3(new_v, s);
    pic = pic + w * nv;
    dv = dv + w * (nv - ld3(old_v, s));
  }
  st3(v, p, (1.0f - flip) * pic + flip * (ld3(v, p) + dv));
  st3(x, p, ld3(x, p) + dt * pic);
}

extern "C" {

int bm_version(void) { return 27; }

int bm_refresh_stress(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s) || !s->F || !s->Fp || !s->stress) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_refresh_stress, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_clear_step(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s)) return -1;
  hipLaunchKernelGGL(k_clear_step, dim3(1), dim3(1), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_rigid_contacts(const bm_params* p, const bm_state* s, int32_t cut_bits, void* stream) {
  if (!state_ok(p, s) || !s->impulses || !s->histories || !s->damage || !s->rigid_partial || !s->counters) return -1;
  if (p->n == 0) return 0;
  if (p->rigid_blocks < blocks_for(p->n)) return -2;
  int nb = blocks_for(p->n);
  hipLaunchKernelGGL(k_rigid, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, *p, *s, int(cut_bits),
                     int(kAccumulate), 0, 0, 0);
  if (p->tools > 1)
    hipLaunchKernelGGL(k_extra_blades, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, *p, *s, int(cut_bits));
  hipLaunchKernelGGL(k_prep, dim3(1), dim3(kPrepBlock), 0, (hipStream_t)stream, *p, *s, nb, int(kAccumulate),
                     ToolTarget{nullptr, nullptr, 0});
  if (p->tools > 1)
    hipLaunchKernelGGL(k_extra_tools, dim3(1), dim3(kPrepBlock), 0, (hipStream_t)stream, *p, *s, nb,
                       int(kAccumulate), ToolTarget{nullptr, nullptr, 0});
  return launch_status();
}

namespace {
static bool pair_args_ok(const bm_params* p, const bm_state* s) {
  return state_ok(p, s) && s->pair_keys && s->pair_history && s->pair_used && s->pair_count && s->pair_dirty &&
         s->pair_builds && s->list_reach && s->built_x && s->built_codes && s->partners && s->partner_count &&
         s->list_heads && s->list_links && s->changed && s->changed_count && s->pair_mark && s->pair_epoch &&
         p->changed_capacity >= 0 && p->pair_capacity > 0 && (p->pair_capacity & (p->pair_capacity - 1)) == 0 &&
         p->list_buckets > 0 && (s->instance != nullptr || (p->list_buckets & (p->list_buckets - 1)) == 0) &&
         p->list_buckets <= (int64_t(1) << 28) && p->list_lanes >= 1 && p->n < (int64_t(1) << 31) &&
         (p->mover_capacity <= 0 || (s->movers && s->displaced && s->refresh_mark && s->refresh_state &&
                                     s->instance == nullptr && p->pair_mode == 0 &&
                                     p->list_lanes <= kRefreshLanes && p->list_buckets >= (int64_t(1) << 17)));
}
static int64_t brick_count(const bm_params* p) { return p->bricks[0] * p->bricks[1] * p->bricks[2]; }
// Workgroups of a brick kernel: one per brick with particles.
static int64_t brick_launch(const bm_params* p) { return p->active_bricks > 0 ? p->active_bricks : brick_count(p); }
static int64_t node_launch(const bm_params* p) { return p->active_nodes > 0 ? p->active_nodes : p->node_count; }
static int64_t window_nodes(const bm_params* p) { return (p->brick[0] + 1) * (p->brick[1] + 1) * (p->brick[2] + 1); }
static Div32 div32(int64_t d) { return Div32{d > 1 ? uint32_t(0xFFFFFFFFu / uint32_t(d)) + 1u : 0u, uint32_t(d)}; }
static Div64 div64(int64_t d) { return Div64{d > 1 ? ~uint64_t(0) / uint64_t(d) + 1u : 0u, uint32_t(d)}; }
// Division constants of a launch (the index ranges they need are checked by brick_args_ok).
static Divisors divisors_of(const bm_params* p) {
  Divisors D;
  D.bx = div32(p->brick[0]);
  D.by = div32(p->brick[1]);
  D.bz = div32(p->brick[2]);
  D.bybz = div32(p->brick[1] * p->brick[2]);
  D.wz = div32(p->brick[2] + 1);
  D.wywz = div32((p->brick[1] + 1) * (p->brick[2] + 1));
  D.nbz = div64(p->bricks[2]);
  D.nbynbz = div64(p->bricks[1] * p->bricks[2]);
  D.dz = div64(p->dims[2]);
  D.dyz = div64(p->dims[1] * p->dims[2]);
  return D;
}
// LDS of a brick kernel: floats_per_fw floats per (field slot, window node) plus
// floats_per_corner per (cell, corner) of the brick.
static size_t lds_bytes(const bm_params* p, int floats_per_fw, int floats_per_corner) {
  int64_t cells = p->brick[0] * p->brick[1] * p->brick[2];
  return size_t(floats_per_fw) * size_t(p->fields) * size_t(window_nodes(p)) * sizeof(float) +
         size_t(floats_per_corner) * size_t(8 * cells) * sizeof(float);
}
static bool brick_args_ok(const bm_params* p, const bm_state* s) {
  if (!state_ok(p, s) || !s->perm || !s->p2g_partial || !s->remap_partial || !s->field_idx || !s->node_nf) return false;
  if (p->brick[0] < 1 || p->brick[1] < 1 || p->brick[2] < 1 || brick_count(p) < 1) return false;
  if ((p->active_bricks > 0 && (!s->brick_list || p->active_bricks > brick_count(p))) ||
      (p->active_nodes > 0 && (!s->node_list || p->active_nodes > p->node_count)))
    return false;
  // One lane per particle slot, and one per window node in k_p2g's field count.
  if (8 * p->brick[0] * p->brick[1] * p->brick[2] > 1024 || window_nodes(p) > 1024) return false;
  // Ranges of the division constants: node coordinates below 2^16 (Div32 of the brick
  // sizes), brick and node indices below 2^32 (Div64).
  if (p->dims[0] >= 65536 || p->dims[1] >= 65536 || p->dims[2] >= 65536 || brick_count(p) >= (int64_t(1) << 32) ||
      p->node_count >= (int64_t(1) << 32))
    return false;
  if (p->bricks[0] * p->brick[0] < p->dims[0] - 1 || p->bricks[1] * p->brick[1] < p->dims[1] - 1 ||
      p->bricks[2] * p->brick[2] < p->dims[2] - 1)
    return false;
  return lds_bytes(p, 7, 7) <= 65536 && lds_bytes(p, 3, 3) <= 65536;
}
// Workgroups of a running F update: kUpdateThreads threads, a multiple of 8 workgroups.
static int update_blocks(int threads) { return (kUpdateThreads / threads) & ~7; }
// Threads of a brick workgroup: one per particle slot (8 per cell) and at least one per window
// node, in whole waves (at most 1024 by brick_args_ok and the LDS bound on the window).
static int brick_threads(const bm_params* p) {
  // Lanes map to slots two cells per row of 16 (lane_cell, lane_slot): whole rows.
  int64_t slots = 16 * ((p->brick[0] * p->brick[1] * p->brick[2] + 1) / 2);
  int64_t t = slots > window_nodes(p) ? slots : window_nodes(p);
  t = ((t + 63) / 64) * 64;
  return int(t > 1024 ? 1024 : t);
}
// The node kernel of the re-scattered momentum (with `update_blocks` deferred-update
// workgroups among `blocks`) and the F update, in the instantiation of the material model.
// `blocks` counts the node and update workgroups; the multi-field workgroups come first.
static void launch_couple_remapped(const bm_params* p, const bm_state* s, const Divisors& D, int blocks,
                                   int update_blocks, hipStream_t st) {
  blocks += multi_blocks(*p);
  hipLaunchKernelGGL(k_couple_remapped, dim3(blocks), dim3(kBlock), 0, st, *p, *s, D, update_blocks);
}
static void launch_update_F(const bm_params* p, const bm_state* s, const Divisors& D, int blocks, int threads,
                            int defer, int nbricks, hipStream_t st) {
  if (p->maxwell)
    hipLaunchKernelGGL(k_update_F<true>, dim3(blocks), dim3(threads), 0, st, *p, *s, D, defer, nbricks);
  else
    hipLaunchKernelGGL(k_update_F<false>, dim3(blocks), dim3(threads), 0, st, *p, *s, D, defer, nbricks);
}
static void launch_register(const bm_params* p, const bm_state* s, hipStream_t st) {
  int64_t total = node_launch(p);
  hipLaunchKernelGGL(k_topo_clear, dim3(blocks_for(total < 1 << 20 ? total : 1 << 20)), dim3(kBlock), 0, st, *p, *s);
  if (p->n > 0) hipLaunchKernelGGL(k_topo_insert, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s);
}
// P2G .. update_F with the brick kernels.  p2g_flags: reset topology_dirty (1) and
// pair_dirty (2) once every reader of this step is done; g2p_flags: assign (1) and the
// pair-list staleness check (2) on the moved particles; defer: leave the F update to the
// next P2G where the field slots allow it.
static void launch_transfers(const bm_params* p, const bm_state* s, int p2g_flags, int64_t digit, int g2p_flags,
                      hipStream_t st, int defer = 0, StepFinish sf = StepFinish{-1, 0, ToolTarget{nullptr, nullptr, 0}}) {
  int nb = int(brick_launch(p)), tb = brick_threads(p);
  const Divisors D = divisors_of(p);
  if (p->maxwell)
    hipLaunchKernelGGL(k_p2g<true>, dim3(nb), dim3(tb), lds_bytes(p, 7, 7), st, *p, *s, p2g_flags, D);
  else
    hipLaunchKernelGGL(k_p2g<false>, dim3(nb), dim3(tb), lds_bytes(p, 7, 7), st, *p, *s, p2g_flags, D);
  hipLaunchKernelGGL(k_couple_velocity, dim3(multi_blocks(*p) + blocks_for(node_launch(p)) + (sf.mode >= 0 ? 1 : 0)),
                     dim3(kBlock), 0, st, *p, *s, D, sf);
  // Knives 1 .. tools-1: their sums and tool steps after knife 0's.
  if (sf.mode >= 0 && p->tools > 1)
    hipLaunchKernelGGL(k_extra_tools, dim3(1), dim3(kPrepBlock), 0, st, *p, *s, sf.rows, sf.mode, sf.tt);
  // Viscosity: the impulse of the coupled old velocities, coupled and added to new_v
  // (solver.py:433-458).
  if (p->viscosity > 0.0) {
    hipLaunchKernelGGL(k_viscous_impulse, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s);
    hipLaunchKernelGGL(k_viscous_couple, dim3(blocks_for(node_launch(p))), dim3(kBlock), 0, st, *p, *s);
  }
  // Held blade: constrain new_v before it moves particles (solver.py:476-477).
  if (p->blade_hold) bm_blade_constraint(p, s, st);
  hipLaunchKernelGGL(k_g2p, dim3(nb), dim3(tb), lds_bytes(p, 3, 3), st, *p, *s, digit, g2p_flags, D);
  // Knives 1 .. tools-1 label their digits (G2P labelled knife 0's) before the node kernel's
  // deferred update reads this step's relabelling.
  if (p->tools > 1 && (g2p_flags & 1))
    hipLaunchKernelGGL(k_assign, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s, digit, 1, 1);
  const int node_blocks = int(blocks_for(node_launch(p)));
  defer = defer && s->update_pending && s->update_done && s->field_idx_prev;
  if (defer) {
    // The F update is deferred to the next P2G: one extra workgroup of the node kernel sets
    // that up, and no F-update launch follows.
    launch_couple_remapped(p, s, D, node_blocks + kDeferredUpdateBlocks, kDeferredUpdateBlocks, st);
    return;
  }
  launch_couple_remapped(p, s, D, node_blocks, 0, st);
  const int ub = update_blocks(tb);
  launch_update_F(p, s, D, nb < ub ? nb : ub, tb, 0, nb, st);
}
}  // namespace

int bm_pair_contacts(const bm_params* p, const bm_state* s, int32_t parity, void* stream) {
  if (!pair_args_ok(p, s)) return -1;
  if (p->n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int old_p = parity & 1, new_p = 1 - old_p;
  hipLaunchKernelGGL(k_check_pair_list, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s);
  hipLaunchKernelGGL(k_clear_pair_slots, dim3(256), dim3(kBlock), 0, st, *s, p->pair_capacity, new_p);
  // Binning tags the bucket heads with the build count after this build's increment.
  hipLaunchKernelGGL(k_list_bin, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s);
  hipLaunchKernelGGL(k_list_reach, dim3(1), dim3(1), 0, st, *p, *s, new_p);
  hipLaunchKernelGGL(k_pairs, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s, old_p, 0,
                     PairExtras{int(blocks_for(p->n)), 0, 0, 0});
  hipLaunchKernelGGL(k_pairs_done, dim3(1), dim3(1), 0, st, *p, *s);
  return launch_status();
}

int bm_register(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s) || !s->field_code || !s->node_nf || !s->field_idx || !s->transfer_codes) return -1;
  if (p->fields < 1 || p->fields > 255) return -2;
  hipStream_t st = (hipStream_t)stream;
  launch_register(p, s, st);
  hipLaunchKernelGGL(k_topo_done, dim3(1), dim3(1), 0, st, *s);
  return launch_status();
}

int bm_p2g(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s)) return -1;
  const int nb = int(brick_launch(p)), tb = brick_threads(p);
  if (p->maxwell)
    hipLaunchKernelGGL(k_p2g<true>, dim3(nb), dim3(tb), lds_bytes(p, 7, 7), (hipStream_t)stream, *p, *s, 0,
                       divisors_of(p));
  else
    hipLaunchKernelGGL(k_p2g<false>, dim3(nb), dim3(tb), lds_bytes(p, 7, 7), (hipStream_t)stream, *p, *s, 0,
                       divisors_of(p));
  return launch_status();
}

int bm_couple_and_velocity(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s)) return -1;
  hipLaunchKernelGGL(k_couple_velocity, dim3(multi_blocks(*p) + blocks_for(node_launch(p))), dim3(kBlock), 0,
                     (hipStream_t)stream, *p, *s, divisors_of(p), StepFinish{-1, 0, ToolTarget{nullptr, nullptr, 0}});
  return launch_status();
}

int bm_g2p_remap(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s)) return -1;
  hipLaunchKernelGGL(k_g2p, dim3(int(brick_launch(p))), dim3(brick_threads(p)), lds_bytes(p, 3, 3), (hipStream_t)stream,
                     *p, *s, int64_t(1), 0, divisors_of(p));
  return launch_status();
}

int bm_couple_remapped(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s)) return -1;
  launch_couple_remapped(p, s, divisors_of(p), int(blocks_for(node_launch(p))), 0, (hipStream_t)stream);
  return launch_status();
}

int bm_update_F(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s)) return -1;
  int nb = int(brick_launch(p)), ub = update_blocks(brick_threads(p));
  launch_update_F(p, s, divisors_of(p), nb < ub ? nb : ub, brick_threads(p), 0, nb, (hipStream_t)stream);
  return launch_status();
}

int bm_material_step(const bm_params* p, const bm_state* s, void* stream) {
  if (!brick_args_ok(p, s) || !s->field_code || !s->node_nf || !s->field_idx || !s->transfer_codes) return -1;
  hipStream_t st = (hipStream_t)stream;
  launch_register(p, s, st);
  launch_transfers(p, s, 1, 1, 0, st);
  return launch_status();
}

int bm_tool_step(const bm_params* p, const bm_state* s, const float* targets, const float* rotations, int32_t index,
                 void* stream) {
  if (!state_ok(p, s) || !targets || !rotations || index < 0) return -1;
  hipLaunchKernelGGL(k_tool_step, dim3(1), dim3(1), 0, (hipStream_t)stream, *p, *s, targets, rotations, int(index));
  return launch_status();
}

int bm_finish_external(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s)) return -1;
  hipLaunchKernelGGL(k_finish_external, dim3(1), dim3(1), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_accumulate_reaction(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s)) return -1;
  hipLaunchKernelGGL(k_accumulate_reaction, dim3(1), dim3(1), 0, (hipStream_t)stream, *s);
  return launch_status();
}

int bm_assign(const bm_params* p, const bm_state* s, int64_t digit, int32_t cutting, void* stream) {
  if (!state_ok(p, s) || digit < 1 || !s->codes || !s->topology_dirty || !s->pair_dirty || !s->changed ||
      !s->changed_count)
    return -1;
  if (p->n == 0 || !cutting) return 0;
  hipLaunchKernelGGL(k_assign, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s, digit,
                     int(cutting), 0);
  // The relabelled particles' new pairs into the partner lists, as in a substep.
  if (pair_args_ok(p, s)) hipLaunchKernelGGL(k_mark_relabelled, dim3(64), dim3(kBlock), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_mark_relabelled(const bm_params* p, const bm_state* s, void* stream) {
  if (!pair_args_ok(p, s)) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_mark_relabelled, dim3(64), dim3(kBlock), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

namespace {
static int substep(const bm_params* p, const bm_state* s, int32_t cut_bits, int32_t cutting, int64_t digit, int32_t parity,
            int32_t accumulate, const float* targets, const float* rotations, int32_t index, void* stream, int defer) {
  if (!brick_args_ok(p, s) || !pair_args_ok(p, s) || !s->rigid_partial || !s->counters || digit < 1) return -1;
  bool external = p->tool_model < 0;
  if (!external && (!targets || !rotations || index < 0)) return -1;
  if (p->n == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  int old_p = parity & 1, new_p = 1 - old_p;
  int mode = kOverwrite | kPairPrep | kTopologyPrep;
  if (external) {
    // With a held blade the grid constraint adds to the reaction first (finish runs last).
    if (!p->blade_hold) mode |= kFinishExternal;
  } else {
    mode |= kToolStep | (accumulate ? kAccumulateInterval : 0);
  }
  int nb = blocks_for(p->n);
  if (p->rigid_blocks < nb + kReduceBlocks) return -2;
  // Registration workgroups of the pair kernel, which exit at once on most substeps: 2048
  // wavefronts, one per (relabelled particle, neighbouring cell).
  const int insert_blocks = int(p->n < 512 * kBlock ? blocks_for(p->n) : 512);
  // 1. contacts with the knife and the board, plus the step's preparation: empty the new
  //    pair table's slots, bin particles for a stale pair list, clear a stale field table.
  // Its first workgroups refresh the pair lists of the last G2P's movers, the next add the
  // pairs of the particles it relabelled (one wave per particle, 64 workgroups of four
  // waves); both exit at once when there is nothing to do.
  const int refresh_blocks = p->mover_capacity > 0 ? kRefreshBlocks : 0;
  const int mark_blocks =
      p->pair_mode != 0 ? 0 : int(p->changed_capacity < 256 ? (p->changed_capacity + 3) / 4 : 64);
  hipLaunchKernelGGL(k_rigid, dim3(refresh_blocks + mark_blocks + nb), dim3(kBlock), 0, st, *p, *s,
                     cutting ? int(cut_bits) : 0, mode, new_p, refresh_blocks, mark_blocks);
  // Knives 1 .. tools-1: their contacts after knife 0's, their sums and tool steps after
  // knife 0's step.
  if (p->tools > 1)
    hipLaunchKernelGGL(k_extra_blades, dim3(nb), dim3(kBlock), 0, st, *p, *s, cutting ? int(cut_bits) : 0);
  // 2. particle pairs across cuts (partner lists rebuilt first when stale); its first
  //    workgroups reduce the contact sums and register the fields of a label change.
  hipLaunchKernelGGL(k_pairs, dim3(nb + kReduceBlocks + insert_blocks), dim3(kBlock), 0, st, *p, *s, old_p, 1,
                     PairExtras{nb, kReduceBlocks, insert_blocks, nb});
  // 3. transfers; the node kernel's last workgroup applies the contact sums and moves the
  //    knife (its update reads only this step's rigid reaction, solver.py:824-864), which
  //    G2P's labelling reads; G2P labels the moved particles and checks the pair list's skin
  //    (the next contact kernel adds the relabelled particles' new pairs to the lists and
  //    refreshes the movers' lists).
  launch_transfers(p, s, 3, digit, (cutting ? 1 : 0) | 2, st, defer,
                   StepFinish{mode, nb, ToolTarget{targets, rotations, int(index)}});
  if (external && p->blade_hold) hipLaunchKernelGGL(k_finish_external, dim3(1), dim3(1), 0, st, *p, *s);
  return launch_status();
}
}  // namespace

int bm_substep(const bm_params* p, const bm_state* s, int32_t cut_bits, int32_t cutting, int64_t digit,
               int32_t parity, int32_t accumulate, const float* targets, const float* rotations, int32_t index,
               void* stream) {
  return substep(p, s, cut_bits, cutting, digit, parity, accumulate, targets, rotations, index, stream, 0);
}

int bm_substep_deferred(const bm_params* p, const bm_state* s, int32_t cut_bits, int32_t cutting, int64_t digit,
                        int32_t parity, int32_t accumulate, const float* targets, const float* rotations,
                        int32_t index, void* stream) {
  return substep(p, s, cut_bits, cutting, digit, parity, accumulate, targets, rotations, index, stream, 1);
}

int bm_load_external_state(const bm_params* p, const bm_state* s, const float* packed49, void* stream) {
  if (!state_ok(p, s) || !packed49) return -1;
  hipLaunchKernelGGL(k_load_external, dim3(1), dim3(1), 0, (hipStream_t)stream, *s, packed49);
  return launch_status();
}

int bm_load_external_batch(const bm_params* p, const bm_state* s, const float* packed, void* stream) {
  if (!state_ok(p, s) || !packed || p->instances < 1 || !s->tool_pos || !s->mobility) return -1;
  hipLaunchKernelGGL(k_load_external_batch, dim3(blocks_for(p->instances)), dim3(kBlock), 0, (hipStream_t)stream, *p,
                     *s, packed);
  return launch_status();
}

int bm_store_interval_batch(const bm_params* p, const bm_state* s, float* out, void* stream) {
  if (!state_ok(p, s) || !out || p->instances < 1 || !s->interval_impulse || !s->interval_angular) return -1;
  hipLaunchKernelGGL(k_store_interval_batch, dim3(blocks_for(p->instances)), dim3(kBlock), 0, (hipStream_t)stream, *p,
                     *s, out);
  return launch_status();
}

int bm_record_pose(const bm_params* p, const bm_state* s, int32_t index, float* out, void* stream) {
  if (!state_ok(p, s) || !out || index < 0) return -1;
  hipLaunchKernelGGL(k_record_pose, dim3(1), dim3(1), 0, (hipStream_t)stream, *s, int(index), out);
  return launch_status();
}

int bm_store_interval(const bm_params* p, const bm_state* s, float* out, void* stream) {
  if (!state_ok(p, s) || !out) return -1;
  hipLaunchKernelGGL(k_store_interval, dim3(1), dim3(1), 0, (hipStream_t)stream, *s, out);
  return launch_status();
}

int bm_validate_interval(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s)) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_validate, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_count_embedded(const bm_params* p, const bm_state* s, int32_t* count, void* stream) {
  if (!state_ok(p, s) || !count) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_count_embedded, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s, count);
  return launch_status();
}

int bm_metrics(const bm_params* p, const bm_state* s, float* out, void* stream) {
  if (!state_ok(p, s) || !out) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_metrics, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s, out);
  return launch_status();
}

int bm_plastic_project(const bm_params* p, const bm_state* s, void* stream) {
  if (!state_ok(p, s)) return -1;
  if (p->n == 0) return 0;
  hipLaunchKernelGGL(k_plastic_project, dim3(blocks_for(p->n)), dim3(kBlock), 0, (hipStream_t)stream, *p, *s);
  return launch_status();
}

int bm_probe_blade_query(const float* local, const float* half, float bevel, float* out, int32_t count,
                         void* stream) {
  if (!local || !half || !out || count < 0) return -1;
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_probe_blade_query, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, local,
                     f3{half[0], half[1], half[2]}, bevel, out, int(count));
  return launch_status();
}

int bm_probe_friction(const float* in, float dt, float mus, float mud, float gk, float gc, float* out, int32_t count,
                      void* stream) {
  if (!in || !out || count < 0) return -1;
  if (count == 0) return 0;
  Contact c{dt, 0.0f, mus, mud, gk, gc};
  hipLaunchKernelGGL(k_probe_friction, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, in, c, out,
                     int(count));
  return launch_status();
}

int bm_probe_stress(const float* F, float mu, float lam, float* out, int32_t* svd_path, int32_t count,
                    void* stream) {
  if (!F || !out || !svd_path || count < 0) return -1;
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_probe_stress, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, F, mu, lam, out,
                     svd_path, count);
  return launch_status();
}

int bm_probe_resist_cut(const float* x, const float* v, const float* tool_velocity, float* damage,
                        const float* params19, int32_t cut_bits, float* out, float* work, int32_t count, void* stream) {
  if (count < 0 || (count > 0 && (!x || !v || !tool_velocity || !damage || !params19 || !out || !work))) return -1;
  if (count == 0) return 0;
  CutProbe c;
  for (int k = 0; k < 19; ++k) c.a[k] = params19[k];
  hipLaunchKernelGGL(k_probe_resist_cut, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, x, v,
                     tool_velocity, damage, c, int(cut_bits), out, work, int(count));
  return launch_status();
}

int bm_couple(const int64_t* keys, const int32_t* heads, const int32_t* links, const float* mass, const float* state,
              float* result, int64_t count, int64_t node_count, void* stream) {
  if (count < 0 || node_count < 1 || count >= (int64_t(1) << 31)) return -1;
  if (count == 0) return 0;
  if (!keys || !heads || !links || !mass || !state || !result) return -1;
  hipLaunchKernelGGL(k_couple_keys, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, keys, heads, links,
                     mass, state, result, count, node_count);
  return launch_status();
}

int bm_g2p(float* x, float* v, const int32_t* slots, const float* weights, const float* old_v, const float* new_v,
           float dt, float flip, int64_t n, int64_t grid_count, void* stream) {
  if (n < 0 || grid_count < 0) return -1;
  if (n == 0) return 0;
  if (!x || !v || !slots || !weights || !old_v || !new_v) return -1;
  hipLaunchKernelGGL(k_g2p_slots, dim3(blocks_for(n)), dim3(kBlock), 0, (hipStream_t)stream, x, v, slots, weights,
                     old_v, new_v, dt, flip, n, grid_count);
  return launch_status();
}

int bm_probe_svd3(const float* A, float* U, float* s, float* V, int32_t count, void* stream) {
  if (!A || !U || !s || !V || count < 0) return -1;
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_probe_svd, dim3(blocks_for(count)), dim3(kBlock), 0, (hipStream_t)stream, A, U, s, V,
                     int(count));
  return launch_status();
}

}  // extern "C"
