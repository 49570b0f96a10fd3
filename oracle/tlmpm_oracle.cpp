// TEST INFRASTRUCTURE ONLY — CPU oracle of the BladeMaster TLMPM material substep.
//
// A sequential restatement of the reference Warp kernels, used by tests/, the smoke check
// and bench.py's cpu_baseline leg as the checker; never linked into the product library.
// Each stage cites the reference function it follows (paths relative to the reference
// repository).  Storage follows the reference's compact path: an open-addressing int64 key
// table (material.py:37-65), compaction to dense field indices (material.py:621-653) and
// per-node linked lists.  Pair contacts enumerate candidate pairs with a uniform cell grid
// and apply pair_at (contact.py:584-640) in ascending (p, q) order; the pair-search mode of
// the reference (all / hash / list) only changes which candidates are visited, never which
// pairs pass pair_at's own tests.
//
// Arithmetic: template parameter R is float (the FP32 mirror of the Warp kernels) or double
// (the FP64 run).  The 3x3 SVD is always evaluated in double by cyclic Jacobi on A^T A
// followed by a Givens QR of A V (the structure of McAdams et al. 2011, "Computing the SVD
// of 3x3 matrices with minimal branching and elementary floating point operations"), and
// rounded to R.
//
// Threads: `threads` > 1 (oracle_set_threads) runs the per-particle and per-field loops with
// OpenMP for the CPU baseline.  Every sum keeps its order: contact sums are added in particle
// order after the parallel loop, and grid scatters become gathers over each field slot's
// (particle, corner) list in particle order, so results equal the one-thread run bit for bit.
//
// Parity status of each function is stated in the header of tests/test_oracle.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class R> struct V3 {
  R x, y, z;
};
template <class R> inline V3<R> mk(R a, R b, R c) { return V3<R>{a, b, c}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> inline R comp(V3<R> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
template <class R> inline void setc(V3<R>& a, int i, R v) {
  if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v;
}

template <class R> struct M3 {
  R m[3][3];
};
template <class R> inline M3<R> zero3() { M3<R> r; for (auto& row : r.m) for (auto& e : row) e = R(0); return r; }
template <class R> inline M3<R> eye3() { M3<R> r = zero3<R>(); r.m[0][0] = r.m[1][1] = r.m[2][2] = R(1); return r; }
template <class R> inline M3<R> mul(const M3<R>& a, const M3<R>& b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
  return r;
}
template <class R> inline V3<R> mulv(const M3<R>& a, V3<R> v) {
  return {a.m[0][0] * v.x + a.m[0][1] * v.y + a.m[0][2] * v.z, a.m[1][0] * v.x + a.m[1][1] * v.y + a.m[1][2] * v.z,
          a.m[2][0] * v.x + a.m[2][1] * v.y + a.m[2][2] * v.z};
}
template <class R> inline M3<R> transpose(const M3<R>& a) {
  M3<R> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
  return r;
}
template <class R> inline M3<R> scale(R s, const M3<R>& a) {
  M3<R> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
  return r;
}
template <class R> inline M3<R> addm(const M3<R>& a, const M3<R>& b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
  return r;
}
template <class R> inline M3<R> outer(V3<R> a, V3<R> b) {
  M3<R> r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = comp(a, i) * comp(b, j);
  return r;
}
template <class R> inline R det(const M3<R>& a) {
  return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
         a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
         a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
template <class R> inline M3<R> inverse(const M3<R>& a) {
  R d = det(a);
  M3<R> r;
  r.m[0][0] = (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) / d;
  r.m[0][1] = (a.m[0][2] * a.m[2][1] - a.m[0][1] * a.m[2][2]) / d;
  r.m[0][2] = (a.m[0][1] * a.m[1][2] - a.m[0][2] * a.m[1][1]) / d;
  r.m[1][0] = (a.m[1][2] * a.m[2][0] - a.m[1][0] * a.m[2][2]) / d;
  r.m[1][1] = (a.m[0][0] * a.m[2][2] - a.m[0][2] * a.m[2][0]) / d;
  r.m[1][2] = (a.m[0][2] * a.m[1][0] - a.m[0][0] * a.m[1][2]) / d;
  r.m[2][0] = (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]) / d;
  r.m[2][1] = (a.m[0][1] * a.m[2][0] - a.m[0][0] * a.m[2][1]) / d;
  r.m[2][2] = (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]) / d;
  return r;
}

// Quaternion (x, y, z, w), Hamilton product, as Warp's wp.quat.
template <class R> struct Q {
  R x, y, z, w;
};
template <class R> inline V3<R> qv(Q<R> q) { return {q.x, q.y, q.z}; }
template <class R> inline V3<R> rotate(Q<R> q, V3<R> v) {
  V3<R> u = qv(q);
  return (R(2) * q.w * q.w - R(1)) * v + (R(2) * dot(u, v)) * u + (R(2) * q.w) * cross(u, v);
}
template <class R> inline V3<R> rotate_inv(Q<R> q, V3<R> v) {
  V3<R> u = qv(q);
  return (R(2) * q.w * q.w - R(1)) * v + (R(2) * dot(u, v)) * u - (R(2) * q.w) * cross(u, v);
}
template <class R> inline Q<R> qmul(Q<R> a, Q<R> b) {
  V3<R> av = qv(a), bv = qv(b);
  V3<R> v = a.w * bv + b.w * av + cross(av, bv);
  return {v.x, v.y, v.z, a.w * b.w - dot(av, bv)};
}
template <class R> inline Q<R> qnormalize(Q<R> q) {
  R l = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return {q.x / l, q.y / l, q.z / l, q.w / l};
}
template <class R> inline Q<R> axis_angle(V3<R> axis, R angle) {
  R s = std::sin(angle / R(2)), c = std::cos(angle / R(2));
  return {axis.x * s, axis.y * s, axis.z * s, c};
}

// ---------------------------------------------------------------------------------------
// 3x3 SVD in double: A = U diag(s) V^T with U, V proper rotations, s sorted by magnitude
// descending, only s[2] may be negative (det A < 0).
void svd3_double(const double A[3][3], double U[3][3], double s[3], double V[3][3]) {
  double S[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
  double W[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
    double diag = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (S[p][q] == 0.0) continue;
        // Exact Jacobi rotation annihilating S[p][q] (Golub & Van Loan, Alg. 8.4.1).
        double tau = (S[q][q] - S[p][p]) / (2.0 * S[p][q]);
        double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        for (int k = 0; k < 3; ++k) {
          double a = S[k][p], b = S[k][q];
          S[k][p] = c * a - sn * b;
          S[k][q] = sn * a + c * b;
        }
        for (int k = 0; k < 3; ++k) {
          double a = S[p][k], b = S[q][k];
          S[p][k] = c * a - sn * b;
          S[q][k] = sn * a + c * b;
        }
        for (int k = 0; k < 3; ++k) {
          double a = W[k][p], b = W[k][q];
          W[k][p] = c * a - sn * b;
          W[k][q] = sn * a + c * b;
        }
      }
  }
  // B = A W; sort columns by norm descending.
  double B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = A[i][0] * W[0][j] + A[i][1] * W[1][j] + A[i][2] * W[2][j];
  int order[3] = {0, 1, 2};
  double norm[3];
  for (int j = 0; j < 3; ++j) norm[j] = B[0][j] * B[0][j] + B[1][j] * B[1][j] + B[2][j] * B[2][j];
  std::sort(order, order + 3, [&](int a, int b) { return norm[a] > norm[b]; });
  double Vs[3][3], Bs[3][3];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      Vs[i][j] = W[i][order[j]];
      Bs[i][j] = B[i][order[j]];
    }
  double dv = Vs[0][0] * (Vs[1][1] * Vs[2][2] - Vs[1][2] * Vs[2][1]) - Vs[0][1] * (Vs[1][0] * Vs[2][2] - Vs[1][2] * Vs[2][0]) +
              Vs[0][2] * (Vs[1][0] * Vs[2][1] - Vs[1][1] * Vs[2][0]);
  if (dv < 0) {
    for (int i = 0; i < 3; ++i) {
      Vs[i][2] = -Vs[i][2];
      Bs[i][2] = -Bs[i][2];
    }
  }
  // Givens QR of Bs: Q^T Bs = Rm; U = Q is a rotation, diag(Rm) are the singular values.
  double Qm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int rows[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  const int cols[3] = {0, 0, 1};
  for (int g = 0; g < 3; ++g) {
    int i = rows[g][0], k = rows[g][1], col = cols[g];
    double a = Bs[i][col], b = Bs[k][col];
    double r = std::hypot(a, b);
    double c = 1.0, sn = 0.0;
    if (r > 0) {
      c = a / r;
      sn = b / r;
    }
    for (int j = 0; j < 3; ++j) {
      double x = Bs[i][j], y = Bs[k][j];
      Bs[i][j] = c * x + sn * y;
      Bs[k][j] = -sn * x + c * y;
    }
    for (int j = 0; j < 3; ++j) {
      double x = Qm[j][i], y = Qm[j][k];
      Qm[j][i] = c * x + sn * y;
      Qm[j][k] = -sn * x + c * y;
    }
  }
  // Make the first two singular values non-negative by flipping paired columns of U and V.
  for (int j = 0; j < 2; ++j) {
    if (Bs[j][j] < 0) {
      // flip column j of U and column 2 of U keeps det U; instead flip U col j and V col j,
      // and compensate the sign on the last singular value by flipping U col 2 and V col 2.
      for (int i = 0; i < 3; ++i) {
        Qm[i][j] = -Qm[i][j];
        Qm[i][2] = -Qm[i][2];
      }
      Bs[j][j] = -Bs[j][j];
      Bs[2][2] = -Bs[2][2];
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      U[i][j] = Qm[i][j];
      V[i][j] = Vs[i][j];
    }
  for (int j = 0; j < 3; ++j) s[j] = Bs[j][j];
}

template <class R> void svd3(const M3<R>& A, M3<R>& U, V3<R>& s, M3<R>& V) {
  double a[3][3], u[3][3], v[3][3], sd[3];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = double(A.m[i][j]);
  svd3_double(a, u, sd, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      U.m[i][j] = R(u[i][j]);
      V.m[i][j] = R(v[i][j]);
    }
  s = {R(sd[0]), R(sd[1]), R(sd[2])};
}

// ---------------------------------------------------------------------------------------
// Ternary side codes (material.py:23-34).
inline bool compatible(int64_t a, int64_t b) {
  bool result = true;
  while (a != 0 || b != 0) {
    int64_t x = a % 3, y = b % 3;
    if (x != 0 && y != 0 && x != y) result = false;
    a /= 3;
    b /= 3;
  }
  return result;
}

inline int64_t hash_slot(int64_t key, int64_t capacity) {
  return int64_t((uint64_t(key) * uint64_t(2654435761ull)) % uint64_t(capacity));
}

// Open-addressing insert with linear probing (material.py:37-47).
inline int64_t insert_key(std::vector<int64_t>& keys, int64_t key) {
  int64_t cap = int64_t(keys.size());
  int64_t slot = hash_slot(key, cap);
  for (int64_t probe = 0; probe < cap; ++probe) {
    if (keys[slot] == -1 || keys[slot] == key) {
      keys[slot] = key;
      return slot;
    }
    slot = (slot + 1) % cap;
  }
  return -1;
}

enum { CUT_ON = 1, CUT_ADVANCE = 2, CUT_HOLD = 4 };

// couple_at (material.py:155-179) over explicit field arrays: for every registered field a
// with positive mass, result_a = state_a + sum_{b compatible, b != a} (m_a s_b - m_b s_a) /
// sum_{b at the node} m_b, walking the node's linked list.
template <class R>
void couple_fields(const int64_t* keys, const int32_t* heads, const int32_t* links, const R* mass, const R* state,
                   R* result, int64_t count, int64_t node_count, int threads = 1) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
  for (int64_t a = 0; a < count; ++a) {
    if (keys[a] < 0 || mass[a] <= R(0)) continue;
    int64_t node = keys[a] % node_count;
    int64_t code_a = keys[a] / node_count;
    R total = R(0);
    R ex[3] = {R(0), R(0), R(0)};
    for (int64_t b = heads[node]; b >= 0; b = links[b]) {
      total += mass[b];
      int64_t code_b = keys[b] / node_count;
      if (b != a && compatible(code_a, code_b))
        for (int i = 0; i < 3; ++i) ex[i] = ex[i] + (mass[a] * state[3 * b + i] - mass[b] * state[3 * a + i]);
    }
    for (int i = 0; i < 3; ++i) result[3 * a + i] = state[3 * a + i] + ex[i] / total;
  }
}

// g2p_at (material.py:343-367) over explicit slots and weights.
template <class R>
void g2p_fields(R* x, R* v, const int32_t* slots, const R* weights, const R* old_v, const R* new_v, R dt, R flip,
                int64_t n, int threads = 1) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
  for (int64_t p = 0; p < n; ++p) {
    R pic[3] = {R(0), R(0), R(0)}, dv[3] = {R(0), R(0), R(0)};
    for (int a = 0; a < 8; ++a) {
      int32_t s = slots[8 * p + a];
      R w = weights[8 * p + a];
      for (int i = 0; i < 3; ++i) {
        pic[i] = pic[i] + w * new_v[3 * s + i];
        dv[i] = dv[i] + w * (new_v[3 * s + i] - old_v[3 * s + i]);
      }
    }
    for (int i = 0; i < 3; ++i) {
      v[3 * p + i] = (R(1) - flip) * pic[i] + flip * (v[3 * p + i] + dv[i]);
      x[3 * p + i] = x[3 * p + i] + dt * pic[i];
    }
  }
}

struct OracleConfig {
  int64_t n, node_count, key_capacity, pair_capacity, max_cut_index;
  double dt, spacing, mass, volume, radius, blade_offset, mu, lam, yield_stress, flip;
  double sn, gk, mus, mud, gc, board_sn, threshold;
  double half[3], bevel, blade_mass, inertia_inv[3], board_half[2];
  double drive_k, drive_c, drive_limit, rot_k, rot_c;
  int32_t free_blade, pose_servo, external_tool, predict, accumulate_tool_impulses, cache_topology;
  int32_t cut_work_advance, cut_hold;
  double viscosity, mu_maxwell, maxwell_relax;  // dynamic viscosity; Maxwell branch constants
  int32_t maxwell, tool_count;                  // tool_count knives (> 1: free knives only)
};

template <class R> struct Sim {
  OracleConfig c;
  int64_t n;
  std::vector<R> x, v, F, Fp, stress, damage, ref, weights, gradients, impulses, histories;
  std::vector<double> viscous;  // Fv - I of the Maxwell branch, float64 at either precision
  std::vector<int64_t> codes, field_codes_override;
  std::vector<int32_t> nodes;
  // Knife k's pose, twist, step reaction and table impulse at [3k .. 3k + 2] ([4k .. 4k + 3]).
  std::vector<R> tool_pos, tool_vel, tool_rot, tool_omega, mobility, reaction, torque, board, work, speed, table;
  // Knives k >= 1 (extra_blade, contact.py:500-566): friction history row and damage of each
  // particle, at [(k - 1) n + p].
  std::vector<R> extra_histories, extra_damage;
  int64_t knives = 1;
  std::vector<R> interval_impulse, interval_angular;
  std::vector<int32_t> errors;
  // Registration / grid (compact path).
  std::vector<int64_t> keys, grid_keys;
  std::vector<int32_t> slots, mapping, heads, links;
  int64_t active_count = 0;
  std::vector<R> grid_mass, momentum, grid_impulse, old_v, new_v, remapped, coupled_remapped;
  bool topology_dirty = true;
  bool single_field = false;
  // Pair history tables (ping-pong), keyed by p * n + q.
  std::unordered_map<int64_t, V3<R>> pair_old, pair_new;
  int64_t pair_count = 0;
  int64_t step_count = 0;
  int threads = 1;
  // (particle, corner) entries 8 p + a of each field slot, in particle order (built at
  // registration): the gather form of the grid scatters for threads > 1.
  std::vector<int64_t> slot_start, slot_items;
  void count_error(int k) {
#pragma omp atomic
    errors[k] += 1;
  }
  int64_t cut_index = 0;
  // Held-blade grid constraint state (blade_constraint.py:221-240).
  std::vector<R> bc_side, bc_normal, bc_lever, bc_rhs, bc_bound, bc_impulse, bc_extrap, bc_grid, bc_coupled;
  R bc_max_violation = R(0);
  int64_t bc_batches = 0;

  explicit Sim(const OracleConfig& cfg) : c(cfg), n(cfg.n) {
    x.assign(3 * n, R(0)); v.assign(3 * n, R(0)); ref.assign(3 * n, R(0));
    F.assign(9 * n, R(0)); Fp.assign(9 * n, R(0)); stress.assign(9 * n, R(0));
    for (int64_t p = 0; p < n; ++p) for (int i = 0; i < 3; ++i) F[9 * p + 4 * i] = Fp[9 * p + 4 * i] = R(1);
    damage.assign(n, R(0)); codes.assign(n, 0); nodes.assign(8 * n, 0);
    viscous.assign(9 * n, 0.0);
    weights.assign(8 * n, R(0)); gradients.assign(24 * n, R(0));
    impulses.assign(3 * n, R(0)); histories.assign(6 * n, R(0));
    knives = std::max<int64_t>(1, cfg.tool_count);
    tool_pos.assign(3 * knives, R(0)); tool_vel.assign(3 * knives, R(0)); tool_rot.assign(4 * knives, R(0));
    for (int64_t k = 0; k < knives; ++k) tool_rot[4 * k + 3] = R(1);
    tool_omega.assign(3 * knives, R(0)); mobility.assign(36, R(0)); reaction.assign(3 * knives, R(0));
    torque.assign(3 * knives, R(0));
    board.assign(3, R(0)); work.assign(1, R(0)); speed.assign(1, R(0)); table.assign(3 * knives, R(0));
    extra_histories.assign(3 * n * (knives - 1), R(0)); extra_damage.assign(n * (knives - 1), R(0));
    interval_impulse.assign(3, R(0)); interval_angular.assign(3, R(0));
    errors.assign(3, 0);
    keys.assign(cfg.key_capacity, -1);
    grid_keys.assign(8 * n, -1);
    slots.assign(8 * n, -1); mapping.assign(cfg.key_capacity, -1);
    heads.assign(cfg.node_count, -1); links.assign(8 * n, -1);
    for (auto* g : {&grid_mass}) g->assign(8 * n, R(0));
    for (auto* g : {&momentum, &grid_impulse, &old_v, &new_v, &remapped, &coupled_remapped}) g->assign(24 * n, R(0));
    bc_side.assign(n, R(0)); bc_normal.assign(3 * n, R(0)); bc_lever.assign(3 * n, R(0));
    bc_rhs.assign(n, R(0)); bc_bound.assign(n, R(0)); bc_impulse.assign(n, R(0)); bc_extrap.assign(n, R(0));
    bc_grid.assign(24 * n, R(0)); bc_coupled.assign(24 * n, R(0));
  }

  V3<R> g3(const std::vector<R>& a, int64_t i) const { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
  void s3(std::vector<R>& a, int64_t i, V3<R> v3) { a[3 * i] = v3.x; a[3 * i + 1] = v3.y; a[3 * i + 2] = v3.z; }
  void add3(std::vector<R>& a, int64_t i, V3<R> v3) { a[3 * i] += v3.x; a[3 * i + 1] += v3.y; a[3 * i + 2] += v3.z; }
  M3<R> g9(const std::vector<R>& a, int64_t p) const {
    M3<R> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a[9 * p + 3 * i + j];
    return r;
  }
  void s9(std::vector<R>& a, int64_t p, const M3<R>& m) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[9 * p + 3 * i + j] = m.m[i][j];
  }
  Q<R> rot(int64_t k = 0) const { return {tool_rot[4 * k], tool_rot[4 * k + 1], tool_rot[4 * k + 2], tool_rot[4 * k + 3]}; }
  V3<R> half() const { return {R(c.half[0]), R(c.half[1]), R(c.half[2])}; }

  // ------------------------------------------------------------------ contact helpers
  // friction (contact.py:21-43)
  void friction(V3<R> nrm, V3<R> relative, R jn, V3<R> history, R me, R& jt_out0, V3<R>& jt, V3<R>& ht) {
    (void)jt_out0;
    R dt = R(c.dt);
    V3<R> vt = relative - dot(relative, nrm) * nrm;
    V3<R> htan = history - dot(history, nrm) * nrm;
    R k = R(c.gk) * me / (dt * dt);
    R damping = R(2) * R(c.gc) * std::sqrt(me * k);
    V3<R> trial = ((-k) * htan - damping * vt) * dt;
    R magnitude = length(trial);
    if (magnitude <= R(c.mus) * jn) {
      jt = trial;
      ht = htan + dt * vt;
      return;
    }
    V3<R> applied = trial * (R(c.mud) * jn / magnitude);
    jt = applied;
    ht = -((applied / dt + damping * vt) / k);
  }

  // penalty_impulse (contact.py:123-147); sn is the normal factor of this contact.
  void penalty(R gap, V3<R> relative, V3<R> nrm, R me, V3<R> history, R sn, V3<R>& j, V3<R>& h) {
    R dt = R(c.dt);
    R predicted = gap + dt * std::min(dot(relative, nrm), R(0));
    if (predicted >= R(0)) {
      j = {R(0), R(0), R(0)};
      h = {R(0), R(0), R(0)};
      return;
    }
    R jn = sn * me / dt * (-predicted);
    V3<R> jt, ht;
    R unused = 0;
    friction(nrm, relative, jn, history, me, unused, jt, ht);
    j = jn * nrm + jt;
    h = ht;
  }

  // blade_query (contact.py:46-120)
  static void blade_query(V3<R> local, V3<R> hf, R bevel, V3<R>& closest, V3<R>& normal, R& distance) {
    closest = {std::clamp(local.x, -hf.x, hf.x), std::clamp(local.y, -hf.y, hf.y), std::clamp(local.z, -hf.z, hf.z)};
    V3<R> delta = local - closest;
    distance = length(delta);
    normal = {R(0), R(0), R(0)};
    if (bevel == R(0)) {
      if (distance > R(1e-12)) {
        normal = delta / distance;
      } else {
        V3<R> margin = {hf.x - std::fabs(local.x), hf.y - std::fabs(local.y), hf.z - std::fabs(local.z)};
        int axis = 0;
        if (margin.y < comp(margin, axis)) axis = 1;
        if (margin.z < comp(margin, axis)) axis = 2;
        setc(normal, axis, R(1));
        if (comp(local, axis) < R(0)) setc(normal, axis, R(-1));
        setc(closest, axis, comp(normal, axis) * comp(hf, axis));
        distance = -comp(margin, axis);
      }
      return;
    }
    // Reflected into the +x half: bevel, side and top faces can be nearest.
    R qx = std::fabs(local.x), qz = local.z;
    R sign = local.x < R(0) ? R(-1) : R(1);
    R ex = hf.x, ez = bevel;
    R edge_len_sq = ex * ex + ez * ez;
    R t = std::clamp(((qx - R(0)) * ex + (qz + hf.z) * ez) / edge_len_sq, R(0), R(1));
    R nx = t * hf.x, nz = -hf.z + t * bevel;
    R best_sq = (qx - nx) * (qx - nx) + (qz - nz) * (qz - nz);
    R il = std::sqrt(edge_len_sq);
    R ox = bevel / il, oz = -hf.x / il;
    R cx = hf.x, cz = std::clamp(qz, -hf.z + bevel, hf.z);
    R dsq = (qx - cx) * (qx - cx) + (qz - cz) * (qz - cz);
    if (dsq < best_sq) { best_sq = dsq; nx = cx; nz = cz; ox = R(1); oz = R(0); }
    cx = std::clamp(qx, R(0), hf.x); cz = hf.z;
    dsq = (qx - cx) * (qx - cx) + (qz - cz) * (qz - cz);
    if (dsq < best_sq) { best_sq = dsq; nx = cx; nz = cz; ox = R(0); oz = R(1); }
    bool inside = qx <= hf.x && qz <= hf.z && hf.x * (qz + hf.z) >= bevel * qx;
    R best = std::sqrt(best_sq);
    nx = sign * nx;
    ox = sign * ox;
    closest = {nx, std::clamp(local.y, -hf.y, hf.y), nz};
    if (inside && std::fabs(local.y) <= hf.y) {
      normal = {ox, R(0), oz};
      distance = -best;
      if (hf.y - std::fabs(local.y) < best) {
        R sy = local.y < R(0) ? R(-1) : R(1);
        closest = local;
        closest.y = sy * hf.y;
        normal = {R(0), sy, R(0)};
        distance = -(hf.y - std::fabs(local.y));
      }
    } else {
      if (inside) {
        closest.x = local.x;
        closest.z = local.z;
      }
      delta = local - closest;
      distance = length(delta);
      normal = {ox, R(0), oz};
      if (distance > R(1e-12)) normal = delta / distance;
    }
  }

  // blade_effective_mass (contact.py:150-178); tool_model -1 external, 0 guided, 1 free.
  R effective_mass(V3<R> lever, V3<R> nrm, Q<R> q, int tool_model) {
    R inv = R(1) / R(c.blade_mass);
    if (tool_model < 0) {
      V3<R> rn = cross(lever, nrm);
      R u[6] = {nrm.x, nrm.y, nrm.z, rn.x, rn.y, rn.z};
      inv = R(0);
      for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 6; ++k) inv += u[r] * mobility[6 * r + k] * u[k];
    } else if (tool_model > 0) {
      V3<R> rn = rotate_inv(q, cross(lever, nrm));
      V3<R> ii = {R(c.inertia_inv[0]), R(c.inertia_inv[1]), R(c.inertia_inv[2])};
      inv += dot(rn, V3<R>{ii.x * rn.x, ii.y * rn.y, ii.z * rn.z});
    }
    return R(1) / (R(1) / R(c.mass) + inv);
  }

  // resist_cut (contact.py:181-232); `rad` is the blade contact offset, `dmg` the damage of
  // particle p against this knife.
  V3<R> resist_cut(int64_t p, V3<R> j, V3<R> origin, V3<R> tvel, Q<R> q, V3<R> tomega, R rad, int cutting) {
    R delta = R(0);
    bool worked = false;
    V3<R> out = resist_cut(p, j, origin, tvel, q, tomega, rad, cutting, damage[p], delta, worked);
    if (worked) work[0] += delta;
    return out;
  }
  // The work done comes back in `delta` (with `worked` set) for the caller to add.
  V3<R> resist_cut(int64_t p, V3<R> j, V3<R> origin, V3<R> tvel, Q<R> q, V3<R> tomega, R rad, int cutting, R& dmg,
                   R& delta_out, bool& worked) {
    V3<R> hf = half();
    V3<R> edge = origin + rotate(q, V3<R>{R(0), R(0), -hf.z});
    V3<R> down = rotate(q, V3<R>{R(0), R(0), R(-1)});
    V3<R> edge_velocity = tvel + cross(tomega, edge - origin);
    V3<R> xp = g3(x, p);
    R front = dot(xp - edge, down);
    R advance = dot(edge_velocity - g3(v, p), down);
    if (cutting & CUT_ADVANCE) advance = dot(edge_velocity, down);
    V3<R> local_edge = rotate_inv(q, xp - edge);
    bool ahead = front >= R(0) && front <= R(2) * rad && std::fabs(local_edge.x) <= R(6) * R(c.spacing) &&
                 std::fabs(local_edge.y) <= hf.y;
    if (!ahead) return j;
    R push = std::max(dot(j, down), R(0));
    if (dmg >= R(1)) return j - push * down;
    if (advance > R(0)) {
      R delta = push * advance;
      delta_out = delta;
      worked = true;
      dmg = std::min(R(1), dmg + delta / R(c.threshold));
    }
    return j;
  }

  // blade_at (contact.py:235-329) of knife k: knife 0 keeps its friction history in row 1 of
  // `histories` and its damage in `damage`; knife k >= 1 (extra_blade, contact.py:500-566) in
  // extra_histories and extra_damage.  The reaction and torque go to knife k's slots.
  // One particle's contribution to a knife's step sums.
  struct KnifeSums {
    V3<R> react{R(0), R(0), R(0)}, torq{R(0), R(0), R(0)};
    R work = R(0);
    bool touched = false, worked = false;
  };
  V3<R> blade_at(int64_t p, int tool_model, int cutting, int64_t k, KnifeSums& sums) {
    V3<R> hf = half();
    R dt = R(c.dt);
    V3<R> xp = g3(x, p), vp = g3(v, p);
    Q<R> q = rot(k);
    V3<R> origin = g3(tool_pos, k), tvel = g3(tool_vel, k), tom = g3(tool_omega, k);
    std::vector<R>& hist = k == 0 ? histories : extra_histories;
    const int64_t row = k == 0 ? n + p : (k - 1) * n + p;
    R& dmg = k == 0 ? damage[p] : extra_damage[(k - 1) * n + p];
    V3<R> result = {R(0), R(0), R(0)};
    R blade_radius = length(hf);
    R max_closing = length(vp) + length(tvel) + length(tom) * blade_radius;
    R offset = R(c.blade_offset);
    if (length(xp - origin) - blade_radius - offset > dt * max_closing) {
      s3(hist, row, {R(0), R(0), R(0)});
      return result;
    }
    V3<R> local = rotate_inv(q, xp - origin);
    V3<R> closest, normal;
    R distance;
    blade_query(local, hf, R(c.bevel), closest, normal, distance);
    R front = -hf.z - local.z;
    if ((cutting & CUT_HOLD) && dmg < R(1) && std::fabs(local.x) <= hf.x + offset && std::fabs(local.y) <= hf.y &&
        front >= R(0)) {
      closest = {local.x, local.y, -hf.z};
      normal = {R(0), R(0), R(-1)};
      distance = front;
    }
    V3<R> lever = rotate(q, closest);
    V3<R> nrm = rotate(q, normal);
    V3<R> relative = vp - tvel - cross(tom, lever);
    if (distance - offset + dt * std::min(dot(relative, nrm), R(0)) >= R(0)) {
      s3(hist, row, {R(0), R(0), R(0)});
      return result;
    }
    R me = effective_mass(lever, nrm, q, tool_model);
    V3<R> j, h;
    penalty(distance - offset, relative, nrm, me, g3(hist, row), R(c.sn), j, h);
    s3(hist, row, h);
    if (dot(j, j) > R(0)) {
      if (cutting != 0) j = resist_cut(p, j, origin, tvel, q, tom, offset, cutting, dmg, sums.work, sums.worked);
      result = j;
      sums.touched = true;
      sums.react = -j;
      sums.torq = -cross(lever, j);
    }
    return result;
  }

  // rigid_and_speed / rigid_at (contact.py:332-497), then extra_blade (contact.py:500-566) for
  // knives 1 .. K-1 after all of knife 0.  The per-particle contacts are independent; the
  // board, reaction, torque and work sums are added afterwards in particle order (knife by
  // knife for the work), the order of a sequential pass.
  void rigid_contacts(int tool_model, int cutting) {
    R dt = R(c.dt);
    std::vector<V3<R>> board_j(n);
    std::vector<char> board_hit(n, 0);
    std::vector<KnifeSums> sums(size_t(n) * size_t(knives));
    R spd = speed[0];
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1) reduction(max : spd)
    for (int64_t p = 0; p < n; ++p) {
      spd = std::max(spd, length(g3(v, p)));
      V3<R> xp = g3(x, p);
      V3<R> total = {R(0), R(0), R(-9.81) * R(c.mass) * dt};
      R gap = xp.z - R(c.radius);
      if (c.board_half[0] > 0 && std::fabs(xp.x) > R(c.board_half[0])) gap = R(1e30);
      if (c.board_half[1] > 0 && std::fabs(xp.y) > R(c.board_half[1])) gap = R(1e30);
      V3<R> jb, hb;
      penalty(gap, g3(v, p), V3<R>{R(0), R(0), R(1)}, R(c.mass), g3(histories, p), R(c.board_sn), jb, hb);
      s3(histories, p, hb);
      if (dot(jb, jb) > R(0)) {
        total = total + jb;
        board_j[p] = jb;
        board_hit[p] = 1;
      }
      total = total + blade_at(p, tool_model, cutting, 0, sums[size_t(p) * knives]);
      s3(impulses, p, total);
    }
    for (int64_t k = 1; k < knives; ++k) {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
      for (int64_t p = 0; p < n; ++p) add3(impulses, p, blade_at(p, tool_model, cutting, k, sums[size_t(p) * knives + k]));
    }
    speed[0] = spd;
    for (int64_t p = 0; p < n; ++p)
      if (board_hit[p]) add3(board, 0, board_j[p]);
    for (int64_t k = 0; k < knives; ++k)
      for (int64_t p = 0; p < n; ++p) {
        const KnifeSums& t = sums[size_t(p) * knives + k];
        if (t.worked) work[0] += t.work;
        if (t.touched) {
          add3(reaction, k, t.react);
          add3(torque, k, t.torq);
        }
      }
  }

  // True when pair_at(p, q) returns at one of its first tests, which have no effect: beyond
  // contact reach this step, or compatible particles that were neighbours in the reference.
  bool pair_inert(int64_t p, int64_t q) const {
    R dt = R(c.dt), rad = R(c.radius);
    V3<R> delta = g3(x, p) - g3(x, q);
    R dist = length(delta);
    if (dist >= R(2) * rad + dt * length(g3(v, p) - g3(v, q))) return true;
    return compatible(codes[p], codes[q]) && length(g3(ref, p) - g3(ref, q)) < R(2) * R(c.spacing);
  }

  // pair_at (contact.py:584-640)
  void pair_at(int64_t p, int64_t q) {
    R dt = R(c.dt), rad = R(c.radius);
    V3<R> delta = g3(x, p) - g3(x, q);
    R dist = length(delta);
    if (dist >= R(2) * rad + dt * length(g3(v, p) - g3(v, q))) return;
    if (compatible(codes[p], codes[q]) && length(g3(ref, p) - g3(ref, q)) < R(2) * R(c.spacing)) return;
    if (dist < R(1e-12)) {
      errors[2] += 1;
      return;
    }
    V3<R> normal = delta / dist;
    V3<R> relative = g3(v, p) - g3(v, q);
    R predicted = R(2) * rad - dist - dt * std::min(dot(relative, normal), R(0));
    if (predicted <= R(0)) return;
    int64_t key = p * n + q;
    V3<R> history = {R(0), R(0), R(0)};
    auto it = pair_old.find(key);
    if (it != pair_old.end()) history = it->second;
    if (int64_t(pair_new.size()) >= c.pair_capacity && pair_new.find(key) == pair_new.end()) {
      errors[0] += 1;
      return;
    }
    R half_mass = R(c.mass) / R(2);
    R jn = R(c.sn) * half_mass / dt * predicted;
    V3<R> jt, ht;
    R unused = 0;
    friction(normal, relative, jn, history, half_mass, unused, jt, ht);
    pair_new[key] = ht;
    pair_count += 1;
    V3<R> impulse = jn * normal + jt;
    add3(impulses, p, impulse);
    add3(impulses, q, -impulse);
  }

  // particle_contacts (solver.py:670-751): candidate pairs within 2r + 2 dt max|v|.
  void particle_contacts() {
    std::swap(pair_old, pair_new);
    pair_new.clear();
    pair_count = 0;
    R dt = R(c.dt);
    double reach = double(R(2) * R(c.radius) + R(2) * dt * speed[0]);
    double cell = reach > 0 ? reach : 1e-3;
    std::unordered_map<int64_t, std::vector<int64_t>> bins;
    auto cell_of = [&](int64_t p, int64_t out[3]) {
      for (int i = 0; i < 3; ++i) out[i] = int64_t(std::floor(double(x[3 * p + i]) / cell));
    };
    auto pack = [](int64_t a, int64_t b, int64_t c3) { return ((a + (1 << 20)) << 42) ^ ((b + (1 << 20)) << 21) ^ (c3 + (1 << 20)); };
    bool finite = true;
    for (int64_t p = 0; p < n && finite; ++p)
      for (int i = 0; i < 3; ++i) finite = finite && std::isfinite(double(x[3 * p + i])) && std::fabs(double(x[3 * p + i])) < 1e5;
    if (!finite) return;
    for (int64_t p = 0; p < n; ++p) {
      int64_t cc[3];
      cell_of(p, cc);
      bins[pack(cc[0], cc[1], cc[2])].push_back(p);
    }
    // Candidates (p, q > p) within reach, in (p, q) order: found in parallel over contiguous
    // particle ranges (one per thread, concatenated in order), then pair_at in that order.
    // Pairs that pair_at would leave without any effect (too far for this step, or
    // compatible neighbours in the reference) are dropped in the parallel pass.
    const int chunks = std::max(1, threads);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> found(chunks);
#pragma omp parallel for schedule(static, 1) num_threads(threads) if (threads > 1)
    for (int t = 0; t < chunks; ++t) {
      std::vector<int64_t> cand;
      const int64_t first = n * t / chunks, last = n * (t + 1) / chunks;
      for (int64_t p = first; p < last; ++p) {
        int64_t cc[3];
        cell_of(p, cc);
        cand.clear();
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
              auto it = bins.find(pack(cc[0] + dx, cc[1] + dy, cc[2] + dz));
              if (it == bins.end()) continue;
              for (int64_t q : it->second)
                if (q > p) cand.push_back(q);
            }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        for (int64_t q : cand) {
          V3<R> d = g3(x, q) - g3(x, p);
          R d2 = dot(d, d);
          R r = R(reach);
          if (d2 >= r * r * R(1.0001) + R(1e-30)) continue;
          if (pair_inert(p, q)) continue;
          found[t].emplace_back(p, q);
        }
      }
    }
    for (const auto& list : found)
      for (const auto& pq : list) pair_at(pq.first, pq.second);
  }

  // ---------------------------------------------------------------- material
  const std::vector<int64_t>& transfer_codes() const { return single_field ? field_codes_override : codes; }

  // register + compact_keys + remap_slots + compact_node_lists (solver.py:391-429)
  void register_topology() {
    std::fill(keys.begin(), keys.end(), int64_t(-1));
    std::fill(heads.begin(), heads.end(), -1);
    const auto& fc = transfer_codes();
    for (int64_t p = 0; p < n; ++p)
      for (int a = 0; a < 8; ++a) {
        int64_t key = fc[p] * c.node_count + int64_t(nodes[8 * p + a]);
        int64_t s = insert_key(keys, key);
        slots[8 * p + a] = int32_t(s);
        if (s < 0) errors[0] += 1;
      }
    active_count = 0;
    for (int64_t h = 0; h < int64_t(keys.size()); ++h)
      if (keys[h] >= 0) {
        mapping[h] = int32_t(active_count);
        grid_keys[active_count] = keys[h];
        active_count += 1;
      }
    for (int64_t i = 0; i < 8 * n; ++i)
      if (slots[i] >= 0) slots[i] = mapping[slots[i]];
    for (int64_t a = 0; a < active_count; ++a) {
      int64_t node = grid_keys[a] % c.node_count;
      links[a] = heads[node];
      heads[node] = int32_t(a);
    }
    // Field slot -> its (particle, corner) entries in particle order (counting sort).
    slot_start.assign(active_count + 1, 0);
    for (int64_t i = 0; i < 8 * n; ++i)
      if (slots[i] >= 0) slot_start[slots[i] + 1] += 1;
    for (int64_t a = 0; a < active_count; ++a) slot_start[a + 1] += slot_start[a];
    slot_items.assign(slot_start[active_count], 0);
    std::vector<int64_t> fill(slot_start.begin(), slot_start.end() - 1);
    for (int64_t i = 0; i < 8 * n; ++i)
      if (slots[i] >= 0) slot_items[fill[slots[i]]++] = i;
    topology_dirty = false;
  }

  // The field slot lists match the current slots (not after host writes to `slots`).
  bool gather_ready() const {
    return threads > 1 && int64_t(slot_start.size()) == active_count + 1;
  }

  // p2g (material.py:127-152)
  void p2g() {
    R dt = R(c.dt), mass = R(c.mass), vol = R(c.volume);
    std::fill(grid_mass.begin(), grid_mass.begin() + active_count, R(0));
    std::fill(momentum.begin(), momentum.begin() + 3 * active_count, R(0));
    std::fill(grid_impulse.begin(), grid_impulse.begin() + 3 * active_count, R(0));
    std::fill(remapped.begin(), remapped.begin() + 3 * active_count, R(0));
    if (gather_ready()) {
      // Each slot sums its entries in particle order: the same sums as the scatter below.
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
      for (int64_t s = 0; s < active_count; ++s)
        for (int64_t e = slot_start[s]; e < slot_start[s + 1]; ++e) {
          int64_t i = slot_items[e], p = i >> 3;
          int a = int(i & 7);
          R w = weights[i];
          V3<R> g = {gradients[24 * p + 3 * a], gradients[24 * p + 3 * a + 1], gradients[24 * p + 3 * a + 2]};
          grid_mass[s] += mass * w;
          add3(momentum, s, (mass * w) * g3(v, p));
          add3(grid_impulse, s, w * g3(impulses, p) - mulv(scale(dt * vol, g9(stress, p)), g));
        }
      return;
    }
    for (int64_t p = 0; p < n; ++p) {
      M3<R> P = g9(stress, p);
      V3<R> vp = g3(v, p), imp = g3(impulses, p);
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s < 0) continue;
        R w = weights[8 * p + a];
        V3<R> g = {gradients[24 * p + 3 * a], gradients[24 * p + 3 * a + 1], gradients[24 * p + 3 * a + 2]};
        grid_mass[s] += mass * w;
        add3(momentum, s, (mass * w) * vp);
        add3(grid_impulse, s, w * imp - mulv(scale(dt * vol, P), g));
      }
    }
  }

  // couple_and_velocity_at (material.py:252-285)
  void couple_and_velocity() {
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (int64_t a = 0; a < active_count; ++a) {
      if (grid_keys[a] < 0 || grid_mass[a] <= R(0)) continue;
      int64_t node = grid_keys[a] % c.node_count;
      int64_t code_a = grid_keys[a] / c.node_count;
      R total = R(0);
      V3<R> dm = {R(0), R(0), R(0)}, di = dm;
      for (int64_t b = heads[node]; b >= 0; b = links[b]) {
        total += grid_mass[b];
        int64_t code_b = grid_keys[b] / c.node_count;
        if (b != a && compatible(code_a, code_b)) {
          dm = dm + (grid_mass[a] * g3(momentum, b) - grid_mass[b] * g3(momentum, a));
          di = di + (grid_mass[a] * g3(grid_impulse, b) - grid_mass[b] * g3(grid_impulse, a));
        }
      }
      V3<R> pm = g3(momentum, a) + dm / total;
      V3<R> jm = g3(grid_impulse, a) + di / total;
      s3(old_v, a, pm / grid_mass[a]);
      s3(new_v, a, (pm + jm) / grid_mass[a]);
    }
  }

  // couple_at (material.py:155-179)
  void couple(const std::vector<R>& state, std::vector<R>& result) {
    couple_fields<R>(grid_keys.data(), heads.data(), links.data(), grid_mass.data(), state.data(), result.data(),
                     active_count, c.node_count, threads);
  }

  // g2p_remap (material.py:343-401)
  void g2p_remap() {
    R mass = R(c.mass);
    g2p_fields<R>(x.data(), v.data(), slots.data(), weights.data(), old_v.data(), new_v.data(), R(c.dt), R(c.flip), n,
                  threads);
    if (gather_ready()) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
      for (int64_t s = 0; s < active_count; ++s)
        for (int64_t e = slot_start[s]; e < slot_start[s + 1]; ++e) {
          int64_t i = slot_items[e];
          add3(remapped, s, (mass * weights[i]) * g3(v, i >> 3));
        }
      return;
    }
    for (int64_t p = 0; p < n; ++p) {
      V3<R> vp = g3(v, p);
      for (int a = 0; a < 8; ++a) add3(remapped, slots[8 * p + a], (mass * weights[8 * p + a]) * vp);
    }
  }

  // corotated_stress (material.py:78-91)
  M3<R> corotated(const M3<R>& U, V3<R> sg, const M3<R>& V, const M3<R>& plastic_inverse) {
    return corotated(U, sg, V, plastic_inverse, R(c.mu), R(c.lam));
  }
  M3<R> corotated(const M3<R>& U, V3<R> sg, const M3<R>& V, const M3<R>& plastic_inverse, R mu, R lam) {
    R J = sg.x * sg.y * sg.z;
    V3<R> d;
    for (int i = 0; i < 3; ++i) setc(d, i, R(2) * mu * (comp(sg, i) - R(1)) + lam * (J - R(1)) * J / comp(sg, i));
    M3<R> D = zero3<R>();
    D.m[0][0] = d.x; D.m[1][1] = d.y; D.m[2][2] = d.z;
    return mul(mul(mul(U, D), transpose(V)), transpose(plastic_inverse));
  }

  // stress_at (material.py:94-112)
  void stress_at(int64_t p) {
    M3<R> pinv = inverse(g9(Fp, p));
    M3<R> f = mul(g9(F, p), pinv);
    R J = det(f);
    if (J <= R(0) || !std::isfinite(double(J))) {
      count_error(1);
      s9(stress, p, zero3<R>());
      return;
    }
    M3<R> U, V;
    V3<R> sg;
    svd3(f, U, sg, V);
    s9(stress, p, corotated(U, sg, V, pinv));
  }

  // plastic_at (material.py:404-431); returns false when inverted.
  bool plastic_at(int64_t p, M3<R>& U, V3<R>& sg, M3<R>& V) {
    svd3(mul(g9(F, p), inverse(g9(Fp, p))), U, sg, V);
    if (std::min(std::min(sg.x, sg.y), sg.z) <= R(0)) {
      count_error(1);
      return false;
    }
    V3<R> e = {std::log(sg.x), std::log(sg.y), std::log(sg.z)};
    R mean = (e.x + e.y + e.z) / R(3);
    V3<R> dev = e - V3<R>{mean, mean, mean};
    R norm = length(dev);
    R limit = R(c.yield_stress) / (R(2) * R(c.mu)) * std::sqrt(R(2) / R(3));
    if (norm > limit) {
      V3<R> ep = e - (R(1) - limit / norm) * dev;
      sg = {std::exp(ep.x), std::exp(ep.y), std::exp(ep.z)};
      M3<R> D = zero3<R>();
      D.m[0][0] = sg.x; D.m[1][1] = sg.y; D.m[2][2] = sg.z;
      M3<R> next_f = mul(mul(U, D), transpose(V));
      s9(Fp, p, mul(inverse(next_f), g9(F, p)));
    }
    return true;
  }

  // maxwell_at (material.py:485-535): generalized standard linear solid on the elastic part
  // f = U diag(sg) V^T.  The Maxwell branch f_m = f (I + D)^-1, D = viscous[p] (float64,
  // rounded to R where it enters f_m as the reference does), has Kirchhoff stress
  // 2 mu_m dev(log s_m) in the basis U_m; relax > 0 first moves that fraction of the
  // deviatoric Hencky strain into Fv: D <- D + delta + delta D with
  // delta = V_m diag(expm1(relax dev)) V_m^T.  (The reference evaluates expm1 by a
  // three-term series below |x| = 1e-3, exact to far below R's resolution there; std::expm1
  // is used here.)  Returns equilibrium(mu - mu_m, lam + 2/3 mu_m) + tau_m F^-T.
  M3<R> maxwell_at(int64_t p, const M3<R>& U, V3<R> sg, const M3<R>& V, R relax) {
    M3<R> S = zero3<R>();
    S.m[0][0] = sg.x; S.m[1][1] = sg.y; S.m[2][2] = sg.z;
    M3<R> f = mul(mul(U, S), transpose(V));
    double* D = &viscous[9 * p];
    M3<R> Fv = eye3<R>();
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Fv.m[i][j] += R(D[3 * i + j]);
    M3<R> Um, Vm;
    V3<R> sm;
    svd3(mul(f, inverse(Fv)), Um, sm, Vm);
    if (std::min(std::min(sm.x, sm.y), sm.z) <= R(0)) {
      count_error(1);
      return zero3<R>();
    }
    V3<R> e = {std::log(sm.x), std::log(sm.y), std::log(sm.z)};
    R mean = (e.x + e.y + e.z) / R(3);
    V3<R> dev = e - V3<R>{mean, mean, mean};
    if (relax > R(0)) {
      V3<R> step = relax * dev;
      dev = dev - step;
      M3<R> G = zero3<R>();
      G.m[0][0] = std::expm1(step.x); G.m[1][1] = std::expm1(step.y); G.m[2][2] = std::expm1(step.z);
      M3<R> delta = mul(mul(Vm, G), transpose(Vm));
      double next[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double dd = 0.0;
          for (int k = 0; k < 3; ++k) dd += double(delta.m[i][k]) * D[3 * k + j];
          next[3 * i + j] = D[3 * i + j] + double(delta.m[i][j]) + dd;
        }
      for (int k = 0; k < 9; ++k) D[k] = next[k];
    }
    R mu_m = R(c.mu_maxwell);
    M3<R> K = zero3<R>();
    K.m[0][0] = R(2) * mu_m * dev.x; K.m[1][1] = R(2) * mu_m * dev.y; K.m[2][2] = R(2) * mu_m * dev.z;
    M3<R> kirchhoff = mul(mul(Um, K), transpose(Um));
    M3<R> equilibrium = corotated(U, sg, V, inverse(g9(Fp, p)), R(c.mu) - mu_m, R(c.lam) + R(2) / R(3) * mu_m);
    return addm(equilibrium, mul(kirchhoff, transpose(inverse(g9(F, p)))));
  }

  // maxwell_stress_at (material.py:538-566) with project; compute_stress_maxwell
  // (material.py:569-587) without it (relax 0 there).
  void maxwell_stress_at(int64_t p, R relax, bool project) {
    M3<R> U, V;
    V3<R> sg;
    if (project && c.yield_stress > 0) {
      if (!plastic_at(p, U, sg, V)) {
        s9(stress, p, zero3<R>());
        return;
      }
    } else {
      M3<R> f = mul(g9(F, p), inverse(g9(Fp, p)));
      R J = det(f);
      if (J <= R(0) || !std::isfinite(double(J))) {
        count_error(1);
        s9(stress, p, zero3<R>());
        return;
      }
      svd3(f, U, sg, V);
    }
    s9(stress, p, maxwell_at(p, U, sg, V, relax));
  }

  // refresh_stress (solver.py:318-329)
  void refresh_stress() {
    for (int64_t p = 0; p < n; ++p) {
      if (c.maxwell) maxwell_stress_at(p, R(0), false);
      else stress_at(p);
    }
  }

  // viscous_impulse (material.py:321-340): with the coupled old velocities at the particle's
  // slots, L = (sum_a old_v_a (x) grad w_a) F^-1, stress = eta J (L + L^T) F^-T, and
  // -dt V0 stress grad w_a added to grid_impulse at each slot.
  void viscous_impulse() {
    R eta = R(c.viscosity), dt = R(c.dt), vol = R(c.volume);
    for (int64_t p = 0; p < n; ++p) {
      M3<R> rate = zero3<R>();
      for (int a = 0; a < 8; ++a) {
        if (slots[8 * p + a] < 0) continue;  // unregistered (registration fills every slot)
        V3<R> g = {gradients[24 * p + 3 * a], gradients[24 * p + 3 * a + 1], gradients[24 * p + 3 * a + 2]};
        rate = addm(rate, outer(g3(old_v, slots[8 * p + a]), g));
      }
      M3<R> Fm = g9(F, p);
      M3<R> inv = inverse(Fm);
      M3<R> L = mul(rate, inv);
      M3<R> P = mul(scale(eta * det(Fm), addm(L, transpose(L))), transpose(inv));
      for (int a = 0; a < 8; ++a) {
        if (slots[8 * p + a] < 0) continue;
        V3<R> g = {gradients[24 * p + 3 * a], gradients[24 * p + 3 * a + 1], gradients[24 * p + 3 * a + 2]};
        add3(grid_impulse, slots[8 * p + a], mulv(scale(-dt * vol, P), g));
      }
    }
  }

  // grid_vel (material.py:212-236) over the active fields.
  void grid_vel(const std::vector<R>& mom, const std::vector<R>& imp) {
    for (int64_t s = 0; s < active_count; ++s) {
      if (grid_mass[s] <= R(0)) continue;
      s3(old_v, s, g3(mom, s) / grid_mass[s]);
      s3(new_v, s, (g3(mom, s) + g3(imp, s)) / grid_mass[s]);
    }
  }

  // update_F_plastic (material.py:445-475), reading the coupled re-scattered momentum.
  void update_F() {
    R dt = R(c.dt);
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (int64_t p = 0; p < n; ++p) {
      M3<R> rate = zero3<R>();
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s >= 0 && grid_mass[s] > R(0)) {
          V3<R> g = {gradients[24 * p + 3 * a], gradients[24 * p + 3 * a + 1], gradients[24 * p + 3 * a + 2]};
          rate = addm(rate, outer(g3(coupled_remapped, s) / grid_mass[s], g));
        }
      }
      s9(F, p, addm(g9(F, p), scale(dt, rate)));
      if (c.maxwell) {
        // update_F_maxwell (material.py:590-618)
        maxwell_stress_at(p, R(c.maxwell_relax), true);
      } else if (c.yield_stress > 0) {
        M3<R> U, V;
        V3<R> sg;
        if (plastic_at(p, U, sg, V))
          s9(stress, p, corotated(U, sg, V, inverse(g9(Fp, p))));
        else
          s9(stress, p, zero3<R>());
      } else {
        stress_at(p);
      }
    }
  }


  // ---------------------------------------------------------------- held-blade constraint
  // BladeConstraint.solve (blade_constraint.py:254-324) with prepare (:13-70), row_bound
  // (:73-90), scatter (:93-111), project (:130-152), record_residual (:155-168),
  // finish_refinement (:171-183), apply_grid (:186-194), apply_reaction (:197-205).
  V3<R> bc_response(int64_t p) {
    V3<R> r = {R(0), R(0), R(0)};
    for (int a = 0; a < 8; ++a) {
      int32_t s = slots[8 * p + a];
      if (s >= 0 && grid_mass[s] > R(0)) r = r + weights[8 * p + a] * g3(bc_coupled, s) / grid_mass[s];
    }
    return r;
  }
  void bc_scatter(const std::vector<R>& values, V3<R>& linear, V3<R>& angular) {
    std::fill(bc_grid.begin(), bc_grid.begin() + 3 * active_count, R(0));
    linear = {R(0), R(0), R(0)};
    angular = linear;
    for (int64_t p = 0; p < n; ++p) {
      V3<R> j = values[p] * g3(bc_normal, p);
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s >= 0) add3(bc_grid, s, weights[8 * p + a] * j);
      }
      linear = linear + j;
      angular = angular + cross(g3(bc_lever, p), j);
    }
    couple(bc_grid, bc_coupled);
  }
  void blade_constraint() {
    std::fill(bc_grid.begin(), bc_grid.begin() + 3 * active_count, R(0));
    std::fill(bc_coupled.begin(), bc_coupled.begin() + 3 * active_count, R(0));
    bool needs = false;
    Q<R> q = rot();
    V3<R> pos = g3(tool_pos, 0), hf = half();
    R bevel = R(c.bevel), radius = R(c.blade_offset), dt = R(c.dt);
    for (int64_t p = 0; p < n; ++p) {
      V3<R> local = rotate_inv(q, g3(x, p) - pos);
      s3(bc_normal, p, {R(0), R(0), R(0)});
      bc_impulse[p] = R(0);
      bc_extrap[p] = R(0);
      if (std::fabs(local.y) >= hf.y || local.z <= -hf.z + bevel || local.z >= hf.z) {
        bc_side[p] = R(0);
        continue;
      }
      V3<R> grid_v = {R(0), R(0), R(0)};
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s >= 0) grid_v = grid_v + weights[8 * p + a] * g3(new_v, s);
      }
      R sd = bc_side[p];
      if (sd == R(0) || std::fabs(local.x) > hf.x + radius) sd = local.x < R(0) ? R(-1) : R(1);
      V3<R> nrm = rotate(q, V3<R>{sd, R(0), R(0)});
      R gap = sd * local.x - hf.x;
      if (gap > radius && gap + dt * dot(grid_v, nrm) > R(0)) {
        bc_side[p] = R(0);
        continue;
      }
      bc_side[p] = sd;
      s3(bc_normal, p, nrm);
      s3(bc_lever, p, rotate(q, V3<R>{sd * hf.x, local.y, local.z}));
      bc_rhs[p] = -std::max(gap, R(0)) / dt - dot(grid_v, nrm);
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s >= 0) bc_grid[3 * s] += weights[8 * p + a];
      }
      if (bc_rhs[p] > R(1e-7)) needs = true;
    }
    if (!needs) return;
    couple(bc_grid, bc_coupled);
    for (int64_t p = 0; p < n; ++p) {
      R total = R(0);
      for (int a = 0; a < 8; ++a) {
        int32_t s = slots[8 * p + a];
        if (s >= 0 && grid_mass[s] > R(0)) total += weights[8 * p + a] * bc_coupled[3 * s] / grid_mass[s];
      }
      bc_bound[p] = total;
    }
    bc_batches = 0;
    V3<R> linear, angular;
    while (true) {
      bc_extrap = bc_impulse;
      for (int it = 0; it < 48; ++it) {
        bc_scatter(bc_extrap, linear, angular);
        R accel = R(it) / R(it + 3);
        for (int64_t p = 0; p < n; ++p) {
          V3<R> nrm = g3(bc_normal, p);
          if (dot(nrm, nrm) == R(0)) continue;
          V3<R> r = bc_response(p);
          R previous = bc_impulse[p];
          R current = std::max(R(0), bc_extrap[p] + (bc_rhs[p] - dot(nrm, r)) / std::max(bc_bound[p], R(1e-12)));
          bc_impulse[p] = current;
          bc_extrap[p] = current + accel * (current - previous);
        }
      }
      bc_scatter(bc_impulse, linear, angular);
      R residual = R(0);
      for (int64_t p = 0; p < n; ++p) {
        V3<R> nrm = g3(bc_normal, p);
        if (dot(nrm, nrm) > R(0)) residual = std::max(residual, std::max(R(0), bc_rhs[p] - dot(nrm, bc_response(p))));
      }
      bc_batches += 1;
      if (residual > R(1e-6) && bc_batches < 8) continue;
      bc_max_violation = std::max(bc_max_violation, residual);
      break;
    }
    for (int64_t s = 0; s < active_count; ++s)
      if (grid_mass[s] > R(0)) s3(new_v, s, g3(new_v, s) + g3(bc_coupled, s) / grid_mass[s]);
    s3(reaction, 0, g3(reaction, 0) - linear);
    s3(torque, 0, g3(torque, 0) - angular);
  }

  void material_step() {
    if (!c.cache_topology || topology_dirty) register_topology();
    p2g();
    if (c.viscosity > 0) {
      // The viscous path of solver.py:433-458: coupled momentum, old_v from it, the viscous
      // impulse on top of the P2G impulse, then both coupled together.
      std::vector<R> coupled_momentum(momentum.size(), R(0)), coupled_impulse(grid_impulse.size(), R(0));
      couple(momentum, coupled_momentum);
      grid_vel(coupled_momentum, grid_impulse);
      viscous_impulse();
      couple(grid_impulse, coupled_impulse);
      grid_vel(coupled_momentum, coupled_impulse);
    } else {
      couple_and_velocity();
    }
    if (c.external_tool && !c.predict) blade_constraint();
    g2p_remap();
    couple(remapped, coupled_remapped);
    update_F();
  }

  // ---------------------------------------------------------------- tool
  // finish_external_step (tool.py:70-115)
  void finish_external() {
    V3<R> j = g3(reaction, 0), tau = g3(torque, 0);
    add3(interval_impulse, 0, j);
    add3(interval_angular, 0, tau + cross(g3(tool_pos, 0), j));
    if (c.predict) {
      R u[6] = {j.x, j.y, j.z, tau.x, tau.y, tau.z};
      V3<R> dv = {R(0), R(0), R(0)}, dw = dv;
      for (int r = 0; r < 3; ++r) {
        R a = R(0), b = R(0);
        for (int k = 0; k < 6; ++k) {
          a += mobility[6 * r + k] * u[k];
          b += mobility[6 * (r + 3) + k] * u[k];
        }
        setc(dv, r, a);
        setc(dw, r, b);
      }
      s3(tool_vel, 0, g3(tool_vel, 0) + dv);
      V3<R> w = g3(tool_omega, 0) + dw;
      s3(tool_omega, 0, w);
      s3(tool_pos, 0, g3(tool_pos, 0) + R(c.dt) * g3(tool_vel, 0));
      R rate = length(w);
      if (rate > R(1e-12)) {
        Q<R> nq = qnormalize(qmul(axis_angle(w / rate, R(c.dt) * rate), rot()));
        tool_rot = {nq.x, nq.y, nq.z, nq.w};
      }
    }
    s3(reaction, 0, {R(0), R(0), R(0)});
    s3(torque, 0, {R(0), R(0), R(0)});
    s3(board, 0, {R(0), R(0), R(0)});
    speed[0] = R(0);
  }

  // accumulate_reaction (tool.py:6-16)
  void accumulate_reaction() {
    V3<R> j = g3(reaction, 0);
    add3(interval_impulse, 0, j);
    add3(interval_angular, 0, g3(torque, 0) + cross(g3(tool_pos, 0), j));
  }

  static V3<R> vertex(int i, V3<R> hf, R bevel) {
    int j = i % 5;
    R vx = R(0), vz = -hf.z;
    if (j == 1) { vx = hf.x; vz += bevel; }
    else if (j == 2) { vx = hf.x; vz = hf.z; }
    else if (j == 3) { vx = -hf.x; vz = hf.z; }
    else if (j == 4) { vx = -hf.x; vz += bevel; }
    R vy = i >= 5 ? hf.y : -hf.y;
    return {vx, vy, vz};
  }
  static V3<R> inv_inertia(Q<R> q, V3<R> ii, V3<R> t) {
    V3<R> l = rotate_inv(q, t);
    return rotate(q, V3<R>{ii.x * l.x, ii.y * l.y, ii.z * l.z});
  }

  // tool.step (tool.py:146-259)
  void tool_step(V3<R> target, Q<R> target_q, int64_t k = 0) {
    R dt = R(c.dt), mass = R(c.blade_mass);
    V3<R> hf = half();
    R bevel = R(c.bevel);
    V3<R> xt = g3(tool_pos, k);
    Q<R> q = qnormalize(rot(k));
    V3<R> vt = g3(tool_vel, k), w = g3(tool_omega, k);
    V3<R> force = {R(0), R(0), R(-9.81) * mass};
    V3<R> servo = {R(0), R(0), R(0)};
    int free_blade = c.free_blade;
    int lock_rotation = !(c.free_blade || c.pose_servo);
    if (free_blade == 0) {
      V3<R> target_center = target + rotate(target_q, V3<R>{R(0), R(0), hf.z});
      force = R(c.drive_k) * (target_center - xt) - R(c.drive_c) * vt;
      R norm = length(force);
      if (norm > R(c.drive_limit)) force = force * (R(c.drive_limit) / norm);
      Q<R> qinv = {-q.x, -q.y, -q.z, q.w};
      Q<R> dq = qnormalize(qmul(target_q, qinv));
      if (dq.w < R(0)) dq = {-dq.x, -dq.y, -dq.z, -dq.w};
      V3<R> vec = qv(dq);
      R s = length(vec);
      V3<R> error = R(2) * vec;
      if (s > R(1e-8)) error = vec * (R(2) * std::atan2(s, dq.w) / s);
      servo = R(c.rot_k) * error - R(c.rot_c) * w;
    }
    vt = vt + (force * dt + g3(reaction, k)) / mass;
    V3<R> ii = {R(c.inertia_inv[0]), R(c.inertia_inv[1]), R(c.inertia_inv[2])};
    if (lock_rotation) {
      ii = {R(0), R(0), R(0)};
      w = {R(0), R(0), R(0)};
    } else {
      V3<R> lw = rotate_inv(q, w);
      V3<R> inertia = {R(1) / ii.x, R(1) / ii.y, R(1) / ii.z};
      V3<R> gyro = rotate(q, cross(lw, V3<R>{inertia.x * lw.x, inertia.y * lw.y, inertia.z * lw.z}));
      w = w + inv_inertia(q, ii, g3(torque, k) + dt * (servo - gyro));
    }
    V3<R> normal = {R(0), R(0), R(1)};
    V3<R> total = {R(0), R(0), R(0)};
    for (int sweep = 0; sweep < 6; ++sweep)
      for (int i = 0; i < 10; ++i) {
        V3<R> r = rotate(q, vertex(i, hf, bevel));
        V3<R> corner = xt + r;
        if ((c.board_half[0] > 0 && std::fabs(corner.x) > R(c.board_half[0])) ||
            (c.board_half[1] > 0 && std::fabs(corner.y) > R(c.board_half[1])))
          continue;
        R gap = corner.z;
        R vn = dot(vt + cross(w, r), normal);
        R desired = -gap / dt;
        if (gap < R(0)) desired = desired * R(0.2);
        if (vn < desired) {
          V3<R> rn = cross(r, normal);
          R effective = R(1) / mass + dot(rn, inv_inertia(q, ii, rn));
          R jn = (desired - vn) / effective;
          V3<R> impulse = jn * normal;
          vt = vt + impulse / mass;
          w = w + inv_inertia(q, ii, cross(r, impulse));
          V3<R> vtan = vt + cross(w, r);
          vtan = vtan - dot(vtan, normal) * normal;
          R sp = length(vtan);
          if (sp > R(1e-10)) {
            V3<R> direction = -vtan / sp;
            V3<R> rt = cross(r, direction);
            R wt = R(1) / mass + dot(rt, inv_inertia(q, ii, rt));
            R jt = std::min(sp / wt, R(c.mud) * jn);
            V3<R> tangent = jt * direction;
            impulse = impulse + tangent;
            vt = vt + tangent / mass;
            w = w + inv_inertia(q, ii, cross(r, tangent));
          }
          total = total + impulse;
        }
      }
    xt = xt + dt * vt;
    R sp = length(w);
    if (sp > R(1e-10)) q = qnormalize(qmul(axis_angle(w / sp, dt * sp), q));
    s3(tool_pos, k, xt);
    s3(tool_vel, k, vt);
    tool_rot[4 * k] = q.x; tool_rot[4 * k + 1] = q.y; tool_rot[4 * k + 2] = q.z; tool_rot[4 * k + 3] = q.w;
    s3(tool_omega, k, w);
    s3(table, k, total);
  }

  // assign (contact.py:996-1039)
  void assign(int64_t digit, int cutting, int64_t k = 0) {
    if (!cutting) return;
    V3<R> hf = half();
    Q<R> q = rot(k);
    V3<R> pos = g3(tool_pos, k), tvel = g3(tool_vel, k), tom = g3(tool_omega, k);
    R dt = R(c.dt), spacing = R(c.spacing), rad = R(c.radius);
    bool relabelled = false;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1) reduction(|| : relabelled)
    for (int64_t p = 0; p < n; ++p) {
      if ((codes[p] / digit) % 3 != 0) continue;
      V3<R> xp = g3(x, p);
      V3<R> edge = pos + rotate(q, V3<R>{R(0), R(0), -hf.z});
      V3<R> local = rotate_inv(q, xp - edge);
      R front = -local.z;
      V3<R> edge_v = tvel + cross(tom, edge - pos);
      V3<R> d = rotate(q, V3<R>{R(0), R(0), R(-1)});
      R advance = dot(edge_v - g3(v, p), d);
      R back = R(2) * spacing + std::max(advance * dt, R(0));
      bool behind = front >= -back && front <= R(0);
      bool on_board = (c.board_half[0] == 0 || std::fabs(xp.x) <= R(c.board_half[0])) &&
                      (c.board_half[1] == 0 || std::fabs(xp.y) <= R(c.board_half[1]));
      V3<R> edge_here = edge + rotate(q, V3<R>{R(0), local.y, R(0)});
      bool pinched = on_board && edge_here.z <= rad && xp.z <= edge_here.z + rad;
      if (std::fabs(local.x) <= R(6) * spacing && std::fabs(local.y) <= hf.y && (behind || pinched)) {
        int64_t side = local.x < R(0) ? 1 : 2;
        codes[p] += digit * side;
        relabelled = true;
      }
    }
    if (relabelled) topology_dirty = true;
  }

  // One substep in the order of _step_kernels (solver.py:920-940).
  void substep(V3<R> target, Q<R> target_q, int cutting_flag, int64_t digit) {
    int tool_model = c.external_tool ? -1 : ((c.free_blade || c.pose_servo) ? 1 : 0);
    int cut_bits = 0;
    if (cutting_flag) cut_bits = CUT_ON | (c.cut_work_advance ? CUT_ADVANCE : 0) | (c.cut_hold ? CUT_HOLD : 0);
    if (!c.external_tool) {
      for (int64_t k = 0; k < knives; ++k) {
        s3(reaction, k, {R(0), R(0), R(0)});
        s3(torque, k, {R(0), R(0), R(0)});
      }
      s3(board, 0, {R(0), R(0), R(0)});
      speed[0] = R(0);
    }
    rigid_contacts(tool_model, cut_bits);
    particle_contacts();
    material_step();
    if (c.external_tool) {
      finish_external();
    } else {
      if (c.accumulate_tool_impulses) accumulate_reaction();
      for (int64_t k = 0; k < knives; ++k) tool_step(target, target_q, k);
    }
    // Knife k labels digit 3^(cut index + k) (solver.py:866-890).
    for (int64_t k = 0, d = digit; k < knives; ++k, d *= 3) assign(d, cutting_flag, k);
    step_count += 1;
  }
};


// ---------------------------------------------------------------- MLS-MPM baseline
struct MlsConfig {
  int64_t dims[3];
  int32_t plastic, pad;
  double origin[3];
  double dx, dt, mass, volume, mu, lam, damaged_shear, softening, compressed, stretched;
};

// MlsSolver._substep (mlsmpm.py:293-393) on the particles of `contact` (x, v, tool state):
// contact impulses with tool_model -1 and no cutting, interval sums, dense-grid P2G,
// grid update and G2P.
template <class R> struct MlsSim {
  MlsConfig m;
  Sim<R> contact;
  std::vector<R> C, F, damage, damage_ref, yield, grid_m, grid_p, grid_v, impulse, angular;
  int64_t inversions = 0;

  MlsSim(const OracleConfig& oc, const MlsConfig& mc) : m(mc), contact(oc) {
    int64_t n = oc.n, nodes = mc.dims[0] * mc.dims[1] * mc.dims[2];
    C.assign(9 * n, R(0));
    F.assign(9 * n, R(0));
    for (int64_t p = 0; p < n; ++p) for (int i = 0; i < 3; ++i) F[9 * p + 4 * i] = R(1);
    damage.assign(n, R(0));
    damage_ref.assign(n, R(1));
    yield.assign(n, R(oc.yield_stress));
    grid_m.assign(nodes, R(0));
    grid_p.assign(3 * nodes, R(0));
    grid_v.assign(3 * nodes, R(0));
    impulse.assign(3, R(0));
    angular.assign(3, R(0));
  }

  bool base_of(const R* x, int64_t b[3], R fx[3]) const {
    for (int i = 0; i < 3; ++i) {
      if (!std::isfinite(double(x[i]))) return false;
      R local = (x[i] - R(m.origin[i])) / R(m.dx);
      R fb = std::floor(local - R(0.5));
      if (!(fb >= R(0) && fb < R(m.dims[i] - 2))) return false;
      b[i] = int64_t(fb);
      fx[i] = local - fb;
    }
    return true;
  }
  static void spline(R f, R w[3]) {
    w[0] = R(0.5) * (R(1.5) - f) * (R(1.5) - f);
    w[1] = R(0.75) - (f - R(1)) * (f - R(1));
    w[2] = R(0.5) * (f - R(0.5)) * (f - R(0.5));
  }
  int64_t node(int64_t i, int64_t j, int64_t k) const { return (i * m.dims[1] + j) * m.dims[2] + k; }

  void p2g() {
    int64_t n = contact.n, nodes = m.dims[0] * m.dims[1] * m.dims[2];
    std::fill(grid_m.begin(), grid_m.end(), R(0));
    std::fill(grid_p.begin(), grid_p.end(), R(0));
    R dx = R(m.dx), mass = R(m.mass);
    for (int64_t p = 0; p < n; ++p) {
      int64_t b[3];
      R fx[3];
      if (!base_of(&contact.x[3 * p], b, fx)) continue;
      R w[3][3];
      for (int a = 0; a < 3; ++a) spline(fx[a], w[a]);
      R mu = R(m.mu);
      if (damage[p] > R(0)) mu = R(m.mu) * R(m.damaged_shear);
      M3<R> Fm;
      for (int i = 0; i < 9; ++i) Fm.m[i / 3][i % 3] = F[9 * p + i];
      M3<R> U, V;
      V3<R> sg;
      svd3(Fm, U, sg, V);
      R J = sg.x * sg.y * sg.z;
      M3<R> D = zero3<R>();
      D.m[0][0] = R(2) * mu * (sg.x - R(1)) + R(m.lam) * (J - R(1)) * J / sg.x;
      D.m[1][1] = R(2) * mu * (sg.y - R(1)) + R(m.lam) * (J - R(1)) * J / sg.y;
      D.m[2][2] = R(2) * mu * (sg.z - R(1)) + R(m.lam) * (J - R(1)) * J / sg.z;
      M3<R> P = mul(mul(U, D), transpose(V));
      R scale_ = -R(m.dt) * R(m.volume) * R(4) / (dx * dx);
      M3<R> stress = mul(scale(scale_, P), transpose(Fm));
      M3<R> Cm;
      for (int i = 0; i < 9; ++i) Cm.m[i / 3][i % 3] = C[9 * p + i];
      M3<R> affine = addm(stress, scale(mass, Cm));
      V3<R> mom = mass * contact.g3(contact.v, p) + contact.g3(contact.impulses, p);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            R weight = w[0][i] * w[1][j] * w[2][k];
            V3<R> dpos = {(R(i) - fx[0]) * dx, (R(j) - fx[1]) * dx, (R(k) - fx[2]) * dx};
            V3<R> c = weight * (mom + mulv(affine, dpos));
            int64_t nd = node(b[0] + i, b[1] + j, b[2] + k);
            grid_p[nd] += c.x;
            grid_p[nodes + nd] += c.y;
            grid_p[2 * nodes + nd] += c.z;
            grid_m[nd] += weight * mass;
          }
    }
  }

  void grid_update() {
    int64_t nodes = m.dims[0] * m.dims[1] * m.dims[2];
    for (int64_t i = 0; i < m.dims[0]; ++i)
      for (int64_t j = 0; j < m.dims[1]; ++j)
        for (int64_t k = 0; k < m.dims[2]; ++k) {
          int64_t nd = node(i, j, k);
          R mm = grid_m[nd];
          R v[3] = {R(0), R(0), R(0)};
          if (mm > R(0))
            for (int c = 0; c < 3; ++c) v[c] = grid_p[c * nodes + nd] / mm;
          if (i < 1 || j < 1 || k < 1 || i >= m.dims[0] - 1 || j >= m.dims[1] - 1 || k >= m.dims[2] - 1)
            v[0] = v[1] = v[2] = R(0);
          for (int c = 0; c < 3; ++c) grid_v[c * nodes + nd] = v[c];
        }
  }

  void g2p() {
    int64_t n = contact.n, nodes = m.dims[0] * m.dims[1] * m.dims[2];
    R dx = R(m.dx), dt = R(m.dt);
    const R minsv = R(0.05);
    for (int64_t p = 0; p < n; ++p) {
      int64_t b[3];
      R fx[3];
      if (!base_of(&contact.x[3 * p], b, fx)) continue;
      R w[3][3];
      for (int a = 0; a < 3; ++a) spline(fx[a], w[a]);
      V3<R> nv = {R(0), R(0), R(0)};
      M3<R> nC = zero3<R>();
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            R weight = w[0][i] * w[1][j] * w[2][k];
            V3<R> dpos = {(R(i) - fx[0]) * dx, (R(j) - fx[1]) * dx, (R(k) - fx[2]) * dx};
            int64_t nd = node(b[0] + i, b[1] + j, b[2] + k);
            V3<R> gv = {grid_v[nd], grid_v[nodes + nd], grid_v[2 * nodes + nd]};
            nv = nv + weight * gv;
            nC = addm(nC, scale(R(4) * weight / (dx * dx), outer(gv, dpos)));
          }
      contact.s3(contact.v, p, nv);
      for (int i = 0; i < 9; ++i) C[9 * p + i] = nC.m[i / 3][i % 3];
      contact.s3(contact.x, p, contact.g3(contact.x, p) + dt * nv);
      M3<R> Fm;
      for (int i = 0; i < 9; ++i) Fm.m[i / 3][i % 3] = F[9 * p + i];
      M3<R> f = mul(addm(eye3<R>(), scale(dt, nC)), Fm);
      M3<R> U, V;
      V3<R> s;
      svd3(f, U, s, V);
      R smin = std::min(std::min(s.x, s.y), s.z);
      if (smin <= R(0)) inversions += 1;
      auto rebuild = [&](V3<R> sv) {
        M3<R> D = zero3<R>();
        D.m[0][0] = sv.x; D.m[1][1] = sv.y; D.m[2][2] = sv.z;
        return mul(mul(U, D), transpose(V));
      };
      if (smin < minsv) {
        s = {std::max(s.x, minsv), std::max(s.y, minsv), std::max(s.z, minsv)};
        f = rebuild(s);
      }
      R ref_J = damage_ref[p];
      if (m.plastic) {
        V3<R> e = {std::log(s.x), std::log(s.y), std::log(s.z)};
        R mean = (e.x + e.y + e.z) / R(3);
        V3<R> dev = e - V3<R>{mean, mean, mean};
        R norm = length(dev);
        R limit = yield[p] / (R(2) * R(m.mu)) * std::sqrt(R(2) / R(3));
        if (norm > limit) {
          e = e - (R(1) - limit / norm) * dev;
          s = {std::exp(e.x), std::exp(e.y), std::exp(e.z)};
          f = rebuild(s);
          if (m.softening > 0 && ref_J > R(0)) {
            yield[p] = std::max(R(0), yield[p] - R(m.softening) * (norm - limit));
            if (yield[p] <= R(0)) damage[p] = R(1);
          }
        }
      }
      for (int i = 0; i < 9; ++i) F[9 * p + i] = f.m[i / 3][i % 3];
      if (ref_J > R(0)) {
        R J = s.x * s.y * s.z / ref_J;
        if (J <= R(m.compressed) || J >= R(m.stretched)) damage[p] = R(1);
      }
    }
  }

  void substep() {
    contact.rigid_contacts(-1, 0);
    for (int i = 0; i < 3; ++i) {
      impulse[i] += contact.reaction[i];
      angular[i] += contact.torque[i];
      contact.reaction[i] = contact.torque[i] = contact.board[i] = R(0);
    }
    contact.speed[0] = R(0);
    p2g();
    grid_update();
    g2p();
  }
};

struct MlsHandle {
  int precision;
  void* sim;
};

// ---------------------------------------------------------------- named array access
struct Handle {
  int precision;
  void* sim;
};

template <class R> bool real_array(Sim<R>& s, const std::string& name, std::vector<R>*& out) {
  std::unordered_map<std::string, std::vector<R>*> table = {
      {"x", &s.x}, {"v", &s.v}, {"F", &s.F}, {"Fp", &s.Fp}, {"stress", &s.stress}, {"damage", &s.damage},
      {"ref", &s.ref}, {"weights", &s.weights}, {"gradients", &s.gradients}, {"impulses", &s.impulses},
      {"histories", &s.histories}, {"tool_pos", &s.tool_pos}, {"tool_vel", &s.tool_vel}, {"tool_rot", &s.tool_rot},
      {"tool_omega", &s.tool_omega}, {"tool_mobility", &s.mobility}, {"reaction", &s.reaction},
      {"torque", &s.torque}, {"board", &s.board}, {"work", &s.work}, {"speed", &s.speed},
      {"tool_table_impulse", &s.table}, {"interval_impulse", &s.interval_impulse},
      {"interval_angular_impulse", &s.interval_angular}, {"grid_mass", &s.grid_mass}, {"momentum", &s.momentum},
      {"grid_impulse", &s.grid_impulse}, {"old_v", &s.old_v}, {"new_v", &s.new_v}, {"remapped", &s.remapped},
      {"coupled_remapped", &s.coupled_remapped}, {"bc_side", &s.bc_side}, {"bc_impulse", &s.bc_impulse},
      {"bc_normal", &s.bc_normal}, {"extra_histories", &s.extra_histories}, {"extra_damage", &s.extra_damage}};
  auto it = table.find(name);
  if (it == table.end()) return false;
  out = it->second;
  return true;
}

template <class R> int get_set(Sim<R>& s, const char* cname, void* data, int64_t count, bool set) {
  std::string name(cname);
  std::vector<R>* arr = nullptr;
  if (name == "viscous") {
    if (count != int64_t(s.viscous.size())) return -2;
    double* d = static_cast<double*>(data);
    for (int64_t i = 0; i < count; ++i) {
      if (set) s.viscous[i] = d[i];
      else d[i] = s.viscous[i];
    }
    return 0;
  }
  if (real_array(s, name, arr)) {
    if (count != int64_t(arr->size())) return -2;
    double* d = static_cast<double*>(data);
    for (int64_t i = 0; i < count; ++i) {
      if (set) (*arr)[i] = R(d[i]);
      else d[i] = double((*arr)[i]);
    }
    return 0;
  }
  int64_t* d = static_cast<int64_t*>(data);
  auto ints = [&](auto& vec) -> int {
    if (count != int64_t(vec.size())) return -2;
    for (int64_t i = 0; i < count; ++i) {
      if (set) vec[i] = static_cast<typename std::decay_t<decltype(vec)>::value_type>(d[i]);
      else d[i] = int64_t(vec[i]);
    }
    return 0;
  };
  if (name == "codes") { int r = ints(s.codes); if (set && r == 0) s.topology_dirty = true; return r; }
  if (name == "nodes") return ints(s.nodes);
  if (name == "errors") return ints(s.errors);
  if (name == "slots") return ints(s.slots);
  if (name == "keys") return ints(s.keys);
  if (name == "grid_keys") return ints(s.grid_keys);
  if (name == "heads") return ints(s.heads);
  if (name == "links") return ints(s.links);
  if (name == "bc_max_violation") {
    if (count != 1) return -2;
    double* dd = static_cast<double*>(data);
    if (set) s.bc_max_violation = R(dd[0]); else dd[0] = double(s.bc_max_violation);
    return 0;
  }
  if (name == "scalars") {
    // [active_count, pair_count, step_count, topology_dirty, single_field]
    if (count != 5) return -2;
    if (set) {
      s.topology_dirty = d[3] != 0;
      s.single_field = d[4] != 0;
      s.field_codes_override.assign(s.n, 0);
      return 0;
    }
    d[0] = s.active_count; d[1] = s.pair_count; d[2] = s.step_count; d[3] = s.topology_dirty; d[4] = s.single_field;
    return 0;
  }
  return -1;
}

}  // namespace

extern "C" {

// Create an oracle simulation; precision 32 (float mirror) or 64 (double).
void* oracle_create(int precision, const OracleConfig* cfg) {
  Handle* h = new Handle{precision, nullptr};
  if (precision == 32) h->sim = new Sim<float>(*cfg);
  else h->sim = new Sim<double>(*cfg);
  return h;
}

void oracle_destroy(void* handle) {
  Handle* h = static_cast<Handle*>(handle);
  if (!h) return;
  if (h->precision == 32) delete static_cast<Sim<float>*>(h->sim);
  else delete static_cast<Sim<double>*>(h->sim);
  delete h;
}

// Copy a named array in (set != 0) or out.  Real arrays travel as float64, integer arrays
// as int64.  Returns 0, -1 for an unknown name, -2 for a size mismatch.
int oracle_array(void* handle, const char* name, void* data, int64_t count, int set) {
  Handle* h = static_cast<Handle*>(handle);
  if (h->precision == 32) return get_set(*static_cast<Sim<float>*>(h->sim), name, data, count, set != 0);
  return get_set(*static_cast<Sim<double>*>(h->sim), name, data, count, set != 0);
}

// One substep; target[3] and target_q[4] (xyzw) drive a servo/guided tool.
int oracle_substep(void* handle, const double* target, const double* target_q, int cutting, int64_t digit) {
  Handle* h = static_cast<Handle*>(handle);
  if (h->precision == 32) {
    auto* s = static_cast<Sim<float>*>(h->sim);
    s->substep({float(target[0]), float(target[1]), float(target[2])},
               {float(target_q[0]), float(target_q[1]), float(target_q[2]), float(target_q[3])}, cutting, digit);
  } else {
    auto* s = static_cast<Sim<double>*>(h->sim);
    s->substep({target[0], target[1], target[2]}, {target_q[0], target_q[1], target_q[2], target_q[3]}, cutting, digit);
  }
  return 0;
}

// Threads of the per-particle and per-field loops (1: sequential).  Returns the count in
// effect, 1 when the library was built without OpenMP.
int oracle_set_threads(void* handle, int threads) {
  Handle* h = static_cast<Handle*>(handle);
#ifdef _OPENMP
  int t = threads < 1 ? 1 : threads;
#else
  int t = 1;
#endif
  if (h->precision == 32) static_cast<Sim<float>*>(h->sim)->threads = t;
  else static_cast<Sim<double>*>(h->sim)->threads = t;
  return t;
}

// Stage-level entry points used by the parity tests (each one stage of the substep).
int oracle_stage(void* handle, const char* stage, int arg0, int64_t arg1) {
  Handle* h = static_cast<Handle*>(handle);
  std::string st(stage);
  auto run = [&](auto* s) -> int {
    int tool_model = s->c.external_tool ? -1 : ((s->c.free_blade || s->c.pose_servo) ? 1 : 0);
    if (st == "rigid") { s->rigid_contacts(tool_model, arg0); return 0; }
    if (st == "pairs") { s->particle_contacts(); return 0; }
    if (st == "material") { s->material_step(); return 0; }
    if (st == "register") { s->register_topology(); return 0; }
    if (st == "p2g") {
      if (!s->c.cache_topology || s->topology_dirty) s->register_topology();
      s->p2g();
      return 0;
    }
    if (st == "assign") { s->assign(arg1, arg0); return 0; }
    if (st == "finish_external") { s->finish_external(); return 0; }
    if (st == "refresh_stress") { s->refresh_stress(); return 0; }
    // The F update alone on the registered slots (grid_mass and coupled_remapped as set).
    if (st == "update_F") { s->update_F(); return 0; }
    // viscous_impulse alone: adds to grid_impulse from old_v on the registered slots.
    if (st == "viscous") { s->viscous_impulse(); return 0; }
    if (st == "plastic_project") {
      for (int64_t p = 0; p < s->n; ++p) {
        using RR = std::decay_t<decltype(s->x[0])>;
        M3<RR> U, V;
        V3<RR> sg;
        s->plastic_at(p, U, sg, V);
      }
      return 0;
    }
    return -1;
  };
  if (h->precision == 32) return run(static_cast<Sim<float>*>(h->sim));
  return run(static_cast<Sim<double>*>(h->sim));
}

// Pure-function probes (float64 in/out, evaluated at the requested precision).
int oracle_blade_query(int precision, const double* local, const double* half, double bevel, double* out7) {
  auto go = [&](auto zero) {
    using RR = decltype(zero);
    V3<RR> c, nrm;
    RR d;
    Sim<RR>::blade_query({RR(local[0]), RR(local[1]), RR(local[2])}, {RR(half[0]), RR(half[1]), RR(half[2])},
                         RR(bevel), c, nrm, d);
    out7[0] = c.x; out7[1] = c.y; out7[2] = c.z; out7[3] = nrm.x; out7[4] = nrm.y; out7[5] = nrm.z; out7[6] = d;
  };
  if (precision == 32) go(0.0f); else go(0.0);
  return 0;
}

int oracle_svd3(const double* A, double* U, double* s, double* V) {
  double a[3][3], u[3][3], v[3][3];
  for (int i = 0; i < 9; ++i) a[i / 3][i % 3] = A[i];
  svd3_double(a, u, s, v);
  for (int i = 0; i < 9; ++i) {
    U[i] = u[i / 3][i % 3];
    V[i] = v[i / 3][i % 3];
  }
  return 0;
}

int oracle_compatible(int64_t a, int64_t b) { return compatible(a, b) ? 1 : 0; }

// friction probe: in[13] = normal 3, relative 3, jn, history 3, me, pad 2; p[5] = dt, mus,
// mud, gk, gc; out[6] = impulse 3, history 3.
int oracle_friction(int precision, const double* in, const double* prm, double* out) {
  auto go = [&](auto zero) {
    using RR = decltype(zero);
    OracleConfig cfg{};
    cfg.dt = prm[0]; cfg.mus = prm[1]; cfg.mud = prm[2]; cfg.gk = prm[3]; cfg.gc = prm[4];
    cfg.n = 0; cfg.node_count = 1; cfg.key_capacity = 16; cfg.pair_capacity = 16;
    Sim<RR> s(cfg);
    V3<RR> jt, ht;
    RR unused = 0;
    s.friction({RR(in[0]), RR(in[1]), RR(in[2])}, {RR(in[3]), RR(in[4]), RR(in[5])}, RR(in[6]),
               {RR(in[7]), RR(in[8]), RR(in[9])}, RR(in[10]), unused, jt, ht);
    out[0] = jt.x; out[1] = jt.y; out[2] = jt.z; out[3] = ht.x; out[4] = ht.y; out[5] = ht.z;
  };
  if (precision == 32) go(0.0f); else go(0.0);
  return 0;
}

// resist_cut probe on particle p of a simulation: j[3], origin[3], tool velocity[3],
// q[4], omega[3]; rad is the contact offset; out[3] the returned impulse.
int oracle_resist_cut(void* handle, int64_t p, const double* j, const double* origin, const double* tvel,
                      const double* q, const double* omega, double rad, int cutting, double* out) {
  Handle* h = static_cast<Handle*>(handle);
  auto go = [&](auto* s) {
    using RR = std::decay_t<decltype(s->x[0])>;
    V3<RR> r = s->resist_cut(p, {RR(j[0]), RR(j[1]), RR(j[2])}, {RR(origin[0]), RR(origin[1]), RR(origin[2])},
                             {RR(tvel[0]), RR(tvel[1]), RR(tvel[2])}, {RR(q[0]), RR(q[1]), RR(q[2]), RR(q[3])},
                             {RR(omega[0]), RR(omega[1]), RR(omega[2])}, RR(rad), cutting);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
  };
  if (h->precision == 32) go(static_cast<Sim<float>*>(h->sim)); else go(static_cast<Sim<double>*>(h->sim));
  return 0;
}

// Eq. 8 probe in double: keys/heads/links/mass/state of `count` fields at node_count nodes.
int oracle_couple(const int64_t* keys, const int32_t* heads, const int32_t* links, const double* mass,
                  const double* state, double* result, int64_t count, int64_t node_count) {
  couple_fields<double>(keys, heads, links, mass, state, result, count, node_count);
  return 0;
}

// g2p probe in double (material.py:343-381).
int oracle_g2p(double* x, double* v, const int32_t* slots, const double* weights, const double* old_v,
               const double* new_v, double dt, double flip, int64_t n) {
  g2p_fields<double>(x, v, slots, weights, old_v, new_v, dt, flip, n);
  return 0;
}

// MLS oracle: created from the contact configuration (tool_model -1 semantics are taken from
// external_tool = 1) and the grid description.
void* oracle_mls_create(int precision, const OracleConfig* cfg, const MlsConfig* mc) {
  MlsHandle* h = new MlsHandle{precision, nullptr};
  if (precision == 32) h->sim = new MlsSim<float>(*cfg, *mc);
  else h->sim = new MlsSim<double>(*cfg, *mc);
  return h;
}

void oracle_mls_destroy(void* handle) {
  MlsHandle* h = static_cast<MlsHandle*>(handle);
  if (!h) return;
  if (h->precision == 32) delete static_cast<MlsSim<float>*>(h->sim);
  else delete static_cast<MlsSim<double>*>(h->sim);
  delete h;
}

int oracle_mls_substep(void* handle, int steps) {
  MlsHandle* h = static_cast<MlsHandle*>(handle);
  for (int i = 0; i < steps; ++i) {
    if (h->precision == 32) static_cast<MlsSim<float>*>(h->sim)->substep();
    else static_cast<MlsSim<double>*>(h->sim)->substep();
  }
  return 0;
}

// Named arrays of the MLS oracle (float64 transfer): x, v, C, F, damage, damage_reference,
// yield_stress, impulse, angular_impulse, tool_pos, tool_vel, tool_rot, tool_omega, inversions.
int oracle_mls_array(void* handle, const char* cname, double* data, int64_t count, int set) {
  MlsHandle* h = static_cast<MlsHandle*>(handle);
  std::string name(cname);
  auto go = [&](auto* s) -> int {
    using RR = std::decay_t<decltype(s->C[0])>;
    std::vector<RR>* a = nullptr;
    if (name == "x") a = &s->contact.x;
    else if (name == "v") a = &s->contact.v;
    else if (name == "C") a = &s->C;
    else if (name == "F") a = &s->F;
    else if (name == "damage") a = &s->damage;
    else if (name == "damage_reference") a = &s->damage_ref;
    else if (name == "yield_stress") a = &s->yield;
    else if (name == "impulse") a = &s->impulse;
    else if (name == "angular_impulse") a = &s->angular;
    else if (name == "tool_pos") a = &s->contact.tool_pos;
    else if (name == "tool_vel") a = &s->contact.tool_vel;
    else if (name == "tool_rot") a = &s->contact.tool_rot;
    else if (name == "tool_omega") a = &s->contact.tool_omega;
    else if (name == "histories") a = &s->contact.histories;
    else if (name == "inversions") {
      if (count != 1) return -2;
      if (set) s->inversions = int64_t(data[0]); else data[0] = double(s->inversions);
      return 0;
    }
    if (!a) return -1;
    if (count != int64_t(a->size())) return -2;
    for (int64_t i = 0; i < count; ++i) {
      if (set) (*a)[i] = RR(data[i]);
      else data[i] = double((*a)[i]);
    }
    return 0;
  };
  if (h->precision == 32) return go(static_cast<MlsSim<float>*>(h->sim));
  return go(static_cast<MlsSim<double>*>(h->sim));
}

}  // extern "C"
