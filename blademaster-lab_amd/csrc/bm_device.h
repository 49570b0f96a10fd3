// Device-side math for the BladeMaster TLMPM kernels on gfx950 (CDNA4, wave64).
//
// Small fixed-size vector/matrix/quaternion types kept in registers, the float32 3x3 SVD
// used by the corotated stress, ternary cut-code helpers and the trilinear stencil that the
// kernels rebuild from its compact form (base node + local fractions) instead of loading
// 8 weights and 24 gradient terms per particle from HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm {

struct f3 {
  float x, y, z;
};

__device__ __forceinline__ f3 make3(float a, float b, float c) { return f3{a, b, c}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

__device__ __forceinline__ f3 ld3(const float* p, int64_t i) { return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
// A load whose address is the same for the whole wavefront, through the scalar cache: no
// vector memory wait holds up the wavefront's other loads.  Only for data that no wavefront
// of the running kernel writes (the scalar cache is not kept coherent with vector stores).
template <typename T>
__device__ __forceinline__ T uniform_load(const T* p) {
  return *(const __attribute__((address_space(4))) T*)(p);
}
__device__ __forceinline__ f3 uniform_ld3(const float* p, int64_t i) {
  return f3{uniform_load(p + 3 * i), uniform_load(p + 3 * i + 1), uniform_load(p + 3 * i + 2)};
}
__device__ __forceinline__ void st3(float* p, int64_t i, f3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ __forceinline__ void atomic_add3(float* p, int64_t i, f3 v) {
  atomicAdd(p + 3 * i, v.x);
  atomicAdd(p + 3 * i + 1, v.y);
  atomicAdd(p + 3 * i + 2, v.z);
}

// Row-major 3x3.
struct m3 {
  float a[9];
};
__device__ __forceinline__ m3 ld9(const float* p, int64_t i) {
  m3 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.a[k] = p[9 * i + k];
  return r;
}
__device__ __forceinline__ void st9(float* p, int64_t i, const m3& m) {
#pragma unroll
  for (int k = 0; k < 9; ++k) p[9 * i + k] = m.a[k];
}
__device__ __forceinline__ m3 mzero() {
  m3 r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.a[k] = 0.0f;
  return r;
}
__device__ __forceinline__ m3 mmul(const m3& x, const m3& y) {
  m3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = x.a[3 * i] * y.a[j] + x.a[3 * i + 1] * y.a[3 + j] + x.a[3 * i + 2] * y.a[6 + j];
  return r;
}
// x * y^T
__device__ __forceinline__ m3 mmul_t(const m3& x, const m3& y) {
  m3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      r.a[3 * i + j] = x.a[3 * i] * y.a[3 * j] + x.a[3 * i + 1] * y.a[3 * j + 1] + x.a[3 * i + 2] * y.a[3 * j + 2];
  return r;
}
__device__ __forceinline__ f3 mvec(const m3& m, f3 v) {
  return f3{m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z, m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
            m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z};
}
__device__ __forceinline__ float mdet(const m3& m) {
  return m.a[0] * (m.a[4] * m.a[8] - m.a[5] * m.a[7]) - m.a[1] * (m.a[3] * m.a[8] - m.a[5] * m.a[6]) +
         m.a[2] * (m.a[3] * m.a[7] - m.a[4] * m.a[6]);
}
__device__ __forceinline__ m3 minv(const m3& m) {
  float inv = 1.0f / mdet(m);
  m3 r;
  r.a[0] = (m.a[4] * m.a[8] - m.a[5] * m.a[7]) * inv;
  r.a[1] = (m.a[2] * m.a[7] - m.a[1] * m.a[8]) * inv;
  r.a[2] = (m.a[1] * m.a[5] - m.a[2] * m.a[4]) * inv;
  r.a[3] = (m.a[5] * m.a[6] - m.a[3] * m.a[8]) * inv;
  r.a[4] = (m.a[0] * m.a[8] - m.a[2] * m.a[6]) * inv;
  r.a[5] = (m.a[2] * m.a[3] - m.a[0] * m.a[5]) * inv;
  r.a[6] = (m.a[3] * m.a[7] - m.a[4] * m.a[6]) * inv;
  r.a[7] = (m.a[1] * m.a[6] - m.a[0] * m.a[7]) * inv;
  r.a[8] = (m.a[0] * m.a[4] - m.a[1] * m.a[3]) * inv;
  return r;
}
__device__ __forceinline__ m3 mtranspose(const m3& m) {
  m3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = m.a[3 * j + i];
  return r;
}

// Quaternion (x, y, z, w); rotation v' = (2w^2 - 1) v + 2 (u.v) u + 2 w (u x v).
struct quat {
  float x, y, z, w;
};
__device__ __forceinline__ quat ldq(const float* p) { return quat{p[0], p[1], p[2], p[3]}; }
__device__ __forceinline__ f3 qvec(quat q) { return f3{q.x, q.y, q.z}; }
__device__ __forceinline__ f3 rotate(quat q, f3 v) {
  f3 u = qvec(q);
  return (2.0f * q.w * q.w - 1.0f) * v + (2.0f * dot(u, v)) * u + (2.0f * q.w) * cross(u, v);
}
__device__ __forceinline__ f3 rotate_inv(quat q, f3 v) {
  f3 u = qvec(q);
  return (2.0f * q.w * q.w - 1.0f) * v + (2.0f * dot(u, v)) * u - (2.0f * q.w) * cross(u, v);
}
__device__ __forceinline__ quat qmul(quat a, quat b) {
  f3 av = qvec(a), bv = qvec(b);
  f3 v = a.w * bv + b.w * av + cross(av, bv);
  return quat{v.x, v.y, v.z, a.w * b.w - dot(av, bv)};
}
__device__ __forceinline__ quat qnormalize(quat q) {
  float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  return quat{q.x / l, q.y / l, q.z / l, q.w / l};
}
__device__ __forceinline__ quat axis_angle(f3 axis, float angle) {
  float s = sinf(0.5f * angle), c = cosf(0.5f * angle);
  return quat{axis.x * s, axis.y * s, axis.z * s, c};
}

// ---------------------------------------------------------------------------------------
// Float32 3x3 SVD, A = U diag(s) V^T, U and V proper rotations, |s0| >= |s1| >= |s2| and
// only s2 may be negative.  Structure of McAdams et al. (2011): symmetric eigenanalysis of
// A^T A by cyclic Jacobi (here with exact rotation angles instead of the paper's
// approximate quaternion rotations, and sweeps until the off-diagonal is below float
// resolution, at most 6), column sort, then QR of A V by three Givens rotations.
// Reciprocals and square roots use the hardware's single-instruction forms (1 ulp).
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ inline void svd3(const m3& A, m3& U, f3& s, m3& V) {
  // S = A^T A (symmetric)
  float s00 = A.a[0] * A.a[0] + A.a[3] * A.a[3] + A.a[6] * A.a[6];
  float s01 = A.a[0] * A.a[1] + A.a[3] * A.a[4] + A.a[6] * A.a[7];
  float s02 = A.a[0] * A.a[2] + A.a[3] * A.a[5] + A.a[6] * A.a[8];
  float s11 = A.a[1] * A.a[1] + A.a[4] * A.a[4] + A.a[7] * A.a[7];
  float s12 = A.a[1] * A.a[2] + A.a[4] * A.a[5] + A.a[7] * A.a[8];
  float s22 = A.a[2] * A.a[2] + A.a[5] * A.a[5] + A.a[8] * A.a[8];
  float W[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
  auto jacobi = [&](float& app, float& aqq, float& apq, float& apr, float& aqr, int p, int q) {
    // A rotation smaller than this changes nothing at float resolution (or spins inside a
    // degenerate eigenspace, which leaves U diag(s) V^T unchanged).
    if (!(fabsf(apq) > 1.0e-30f)) return;
    float tau = (aqq - app) * fast_rcp(2.0f * apq);
    float t = copysignf(fast_rcp(fabsf(tau) + fast_sqrt(1.0f + tau * tau)), tau);
    if (!(fabsf(tau) < 1.0e18f)) t = 0.5f * copysignf(fast_rcp(fabsf(tau)), tau);  // tau^2 overflow
    float c = fast_rsqrt(1.0f + t * t), sn = t * c;
    app = app - t * apq;
    aqq = aqq + t * apq;
    apq = 0.0f;
    float r1 = apr, r2 = aqr;
    apr = c * r1 - sn * r2;
    aqr = sn * r1 + c * r2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      float a = W[3 * k + p], b = W[3 * k + q];
      W[3 * k + p] = c * a - sn * b;
      W[3 * k + q] = sn * a + c * b;
    }
  };
#pragma unroll 1
  for (int sweep = 0; sweep < 6; ++sweep) {
    // Converged when the off-diagonal is at float rounding level of the diagonal (relative
    // 3e-7): further sweeps only exchange rounding errors.
    float off = s01 * s01 + s02 * s02 + s12 * s12;
    float diag = s00 * s00 + s11 * s11 + s22 * s22;
    if (off <= 1.0e-13f * diag) break;
    jacobi(s00, s11, s01, s02, s12, 0, 1);
    jacobi(s00, s22, s02, s01, s12, 0, 2);
    jacobi(s11, s22, s12, s01, s02, 1, 2);
  }
  // B = A W, columns sorted by norm descending.
  float B[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      B[3 * i + j] = A.a[3 * i] * W[j] + A.a[3 * i + 1] * W[3 + j] + A.a[3 * i + 2] * W[6 + j];
  float n0 = B[0] * B[0] + B[3] * B[3] + B[6] * B[6];
  float n1 = B[1] * B[1] + B[4] * B[4] + B[7] * B[7];
  float n2 = B[2] * B[2] + B[5] * B[5] + B[8] * B[8];
  auto swapcols = [&](int a, int b, float& na, float& nb) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      float t = B[3 * i + a];
      B[3 * i + a] = B[3 * i + b];
      B[3 * i + b] = -t;
      t = W[3 * i + a];
      W[3 * i + a] = W[3 * i + b];
      W[3 * i + b] = -t;  // negation keeps det W = +1
    }
    float t = na;
    na = nb;
    nb = t;
  };
  if (n0 < n1) swapcols(0, 1, n0, n1);
  if (n0 < n2) swapcols(0, 2, n0, n2);
  if (n1 < n2) swapcols(1, 2, n1, n2);
  // Givens QR: zero B[1][0], B[2][0], B[2][1]; Q accumulates the rotations.
  float Q[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
  auto givens = [&](int i, int k, int col) {
    float a = B[3 * i + col], b = B[3 * k + col];
    float r2 = a * a + b * b;
    float c = 1.0f, sn = 0.0f;
    if (r2 > 1.0e-36f) {
      float inv = fast_rsqrt(r2);
      c = a * inv;
      sn = b * inv;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float x = B[3 * i + j], y = B[3 * k + j];
      B[3 * i + j] = c * x + sn * y;
      B[3 * k + j] = -sn * x + c * y;
      x = Q[3 * j + i];
      y = Q[3 * j + k];
      Q[3 * j + i] = c * x + sn * y;
      Q[3 * j + k] = -sn * x + c * y;
    }
  };
  givens(0, 1, 0);
  givens(0, 2, 0);
  givens(1, 2, 1);
  // Non-negative first two singular values: flip U column j together with U column 2.
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (B[4 * j] < 0.0f) {
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        Q[3 * i + j] = -Q[3 * i + j];
        Q[3 * i + 2] = -Q[3 * i + 2];
      }
      B[4 * j] = -B[4 * j];
      B[8] = -B[8];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    U.a[k] = Q[k];
    V.a[k] = W[k];
  }
  s = f3{B[0], B[4], B[8]};
}

// First Piola stress of fixed corotated elasticity from the SVD of f = F Fp^-1,
// P = U diag(2 mu (s - 1) + lam (J - 1) J / s) V^T Fp^-T  (material.py:78-91).
__device__ __forceinline__ m3 corotated(const m3& U, f3 s, const m3& V, const m3& pinv, float mu, float lam) {
  // J / s_i is the product of the other two singular values.
  float J = s.x * s.y * s.z;
  float lj = lam * (J - 1.0f);
  float d0 = 2.0f * mu * (s.x - 1.0f) + lj * (s.y * s.z);
  float d1 = 2.0f * mu * (s.y - 1.0f) + lj * (s.x * s.z);
  float d2 = 2.0f * mu * (s.z - 1.0f) + lj * (s.x * s.y);
  m3 UD;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    UD.a[3 * i] = U.a[3 * i] * d0;
    UD.a[3 * i + 1] = U.a[3 * i + 1] * d1;
    UD.a[3 * i + 2] = U.a[3 * i + 2] * d2;
  }
  m3 P = mmul_t(UD, V);     // U D V^T
  return mmul_t(P, pinv);   // (U D V^T) Fp^-T
}
// Cofactor matrix of m (= det(m) m^-T).
__device__ __forceinline__ m3 cofactor(const m3& m) {
  m3 c;
  c.a[0] = m.a[4] * m.a[8] - m.a[5] * m.a[7];
  c.a[1] = m.a[5] * m.a[6] - m.a[3] * m.a[8];
  c.a[2] = m.a[3] * m.a[7] - m.a[4] * m.a[6];
  c.a[3] = m.a[2] * m.a[7] - m.a[1] * m.a[8];
  c.a[4] = m.a[0] * m.a[8] - m.a[2] * m.a[6];
  c.a[5] = m.a[1] * m.a[6] - m.a[0] * m.a[7];
  c.a[6] = m.a[1] * m.a[5] - m.a[2] * m.a[4];
  c.a[7] = m.a[2] * m.a[3] - m.a[0] * m.a[5];
  c.a[8] = m.a[0] * m.a[4] - m.a[1] * m.a[3];
  return c;
}

// Fixed-corotated first Piola stress of f (det f = J > 0) without an SVD:
// U diag(2 mu (s - 1) + lam (J - 1) J / s) V^T = 2 mu (f - R) + lam (J - 1) cof(f), where
// R = U V^T is the rotation of the polar decomposition f = R S and cof(f) = J f^-T.  R comes
// from Newton's iteration X <- (X + X^-T) / 2 from X = f (Higham 1986), which converges
// quadratically: an update that changes X by c leaves an error of about c^2 / 2 in the new
// iterate, so the iteration stops once c <= 1e-4 (Frobenius), when that error is below float
// resolution.  Returns false when that has not happened within 12 iterations (a nearly
// singular f): the caller then takes the SVD path.
__device__ inline bool corotated_polar(const m3& f, float J, float mu, float lam, m3& P) {
  m3 C = cofactor(f);
  m3 X = f, Cx = C;
  float det = J;
  bool converged = false;
#pragma unroll 1
  for (int it = 0; it < 12; ++it) {
    float half_inv = 0.5f * __builtin_amdgcn_rcpf(det);
    float change = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      float y = 0.5f * X.a[k] + Cx.a[k] * half_inv;
      float d = y - X.a[k];
      change += d * d;
      X.a[k] = y;
    }
    if (change <= 1.0e-8f) {
      converged = true;
      break;
    }
    Cx = cofactor(X);
    det = X.a[0] * Cx.a[0] + X.a[1] * Cx.a[1] + X.a[2] * Cx.a[2];
  }
  float tm = 2.0f * mu, lj = lam * (J - 1.0f);
#pragma unroll
  for (int k = 0; k < 9; ++k) P.a[k] = tm * (f.a[k] - X.a[k]) + lj * C.a[k];
  return converged;
}

__device__ __forceinline__ bool is_identity(const m3& m) {
  return m.a[0] == 1.0f && m.a[4] == 1.0f && m.a[8] == 1.0f && m.a[1] == 0.0f && m.a[2] == 0.0f && m.a[3] == 0.0f &&
         m.a[5] == 0.0f && m.a[6] == 0.0f && m.a[7] == 0.0f;
}

// ---------------------------------------------------------------------------------------
// Ternary cut codes: digits agree wherever both are non-zero (material.py:23-34).  Codes
// are non-negative int64; 0 (uncut) is compatible with everything.
__device__ __forceinline__ bool compatible(int64_t a, int64_t b) {
  if (a == 0 || b == 0 || a == b) return true;
  uint64_t x = uint64_t(a), y = uint64_t(b);
  while (x != 0 && y != 0) {
    uint64_t qx = x / 3u, qy = y / 3u;
    uint64_t dx = x - 3u * qx, dy = y - 3u * qy;
    if (dx != 0 && dy != 0 && dx != dy) return false;
    x = qx;
    y = qy;
  }
  return true;
}

// ---------------------------------------------------------------------------------------
// Trilinear stencil of one particle from its local cell fractions l (0..1 per axis).
// Corner a has offsets (a >> 2 & 1, a >> 1 & 1, a & 1); weight = product of factors,
// gradient component = sign * product of the two other factors / h.
struct Stencil {
  float w[8];
  f3 g[8];
};
__device__ __forceinline__ void stencil_weights(f3 l, float* w) {
  float fx[2] = {1.0f - l.x, l.x}, fy[2] = {1.0f - l.y, l.y}, fz[2] = {1.0f - l.z, l.z};
#pragma unroll
  for (int a = 0; a < 8; ++a) w[a] = fx[(a >> 2) & 1] * fy[(a >> 1) & 1] * fz[a & 1];
}
__device__ __forceinline__ void stencil_full(f3 l, float inv_h, Stencil& st) {
  float fx[2] = {1.0f - l.x, l.x}, fy[2] = {1.0f - l.y, l.y}, fz[2] = {1.0f - l.z, l.z};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    int ix = (a >> 2) & 1, iy = (a >> 1) & 1, iz = a & 1;
    float sx = ix ? 1.0f : -1.0f, sy = iy ? 1.0f : -1.0f, sz = iz ? 1.0f : -1.0f;
    st.w[a] = fx[ix] * fy[iy] * fz[iz];
    st.g[a] = f3{sx * (fy[iy] * fz[iz]) * inv_h, sy * (fx[ix] * fz[iz]) * inv_h, sz * (fx[ix] * fy[iy]) * inv_h};
  }
}
// Weight and reference gradient of one corner (the same products as stencil_full).
__device__ __forceinline__ void stencil_corner(f3 l, float inv_h, int a, float& w, f3& g) {
  int ix = (a >> 2) & 1, iy = (a >> 1) & 1, iz = a & 1;
  float fx = ix ? l.x : 1.0f - l.x, fy = iy ? l.y : 1.0f - l.y, fz = iz ? l.z : 1.0f - l.z;
  float sx = ix ? 1.0f : -1.0f, sy = iy ? 1.0f : -1.0f, sz = iz ? 1.0f : -1.0f;
  w = fx * fy * fz;
  g = f3{sx * (fy * fz) * inv_h, sy * (fx * fz) * inv_h, sz * (fx * fy) * inv_h};
}
// Node offset of corner a in a grid with (dy, dz) node counts along y and z.
__device__ __forceinline__ int corner_offset(int a, int dyz, int dz) {
  return ((a >> 2) & 1) * dyz + ((a >> 1) & 1) * dz + (a & 1);
}

}  // namespace bm
