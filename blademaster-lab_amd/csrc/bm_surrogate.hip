// Interval-wrench surrogate on gfx950: a 3-layer ReLU network trained and evaluated by
// hand-written MFMA kernels, and the occupancy/damage patch around each knife edge that feeds
// it.  The reference has no learned component; its intra-interval predictor is the mobility
// matrix of tool_substep "predict" (tool.py:97-111).
//
// Training (float32).  One GEMM kernel serves every product of a training step: the forward
// layer [x, 1] Wa, the weight gradient [x, 1]^T dY and the input gradient dY Wa^T.  A
// workgroup of 4 waves owns a 64 x 64 tile of C; each wave a 32 x 32 quarter, as 2 x 2 tiles
// of v_mfma_f32_16x16x4_f32.  K is staged through LDS in slices of 16, loaded so that global
// reads run along the contiguous axis of each operand (both transposes).  The implicit column
// (or row) of ones of the augmented layers is produced by the loader, and the ReLU of the
// forward pass and the ReLU mask of the backward pass are applied in the epilogue, so a
// training step is 9 GEMMs, the loss and one Adam launch over the flat parameter vector.
//
// Inference (bfloat16 weights and activations, float32 accumulation).  One fused launch runs
// all three layers for 16 rows per workgroup: the activations stay in LDS, each of the 4 waves
// computes 64 hidden columns with v_mfma_f32_16x16x32_bf16 from weights packed transposed
// (k contiguous, 16 bytes per lane and step), and the last layer runs on wave 0.
//
// MFMA 16x16 operand layout (wave64): lane l supplies row l % 16 of A and column l % 16 of B
// for the k values of its group l / 16; the result holds C[4 (l / 16) + r][l % 16], r = 0..3.
// A and B use the same k assignment, so the sum over k does not depend on its order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "bm_device.h"
#include "bm_hip.h"

using namespace bm;

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 64;
constexpr int kSlice = 16;
constexpr int kPitch = kTile + 4;
constexpr int kThreads = 256;

// Inference kernel shape: hidden width 256, inputs + 1 padded to 128, hidden + 1 to 288.
constexpr int kHidden = 256;
constexpr int kK1 = 128;
constexpr int kK2 = 288;
constexpr int kOutPad = 16;
constexpr int kRows = 16;
// LDS row pitch of the activations, in bf16: +8 staggers the rows of one 16-lane group over
// different banks for the 16-byte fragment reads.
constexpr int kActPitch1 = kK1 + 8;
constexpr int kActPitch2 = kK2 + 8;

__device__ __forceinline__ float bf16_to_float(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
// Round to nearest, ties to even (finite values).
__device__ __forceinline__ uint16_t float_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Operand A of a GEMM: op(A) [M][K] with `rows` x `cols` stored values; `ones` 1 adds a
// column of ones at k = cols, 2 a row of ones at m = rows; everything else reads as zero.
struct OperandA {
  const float* a;
  int lda, trans, rows, cols, ones;
  __device__ __forceinline__ float at(int r, int k) const {
    if (r < rows && k < cols) return trans ? a[int64_t(k) * lda + r] : a[int64_t(r) * lda + k];
    if (ones == 1 && k == cols && r < rows) return 1.0f;
    if (ones == 2 && r == rows && k < cols) return 1.0f;
    return 0.0f;
  }
};

// C [M][N] = op(A) op(B), op(B) [K][N] (trans_b: B stored [N][K]).  relu: max(., 0);
// mask (layout of C): zero where mask <= 0.  Split K: workgroup z sums k in
// [z k_chunk, (z + 1) k_chunk) into C + z split_stride.
__global__ void __launch_bounds__(kThreads) k_gemm(OperandA A, const float* B, float* C, const float* mask, int M,
                                                   int N, int K, int ldb, int ldc, int trans_b, int relu, int k_chunk,
                                                   int64_t split_stride) {
  __shared__ float As[kSlice][kPitch];
  __shared__ float Bs[kSlice][kPitch];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row0 = blockIdx.y * kTile, col0 = blockIdx.x * kTile;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int k_lo = blockIdx.z * k_chunk, k_hi = min(K, k_lo + k_chunk);
  C += blockIdx.z * split_stride;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = k_lo; k0 < k_hi; k0 += kSlice) {
#pragma unroll
    for (int q = 0; q < (kSlice * kTile) / kThreads; ++q) {
      int e = tid + q * kThreads;
      // op(A) tile: read along k when A is row-major, along m when it is transposed.
      int ar = A.trans ? e % kTile : e / kSlice;
      int ak = A.trans ? e / kTile : e % kSlice;
      As[ak][ar] = (row0 + ar < M && k0 + ak < k_hi) ? A.at(row0 + ar, k0 + ak) : 0.0f;
      int bc = trans_b ? e / kSlice : e % kTile;
      int bk = trans_b ? e % kSlice : e / kTile;
      float bv = 0.0f;
      if (col0 + bc < N && k0 + bk < k_hi)
        bv = trans_b ? B[int64_t(col0 + bc) * ldb + k0 + bk] : B[int64_t(k0 + bk) * ldb + col0 + bc];
      Bs[bk][bc] = bv;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kSlice; ks += 4) {
      int k = ks + (lane >> 4);
      float a0 = As[k][wr + (lane & 15)], a1 = As[k][wr + 16 + (lane & 15)];
      float b0 = Bs[k][wc + (lane & 15)], b1 = Bs[k][wc + 16 + (lane & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int c = col0 + wc + 16 * j + (lane & 15);
      if (c >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = row0 + wr + 16 * i + 4 * (lane >> 4) + r;
        if (row >= M) continue;
        float v = acc[i][j][r];
        if (relu) v = fmaxf(v, 0.0f);
        if (mask && !(mask[int64_t(row) * ldc + c] > 0.0f)) v = 0.0f;
        C[int64_t(row) * ldc + c] = v;
      }
    }
}

// C[i] = sum over z of partial[z][i], in split order.
__global__ void k_sum_splits(const float* partial, int splits, int64_t count, float* C) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
  for (int z = 0; z < splits; ++z) s += partial[z * count + i];
  C[i] = s;
}

// Mean squared error over n values, d = 2 (p - t) / n, in one workgroup (fixed order); thread 0
// also counts the training step.
constexpr int kLossThreads = 1024;
__global__ void __launch_bounds__(kLossThreads) k_mse(const float* P, const float* T, float* dP, float* loss,
                                                      int32_t* step, int64_t n) {
  __shared__ float part[kLossThreads / 64];
  float inv = 1.0f / float(n), s = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += kLossThreads) {
    float d = P[i] - T[i];
    dP[i] = 2.0f * d * inv;
    s += d * d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float total = 0.0f;
    for (int w = 0; w < kLossThreads / 64; ++w) total += part[w];
    loss[0] = total * inv;
    if (step) step[0] += 1;
  }
}

// Adam (Kingma and Ba, 2015) with bias correction and decoupled weight decay (Loshchilov and
// Hutter, 2019): w -= lr (m^ / (sqrt(v^) + eps) + decay w); hyper = {lr, decay}; step[0] is the
// 1-based step of this update.
__global__ void k_adam(float* w, const float* g, float* m, float* v, int64_t n, const float* hyper,
                       const int32_t* step) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1.0e-8f;
  float t = float(step[0]);
  float c1 = 1.0f - powf(beta1, t), c2 = 1.0f - powf(beta2, t);
  float gi = g[i];
  float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i];
  w[i] = wi - hyper[0] * ((mi / c1) / (sqrtf(vi / c2) + eps) + hyper[1] * wi);
}

// Wt [cols_pad][rows_pad] = bf16(Wa [rows][cols])^T, zero padded.
__global__ void k_pack(const float* Wa, int rows, int cols, uint16_t* Wt, int rows_pad, int cols_pad) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(rows_pad) * cols_pad) return;
  int c = int(i / rows_pad), r = int(i % rows_pad);
  Wt[i] = (r < rows && c < cols) ? float_to_bf16(Wa[int64_t(r) * cols + c]) : uint16_t(0);
}

// 16 rows of act (pitch in bf16) times the packed layer Wt [cols][kpad], columns
// [c0, c0 + 16 tiles): the wave's 16 x 16 tile at column block c0.
template <int KPAD, int PITCH>
__device__ __forceinline__ floatx4 mfma_tile(const uint16_t* act, const uint16_t* Wt, int c0, int lane) {
  floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
  const uint16_t* arow = act + (lane & 15) * PITCH + 8 * (lane >> 4);
  const uint16_t* brow = Wt + int64_t(c0 + (lane & 15)) * KPAD + 8 * (lane >> 4);
#pragma unroll
  for (int k = 0; k < KPAD; k += 32) {
    uint4 a = *reinterpret_cast<const uint4*>(arow + k);
    uint4 b = *reinterpret_cast<const uint4*>(brow + k);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
  }
  return acc;
}

// Hidden layer: ReLU(act W) for 16 rows into the next activation buffer, as bf16.
template <int KPAD, int PITCH_IN, int PITCH_OUT>
__device__ __forceinline__ void hidden_layer(const uint16_t* act, const uint16_t* Wt, uint16_t* next, int wave,
                                             int lane) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    int c0 = 64 * wave + 16 * t;
    floatx4 acc = mfma_tile<KPAD, PITCH_IN>(act, Wt, c0, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      next[(4 * (lane >> 4) + r) * PITCH_OUT + c0 + (lane & 15)] = float_to_bf16(fmaxf(acc[r], 0.0f));
  }
}

__global__ void __launch_bounds__(kThreads) k_infer(const float* x, float* y, const uint16_t* packed, int M,
                                                    int inputs, int outputs, const float* x_shift,
                                                    const float* x_scale, const float* y_shift, const float* y_scale) {
  __shared__ __align__(16) uint16_t act0[kRows * kActPitch1];
  __shared__ __align__(16) uint16_t act1[kRows * kActPitch2];
  __shared__ __align__(16) uint16_t act2[kRows * kActPitch2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row0 = blockIdx.x * kRows;
  const uint16_t one = 0x3f80;  // 1.0 in bf16
  for (int e = tid; e < kRows * kK1; e += kThreads) {
    int r = e / kK1, k = e % kK1;
    uint16_t v = 0;
    if (k < inputs && row0 + r < M) {
      float xv = x[int64_t(row0 + r) * inputs + k];
      if (x_shift) xv -= x_shift[k];
      if (x_scale) xv *= x_scale[k];
      v = float_to_bf16(xv);
    }
    else if (k == inputs) v = one;
    act0[r * kActPitch1 + k] = v;
  }
  // Bias column and zero padding of the hidden activations (columns 256..287).
  for (int e = tid; e < kRows * (kK2 - kHidden); e += kThreads) {
    int r = e / (kK2 - kHidden), k = kHidden + e % (kK2 - kHidden);
    uint16_t v = k == kHidden ? one : uint16_t(0);
    act1[r * kActPitch2 + k] = v;
    act2[r * kActPitch2 + k] = v;
  }
  __syncthreads();
  const uint16_t* W1 = packed;
  const uint16_t* W2 = W1 + kHidden * kK1;
  const uint16_t* W3 = W2 + kHidden * kK2;
  hidden_layer<kK1, kActPitch1, kActPitch2>(act0, W1, act1, wave, lane);
  __syncthreads();
  hidden_layer<kK2, kActPitch2, kActPitch2>(act1, W2, act2, wave, lane);
  __syncthreads();
  if (wave == 0) {
    floatx4 acc = mfma_tile<kK2, kActPitch2>(act2, W3, 0, lane);
    int c = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int row = row0 + 4 * (lane >> 4) + r;
      if (c < outputs && row < M)
        y[int64_t(row) * outputs + c] = acc[r] * (y_scale ? y_scale[c] : 1.0f) + (y_shift ? y_shift[c] : 0.0f);
    }
  }
}

// ------------------------------------------------------------------ knife-edge patch
constexpr int kPatch = 8;
constexpr int kPatchCells = kPatch * kPatch * kPatch;

__global__ void k_patch_clear(float* out, int64_t count) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = 0.0f;
}

__global__ void k_patch(const bm_params P, const bm_state S, float* out) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  int inst = S.instance ? S.instance[p] : 0;
  quat q = ldq(S.tool_rot + 4 * inst);
  f3 local = rotate_inv(q, ld3(S.x, p) - ld3(S.tool_pos, inst));
  float h = float(P.spacing), hy = float(P.half[1]), hz = float(P.half[2]);
  // Cells: h across the blade and about the edge, hy / 4 along it.
  int ix = int(floorf((local.x + 4.0f * h) / h));
  int iy = int(floorf((local.y + hy) * (4.0f / hy)));
  int iz = int(floorf((local.z + hz + 4.0f * h) / h));
  if (ix < 0 || ix >= kPatch || iy < 0 || iy >= kPatch || iz < 0 || iz >= kPatch) return;
  int64_t cell = int64_t(inst) * 2 * kPatchCells + (ix * kPatch + iy) * kPatch + iz;
  atomicAdd(out + cell, 1.0f);
  atomicAdd(out + cell + kPatchCells, S.damage[p]);
}

inline int blocks_for(int64_t n, int b = 256) { return int((n + b - 1) / b); }

int status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

struct Layers {
  int64_t off[3];
  int rows[3], cols[3];
};

Layers layers_of(const bm_mlp& m) {
  Layers L;
  L.rows[0] = m.inputs;
  L.cols[0] = m.hidden;
  L.rows[1] = m.hidden;
  L.cols[1] = m.hidden;
  L.rows[2] = m.hidden;
  L.cols[2] = m.outputs;
  int64_t o = 0;
  for (int i = 0; i < 3; ++i) {
    L.off[i] = o;
    o += int64_t(L.rows[i] + 1) * L.cols[i];
  }
  return L;
}

bool shape_ok(const bm_mlp* m) {
  return m && m->inputs > 0 && m->hidden > 0 && m->outputs > 0 && m->params && m->h1 && m->h2 && m->out;
}

// Largest number of K splits of a weight-gradient product, and the K each split covers at least.
constexpr int kMaxSplits = 16;
constexpr int kSplitK = 256;

void gemm(OperandA A, const float* B, float* C, const float* mask, int M, int N, int K, int ldb, int ldc, int trans_b,
          int relu, hipStream_t st) {
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_gemm, grid, dim3(kThreads), 0, st, A, B, C, mask, M, N, K, ldb, ldc, trans_b, relu, K,
                     int64_t(0));
}

// C [M][N] (ldc = N) = op(A) op(B) with K split over workgroups when the tile grid alone is
// too small to fill the device: partial products in `work`, then a fixed-order sum.
void gemm_split(OperandA A, const float* B, float* C, int M, int N, int K, int ldb, int trans_b, float* work,
                int64_t work_len, hipStream_t st) {
  int tiles = ((N + kTile - 1) / kTile) * ((M + kTile - 1) / kTile);
  int splits = std::min(kMaxSplits, std::max(1, K / kSplitK));
  while (splits > 1 && int64_t(splits) * M * N > work_len) --splits;
  if (!work || splits < 2 || tiles >= 256) {
    gemm(A, B, C, nullptr, M, N, K, ldb, N, trans_b, 0, st);
    return;
  }
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kSlice - 1) / kSlice * kSlice;
  splits = (K + chunk - 1) / chunk;
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile, splits);
  int64_t count = int64_t(M) * N;
  hipLaunchKernelGGL(k_gemm, grid, dim3(kThreads), 0, st, A, B, work, nullptr, M, N, K, ldb, N, trans_b, 0, chunk,
                     count);
  hipLaunchKernelGGL(k_sum_splits, dim3(blocks_for(count)), dim3(256), 0, st, work, splits, count, C);
}

// out = net(x): [x,1] Wa1 -> ReLU -> [h1,1] Wa2 -> ReLU -> [h2,1] Wa3.
void forward(const bm_mlp& m, const float* x, int batch, hipStream_t st) {
  Layers L = layers_of(m);
  const float* in[3] = {x, m.h1, m.h2};
  float* outs[3] = {m.h1, m.h2, m.out};
  for (int l = 0; l < 3; ++l) {
    OperandA A{in[l], L.rows[l], 0, batch, L.rows[l], 1};
    gemm(A, m.params + L.off[l], outs[l], nullptr, batch, L.cols[l], L.rows[l] + 1, L.cols[l], L.cols[l], 0, l < 2,
         st);
  }
}

void backward(const bm_mlp& m, const float* x, const float* target, int batch, hipStream_t st) {
  Layers L = layers_of(m);
  forward(m, x, batch, st);
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(kLossThreads), 0, st, m.out, target, m.d_out, m.loss, m.step,
                     int64_t(batch) * m.outputs);
  const float* in[3] = {x, m.h1, m.h2};
  const float* d_of[3] = {m.d_h1, m.d_h2, m.d_out};
  float* d_in[3] = {nullptr, m.d_h1, m.d_h2};
  for (int l = 2; l >= 0; --l) {
    // dWa_l = [in, 1]^T d_l: op(A) [rows + 1][batch] from in [batch][rows].
    OperandA At{in[l], L.rows[l], 1, L.rows[l], batch, 2};
    gemm_split(At, d_of[l], m.grads + L.off[l], L.rows[l] + 1, L.cols[l], batch, L.cols[l], 0, m.work, m.work_len,
               st);
    if (l > 0) {
      // d_in = (d_l Wa_l[:rows]^T) masked by ReLU(in) > 0.
      OperandA D{d_of[l], L.cols[l], 0, batch, L.cols[l], 0};
      gemm(D, m.params + L.off[l], d_in[l], in[l], batch, L.rows[l], L.cols[l], L.cols[l], L.rows[l], 1, 0, st);
    }
  }
}

}  // namespace

extern "C" {

int bm_mlp_sizes(int32_t inputs, int32_t hidden, int32_t outputs, int64_t* params, int64_t* packed,
                 int64_t* work) {
  if (inputs < 1 || hidden < 1 || outputs < 1 || !params || !packed || !work) return -1;
  int64_t l1 = int64_t(inputs + 1) * hidden, l2 = int64_t(hidden + 1) * hidden, l3 = int64_t(hidden + 1) * outputs;
  *params = l1 + l2 + l3;
  *packed = int64_t(kHidden) * kK1 + int64_t(kHidden) * kK2 + int64_t(kOutPad) * kK2;
  *work = kMaxSplits * std::max(l1, std::max(l2, l3));
  return 0;
}

int bm_mlp_gemm(const float* A, const float* B, float* C, int32_t M, int32_t N, int32_t K, int32_t lda, int32_t ldb,
                int32_t ldc, int32_t trans_a, int32_t trans_b, void* stream) {
  if (!A || !B || !C || M < 0 || N < 0 || K < 0) return -1;
  if (M == 0 || N == 0) return 0;
  OperandA op{A, lda, trans_a ? 1 : 0, M, K, 0};
  gemm(op, B, C, nullptr, M, N, K, ldb, ldc, trans_b ? 1 : 0, 0, (hipStream_t)stream);
  return status();
}

int bm_mlp_gemm_split(const float* A, const float* B, float* C, int32_t M, int32_t N, int32_t K, int32_t lda,
                      int32_t ldb, int32_t trans_a, int32_t trans_b, float* work, int64_t work_len, void* stream) {
  if (!A || !B || !C || !work || M < 0 || N < 0 || K < 0 || work_len < 0) return -1;
  if (M == 0 || N == 0) return 0;
  OperandA op{A, lda, trans_a ? 1 : 0, M, K, 0};
  gemm_split(op, B, C, M, N, K, ldb, trans_b ? 1 : 0, work, work_len, (hipStream_t)stream);
  return status();
}

int bm_mlp_forward(const bm_mlp* m, const float* x, int32_t batch, void* stream) {
  if (!shape_ok(m) || !x || batch < 0) return -1;
  if (batch == 0) return 0;
  forward(*m, x, batch, (hipStream_t)stream);
  return status();
}

int bm_mlp_backward(const bm_mlp* m, const float* x, const float* target, int32_t batch, void* stream) {
  if (!shape_ok(m) || !x || !target || !m->grads || !m->d_out || !m->d_h1 || !m->d_h2 || !m->loss || batch < 1)
    return -1;
  bm_mlp copy = *m;
  copy.step = nullptr;  // a gradient evaluation is not a training step
  backward(copy, x, target, batch, (hipStream_t)stream);
  return status();
}

int bm_mlp_train_step(const bm_mlp* m, const float* x, const float* target, int32_t batch, void* stream) {
  if (!shape_ok(m) || !x || !target || !m->grads || !m->d_out || !m->d_h1 || !m->d_h2 || !m->loss || !m->step ||
      !m->lr || !m->adam_m || !m->adam_v || batch < 1)
    return -1;
  hipStream_t st = (hipStream_t)stream;
  backward(*m, x, target, batch, st);
  int64_t count = 0, packed = 0, work = 0;
  bm_mlp_sizes(m->inputs, m->hidden, m->outputs, &count, &packed, &work);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(count)), dim3(256), 0, st, m->params, m->grads, m->adam_m, m->adam_v,
                     count, m->lr, m->step);
  return status();
}

int bm_mlp_pack(const bm_mlp* m, void* stream) {
  if (!m || !m->params || !m->packed) return -1;
  if (m->hidden != kHidden || m->inputs < 1 || m->inputs >= kK1 || m->outputs < 1 || m->outputs > kOutPad) return -2;
  hipStream_t st = (hipStream_t)stream;
  Layers L = layers_of(*m);
  const int rows_pad[3] = {kK1, kK2, kK2}, cols_pad[3] = {kHidden, kHidden, kOutPad};
  int64_t o = 0;
  for (int l = 0; l < 3; ++l) {
    int64_t count = int64_t(rows_pad[l]) * cols_pad[l];
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(count)), dim3(256), 0, st, m->params + L.off[l], L.rows[l] + 1,
                       L.cols[l], m->packed + o, rows_pad[l], cols_pad[l]);
    o += count;
  }
  return status();
}

int bm_mlp_infer(const bm_mlp* m, const float* x, float* y, int32_t batch, void* stream) {
  if (!m || !m->packed || !x || !y || batch < 0) return -1;
  if (m->hidden != kHidden || m->inputs < 1 || m->inputs >= kK1 || m->outputs < 1 || m->outputs > kOutPad) return -2;
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_infer, dim3((batch + kRows - 1) / kRows), dim3(kThreads), 0, (hipStream_t)stream, x, y,
                     m->packed, batch, m->inputs, m->outputs, m->x_shift, m->x_scale, m->y_shift, m->y_scale);
  return status();
}

int bm_blade_patch(const bm_params* p, const bm_state* s, float* out, void* stream) {
  if (!p || !s || !out || !s->x || !s->damage || !s->tool_pos || !s->tool_rot || p->instances < 1) return -1;
  if (p->instances > 1 && !s->instance) return -1;
  hipStream_t st = (hipStream_t)stream;
  int64_t count = p->instances * 2 * kPatchCells;
  hipLaunchKernelGGL(k_patch_clear, dim3(blocks_for(count)), dim3(256), 0, st, out, count);
  if (p->n > 0) hipLaunchKernelGGL(k_patch, dim3(blocks_for(p->n)), dim3(256), 0, st, *p, *s, out);
  return status();
}

}  // extern "C"
