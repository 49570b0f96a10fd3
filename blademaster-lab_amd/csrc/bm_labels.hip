// Piece labels (topology.components, topology.py:17-57) on gfx950.
//
// Graph vertices are the particles plus one vertex per (grid node, cut code) field.  A
// particle joins the fields of its stencil corners with positive weight; two fields of one
// node join when their codes are compatible.  The fields come from a registration of the
// particles' codes into a scratch field table of L slots per node (slot f of node v is
// vertex n + f N + v).
//
// Connected components by lock-free union-find: unite() hooks the larger root under the
// smaller with a CAS, so every component ends with its smallest vertex as root, which is
// its lowest particle index (every field vertex touches a particle).  Parent reads and the
// path-halving writes are agent-scope atomics (no stale L1 copies); every loop is bounded
// and counts an error instead of spinning if the bound is hit.  Labels are the ranks of the
// roots in particle order (an exclusive scan of "is a root"), so they equal the reference's
// ordering by lowest particle index bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "bm_device.h"
#include "bm_hip.h"

using namespace bm;

// Scratch buffers of bm_piece_labels_static: one file-scope array on every GPU the library
// runs on, held until the process ends.  On the host its name is only a stand-in; host code
// uses the device address that hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchBytes = size_t(BM_LABELS_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kBlock = 256;
constexpr int kScanBlock = 1024;
constexpr int kMaxIterations = 1 << 22;

inline int blocks_for(int64_t n, int block = kBlock) {
  int64_t b = (n + block - 1) / block;
  return int(b < 1 ? 1 : b);
}

__device__ __forceinline__ int32_t load_parent(const int32_t* parent, int32_t x) {
  return __hip_atomic_load(parent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving; -1 if the bound is hit.
__device__ int32_t find_root(int32_t* parent, int32_t x) {
  for (int it = 0; it < kMaxIterations; ++it) {
    int32_t px = load_parent(parent, x);
    if (px == x) return x;
    int32_t gx = load_parent(parent, px);
    if (gx != px) __hip_atomic_store(parent + x, gx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = px;
  }
  return -1;
}

__device__ void unite(int32_t* parent, int32_t a, int32_t b, int32_t* errors) {
  for (int it = 0; it < kMaxIterations; ++it) {
    a = find_root(parent, a);
    b = find_root(parent, b);
    if (a < 0 || b < 0) break;
    if (a == b) return;
    if (a < b) {
      int32_t t = a;
      a = b;
      b = t;
    }
    // a is a root larger than b: hook it under b if it is still a root.
    if (atomicCAS(parent + a, a, b) == a) return;
  }
  atomicAdd(errors, 1);
}

__global__ void k_lbl_clear_errors(int32_t* errors) { errors[0] = 0; }

__global__ void k_lbl_init(bm_labels L, int64_t n, int64_t FN, int64_t N) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n + FN) L.parent[i] = int32_t(i);
  if (i < FN) L.field_code[i] = -1;
  if (i < N) L.node_nf[i] = 0;
}

// Field slot of each (particle, corner) for the particle's code: the first slot holding the
// code, claiming the first empty one (CAS).  A node with more than F codes counts an error.
__global__ void k_lbl_register(const bm_params P, const bm_state S, bm_labels L) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  int64_t code = S.codes[p];
  int32_t base = S.base_node[p];
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  int64_t N = P.node_count;
  int F = int(L.fields);
  for (int a = 0; a < 8; ++a) {
    int64_t node = int64_t(base) + corner_offset(a, dyz, dz);
    int found = -1;
    for (int f = 0; f < F; ++f) {
      int64_t* slot = L.field_code + int64_t(f) * N + node;
      unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)(-1LL),
                                          (unsigned long long)code);
      if (int64_t(prev) == -1) atomicMax(L.node_nf + node, f + 1);
      if (int64_t(prev) == -1 || int64_t(prev) == code) {
        found = f;
        break;
      }
    }
    if (found < 0) {
      atomicAdd(L.errors, 1);
      found = 0;
    }
    L.field_idx[8 * p + a] = uint8_t(found);
  }
}

// Particle p joins the fields of its positive-weight corners.
__global__ void k_lbl_join_particles(const bm_params P, const bm_state S, bm_labels L) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  int32_t base = S.base_node[p];
  int dz = int(P.dims[2]), dyz = int(P.dims[1] * P.dims[2]);
  int64_t N = P.node_count;
  uint8_t mask = L.corner_mask[p];
  for (int a = 0; a < 8; ++a) {
    if (!((mask >> a) & 1)) continue;
    int64_t node = int64_t(base) + corner_offset(a, dyz, dz);
    int64_t v = P.n + int64_t(L.field_idx[8 * p + a]) * N + node;
    unite(L.parent, int32_t(p), int32_t(v), L.errors);
  }
}

// Compatible fields of one node join.
__global__ void k_lbl_join_fields(const bm_params P, bm_labels L) {
  int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (node >= P.node_count) return;
  int nf = L.node_nf[node];
  int64_t N = P.node_count;
  for (int i = 0; i < nf; ++i)
    for (int j = i + 1; j < nf; ++j)
      if (compatible(L.field_code[int64_t(i) * N + node], L.field_code[int64_t(j) * N + node]))
        unite(L.parent, int32_t(P.n + int64_t(i) * N + node), int32_t(P.n + int64_t(j) * N + node), L.errors);
}

// root[p]; per-block counts of roots.
__global__ void __launch_bounds__(kScanBlock) k_lbl_roots(const bm_params P, bm_labels L) {
  __shared__ int32_t count;
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < P.n) {
    int32_t r = find_root(L.parent, int32_t(p));
    if (r < 0) {
      atomicAdd(L.errors, 1);
      r = int32_t(p);
    }
    L.root[p] = r;
    if (r == p) atomicAdd(&count, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) L.block_sum[blockIdx.x] = count;
}

// Exclusive scan of the block counts (one workgroup); the total is the number of pieces.
__global__ void __launch_bounds__(kScanBlock) k_lbl_scan(bm_labels L, int nblocks) {
  __shared__ int32_t buf[kScanBlock];
  int32_t carry = 0;
  for (int base = 0; base < nblocks; base += kScanBlock) {
    int i = base + int(threadIdx.x);
    int32_t v = i < nblocks ? L.block_sum[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      int32_t t = threadIdx.x >= unsigned(off) ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblocks) L.block_sum[i] = carry + buf[threadIdx.x] - v;
    int32_t total = buf[kScanBlock - 1];
    __syncthreads();
    carry += total;
  }
  if (threadIdx.x == 0) L.count[0] = carry;
}

// Rank of each root in particle order: block offset + rank inside the block.
__global__ void __launch_bounds__(kScanBlock) k_lbl_rank(const bm_params P, bm_labels L) {
  __shared__ int32_t buf[kScanBlock];
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int32_t is_root = (p < P.n && L.root[p] == p) ? 1 : 0;
  buf[threadIdx.x] = is_root;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int32_t t = threadIdx.x >= unsigned(off) ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  if (is_root) L.labels[p] = L.block_sum[blockIdx.x] + buf[threadIdx.x] - 1;
}

__global__ void k_lbl_label(const bm_params P, bm_labels L) {
  int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P.n) return;
  int32_t r = L.root[p];
  if (r != p) L.labels[p] = L.labels[r];
}

int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : int(e);
}

// The launch sequence shared by both entry points; l holds every buffer.
int launch_labels(const bm_params* p, const bm_state* s, const bm_labels* l, hipStream_t st) {
  int64_t FN = l->fields * p->node_count;
  int64_t init = p->n + FN;
  hipLaunchKernelGGL(k_lbl_init, dim3(blocks_for(init)), dim3(kBlock), 0, st, *l, p->n, FN, p->node_count);
  hipLaunchKernelGGL(k_lbl_register, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s, *l);
  hipLaunchKernelGGL(k_lbl_join_particles, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *s, *l);
  hipLaunchKernelGGL(k_lbl_join_fields, dim3(blocks_for(p->node_count)), dim3(kBlock), 0, st, *p, *l);
  int nb = blocks_for(p->n, kScanBlock);
  hipLaunchKernelGGL(k_lbl_roots, dim3(nb), dim3(kScanBlock), 0, st, *p, *l);
  hipLaunchKernelGGL(k_lbl_scan, dim3(1), dim3(kScanBlock), 0, st, *l, nb);
  hipLaunchKernelGGL(k_lbl_rank, dim3(nb), dim3(kScanBlock), 0, st, *p, *l);
  hipLaunchKernelGGL(k_lbl_label, dim3(blocks_for(p->n)), dim3(kBlock), 0, st, *p, *l);
  return launch_status();
}

// ------------------------------------------------------------------ no-allocation variant
constexpr int kMaxDevices = 64;
unsigned char* scratch_address[kMaxDevices] = {};
std::mutex scratch_lock;

// Device address of the current device's scratch array, looked up on its first use there.
int scratch_of_current_device(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -4;
  std::lock_guard<std::mutex> guard(scratch_lock);
  if (scratch_address[dev] == nullptr) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) return -4;
    scratch_address[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = scratch_address[dev];
  return 0;
}

inline int64_t aligned(int64_t bytes) { return (bytes + 255) & ~int64_t(255); }

}  // namespace

extern "C" int bm_piece_labels(const bm_params* p, const bm_state* s, const bm_labels* l, void* stream) {
  if (!p || !s || !l || !s->codes || !s->base_node || !l->corner_mask || !l->field_code || !l->node_nf ||
      !l->field_idx || !l->parent || !l->root || !l->block_sum || !l->labels || !l->count || !l->errors)
    return -1;
  if (p->n < 0 || l->fields < 1 || l->fields > 255) return -1;
  int64_t FN = l->fields * p->node_count;
  if (p->n + FN >= (int64_t(1) << 31)) return -2;
  if (p->n == 0) return 0;
  return launch_labels(p, s, l, (hipStream_t)stream);
}

extern "C" int64_t bm_piece_labels_static_bytes(int64_t n, int64_t node_count, int64_t fields) {
  if (n < 0 || node_count < 0 || fields < 1 || fields > 255) return -1;
  int64_t FN = fields * node_count;
  return aligned(FN * 8) + aligned(node_count * 4) + aligned(n * 8) + aligned((n + FN) * 4) + aligned(n * 4) +
         aligned(((n + kScanBlock - 1) / kScanBlock) * 4);
}

extern "C" int bm_piece_labels_static(const bm_params* p, const bm_state* s, int64_t fields,
                                      const uint8_t* corner_mask, int32_t* labels, int32_t* count, int32_t* errors,
                                      void* stream) {
  if (!p || !s || !s->codes || !s->base_node || !corner_mask || !labels || !count || !errors) return -1;
  if (p->n < 0 || p->node_count < 0 || fields < 1 || fields > 255) return -1;
  int64_t FN = fields * p->node_count;
  if (p->n + FN >= (int64_t(1) << 31)) return -2;
  if (bm_piece_labels_static_bytes(p->n, p->node_count, fields) > int64_t(kScratchBytes)) return -3;
  if (p->n == 0) return 0;
  unsigned char* base = nullptr;
  int code = scratch_of_current_device(&base);
  if (code) return code;
  bm_labels l;
  l.fields = fields;
  l.corner_mask = corner_mask;
  int64_t at = 0;
  l.field_code = reinterpret_cast<int64_t*>(base + at);
  at += aligned(FN * 8);
  l.node_nf = reinterpret_cast<int32_t*>(base + at);
  at += aligned(p->node_count * 4);
  l.field_idx = base + at;
  at += aligned(p->n * 8);
  l.parent = reinterpret_cast<int32_t*>(base + at);
  at += aligned((p->n + FN) * 4);
  l.root = reinterpret_cast<int32_t*>(base + at);
  at += aligned(p->n * 4);
  l.block_sum = reinterpret_cast<int32_t*>(base + at);
  l.labels = labels;
  l.count = count;
  l.errors = errors;
  // The error count is an output that starts at zero: cleared by a kernel (no memset node
  // when the caller records this call into a graph).
  hipLaunchKernelGGL(k_lbl_clear_errors, dim3(1), dim3(1), 0, (hipStream_t)stream, errors);
  return launch_labels(p, s, &l, (hipStream_t)stream);
}
