// Dij products on the device: the planning loop's D = Dij @ W for many weight vectors
// at once (the reference's DijResult.dose_for_weights, pyradmc/scoring/dij.py:251-260,
// forms one vector's dose on the host). C ABI in include/hipsp.h.
//
// Set-up (hsp_create) turns the CSC Dij into two row-major forms on the device:
//  * CSR with every row in ascending column order (repeated entries in CSC order):
//    the columns are scattered one launch per column, in column order, so rows come
//    out sorted, and a per-row insertion pass settles ties by CSC position;
//  * row blocks: for each block of 8 consecutive rows (voxels adjacent along the
//    fastest grid axis, which see nearly the same beamlets), the union of the 8 rows'
//    columns in ascending order, each with its 8 values (0 where a row has no entry).
// Values stay in the type the caller gave (float64, as the reference's Dij holds them,
// or float32); a float64 product reads either exactly.
// A product walks the row blocks. One weight row W[j][.] loaded from cache serves up
// to 8 voxels, which cuts the L2 traffic for the weights by the block's overlap. Every
// voxel still adds its own entries in ascending column order: in float64 with the
// product rounded before the add, this is the reference's sum bit for bit.
//
// Roofline (HBM): per product, 4 + value bytes per CSR nonzero (as row blocks: 4 + 8
// values per block entry, read once per vector pass), the output (8 or 4 bytes per
// (row, vector)) and the weights (n_cols * n_w elements, cache-resident).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/hipsp.h"

namespace {
constexpr int kBlockRows = 8;     // rows per row block
constexpr int kHalfRows = kBlockRows / 2;  // rows per lane in the chunked kernels (a lane pair per entry)
constexpr int kSuperBlocks = 16;  // row blocks per super-block (one workgroup, a wave per block)
}

struct hsp_matrix {
  int device = 0;
  hipStream_t stream = nullptr;
  long long n_rows = 0, n_cols = 0, nnz = 0;
  int value_type = HSP_F64;      // element type of val / blk_val
  size_t value_bytes = 8;
  long long* row_ptr = nullptr;  // CSR [n_rows + 1]
  int* col = nullptr;            // [nnz]
  void* val = nullptr;           // [nnz]
  long long n_blocks = 0;        // ceil(n_rows / 8)
  long long n_block_entries = 0;
  long long* blk_ptr = nullptr;  // [n_blocks + 1]
  int* blk_col = nullptr;        // [n_block_entries]
  void* blk_val = nullptr;       // [n_block_entries][8]
  void* w = nullptr;             // weights on the device (capacity w_bytes)
  size_t w_bytes = 0;
  void* out = nullptr;           // internal product buffer (capacity out_bytes)
  size_t out_bytes = 0;
  long long out_r0 = 0, out_r1 = 0;
  int out_nw = 0, out_precision = HSP_F64;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // CSC copy for the transpose product, and its vector / output buffers
  long long* csc_ptr = nullptr;  // [n_cols + 1]
  int* csc_row = nullptr;        // [nnz]
  void* csc_val = nullptr;       // [nnz] of value_type
  double* tv = nullptr;          // [n_rows]
  double* tout = nullptr;        // [n_cols]
  // Super-blocks (kSuperBlocks consecutive row blocks) for the tiled product; built
  // only when n_cols <= 2^24.
  long long n_super = 0, n_super_entries = 0;
  long long* sb_ptr = nullptr;   // [n_super + 1]
  int* sb_col = nullptr;         // [n_super_entries]: distinct columns of the super-block, ascending
  int* blk_word = nullptr;       // [n_block_entries]: (index in the super-block's columns << 8) | row mask
  std::string error;
};

namespace {

int fail(hsp_matrix* m, int code, const std::string& msg) {
  if (m) m->error = msg;
  return code;
}

#define HSP_HIP(m, call)                                                                       \
  do {                                                                                         \
    hipError_t _e = (call);                                                                    \
    if (_e != hipSuccess) {                                                                    \
      (void)hipGetLastError(); /* reported here, not again by a later call */                  \
      return fail(m, _e == hipErrorOutOfMemory ? HSP_ENOMEM : HSP_EHIP,                        \
                  std::string(#call) + ": " + hipGetErrorString(_e));                          \
    }                                                                                          \
  } while (0)

// Largest grid of 256-thread workgroups one launch takes (2^32 - 1 work-items).
constexpr long long kMaxWorkgroups = 16777215;

unsigned int grid_for(long long n, long long cap = 65536) {
  long long b = (n + 255) / 256;
  return static_cast<unsigned int>(std::max(1ll, std::min(b, cap)));
}

// ---------------------------------------------------------------------------------------
// Set-up kernels.

__global__ void k_count_rows(const long long* __restrict__ rows, long long nnz,
                             unsigned int* __restrict__ count) {
  for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < nnz;
       e += static_cast<long long>(gridDim.x) * blockDim.x)
    atomicAdd(count + rows[e], 1u);
}

// Exclusive scan of n counts into out[0..n] (out[n] = total): per-tile sums, a serial
// scan of the tile sums, then a scan inside each 1024-entry tile.
constexpr int kScanTile = 1024;

__global__ void k_tile_sums(const unsigned int* __restrict__ count, long long n,
                            long long* __restrict__ tile_sum) {
  __shared__ long long red[256];
  const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
  long long s = 0;
  for (int i = threadIdx.x; i < kScanTile; i += blockDim.x)
    if (base + i < n) s += count[base + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = red[0];
}

__global__ void k_scan_tile_sums(long long* __restrict__ tile_sum, long long tiles,
                                 long long* __restrict__ total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    long long run = 0;
    for (long long t = 0; t < tiles; ++t) {
      const long long v = tile_sum[t];
      tile_sum[t] = run;
      run += v;
    }
    *total = run;
  }
}

__global__ void k_tile_scan(const unsigned int* __restrict__ count, long long n,
                            const long long* __restrict__ tile_off, long long* __restrict__ out) {
  __shared__ long long vals[kScanTile];
  const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += blockDim.x)
    vals[i] = base + i < n ? count[base + i] : 0;
  __syncthreads();
  if (threadIdx.x == 0) {  // serial scan of one tile (1024 adds)
    long long run = tile_off[blockIdx.x];
    for (int i = 0; i < kScanTile; ++i) {
      const long long v = vals[i];
      vals[i] = run;
      run += v;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kScanTile; i += blockDim.x)
    if (base + i < n) out[base + i] = vals[i];
}

// Entries [e_lo, e_hi) of CSC column c into their rows. One launch per column, in
// column order, so each row's entries arrive in ascending column order.
template <typename V>
__global__ void k_scatter_column(long long e_lo, long long e_hi, int c, const long long* __restrict__ rows,
                                 const V* __restrict__ values, const long long* __restrict__ row_ptr,
                                 unsigned int* __restrict__ cursor, int* __restrict__ col,
                                 V* __restrict__ val, long long* __restrict__ src) {
  const long long e = e_lo + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= e_hi) return;
  const long long r = rows[e];
  const long long pos = row_ptr[r] + atomicAdd(cursor + r, 1u);
  col[pos] = c;
  val[pos] = values[e];
  src[pos] = e;
}

// Order every row by (column, CSC position). Rows arrive sorted by column, so this
// only reorders repeated (row, column) entries; it is linear in the row length.
template <typename V>
__global__ void k_sort_rows(const long long* __restrict__ row_ptr, long long n_rows,
                            int* __restrict__ col, V* __restrict__ val, long long* __restrict__ src) {
  for (long long r = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; r < n_rows;
       r += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long lo = row_ptr[r], hi = row_ptr[r + 1];
    for (long long i = lo + 1; i < hi; ++i) {
      const int c = col[i];
      const V v = val[i];
      const long long s = src[i];
      long long j = i - 1;
      while (j >= lo && (col[j] > c || (col[j] == c && src[j] > s))) {
        col[j + 1] = col[j];
        val[j + 1] = val[j];
        src[j + 1] = src[j];
        --j;
      }
      col[j + 1] = c;
      val[j + 1] = v;
      src[j + 1] = s;
    }
  }
}

// Merge of the (up to 8) sorted rows of block b: calls emit(column, values[8]) once
// per union entry, in ascending column order. A repeated (row, column) pair becomes
// a second union entry for the same column, so each row keeps its own order.
template <typename V, typename Emit>
__device__ void merge_block(const long long* __restrict__ row_ptr, const int* __restrict__ col,
                            const V* __restrict__ val, long long n_rows, long long b, Emit emit) {
  long long cur[kBlockRows], end[kBlockRows];
  for (int i = 0; i < kBlockRows; ++i) {
    const long long r = b * kBlockRows + i;
    cur[i] = r < n_rows ? row_ptr[r] : 0;
    end[i] = r < n_rows ? row_ptr[r + 1] : 0;
  }
  for (;;) {
    int c_min = 0x7fffffff;
    bool any = false;
    for (int i = 0; i < kBlockRows; ++i)
      if (cur[i] < end[i]) {
        c_min = min(c_min, col[cur[i]]);
        any = true;
      }
    if (!any) return;
    V v[kBlockRows];
    for (int i = 0; i < kBlockRows; ++i) {
      v[i] = V(0);
      if (cur[i] < end[i] && col[cur[i]] == c_min) v[i] = val[cur[i]++];
    }
    emit(c_min, v);
  }
}

template <typename V>
__global__ void k_block_count(const long long* __restrict__ row_ptr, const int* __restrict__ col,
                              const V* __restrict__ val, long long n_rows, long long n_blocks,
                              unsigned int* __restrict__ count) {
  for (long long b = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; b < n_blocks;
       b += static_cast<long long>(gridDim.x) * blockDim.x) {
    unsigned int n = 0;
    merge_block(row_ptr, col, val, n_rows, b, [&](int, const V*) { ++n; });
    count[b] = n;
  }
}

template <typename V>
__global__ void k_block_fill(const long long* __restrict__ row_ptr, const int* __restrict__ col,
                             const V* __restrict__ val, long long n_rows, long long n_blocks,
                             const long long* __restrict__ blk_ptr, int* __restrict__ blk_col,
                             V* __restrict__ blk_val) {
  for (long long b = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; b < n_blocks;
       b += static_cast<long long>(gridDim.x) * blockDim.x) {
    long long u = blk_ptr[b];
    merge_block(row_ptr, col, val, n_rows, b, [&](int c, const V* v) {
      blk_col[u] = c;
      V* dst = blk_val + u * kBlockRows;
      for (int i = 0; i < kBlockRows; ++i) dst[i] = v[i];
      ++u;
    });
  }
}

// Merge of the union columns of super-block sb's row blocks: calls emit(column, cur,
// end) once per distinct column, in ascending order; emit advances every block's
// cursor past its entries of that column (a repeated (row, column) pair gives a block
// two entries of one column).
template <typename Emit>
__device__ void merge_super(const long long* __restrict__ blk_ptr, const int* __restrict__ blk_col,
                            long long n_blocks, long long sb, Emit emit) {
  long long cur[kSuperBlocks], end[kSuperBlocks];
#pragma unroll
  for (int g = 0; g < kSuperBlocks; ++g) {
    const long long b = sb * kSuperBlocks + g;
    cur[g] = b < n_blocks ? blk_ptr[b] : 0;
    end[g] = b < n_blocks ? blk_ptr[b + 1] : 0;
  }
  for (;;) {
    int c_min = 0x7fffffff;
#pragma unroll
    for (int g = 0; g < kSuperBlocks; ++g)
      if (cur[g] < end[g]) c_min = min(c_min, blk_col[cur[g]]);
    if (c_min == 0x7fffffff) return;
    emit(c_min, cur, end);
  }
}

__global__ void k_super_count(const long long* __restrict__ blk_ptr, const int* __restrict__ blk_col,
                              long long n_blocks, long long n_super, unsigned int* __restrict__ count) {
  for (long long sb = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; sb < n_super;
       sb += static_cast<long long>(gridDim.x) * blockDim.x) {
    unsigned int n = 0;
    merge_super(blk_ptr, blk_col, n_blocks, sb, [&](int c, long long* cur, const long long* end) {
#pragma unroll
      for (int g = 0; g < kSuperBlocks; ++g)
        while (cur[g] < end[g] && blk_col[cur[g]] == c) ++cur[g];
      ++n;
    });
    count[sb] = n;
  }
}

// The super-block's columns, and for each of its blocks' entries the word
// (position of the entry's column among them << 8) | (bit i set: row i's value is not 0).
template <typename V>
__global__ void k_super_fill(const long long* __restrict__ blk_ptr, const int* __restrict__ blk_col,
                             const V* __restrict__ blk_val, long long n_blocks, long long n_super,
                             const long long* __restrict__ sb_ptr, int* __restrict__ sb_col,
                             int* __restrict__ blk_word) {
  for (long long sb = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; sb < n_super;
       sb += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long o = sb_ptr[sb];
    int k = 0;
    merge_super(blk_ptr, blk_col, n_blocks, sb, [&](int c, long long* cur, const long long* end) {
      sb_col[o + k] = c;
#pragma unroll
      for (int g = 0; g < kSuperBlocks; ++g)
        while (cur[g] < end[g] && blk_col[cur[g]] == c) {
          int mask = 0;
          for (int i = 0; i < kBlockRows; ++i)
            if (blk_val[cur[g] * kBlockRows + i] != V(0)) mask |= 1 << i;
          blk_word[cur[g]] = static_cast<int>((static_cast<unsigned int>(k) << 8) | static_cast<unsigned int>(mask));
          ++cur[g];
        }
      ++k;
    });
  }
}

// ---------------------------------------------------------------------------------------
// Product kernels.

// acc + v * w. float64: product rounded, then the add (the reference's two numpy
// operations); float32: one fused multiply-add.
__device__ __forceinline__ double madd(double acc, double v, double w) {
#pragma clang fp contract(off)
  const double p = v * w;
  return acc + p;
}
__device__ __forceinline__ float madd(float acc, float v, float w) { return fmaf(v, w, acc); }

// Workgroup index remapped so that each XCD (workgroups go round-robin over the 8)
// takes one contiguous range of row blocks: its L2 then holds the weight rows of one
// region of the grid.
__device__ __forceinline__ long long xcd_local_group() {
  const long long g = gridDim.x, bid = blockIdx.x;
  const long long xcd = bid & 7, idx = bid >> 3;
  long long start = 0;
  for (long long x = 0; x < xcd; ++x) start += (g - x + 7) >> 3;
  return start + idx;
}

// Many vectors: a wave per row block, lanes over vectors, 64 * KL vectors per pass.
// The pass is the outer grid index (groups_per_pass workgroups each, a multiple of
// 8): the whole grid finishes one slice of W (n_cols x 64 KL, 2.3 MB for 1125
// columns in float64) before the next, so the slice in use stays L2-resident, at the
// price of reading the row blocks once per pass.
template <typename T, typename V, int KL>
__global__ void __launch_bounds__(256) k_spmm_wide(const long long* __restrict__ blk_ptr,
                                                   const int* __restrict__ blk_col,
                                                   const V* __restrict__ blk_val,
                                                   const T* __restrict__ w, int n_w, long long r0,
                                                   long long r1, unsigned int groups_per_pass,
                                                   T* __restrict__ out) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  // The wave index is uniform; saying so keeps the block's columns and values in
  // scalar registers and its empty-entry tests on the scalar unit.
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const unsigned int pass = blockIdx.x / groups_per_pass, local = blockIdx.x % groups_per_pass;
  // Within a pass, each XCD (workgroups go round-robin over the 8) takes one
  // contiguous range of row blocks.
  const long long group = static_cast<long long>(local & 7u) * (groups_per_pass >> 3) + (local >> 3);
  const long long b = r0 / kBlockRows + group * 4 + wave;
  if (b * kBlockRows >= r1) return;
  const long long u0 = blk_ptr[b], u1 = blk_ptr[b + 1];
  {
    const int k0 = static_cast<int>(pass) * 64 * KL;
    T acc[kBlockRows][KL];
#pragma unroll
    for (int i = 0; i < kBlockRows; ++i)
#pragma unroll
      for (int q = 0; q < KL; ++q) acc[i][q] = T(0);
    for (long long u = u0; u < u1; ++u) {
      const long long j = blk_col[u];
      V v[kBlockRows];
#pragma unroll
      for (int i = 0; i < kBlockRows; ++i) v[i] = blk_val[u * kBlockRows + i];
      T wv[KL];
#pragma unroll
      for (int q = 0; q < KL; ++q) {
        const int k = k0 + q * 64 + lane;
        wv[q] = k < n_w ? w[j * n_w + k] : T(0);
      }
#pragma unroll
      for (int i = 0; i < kBlockRows; ++i)
        if (v[i] != V(0)) {
#pragma unroll
          for (int q = 0; q < KL; ++q) acc[i][q] = madd(acc[i][q], static_cast<T>(v[i]), wv[q]);
        }
    }
#pragma unroll
    for (int i = 0; i < kBlockRows; ++i) {
      const long long r = b * kBlockRows + i;
      if (r < r0 || r >= r1) continue;
      T* dst = out + (r - r0) * n_w;
#pragma unroll
      for (int q = 0; q < KL; ++q) {
        const int k = k0 + q * 64 + lane;
        if (k < n_w) dst[k] = acc[i][q];
      }
    }
  }
}

// Word 3 of a raw (untyped, stride 0) buffer resource on gfx9-family GPUs.
constexpr unsigned int kRawBufferWord3 = 0x00020000u;

__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r, unsigned int voff, unsigned int soff,
                                             float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, unsigned int voff, unsigned int soff,
                                              double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// A value's bits as loaded (wave-uniform, so scalar loads), the test for a value other
// than +0 / -0 made on those bits on the scalar unit (gfx950 has no scalar float
// compare), and the value itself.
template <typename V> struct ValueBits;
template <> struct ValueBits<float> {
  using U = uint32_t;
  __device__ static bool nonzero(U b) { return (b << 1) != 0u; }
  __device__ static float value(U b) { return __builtin_bit_cast(float, b); }
};
template <> struct ValueBits<double> {
  using U = uint64_t;
  __device__ static bool nonzero(U b) { return (b << 1) != 0ull; }
  __device__ static double value(U b) { return __builtin_bit_cast(double, b); }
};

// k_spmm_wide for at most 2^24 columns and weights of at most 2 GiB, built so that a
// wave never waits on memory latency in the entry loop:
//  * the row blocks are read with vector loads, 32 entries at a time and one chunk
//    ahead: lane 2c + h holds entry c's column and its rows of half h. Once a chunk has
//    arrived, each lane pair forms the word (column << 8) | mask, where bit i of the
//    mask says row i's value is not zero. Per entry, one readlane gives that word to
//    the scalar unit, and one readlane per non-zero row gives the value. (Scalar loads
//    of the row blocks, which stream from HBM, stalled every wave for a full memory
//    round trip per entry: all scalar loads retire under one counter, so they cannot
//    be issued ahead.)
//  * the weight rows are read through a buffer resource: the row's byte offset is one
//    scalar multiply per entry and the lane's offsets within the row are fixed for the
//    pass, so an entry's loads need no vector address arithmetic and no bounds
//    branches. Lanes past n_w read the next row's first elements, or 0 past the
//    buffer's end (the resource's range check), into accumulators that are never
//    stored. They run one entry ahead of the multiply-adds, in two alternating register
//    sets (no copies between them, so the wait before an entry's multiply-adds covers
//    only that entry's loads); the prefetch past the block's last entry re-reads it.
// Each row's additions keep ascending column order, as in k_spmm_wide.
constexpr int kChunk = 32;  // row-block entries per vector-loaded chunk

template <typename T, typename V, int KL>
__global__ void __launch_bounds__(256) k_spmm_wide_rsrc(const long long* __restrict__ blk_ptr,
                                                        const int* __restrict__ blk_col,
                                                        const V* __restrict__ blk_val,
                                                        const T* __restrict__ w, unsigned int w_bytes, int n_w,
                                                        long long r0, long long r1, unsigned int groups_per_pass,
                                                        T* __restrict__ out) {
  using B = ValueBits<V>;
  using U = typename B::U;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const unsigned int pass = blockIdx.x / groups_per_pass, local = blockIdx.x % groups_per_pass;
  const long long group = static_cast<long long>(local & 7u) * (groups_per_pass >> 3) + (local >> 3);
  const long long b = r0 / kBlockRows + group * 4 + wave;
  if (b * kBlockRows >= r1) return;
  const long long u0 = blk_ptr[b];
  const int n_e = static_cast<int>(blk_ptr[b + 1] - u0);  // a block's union holds at most n_cols < 2^24
  const int* __restrict__ cols = blk_col + u0;
  const U* __restrict__ bits = reinterpret_cast<const U*>(blk_val) + u0 * kBlockRows;
  const int k0 = static_cast<int>(pass) * 64 * KL;
  const __amdgpu_buffer_rsrc_t rw =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(w), 0, static_cast<int>(w_bytes), kRawBufferWord3);
  const unsigned int row_bytes = static_cast<unsigned int>(n_w) * static_cast<unsigned int>(sizeof(T));
  const unsigned int voff = static_cast<unsigned int>(k0 + lane) * static_cast<unsigned int>(sizeof(T));
  const int half = lane & 1;
  T acc[kBlockRows][KL];
#pragma unroll
  for (int i = 0; i < kBlockRows; ++i)
#pragma unroll
    for (int q = 0; q < KL; ++q) acc[i][q] = T(0);

  // This lane's part of the chunk starting at entry c0 (past the last entry: the last).
  auto load_chunk = [&](int c0, int& col, U (&v4)[kHalfRows]) {
    const int e = min(c0 + (lane >> 1), n_e - 1);
    col = cols[e];
    const U* p = bits + static_cast<long long>(e) * kBlockRows + half * kHalfRows;
#pragma unroll
    for (int i = 0; i < kHalfRows; ++i) v4[i] = p[i];
  };
  // (column << 8) | mask of the lane pair's entry, on both lanes of the pair.
  auto chunk_words = [&](int col, const U (&v4)[kHalfRows]) {
    int nib = 0;
#pragma unroll
    for (int i = 0; i < kHalfRows; ++i) nib |= B::nonzero(v4[i]) ? 1 << i : 0;
    const int other = __shfl_xor(nib, 1);
    return (col << 8) | (half ? other | nib << kHalfRows : nib | other << kHalfRows);
  };
  auto fetch = [&](int word, T (&wts)[KL]) {
    const unsigned int soff = (static_cast<unsigned int>(word) >> 8) * row_bytes;
    // An empty asm makes the offset opaque here, so that vo + constant is not hoisted
    // into one register per load but folded into each load's immediate offset.
    unsigned int vo = voff;
    __asm__ volatile("" : "+v"(vo));
#pragma unroll
    for (int q = 0; q < KL; ++q) wts[q] = buffer_load(rw, vo + static_cast<unsigned int>(q * 64 * sizeof(T)), soff, T(0));
  };
  auto value = [&](const U (&v4)[kHalfRows], int i, int t) -> T {
    const int src = 2 * t + i / kHalfRows;
    if constexpr (sizeof(U) == 4) {
      return static_cast<T>(B::value(static_cast<U>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows]), src))));
    } else {
      const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows]), src));
      const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows] >> 32), src));
      return static_cast<T>(B::value((static_cast<U>(hi) << 32) | lo));
    }
  };
  auto accumulate = [&](int word, const U (&v4)[kHalfRows], int t, const T (&wts)[KL]) {
#pragma unroll
    for (int i = 0; i < kBlockRows; ++i)
      if (word >> i & 1) {
        const T v = value(v4, i, t);
#pragma unroll
        for (int q = 0; q < KL; ++q) acc[i][q] = madd(acc[i][q], v, wts[q]);
      }
  };

  if (n_e > 0) {
    int col_n;
    U v4[kHalfRows], v4_n[kHalfRows];
    load_chunk(0, col_n, v4_n);
    int words_n = chunk_words(col_n, v4_n);
    T wa[KL], wb[KL];
    fetch(__builtin_amdgcn_readlane(words_n, 0), wa);
    for (int c0 = 0; c0 < n_e; c0 += kChunk) {
      const int words = words_n;
#pragma unroll
      for (int i = 0; i < kHalfRows; ++i) v4[i] = v4_n[i];
      load_chunk(c0 + kChunk, col_n, v4_n);
      const int ce = min(kChunk, n_e - c0);
      const bool more = c0 + kChunk < n_e;
      // Entry t: the next entry's weights into `nxt` (past the chunk: the next chunk's
      // first entry, whose words are formed here; past the block: entry t again), then
      // entry t's multiply-adds from `cur`. Every path issues the same loads, so the
      // wait before the multiply-adds is exact.
      auto step = [&](int t, const T (&cur)[KL], T (&nxt)[KL]) {
        int next_word;
        if (t + 1 < ce) {
          next_word = __builtin_amdgcn_readlane(words, 2 * (t + 1));
        } else if (more) {
          words_n = chunk_words(col_n, v4_n);
          next_word = __builtin_amdgcn_readlane(words_n, 0);
        } else {
          next_word = __builtin_amdgcn_readlane(words, 2 * t);
        }
        fetch(next_word, nxt);
        accumulate(__builtin_amdgcn_readlane(words, 2 * t), v4, t, cur);
      };
#pragma unroll
      for (int t = 0; t < kChunk; t += 2) {
        step(t, wa, wb);
        if (t + 1 >= ce) break;
        step(t + 1, wb, wa);
        if (t + 2 >= ce) break;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kBlockRows; ++i) {
    const long long r = b * kBlockRows + i;
    if (r < r0 || r >= r1) continue;
    T* dst = out + (r - r0) * n_w;
#pragma unroll
    for (int q = 0; q < KL; ++q) {
      const int k = k0 + q * 64 + lane;
      if (k < n_w) dst[k] = acc[i][q];
    }
  }
}

// The product for many vectors with the weights staged in LDS: a workgroup takes one
// super-block (16 consecutive row blocks, a wave per block; 128 voxels) and one vector
// pass. It walks the distinct columns of its blocks in ascending order, a 64 KiB tile
// of weight rows at a time: the rows W[j][k0 : k0 + 64 KL] of a tile come from L2 once
// per workgroup, straight into LDS (buffer loads with the LDS destination, 1 KiB per
// wave load), and every wave then reads from LDS the rows its own entries need.
// Tiles are double-buffered: the next tile's loads are issued right after the barrier
// that publishes the current one. The columns being staged come from LDS too (up to
// 1024 at a time), so that the loop makes no scalar loads: those retire under the same
// counter as LDS reads and would serialize them.
// Each wave's entries arrive in chunks of 32 with vector loads (lane 2c + h: entry c's
// word and the rows of half h); readlane hands the word and each non-zero row's value to
// the scalar unit. Three consecutive chunks sit in registers and a tile reads only
// chunks requested before its barrier (which waits for them), so no chunk load is
// waited for inside a tile. A tile's entries are a run of the chunk (the words
// ascend), counted with a ballot. LDS reads of the weights run one entry ahead of the
// multiply-adds in two alternating register sets, as in k_spmm_wide_rsrc.
// Each row's additions keep ascending column order: tiles ascend, and so do the
// entries within a tile.
constexpr int kTileBytes = 65536;  // one LDS tile buffer: 32 weight rows of 2 KiB, 16 of 4 KiB

template <typename T, typename V, int KL>
__global__ void __launch_bounds__(64 * kSuperBlocks) k_spmm_tiled(
    const long long* __restrict__ blk_ptr, const int* __restrict__ blk_word, const V* __restrict__ blk_val,
    const long long* __restrict__ sb_ptr, const int* __restrict__ sb_col, long long n_blocks,
    const T* __restrict__ w, unsigned int w_bytes, int n_w, long long r0, long long r1,
    unsigned int groups_per_pass, T* __restrict__ out) {
  using B = ValueBits<V>;
  using U = typename B::U;
  constexpr int kSlotBytes = 64 * KL * static_cast<int>(sizeof(T));
  static_assert(kSlotBytes % 1024 == 0, "a slot is a whole number of 1 KiB wave loads");
  constexpr int kSlotLoads = kSlotBytes / 1024;
  constexpr int kTileSlots = kTileBytes / kSlotBytes;  // weight rows per tile
  constexpr int kThreads = 64 * kSuperBlocks;
  constexpr int kChunkEntries = 32;
  __shared__ __align__(16) unsigned char tiles[2][kTileSlots * kSlotBytes];
  __shared__ int stage[kThreads];

  const int lane = static_cast<int>(threadIdx.x & 63);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const unsigned int pass = blockIdx.x / groups_per_pass, local = blockIdx.x % groups_per_pass;
  const long long group = static_cast<long long>(local & 7u) * (groups_per_pass >> 3) + (local >> 3);
  const long long sb = r0 / (kBlockRows * kSuperBlocks) + group;
  const long long sb_end = ((r1 + kBlockRows - 1) / kBlockRows + kSuperBlocks - 1) / kSuperBlocks;
  if (sb >= sb_end) return;  // the whole workgroup
  const long long s_lo = sb_ptr[sb];
  const int n_s = static_cast<int>(sb_ptr[sb + 1] - s_lo);
  const long long b = sb * kSuperBlocks + wave;
  const bool active = b < n_blocks && b * kBlockRows < r1 && (b + 1) * kBlockRows > r0;
  const long long u0 = active ? blk_ptr[b] : 0;
  const int n_e = active ? static_cast<int>(blk_ptr[b + 1] - u0) : 0;
  const int* __restrict__ words_g = blk_word + u0;
  const U* __restrict__ bits = reinterpret_cast<const U*>(blk_val) + u0 * kBlockRows;
  const int k0 = static_cast<int>(pass) * 64 * KL;
  const __amdgpu_buffer_rsrc_t rw =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(w), 0, static_cast<int>(w_bytes), kRawBufferWord3);
  const unsigned int row_bytes = static_cast<unsigned int>(n_w) * static_cast<unsigned int>(sizeof(T));
  const int half = lane & 1, pair = lane >> 1;
  T acc[kBlockRows][KL];
#pragma unroll
  for (int i = 0; i < kBlockRows; ++i)
#pragma unroll
    for (int q = 0; q < KL; ++q) acc[i][q] = T(0);

  auto fill_stage = [&](int c_base) {
    const int c = c_base + static_cast<int>(threadIdx.x);
    stage[threadIdx.x] = c < n_s ? sb_col[s_lo + c] : 0;
  };
  // Tile t's weight rows into buffer buf: wave loads j = wave, wave + 16, ... of the
  // tile's kTileSlots * kSlotLoads.
  auto issue_tile = [&](int t, int buf) {
#pragma unroll
    for (int r = 0; r < (kTileSlots * kSlotLoads + kSuperBlocks - 1) / kSuperBlocks; ++r) {
      const int j = wave + r * kSuperBlocks;
      const int s = j / kSlotLoads, h = j % kSlotLoads;
      if (j >= kTileSlots * kSlotLoads || t + s >= n_s) break;
      const int col = __builtin_amdgcn_readfirstlane(stage[(t + s) % kThreads]);
      auto* dst = reinterpret_cast<__attribute__((address_space(3))) void*>(
          reinterpret_cast<uintptr_t>(&tiles[buf][s * kSlotBytes + h * 1024]));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(
          rw, dst, 16, static_cast<int>(k0 * sizeof(T)) + h * 1024 + lane * 16,
          static_cast<int>(static_cast<unsigned int>(col) * row_bytes), 0, 0);
    }
  };
  // Entries c0 .. c0 + 31 of this wave (past the last: the last). Only when n_e > 0.
  auto load_chunk = [&](int c0, int& word, U (&v4)[kHalfRows]) {
    const int e = min(c0 + pair, n_e - 1);
    word = words_g[e];
    const U* p = bits + static_cast<long long>(e) * kBlockRows + half * kHalfRows;
#pragma unroll
    for (int i = 0; i < kHalfRows; ++i) v4[i] = p[i];
  };
  // Entries of tile t in the chunk that starts at entry c0, from its entry o on (they
  // form a run: the words ascend).
  auto count_in_tile = [&](int word, int c0, int o, int t) {
    const bool in = half == 0 && pair >= o && c0 + pair < n_e &&
                    static_cast<int>(static_cast<unsigned int>(word) >> 8) < t + kTileSlots;
    return __popcll(__ballot(in));
  };
  auto value = [&](const U (&v4)[kHalfRows], int i, int c) -> T {
    const int src = 2 * c + i / kHalfRows;
    if constexpr (sizeof(U) == 4) {
      return static_cast<T>(B::value(static_cast<U>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows]), src))));
    } else {
      const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows]), src));
      const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v4[i % kHalfRows] >> 32), src));
      return static_cast<T>(B::value((static_cast<U>(hi) << 32) | lo));
    }
  };
  // Entries o .. o + cnt - 1 of a chunk against tile t in buffer buf.
  auto process = [&](int word, const U (&v4)[kHalfRows], int o, int cnt, int t, int buf) {
    const T* tile = reinterpret_cast<const T*>(tiles[buf]);
    // The weights of entry c into wts; returns the entry's word (one readlane serves
    // both the LDS address and, an entry later, the row tests).
    auto read_w = [&](int c, T (&wts)[KL]) {
      const int wd = __builtin_amdgcn_readlane(word, 2 * c);
      const int slot = static_cast<int>(static_cast<unsigned int>(wd) >> 8) - t;
      const T* src = tile + slot * (64 * KL) + lane;
#pragma unroll
      for (int q = 0; q < KL; ++q) wts[q] = src[q * 64];
      return wd;
    };
    auto accumulate = [&](int c, int wd, const T (&wts)[KL]) {
#pragma unroll
      for (int i = 0; i < kBlockRows; ++i)
        if (wd >> i & 1) {
          const T v = value(v4, i, c);
#pragma unroll
          for (int q = 0; q < KL; ++q) acc[i][q] = madd(acc[i][q], v, wts[q]);
        }
    };
    T wa[KL], wb[KL];
    const int last = o + cnt - 1;
    int da = read_w(o, wa);
    for (int c = o;; c += 2) {
      const int db = read_w(min(c + 1, last), wb);
      accumulate(c, da, wa);
      if (c + 1 > last) break;
      da = read_w(min(c + 2, last), wa);
      accumulate(c + 1, db, wb);
      if (c + 2 > last) break;
    }
  };

  // Chunks k, k + 1 and k + 2 of this wave's entries (32 each) sit in registers. A tile
  // reads only chunks whose loads were issued before its barrier, which waits for
  // them: chunk k + 2 is requested when chunk k becomes current, and a tile may run on
  // from chunk k into chunk k + 1. Only a tile with more entries than that (repeated
  // entries) reloads the window with waits.
  const int n_tiles = (n_s + kTileSlots - 1) / kTileSlots;
  int w0 = 0, w1 = 0, w2 = 0;
  U v0[kHalfRows] = {}, v1[kHalfRows] = {}, v2[kHalfRows] = {};
  int k = 0, e = 0;  // current chunk; this wave's next entry
  auto load_window = [&]() {
    if (kChunkEntries * k < n_e) load_chunk(kChunkEntries * k, w0, v0);
    if (kChunkEntries * (k + 1) < n_e) load_chunk(kChunkEntries * (k + 1), w1, v1);
    if (kChunkEntries * (k + 2) < n_e) load_chunk(kChunkEntries * (k + 2), w2, v2);
  };
  fill_stage(0);
  load_window();
  __syncthreads();
  issue_tile(0, 0);
  for (int ti = 0; ti < n_tiles; ++ti) {
    const int t = ti * kTileSlots, buf = ti & 1;
    // This wave's tile loads (and chunk loads) have landed; so have every wave's, and
    // every wave is done with the other buffer.
    __syncthreads();
    if (e >= kChunkEntries * (k + 1)) {  // chunk k is used up: move the window on
      w0 = w1;
      w1 = w2;
#pragma unroll
      for (int i = 0; i < kHalfRows; ++i) {
        v0[i] = v1[i];
        v1[i] = v2[i];
      }
      ++k;
      if (kChunkEntries * (k + 2) < n_e) load_chunk(kChunkEntries * (k + 2), w2, v2);
    }
    if (ti + 1 < n_tiles) {
      if ((t + kTileSlots) % kThreads == 0) {
        fill_stage(t + kTileSlots);
        __syncthreads();
      }
      issue_tile(t + kTileSlots, buf ^ 1);
    }
    if (e < n_e) {
      const int o = e - kChunkEntries * k;
      const int cnt = count_in_tile(w0, kChunkEntries * k, o, t);
      if (cnt > 0) process(w0, v0, o, cnt, t, buf);
      e += cnt;
      if (o + cnt == kChunkEntries && e < n_e) {  // the tile may run on into chunk k + 1
        const int cnt1 = count_in_tile(w1, kChunkEntries * (k + 1), 0, t);
        if (cnt1 > 0) process(w1, v1, 0, cnt1, t, buf);
        e += cnt1;
        // Chunk k + 1 used up as well: the tile's remaining entries one reloaded chunk
        // at a time, then the window reloaded at the new position.
        if (cnt1 == kChunkEntries && e < n_e) {
          for (;;) {
            k = e / kChunkEntries;
            load_chunk(kChunkEntries * k, w0, v0);
            const int c = count_in_tile(w0, kChunkEntries * k, 0, t);
            if (c > 0) process(w0, v0, 0, c, t, buf);
            e += c;
            if (c < kChunkEntries || e >= n_e) break;
          }
          k = e / kChunkEntries;
          load_window();
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kBlockRows; ++i) {
    const long long r = b * kBlockRows + i;
    if (r < r0 || r >= r1) continue;
    T* dst = out + (r - r0) * n_w;
#pragma unroll
    for (int q = 0; q < KL; ++q) {
      const int k = k0 + q * 64 + lane;
      if (k < n_w) dst[k] = acc[i][q];
    }
  }
}

// Few vectors (n_w <= 8): 8 lanes per row block, a lane per row.
template <typename T, typename V>
__global__ void __launch_bounds__(256) k_spmm_narrow(const long long* __restrict__ blk_ptr,
                                                     const int* __restrict__ blk_col,
                                                     const V* __restrict__ blk_val,
                                                     const T* __restrict__ w, int n_w, long long r0,
                                                     long long r1, T* __restrict__ out) {
  const int i = static_cast<int>(threadIdx.x & (kBlockRows - 1));
  const long long b = r0 / kBlockRows + xcd_local_group() * (256 / kBlockRows) + (threadIdx.x / kBlockRows);
  if (b * kBlockRows >= r1) return;
  T acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = T(0);
  const long long u1 = blk_ptr[b + 1];
  for (long long u = blk_ptr[b]; u < u1; ++u) {
    const V v = blk_val[u * kBlockRows + i];
    if (v == V(0)) continue;
    const T* wr = w + static_cast<long long>(blk_col[u]) * n_w;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < n_w) acc[q] = madd(acc[q], static_cast<T>(v), wr[q]);
  }
  const long long r = b * kBlockRows + i;
  if (r < r0 || r >= r1) return;
  T* dst = out + (r - r0) * n_w;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < n_w) dst[q] = acc[q];
}

// Launch shape of a product over rows [r0, r1): workgroups per vector pass, passes,
// and vectors per lane (0 for the narrow kernel, whose passes count is 1).
struct ProductShape {
  long long groups_per_pass, passes;
  int kl;
};

// Vectors per lane: the fewest passes whose accumulators (rows x vectors per lane)
// stay within 64 registers, so that 4 waves fit per SIMD: with 8-row blocks, 8 per
// lane in float32 and 4 in float64 (measured with the per-wave kernel on the full plan
// with 1000 vectors: float32 117 ms against 168 ms at 4 per lane and 134 ms at 16;
// float64 at 8 per lane 271 against 241 ms).
template <typename T>
constexpr int kMaxVectorsPerLane = 64 / (kBlockRows * static_cast<int>(sizeof(T) / 4));

ProductShape product_shape(int n_w, long long r0, long long r1, bool f32) {
  const long long blocks = (r1 + kBlockRows - 1) / kBlockRows - r0 / kBlockRows;
  if (n_w <= 8) return {(blocks + 256 / kBlockRows - 1) / (256 / kBlockRows), 1, 0};
  const int kl_max = f32 ? kMaxVectorsPerLane<float> : kMaxVectorsPerLane<double>;
  int kl = 1;
  while (kl < kl_max && 64 * kl < n_w) kl *= 2;
  return {((blocks + 3) / 4 + 7) / 8 * 8, (n_w + 64 * kl - 1) / (64 * kl), kl};
}

// Launches Kernel<KL> for the run-time kl (a power of two up to the type's maximum).
template <typename T, typename F>
void with_kl(int kl, F launch) {
  if (kl == 1) launch(std::integral_constant<int, 1>());
  if (kl == 2) launch(std::integral_constant<int, 2>());
  if (kl == 4) launch(std::integral_constant<int, 4>());
  if constexpr (kMaxVectorsPerLane<T> >= 8)
    if (kl == 8) launch(std::integral_constant<int, 8>());
  if constexpr (kMaxVectorsPerLane<T> >= 16)
    if (kl == 16) launch(std::integral_constant<int, 16>());
}

template <typename T, typename V>
void launch_product(hsp_matrix* m, const T* w, int n_w, long long r0, long long r1, T* out) {
  const V* bv = static_cast<const V*>(m->blk_val);
  const ProductShape shape = product_shape(n_w, r0, r1, sizeof(T) == 4);
  if (shape.kl == 0) {
    hipLaunchKernelGGL((k_spmm_narrow<T, V>), dim3(static_cast<unsigned int>(shape.groups_per_pass)), dim3(256), 0,
                       m->stream, m->blk_ptr, m->blk_col, bv, w, n_w, r0, r1, out);
    return;
  }
  const unsigned int gpp = static_cast<unsigned int>(shape.groups_per_pass);
  const int kl = shape.kl;
  const dim3 grid(static_cast<unsigned int>(shape.groups_per_pass * shape.passes));
  const unsigned long long w_bytes = static_cast<unsigned long long>(m->n_cols) * n_w * sizeof(T);
  const int slot_bytes = 64 * kl * static_cast<int>(sizeof(T));
  if (w_bytes <= 0x7fffffffull && m->sb_ptr && slot_bytes >= 1024) {
    const long long sb_lo = r0 / (kBlockRows * kSuperBlocks);
    const long long sb_hi = ((r1 + kBlockRows - 1) / kBlockRows + kSuperBlocks - 1) / kSuperBlocks;
    const long long groups = (sb_hi - sb_lo + 7) / 8 * 8;
    const unsigned int wb = static_cast<unsigned int>(w_bytes);
    const dim3 tgrid(static_cast<unsigned int>(groups * shape.passes));
    const unsigned int g = static_cast<unsigned int>(groups);
    with_kl<T>(kl, [&](auto c) {
      constexpr int K = decltype(c)::value;
      if constexpr (64 * K * sizeof(T) >= 1024)
        hipLaunchKernelGGL((k_spmm_tiled<T, V, K>), tgrid, dim3(64 * kSuperBlocks), 0, m->stream, m->blk_ptr,
                           m->blk_word, bv, m->sb_ptr, m->sb_col, m->n_blocks, w, wb, n_w, r0, r1, g, out);
    });
    return;
  }
  if (w_bytes <= 0x7fffffffull && m->n_cols <= (1ll << 24)) {
    const unsigned int wb = static_cast<unsigned int>(w_bytes);
    with_kl<T>(kl, [&](auto c) {
      hipLaunchKernelGGL((k_spmm_wide_rsrc<T, V, decltype(c)::value>), grid, dim3(256), 0, m->stream, m->blk_ptr,
                         m->blk_col, bv, w, wb, n_w, r0, r1, gpp, out);
    });
    return;
  }
  with_kl<T>(kl, [&](auto c) {
    hipLaunchKernelGGL((k_spmm_wide<T, V, decltype(c)::value>), grid, dim3(256), 0, m->stream, m->blk_ptr,
                       m->blk_col, bv, w, n_w, r0, r1, gpp, out);
  });
}

// Exclusive scan of n device counts into out[0..n]; returns the total.
int scan_counts(hsp_matrix* m, const unsigned int* count, long long n, long long* out, long long* total) {
  const long long tiles = std::max((n + kScanTile - 1) / kScanTile, 1ll);
  long long* d_tile = nullptr;
  HSP_HIP(m, hipMalloc(&d_tile, sizeof(long long) * tiles));
  hipLaunchKernelGGL(k_tile_sums, dim3(static_cast<unsigned int>(tiles)), dim3(256), 0, m->stream, count, n,
                     d_tile);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(64), 0, m->stream, d_tile, tiles, out + n);
  hipLaunchKernelGGL(k_tile_scan, dim3(static_cast<unsigned int>(tiles)), dim3(256), 0, m->stream, count, n,
                     d_tile, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(total, out + n, sizeof(long long), hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  (void)hipFree(d_tile);
  HSP_HIP(m, e);
  return HSP_OK;
}

// Set-up after the matrix's arrays are allocated; temporaries freed on every path.
template <typename V>
int build_forms(hsp_matrix* m, const int64_t* indptr, const int64_t* indices, const V* values) {
  V* val = static_cast<V*>(m->val);
  const long long n_rows = m->n_rows, n_cols = m->n_cols, nnz = m->nnz;
  const size_t nz = static_cast<size_t>(std::max(nnz, 1ll));
  long long *d_rows = nullptr, *d_src = nullptr;
  V* d_values = nullptr;
  unsigned int* d_count = nullptr;
  auto release = [&]() {
    for (void* p : {static_cast<void*>(d_rows), static_cast<void*>(d_src), static_cast<void*>(d_values),
                    static_cast<void*>(d_count)})
      if (p) (void)hipFree(p);
    d_rows = d_src = nullptr;
    d_values = nullptr;
    d_count = nullptr;
  };
  auto run = [&]() -> int {
    HSP_HIP(m, hipMalloc(&d_rows, sizeof(long long) * nz));
    HSP_HIP(m, hipMalloc(&d_src, sizeof(long long) * nz));
    HSP_HIP(m, hipMalloc(&d_values, sizeof(V) * nz));
    HSP_HIP(m, hipMalloc(&d_count, sizeof(unsigned int) * std::max(n_rows, m->n_blocks)));
    if (nnz) {
      HSP_HIP(m, hipMemcpy(d_rows, indices, sizeof(long long) * nnz, hipMemcpyHostToDevice));
      HSP_HIP(m, hipMemcpy(d_values, values, sizeof(V) * nnz, hipMemcpyHostToDevice));
    }
    HSP_HIP(m, hipMemsetAsync(d_count, 0, sizeof(unsigned int) * n_rows, m->stream));
    hipLaunchKernelGGL(k_count_rows, dim3(grid_for(nnz)), dim3(256), 0, m->stream, d_rows, nnz, d_count);
    HSP_HIP(m, hipGetLastError());
    long long total = 0;
    if (int rc = scan_counts(m, d_count, n_rows, m->row_ptr, &total)) return rc;
    if (total != nnz) return fail(m, HSP_EHIP, "row count scan disagrees with nnz");
    HSP_HIP(m, hipMemsetAsync(d_count, 0, sizeof(unsigned int) * n_rows, m->stream));  // cursors
    for (long long c = 0; c < n_cols; ++c) {
      const long long lo = indptr[c], hi = indptr[c + 1];
      if (hi == lo) continue;
      hipLaunchKernelGGL(k_scatter_column<V>, dim3(static_cast<unsigned int>((hi - lo + 255) / 256)), dim3(256),
                         0, m->stream, lo, hi, static_cast<int>(c), d_rows, d_values, m->row_ptr, d_count, m->col,
                         val, d_src);
    }
    hipLaunchKernelGGL(k_sort_rows<V>, dim3(grid_for(n_rows)), dim3(256), 0, m->stream, m->row_ptr, n_rows, m->col,
                       val, d_src);
    HSP_HIP(m, hipGetLastError());
    HSP_HIP(m, hipMemsetAsync(d_count, 0, sizeof(unsigned int) * m->n_blocks, m->stream));
    hipLaunchKernelGGL(k_block_count<V>, dim3(grid_for(m->n_blocks)), dim3(256), 0, m->stream, m->row_ptr,
                       m->col, val, n_rows, m->n_blocks, d_count);
    HSP_HIP(m, hipGetLastError());
    if (int rc = scan_counts(m, d_count, m->n_blocks, m->blk_ptr, &m->n_block_entries)) return rc;
    const size_t ne = static_cast<size_t>(std::max(m->n_block_entries, 1ll));
    HSP_HIP(m, hipMalloc(&m->blk_col, sizeof(int) * ne));
    HSP_HIP(m, hipMalloc(&m->blk_val, sizeof(V) * kBlockRows * ne));
    hipLaunchKernelGGL(k_block_fill<V>, dim3(grid_for(m->n_blocks)), dim3(256), 0, m->stream, m->row_ptr,
                       m->col, val, n_rows, m->n_blocks, m->blk_ptr, m->blk_col, static_cast<V*>(m->blk_val));
    HSP_HIP(m, hipGetLastError());
    if (n_cols <= (1ll << 24)) {  // super-blocks: a column's position among them fits the word
      m->n_super = (m->n_blocks + kSuperBlocks - 1) / kSuperBlocks;
      HSP_HIP(m, hipMalloc(&m->sb_ptr, sizeof(long long) * (m->n_super + 1)));
      hipLaunchKernelGGL(k_super_count, dim3(grid_for(m->n_super)), dim3(256), 0, m->stream, m->blk_ptr,
                         m->blk_col, m->n_blocks, m->n_super, d_count);
      HSP_HIP(m, hipGetLastError());
      if (int rc = scan_counts(m, d_count, m->n_super, m->sb_ptr, &m->n_super_entries)) return rc;
      HSP_HIP(m, hipMalloc(&m->sb_col, sizeof(int) * static_cast<size_t>(std::max(m->n_super_entries, 1ll))));
      HSP_HIP(m, hipMalloc(&m->blk_word, sizeof(int) * ne));
      hipLaunchKernelGGL(k_super_fill<V>, dim3(grid_for(m->n_super)), dim3(256), 0, m->stream, m->blk_ptr,
                         m->blk_col, static_cast<const V*>(m->blk_val), m->n_blocks, m->n_super, m->sb_ptr,
                         m->sb_col, m->blk_word);
      HSP_HIP(m, hipGetLastError());
    }
    HSP_HIP(m, hipStreamSynchronize(m->stream));
    return HSP_OK;
  };
  const int rc = run();
  release();
  return rc;
}

int apply_on_device(hsp_matrix* m, const void* w, int n_w, long long r0, long long r1, int precision,
                    void* out_device, double* kernel_ms) {
  const size_t elem = precision == HSP_F64 ? sizeof(double) : sizeof(float);
  if (!out_device) {
    const size_t need = static_cast<size_t>(r1 - r0) * static_cast<size_t>(n_w) * elem;
    if (need > m->out_bytes) {
      if (m->out) (void)hipFree(m->out);
      m->out = nullptr;
      m->out_bytes = 0;
      HSP_HIP(m, hipMalloc(&m->out, need));
      m->out_bytes = need;
    }
    out_device = m->out;
    m->out_r0 = r0;
    m->out_r1 = r1;
    m->out_nw = n_w;
    m->out_precision = precision;
  }
  if (kernel_ms && !m->ev[0]) {
    HSP_HIP(m, hipEventCreate(&m->ev[0]));
    HSP_HIP(m, hipEventCreate(&m->ev[1]));
  }
  if (kernel_ms) HSP_HIP(m, hipEventRecord(m->ev[0], m->stream));
  const bool v64 = m->value_type == HSP_F64;
  if (precision == HSP_F64) {
    const auto* wd = static_cast<const double*>(w);
    auto* od = static_cast<double*>(out_device);
    if (v64)
      launch_product<double, double>(m, wd, n_w, r0, r1, od);
    else
      launch_product<double, float>(m, wd, n_w, r0, r1, od);
  } else {
    const auto* wf = static_cast<const float*>(w);
    auto* of = static_cast<float*>(out_device);
    if (v64)
      launch_product<float, double>(m, wf, n_w, r0, r1, of);
    else
      launch_product<float, float>(m, wf, n_w, r0, r1, of);
  }
  HSP_HIP(m, hipGetLastError());
  if (kernel_ms) HSP_HIP(m, hipEventRecord(m->ev[1], m->stream));
  HSP_HIP(m, hipStreamSynchronize(m->stream));
  if (kernel_ms) {
    float ms = 0.0f;
    HSP_HIP(m, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    *kernel_ms = ms;
  }
  return HSP_OK;
}

// out[j] = sum over column j's entries of val * v[row]: one 256-lane workgroup per
// column (grid-stride over columns); lane t adds entries t, t + 256, ... in order, then
// a fixed tree over the lanes. Products are rounded before the add (no contraction in
// this file), so the result depends only on the matrix and v.
template <typename V>
__global__ void __launch_bounds__(256) k_transpose(const long long* __restrict__ ptr, const int* __restrict__ row,
                                                   const V* __restrict__ val, long long n_cols,
                                                   const double* __restrict__ v, double* __restrict__ out) {
  __shared__ double red[256];
  for (long long c = blockIdx.x; c < n_cols; c += gridDim.x) {
    double acc = 0.0;
    for (long long e = ptr[c] + threadIdx.x; e < ptr[c + 1]; e += blockDim.x)
      acc += static_cast<double>(val[e]) * v[row[e]];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
    __syncthreads();  // red is reused by the next column
  }
}

int check_apply(hsp_matrix* m, int n_w, long long r0, long long r1, int precision) {
  if (n_w < 1) return fail(m, HSP_EINVAL, "n_w must be at least 1");
  if (precision != HSP_F64 && precision != HSP_F32) return fail(m, HSP_EINVAL, "unknown precision");
  if (r0 < 0 || r1 > m->n_rows || r0 >= r1) return fail(m, HSP_EINVAL, "row range outside the matrix");
  const ProductShape shape = product_shape(n_w, r0, r1, precision == HSP_F32);
  if (shape.groups_per_pass * shape.passes > kMaxWorkgroups)
    return fail(m, HSP_EINVAL, "row range and vector count too large for one launch: split the rows");
  return HSP_OK;
}

}  // namespace

extern "C" {

const char* hsp_last_error(hsp_matrix* m) { return m ? m->error.c_str() : "null matrix"; }

int hsp_create(int device, long long n_rows, long long n_cols, const int64_t* indptr,
               const int64_t* indices, const void* values, int value_type, hsp_matrix** out) {
  if (!out || !indptr || n_rows < 1 || n_cols < 1) return HSP_EINVAL;
  if (value_type != HSP_F64 && value_type != HSP_F32) return HSP_EINVAL;
  *out = nullptr;
  if (n_cols >= (1ll << 31)) return HSP_EINVAL;  // column indices are int32 on the device
  if (n_rows >= (1ll << 31)) return HSP_EINVAL;  // and so are the CSC copy's row indices
  if (indptr[0] != 0) return HSP_EINVAL;
  for (long long c = 0; c < n_cols; ++c)
    if (indptr[c + 1] < indptr[c]) return HSP_EINVAL;
  const long long nnz = indptr[n_cols];
  if (nnz > 0 && (!indices || !values)) return HSP_EINVAL;
  for (long long e = 0; e < nnz; ++e)
    if (indices[e] < 0 || indices[e] >= n_rows) return HSP_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return HSP_EHIP;
  auto* m = new hsp_matrix();
  m->device = device;
  m->n_rows = n_rows;
  m->n_cols = n_cols;
  m->nnz = nnz;
  m->value_type = value_type;
  m->value_bytes = value_type == HSP_F64 ? sizeof(double) : sizeof(float);
  m->n_blocks = (n_rows + kBlockRows - 1) / kBlockRows;
  const size_t nz = static_cast<size_t>(std::max(nnz, 1ll));
  int rc = HSP_OK;
  auto step = [&](hipError_t e) {
    if (e != hipSuccess && rc == HSP_OK) {
      rc = e == hipErrorOutOfMemory ? HSP_ENOMEM : HSP_EHIP;
      m->error = hipGetErrorString(e);
    }
  };
  step(hipStreamCreate(&m->stream));
  step(hipMalloc(&m->row_ptr, sizeof(long long) * (n_rows + 1)));
  step(hipMalloc(&m->col, sizeof(int) * nz));
  step(hipMalloc(&m->val, m->value_bytes * nz));
  step(hipMalloc(&m->blk_ptr, sizeof(long long) * (m->n_blocks + 1)));
  // the CSC copy of the transpose product (row indices narrowed to int32)
  step(hipMalloc(&m->csc_ptr, sizeof(long long) * (n_cols + 1)));
  step(hipMalloc(&m->csc_row, sizeof(int) * nz));
  step(hipMalloc(&m->csc_val, m->value_bytes * nz));
  if (rc == HSP_OK) {
    std::vector<int> rows32(nz, 0);
    for (long long e = 0; e < nnz; ++e) rows32[e] = static_cast<int>(indices[e]);
    step(hipMemcpy(m->csc_ptr, indptr, sizeof(long long) * (n_cols + 1), hipMemcpyHostToDevice));
    step(hipMemcpy(m->csc_row, rows32.data(), sizeof(int) * nz, hipMemcpyHostToDevice));
    if (nnz > 0) step(hipMemcpy(m->csc_val, values, m->value_bytes * nnz, hipMemcpyHostToDevice));
  }
  if (rc == HSP_OK)
    rc = value_type == HSP_F64 ? build_forms(m, indptr, indices, static_cast<const double*>(values))
                               : build_forms(m, indptr, indices, static_cast<const float*>(values));
  if (rc != HSP_OK) {
    hsp_destroy(m);
    return rc;
  }
  *out = m;
  return HSP_OK;
}

void hsp_destroy(hsp_matrix* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  for (void* p : {static_cast<void*>(m->row_ptr), static_cast<void*>(m->col), static_cast<void*>(m->val),
                  static_cast<void*>(m->blk_ptr), static_cast<void*>(m->blk_col), m->blk_val, m->w, m->out,
                  static_cast<void*>(m->csc_ptr), static_cast<void*>(m->csc_row), m->csc_val,
                  static_cast<void*>(m->tv), static_cast<void*>(m->tout), static_cast<void*>(m->sb_ptr),
                  static_cast<void*>(m->sb_col), static_cast<void*>(m->blk_word)})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : m->ev)
    if (e) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int hsp_apply_transpose(hsp_matrix* m, const double* v, double* out, double* kernel_ms) {
  if (!m) return HSP_EINVAL;
  if (!v || !out) return fail(m, HSP_EINVAL, "null vector");
  HSP_HIP(m, hipSetDevice(m->device));
  if (!m->tv) HSP_HIP(m, hipMalloc(&m->tv, sizeof(double) * m->n_rows));
  if (!m->tout) HSP_HIP(m, hipMalloc(&m->tout, sizeof(double) * m->n_cols));
  HSP_HIP(m, hipMemcpyAsync(m->tv, v, sizeof(double) * m->n_rows, hipMemcpyHostToDevice, m->stream));
  if (kernel_ms && !m->ev[0]) {
    HSP_HIP(m, hipEventCreate(&m->ev[0]));
    HSP_HIP(m, hipEventCreate(&m->ev[1]));
  }
  if (kernel_ms) HSP_HIP(m, hipEventRecord(m->ev[0], m->stream));
  const unsigned int grid = static_cast<unsigned int>(std::min<long long>(m->n_cols, 65536));
  if (m->value_type == HSP_F64)
    hipLaunchKernelGGL(k_transpose<double>, dim3(grid), dim3(256), 0, m->stream, m->csc_ptr, m->csc_row,
                       static_cast<const double*>(m->csc_val), m->n_cols, m->tv, m->tout);
  else
    hipLaunchKernelGGL(k_transpose<float>, dim3(grid), dim3(256), 0, m->stream, m->csc_ptr, m->csc_row,
                       static_cast<const float*>(m->csc_val), m->n_cols, m->tv, m->tout);
  HSP_HIP(m, hipGetLastError());
  if (kernel_ms) HSP_HIP(m, hipEventRecord(m->ev[1], m->stream));
  HSP_HIP(m, hipMemcpyAsync(out, m->tout, sizeof(double) * m->n_cols, hipMemcpyDeviceToHost, m->stream));
  HSP_HIP(m, hipStreamSynchronize(m->stream));
  if (kernel_ms) {
    float ms = 0.0f;
    HSP_HIP(m, hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    *kernel_ms = ms;
  }
  return HSP_OK;
}

int hsp_apply(hsp_matrix* m, const double* weights, int n_w, long long row_begin, long long row_end,
              int precision, void* out_device, double* kernel_ms) {
  if (!m) return HSP_EINVAL;
  if (!weights) return fail(m, HSP_EINVAL, "null weights");
  if (int rc = check_apply(m, n_w, row_begin, row_end, precision)) return rc;
  HSP_HIP(m, hipSetDevice(m->device));
  const size_t count = static_cast<size_t>(m->n_cols) * static_cast<size_t>(n_w);
  const size_t bytes = count * (precision == HSP_F64 ? sizeof(double) : sizeof(float));
  if (bytes > m->w_bytes) {
    if (m->w) (void)hipFree(m->w);
    m->w = nullptr;
    m->w_bytes = 0;
    HSP_HIP(m, hipMalloc(&m->w, bytes));
    m->w_bytes = bytes;
  }
  if (precision == HSP_F64) {
    HSP_HIP(m, hipMemcpy(m->w, weights, bytes, hipMemcpyHostToDevice));
  } else {
    std::vector<float> w32(weights, weights + count);
    HSP_HIP(m, hipMemcpy(m->w, w32.data(), bytes, hipMemcpyHostToDevice));
  }
  return apply_on_device(m, m->w, n_w, row_begin, row_end, precision, out_device, kernel_ms);
}

int hsp_apply_device(hsp_matrix* m, const void* weights_device, int n_w, long long row_begin,
                     long long row_end, int precision, void* out_device, double* kernel_ms) {
  if (!m) return HSP_EINVAL;
  if (!weights_device) return fail(m, HSP_EINVAL, "null weights");
  if (int rc = check_apply(m, n_w, row_begin, row_end, precision)) return rc;
  HSP_HIP(m, hipSetDevice(m->device));
  return apply_on_device(m, weights_device, n_w, row_begin, row_end, precision, out_device, kernel_ms);
}

int hsp_fetch(hsp_matrix* m, long long row_begin, long long row_end, void* host) {
  if (!m) return HSP_EINVAL;
  if (!host) return fail(m, HSP_EINVAL, "null host buffer");
  if (!m->out || row_begin < m->out_r0 || row_end > m->out_r1 || row_begin >= row_end)
    return fail(m, HSP_EINVAL, "rows not in the last internal product");
  HSP_HIP(m, hipSetDevice(m->device));
  const size_t elem = m->out_precision == HSP_F64 ? sizeof(double) : sizeof(float);
  const size_t off = static_cast<size_t>(row_begin - m->out_r0) * m->out_nw * elem;
  HSP_HIP(m, hipMemcpy(host, static_cast<char*>(m->out) + off,
                       elem * static_cast<size_t>(row_end - row_begin) * m->out_nw, hipMemcpyDeviceToHost));
  return HSP_OK;
}

int hsp_csr(hsp_matrix* m, int64_t* row_ptr, int32_t* col, void* val) {
  if (!m) return HSP_EINVAL;
  if (!row_ptr) return fail(m, HSP_EINVAL, "null row_ptr");
  HSP_HIP(m, hipSetDevice(m->device));
  HSP_HIP(m, hipMemcpy(row_ptr, m->row_ptr, sizeof(long long) * (m->n_rows + 1), hipMemcpyDeviceToHost));
  if (m->nnz) {
    if (!col || !val) return fail(m, HSP_EINVAL, "null col or val");
    HSP_HIP(m, hipMemcpy(col, m->col, sizeof(int) * m->nnz, hipMemcpyDeviceToHost));
    HSP_HIP(m, hipMemcpy(val, m->val, m->value_bytes * m->nnz, hipMemcpyDeviceToHost));
  }
  return HSP_OK;
}

long long hsp_block_entries(hsp_matrix* m) { return m ? m->n_block_entries : -1; }

}  // extern "C"
