// Kernels of the 3D convolutional Dij-column denoiser: BF16 MFMA convolutions
// (forward, data gradient, weight gradient) with FP32 accumulation, and the
// elementwise, pooling, loss, reduction and optimizer kernels of the network, behind
// the C ABI of include/hipdn.h. The reference has no learned component; these
// kernels serve the denoised Dij builder (pyradmc_amd/denoise.py).
//
// Layouts. Activations are BF16, position-major "nd" = [N][D][H][W][C] with C a
// multiple of 16; a tensor that feeds a weight gradient also has a channel-major copy
// "cm" = [C][N*D*H*W]. Weights: FP32 master [27][Cout][Cin] (tap = (kz*3+ky)*3+kx);
// BF16 forward copy wf = [27][Cout][Cin]; BF16 data-gradient copy wd = [27][Cin][Cout]
// with the taps flipped (tap -> 26 - tap), so the data gradient is a forward
// convolution of the output gradient.
//
// MFMA: v_mfma_f32_16x16x32_bf16 (gfx950). For D = A x B with A 16x32, B 32x16:
// lane l holds A[l % 16][8 (l / 16) + e], B[8 (l / 16) + e][l % 16] (e = 0..7), and
// D[4 (l / 16) + r][l % 16] (r = 0..3). Every operand is one aligned 16-byte load.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/hipdn.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int e = 0; e < 8; ++e) z[e] = static_cast<__bf16>(0.0f);
  return z;
}

__device__ inline float act(float v, float slope) { return v >= 0.0f ? v : v * slope; }

// Where a convolution reads its input from (inference fuses the preceding element-wise
// step into the operand loads; the values are those the separate kernels would store):
// kSrcPlain x itself; kSrcUp the nearest 2x upsampling of x, which has half the output
// dims (k_up_fwd); kSrcAdd x + x2 rounded to bf16 per element (k_add); kSrcLow8 x whose
// channels beyond the first 8 are zero (the network's padded input), which the
// reduction then skips.
constexpr int kSrcPlain = 0, kSrcUp = 1, kSrcAdd = 2, kSrcLow8 = 3;

__device__ inline bf16x8 add8(bf16x8 a, bf16x8 b) {
  bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = static_cast<__bf16>(static_cast<float>(a[e]) + static_cast<float>(b[e]));
  return r;
}

// ------------------------------------------------------------------ convolution
//
// y[p][co] = act(bias[co] + sum_{tap, ci} x[p + off(tap)][ci] wt[tap][co][ci]) with
// zero padding. One wave computes 16*TM output channels x 64 positions (4 blocks of
// 16); a workgroup of 4 waves covers 256 consecutive positions (P % 64 == 0). K = 27*cin runs in
// chunks of 32; each lane's 8 k values lie inside one tap since cin % 8 == 0.
template <int TM, int SRC>
__global__ void __launch_bounds__(256) k_conv_fwd(const __bf16* __restrict__ x,
                                                  const __bf16* __restrict__ x2,
                                                  const __bf16* __restrict__ wt,
                                                  const float* __restrict__ bias,
                                                  __bf16* __restrict__ y_nd,
                                                  __bf16* __restrict__ y_cm, int n, int d, int h,
                                                  int w, int cin, int cout, float slope) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long P = static_cast<long long>(n) * d * h * w;
  const long long p0 = (static_cast<long long>(blockIdx.x) * 4 + wave) * 64;
  if (p0 >= P) return;  // whole wave past the end (P % 64 == 0); no barriers follow
  const int co0 = blockIdx.y * 16 * TM;
  const int col = lane & 15;
  const int kg = lane >> 4;

  int pn[4], pz[4], py[4], px[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    long long t = p0 + 16 * nb + col;
    px[nb] = static_cast<int>(t % w);
    t /= w;
    py[nb] = static_cast<int>(t % h);
    t /= h;
    pz[nb] = static_cast<int>(t % d);
    pn[nb] = static_cast<int>(t / d);
  }

  f32x4 acc[TM][4];
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[m][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  const int kc = SRC == kSrcLow8 ? 8 : cin;  // channels read per tap
  const int K = 27 * kc;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const int k = k0 + 8 * kg;
    bf16x8 a[TM], b[4];
    if (k < K) {
      const int tap = k / kc;
      const int c = k - tap * kc;
      const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
      for (int m = 0; m < TM; ++m)
        a[m] = load8(wt + (static_cast<long long>(tap) * cout + co0 + 16 * m + col) * cin + c);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const int zz = pz[nb] + dz, yy = py[nb] + dy, xx = px[nb] + dx;
        if (zz >= 0 && zz < d && yy >= 0 && yy < h && xx >= 0 && xx < w) {
          if constexpr (SRC == kSrcUp) {
            const long long q =
                ((static_cast<long long>(pn[nb]) * (d >> 1) + (zz >> 1)) * (h >> 1) + (yy >> 1)) * (w >> 1) + (xx >> 1);
            b[nb] = load8(x + q * cin + c);
          } else {
            const long long q = ((static_cast<long long>(pn[nb]) * d + zz) * h + yy) * w + xx;
            b[nb] = load8(x + q * cin + c);
            if constexpr (SRC == kSrcAdd) b[nb] = add8(b[nb], load8(x2 + q * cin + c));
          }
        } else {
          b[nb] = zero8();
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < TM; ++m) a[m] = zero8();
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) b[nb] = zero8();
    }
#pragma unroll
    for (int m = 0; m < TM; ++m)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) acc[m][nb] = mfma16(a[m], b[nb], acc[m][nb]);
  }

#pragma unroll
  for (int m = 0; m < TM; ++m) {
    const int cb = co0 + 16 * m + 4 * kg;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const long long p = p0 + 16 * nb + col;
      __bf16 out[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = acc[m][nb][r] + (bias ? bias[cb + r] : 0.0f);
        out[r] = static_cast<__bf16>(act(v, slope));
      }
      uint2 packed;
      std::memcpy(&packed, out, 8);
      *reinterpret_cast<uint2*>(y_nd + p * cout + cb) = packed;
      if (y_cm) {
#pragma unroll
        for (int r = 0; r < 4; ++r) y_cm[static_cast<long long>(cb + r) * P + p] = out[r];
      }
    }
  }
}

// The same convolution for rows of KW = 64 or 32 voxels (w == KW, h % (256 / KW) == 0,
// cout 16 x TM): a workgroup computes 256 / KW consecutive h-rows of one z-slice (256
// positions; each wave 64 of them) and first stages the input those rows read - 3 z x
// (rows + 2) h x (KW + 2) w positions with their one-voxel halo, zero outside the
// volume - in LDS (KW 64: 38 KB for cin 16, 76 KB for cin 32; KW 32: 33, 65 or 130 KB
// for cin 16, 32 or 64). Every tap's operand then comes from LDS instead of a gathered
// global load: the kernel reads each input element from HBM once per workgroup instead
// of once per tap. The k-loop, MFMA operands and output are those of k_conv_fwd, so
// the two kernels give identical sums.
template <int TM, int CIN, int SRC, int KW = 64>
__global__ void __launch_bounds__(256) k_conv_fwd_tile(const __bf16* __restrict__ x,
                                                       const __bf16* __restrict__ x2,
                                                       const __bf16* __restrict__ wt,
                                                       const float* __restrict__ bias,
                                                       __bf16* __restrict__ y_nd,
                                                       __bf16* __restrict__ y_cm, int n, int d, int h,
                                                       int cout, float slope) {
  constexpr int kW = KW, kRows = 256 / KW, kTW = kW + 2, kTH = kRows + 2, kTZ = 3, kC8 = CIN / 8;
  __shared__ u32x4 tile[kTZ * kTH * kTW * kC8];
  const int rows = h / kRows;
  const int hb = static_cast<int>(blockIdx.x % rows);
  const long long t = blockIdx.x / rows;
  const int z = static_cast<int>(t % d);
  const int nn = static_cast<int>(t / d);
  const int h0 = kRows * hb;
  // The tile fill: each thread issues kFill loads before it stores any, so their
  // latencies overlap. A load whose position lies outside the volume (or past the tile)
  // reads the tensor's first element instead, always valid, and is replaced by zero.
  constexpr int kUnits = kTZ * kTH * kTW * kC8;
  constexpr int kFill = 8;
  for (int i0 = threadIdx.x; i0 < kUnits; i0 += 256 * kFill) {
    u32x4 v[kFill], v2[kFill];
    bool ok[kFill];
#pragma unroll
    for (int j = 0; j < kFill; ++j) {
      const int i = i0 + 256 * j;
      const int c8 = i % kC8;
      const int pos = i / kC8;
      const int xx = pos % kTW, yy = (pos / kTW) % kTH, zz = pos / (kTW * kTH);
      const int gz = z + zz - 1, gy = h0 + yy - 1, gx = xx - 1;
      ok[j] = i < kUnits && gz >= 0 && gz < d && gy >= 0 && gy < h && gx >= 0 && gx < kW;
      long long q;
      if constexpr (SRC == kSrcUp)
        q = ((static_cast<long long>(nn) * (d >> 1) + (gz >> 1)) * (h >> 1) + (gy >> 1)) * (kW >> 1) + (gx >> 1);
      else
        q = ((static_cast<long long>(nn) * d + gz) * h + gy) * kW + gx;
      q = ok[j] ? q : 0;
      v[j] = *reinterpret_cast<const u32x4*>(x + q * CIN + 8 * c8);
      if constexpr (SRC == kSrcAdd) v2[j] = *reinterpret_cast<const u32x4*>(x2 + q * CIN + 8 * c8);
    }
#pragma unroll
    for (int j = 0; j < kFill; ++j) {
      const int i = i0 + 256 * j;
      u32x4 val = v[j];
      if constexpr (SRC == kSrcAdd) {
        bf16x8 a8, b8;
        std::memcpy(&a8, &v[j], 16);
        std::memcpy(&b8, &v2[j], 16);
        a8 = add8(a8, b8);
        std::memcpy(&val, &a8, 16);
      }
      if (!ok[j]) val = u32x4{0u, 0u, 0u, 0u};
      if (i < kUnits) tile[i] = val;
    }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane & 15;
  const int kg = lane >> 4;
  // Position block nb (16 positions) of this wave: row and first x within the rows.
  int blk_row[4], blk_x[4];
#pragma unroll
  for (int nb = 0; nb < 4; ++nb) {
    blk_row[nb] = wave * (64 / kW) + (16 * nb) / kW;
    blk_x[nb] = (16 * nb) % kW;
  }
  f32x4 acc[TM][4];
#pragma unroll
  for (int m = 0; m < TM; ++m)
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) acc[m][nb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  constexpr int kKC = SRC == kSrcLow8 ? 8 : CIN;  // channels read per tap
  constexpr int K = 27 * kKC;
  // Weight fragments of k-step k0 (zero past K): global loads, issued one k-step ahead
  // of their MFMAs so that their latency overlaps the previous step's work.
  auto load_a = [&](int k0, bf16x8 (&a)[TM]) {
    const int k = k0 + 8 * kg;
    if (k < K) {
      const int tap = k / kKC;
      const int c = k - tap * kKC;
#pragma unroll
      for (int m = 0; m < TM; ++m)
        a[m] = load8(wt + (static_cast<long long>(tap) * cout + 16 * m + col) * CIN + c);
    } else {
#pragma unroll
      for (int m = 0; m < TM; ++m) a[m] = zero8();
    }
  };
  // Input fragments of k-step k0 from the LDS tile (zero past K).
  auto load_b = [&](int k0, bf16x8 (&b)[4]) {
    const int k = k0 + 8 * kg;
    if (k < K) {
      const int tap = k / kKC;
      const int c = k - tap * kKC;
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // halo-shifted: 0..2
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) {
        const u32x4 v = tile[((dz * kTH + blk_row[nb] + dy) * kTW + dx + blk_x[nb] + col) * kC8 + c / 8];
        std::memcpy(&b[nb], &v, 16);
      }
    } else {
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) b[nb] = zero8();
    }
  };
  auto mfmas = [&](const bf16x8 (&a)[TM], const bf16x8 (&b)[4]) {
#pragma unroll
    for (int m = 0; m < TM; ++m)
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) acc[m][nb] = mfma16(a[m], b[nb], acc[m][nb]);
  };
  // k-steps in order, two per trip: the next step's fragments load into the other
  // register set while this step's MFMAs run (ping-pong, no copies).
  bf16x8 a0[TM], a1[TM], b0[4], b1[4];
  load_a(0, a0);
  load_b(0, b0);
  for (int k0 = 0; k0 < K; k0 += 64) {
    load_a(k0 + 32, a1);
    load_b(k0 + 32, b1);
    mfmas(a0, b0);
    if (k0 + 32 >= K) break;
    load_a(k0 + 64, a0);
    load_b(k0 + 64, b0);
    mfmas(a1, b1);
  }

  const long long P = static_cast<long long>(n) * d * h * kW;
  const long long row0 = (static_cast<long long>(nn) * d + z) * h + h0;
#pragma unroll
  for (int m = 0; m < TM; ++m) {
    const int cb = 16 * m + 4 * kg;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
      const long long p = (row0 + blk_row[nb]) * kW + blk_x[nb] + col;
      __bf16 out[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = acc[m][nb][r] + (bias ? bias[cb + r] : 0.0f);
        out[r] = static_cast<__bf16>(act(v, slope));
      }
      uint2 packed;
      std::memcpy(&packed, out, 8);
      *reinterpret_cast<uint2*>(y_nd + p * cout + cb) = packed;
      if (y_cm) {
#pragma unroll
        for (int r = 0; r < 4; ++r) y_cm[static_cast<long long>(cb + r) * P + p] = out[r];
      }
    }
  }
}

// 8 channel-major elements x[c][row + x0 + dx .. + 7] of one W-row (w % 8 == 0,
// x0 % 8 == 0), zero outside [0, w): the tap's x shift as a funnel shift of aligned
// 16-byte loads.
__device__ inline bf16x8 load_row8(const __bf16* row, int x0, int w, int dx) {
  const u32x4 cur = *reinterpret_cast<const u32x4*>(row + x0);
  u32x4 out;
  if (dx == 0) {
    out = cur;
  } else if (dx > 0) {
    const u32x4 nxt = x0 + 8 < w ? *reinterpret_cast<const u32x4*>(row + x0 + 8) : u32x4{0, 0, 0, 0};
    out[0] = (cur[0] >> 16) | (cur[1] << 16);
    out[1] = (cur[1] >> 16) | (cur[2] << 16);
    out[2] = (cur[2] >> 16) | (cur[3] << 16);
    out[3] = (cur[3] >> 16) | (nxt[0] << 16);
  } else {
    const u32x4 prv = x0 > 0 ? *reinterpret_cast<const u32x4*>(row + x0 - 8) : u32x4{0, 0, 0, 0};
    out[0] = (cur[0] << 16) | (prv[3] >> 16);
    out[1] = (cur[1] << 16) | (cur[0] >> 16);
    out[2] = (cur[2] << 16) | (cur[1] >> 16);
    out[3] = (cur[3] << 16) | (cur[2] >> 16);
  }
  bf16x8 r;
  std::memcpy(&r, &out, 16);
  return r;
}

// Weight-gradient partials: part[s][tap][co][ci] = sum over split s's positions of
// dz[p][co] x[p + off(tap)][ci]. One wave owns one 16x16 (ci, co) block for all 27
// taps (108 accumulators) and a contiguous range of positions, in chunks of 32.
__global__ void __launch_bounds__(64) k_conv_wgrad(const __bf16* __restrict__ x_cm,
                                                   const __bf16* __restrict__ dz_cm,
                                                   float* __restrict__ part, int n, int d, int h,
                                                   int w, int cin, int cout, long long per_split) {
  const int lane = threadIdx.x;
  const int col = lane & 15;
  const int kg = lane >> 4;
  const int ci0 = 16 * blockIdx.x;
  const int co0 = 16 * blockIdx.y;
  const int split = blockIdx.z;
  const long long P = static_cast<long long>(n) * d * h * w;
  const long long pbeg = static_cast<long long>(split) * per_split;
  const long long pend = pbeg + per_split < P ? pbeg + per_split : P;

  f32x4 acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  const __bf16* xrow_c = x_cm + static_cast<long long>(ci0 + col) * P;
  const __bf16* dz_c = dz_cm + static_cast<long long>(co0 + col) * P;
  for (long long pb = pbeg; pb < pend; pb += 32) {
    const long long p = pb + 8 * kg;  // this lane's 8 positions: one aligned W-row segment
    const bf16x8 b = load8(dz_c + p);
    long long t = p;
    const int x0 = static_cast<int>(t % w);
    t /= w;
    const int y = static_cast<int>(t % h);
    t /= h;
    const int z = static_cast<int>(t % d);
    const long long nn = t / d;
#pragma unroll
    for (int dzz = -1; dzz <= 1; ++dzz) {
#pragma unroll
      for (int dyy = -1; dyy <= 1; ++dyy) {
        const int zz = z + dzz, yy = y + dyy;
        const bool row_ok = zz >= 0 && zz < d && yy >= 0 && yy < h;
        const long long row_off = row_ok ? ((nn * d + zz) * h + yy) * w : 0;
#pragma unroll
        for (int dxx = -1; dxx <= 1; ++dxx) {
          const int tap = ((dzz + 1) * 3 + (dyy + 1)) * 3 + (dxx + 1);
          const bf16x8 a = row_ok ? load_row8(xrow_c + row_off, x0, w, dxx) : zero8();
          acc[tap] = mfma16(a, b, acc[tap]);
        }
      }
    }
  }
  // D[i = ci][j = co]: lane holds ci = ci0 + 4 kg + r, co = co0 + col.
  float* out = part + static_cast<long long>(split) * 27 * cout * cin;
#pragma unroll
  for (int tap = 0; tap < 27; ++tap) {
    float4 v = make_float4(acc[tap][0], acc[tap][1], acc[tap][2], acc[tap][3]);
    *reinterpret_cast<float4*>(out + (static_cast<long long>(tap) * cout + co0 + col) * cin + ci0 +
                               4 * kg) = v;
  }
}

// out[i] = sum_s part[s][i], in split order (deterministic).
__global__ void k_reduce_splits(const float* __restrict__ part, float* __restrict__ out,
                                long long len, int splits) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    float s = 0.0f;
    for (int k = 0; k < splits; ++k) s += part[static_cast<long long>(k) * len + i];
    out[i] = s;
  }
}

// Per-channel sums of a channel-major tensor: part[s][c] over split s of the positions.
__global__ void __launch_bounds__(256) k_channel_sum(const __bf16* __restrict__ cm,
                                                     float* __restrict__ part, long long P,
                                                     long long per_split) {
  const int c = blockIdx.x;
  const int s = blockIdx.y;
  const long long beg = static_cast<long long>(s) * per_split;
  const long long end = beg + per_split < P ? beg + per_split : P;
  float acc = 0.0f;
  for (long long p = beg + threadIdx.x; p < end; p += blockDim.x)
    acc += static_cast<float>(cm[static_cast<long long>(c) * P + p]);
  __shared__ float red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[static_cast<long long>(s) * gridDim.x + c] = red[0];  // [split][C]
}

// ------------------------------------------------------------------- elementwise
//
// Thread = (position p, 8-channel group g); p runs fastest so the channel-major
// writes of a wave are contiguous.

__device__ inline void store_both(__bf16* nd, __bf16* cm, long long P, int C, long long p, int g,
                                  const float* v) {
  __bf16 o[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = static_cast<__bf16>(v[e]);
  bf16x8 packed;
  std::memcpy(&packed, o, 16);
  *reinterpret_cast<bf16x8*>(nd + p * C + 8 * g) = packed;
  if (cm) {
#pragma unroll
    for (int e = 0; e < 8; ++e) cm[static_cast<long long>(8 * g + e) * P + p] = o[e];
  }
}

// dz = dy * act'(y) (act'(y) = 1 for y > 0, slope otherwise), both layouts.
__global__ void k_act_bwd(const __bf16* __restrict__ y, const __bf16* __restrict__ dy,
                          __bf16* __restrict__ dz_nd, __bf16* __restrict__ dz_cm, long long P,
                          int C, float slope) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  const bf16x8 yv = load8(y + p * C + 8 * g);
  const bf16x8 gv = load8(dy + p * C + 8 * g);
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float gy = static_cast<float>(gv[e]);
    v[e] = static_cast<float>(yv[e]) > 0.0f ? gy : gy * slope;
  }
  store_both(dz_nd, dz_cm, P, C, p, g, v);
}

// y = a + b, both layouts.
__global__ void k_add(const __bf16* __restrict__ a, const __bf16* __restrict__ b,
                      __bf16* __restrict__ y_nd, __bf16* __restrict__ y_cm, long long P, int C) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  const bf16x8 av = load8(a + p * C + 8 * g);
  const bf16x8 bv = load8(b + p * C + 8 * g);
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = static_cast<float>(av[e]) + static_cast<float>(bv[e]);
  store_both(y_nd, y_cm, P, C, p, g, v);
}

// 2x2x2 average pooling: x [n][2d][2h][2w][C] -> y [n][d][h][w][C] (nd and cm).
__global__ void k_pool_fwd(const __bf16* __restrict__ x, __bf16* __restrict__ y_nd,
                           __bf16* __restrict__ y_cm, int n, int d, int h, int w, int C) {
  const long long P = static_cast<long long>(n) * d * h * w;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  long long t = p;
  const int xx = static_cast<int>(t % w);
  t /= w;
  const int yy = static_cast<int>(t % h);
  t /= h;
  const int zz = static_cast<int>(t % d);
  const long long nn = t / d;
  float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const long long q = ((nn * 2 * d + 2 * zz + a) * 2 * h + 2 * yy + b) * 2 * w + 2 * xx + c;
        const bf16x8 s = load8(x + q * C + 8 * g);
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] += static_cast<float>(s[e]);
      }
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] *= 0.125f;
  store_both(y_nd, y_cm, P, C, p, g, v);
}

// Gradient of the pooling: dx[fine] = dy[coarse] / 8 (nd only). Dims are the fine
// ones (2d, 2h, 2w).
__global__ void k_pool_bwd(const __bf16* __restrict__ dy, __bf16* __restrict__ dx, int n, int d2,
                           int h2, int w2, int C) {
  const long long P = static_cast<long long>(n) * d2 * h2 * w2;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  long long t = p;
  const int xx = static_cast<int>(t % w2);
  t /= w2;
  const int yy = static_cast<int>(t % h2);
  t /= h2;
  const int zz = static_cast<int>(t % d2);
  const long long nn = t / d2;
  const long long q = ((nn * (d2 / 2) + zz / 2) * (h2 / 2) + yy / 2) * (w2 / 2) + xx / 2;
  const bf16x8 s = load8(dy + q * C + 8 * g);
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = static_cast<float>(s[e]) * 0.125f;
  store_both(dx, nullptr, P, C, p, g, v);
}

// Nearest 2x upsampling: x [n][d][h][w][C] -> y [n][2d][2h][2w][C]. Dims are the fine
// ones.
__global__ void k_up_fwd(const __bf16* __restrict__ x, __bf16* __restrict__ y_nd,
                         __bf16* __restrict__ y_cm, int n, int d2, int h2, int w2, int C) {
  const long long P = static_cast<long long>(n) * d2 * h2 * w2;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  long long t = p;
  const int xx = static_cast<int>(t % w2);
  t /= w2;
  const int yy = static_cast<int>(t % h2);
  t /= h2;
  const int zz = static_cast<int>(t % d2);
  const long long nn = t / d2;
  const long long q = ((nn * (d2 / 2) + zz / 2) * (h2 / 2) + yy / 2) * (w2 / 2) + xx / 2;
  const bf16x8 s = load8(x + q * C + 8 * g);
  float v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = static_cast<float>(s[e]);
  store_both(y_nd, y_cm, P, C, p, g, v);
}

// Gradient of the upsampling: dx[coarse] = sum of dy over its 8 fine cells (nd only).
// Dims are the coarse ones.
__global__ void k_up_bwd(const __bf16* __restrict__ dy, __bf16* __restrict__ dx, int n, int d,
                         int h, int w, int C) {
  const long long P = static_cast<long long>(n) * d * h * w;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int groups = C / 8;
  if (i >= P * groups) return;
  const long long p = i % P;
  const int g = static_cast<int>(i / P);
  long long t = p;
  const int xx = static_cast<int>(t % w);
  t /= w;
  const int yy = static_cast<int>(t % h);
  t /= h;
  const int zz = static_cast<int>(t % d);
  const long long nn = t / d;
  float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const long long q = ((nn * 2 * d + 2 * zz + a) * 2 * h + 2 * yy + b) * 2 * w + 2 * xx + c;
        const bf16x8 s = load8(dy + q * C + 8 * g);
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] += static_cast<float>(s[e]);
      }
  store_both(dx, nullptr, P, C, p, g, v);
}

// Residual head and sigma-weighted squared error. pred = inp[p][0] + out[p][0];
// per-block partial of sum weight * (pred - target)^2; gradient with respect to out
// (channel 0; the other channels zero) = 2 weight (pred - target) * grad_scale, in
// both layouts. Also writes pred (float) when pred_out is not null.
__global__ void __launch_bounds__(256) k_loss(const __bf16* __restrict__ out,
                                              const __bf16* __restrict__ inp, int c_out, int c_in,
                                              const float* __restrict__ target,
                                              const float* __restrict__ weight,
                                              __bf16* __restrict__ dout_nd,
                                              __bf16* __restrict__ dout_cm,
                                              float* __restrict__ pred_out,
                                              float* __restrict__ partial, long long P,
                                              float grad_scale) {
  const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (p < P) {
    const float pred = static_cast<float>(inp[p * c_in]) + static_cast<float>(out[p * c_out]);
    if (pred_out) pred_out[p] = pred;
    if (target) {
      const float diff = pred - target[p];
      const float wt = weight ? weight[p] : 1.0f;
      l = wt * diff * diff;
      if (dout_nd) {
        const __bf16 g = static_cast<__bf16>(2.0f * wt * diff * grad_scale);
        const __bf16 zero = static_cast<__bf16>(0.0f);
        for (int c = 0; c < c_out; ++c) {
          dout_nd[p * c_out + c] = c == 0 ? g : zero;
          if (dout_cm) dout_cm[static_cast<long long>(c) * P + p] = c == 0 ? g : zero;
        }
      }
    }
  }
  __shared__ float red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (static_cast<int>(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = red[0];
}

// FP32 master -> BF16 forward [27][cout][cin] and flipped, transposed data-gradient
// [27][cin][cout] copies.
__global__ void k_prepare_weights(const float* __restrict__ master, __bf16* __restrict__ wf,
                                  __bf16* __restrict__ wd, int cin, int cout) {
  const long long total = 27LL * cin * cout;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int ci = static_cast<int>(i % cin);
  const long long t = i / cin;
  const int co = static_cast<int>(t % cout);
  const int tap = static_cast<int>(t / cout);
  const __bf16 v = static_cast<__bf16>(master[i]);
  wf[i] = v;
  wd[(static_cast<long long>(26 - tap) * cin + ci) * cout + co] = v;
}

// Adam (Kingma and Ba 2015) with bias correction; zero gradients leave padding zero.
__global__ void k_adam(float* __restrict__ param, const float* __restrict__ grad,
                       float* __restrict__ m, float* __restrict__ v, long long len, float lr,
                       float beta1, float beta2, float eps, float corr1, float corr2,
                       float weight_decay) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const float g = grad[i] + weight_decay * param[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * g;
  const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
  m[i] = mi;
  v[i] = vi;
  param[i] -= lr * (mi / corr1) / (sqrtf(vi / corr2) + eps);
}

unsigned int blocks_for(long long n, int threads) {
  return static_cast<unsigned int>((n + threads - 1) / threads);
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int launch_status() { return hipGetLastError() == hipSuccess ? HDN_OK : HDN_EHIP; }

bool shape_ok(int n, int d, int h, int w) { return n > 0 && d > 0 && h > 0 && w > 0; }

}  // namespace

extern "C" {

}  // extern "C"

namespace {

template <int SRC>
int conv_fwd(const __bf16* xb, const __bf16* x2b, const __bf16* wb, const float* bias, __bf16* yb, __bf16* cb,
             int n, int d, int h, int w, int cin, int cout, float slope, hipStream_t stream) {
  if (w == 64 && h % 4 == 0 && (cin == 16 || cin == 32) && (cout == 16 || (cout == 32 && cin == 16))) {
    // Full-resolution layers: input staged in LDS (k_conv_fwd_tile).
    const dim3 tiles(static_cast<unsigned int>(static_cast<long long>(n) * d * (h / 4)));
    if (cin == 16 && cout == 32)
      hipLaunchKernelGGL((k_conv_fwd_tile<2, 16, SRC>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n,
                         d, h, cout, slope);
    else if (cin == 16)
      hipLaunchKernelGGL((k_conv_fwd_tile<1, 16, SRC>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n,
                         d, h, cout, slope);
    else
      hipLaunchKernelGGL((k_conv_fwd_tile<1, 32, SRC>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n,
                         d, h, cout, slope);
    return launch_status();
  }
  if (w == 32 && h % 8 == 0 && (cin == 16 || cin == 32 || cin == 64) && cout == 32) {
    // Half-resolution layers: the same staging, 8 rows of 32 per workgroup.
    const dim3 tiles(static_cast<unsigned int>(static_cast<long long>(n) * d * (h / 8)));
    if (cin == 16)
      hipLaunchKernelGGL((k_conv_fwd_tile<2, 16, SRC, 32>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb,
                         n, d, h, cout, slope);
    else if (cin == 32)
      hipLaunchKernelGGL((k_conv_fwd_tile<2, 32, SRC, 32>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb,
                         n, d, h, cout, slope);
    else
      hipLaunchKernelGGL((k_conv_fwd_tile<2, 64, SRC, 32>), tiles, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb,
                         n, d, h, cout, slope);
    return launch_status();
  }
  const long long P = static_cast<long long>(n) * d * h * w;
  const int tm = cout % 64 == 0 ? 4 : (cout % 32 == 0 ? 2 : 1);
  const dim3 grid(static_cast<unsigned int>((P + 255) / 256), static_cast<unsigned int>(cout / (16 * tm)));
  if (tm == 4)
    hipLaunchKernelGGL((k_conv_fwd<4, SRC>), grid, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n, d, h, w,
                       cin, cout, slope);
  else if (tm == 2)
    hipLaunchKernelGGL((k_conv_fwd<2, SRC>), grid, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n, d, h, w,
                       cin, cout, slope);
  else
    hipLaunchKernelGGL((k_conv_fwd<1, SRC>), grid, dim3(256), 0, stream, xb, x2b, wb, bias, yb, cb, n, d, h, w,
                       cin, cout, slope);
  return launch_status();
}

// Weight-gradient partials of hdn_conv3d_wgrad_noalloc live here: a library-owned array
// each GPU holds from the code object's load until the process ends. In host code the
// name stands for a host-side shadow, not a device address: kernels get the address
// hipGetSymbolAddress returns for the current device.
__device__ __align__(256) float scratch[HDN_SCRATCH_FLOATS];

// Validates a weight-gradient call and returns the positions each split owns.
int wgrad_split(const void* x_cm, const void* dz_cm, int splits, int n, int d, int h, int w, int cin,
                int cout, long long* per_split_out) {
  if (!x_cm || !dz_cm || splits < 1 || !shape_ok(n, d, h, w)) return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d * h * w;
  if (w % 8 != 0 || P % 32 != 0 || cin % 16 != 0 || cout % 16 != 0 || cin < 16 || cout < 16)
    return HDN_EINVAL;
  long long per_split = (P + splits - 1) / splits;
  per_split = (per_split + 31) / 32 * 32;
  if (per_split * (splits - 1) >= P) return HDN_EINVAL;  // every split must own positions
  *per_split_out = per_split;
  return HDN_OK;
}

void launch_wgrad(const void* x_cm, const void* dz_cm, float* part, int splits, int n, int d, int h,
                  int w, int cin, int cout, long long per_split, hipStream_t stream) {
  const dim3 grid(static_cast<unsigned int>(cin / 16), static_cast<unsigned int>(cout / 16),
                  static_cast<unsigned int>(splits));
  hipLaunchKernelGGL(k_conv_wgrad, grid, dim3(64), 0, stream, static_cast<const __bf16*>(x_cm),
                     static_cast<const __bf16*>(dz_cm), part, n, d, h, w, cin, cout, per_split);
}

void launch_reduce(const float* part, float* out, long long len, int splits, hipStream_t stream) {
  const unsigned int blocks = static_cast<unsigned int>(std::min<long long>(blocks_for(len, 256), 65535));
  hipLaunchKernelGGL(k_reduce_splits, dim3(blocks), dim3(256), 0, stream, part, out, len, splits);
}

// Device address of `scratch` on each GPU, looked up once per device with that device
// current (each GPU holds its own copy of the array at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_address[kMaxDevices] = {};

int scratch_for_current_device(void* stream, float** out) {
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) return HDN_EHIP;
  if (stream) {
    // The kernels run on the stream's device: it must be the current one, whose copy of
    // the array is the one looked up.
    int stream_device = -1;
    if (hipStreamGetDevice(as_stream(stream), &stream_device) != hipSuccess) return HDN_EHIP;
    if (stream_device != device) return HDN_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_address[device]) {
    void* address = nullptr;
    if (hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)) != hipSuccess || !address) return HDN_EHIP;
    g_scratch_address[device] = static_cast<float*>(address);
  }
  *out = g_scratch_address[device];
  return HDN_OK;
}

}  // namespace

extern "C" {

int hdn_conv3d_fwd_src(const void* x, const void* x2, int src, const void* wt, const float* bias, void* y_nd,
                       void* y_cm, int n, int d, int h, int w, int cin, int cout, float slope, void* stream) {
  if (!x || !wt || !y_nd || !shape_ok(n, d, h, w)) return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d * h * w;
  if (P % 64 != 0 || cin % 16 != 0 || cin < 16 || cout % 16 != 0 || cout < 16) return HDN_EINVAL;
  if (src == kSrcUp && (d % 2 || h % 2 || w % 2)) return HDN_EINVAL;
  if (src == kSrcAdd && !x2) return HDN_EINVAL;
  if (src != kSrcPlain && src != kSrcUp && src != kSrcAdd && src != kSrcLow8) return HDN_EINVAL;
  const __bf16* xb = static_cast<const __bf16*>(x);
  const __bf16* x2b = static_cast<const __bf16*>(x2);
  const __bf16* wb = static_cast<const __bf16*>(wt);
  __bf16* yb = static_cast<__bf16*>(y_nd);
  __bf16* cb = static_cast<__bf16*>(y_cm);
  if (src == kSrcUp) return conv_fwd<kSrcUp>(xb, nullptr, wb, bias, yb, cb, n, d, h, w, cin, cout, slope, as_stream(stream));
  if (src == kSrcAdd) return conv_fwd<kSrcAdd>(xb, x2b, wb, bias, yb, cb, n, d, h, w, cin, cout, slope, as_stream(stream));
  if (src == kSrcLow8) return conv_fwd<kSrcLow8>(xb, nullptr, wb, bias, yb, cb, n, d, h, w, cin, cout, slope, as_stream(stream));
  return conv_fwd<kSrcPlain>(xb, nullptr, wb, bias, yb, cb, n, d, h, w, cin, cout, slope, as_stream(stream));
}

int hdn_conv3d_fwd(const void* x, const void* wt, const float* bias, void* y_nd, void* y_cm, int n,
                   int d, int h, int w, int cin, int cout, float slope, void* stream) {
  return hdn_conv3d_fwd_src(x, nullptr, kSrcPlain, wt, bias, y_nd, y_cm, n, d, h, w, cin, cout, slope, stream);
}

int hdn_conv3d_wgrad(const void* x_cm, const void* dz_cm, float* part, int splits, int n, int d,
                     int h, int w, int cin, int cout, void* stream) {
  long long per_split = 0;
  const int rc = wgrad_split(x_cm, dz_cm, splits, n, d, h, w, cin, cout, &per_split);
  if (rc != HDN_OK || !part) return rc != HDN_OK ? rc : HDN_EINVAL;
  launch_wgrad(x_cm, dz_cm, part, splits, n, d, h, w, cin, cout, per_split, as_stream(stream));
  return launch_status();
}

int hdn_reduce_splits(const float* part, float* out, long long len, int splits, void* stream) {
  if (!part || !out || len < 1 || splits < 1) return HDN_EINVAL;
  launch_reduce(part, out, len, splits, as_stream(stream));
  return launch_status();
}

int hdn_conv3d_wgrad_noalloc(const void* x_cm, const void* dz_cm, float* grad, int splits, int n,
                             int d, int h, int w, int cin, int cout, void* stream) {
  long long per_split = 0;
  const int rc = wgrad_split(x_cm, dz_cm, splits, n, d, h, w, cin, cout, &per_split);
  if (rc != HDN_OK || !grad) return rc != HDN_OK ? rc : HDN_EINVAL;
  const long long len = 27LL * cout * cin;
  if (static_cast<long long>(splits) * len > HDN_SCRATCH_FLOATS) return HDN_ESIZE;
  float* part = nullptr;
  const int srv = scratch_for_current_device(stream, &part);
  if (srv != HDN_OK) return srv;
  // Every partial is written before it is read, so the array needs no clearing.
  const hipStream_t s = as_stream(stream);
  launch_wgrad(x_cm, dz_cm, part, splits, n, d, h, w, cin, cout, per_split, s);
  launch_reduce(part, grad, len, splits, s);
  return launch_status();
}

int hdn_channel_sum(const void* cm, float* part, long long P, int C, int splits, void* stream) {
  if (!cm || !part || P < 1 || C < 1 || splits < 1) return HDN_EINVAL;
  const long long per_split = (P + splits - 1) / splits;
  hipLaunchKernelGGL(k_channel_sum, dim3(static_cast<unsigned int>(C), static_cast<unsigned int>(splits)),
                     dim3(256), 0, as_stream(stream), static_cast<const __bf16*>(cm), part, P,
                     per_split);
  return launch_status();
}

int hdn_act_bwd(const void* y, const void* dy, void* dz_nd, void* dz_cm, long long P, int C,
                float slope, void* stream) {
  if (!y || !dy || !dz_nd || P < 1 || C % 8 != 0) return HDN_EINVAL;
  hipLaunchKernelGGL(k_act_bwd, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(y), static_cast<const __bf16*>(dy),
                     static_cast<__bf16*>(dz_nd), static_cast<__bf16*>(dz_cm), P, C, slope);
  return launch_status();
}

int hdn_add(const void* a, const void* b, void* y_nd, void* y_cm, long long P, int C, void* stream) {
  if (!a || !b || !y_nd || P < 1 || C % 8 != 0) return HDN_EINVAL;
  hipLaunchKernelGGL(k_add, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(a), static_cast<const __bf16*>(b),
                     static_cast<__bf16*>(y_nd), static_cast<__bf16*>(y_cm), P, C);
  return launch_status();
}

int hdn_pool_fwd(const void* x, void* y_nd, void* y_cm, int n, int d, int h, int w, int C,
                 void* stream) {
  if (!x || !y_nd || !shape_ok(n, d, h, w) || C % 8 != 0) return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d * h * w;
  hipLaunchKernelGGL(k_pool_fwd, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(x), static_cast<__bf16*>(y_nd),
                     static_cast<__bf16*>(y_cm), n, d, h, w, C);
  return launch_status();
}

int hdn_pool_bwd(const void* dy, void* dx, int n, int d2, int h2, int w2, int C, void* stream) {
  if (!dy || !dx || !shape_ok(n, d2, h2, w2) || C % 8 != 0 || d2 % 2 || h2 % 2 || w2 % 2)
    return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d2 * h2 * w2;
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(dy), static_cast<__bf16*>(dx), n, d2, h2, w2, C);
  return launch_status();
}

int hdn_up_fwd(const void* x, void* y_nd, void* y_cm, int n, int d2, int h2, int w2, int C,
               void* stream) {
  if (!x || !y_nd || !shape_ok(n, d2, h2, w2) || C % 8 != 0 || d2 % 2 || h2 % 2 || w2 % 2)
    return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d2 * h2 * w2;
  hipLaunchKernelGGL(k_up_fwd, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(x), static_cast<__bf16*>(y_nd),
                     static_cast<__bf16*>(y_cm), n, d2, h2, w2, C);
  return launch_status();
}

int hdn_up_bwd(const void* dy, void* dx, int n, int d, int h, int w, int C, void* stream) {
  if (!dy || !dx || !shape_ok(n, d, h, w) || C % 8 != 0) return HDN_EINVAL;
  const long long P = static_cast<long long>(n) * d * h * w;
  hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(P * (C / 8), 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(dy), static_cast<__bf16*>(dx), n, d, h, w, C);
  return launch_status();
}

int hdn_loss(const void* out, const void* inp, int c_out, int c_in, const float* target,
             const float* weight, void* dout_nd, void* dout_cm, float* pred, float* partial,
             long long P, float grad_scale, void* stream) {
  if (!out || !inp || P < 1 || c_out < 1 || c_in < 1) return HDN_EINVAL;
  if ((dout_nd || partial) && !target) return HDN_EINVAL;
  hipLaunchKernelGGL(k_loss, dim3(blocks_for(P, 256)), dim3(256), 0, as_stream(stream),
                     static_cast<const __bf16*>(out), static_cast<const __bf16*>(inp), c_out, c_in,
                     target, weight, static_cast<__bf16*>(dout_nd), static_cast<__bf16*>(dout_cm),
                     pred, partial, P, grad_scale);
  return launch_status();
}

int hdn_prepare_weights(const float* master, void* wf, void* wd, int cin, int cout, void* stream) {
  if (!master || !wf || !wd || cin < 1 || cout < 1) return HDN_EINVAL;
  hipLaunchKernelGGL(k_prepare_weights, dim3(blocks_for(27LL * cin * cout, 256)), dim3(256), 0,
                     as_stream(stream), master, static_cast<__bf16*>(wf), static_cast<__bf16*>(wd),
                     cin, cout);
  return launch_status();
}

int hdn_adam(float* param, const float* grad, float* m, float* v, long long len, float lr,
             float beta1, float beta2, float eps, int step, float weight_decay, void* stream) {
  if (!param || !grad || !m || !v || len < 1 || step < 1) return HDN_EINVAL;
  const float corr1 = 1.0f - std::pow(beta1, static_cast<float>(step));
  const float corr2 = 1.0f - std::pow(beta2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(len, 256)), dim3(256), 0, as_stream(stream), param, grad,
                     m, v, len, lr, beta1, beta2, eps, corr1, corr2, weight_decay);
  return launch_status();
}

}  // extern "C"
