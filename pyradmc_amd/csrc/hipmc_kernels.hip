// Transport, tally, fold and truncation kernels of the MI355X engine, and the
// C ABI that drives them (include/hipmc.h).
//
// Layout in HBM:
//   voxel maps      density f32 [nx][ny][nz], material u8 [nx][ny][nz]
//   tables          float4 rows per material: photon (compton, photo, pair, rayleigh),
//                   electron A (csda, moller, restricted range, radiative),
//                   electron B (scattering power, log eta, restricted stopping, 0)
//   GS grid         f32 [n_eta][n_theta2][512]
//   particles       48-byte records (3 x 16 B) in four ping-pong queues
//   tallies         int64 quanta [column][scoring voxel]; f64 sums s1, s2 alike
//
// Transport runs in rounds like the reference's queue ping-pong
// (pyradmc/backends/warp/engine.py:2732-2831): a photon round consumes the photon
// queue, an electron round the electron queue, each pushing secondaries into the
// other or the next round's queue. Within a round every lane is persistent and
// refills itself: it transports its particle one Woodcock flight or one condensed
// substep per loop iteration and pops the next particle the moment its own ends,
// so a wave does not idle behind its longest history. Ledger sums are kept per
// lane in registers and flushed once per lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <type_traits>
#include <vector>

#include "../../include/hipmc.h"
#include "hipmc_collimation.h"
#include "hipmc_device.h"

using namespace hmc;

namespace {

// ------------------------------------------------------------------- particles

struct __align__(16) Particle {
  float x, y, z, e;
  float ux, uy, uz, w;
  uint64_t rng;
  uint32_t meta;  // kind (2 bits) | primary (1 bit) << 2 | column << 3
  uint32_t pad;
};
static_assert(sizeof(Particle) == 48, "particle record must stay 48 bytes");

__host__ __device__ inline uint32_t pack_meta(uint32_t kind, uint32_t primary, uint32_t column) {
  return kind | (primary << 2) | (column << 3);
}

// A particle queue: records, the shared count (reserved slots), capacity, and - in a
// round kernel - this wave's chunk state in LDS ({next, end} of the slots the wave has
// reserved but not yet filled; see chunk_reserve).
struct Queue {
  Particle* data;
  unsigned int* count;
  unsigned int capacity;
  unsigned int push_chunk = 64;  // slots a wave reserves per atomic (multiple of 64)
  int lds = -1;                  // which of the wave's two output-chunk states (LDS) is this queue's
};

// Per-wave state of a round kernel in LDS, at file scope so that the device functions
// address it directly (LDS instructions; a pointer to it kept in a kernel argument
// would become a generic address): the two output queues' current chunks
// [wave][queue][next, end], and the block's ledger reduction slots [book][wave]
// (books 6 and 7: the wave's pops and pushes).
__shared__ unsigned int g_chunk[4][2][2];
__shared__ long long g_books[8 * 4];

// Accesses to that state: relaxed atomics at wave scope, so that every access is a real
// LDS access (lanes of divergent branches see one another's updates) and the compiler
// still knows the address space.
__device__ inline unsigned int lds_load(unsigned int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ inline void lds_store(unsigned int* p, unsigned int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ inline void lds_add(long long* p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Kind tag of a padding record: a slot a producer wave reserved but did not fill.
constexpr uint32_t kEmptyKind = 3u;

// Slot reservation for the lanes of one wave that want one: a single atomic per
// wave, ranks by lane order. Must be reached by every active lane of the wave.
__device__ inline unsigned int wave_reserve(unsigned int* counter, bool want) {
  const unsigned long long mask = __ballot(want ? 1 : 0);
  if (mask == 0ull) return 0u;
  const int lane = static_cast<int>(__lane_id());
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned int base = 0u;
  if (lane == leader) base = atomicAdd(counter, static_cast<unsigned int>(__popcll(mask)));
  base = __shfl(base, leader);
  return base + static_cast<unsigned int>(__popcll(mask & ((1ull << lane) - 1ull)));
}

__device__ inline long long wave_sum(long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Queue indices handed to the idle lanes of a wave, in blocks of R (the refill block,
// a multiple of 64 set per transport call) given to the idle lanes in lane order. A
// wave's first block is its own by position (wave w of the grid takes [R w, R w + R)),
// with no atomic; later blocks come from one atomic per block on the queue head,
// offset past those static blocks. So the single
// head counter sees one atomic per block instead of one per particle, and a wave that
// starts beyond an exhausted queue exits without touching it: a round can launch at
// full width whatever its queue holds. (Reading the head before claiming, instead,
// queues that read behind the head's atomics in L2 and costs far more than it saves.)
// Called by every lane still in the round loop, at the loop head (converged); the
// pool state is therefore identical in all of them. Which lane transports which
// particle does not change any result: each particle carries its own random stream
// and the tallies are integer sums.
struct RefillPool {
  unsigned int next, end;
  unsigned int block;     // R
  unsigned int n_static;  // indices below this are the waves' static first blocks
  // Every wave of the grid owns one static block.
  __device__ explicit RefillPool(unsigned int r)
      : next((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * r),
        end((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * r + r),
        block(r),
        n_static(gridDim.x * (blockDim.x >> 6) * r) {}
};

__device__ inline unsigned int refill_index(RefillPool& pool, unsigned int* head, bool idle) {
  const unsigned long long mask = __ballot(idle ? 1 : 0);
  if (mask == 0ull) return 0u;
  const unsigned int want = static_cast<unsigned int>(__popcll(mask));
  const int lane = static_cast<int>(__lane_id());
  const unsigned int rank =
      static_cast<unsigned int>(__popcll(mask & ((1ull << lane) - 1ull)));
  const unsigned int avail = pool.end - pool.next;
  unsigned int idx;
  if (want <= avail) {
    idx = pool.next + rank;
    pool.next += want;
  } else {
    const unsigned long long present = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(present)) - 1;
    unsigned int base = 0u;
    if (lane == leader) base = pool.n_static + atomicAdd(head, pool.block);
    base = __shfl(base, leader);
    idx = rank < avail ? pool.next + rank : base + (rank - avail);
    pool.next = base + (want - avail);
    pool.end = base + pool.block;
  }
  return idx;
}

enum LedgerSlot {
  L_ESCAPED = 0,
  L_UNSCORED,
  L_DEPOSITED,
  L_EMITTED,
  L_VIOLATIONS,
  L_TRAPPED,
  L_OVERFLOW,  // largest queue count seen above capacity (atomicMax), 0 if none
  // per-kernel event counts, photon round then electron round: voxel fetches,
  // tally atomics, steps, particles popped (padding excluded), particles pushed
  L_P_VOX,
  L_P_DEP,
  L_P_STEPS,
  L_P_POP,
  L_P_PUSH,
  L_E_VOX,
  L_E_DEP,
  L_E_STEPS,
  L_E_POP,
  L_E_PUSH,
  L_COUNT
};

// Word offsets in the context's page-locked readback block.
constexpr int kStageLedger = 2;                     // after four u32 queue counts
constexpr int kStageFolded = kStageLedger + L_COUNT;
constexpr int kStageWords = kStageFolded + 1;

struct TransportParams {
  DevGrid g;
  DevScoring s;
  DevTables t;
  long long* edep;
  unsigned long long* ledger;
  float pcut, ecut, step_fraction, resolution;
  float log_pcut;  // logf(pcut), for the bremsstrahlung spectrum
  int transport_electrons, msc_gs, dose_to_water, max_steps;
  unsigned int refill_block;  // queue indices a wave claims at a time (multiple of 64)
};

// Per-lane books kept in registers and flushed once per lane at the end of a round;
// event counts are 32-bit (a lane's count within one launch stays far below 2^32).
// Rare events (majorant violations, trapped particles) go straight to the ledger.
struct LaneTally {
  long long escaped = 0, unscored = 0, deposited = 0;
  unsigned int vox = 0, dep = 0, steps = 0;
  // One pending tally cell per lane: consecutive deposits into the same cell are
  // summed here and written with one atomic when the cell changes (integer
  // addition is associative, so the maps are unchanged bit for bit).
  long long pend_cell = -1;
  long long pend_q = 0;

  __device__ void flush_pending(long long* edep) {
    if (pend_q != 0) {
      atomicAdd(reinterpret_cast<unsigned long long*>(edep + pend_cell),
                static_cast<unsigned long long>(pend_q));
      ++dep;
    }
    pend_q = 0;
    pend_cell = -1;
  }

  // base = L_P_VOX for the photon round, L_E_VOX for the electron round. Every thread
  // of the block must call this after its lane loop (it synchronises the block): the
  // books are summed over the wave, then over the block in LDS (red: 8 x 4 slots; the
  // wave's pop and push counts are already in red[6*4 + wave] and red[7*4 + wave]),
  // and written with one atomic per book per block.
  __device__ void flush(unsigned long long* ledger, int base, long long* red) const {
    const long long books[6] = {escaped, unscored, deposited, static_cast<long long>(vox),
                                static_cast<long long>(dep), static_cast<long long>(steps)};
    const int wave = static_cast<int>(threadIdx.x >> 6);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const long long total = wave_sum(books[k]);
      if (__lane_id() == 0) red[k * 4 + wave] = total;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
      const int slot[8] = {L_ESCAPED, L_UNSCORED, L_DEPOSITED, base,
                           base + 1,  base + 2,   base + 3,    base + 4};
      long long total = 0;
      for (int wv = 0; wv < static_cast<int>(blockDim.x >> 6); ++wv) total += red[threadIdx.x * 4 + wv];
      if (total != 0)
        atomicAdd(ledger + slot[threadIdx.x], static_cast<unsigned long long>(total));
    }
  }
};

// Queue slots for the lanes of a wave that push at one call site. A wave reserves
// q.push_chunk slots at a time with one atomic on the shared count and hands them out
// in lane order, finishing its current chunk before starting the next, so the only
// unfilled slots are each wave's tail at the end of the round (seal_chunk pads them).
// The chunk state lives in LDS, so the lanes executing a divergent call site always
// see the wave's current state; only the first active lane reads and writes it.

__device__ inline unsigned int chunk_reserve(const Queue& q) {
  const unsigned long long mask = __ballot(1);
  const int lane = static_cast<int>(__lane_id());
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  const unsigned int cnt = static_cast<unsigned int>(__popcll(mask));
  const unsigned int rank =
      static_cast<unsigned int>(__popcll(mask & ((1ull << lane) - 1ull)));
  unsigned int next = 0u, avail = 0u, fresh = 0u;
  if (lane == leader) {
    const int wave = static_cast<int>(threadIdx.x >> 6);
    unsigned int* chunk = g_chunk[wave][q.lds];
    lds_add(&g_books[7 * 4 + wave], cnt);
    next = lds_load(chunk);
    avail = lds_load(chunk + 1) - next;
    if (cnt <= avail) {
      lds_store(chunk, next + cnt);
    } else {
      fresh = atomicAdd(q.count, q.push_chunk);  // cnt - avail <= 64 <= q.push_chunk
      lds_store(chunk, fresh + (cnt - avail));
      lds_store(chunk + 1, fresh + q.push_chunk);
    }
  }
  next = __shfl(next, leader);
  avail = __shfl(avail, leader);
  fresh = __shfl(fresh, leader);
  return rank < avail ? next + rank : fresh + (rank - avail);
}

// Pad the wave's unfilled reserved slots with empty records. Every lane of the wave
// calls this after its round loop.
__device__ inline void seal_chunk(const Queue& q) {
  unsigned int* chunk = g_chunk[threadIdx.x >> 6][q.lds];
  const unsigned int next = lds_load(chunk), end = lds_load(chunk + 1);
  for (unsigned int sl = next + __lane_id(); sl < end; sl += 64u)
    if (sl < q.capacity) q.data[sl].meta = kEmptyKind;
}

// Add the lanes (of those executing the call) that popped a real particle to the
// wave's pop count in LDS.
__device__ inline void count_pops(bool real) {
  const unsigned long long mask = __ballot(real ? 1 : 0);
  const unsigned long long active = __ballot(1);
  if (static_cast<int>(__lane_id()) == __ffsll(static_cast<long long>(active)) - 1) {
    lds_add(&g_books[6 * 4 + (threadIdx.x >> 6)], __popcll(mask));
  }
}

// Append one particle (chunk_reserve over the lanes executing the call).
__device__ inline void push(const Queue& q, uint32_t kind, uint32_t column, float e, float w,
                            float x, float y, float z, float ux, float uy, float uz,
                            uint64_t rng) {
  const unsigned int idx = chunk_reserve(q);
  if (idx < q.capacity) {
    Particle p;
    p.x = x;
    p.y = y;
    p.z = z;
    p.e = e;
    p.ux = ux;
    p.uy = uy;
    p.uz = uz;
    p.w = w;
    p.rng = rng;
    p.meta = pack_meta(kind, 0u, column);
    p.pad = 0u;
    q.data[idx] = p;
  }
}

// The round kernels come in two builds. kFast is the common configuration: dose to
// medium (no stopping-power weight), no deposit spreading, and a scoring grid that is
// the transport grid, so a position inside the transport grid needs no second
// containment test. The general build covers everything else; both give the same
// bits where both apply.

// Dose-to-water weight of a deposit (1 under dose to medium).
template <bool kFast>
__device__ inline float spr(const TransportParams& P, float e, int mat) {
  if constexpr (kFast) return 1.0f;
  else return spr_factor(P.t, P.dose_to_water, e, mat, P.ecut);
}

// Fixed-point deposit at a position inside the transport grid, routed into the
// scoring grid or the unscored book (pyradmc/backends/warp/kernels.py:535-584). As there,
// quanta are x + 1/2 truncated toward zero (an amount a rounding step left slightly
// negative rounds toward zero, unlike an escape's floor).
template <bool kFast>
__device__ inline void deposit(const TransportParams& P, LaneTally& tal, long long base, float x,
                               float y, float z, float amount, float factor) {
  const long long idx = kFast ? grid_index(P.s, x, y, z) : scoring_index(P.s, x, y, z);
  if (idx >= 0) {
    const long long q = whole_to_ll(trunc(static_cast<double>(amount) * static_cast<double>(factor) *
                                          kInvQuantum + 0.5));
    const long long cell = base + idx;
    if (cell != tal.pend_cell) {
      tal.flush_pending(P.edep);
      tal.pend_cell = cell;
    }
    tal.pend_q += q;
    if (!kFast && P.dose_to_water)
      tal.deposited += whole_to_ll(trunc(static_cast<double>(amount) * kInvQuantum + 0.5));
  } else {
    tal.unscored += whole_to_ll(trunc(static_cast<double>(amount) * kInvQuantum + 0.5));
  }
}

__device__ inline void escape(LaneTally& tal, float amount) {
  tal.escaped += quanta_round(static_cast<double>(amount));
}

// Deposit if inside the transport grid, else escaped (kernels.py:600-633).
template <bool kFast>
__device__ inline void deposit_or_escape(const TransportParams& P, LaneTally& tal, long long base,
                                         float x, float y, float z, float amount, float factor) {
  if (amount <= 0.0f) return;
  if (inside_grid(P.g, x, y, z)) deposit<kFast>(P, tal, base, x, y, z, amount, factor);
  else escape(tal, amount);
}

// One half-substep's continuous loss, in equal pieces no longer than the
// resolution (pyradmc/transport/electron.py:169-191, 361-380).
template <bool kFast>
__device__ inline void deposit_spread(const TransportParams& P, LaneTally& tal, long long base,
                                      float x, float y, float z, float ux, float uy, float uz,
                                      float length, float amount, float factor) {
  if constexpr (kFast) {  // one piece: the midpoint takes the whole amount
    const float t = 0.5f * length;
    deposit_or_escape<kFast>(P, tal, base, x + ux * t, y + uy * t, z + uz * t, amount, factor);
    return;
  }
  int n = 1;
  if (P.resolution > 0.0f) {
    n = static_cast<int>(ceilf(length / P.resolution));
    n = min(max(n, 1), kSubstepDepositCap);
  }
  const float share = amount / static_cast<float>(n);
  const float piece = length / static_cast<float>(n);
  for (int k = 0; k < n; ++k) {
    const float t = (static_cast<float>(k) + 0.5f) * piece;
    deposit_or_escape<kFast>(P, tal, base, x + ux * t, y + uy * t, z + uz * t, share, factor);
  }
}

// Positron annihilation at rest: two back-to-back 511 keV photons, isotropic
// (pyradmc/transport/particles.py:75-102).
template <bool kFast>
__device__ inline void annihilate(const TransportParams& P, LaneTally& tal, const Queue& qp,
                                  long long base, uint32_t column, float x, float y, float z,
                                  float w, float factor, uint64_t& rng) {
  if (P.pcut >= kElectronMass) {
    deposit<kFast>(P, tal, base, x, y, z, w * 2.0f * kElectronMass, factor);
    return;
  }
  float ax, ay, az;
  sample_isotropic(rng, ax, ay, az);
  const uint64_t r1 = spawn_stream(rng);
  push(qp, PHOTON, column, kElectronMass, w, x, y, z, ax, ay, az, r1);
  const uint64_t r2 = spawn_stream(rng);
  push(qp, PHOTON, column, kElectronMass, w, x, y, z, -ax, -ay, -az, r2);
}

__device__ inline void load_particle(const Particle* data, unsigned int i, float& x, float& y,
                                     float& z, float& e, float& ux, float& uy, float& uz,
                                     float& w, uint64_t& rng, uint32_t& meta) {
  const float4* p = reinterpret_cast<const float4*>(data + i);
  const float4 a = p[0];
  const float4 b = p[1];
  const uint4 c = reinterpret_cast<const uint4*>(p)[2];
  x = a.x;
  y = a.y;
  z = a.z;
  e = a.w;
  ux = b.x;
  uy = b.y;
  uz = b.z;
  w = b.w;
  rng = static_cast<uint64_t>(c.x) | (static_cast<uint64_t>(c.y) << 32);
  meta = c.z;
}

// Vacuum flight to the grid for a particle born outside; false when it misses
// (pyradmc/backends/warp/kernels.py:1214-1229).
__device__ inline bool enter_grid(const DevGrid& g, float& x, float& y, float& z, float ux,
                                  float uy, float uz) {
  if (inside_grid(g, x, y, z)) return true;
  float t = slab_entry(g, x, y, z, ux, uy, uz);
  if (t >= kInfinityCm) return false;
  t += kEntryNudge;
  x += t * ux;
  y += t * uy;
  z += t * uz;
  return inside_grid(g, x, y, z);
}

// ----------------------------------------------------------------- photon round
//
// Woodcock flight on the majorant, real/virtual test, channel selection; Compton
// continues in-lane, photo and pair hand electrons to the electron queue
// (pyradmc/transport/photon.py:161-282, kernels.py:1171-1529).
// Waves per SIMD the round kernels are register-budgeted for (a tuning constant:
// more resident waves hide more of the dependent-load latency of a transport step).
#ifndef HMC_ROUND_WAVES
#define HMC_ROUND_WAVES 5
#endif

// Idle lanes of a wave in the electron round that make a refill worth its pass (of 64;
// measured on the bench's slab step: 8 and 32 were slower than 16, and the photon
// round gained nothing from the same rule).
constexpr int kRefillIdleLanes = 16;

// Energy bins of the electron queue's sort (a quarter octave each above the cutoff),
// and the blocks that sort it, each over one contiguous range of the queue.
constexpr int kEnergyBins = 32;
constexpr int kSortBlocks = 1024;

template <bool kFast>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HMC_ROUND_WAVES, 8)))
k_photon_round(TransportParams P, Queue qin,
                                                      unsigned int* head, Queue qe, Queue qp) {
  const unsigned int n_in = min(*qin.count, qin.capacity);
  LaneTally tal;
  bool active = false;
  float x = 0, y = 0, z = 0, e = 0, ux = 0, uy = 0, uz = 1, w = 0;
  uint64_t rng = 0;
  uint32_t column = 0;
  long long base = 0;
  float mu_maj = 0.0f;
  int pi = 0;
  float pw = 0.0f;
  int steps = 0;
  const int n = P.t.n_points;

  // Per-wave chunk state of the two output queues (chunk_reserve), and the block's
  // ledger reduction slots.
  const int wave_in_block = static_cast<int>(threadIdx.x >> 6);
  if (__lane_id() < 4) g_chunk[wave_in_block][__lane_id() >> 1][__lane_id() & 1] = 0u;
  if (__lane_id() < 2) g_books[(6 + __lane_id()) * 4 + wave_in_block] = 0;
  qe.lds = 0;
  qp.lds = 1;
  RefillPool pool(P.refill_block);
  while (true) {
    const unsigned int i = refill_index(pool, head, !active);
    if (!active) {
      if (i >= n_in) break;
      uint32_t meta;
      load_particle(qin.data, i, x, y, z, e, ux, uy, uz, w, rng, meta);
      const bool real = (meta & 3u) != kEmptyKind;  // else padding left by a producer wave
      count_pops(real);
      if (!real) continue;
      column = meta >> 3;
      base = static_cast<long long>(column) * P.s.n_voxels;
      if (!enter_grid(P.g, x, y, z, ux, uy, uz)) {
        escape(tal, w * e);
        continue;
      }
      active = true;
      steps = 0;
      grid_locate(__logf(e), P.t.p_log_min, P.t.p_inv_dlog, n, pi, pw);
      mu_maj = lerp1(P.t.majorant, pi, pw);
    }
    tal.steps++;
    if (++steps > P.max_steps) {
      atomicAdd(P.ledger + L_TRAPPED, 1ull);
      escape(tal, w * e);
      active = false;
      continue;
    }

    const float step = sample_path(mu_maj, rng);
    x += step * ux;
    y += step * uy;
    z += step * uz;
    if (!inside_grid(P.g, x, y, z)) {
      escape(tal, w * e);
      active = false;
      continue;
    }
    const int ix = axis_index(x, P.g.x_lo, P.g.inv_sx, P.g.nx);
    const int iy = axis_index(y, P.g.y_lo, P.g.inv_sy, P.g.ny);
    const int iz = axis_index(z, P.g.z_lo, P.g.inv_sz, P.g.nz);
    const long long vox = (static_cast<long long>(ix) * P.g.ny + iy) * P.g.nz + iz;
    const float rho = P.g.density[vox];
    const int mat = P.g.material[vox];
    tal.vox++;
    const float4 c = lerp4(P.t.photon + static_cast<long long>(mat) * n, pi, pw);
    const float mu_c = rho * c.x;
    const float mu_p = rho * c.y;
    const float mu_pp = rho * c.z;
    const float mu_r = rho * c.w;
    const float mu_real = mu_c + mu_p + mu_pp + mu_r;
    if (mu_real > mu_maj) {
      atomicAdd(P.ledger + L_VIOLATIONS, 1ull);
      escape(tal, w * e);
      active = false;
      continue;
    }
    if (uniform(rng) * mu_maj >= mu_real) continue;  // delta scattering

    const int process = select_process(mu_c, mu_p, mu_pp, mu_r, rng);

    if (process == RAYLEIGH) {
      const float cos_r = sample_coherent(P.t, mat, e, rng);
      const float phi = kTwoPi * uniform(rng);
      rotate(ux, uy, uz, cos_r, phi);
    } else if (process == COMPTON) {
      const float ratio = sample_compton_ratio(e, rng);
      const float recoil = e * (1.0f - ratio);
      const float phi = kTwoPi * uniform(rng);
      if (P.transport_electrons && recoil > P.ecut) {
        const float ce = compton_electron_cos(e, ratio);
        float ex = ux, ey = uy, ez = uz;
        rotate(ex, ey, ez, ce, phi + kPi);
        const uint64_t child = spawn_stream(rng);
        push(qe, ELECTRON, column, recoil, w, x, y, z, ex, ey, ez, child);
      } else {
        deposit<kFast>(P, tal, base, x, y, z, w * recoil,
                spr<kFast>(P, recoil, mat));
      }
      const float cg = compton_cos(e, ratio);
      rotate(ux, uy, uz, cg, phi);
      e *= ratio;
      if (e <= P.pcut) {
        deposit<kFast>(P, tal, base, x, y, z, w * e, spr<kFast>(P, e, mat));
        active = false;
        continue;
      }
      if (e < kRouletteMeV && w < kRouletteWeightCap) {
        const float new_w = uniform(rng) < kRouletteSurvival ? w / kRouletteSurvival : 0.0f;
        escape(tal, (w - new_w) * e);
        if (new_w == 0.0f) {
          active = false;
          continue;
        }
        w = new_w;
      }
      grid_locate(__logf(e), P.t.p_log_min, P.t.p_inv_dlog, n, pi, pw);
      mu_maj = lerp1(P.t.majorant, pi, pw);
    } else if (process == PHOTOELECTRIC) {
      if (P.transport_electrons && e > P.ecut) {
        const uint64_t child = spawn_stream(rng);
        push(qe, ELECTRON, column, e, w, x, y, z, ux, uy, uz, child);
      } else {
        deposit<kFast>(P, tal, base, x, y, z, w * e, spr<kFast>(P, e, mat));
      }
      active = false;
    } else {  // pair production
      const float kinetic = e - 2.0f * kElectronMass;
      const float f_ann = spr<kFast>(P, kElectronMass, mat);
      if (P.transport_electrons) {
        const float fraction = uniform(rng);
        const float share_e = fraction * kinetic;
        const float share_p = (1.0f - fraction) * kinetic;
        if (share_e > P.ecut) {
          const uint64_t child = spawn_stream(rng);
          push(qe, ELECTRON, column, share_e, w, x, y, z, ux, uy, uz, child);
        } else {
          deposit<kFast>(P, tal, base, x, y, z, w * share_e,
                  spr<kFast>(P, share_e, mat));
        }
        if (share_p > P.ecut) {
          const uint64_t child = spawn_stream(rng);
          push(qe, POSITRON, column, share_p, w, x, y, z, ux, uy, uz, child);
        } else {
          deposit<kFast>(P, tal, base, x, y, z, w * share_p,
                  spr<kFast>(P, share_p, mat));
          annihilate<kFast>(P, tal, qp, base, column, x, y, z, w, f_ann, rng);
        }
      } else {
        deposit<kFast>(P, tal, base, x, y, z, w * kinetic,
                spr<kFast>(P, kinetic, mat));
        annihilate<kFast>(P, tal, qp, base, column, x, y, z, w, f_ann, rng);
      }
      active = false;
    }
  }
  seal_chunk(qe);
  seal_chunk(qp);
  tal.flush_pending(P.edep);
  tal.flush(P.ledger, L_P_VOX, g_books);
}

// --------------------------------------------------------------- electron round
//
// Class-II condensed-history substep: range/boundary/Moller-limited step, exact
// CSDA loss through the restricted range and its inverse, Bernoulli brems, random
// hinge with GS (or Gaussian) deflection, two half-step deposits, Moller delta
// (pyradmc/transport/electron.py:292-481, kernels.py:1532-1885).
template <bool kFast>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HMC_ROUND_WAVES, 8)))
k_electron_round(TransportParams P, Queue qin,
                                                        unsigned int* head, Queue qp, Queue qe) {
  const unsigned int n_in = min(*qin.count, qin.capacity);
  LaneTally tal;
  bool active = false;
  float x = 0, y = 0, z = 0, e = 0, ux = 0, uy = 0, uz = 1, w = 0;
  uint64_t rng = 0;
  uint32_t column = 0;
  long long base = 0;
  bool positron = false;
  int steps = 0;
  const int n = P.t.n_points;

  // Per-wave chunk state of the two output queues (chunk_reserve), and the block's
  // ledger reduction slots.
  const int wave_in_block = static_cast<int>(threadIdx.x >> 6);
  if (__lane_id() < 4) g_chunk[wave_in_block][__lane_id() >> 1][__lane_id() & 1] = 0u;
  if (__lane_id() < 2) g_books[(6 + __lane_id()) * 4 + wave_in_block] = 0;
  qp.lds = 0;
  qe.lds = 1;
  RefillPool pool(P.refill_block);
  while (true) {
    // Idle lanes take new particles together, once kRefillIdleLanes of the wave's
    // looping lanes are idle (or all of them): the refill block then runs for more
    // lanes at a time and in fewer passes. Each particle's history is its own, so
    // the results do not depend on when a lane takes it.
    const int n_idle = __popcll(__ballot(!active));
    const bool refill_now = !active && (n_idle >= kRefillIdleLanes || n_idle == __popcll(__ballot(1)));
    const unsigned int i = refill_index(pool, head, refill_now);
    if (!active) {
      if (!refill_now) continue;
      if (i >= n_in) break;
      uint32_t meta;
      load_particle(qin.data, i, x, y, z, e, ux, uy, uz, w, rng, meta);
      const bool real = (meta & 3u) != kEmptyKind;  // else padding left by a producer wave
      count_pops(real);
      if (!real) continue;
      column = meta >> 3;
      positron = (meta & 3u) == POSITRON;
      base = static_cast<long long>(column) * P.s.n_voxels;
      const float latent = positron ? 2.0f * kElectronMass : 0.0f;
      if (!enter_grid(P.g, x, y, z, ux, uy, uz)) {
        escape(tal, w * (e + latent));
        continue;
      }
      active = true;
      steps = 0;
    }
    tal.steps++;
    const float latent = positron ? 2.0f * kElectronMass : 0.0f;
    if (++steps > P.max_steps) {
      atomicAdd(P.ledger + L_TRAPPED, 1ull);
      escape(tal, w * (e + latent));
      active = false;
      continue;
    }

    const int ix = axis_index(x, P.g.x_lo, P.g.inv_sx, P.g.nx);
    const int iy = axis_index(y, P.g.y_lo, P.g.inv_sy, P.g.ny);
    const int iz = axis_index(z, P.g.z_lo, P.g.inv_sz, P.g.nz);
    const long long vox = (static_cast<long long>(ix) * P.g.ny + iy) * P.g.nz + iz;
    const int mat = P.g.material[vox];
    tal.vox++;

    if (e <= P.ecut) {
      // Terminal: local deposit, annihilation for a positron.
      deposit<kFast>(P, tal, base, x, y, z, w * e, spr<kFast>(P, e, mat));
      if (positron)
        annihilate<kFast>(P, tal, qp, base, column, x, y, z, w,
                   spr<kFast>(P, kElectronMass, mat), rng);
      active = false;
      continue;
    }
    const float rho = P.g.density[vox];
    const long long row = static_cast<long long>(mat) * n;

    int ei;
    float ew;
    const float log_e = __logf(e);
    grid_locate(log_e, P.t.e_log_min, P.t.e_inv_dlog, n, ei, ew);
    const float4 ea = lerp4(P.t.elec_a + row, ei, ew);  // csda, moller, r_range, radiative

    const float range_cm = ea.x / rho;
    float s_max = P.step_fraction * range_cm;
    if (!P.msc_gs) {
      const float tp = lerp4(P.t.elec_b + row, ei, ew).x;
      s_max = fminf(s_max, kHingeTheta2Max / (tp * rho));
    }
    const float s_boundary = boundary_distance(P.g, x, y, z, ux, uy, uz) + kBoundaryNudge;
    const float s_geometry = fminf(s_max, s_boundary);
    const float sigma_moller = rho * ea.y;
    float s = s_geometry;
    bool moller_pending = false;
    if (sigma_moller > 0.0f) {
      const float s_int = sample_path(sigma_moller, rng);
      if (s_int <= s_geometry) {
        s = s_int;
        moller_pending = true;
      }
    }

    const float available = ea.z;
    float de;
    if (rho * s >= available) {
      s = available / rho;
      de = e - P.ecut;
      moller_pending = false;
    } else {
      int ri;
      float rw;
      grid_locate(__logf(available - rho * s), P.t.r_log_min, P.t.r_inv_dlog, n, ri, rw);
      float e_end = lerp1(P.t.inv_range + row, ri, rw);
      e_end = fminf(fmaxf(e_end, P.ecut), e);
      de = e - e_end;
    }
    // Bernoulli bremsstrahlung matched to the radiative stopping power
    // (pyradmc/physics/brems.py:30-66).
    const float radiative_loss = ea.w * rho * s;
    float emit_probability, local_brems;
    brems_step(e, radiative_loss, P.pcut, log_e, P.log_pcut, emit_probability, local_brems);
    const float continuous = de + local_brems;
    const bool emit_brems = uniform(rng) < emit_probability;
    const float factor = spr<kFast>(P, e, mat);

    const float s1 = uniform(rng) * s;
    const float s2 = s - s1;
    const float d1 = s > 0.0f ? continuous * (s1 / s) : 0.0f;
    const float d2 = continuous - d1;

    deposit_spread<kFast>(P, tal, base, x, y, z, ux, uy, uz, s1, w * d1, factor);
    e -= d1;
    x += ux * s1;
    y += uy * s1;
    z += uz * s1;
    if (!inside_grid(P.g, x, y, z)) {
      escape(tal, w * (e + latent));
      active = false;
      continue;
    }

    const float log_e_mid = __logf(e);
    grid_locate(log_e_mid, P.t.e_log_min, P.t.e_inv_dlog, n, ei, ew);
    const float4 eb = lerp4(P.t.elec_b + row, ei, ew);  // scattering power, log eta
    const float mean_square = eb.x * rho * s;
    float cos_hinge;
    if (P.msc_gs) {
      const float fx = eb.y * kGsBinsPerLog;
      const float fy = __logf(fmaxf(mean_square, kGsTheta2Min)) * kGsBinsPerLog;
      cos_hinge = sample_gs(P.t, fx, fy, mean_square, rng);
    } else {
      cos_hinge = sample_gaussian_hinge(mean_square, rng);
    }
    const float phi = kTwoPi * uniform(rng);
    rotate(ux, uy, uz, cos_hinge, phi);

    if (emit_brems && e > P.pcut) {
      float k = sample_brems_energy(P.pcut, log_e_mid, P.log_pcut, rng);
      k = fminf(k, e - d2);
      if (k > P.pcut) {
        float photon_w = w;
        if (k < kRouletteMeV && w < kRouletteWeightCap) {
          photon_w = uniform(rng) < kRouletteSurvival ? w / kRouletteSurvival : 0.0f;
          escape(tal, (w - photon_w) * k);
        }
        if (photon_w > 0.0f) {
          const uint64_t child = spawn_stream(rng);
          push(qp, PHOTON, column, k, photon_w, x, y, z, ux, uy, uz, child);
        }
      } else {
        deposit_or_escape<kFast>(P, tal, base, x, y, z, w * k,
                          spr<kFast>(P, k, mat));
      }
      e -= k;
    }

    deposit_spread<kFast>(P, tal, base, x, y, z, ux, uy, uz, s2, w * d2, factor);
    e -= d2;
    x += ux * s2;
    y += uy * s2;
    z += uz * s2;
    if (!inside_grid(P.g, x, y, z)) {
      escape(tal, w * (e + latent));
      active = false;
      continue;
    }

    if (moller_pending && e > 2.0f * P.ecut) {
      const float delta = sample_moller(e, P.ecut, rng);
      float cd, cp;
      moller_cosines(e, delta, cd, cp);
      const float mphi = kTwoPi * uniform(rng);
      float dx = ux, dy = uy, dz = uz;
      rotate(dx, dy, dz, cd, mphi);
      const uint64_t child = spawn_stream(rng);
      push(qe, ELECTRON, column, delta, w, x, y, z, dx, dy, dz, child);
      rotate(ux, uy, uz, cp, mphi + kPi);
      e -= delta;
    }

    // Ended at or below the cutoff (a range-limited step ends exactly there): the
    // terminal deposit the next pass would make, made now with the same position,
    // energy, random state and books, so that the lane is idle from the next pass on
    // instead of spending it on the deposit alone. A particle the next pass would
    // count as trapped is left to it.
    if (e <= P.ecut && steps < P.max_steps) {
      tal.steps++;
      tal.vox++;
      int mat_end = 0;
      if constexpr (!kFast) {
        const int jx = axis_index(x, P.g.x_lo, P.g.inv_sx, P.g.nx);
        const int jy = axis_index(y, P.g.y_lo, P.g.inv_sy, P.g.ny);
        const int jz = axis_index(z, P.g.z_lo, P.g.inv_sz, P.g.nz);
        mat_end = P.g.material[(static_cast<long long>(jx) * P.g.ny + jy) * P.g.nz + jz];
      }
      deposit<kFast>(P, tal, base, x, y, z, w * e, spr<kFast>(P, e, mat_end));
      if (positron)
        annihilate<kFast>(P, tal, qp, base, column, x, y, z, w, spr<kFast>(P, kElectronMass, mat_end), rng);
      active = false;
    }
  }
  seal_chunk(qp);
  seal_chunk(qe);
  tal.flush_pending(P.edep);
  tal.flush(P.ledger, L_E_VOX, g_books);
}

// -------------------------------------------------------------------- generate

struct DevSource {
  int kind, particle;
  float energy;
  float px, py, pz, dx, dy, dz;
  float z0, x0, xe, y0, ye;
  float fx, fy, fz, uax, uay, uaz, vax, vay, vaz, width_u, width_v;
  float sigma_u, sigma_v;  // spot fan
  int n_bins;
  const float* edges;
  const float* cdf;
  const float* beamlet;  // lattice [n][4], fan [n][3]
  // uploaded primaries
  const int32_t* up_particle;
  const float* up;  // [9][n]: energy, x, y, z, ux, uy, uz, weight
  long long up_n;
};

struct DevItem {
  long long key_base;
  long long start;  // exclusive prefix of counts
  int column;       // flat tally column (batch slot * columns + column)
  int beamlet;      // index into the source's per-beamlet parameters
};

// Spectrum inversion: first cdf entry >= u, clamped to the top bin
// (pyradmc/geometry/spectrum.py:120-143, kernels.py:2110-2117).
__device__ inline float sample_spectrum(const DevSource& S, uint64_t& rng) {
  const float u_bin = uniform(rng);
  const float u_within = uniform(rng);
  int lo = 0, hi = S.n_bins;  // lower_bound over [0, n_bins)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (S.cdf[mid] < u_bin) lo = mid + 1;
    else hi = mid;
  }
  const int k = min(lo, S.n_bins - 1);
  const float a = S.edges[k];
  return a + (S.edges[k + 1] - a) * u_within;
}


// One primary of the source from the history's stream rng (the source's draws come
// first on it): beamlet indexes per-beamlet parameters, g is the flat primary index of
// host-sampled primaries. Mono sources leave book false: their emitted energy is
// analytic (pyradmc/backends/warp/kernels.py:1888-2401 for the generators).
struct Primary {
  float e, x, y, z, ux, uy, uz, w;
  uint32_t kind;
  bool book;
};

__device__ inline Primary make_primary(const DevSource& S, int beamlet, long long g, uint64_t& rng) {
  float e = S.energy, w = 1.0f;
  float x = 0, y = 0, z = 0, ux = 0, uy = 0, uz = 1;
  bool book = false;
  uint32_t kind = static_cast<uint32_t>(S.particle);
  switch (S.kind) {
    case HMC_SRC_PENCIL:
      x = S.px; y = S.py; z = S.pz;
      ux = S.dx; uy = S.dy; uz = S.dz;
      break;
    case HMC_SRC_PARALLEL:
      x = S.x0 + S.xe * uniform(rng);
      y = S.y0 + S.ye * uniform(rng);
      z = S.z0;
      break;
    case HMC_SRC_PARALLEL_SPECTRAL:
      e = sample_spectrum(S, rng);
      x = S.x0 + S.xe * uniform(rng);
      y = S.y0 + S.ye * uniform(rng);
      z = S.z0;
      kind = PHOTON;
      book = true;
      break;
    case HMC_SRC_LATTICE: {
      const float* b = S.beamlet + 4 * static_cast<long long>(beamlet);
      x = b[0] + b[1] * uniform(rng);
      y = b[2] + b[3] * uniform(rng);
      z = S.z0;
      kind = PHOTON;
      break;
    }
    case HMC_SRC_SPECTRAL_FAN: {
      e = sample_spectrum(S, rng);
      const float du = (uniform(rng) - 0.5f) * S.width_u;
      const float dv = (uniform(rng) - 0.5f) * S.width_v;
      const float* c = S.beamlet + 3 * static_cast<long long>(beamlet);
      const float dx = c[0] + du * S.uax + dv * S.vax - S.fx;
      const float dy = c[1] + du * S.uay + dv * S.vay - S.fy;
      const float dz = c[2] + du * S.uaz + dv * S.vaz - S.fz;
      const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
      x = S.fx; y = S.fy; z = S.fz;
      ux = dx * inv; uy = dy * inv; uz = dz * inv;
      kind = PHOTON;
      book = true;
      break;
    }
    case HMC_SRC_SPOT_FAN: {
      // Six uniforms: energy (two), rectangle offset (two), Box-Muller spot (two);
      // the primary starts on the rectangle, aimed from the sampled spot point
      // (pyradmc/geometry/source.py:760-792).
      e = sample_spectrum(S, rng);
      const float du = (uniform(rng) - 0.5f) * S.width_u;
      const float dv = (uniform(rng) - 0.5f) * S.width_v;
      const float radius = sqrtf(-2.0f * logf(fmaxf(uniform(rng), 1.0e-12f)));
      const float angle = kTwoPi * uniform(rng);
      const float su = S.sigma_u * radius * cosf(angle), sv = S.sigma_v * radius * sinf(angle);
      const float* c = S.beamlet + 3 * static_cast<long long>(beamlet);
      x = c[0] + du * S.uax + dv * S.vax;
      y = c[1] + du * S.uay + dv * S.vay;
      z = c[2] + du * S.uaz + dv * S.vaz;
      const float dx = x - (S.fx + su * S.uax + sv * S.vax);
      const float dy = y - (S.fy + su * S.uay + sv * S.vay);
      const float dz = z - (S.fz + su * S.uaz + sv * S.vaz);
      const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
      ux = dx * inv; uy = dy * inv; uz = dz * inv;
      kind = PHOTON;
      book = true;
      break;
    }
    default: {  // host-sampled primaries by flat primary index, or a record of the
                // resident population drawn with 64 bits of the history's stream
      long long j = g;
      if (S.kind == HMC_SRC_RESIDENT) {
        const uint64_t hi = rng_next(rng);
        const uint64_t u = (hi << 32) | rng_next(rng);
        j = static_cast<long long>(__umul64hi(u, static_cast<unsigned long long>(S.up_n)));
      }
      kind = static_cast<uint32_t>(S.up_particle[j]);
      e = S.up[j];
      x = S.up[S.up_n + j];
      y = S.up[2 * S.up_n + j];
      z = S.up[3 * S.up_n + j];
      ux = S.up[4 * S.up_n + j];
      uy = S.up[5 * S.up_n + j];
      uz = S.up[6 * S.up_n + j];
      w = S.up[7 * S.up_n + j];
      book = true;
      break;
    }
  }
  return Primary{e, x, y, z, ux, uy, uz, w, kind, book};
}

// Primaries [p0, p0 + n_gen) of the flattened work items, grid-stride (at most a few
// thousand blocks). Single-kind sources write fixed slots of the target queue
// (primary t -> slot t; the count is set once); host-sampled mixed-kind primaries
// reserve slots in the queue of their kind with one atomic per queue per block and
// pass. The emitted energy is reduced over the block and added with one atomic per
// block (pyradmc/backends/warp/kernels.py:1888-2401 for the generators). kRays: the
// source carries a collimator or mask factor (a separate build, so that sources
// without one keep the smaller register budget).
template <bool kRays>
__global__ void __launch_bounds__(256) k_generate(DevSource S, DevRayWeights RW, const DevItem* items,
                                                  int n_items, long long p0, long long n_gen,
                                                  uint64_t seed, Queue qp, Queue qe,
                                                  unsigned long long* ledger) {
  __shared__ unsigned int wave_count[2][4];
  __shared__ unsigned int block_base[2];
  __shared__ long long red[4];
  // The work items and the spectrum's cdf and bin edges, when small enough, are read
  // from LDS: each primary's two binary searches are chains of dependent loads.
  constexpr int kLdsItems = 512, kLdsBins = 256;
  __shared__ DevItem s_items[kLdsItems];
  __shared__ float s_cdf[kLdsBins], s_edges[kLdsBins + 1];
  const bool items_in_lds = n_items <= kLdsItems;
  if (items_in_lds)
    for (int k = static_cast<int>(threadIdx.x); k < n_items; k += static_cast<int>(blockDim.x)) s_items[k] = items[k];
  const DevItem* item_table = items_in_lds ? s_items : items;
  DevSource Sl = S;
  if (S.cdf && S.edges && S.n_bins > 0 && S.n_bins <= kLdsBins) {
    for (int k = static_cast<int>(threadIdx.x); k < S.n_bins; k += static_cast<int>(blockDim.x)) s_cdf[k] = S.cdf[k];
    for (int k = static_cast<int>(threadIdx.x); k <= S.n_bins; k += static_cast<int>(blockDim.x)) s_edges[k] = S.edges[k];
    Sl.cdf = s_cdf;
    Sl.edges = s_edges;
  }
  __syncthreads();
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const int lane = static_cast<int>(__lane_id());
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  const long long passes = (n_gen + stride - 1) / stride;
  long long emitted_sum = 0;
  for (long long pass = 0; pass < passes; ++pass) {
    const long long t =
        pass * stride + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = t < n_gen;
    long long emitted = 0;
    Particle p;
    uint32_t kind = static_cast<uint32_t>(S.particle);
    if (live) {
      const long long g = p0 + t;
      int lo = 0, hi = n_items;  // last item with start <= g
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (item_table[mid].start <= g) lo = mid;
        else hi = mid;
      }
      const DevItem item = item_table[lo];
      const long long i = g - item.start;
      const uint32_t column = static_cast<uint32_t>(item.column);
      uint64_t rng = rng_init(seed, static_cast<uint64_t>(item.key_base + i));
      Primary q = make_primary(Sl, item.beamlet, g, rng);
      float e = q.e, w = q.w, x = q.x, y = q.y, z = q.z, ux = q.ux, uy = q.uy, uz = q.uz;
      bool book = q.book;
      kind = q.kind;
      if constexpr (kRays) {
        if (S.kind != HMC_SRC_UPLOADED) {
          // collimator or mask factor of the primary's ray (no random number drawn)
          w *= static_cast<float>(ray_weight(RW, e, x, y, z, ux, uy, uz));
          book = true;
        }
      }
      if (book) {
        const float latent = kind == POSITRON ? 2.0f * kElectronMass : 0.0f;
        emitted = quanta_round(static_cast<double>(w * (e + latent)));
      }
      p.x = x; p.y = y; p.z = z; p.e = e;
      p.ux = ux; p.uy = uy; p.uz = uz; p.w = w;
      p.rng = rng;
      p.meta = pack_meta(kind, 1u, column);
      p.pad = 0u;
    }
    if (S.kind != HMC_SRC_UPLOADED && S.kind != HMC_SRC_RESIDENT) {
      // One kind per launch: fixed slots, no atomics.
      const Queue& q = S.particle == PHOTON || S.kind >= HMC_SRC_PARALLEL_SPECTRAL ? qp : qe;
      if (live && t < q.capacity) q.data[t] = p;
    } else {
      // Block-level slot reservation: wave counts in LDS, one atomic per queue.
      const bool to_p = live && kind == PHOTON;
      const bool to_e = live && kind != PHOTON;
      const unsigned long long mp = __ballot(to_p ? 1 : 0);
      const unsigned long long me = __ballot(to_e ? 1 : 0);
      if (lane == 0) {
        wave_count[0][wave] = static_cast<unsigned int>(__popcll(mp));
        wave_count[1][wave] = static_cast<unsigned int>(__popcll(me));
      }
      __syncthreads();
      if (threadIdx.x < 2) {
        unsigned int total = 0u;
        for (int wv = 0; wv < static_cast<int>(blockDim.x >> 6); ++wv) total += wave_count[threadIdx.x][wv];
        block_base[threadIdx.x] = total ? atomicAdd(threadIdx.x == 0 ? qp.count : qe.count, total) : 0u;
      }
      __syncthreads();
      unsigned int off_p = block_base[0], off_e = block_base[1];
      for (int wv = 0; wv < wave; ++wv) {
        off_p += wave_count[0][wv];
        off_e += wave_count[1][wv];
      }
      const unsigned long long below = (1ull << lane) - 1ull;
      const unsigned int sp = off_p + static_cast<unsigned int>(__popcll(mp & below));
      const unsigned int se = off_e + static_cast<unsigned int>(__popcll(me & below));
      if (to_p && sp < qp.capacity) qp.data[sp] = p;
      if (to_e && se < qe.capacity) qe.data[se] = p;
      __syncthreads();  // wave_count and block_base are reused by the next pass
    }
    emitted_sum += emitted;
  }
  if (S.kind != HMC_SRC_UPLOADED && S.kind != HMC_SRC_RESIDENT && blockIdx.x == 0 && threadIdx.x == 0) {
    const Queue& q = S.particle == PHOTON || S.kind >= HMC_SRC_PARALLEL_SPECTRAL ? qp : qe;
    *q.count = static_cast<unsigned int>(n_gen);
  }
  const long long wave_total = wave_sum(emitted_sum);
  if (lane == 0) red[wave] = wave_total;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long total = 0;
    for (int wv = 0; wv < static_cast<int>(blockDim.x >> 6); ++wv) total += red[wv];
    if (total != 0) atomicAdd(ledger + L_EMITTED, static_cast<unsigned long long>(total));
  }
}

__global__ void k_fill_u32(unsigned int* a, int n, unsigned int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Reset the two queues a round consumed (counts and heads), recording a count above
// capacity (a dropped push) in the ledger's overflow slot first.
__global__ void k_reset_consumed(unsigned int* counts, unsigned int* heads, int qa, int qb,
                                 unsigned int capacity, unsigned long long* ledger) {
  if (threadIdx.x == 0) {
    const unsigned int ca = counts[qa], cb = counts[qb];
    const unsigned int worst = ca > cb ? ca : cb;
    if (worst > capacity) atomicMax(ledger + L_OVERFLOW, static_cast<unsigned long long>(worst));
    counts[qa] = 0u;
    counts[qb] = 0u;
    heads[qa] = 0u;
    heads[qb] = 0u;
  }
}

// Float32 copies of the compacted dose and sigma, rounded to nearest (as a float64 array's
// conversion to float32 on the host), so that a float32 CSC crosses PCIe at half the bytes.
__global__ void k_narrow_f32(const double* __restrict__ dose, const double* __restrict__ sigma, long long n,
                             float* __restrict__ out) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    out[i] = static_cast<float>(dose[i]);
    out[n + i] = static_cast<float>(sigma[i]);
  }
}

// Beam-aligned windows of finalized columns, for the denoiser: out[c][z][x][y] =
// mean[c][((x0_c + x) * ny + (y0_c + y)) * nz + z] as float (x < wx, y < wy, all z).
__global__ void k_gather_windows(const double* __restrict__ mean, const int* __restrict__ origins,
                                 long long n_cols, int wx, int wy, int ny, int nz, long long nv,
                                 float* __restrict__ out) {
  const long long total = n_cols * nz * wx * wy;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    long long t = i;
    const int y = static_cast<int>(t % wy);
    t /= wy;
    const int x = static_cast<int>(t % wx);
    t /= wx;
    const int z = static_cast<int>(t % nz);
    const long long c = t / nz;
    const long long v = (static_cast<long long>(origins[2 * c] + x) * ny + origins[2 * c + 1] + y) * nz + z;
    out[i] = static_cast<float>(mean[c * nv + v]);
  }
}

// The inverse for denoised results: mean and sigma of each column become the window
// values (zero outside the window; the caller cleared them first).
__global__ void k_scatter_windows(double* __restrict__ mean, double* __restrict__ sigma,
                                  const int* __restrict__ origins, long long n_cols, int wx, int wy,
                                  int ny, int nz, long long nv, const float* __restrict__ in_mean,
                                  const float* __restrict__ in_sigma) {
  const long long total = n_cols * nz * wx * wy;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    long long t = i;
    const int y = static_cast<int>(t % wy);
    t /= wy;
    const int x = static_cast<int>(t % wx);
    t /= wx;
    const int z = static_cast<int>(t % nz);
    const long long c = t / nz;
    const long long v = (static_cast<long long>(origins[2 * c] + x) * ny + origins[2 * c + 1] + y) * nz + z;
    mean[c * nv + v] = static_cast<double>(in_mean[i]);
    sigma[c * nv + v] = in_sigma ? static_cast<double>(in_sigma[i]) : 0.0;
  }
}

// mean / sigma of one column at the given voxels (indices checked on the host).
__global__ void k_set_entries(double* __restrict__ mean, double* __restrict__ sigma, long long base,
                              long long n, const long long* __restrict__ idx, const double* __restrict__ dose,
                              const double* __restrict__ sig) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    mean[base + idx[i]] = dose[i];
    sigma[base + idx[i]] = sig[i];
  }
}

__global__ void k_fill_u64(unsigned long long* a, long long n, unsigned long long v) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    a[i] = v;
}

// ---------------------------------------------------------------- fold/finalize
//
// Exact IEEE float64, contraction off, so the device fold reproduces the host
// scorer arithmetic bit for bit (pyradmc/backends/warp/kernels.py:771-909,
// pyradmc/scoring/dose.py:169-207).
#pragma clang fp contract(off)

constexpr int kFoldSlots = 4;

__global__ void __launch_bounds__(256) k_fold(long long* edep, const double* mass,
                                              const double* nhist, long long n_vox,
                                              long long n_cols, int n_slots, double* s1,
                                              double* s2, unsigned long long* folded) {
  const long long total = n_vox * n_cols;
  const long long slot_stride = total;
  long long local = 0;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    long long col = 0;
    double m = 0.0, a = 0.0, b = 0.0;
    bool touched = false;
    for (int s0 = 0; s0 < n_slots; s0 += kFoldSlots) {
      // The next kFoldSlots slots' tallies are loaded together (more bytes in flight
      // per wave), then taken in batch order.
      long long qs[kFoldSlots];
#pragma unroll
      for (int k = 0; k < kFoldSlots; ++k)
        qs[k] = s0 + k < n_slots ? edep[(s0 + k) * slot_stride + i] : 0;
#pragma unroll
      for (int k = 0; k < kFoldSlots; ++k) {
        const long long q = qs[k];
        if (q == 0) continue;
        const int slot = s0 + k;
        edep[slot * slot_stride + i] = 0;
        if (!touched) {
          // Most cells of a beamlet column see no deposit in a fold: their column,
          // mass and sums are looked up only once a slot holds one.
          col = i / n_vox;
          m = mass[i - col * n_vox];
          a = s1[i];
          b = s2[i];
        }
        double d = 0.0;
        if (m > 0.0) {
          const double energy = static_cast<double>(q) * 1.0e-9;
          d = energy / (m * nhist[slot * n_cols + col]);
        }
        a = a + d;
        b = b + d * d;
        touched = true;
        local += q;
      }
    }
    if (touched) {
      s1[i] = a;
      s2[i] = b;
    }
  }
  const long long sum = wave_sum(local);
  if (__lane_id() == 0 && sum != 0) atomicAdd(folded, static_cast<unsigned long long>(sum));
}

__global__ void __launch_bounds__(256) k_finalize(double* s1, double* s2, long long total,
                                                  int n_batches) {
  const double n = static_cast<double>(n_batches);
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double a = s1[i];
    const double b = s2[i];
    if (a == 0.0 && b == 0.0) continue;  // exact zeros stay
    const double mean = a / n;
    double sigma = 0.0;
    if (n_batches > 1) {
      const double var = fmax(0.0, b / n - mean * mean) * n / (n - 1.0);
      sigma = sqrt(var / n);
    }
    s1[i] = mean;
    s2[i] = sigma;
  }
}

// ------------------------------------------------------- truncation/compaction
//
// Keep an entry when m > 0 and m >= truncation * column max (one f64 product, as
// the host assembler, pyradmc/scoring/dij.py:351-360). Each block owns one chunk of
// one column; ascending scans within chunks and ascending chunk offsets reproduce
// np.flatnonzero order, so the CSC row indices come out sorted.
constexpr int kTruncBlock = 256;
constexpr int kTruncPerThread = 16;
constexpr int kTruncChunk = kTruncBlock * kTruncPerThread;

__global__ void __launch_bounds__(kTruncBlock) k_column_max(const double* mean, long long n_vox,
                                                            int chunks_per_col,
                                                            unsigned long long* col_max) {
  const long long col = blockIdx.x / chunks_per_col;
  const long long chunk = blockIdx.x - col * chunks_per_col;
  const long long v0 = chunk * kTruncChunk;
  const long long v1 = min(v0 + kTruncChunk, n_vox);
  double m = 0.0;
  for (long long v = v0 + threadIdx.x; v < v1; v += kTruncBlock)
    m = fmax(m, mean[col * n_vox + v]);
  __shared__ double red[kTruncBlock];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = kTruncBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // Non-negative doubles order like their bit patterns.
    unsigned long long bits;
    const double v = red[0];
    memcpy(&bits, &v, sizeof(bits));
    atomicMax(col_max + col, bits);
  }
}

__device__ inline bool keep_entry(double m, double thr) { return m > 0.0 && m >= thr; }

__global__ void __launch_bounds__(kTruncBlock) k_count_kept(const double* mean, long long n_vox,
                                                            int chunks_per_col,
                                                            const unsigned long long* col_max,
                                                            double truncation, long long* counts) {
  const long long col = blockIdx.x / chunks_per_col;
  const long long chunk = blockIdx.x - col * chunks_per_col;
  const long long v0 = chunk * kTruncChunk;
  const long long v1 = min(v0 + kTruncChunk, n_vox);
  double cm;
  const unsigned long long bits = col_max[col];
  memcpy(&cm, &bits, sizeof(cm));
  const double thr = truncation * cm;
  int c = 0;
  for (long long v = v0 + threadIdx.x; v < v1; v += kTruncBlock)
    c += keep_entry(mean[col * n_vox + v], thr) ? 1 : 0;
  __shared__ int red[kTruncBlock];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = kTruncBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// One block per chunk, in kTruncPerThread passes of kTruncBlock consecutive voxels (one
// voxel per thread, coalesced): a ballot prefix within each wave and the wave totals
// in LDS give every kept entry its slot, in ascending voxel order.
__global__ void __launch_bounds__(kTruncBlock) k_compact(const double* mean, const double* sigma,
                                                         long long n_vox, int chunks_per_col,
                                                         const unsigned long long* col_max,
                                                         double truncation,
                                                         const long long* offsets,
                                                         long long* out_idx, double* out_dose,
                                                         double* out_sigma) {
  constexpr int kWaves = kTruncBlock / 64;
  const long long col = blockIdx.x / chunks_per_col;
  const long long chunk = blockIdx.x - col * chunks_per_col;
  const long long v0 = chunk * kTruncChunk;
  double cm;
  const unsigned long long bits = col_max[col];
  memcpy(&cm, &bits, sizeof(cm));
  const double thr = truncation * cm;
  const int lane = static_cast<int>(__lane_id());
  const int wave = static_cast<int>(threadIdx.x >> 6);
  __shared__ int wave_kept[2][kWaves];
  long long cursor = offsets[blockIdx.x];
  for (int pass = 0; pass < kTruncPerThread; ++pass) {
    const long long v = v0 + static_cast<long long>(pass) * kTruncBlock + threadIdx.x;
    double m = 0.0;
    bool keep = false;
    if (v < n_vox) {
      m = mean[col * n_vox + v];
      keep = keep_entry(m, thr);
    }
    const unsigned long long mask = __ballot(keep ? 1 : 0);
    int* kept = wave_kept[pass & 1];  // alternating rows: one barrier per pass
    if (lane == 0) kept[wave] = __popcll(mask);
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      before += w < wave ? kept[w] : 0;
      total += kept[w];
    }
    if (keep) {
      const long long at = cursor + before + __popcll(mask & ((1ull << lane) - 1ull));
      out_idx[at] = v;
      out_dose[at] = m;
      out_sigma[at] = sigma[col * n_vox + v];
    }
    cursor += total;
  }
}

#pragma clang fp contract(on)

}  // namespace

// =============================================================== host context

struct hmc_ctx {
  int device = 0;
  std::string error;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  std::unordered_map<void*, size_t> alloc_bytes;  // dev_alloc capacities

  // transport grid
  DevGrid g{};
  float* d_density = nullptr;
  uint8_t* d_material = nullptr;
  int max_material = 0;
  bool has_grid = false;

  // scoring
  DevScoring s{};
  double* d_mass = nullptr;
  long long n_columns = 0;
  int n_slots = 1;
  long long* d_edep = nullptr;
  double* d_s1 = nullptr;
  double* d_s2 = nullptr;
  double* d_nhist = nullptr;
  // Leading map slots that may hold quanta: transport raises it, a fold of those slots
  // (which consumes them) clears it, and a reset zeroes only that many slots.
  int dirty_slots = 0;
  int* d_origins = nullptr;  // window origins for hmc_gather_windows / hmc_scatter_windows
  bool has_scoring = false;
  bool finalized = false;
  // Folds made without a host wait (hmc_fold with total_quanta NULL) add their quanta to
  // the device counter past the ledger; the next hmc_finalize reads it back with the
  // means, and hmc_folded_quanta hands it out.
  bool fold_pending = false;
  long long folded_quanta = 0;

  // tables
  DevTables t{};
  std::vector<void*> table_buffers;
  bool has_tables = false;
  float* d_gs = nullptr;
  bool has_gs = false;

  // source
  DevSource src{};
  float* d_edges = nullptr;
  float* d_cdf = nullptr;
  float* d_beamlet = nullptr;
  int32_t* d_up_particle = nullptr;
  float* d_up = nullptr;
  // ray weight model of the generator and the device buffers it points into
  DevRayWeights rw{};
  std::vector<void*> rw_buffers;
  // resident population: [n] kinds and [8][n] columns, rows [0, pop_written) written
  int32_t* d_pop_particle = nullptr;
  float* d_pop = nullptr;
  long long pop_n = 0, pop_written = 0;
  int src_beamlets = 0;
  bool has_source = false;
  // What d_edges, d_cdf and d_beamlet hold: a source set again with the same arrays
  // (every run of a loop over seeds) uploads nothing.
  std::vector<float> h_edges, h_cdf, h_beamlet;

  hmc_physics phys{};
  bool has_physics = false;

  // queues: 4 queues + counts (4) + heads (4)
  Particle* d_q[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t q_capacity = 0;
  unsigned int* d_counts = nullptr;  // [0..3] counts, [4..7] heads, then the energy sort's histogram
  unsigned long long* d_ledger = nullptr;
  // Page-locked readback of the per-group counts (words 0-1: four u32), the ledger
  // (words 2..) and the fold's quanta (the last word). Copies into page-locked memory
  // need no host-side staging, so one wait covers them.
  unsigned long long* h_stage = nullptr;
  DevItem* d_items = nullptr;
  size_t items_cap = 0;
  bool profile = false;
  std::vector<hipEvent_t> ev_pool;  // per-launch timing events (profiling)

  // launch shape of the transport rounds (hmc_set_launch)
  int round_lanes_per_cu = 2048;
  int refill_block = 64;
  int push_chunk = 64;
  bool use_graphs = true;
  // the captured group of rounds and what it was captured with
  hipGraphExec_t round_graph = nullptr;
  std::vector<unsigned char> round_graph_key;

  // truncation
  unsigned long long* d_colmax = nullptr;
  long long* d_chunk_counts = nullptr;
  long long* d_chunk_offsets = nullptr;
  long long* d_out_idx = nullptr;
  double* d_out_dose = nullptr;
  double* d_out_sigma = nullptr;
  float* d_out32 = nullptr;  // float32 copies of the compacted dose and sigma (hmc_fetch_csc_f32)
  float* d_samples = nullptr;  // hmc_sample_physics output
  long long* d_entry_idx = nullptr;  // hmc_set_column_entries staging
  double* d_entry_val = nullptr;
  long long trunc_total = 0;
  size_t out_cap = 0;
};

namespace {

int fail(hmc_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

// A failed call is reported once: the runtime's last-error slot is cleared so a later,
// unrelated hipGetLastError() does not return it again.
#define HMC_HIP(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      (void)hipGetLastError();                                                          \
      return fail(ctx, _e == hipErrorOutOfMemory ? HMC_ENOMEM : HMC_EHIP,               \
                  std::string(#call) + ": " + hipGetErrorString(_e));                   \
    }                                                                                   \
  } while (0)

// Device buffer of at least `count` elements: an existing allocation that is large
// enough is kept (hipMalloc/hipFree synchronise the device, so repeated runs with the
// same shapes must not reallocate), otherwise it is replaced.
template <typename T>
int dev_alloc(hmc_ctx* ctx, T** ptr, size_t count) {
  if (count == 0) count = 1;
  const size_t bytes = count * sizeof(T);
  if (*ptr) {
    auto it = ctx->alloc_bytes.find(static_cast<void*>(*ptr));
    if (it != ctx->alloc_bytes.end() && it->second >= bytes) return HMC_OK;
    if (it != ctx->alloc_bytes.end()) ctx->alloc_bytes.erase(it);
    (void)hipFree(*ptr);
    *ptr = nullptr;
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(ptr), bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(ctx, e == hipErrorOutOfMemory ? HMC_ENOMEM : HMC_EHIP,
                std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  ctx->alloc_bytes[static_cast<void*>(*ptr)] = bytes;
  return HMC_OK;
}

template <typename T>
int upload(hmc_ctx* ctx, T** ptr, const T* host, size_t count) {
  int rc = dev_alloc(ctx, ptr, count);
  if (rc) return rc;
  if (host && count) HMC_HIP(ctx, hipMemcpy(*ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
  return HMC_OK;
}

void free_ptr(void* p) {
  if (p) (void)hipFree(p);
}

int blocks_for(long long n, int block = 256, long long cap = 1 << 20) {
  long long b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<int>(b);
}

}  // namespace

// Draws of every scalar sampler the transport rounds use, one independent stream per
// sample (key = sample index), for the sampler tests: out[i][column] as documented at
// hmc_sample_physics. GS and coherent columns need the context's tables (NaN without).
__global__ void __launch_bounds__(256) k_sample_physics(uint64_t seed, long long n, hmc_sampler_args a,
                                                        DevTables t, int have_tables, int have_gs,
                                                        float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t rng = rng_init(seed, static_cast<uint64_t>(i));
  float* o = out + i * HMC_SAMPLE_COLUMNS;
  const float nan = __int_as_float(0x7fc00000);
  o[0] = sample_path(static_cast<float>(a.path_mu), rng);
  o[1] = static_cast<float>(select_process(static_cast<float>(a.channel_mu[0]), static_cast<float>(a.channel_mu[1]),
                                           static_cast<float>(a.channel_mu[2]), static_cast<float>(a.channel_mu[3]),
                                           rng));
  const float e0 = static_cast<float>(a.compton_energy[0]);
  const float r = sample_compton_ratio(e0, rng);
  o[2] = r;
  o[3] = compton_cos(e0, r);
  o[4] = compton_electron_cos(e0, r);
  o[5] = sample_compton_ratio(static_cast<float>(a.compton_energy[1]), rng);
  float ax, ay, az;
  sample_isotropic(rng, ax, ay, az);
  o[6] = ax * ax + ay * ay + az * az;
  o[7] = az;
  float vx = ax, vy = ay, vz = az;
  rotate(vx, vy, vz, static_cast<float>(a.rotate_cos), static_cast<float>(a.rotate_phi));
  o[8] = vx * vx + vy * vy + vz * vz;
  const float em = static_cast<float>(a.moller_energy);
  const float w = sample_moller(em, static_cast<float>(a.moller_cut), rng);
  o[9] = w;
  float cd, cp;
  moller_cosines(em, w, cd, cp);
  o[10] = cd;
  o[11] = cp;
  o[12] = sample_gaussian_hinge(static_cast<float>(a.hinge_ms), rng);
  const float eb = static_cast<float>(a.brems_energy), pc = static_cast<float>(a.brems_pcut);
  const float log_eb = __logf(eb), log_pc = __logf(pc);
  float prob, local;
  brems_step(eb, static_cast<float>(a.brems_radiative_loss), pc, log_eb, log_pc, prob, local);
  o[13] = prob;
  o[14] = local;
  o[15] = sample_brems_energy(pc, log_eb, log_pc, rng);
  o[16] = have_tables ? sample_coherent(t, a.coherent_material, static_cast<float>(a.coherent_energy), rng) : nan;
  // GS at (log eta, <theta^2>) with the round's keying; the draw after it and the draw
  // after a Gaussian hinge from the same state (each consumes one uniform).
  const float fx = static_cast<float>(a.gs_log_eta) * kGsBinsPerLog;
  for (int k = 0; k < 2; ++k) {
    const float ms = static_cast<float>(a.gs_ms[k]);
    const float fy = __logf(fmaxf(ms, kGsTheta2Min)) * kGsBinsPerLog;
    o[17 + k] = have_gs ? sample_gs(t, fx, fy, ms, rng) : nan;
  }
  uint64_t fork = rng;
  const float ms0 = static_cast<float>(a.gs_ms[0]);
  if (have_gs) {
    (void)sample_gs(t, fx, __logf(fmaxf(ms0, kGsTheta2Min)) * kGsBinsPerLog, ms0, fork);
    o[19] = uniform(fork);
  } else {
    o[19] = nan;
  }
  fork = rng;
  (void)sample_gaussian_hinge(ms0, fork);
  o[20] = uniform(fork);
  // the pole guard: (0, 0, 1) turned by (rotate_cos, rotate_phi)
  float px = 0.0f, py = 0.0f, pz = 1.0f;
  rotate(px, py, pz, static_cast<float>(a.rotate_cos), static_cast<float>(a.rotate_phi));
  o[21] = pz;
}

// ======================================================================== ABI

extern "C" {

int hmc_create(int device, hmc_ctx** out) {
  if (!out) return HMC_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HMC_EINVAL;
  auto* ctx = new hmc_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    delete ctx;
    return HMC_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
  if (hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return HMC_EHIP;
  }
  if (dev_alloc(ctx, &ctx->d_counts, 8 + kEnergyBins * kSortBlocks) || dev_alloc(ctx, &ctx->d_ledger, L_COUNT + 1) ||
      hipHostMalloc(reinterpret_cast<void**>(&ctx->h_stage), sizeof(unsigned long long) * kStageWords,
                    hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    free_ptr(ctx->d_counts);
    free_ptr(ctx->d_ledger);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HMC_ENOMEM;
  }
  *out = ctx;
  return HMC_OK;
}

void hmc_destroy(hmc_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  void* ptrs[] = {ctx->d_density,      ctx->d_material,    ctx->d_mass,       ctx->d_edep,
                  ctx->d_s1,           ctx->d_s2,          ctx->d_nhist,      ctx->d_gs,
                  ctx->d_edges,        ctx->d_cdf,         ctx->d_beamlet,    ctx->d_up_particle,
                  ctx->d_up,           ctx->d_q[0],        ctx->d_q[1],       ctx->d_q[2],
                  ctx->d_q[3],         ctx->d_counts,      ctx->d_ledger,     ctx->d_items,
                  ctx->d_colmax,       ctx->d_chunk_counts, ctx->d_chunk_offsets,
                  ctx->d_out_idx,      ctx->d_out_dose,    ctx->d_out_sigma,  ctx->d_origins,
                  ctx->d_out32,        ctx->d_samples,     ctx->d_entry_idx,  ctx->d_entry_val,
                  ctx->d_pop_particle, ctx->d_pop};
  for (void* p : ptrs) free_ptr(p);
  for (void* p : ctx->table_buffers) free_ptr(p);
  for (void* p : ctx->rw_buffers) free_ptr(p);
  for (hipEvent_t e : ctx->ev_pool)
    if (e) (void)hipEventDestroy(e);
  if (ctx->round_graph) (void)hipGraphExecDestroy(ctx->round_graph);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* hmc_last_error(const hmc_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int hmc_device_info(int device, char* name, int name_len, int* cu_count, int64_t* total_bytes,
                    int64_t* free_bytes) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return HMC_EINVAL;
  if ((name && name_len > 0) || cu_count || total_bytes) {  // the properties query is the slow part
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return HMC_EHIP;
    if (name && name_len > 0) {
      // The marketing name can be empty where the ids table is not readable; the
      // architecture and CU count always identify the part.
      std::snprintf(name, static_cast<size_t>(name_len), "%s%s%s, %d CUs", prop.name,
                    prop.name[0] ? " " : "", prop.gcnArchName, prop.multiProcessorCount);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (total_bytes) *total_bytes = static_cast<int64_t>(prop.totalGlobalMem);
  }
  if (free_bytes) {
    size_t f = 0, tot = 0;
    (void)hipSetDevice(device);
    if (hipMemGetInfo(&f, &tot) != hipSuccess) return HMC_EHIP;
    *free_bytes = static_cast<int64_t>(f);
  }
  return HMC_OK;
}

int hmc_set_grid(hmc_ctx* ctx, int nx, int ny, int nz, const double* origin,
                 const double* spacing, const float* density, const uint8_t* material) {
  if (!ctx || !origin || !spacing || !density || !material) return fail(ctx, HMC_EINVAL, "null");
  if (nx < 1 || ny < 1 || nz < 1) return fail(ctx, HMC_EINVAL, "empty grid shape");
  if (spacing[0] <= 0 || spacing[1] <= 0 || spacing[2] <= 0)
    return fail(ctx, HMC_EINVAL, "non-positive spacing");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const size_t n = static_cast<size_t>(nx) * ny * nz;
  int rc = upload(ctx, &ctx->d_density, density, n);
  if (rc) return rc;
  rc = upload(ctx, &ctx->d_material, material, n);
  if (rc) return rc;
  DevGrid& g = ctx->g;
  g.x_lo = static_cast<float>(origin[0]);
  g.y_lo = static_cast<float>(origin[1]);
  g.z_lo = static_cast<float>(origin[2]);
  g.x_hi = static_cast<float>(origin[0] + nx * spacing[0]);
  g.y_hi = static_cast<float>(origin[1] + ny * spacing[1]);
  g.z_hi = static_cast<float>(origin[2] + nz * spacing[2]);
  g.sx = static_cast<float>(spacing[0]);
  g.sy = static_cast<float>(spacing[1]);
  g.sz = static_cast<float>(spacing[2]);
  g.inv_sx = static_cast<float>(1.0 / spacing[0]);
  g.inv_sy = static_cast<float>(1.0 / spacing[1]);
  g.inv_sz = static_cast<float>(1.0 / spacing[2]);
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.density = ctx->d_density;
  g.material = ctx->d_material;
  ctx->max_material = static_cast<int>(*std::max_element(material, material + n));
  ctx->has_grid = true;
  return HMC_OK;
}

int hmc_set_scoring(hmc_ctx* ctx, int nx, int ny, int nz, const double* origin,
                    const double* spacing, const double* voxel_mass, int64_t n_columns,
                    int n_slots) {
  if (!ctx || !origin || !spacing || !voxel_mass) return fail(ctx, HMC_EINVAL, "null");
  if (nx < 1 || ny < 1 || nz < 1 || n_columns < 1 || n_slots < 1 ||
      n_columns * n_slots >= (1ll << 28))
    return fail(ctx, HMC_EINVAL, "bad scoring shape, column or slot count");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const long long nv = static_cast<long long>(nx) * ny * nz;
  int rc = upload(ctx, &ctx->d_mass, voxel_mass, static_cast<size_t>(nv));
  if (rc) return rc;
  const size_t total = static_cast<size_t>(nv) * static_cast<size_t>(n_columns);
  if ((rc = dev_alloc(ctx, &ctx->d_edep, total * static_cast<size_t>(n_slots)))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_s1, total))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_s2, total))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_nhist, static_cast<size_t>(n_columns) * n_slots))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_colmax, static_cast<size_t>(n_columns)))) return rc;
  DevScoring& s = ctx->s;
  s.x_lo = static_cast<float>(origin[0]);
  s.y_lo = static_cast<float>(origin[1]);
  s.z_lo = static_cast<float>(origin[2]);
  s.x_hi = static_cast<float>(origin[0] + nx * spacing[0]);
  s.y_hi = static_cast<float>(origin[1] + ny * spacing[1]);
  s.z_hi = static_cast<float>(origin[2] + nz * spacing[2]);
  s.inv_sx = static_cast<float>(1.0 / spacing[0]);
  s.inv_sy = static_cast<float>(1.0 / spacing[1]);
  s.inv_sz = static_cast<float>(1.0 / spacing[2]);
  s.nx = nx;
  s.ny = ny;
  s.nz = nz;
  s.n_voxels = nv;
  ctx->n_columns = n_columns;
  ctx->n_slots = n_slots;
  ctx->has_scoring = true;
  ctx->dirty_slots = n_slots;  // fresh or reused buffers: clear every slot once
  return hmc_reset_maps(ctx);
}

int hmc_reset_maps(hmc_ctx* ctx) {
  if (!ctx || !ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const long long total = ctx->s.n_voxels * ctx->n_columns;
  const int blocks = blocks_for(total, 256, 65536);
  if (ctx->dirty_slots > 0)
    hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(total * ctx->dirty_slots, 256, 65536)), dim3(256), 0,
                       ctx->stream, reinterpret_cast<unsigned long long*>(ctx->d_edep),
                       total * ctx->dirty_slots, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(ctx->d_s1), total, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(ctx->d_s2), total, 0ull);
  HMC_HIP(ctx, hipGetLastError());  // stream-ordered before the next transport: no host wait
  ctx->dirty_slots = 0;
  ctx->finalized = false;
  ctx->fold_pending = false;
  ctx->folded_quanta = 0;
  return HMC_OK;
}

int hmc_set_quanta(hmc_ctx* ctx, const int64_t* quanta) {
  if (!ctx || !quanta) return fail(ctx, HMC_EINVAL, "null");
  if (!ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const size_t total = static_cast<size_t>(ctx->s.n_voxels * ctx->n_columns * ctx->n_slots);
  HMC_HIP(ctx, hipMemcpy(ctx->d_edep, quanta, sizeof(long long) * total, hipMemcpyHostToDevice));
  ctx->dirty_slots = ctx->n_slots;
  return HMC_OK;
}

int hmc_get_quanta(hmc_ctx* ctx, int64_t* quanta) {
  if (!ctx || !quanta) return fail(ctx, HMC_EINVAL, "null");
  if (!ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t total = static_cast<size_t>(ctx->s.n_voxels * ctx->n_columns * ctx->n_slots);
  HMC_HIP(ctx, hipMemcpy(quanta, ctx->d_edep, sizeof(long long) * total, hipMemcpyDeviceToHost));
  return HMC_OK;
}

int hmc_set_tables(hmc_ctx* ctx, const hmc_tables* tb) {
  if (!ctx || !tb) return fail(ctx, HMC_EINVAL, "null");
  const int n = tb->n_points, nm = tb->n_mat, nc = tb->n_coherent;
  if (n < 2 || nm < 1 || nm > 256 || nc < 2)
    return fail(ctx, HMC_EINVAL, "table sizes out of range (n_points>=2, 1<=n_mat<=256)");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  for (void* p : ctx->table_buffers) free_ptr(p);
  ctx->table_buffers.clear();
  const size_t rows = static_cast<size_t>(nm) * n;
  std::vector<float4> photon(rows), ea(rows), eb(rows);
  for (size_t i = 0; i < rows; ++i) {
    photon[i] = make_float4(tb->mu_compton[i], tb->mu_photo[i], tb->mu_pair[i], tb->mu_rayleigh[i]);
    ea[i] = make_float4(tb->csda_range[i], tb->moller[i], tb->restricted_range[i],
                        tb->stopping_radiative[i]);
    eb[i] = make_float4(tb->scattering_power[i], tb->log_eta[i], tb->stopping_restricted[i], 0.0f);
  }
  float* d_maj = nullptr;
  float4 *d_ph = nullptr, *d_ea = nullptr, *d_eb = nullptr;
  float *d_inv = nullptr, *d_cx = nullptr, *d_cc = nullptr;
  int rc;
  if ((rc = upload(ctx, &d_maj, tb->majorant, n))) return rc;
  ctx->table_buffers.push_back(d_maj);
  if ((rc = upload(ctx, &d_ph, photon.data(), rows))) return rc;
  ctx->table_buffers.push_back(d_ph);
  if ((rc = upload(ctx, &d_ea, ea.data(), rows))) return rc;
  ctx->table_buffers.push_back(d_ea);
  if ((rc = upload(ctx, &d_eb, eb.data(), rows))) return rc;
  ctx->table_buffers.push_back(d_eb);
  if ((rc = upload(ctx, &d_inv, tb->energy_of_restricted_range, rows))) return rc;
  ctx->table_buffers.push_back(d_inv);
  if ((rc = upload(ctx, &d_cx, tb->coherent_x, static_cast<size_t>(nc)))) return rc;
  ctx->table_buffers.push_back(d_cx);
  if ((rc = upload(ctx, &d_cc, tb->coherent_cumulative, static_cast<size_t>(nm) * nc))) return rc;
  ctx->table_buffers.push_back(d_cc);
  DevTables& t = ctx->t;
  t.majorant = d_maj;
  t.photon = d_ph;
  t.elec_a = d_ea;
  t.elec_b = d_eb;
  t.inv_range = d_inv;
  t.coherent_x = d_cx;
  t.coherent_cumulative = d_cc;
  t.p_log_min = static_cast<float>(tb->p_log_min);
  t.p_inv_dlog = static_cast<float>(tb->p_inv_dlog);
  t.e_log_min = static_cast<float>(tb->e_log_min);
  t.e_inv_dlog = static_cast<float>(tb->e_inv_dlog);
  t.r_log_min = static_cast<float>(tb->r_log_min);
  t.r_inv_dlog = static_cast<float>(tb->r_inv_dlog);
  t.n_points = n;
  t.n_mat = nm;
  t.n_coherent = nc;
  ctx->has_tables = true;
  return HMC_OK;
}

int hmc_set_gs(hmc_ctx* ctx, const float* values, int n_eta, int n_theta2, int n_u, int ix0,
               int iy0) {
  if (!ctx || !values) return fail(ctx, HMC_EINVAL, "null");
  if (n_eta < 2 || n_theta2 < 2 || n_u < 1)
    return fail(ctx, HMC_EINVAL, "GS grid needs at least 2x2 nodes");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const size_t count = static_cast<size_t>(n_eta) * n_theta2 * n_u;
  int rc = upload(ctx, &ctx->d_gs, values, count);
  if (rc) return rc;
  ctx->t.gs = ctx->d_gs;
  ctx->t.gs_ix0 = ix0;
  ctx->t.gs_iy0 = iy0;
  ctx->t.gs_n_eta = n_eta;
  ctx->t.gs_n_theta2 = n_theta2;
  ctx->t.gs_n_u = n_u;
  ctx->has_gs = true;
  return HMC_OK;
}

int hmc_set_source(hmc_ctx* ctx, const hmc_source* s) {
  if (!ctx || !s) return fail(ctx, HMC_EINVAL, "null");
  if (s->kind < HMC_SRC_PENCIL || s->kind > HMC_SRC_SPOT_FAN)
    return fail(ctx, HMC_EINVAL, "unknown source kind");
  if (s->kind == HMC_SRC_SPOT_FAN && !(s->sigma_u >= 0.0 && s->sigma_v >= 0.0))
    return fail(ctx, HMC_EINVAL, "spot sigma must be non-negative");
  if (s->kind == HMC_SRC_RESIDENT && (ctx->pop_n < 1 || ctx->pop_written != ctx->pop_n))
    return fail(ctx, HMC_EINVAL, "resident source without a complete population");
  if (s->particle < 0 || s->particle > 2) return fail(ctx, HMC_EINVAL, "unknown particle kind");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  DevSource& d = ctx->src;
  d.kind = s->kind;
  d.particle = s->particle;
  d.energy = static_cast<float>(s->energy);
  d.px = static_cast<float>(s->position[0]);
  d.py = static_cast<float>(s->position[1]);
  d.pz = static_cast<float>(s->position[2]);
  d.dx = static_cast<float>(s->direction[0]);
  d.dy = static_cast<float>(s->direction[1]);
  d.dz = static_cast<float>(s->direction[2]);
  d.z0 = static_cast<float>(s->z0);
  d.x0 = static_cast<float>(s->x0);
  d.xe = static_cast<float>(s->x_extent);
  d.y0 = static_cast<float>(s->y0);
  d.ye = static_cast<float>(s->y_extent);
  d.fx = static_cast<float>(s->focal[0]);
  d.fy = static_cast<float>(s->focal[1]);
  d.fz = static_cast<float>(s->focal[2]);
  d.uax = static_cast<float>(s->u_axis[0]);
  d.uay = static_cast<float>(s->u_axis[1]);
  d.uaz = static_cast<float>(s->u_axis[2]);
  d.vax = static_cast<float>(s->v_axis[0]);
  d.vay = static_cast<float>(s->v_axis[1]);
  d.vaz = static_cast<float>(s->v_axis[2]);
  d.width_u = static_cast<float>(s->width_u);
  d.width_v = static_cast<float>(s->width_v);
  d.n_bins = s->n_bins;
  int rc;
  d.sigma_u = static_cast<float>(s->sigma_u);
  d.sigma_v = static_cast<float>(s->sigma_v);
  const bool spectral = s->kind == HMC_SRC_PARALLEL_SPECTRAL || s->kind == HMC_SRC_SPECTRAL_FAN ||
                        s->kind == HMC_SRC_SPOT_FAN;
  if (spectral) {
    if (s->n_bins < 1 || !s->edges || !s->cdf) return fail(ctx, HMC_EINVAL, "spectrum missing");
    std::vector<float> e(s->edges, s->edges + s->n_bins + 1), c(s->cdf, s->cdf + s->n_bins);
    if (!ctx->d_edges || e != ctx->h_edges) {
      ctx->h_edges.clear();
      if ((rc = upload(ctx, &ctx->d_edges, e.data(), e.size()))) return rc;
      ctx->h_edges = std::move(e);
    }
    if (!ctx->d_cdf || c != ctx->h_cdf) {
      ctx->h_cdf.clear();
      if ((rc = upload(ctx, &ctx->d_cdf, c.data(), c.size()))) return rc;
      ctx->h_cdf = std::move(c);
    }
    d.edges = ctx->d_edges;
    d.cdf = ctx->d_cdf;
  }
  if (s->kind == HMC_SRC_LATTICE || s->kind == HMC_SRC_SPECTRAL_FAN || s->kind == HMC_SRC_SPOT_FAN) {
    const int stride = s->kind == HMC_SRC_LATTICE ? 4 : 3;
    if (s->n_beamlets < 1 || !s->beamlet_params)
      return fail(ctx, HMC_EINVAL, "beamlet parameters missing");
    std::vector<float> b(s->beamlet_params, s->beamlet_params + s->n_beamlets * stride);
    if (!ctx->d_beamlet || b != ctx->h_beamlet) {
      ctx->h_beamlet.clear();
      if ((rc = upload(ctx, &ctx->d_beamlet, b.data(), b.size()))) return rc;
      ctx->h_beamlet = std::move(b);
    }
    d.beamlet = ctx->d_beamlet;
    ctx->src_beamlets = s->n_beamlets;
  }
  if (s->kind == HMC_SRC_UPLOADED) d.up_n = 0;  // primaries must be uploaded for this source
  if (s->kind == HMC_SRC_RESIDENT) {
    d.up_particle = ctx->d_pop_particle;
    d.up = ctx->d_pop;
    d.up_n = ctx->pop_n;
  }
  ctx->has_source = true;
  return HMC_OK;
}

int hmc_set_population(hmc_ctx* ctx, int64_t n, int64_t offset, int64_t count, const int32_t* particle,
                       const float* energy, const float* x, const float* y, const float* z,
                       const float* ux, const float* uy, const float* uz, const float* weight) {
  if (!ctx || !particle || !energy || !x || !y || !z || !ux || !uy || !uz || !weight)
    return fail(ctx, HMC_EINVAL, "population columns incomplete");
  if (n < 1 || count < 1 || offset < 0 || offset > n - count)
    return fail(ctx, HMC_EINVAL, "population rows out of range");
  if (offset != (offset == 0 ? 0 : ctx->pop_written) || (offset > 0 && n != ctx->pop_n))
    return fail(ctx, HMC_EINVAL, "population rows must be written in order");
  for (int64_t i = 0; i < count; ++i)
    if (particle[i] < 0 || particle[i] > 2) return fail(ctx, HMC_EINVAL, "bad particle kind");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no generator may be reading the rows
  int rc;
  if (offset == 0) {
    ctx->pop_n = ctx->pop_written = 0;
    if (ctx->src.kind == HMC_SRC_RESIDENT) ctx->src.up_n = 0;
    if ((rc = dev_alloc(ctx, &ctx->d_pop_particle, static_cast<size_t>(n)))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_pop, static_cast<size_t>(n) * 8))) return rc;
    ctx->pop_n = n;
  }
  HMC_HIP(ctx, hipMemcpy(ctx->d_pop_particle + offset, particle, sizeof(int32_t) * count,
                         hipMemcpyHostToDevice));
  const float* cols[8] = {energy, x, y, z, ux, uy, uz, weight};
  for (int c = 0; c < 8; ++c)
    HMC_HIP(ctx, hipMemcpy(ctx->d_pop + c * n + offset, cols[c], sizeof(float) * count,
                           hipMemcpyHostToDevice));
  ctx->pop_written = offset + count;
  return HMC_OK;
}

int hmc_set_uploaded(hmc_ctx* ctx, int64_t n, const int32_t* particle, const float* energy,
                     const float* x, const float* y, const float* z, const float* ux,
                     const float* uy, const float* uz, const float* weight) {
  if (!ctx || n < 1 || !particle || !energy || !x || !y || !z || !ux || !uy || !uz || !weight)
    return fail(ctx, HMC_EINVAL, "uploaded primaries incomplete");
  if (ctx->src.kind != HMC_SRC_UPLOADED)
    return fail(ctx, HMC_EINVAL, "primaries uploaded for a source that is not host-sampled");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  for (int64_t i = 0; i < n; ++i)
    if (particle[i] < 0 || particle[i] > 2) return fail(ctx, HMC_EINVAL, "bad particle kind");
  int rc = upload(ctx, &ctx->d_up_particle, particle, static_cast<size_t>(n));
  if (rc) return rc;
  std::vector<float> cols(static_cast<size_t>(n) * 8);
  const float* src[8] = {energy, x, y, z, ux, uy, uz, weight};
  for (int c = 0; c < 8; ++c) std::memcpy(cols.data() + c * n, src[c], sizeof(float) * n);
  if ((rc = upload(ctx, &ctx->d_up, cols.data(), cols.size()))) return rc;
  ctx->src.up_particle = ctx->d_up_particle;
  ctx->src.up = ctx->d_up;
  ctx->src.up_n = n;
  return HMC_OK;
}

__global__ void k_ray_weights(DevRayWeights RW, long long n, const float* e, const float* pos,
                              const float* dir, double* out) {
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = ray_weight(RW, e[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], dir[3 * i], dir[3 * i + 1],
                        dir[3 * i + 2]);
}

// Validate a ray weight model and copy its arrays to the device (buffers appended to
// bufs, which the caller frees); with need_tables false a stack's tables are not read.
int build_ray_weights(hmc_ctx* ctx, const hmc_ray_weights* w, bool need_tables, DevRayWeights& r,
                      std::vector<void*>& bufs) {
  r = DevRayWeights{};
  int rc;
  auto put = [&](const auto* host, size_t count, const auto** dst) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(host)>>;
    T* d = nullptr;
    const int code = upload(ctx, &d, host, count);
    if (d) {
      bufs.push_back(d);
      ctx->alloc_bytes.erase(d);  // owned by bufs, freed with free_ptr
    }
    *dst = d;
    return code;
  };
  if (w->kind == 1) {
    const int n = w->n_devices;
    if (n < 1 || !w->device_kind || !w->flags || !w->n_pairs || !w->edge_offset || !w->pair_offset ||
        !w->z_top || !w->z_bottom || !w->edge_neg || !w->edge_pos || !w->density || w->n_leaf_edges < 1 ||
        w->n_tips < 1 || !w->leaf_edges || !w->tips_neg || !w->tips_pos ||
        (need_tables && (w->n_table < 2 || !w->log_energies || !w->log_mu)))
      return fail(ctx, HMC_EINVAL, "beam-limiting stack incomplete");
    for (int j = 0; j < n; ++j) {
      if (w->device_kind[j] != 0 && w->device_kind[j] != 1) return fail(ctx, HMC_EINVAL, "unknown device kind");
      if (!(w->z_top[j] > 0.0) || !(w->z_bottom[j] > w->z_top[j]))
        return fail(ctx, HMC_EINVAL, "device slab must lie downstream of the focal spot");
      if (w->device_kind[j] == 1) {
        const int np = w->n_pairs[j];
        if (np < 1 || w->edge_offset[j] < 0 || w->pair_offset[j] < 0 ||
            w->edge_offset[j] + np + 1 > w->n_leaf_edges || w->pair_offset[j] + np > w->n_tips)
          return fail(ctx, HMC_EINVAL, "MLC leaf arrays out of range");
        for (int i = 0; i < np; ++i)
          if (!(w->leaf_edges[w->edge_offset[j] + i + 1] > w->leaf_edges[w->edge_offset[j] + i]))
            return fail(ctx, HMC_EINVAL, "MLC strip edges must increase");
      }
    }
    if (need_tables)
      for (int k = 1; k < w->n_table; ++k)
        if (!(w->log_energies[k] > w->log_energies[k - 1])) return fail(ctx, HMC_EINVAL, "energies must increase");
    r.n_dev = n;
    for (int k = 0; k < 3; ++k) r.o[k] = w->origin[k];
    for (int k = 0; k < 9; ++k) r.b[k] = w->basis[k];
    if ((rc = put(w->device_kind, n, &r.dkind)) || (rc = put(w->flags, n, &r.flags)) ||
        (rc = put(w->n_pairs, n, &r.n_pairs)) || (rc = put(w->edge_offset, n, &r.off)) ||
        (rc = put(w->pair_offset, n, &r.poff)) || (rc = put(w->z_top, n, &r.z_top)) ||
        (rc = put(w->z_bottom, n, &r.z_bot)) || (rc = put(w->edge_neg, n, &r.e_neg)) ||
        (rc = put(w->edge_pos, n, &r.e_pos)) || (rc = put(w->density, n, &r.dens)) ||
        (rc = put(w->leaf_edges, w->n_leaf_edges, &r.edges)) || (rc = put(w->tips_neg, w->n_tips, &r.tn)) ||
        (rc = put(w->tips_pos, w->n_tips, &r.tp)))
      return rc;
    if (need_tables) {
      r.n_table = w->n_table;
      if ((rc = put(w->log_energies, w->n_table, &r.log_e)) ||
          (rc = put(w->log_mu, static_cast<size_t>(n) * w->n_table, &r.log_mu)))
        return rc;
      r.le0 = w->log_energies[0];
      r.inv_dle = (w->n_table - 1) / (w->log_energies[w->n_table - 1] - w->log_energies[0]);
    }
  } else if (w->kind == 2) {
    if (w->mask_nu < 1 || w->mask_nv < 1 || !w->mask || !(w->width_u > 0.0) || !(w->width_v > 0.0))
      return fail(ctx, HMC_EINVAL, "mask incomplete");
    const size_t cells = static_cast<size_t>(w->mask_nu) * w->mask_nv;
    for (size_t i = 0; i < cells; ++i)
      if (!(w->mask[i] >= 0.0 && w->mask[i] <= 1.0)) return fail(ctx, HMC_EINVAL, "mask values must lie in 0..1");
    if ((rc = put(w->mask, cells, &r.mask))) return rc;
    r.nu = w->mask_nu;
    r.nv = w->mask_nv;
    r.wu = w->width_u;
    r.wv = w->width_v;
    for (int k = 0; k < 3; ++k) {
      r.c[k] = w->center[k];
      r.u[k] = w->u_axis[k];
      r.v[k] = w->v_axis[k];
    }
    r.nrm[0] = r.u[1] * r.v[2] - r.u[2] * r.v[1];
    r.nrm[1] = r.u[2] * r.v[0] - r.u[0] * r.v[2];
    r.nrm[2] = r.u[0] * r.v[1] - r.u[1] * r.v[0];
  } else {
    return fail(ctx, HMC_EINVAL, "unknown ray weight kind");
  }
  r.kind = w->kind;
  return HMC_OK;
}

int hmc_set_ray_weights(hmc_ctx* ctx, const hmc_ray_weights* w) {
  if (!ctx) return fail(ctx, HMC_EINVAL, "null");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no generator may be reading the buffers
  for (void* p : ctx->rw_buffers) free_ptr(p);
  ctx->rw_buffers.clear();
  ctx->rw = DevRayWeights{};
  if (!w || w->kind == 0) return HMC_OK;
  DevRayWeights r{};
  const int rc = build_ray_weights(ctx, w, true, r, ctx->rw_buffers);
  if (rc) {
    for (void* p : ctx->rw_buffers) free_ptr(p);
    ctx->rw_buffers.clear();
    return rc;
  }
  ctx->rw = r;
  return HMC_OK;
}

int hmc_ray_weight_factors(hmc_ctx* ctx, int64_t n, const float* energy, const float* position,
                           const float* direction, double* out) {
  if (!ctx || n < 1 || !energy || !position || !direction || !out) return fail(ctx, HMC_EINVAL, "null or empty");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  float *d_e = nullptr, *d_p = nullptr, *d_d = nullptr;
  double* d_out = nullptr;
  int rc = HMC_OK;
  if ((rc = upload(ctx, &d_e, energy, static_cast<size_t>(n))) ||
      (rc = upload(ctx, &d_p, position, static_cast<size_t>(n) * 3)) ||
      (rc = upload(ctx, &d_d, direction, static_cast<size_t>(n) * 3)) ||
      (rc = dev_alloc(ctx, &d_out, static_cast<size_t>(n)))) {
  } else {
    hipLaunchKernelGGL(k_ray_weights, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, ctx->rw,
                       static_cast<long long>(n), d_e, d_p, d_d, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fail(ctx, HMC_EHIP, hipGetErrorString(e));
  }
  for (void* p : {static_cast<void*>(d_e), static_cast<void*>(d_p), static_cast<void*>(d_d),
                  static_cast<void*>(d_out)}) {
    if (p) {
      ctx->alloc_bytes.erase(p);
      free_ptr(p);
    }
  }
  return rc;
}

// ------------------------------------------------------------ head pre-solve

constexpr int kPresolveMaxDevices = 16;
constexpr int kPresolveSlots = 4;           // stack photon, air photon, air electron, transmitted
constexpr uint64_t kPresolveSalt = 0x48454144ULL;

struct DevPresolve {
  int first_compton, has_air, n_table, air_row;
  double exit_z, air_z0, air_z1, pcut, ecut, floor;
  double o[3], b[9];
  const double* log_e;
  const double* log_tot;  // [rows][n_table]
  const double* log_com;
  double le0, inv_dle;
};

struct PresolveRecord {
  float e, x, y, z, ux, uy, uz, w;
  int kind;  // -1: not kept
};

__device__ inline double presolve_mu(const DevPresolve& P, const double* tab, int row, double e) {
  return log_table(P.log_e, tab + static_cast<long long>(row) * P.n_table, P.n_table, P.le0, P.inv_dle, log(e));
}

__device__ inline double roulette_below(double w, double floor, uint64_t& rng) {
  if (floor <= 0.0 || w <= 0.0 || w >= floor) return w;
  return static_cast<double>(uniform(rng)) < w / floor ? floor : 0.0;
}

// Move a particle along its ray to the exit plane; kept when forward (local w), at or
// above the cutoff, with positive weight and the plane ahead (head.py:563-583).
__device__ inline int exit_record(const DevPresolve& P, PresolveRecord& r, int kind, double e, const double* s,
                                  float dx, float dy, float dz, double w, double cutoff) {
  r.kind = -1;
  if (!(w > 0.0) || !(e >= cutoff)) return 0;
  const double d_w = P.b[6] * dx + P.b[7] * dy + P.b[8] * dz;
  if (!(d_w > 0.0)) return 0;
  const double p_w = P.b[6] * (s[0] - P.o[0]) + P.b[7] * (s[1] - P.o[1]) + P.b[8] * (s[2] - P.o[2]);
  const double t = (P.exit_z - p_w) / d_w;
  if (!(t >= 0.0)) return 0;
  r.e = static_cast<float>(e);
  r.x = static_cast<float>(s[0] + t * dx);
  r.y = static_cast<float>(s[1] + t * dy);
  r.z = static_cast<float>(s[2] + t * dz);
  r.ux = dx;
  r.uy = dy;
  r.uz = dz;
  r.w = static_cast<float>(w);
  r.kind = kind;
  return 1;
}

__device__ inline void to_beam(const DevPresolve& P, const double* q, const double* d, double* pl, double* dl) {
  const double r[3] = {q[0] - P.o[0], q[1] - P.o[1], q[2] - P.o[2]};
  for (int i = 0; i < 3; ++i) {
    pl[i] = P.b[3 * i] * r[0] + P.b[3 * i + 1] * r[1] + P.b[3 * i + 2] * r[2];
    dl[i] = P.b[3 * i] * d[0] + P.b[3 * i + 1] * d[1] + P.b[3 * i + 2] * d[2];
  }
}

// One lane per primary (pyradmc/backends/warp/presolve.py:356-541). Records go to the
// primary's fixed slots; the block's kept count goes to block_counts for the
// deterministic compaction that follows.
__global__ void __launch_bounds__(256) k_presolve(DevRayWeights G, DevPresolve P, DevSource S, int generate,
                                                  long long n, uint64_t seed, const float* energy,
                                                  const float* px, const float* py, const float* pz,
                                                  const float* pux, const float* puy, const float* puz,
                                                  const float* pw, PresolveRecord* out, long long* block_counts) {
  __shared__ int kept_block;
  if (threadIdx.x == 0) kept_block = 0;
  __syncthreads();
  const long long h = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  int kept = 0;
  if (h < n) {
    PresolveRecord* rec = out + kPresolveSlots * h;
    uint64_t rng;
    double e, w0, p[3], d[3];
    if (generate) {  // the source's draws first, then the head's, on the history's stream
      rng = rng_init(seed, static_cast<uint64_t>(h));
      const Primary q = make_primary(S, 0, h, rng);
      e = q.e;
      w0 = q.w;
      p[0] = q.x; p[1] = q.y; p[2] = q.z;
      d[0] = q.ux; d[1] = q.uy; d[2] = q.uz;
    } else {
      rng = rng_init(seed ^ kPresolveSalt, static_cast<uint64_t>(h));
      e = energy[h];
      w0 = pw[h];
      p[0] = px[h]; p[1] = py[h]; p[2] = pz[h];
      d[0] = pux[h]; d[1] = puy[h]; d[2] = puz[h];
    }
    double pl[3], dl[3];
    to_beam(P, p, d, pl, dl);
    double tau[kPresolveMaxDevices];
    double tau_total = 0.0;
    for (int j = 0; j < G.n_dev; ++j) {
      const double t = device_chord(G, j, pl, dl, false);
      tau[j] = t > 0.0 ? presolve_mu(P, P.log_tot, j, e) * t : 0.0;
      tau_total += tau[j];
    }
    double transmitted = w0 * exp(-tau_total);
    rec[0].kind = rec[1].kind = rec[2].kind = -1;
    if (P.first_compton && tau_total > 0.0 && isfinite(tau_total) && dl[2] != 0.0) {
      // the device by its share of the optical depth (one the ray crosses)
      const double target = static_cast<double>(uniform(rng)) * tau_total;
      double cum = 0.0;
      int jd = -1;
      for (int j = 0; j < G.n_dev; ++j) {
        cum += tau[j];
        if (jd < 0 && tau[j] > 0.0 && cum >= target) jd = j;
      }
      if (jd < 0) jd = G.n_dev - 1;
      const double cf = presolve_mu(P, P.log_com, jd, e) / presolve_mu(P, P.log_tot, jd, e);
      const double t_mid = (0.5 * (G.z_top[jd] + G.z_bot[jd]) - pl[2]) / dl[2];
      const double site[3] = {p[0] + t_mid * d[0], p[1] + t_mid * d[1], p[2] + t_mid * d[2]};
      const float ratio = sample_compton_ratio(static_cast<float>(e), rng);
      const double e_out = static_cast<double>(ratio) * e;
      float ox = static_cast<float>(d[0]), oy = static_cast<float>(d[1]), oz = static_cast<float>(d[2]);
      rotate(ox, oy, oz, compton_cos(static_cast<float>(e), ratio), kTwoPi * uniform(rng));
      const double dout[3] = {ox, oy, oz};
      double sl[3], sdl[3];
      to_beam(P, site, dout, sl, sdl);
      double tau_out = 0.0;
      for (int j = 0; j < G.n_dev; ++j) {
        const double t = device_chord(G, j, sl, sdl, true);
        if (t > 0.0) tau_out += presolve_mu(P, P.log_tot, j, e_out) * t;
      }
      const double ws = roulette_below(w0 * (1.0 - exp(-tau_total)) * cf * exp(-tau_out), P.floor, rng);
      kept += exit_record(P, rec[0], PHOTON, e_out, site, ox, oy, oz, ws, P.pcut);
    }
    if (P.has_air) {
      const double length = dl[2] > 0.0 ? (P.air_z1 - P.air_z0) / dl[2] : 0.0;
      const double mu_air = presolve_mu(P, P.log_tot, P.air_row, e);
      const double tau_air = mu_air * length;
      const double before = transmitted;
      transmitted *= exp(-tau_air);
      if (P.first_compton && dl[2] > 0.0 && tau_air > 0.0) {
        const double pair_w = before * (1.0 - exp(-tau_air)) * presolve_mu(P, P.log_com, P.air_row, e) / mu_air;
        const double u = uniform(rng);
        const double depth = -log(1.0 - u * (1.0 - exp(-tau_air))) / mu_air;
        const double ts = (P.air_z0 - pl[2]) / dl[2] + depth;
        const double site[3] = {p[0] + ts * d[0], p[1] + ts * d[1], p[2] + ts * d[2]};
        const float fe = static_cast<float>(e);
        const float ratio = sample_compton_ratio(fe, rng);
        const double e_ph = static_cast<double>(ratio) * e;
        const float phi = kTwoPi * uniform(rng);
        float ax = static_cast<float>(d[0]), ay = static_cast<float>(d[1]), az = static_cast<float>(d[2]);
        float bx = ax, by = ay, bz = az;
        rotate(ax, ay, az, compton_cos(fe, ratio), phi);
        rotate(bx, by, bz, compton_electron_cos(fe, ratio), phi + kPi);
        kept += exit_record(P, rec[1], PHOTON, e_ph, site, ax, ay, az, roulette_below(pair_w, P.floor, rng), P.pcut);
        kept += exit_record(P, rec[2], ELECTRON, e - e_ph, site, bx, by, bz, roulette_below(pair_w, P.floor, rng),
                            P.ecut);
      }
    }
    kept += exit_record(P, rec[3], PHOTON, e, p, static_cast<float>(d[0]), static_cast<float>(d[1]),
                        static_cast<float>(d[2]), transmitted, P.pcut);
  }
  if (kept) atomicAdd(&kept_block, kept);
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = kept_block;
}

// Kept records of each block, in (primary, slot) order, from the block's offset on;
// weights scaled per primary. Same grid as k_presolve.
__global__ void __launch_bounds__(256) k_presolve_compact(const PresolveRecord* rec, long long n,
                                                          const long long* block_offsets, int32_t* out_kind,
                                                          float* out, long long pop_n, float scale) {
  __shared__ int scan[256];
  const int t = threadIdx.x;
  const long long h = static_cast<long long>(blockIdx.x) * blockDim.x + t;
  int mine = 0;
  if (h < n)
    for (int k = 0; k < kPresolveSlots; ++k) mine += rec[kPresolveSlots * h + k].kind >= 0;
  scan[t] = mine;
  __syncthreads();
  for (int step = 1; step < 256; step <<= 1) {  // inclusive Hillis-Steele scan
    const int add = t >= step ? scan[t - step] : 0;
    __syncthreads();
    scan[t] += add;
    __syncthreads();
  }
  if (h >= n || mine == 0) return;
  long long j = block_offsets[blockIdx.x] + scan[t] - mine;
  for (int k = 0; k < kPresolveSlots; ++k) {
    const PresolveRecord& r = rec[kPresolveSlots * h + k];
    if (r.kind < 0) continue;
    out_kind[j] = r.kind;
    const float cols[8] = {r.e, r.x, r.y, r.z, r.ux, r.uy, r.uz, r.w * scale};
    for (int c = 0; c < 8; ++c) out[c * pop_n + j] = cols[c];
    ++j;
  }
}

int hmc_presolve_head(hmc_ctx* ctx, uint64_t seed, int64_t n, const hmc_source* generator, const float* energy,
                      const float* x, const float* y, const float* z, const float* ux, const float* uy,
                      const float* uz, const float* weight, const hmc_ray_weights* stack, const hmc_presolve* q,
                      int64_t* n_out) {
  if (!ctx || !q || !n_out) return fail(ctx, HMC_EINVAL, "null");
  if (!generator && (!energy || !x || !y || !z || !ux || !uy || !uz || !weight))
    return fail(ctx, HMC_EINVAL, "pre-solve primaries need a generator or all eight columns");
  if (n < 1 || n > (static_cast<int64_t>(1) << 31) - 1) return fail(ctx, HMC_EINVAL, "primaries out of range");
  const int n_dev = stack && stack->kind == 1 ? stack->n_devices : 0;
  if (stack && stack->kind != 0 && stack->kind != 1) return fail(ctx, HMC_EINVAL, "the stack must be kind 1");
  if (n_dev > kPresolveMaxDevices) return fail(ctx, HMC_EINVAL, "at most 16 devices in a pre-solve stack");
  const int rows = n_dev + (q->has_air ? 1 : 0);
  if (q->n_rows != rows) return fail(ctx, HMC_EINVAL, "table rows must be the devices then the air column");
  if (rows > 0 && (q->n_table < 2 || !q->log_energies || !q->log_mu_total || !q->log_mu_compton))
    return fail(ctx, HMC_EINVAL, "attenuation tables incomplete");
  for (int k = 1; k < q->n_table && rows > 0; ++k)
    if (!(q->log_energies[k] > q->log_energies[k - 1])) return fail(ctx, HMC_EINVAL, "energies must increase");
  if (q->has_air && !(q->air_z_end > q->air_z_start)) return fail(ctx, HMC_EINVAL, "empty air column");
  // Kahn's rejection accepts only for a positive finite energy
  if (generator) {
    const int k = generator->kind;
    const bool mono = k == HMC_SRC_PENCIL || k == HMC_SRC_PARALLEL;
    const bool fan = k == HMC_SRC_SPECTRAL_FAN || k == HMC_SRC_SPOT_FAN;
    if (!mono && !fan && k != HMC_SRC_PARALLEL_SPECTRAL)
      return fail(ctx, HMC_EINVAL, "the pre-solve generates open-field photon sources only");
    if (mono && (generator->particle != 0 || !(generator->energy > 0.0) || !std::isfinite(generator->energy)))
      return fail(ctx, HMC_EINVAL, "the pre-solve needs photons of positive energy");
    if (fan && generator->n_beamlets != 1) return fail(ctx, HMC_EINVAL, "the pre-solve takes one aperture");
    if (!mono && (generator->n_bins < 1 || !generator->edges || !(generator->edges[0] > 0.0)))
      return fail(ctx, HMC_EINVAL, "the pre-solve needs a spectrum of positive energies");
  } else {
    for (int64_t i = 0; i < n; ++i)
      if (!(energy[i] > 0.0f) || !std::isfinite(energy[i]) || !std::isfinite(weight[i]) || !(weight[i] >= 0.0f))
        return fail(ctx, HMC_EINVAL, "pre-solve primaries need positive finite energies and weights");
  }
  if (generator) {
    const int src_rc = hmc_set_source(ctx, generator);
    if (src_rc) return src_rc;
  }
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<void*> bufs;
  auto release = [&]() {
    for (void* p : bufs) free_ptr(p);
    bufs.clear();
  };
  auto take = [&](auto** ptr, const auto* host, size_t count) -> int {
    const int code = upload(ctx, ptr, host, count);
    if (*ptr) {
      bufs.push_back(*ptr);
      ctx->alloc_bytes.erase(*ptr);
    }
    return code;
  };
  DevRayWeights G{};
  int rc = HMC_OK;
  if (n_dev > 0 && (rc = build_ray_weights(ctx, stack, false, G, bufs))) {
    release();
    return rc;
  }
  DevPresolve P{};
  P.first_compton = q->first_compton ? 1 : 0;
  P.has_air = q->has_air ? 1 : 0;
  P.air_row = n_dev;
  P.exit_z = q->exit_z;
  P.air_z0 = q->air_z_start;
  P.air_z1 = q->air_z_end;
  P.pcut = q->pcut;
  P.ecut = q->ecut;
  P.floor = q->weight_floor;
  for (int k = 0; k < 3; ++k) P.o[k] = q->origin[k];
  for (int k = 0; k < 9; ++k) P.b[k] = q->basis[k];
  double *d_le = nullptr, *d_tot = nullptr, *d_com = nullptr;
  if (rows > 0) {
    P.n_table = q->n_table;
    const size_t cells = static_cast<size_t>(rows) * q->n_table;
    if ((rc = take(&d_le, q->log_energies, q->n_table)) || (rc = take(&d_tot, q->log_mu_total, cells)) ||
        (rc = take(&d_com, q->log_mu_compton, cells))) {
      release();
      return rc;
    }
    P.log_e = d_le;
    P.log_tot = d_tot;
    P.log_com = d_com;
    P.le0 = q->log_energies[0];
    P.inv_dle = (q->n_table - 1) / (q->log_energies[q->n_table - 1] - q->log_energies[0]);
  }
  const float* host_cols[8] = {energy, x, y, z, ux, uy, uz, weight};
  float* dev_cols[8] = {};
  for (int c = 0; c < 8 && !generator; ++c)
    if ((rc = take(&dev_cols[c], host_cols[c], static_cast<size_t>(n)))) {
      release();
      return rc;
    }
  const long long blocks = (n + 255) / 256;
  PresolveRecord* d_rec = nullptr;
  long long* d_counts = nullptr;
  if ((rc = take(&d_rec, static_cast<const PresolveRecord*>(nullptr), static_cast<size_t>(n) * kPresolveSlots)) ||
      (rc = take(&d_counts, static_cast<const long long*>(nullptr), static_cast<size_t>(blocks)))) {
    release();
    return rc;
  }
  hipLaunchKernelGGL(k_presolve, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, ctx->stream, G, P, ctx->src,
                     generator ? 1 : 0, static_cast<long long>(n), seed, dev_cols[0], dev_cols[1], dev_cols[2],
                     dev_cols[3],
                     dev_cols[4], dev_cols[5], dev_cols[6], dev_cols[7], d_rec, d_counts);
  hipError_t e = hipGetLastError();
  std::vector<long long> offsets(static_cast<size_t>(blocks));
  if (e == hipSuccess)
    e = hipMemcpyAsync(offsets.data(), d_counts, sizeof(long long) * blocks, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    release();
    return fail(ctx, HMC_EHIP, hipGetErrorString(e));
  }
  long long total = 0;
  for (auto& c : offsets) {
    const long long k = c;
    c = total;
    total += k;
  }
  if (total == 0) {
    release();
    return fail(ctx, HMC_EINVAL, "the pre-solve kept no particles; check the geometry and cutoffs");
  }
  // the population replaces the resident one
  ctx->pop_n = ctx->pop_written = 0;
  if (ctx->src.kind == HMC_SRC_RESIDENT) ctx->src.up_n = 0;
  if ((rc = dev_alloc(ctx, &ctx->d_pop_particle, static_cast<size_t>(total))) ||
      (rc = dev_alloc(ctx, &ctx->d_pop, static_cast<size_t>(total) * 8))) {
    release();
    return rc;
  }
  long long* d_offsets = nullptr;
  if ((rc = take(&d_offsets, offsets.data(), offsets.size()))) {
    release();
    return rc;
  }
  hipLaunchKernelGGL(k_presolve_compact, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, ctx->stream, d_rec,
                     static_cast<long long>(n), d_offsets, ctx->d_pop_particle, ctx->d_pop, total,
                     static_cast<float>(static_cast<double>(total) / static_cast<double>(n)));
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  release();
  if (e != hipSuccess) return fail(ctx, HMC_EHIP, hipGetErrorString(e));
  ctx->pop_n = ctx->pop_written = total;
  *n_out = total;
  return HMC_OK;
}

int hmc_get_population(hmc_ctx* ctx, int64_t offset, int64_t count, int32_t* particle, float* energy, float* x,
                       float* y, float* z, float* ux, float* uy, float* uz, float* weight) {
  if (!ctx || !particle || !energy || !x || !y || !z || !ux || !uy || !uz || !weight)
    return fail(ctx, HMC_EINVAL, "null");
  if (count < 0 || offset < 0 || offset > ctx->pop_written - count)
    return fail(ctx, HMC_EINVAL, "population rows out of range");
  if (count == 0) return HMC_OK;
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  HMC_HIP(ctx, hipMemcpy(particle, ctx->d_pop_particle + offset, sizeof(int32_t) * count, hipMemcpyDeviceToHost));
  float* cols[8] = {energy, x, y, z, ux, uy, uz, weight};
  for (int c = 0; c < 8; ++c)
    HMC_HIP(ctx, hipMemcpy(cols[c], ctx->d_pop + c * ctx->pop_n + offset, sizeof(float) * count,
                           hipMemcpyDeviceToHost));
  return HMC_OK;
}

int hmc_set_physics(hmc_ctx* ctx, const hmc_physics* p) {
  if (!ctx || !p) return fail(ctx, HMC_EINVAL, "null");
  if (p->pcut <= 0 || p->ecut <= 0) return fail(ctx, HMC_EINVAL, "non-positive cutoff");
  if (p->step_energy_fraction <= 0) return fail(ctx, HMC_EINVAL, "non-positive step fraction");
  ctx->phys = *p;
  ctx->has_physics = true;
  return HMC_OK;
}

int hmc_set_profiling(hmc_ctx* ctx, int enabled) {
  if (!ctx) return HMC_EINVAL;
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  ctx->profile = enabled != 0;  // events are created by the first profiled transport
  return HMC_OK;
}

int hmc_set_launch(hmc_ctx* ctx, int round_lanes_per_cu, int refill_block, int push_chunk, int use_graphs) {
  if (!ctx) return HMC_EINVAL;
  if (round_lanes_per_cu < 256 || round_lanes_per_cu > 8192 || round_lanes_per_cu % 256 != 0)
    return fail(ctx, HMC_EINVAL, "round lanes per CU must be a multiple of 256 in [256, 8192]");
  if (refill_block < 64 || refill_block > 1024 || refill_block % 64 != 0)
    return fail(ctx, HMC_EINVAL, "refill block must be a multiple of 64 in [64, 1024]");
  if (push_chunk < 64 || push_chunk > 1024 || push_chunk % 64 != 0)
    return fail(ctx, HMC_EINVAL, "push chunk must be a multiple of 64 in [64, 1024]");
  ctx->round_lanes_per_cu = round_lanes_per_cu;
  ctx->refill_block = refill_block;
  ctx->push_chunk = push_chunk;
  ctx->use_graphs = use_graphs != 0;
  return HMC_OK;
}

// Enqueue one group of kRoundsPerSync photon / electron rounds on the context's stream,
// starting from queue roles (0, 1 | 2, 3): photons and electrons in queues 0 and 1,
// outputs to 2 and 3, roles swapped after each round. The group has an even number of
// rounds, so it ends with the starting roles and every group is the same sequence of
// launches. With events (profiling), each round's two kernels are bracketed.
constexpr int kRoundsPerSync = 4;

// ------------------------------------------------------------- electron energy sort
//
// Before the first electron round of a group, its queue is reordered by energy into
// kEnergyBins bins of a quarter octave above the cutoff, highest energies first, with
// padding records dropped. The persistent lanes then take the longest histories first
// and finish the round on short ones: the round's tail, where most lanes of the grid
// are idle, gets shorter (ascending order instead made the round 8 % slower than no
// sort). Each particle carries its own random stream and the tallies are integer sums,
// so the results do not depend on the order.
// A counting sort without global atomics: kSortBlocks blocks each take one contiguous
// range of the queue, count it per bin (k_energy_count), one block scans the counts in
// (bin, block) order into each block's first slot per bin (k_energy_scan), and each
// block moves its records to the slots that LDS cursors hand out (k_energy_scatter).
__device__ inline int energy_bin(float e, float inv_ecut) {
  const int b = static_cast<int>(4.0f * __log2f(fmaxf(e * inv_ecut, 1.0f)));
  return kEnergyBins - 1 - min(max(b, 0), kEnergyBins - 1);  // 0: the highest energies
}

// Block blockIdx.x's range [lo, hi) of the queue's n records.
__device__ inline void sort_range(unsigned int n, unsigned int& lo, unsigned int& hi) {
  const unsigned int per = (n + kSortBlocks - 1) / kSortBlocks;
  lo = min(blockIdx.x * per, n);
  hi = min(lo + per, n);
}

// hist[bin * kSortBlocks + block]: the block's records of that bin.
__global__ void __launch_bounds__(256) k_energy_count(const Particle* __restrict__ q, const unsigned int* count,
                                                      unsigned int cap, float inv_ecut, unsigned int* hist) {
  __shared__ unsigned int h[kEnergyBins];
  if (threadIdx.x < kEnergyBins) h[threadIdx.x] = 0u;
  __syncthreads();
  unsigned int lo, hi;
  sort_range(min(*count, cap), lo, hi);
  for (unsigned int i = lo + threadIdx.x; i < hi; i += blockDim.x)
    if ((q[i].meta & 3u) != kEmptyKind) atomicAdd(&h[energy_bin(q[i].e, inv_ecut)], 1u);
  __syncthreads();
  if (threadIdx.x < kEnergyBins) hist[threadIdx.x * kSortBlocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of hist in place (one block of 1024 threads, 32 entries each), the
// sorted queue's count and head. The count it replaces is a consumed queue's, which
// k_reset_consumed would have checked against the capacity: that check is made here.
__global__ void __launch_bounds__(1024) k_energy_scan(unsigned int* hist, unsigned int* sorted_count,
                                                      unsigned int* sorted_head, unsigned int capacity,
                                                      unsigned long long* ledger) {
  constexpr int kPer = kEnergyBins * kSortBlocks / 1024;
  __shared__ unsigned int part[1024];
  const int t = static_cast<int>(threadIdx.x);
  unsigned int v[kPer];
  unsigned int sum = 0u;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    v[j] = hist[t * kPer + j];
    sum += v[j];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned int x = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  unsigned int run = part[t] - sum;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    hist[t * kPer + j] = run;
    run += v[j];
  }
  if (t == 1023) {
    if (*sorted_count > capacity) atomicMax(ledger + L_OVERFLOW, static_cast<unsigned long long>(*sorted_count));
    *sorted_count = part[1023];
    *sorted_head = 0u;
  }
}

// Per 256 records: the block's bins for them in LDS (rank within the slice's bin by
// LDS atomics, the slice's per-bin offsets by one wave's scan), the records regrouped
// by bin in LDS, and then written out by consecutive threads, so that the lanes of a
// wave write runs of consecutive slots.
__global__ void __launch_bounds__(256) k_energy_scatter(const Particle* __restrict__ q, const unsigned int* count,
                                                        unsigned int cap, float inv_ecut,
                                                        const unsigned int* __restrict__ first,
                                                        Particle* __restrict__ dst) {
  __shared__ unsigned int cursor[kEnergyBins], slice_count[kEnergyBins], slice_off[kEnergyBins];
  __shared__ uint4 stage[256][3];  // records as three 16-byte words (x y z e | ux uy uz w | rng meta pad)
  __shared__ int stage_bin[256];
  if (threadIdx.x < kEnergyBins) cursor[threadIdx.x] = first[threadIdx.x * kSortBlocks + blockIdx.x];
  unsigned int lo, hi;
  sort_range(min(*count, cap), lo, hi);
  for (unsigned int s0 = lo; s0 < hi; s0 += blockDim.x) {
    if (threadIdx.x < kEnergyBins) slice_count[threadIdx.x] = 0u;
    __syncthreads();
    const unsigned int i = s0 + threadIdx.x;
    int bin = -1;
    uint4 w0 = make_uint4(0u, 0u, 0u, 0u), w1 = w0, w2 = w0;
    if (i < hi) {
      const uint4* rec = reinterpret_cast<const uint4*>(q + i);
      w0 = rec[0];
      w1 = rec[1];
      w2 = rec[2];
      if ((w2.z & 3u) != kEmptyKind) bin = energy_bin(__uint_as_float(w0.w), inv_ecut);
    }
    const unsigned int rank = bin >= 0 ? atomicAdd(&slice_count[bin], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the slice's 32 bin counts by the first wave
      const unsigned int c = threadIdx.x < kEnergyBins ? slice_count[threadIdx.x] : 0u;
      unsigned int incl = c;
      for (int off = 1; off < kEnergyBins; off <<= 1) {
        const unsigned int x = __shfl_up(incl, off);
        if (static_cast<int>(threadIdx.x) >= off) incl += x;
      }
      if (threadIdx.x < kEnergyBins) slice_off[threadIdx.x] = incl - c;
    }
    __syncthreads();
    const unsigned int n_real = slice_off[kEnergyBins - 1] + slice_count[kEnergyBins - 1];
    if (bin >= 0) {
      const unsigned int at = slice_off[bin] + rank;
      stage[at][0] = w0;
      stage[at][1] = w1;
      stage[at][2] = w2;
      stage_bin[at] = bin;
    }
    __syncthreads();
    if (threadIdx.x < n_real) {
      const int b = stage_bin[threadIdx.x];
      uint4* out = reinterpret_cast<uint4*>(dst + cursor[b] + (threadIdx.x - slice_off[b]));
      out[0] = stage[threadIdx.x][0];
      out[1] = stage[threadIdx.x][1];
      out[2] = stage[threadIdx.x][2];
    }
    __syncthreads();
    if (threadIdx.x < kEnergyBins) cursor[threadIdx.x] += slice_count[threadIdx.x];
  }
}

// The rounds' fast build applies: dose to medium, no deposit spreading, and a scoring
// grid with the transport grid's bounds, voxel sizes and counts.
static bool fast_rounds(const TransportParams& P) {
  const DevGrid& g = P.g;
  const DevScoring& s = P.s;
  return !P.dose_to_water && P.resolution == 0.0f && g.x_lo == s.x_lo && g.y_lo == s.y_lo &&
         g.z_lo == s.z_lo && g.x_hi == s.x_hi && g.y_hi == s.y_hi && g.z_hi == s.z_hi &&
         g.inv_sx == s.inv_sx && g.inv_sy == s.inv_sy && g.inv_sz == s.inv_sz && g.nx == s.nx &&
         g.ny == s.ny && g.nz == s.nz;
}

static int enqueue_round_group(hmc_ctx* ctx, const TransportParams& P, unsigned int round_blocks,
                               unsigned int cap, hipEvent_t* events) {
  const bool fast = fast_rounds(P);
  unsigned int* counts = ctx->d_counts;
  unsigned int* heads = ctx->d_counts + 4;
  int pa = 0, ea = 1, pb = 2, eb = 3;
  for (int r = 0; r < kRoundsPerSync; ++r) {
    Queue qa_p{ctx->d_q[pa], counts + pa, cap}, qa_e{ctx->d_q[ea], counts + ea, cap};
    Queue qb_p{ctx->d_q[pb], counts + pb, cap}, qb_e{ctx->d_q[eb], counts + eb, cap};
    const unsigned int pc = static_cast<unsigned int>(ctx->push_chunk);
    qa_p.push_chunk = qa_e.push_chunk = qb_p.push_chunk = qb_e.push_chunk = pc;
    hipEvent_t* ev = events ? events + 4 * r : nullptr;
    // Photon round: consume pa; electrons append to ea, photons to pb.
    if (ev) HMC_HIP(ctx, hipEventRecord(ev[0], ctx->stream));
    if (fast)
      hipLaunchKernelGGL((k_photon_round<true>), dim3(round_blocks), dim3(256), 0, ctx->stream, P, qa_p,
                         heads + pa, qa_e, qb_p);
    else
      hipLaunchKernelGGL((k_photon_round<false>), dim3(round_blocks), dim3(256), 0, ctx->stream, P, qa_p,
                         heads + pa, qa_e, qb_p);
    HMC_HIP(ctx, hipGetLastError());
    if (ev) HMC_HIP(ctx, hipEventRecord(ev[1], ctx->stream));
    // Electron round: consume ea (including the photon round's pushes). In the first
    // round of the group the queue is first sorted by energy into pa's records, which
    // the photon round has consumed; pa's count and head then serve the sorted queue
    // (both are reset with ea's after the round).
    Queue qin_e = qa_e;
    unsigned int* head_e = heads + ea;
    if (r == 0) {
      const float inv_ecut = 1.0f / P.ecut;
      unsigned int* hist = counts + 8;
      hipLaunchKernelGGL(k_energy_count, dim3(kSortBlocks), dim3(256), 0, ctx->stream, qa_e.data, qa_e.count, cap,
                         inv_ecut, hist);
      hipLaunchKernelGGL(k_energy_scan, dim3(1), dim3(1024), 0, ctx->stream, hist, counts + pa, heads + pa, cap,
                         ctx->d_ledger);
      hipLaunchKernelGGL(k_energy_scatter, dim3(kSortBlocks), dim3(256), 0, ctx->stream, qa_e.data, qa_e.count,
                         cap, inv_ecut, hist, qa_p.data);
      HMC_HIP(ctx, hipGetLastError());
      qin_e = Queue{qa_p.data, counts + pa, cap};
      head_e = heads + pa;
    }
    if (ev) HMC_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
    if (fast)
      hipLaunchKernelGGL((k_electron_round<true>), dim3(round_blocks), dim3(256), 0, ctx->stream, P, qin_e,
                         head_e, qb_p, qb_e);
    else
      hipLaunchKernelGGL((k_electron_round<false>), dim3(round_blocks), dim3(256), 0, ctx->stream, P, qin_e,
                         head_e, qb_p, qb_e);
    HMC_HIP(ctx, hipGetLastError());
    if (ev) HMC_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
    hipLaunchKernelGGL(k_reset_consumed, dim3(1), dim3(64), 0, ctx->stream, counts, heads, pa, ea, cap,
                       ctx->d_ledger);
    HMC_HIP(ctx, hipGetLastError());
    std::swap(pa, pb);
    std::swap(ea, eb);
  }
  return HMC_OK;
}

// Launch one group of rounds: as a replay of the captured graph when graphs are on
// (captured again whenever anything the launches take by value has changed), else
// launch by launch. Profiled groups always launch one by one: events recorded inside
// a captured graph cannot be timed (hipEventElapsedTime rejects them).
static int launch_round_group(hmc_ctx* ctx, const TransportParams& P, unsigned int round_blocks,
                              unsigned int cap, hipEvent_t* events, bool* replayed) {
  *replayed = false;
  if (!ctx->use_graphs || events) return enqueue_round_group(ctx, P, round_blocks, cap, events);
  std::vector<unsigned char> key(sizeof(P) + sizeof(ctx->d_q) + 3 * sizeof(unsigned int) + sizeof(void*) * 2);
  unsigned char* k = key.data();
  std::memcpy(k, &P, sizeof(P));
  k += sizeof(P);
  std::memcpy(k, ctx->d_q, sizeof(ctx->d_q));
  k += sizeof(ctx->d_q);
  const unsigned int ints[3] = {round_blocks, cap, static_cast<unsigned int>(ctx->push_chunk)};
  std::memcpy(k, ints, sizeof(ints));
  k += sizeof(ints);
  const void* ptrs[2] = {ctx->d_counts, ctx->d_ledger};
  std::memcpy(k, ptrs, sizeof(ptrs));
  if (!ctx->round_graph || key != ctx->round_graph_key) {
    if (ctx->round_graph) {
      HMC_HIP(ctx, hipGraphExecDestroy(ctx->round_graph));
      ctx->round_graph = nullptr;
    }
    hipGraph_t graph = nullptr;
    HMC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_round_group(ctx, P, round_blocks, cap, events);
    const hipError_t end = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != HMC_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HMC_HIP(ctx, end);
    const hipError_t inst = hipGraphInstantiate(&ctx->round_graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
      ctx->round_graph = nullptr;
      HMC_HIP(ctx, inst);
    }
    ctx->round_graph_key = key;
  }
  HMC_HIP(ctx, hipGraphLaunch(ctx->round_graph, ctx->stream));
  *replayed = true;
  return HMC_OK;
}

int hmc_transport(hmc_ctx* ctx, uint64_t seed, const hmc_work_item* items, int n_items,
                  int64_t chunk, int queue_factor, hmc_ledger* ledger) {
  if (!ctx || !items || n_items < 1 || chunk < 1 || queue_factor < 1 || !ledger)
    return fail(ctx, HMC_EINVAL, "bad transport arguments");
  if (!ctx->has_grid || !ctx->has_scoring || !ctx->has_tables || !ctx->has_source ||
      !ctx->has_physics)
    return fail(ctx, HMC_ESTATE, "grid, scoring, tables, source and physics must be set");
  if (ctx->phys.msc_gs && !ctx->has_gs) return fail(ctx, HMC_ESTATE, "GS grid not set");
  if (ctx->max_material >= ctx->t.n_mat)
    return fail(ctx, HMC_EINVAL, "voxel material index " + std::to_string(ctx->max_material) +
                                     " beyond the " + std::to_string(ctx->t.n_mat) +
                                     " material rows of the tables");
  HMC_HIP(ctx, hipSetDevice(ctx->device));

  std::vector<DevItem> dev_items(static_cast<size_t>(n_items));
  long long total = 0;
  int slots_written = 0;
  for (int i = 0; i < n_items; ++i) {
    if (items[i].count < 0 || items[i].column < 0 ||
        items[i].column >= ctx->n_columns * ctx->n_slots)
      return fail(ctx, HMC_EINVAL, "work item count/column out of range");
    slots_written = std::max(slots_written, static_cast<int>(items[i].column / ctx->n_columns) + 1);
    const bool per_beamlet = ctx->src.kind == HMC_SRC_LATTICE || ctx->src.kind == HMC_SRC_SPECTRAL_FAN ||
                             ctx->src.kind == HMC_SRC_SPOT_FAN;
    if (items[i].beamlet < 0 || (per_beamlet && items[i].beamlet >= ctx->src_beamlets))
      return fail(ctx, HMC_EINVAL, "work item beamlet out of range");
    dev_items[i].key_base = items[i].key_base;
    dev_items[i].start = total;
    dev_items[i].column = items[i].column;
    dev_items[i].beamlet = items[i].beamlet;
    total += items[i].count;
  }
  if (ctx->src.kind == HMC_SRC_UPLOADED && total > ctx->src.up_n)
    return fail(ctx, HMC_EINVAL, "more histories than uploaded primaries");
  if (ctx->src.kind == HMC_SRC_RESIDENT && ctx->src.up_n < 1)
    return fail(ctx, HMC_EINVAL, "resident source not set since its population was replaced");
  ctx->dirty_slots = std::max(ctx->dirty_slots, slots_written);
  int rc;
  if (static_cast<size_t>(n_items) > ctx->items_cap) {
    if ((rc = dev_alloc(ctx, &ctx->d_items, static_cast<size_t>(n_items)))) return rc;
    ctx->items_cap = static_cast<size_t>(n_items);
  }
  HMC_HIP(ctx, hipMemcpyAsync(ctx->d_items, dev_items.data(), sizeof(DevItem) * n_items,
                              hipMemcpyHostToDevice, ctx->stream));

  const long long chunk_n = std::min<long long>(chunk, std::max<long long>(total, 1));
  const size_t capacity = static_cast<size_t>(chunk_n) * static_cast<size_t>(queue_factor);
  // Queue heads overshoot the queue length by up to two refill blocks per wave.
  if (capacity > (1ull << 31)) return fail(ctx, HMC_EINVAL, "queue capacity exceeds 2^31");
  if (capacity > ctx->q_capacity) {
    for (int q = 0; q < 4; ++q)
      if ((rc = dev_alloc(ctx, &ctx->d_q[q], capacity))) return rc;
    ctx->q_capacity = capacity;
  }

  TransportParams P;
  std::memset(&P, 0, sizeof(P));  // padding too: the bytes key the captured graph
  P.g = ctx->g;
  P.s = ctx->s;
  P.t = ctx->t;
  P.edep = ctx->d_edep;
  P.ledger = ctx->d_ledger;
  P.pcut = static_cast<float>(ctx->phys.pcut);
  P.log_pcut = std::log(P.pcut);
  P.ecut = static_cast<float>(ctx->phys.ecut);
  P.step_fraction = static_cast<float>(ctx->phys.step_energy_fraction);
  P.resolution = static_cast<float>(ctx->phys.deposit_resolution_cm);
  P.transport_electrons = ctx->phys.transport_electrons;
  P.msc_gs = ctx->phys.msc_gs;
  P.dose_to_water = ctx->phys.dose_to_water;
  P.max_steps = 1 << 20;
  P.refill_block = static_cast<unsigned int>(ctx->refill_block);

  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, ctx->d_ledger,
                     static_cast<long long>(L_COUNT), 0ull);

  const unsigned int cap = static_cast<unsigned int>(ctx->q_capacity);
  // Rounds launch at full width: lanes of an exhausted queue exit at once, so the host
  // never needs a queue's size to launch it, and it enqueues kRoundsPerSync photon /
  // electron rounds between count reads.
  const unsigned int round_blocks = static_cast<unsigned int>(
      blocks_for(static_cast<long long>(std::max(ctx->cu_count, 1)) * ctx->round_lanes_per_cu));
  if (static_cast<long long>(round_blocks) * 4 * ctx->refill_block > (1ll << 30))
    return fail(ctx, HMC_EINVAL, "round width times refill block too large");
  long long overflow = 0, rounds = 0, launches[2] = {0, 0}, graph_launches = 0;
  double kernel_ms[2] = {0.0, 0.0};
  if (ctx->profile && ctx->ev_pool.size() < 4 * kRoundsPerSync) {
    for (size_t k = ctx->ev_pool.size(); k < 4 * kRoundsPerSync; ++k) {
      hipEvent_t e = nullptr;
      HMC_HIP(ctx, hipEventCreate(&e));
      ctx->ev_pool.push_back(e);
    }
  }
  unsigned int* counts = ctx->d_counts;
  // The ledger comes back with every count read (one host wait per group of rounds).
  const unsigned long long* host_ledger = ctx->h_stage + kStageLedger;

  for (long long p0 = 0; p0 < total && !overflow; p0 += chunk_n) {
    const long long n_gen = std::min(chunk_n, total - p0);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, counts, 8, 0u);
    const int pa = 0, ea = 1;  // queue roles at the start of every group of rounds
    Queue qpa{ctx->d_q[pa], counts + pa, cap}, qea{ctx->d_q[ea], counts + ea, cap};
    if (ctx->rw.kind != 0)
      hipLaunchKernelGGL(k_generate<true>, dim3(blocks_for(n_gen, 256, 2048)), dim3(256), 0,
                         ctx->stream, ctx->src, ctx->rw, ctx->d_items, n_items, p0, n_gen,
                         static_cast<uint64_t>(seed), qpa, qea, ctx->d_ledger);
    else
      hipLaunchKernelGGL(k_generate<false>, dim3(blocks_for(n_gen, 256, 2048)), dim3(256), 0,
                         ctx->stream, ctx->src, ctx->rw, ctx->d_items, n_items, p0, n_gen,
                         static_cast<uint64_t>(seed), qpa, qea, ctx->d_ledger);
    HMC_HIP(ctx, hipGetLastError());
    while (true) {
      bool replayed = false;
      if ((rc = launch_round_group(ctx, P, round_blocks, cap, ctx->profile ? ctx->ev_pool.data() : nullptr,
                                   &replayed)))
        return rc;
      graph_launches += replayed ? 1 : 0;
      rounds += kRoundsPerSync;
      launches[0] += kRoundsPerSync;
      launches[1] += kRoundsPerSync;
      HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage, counts, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                                  ctx->stream));
      HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage + kStageLedger, ctx->d_ledger, sizeof(unsigned long long) * L_COUNT,
                                  hipMemcpyDeviceToHost, ctx->stream));
      HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      unsigned int host_counts[4];
      std::memcpy(host_counts, ctx->h_stage, sizeof(host_counts));
      if (ctx->profile) {
        for (int r = 0; r < kRoundsPerSync; ++r) {
          for (int k = 0; k < 2; ++k) {
            float ms = 0.0f;
            HMC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_pool[4 * r + 2 * k], ctx->ev_pool[4 * r + 2 * k + 1]));
            kernel_ms[k] += ms;
          }
        }
      }
      for (int q = 0; q < 4; ++q)
        if (host_counts[q] > cap) overflow = std::max<long long>(overflow, host_counts[q]);
      if (overflow || (host_counts[pa] == 0 && host_counts[ea] == 0)) break;
      if (rounds > 100000) return fail(ctx, HMC_EHIP, "transport did not drain in 100000 rounds");
    }
  }
  if (rounds == 0) {  // no histories: nothing ran, the ledger holds its cleared state
    HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage + kStageLedger, ctx->d_ledger, sizeof(unsigned long long) * L_COUNT,
                                hipMemcpyDeviceToHost, ctx->stream));
    HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  overflow = std::max<long long>(overflow, static_cast<long long>(host_ledger[L_OVERFLOW]));
  auto L = [&](int slot) { return static_cast<int64_t>(host_ledger[slot]); };
  ledger->escaped = L(L_ESCAPED);
  ledger->unscored = L(L_UNSCORED);
  ledger->deposited = L(L_DEPOSITED);
  ledger->emitted = L(L_EMITTED);
  ledger->violations = L(L_VIOLATIONS);
  ledger->trapped = L(L_TRAPPED);
  ledger->overflow = overflow;
  ledger->rounds = rounds;
  ledger->photon_voxels = L(L_P_VOX);
  ledger->photon_deposits = L(L_P_DEP);
  ledger->photon_pushes = L(L_P_PUSH);
  ledger->photon_pops = L(L_P_POP);
  ledger->photon_steps = L(L_P_STEPS);
  ledger->electron_voxels = L(L_E_VOX);
  ledger->electron_deposits = L(L_E_DEP);
  ledger->electron_pushes = L(L_E_PUSH);
  ledger->electron_pops = L(L_E_POP);
  ledger->electron_steps = L(L_E_STEPS);
  ledger->photon_launches = launches[0];
  ledger->electron_launches = launches[1];
  ledger->photon_kernel_ms = kernel_ms[0];
  ledger->electron_kernel_ms = kernel_ms[1];
  ledger->graph_launches = graph_launches;
  if (overflow)
    return fail(ctx, HMC_EOVERFLOW,
                "secondary queue overflow (" + std::to_string(overflow) + " > capacity " +
                    std::to_string(cap) + "); raise queue_factor or lower chunk_size");
  if (ledger->violations)
    return fail(ctx, HMC_EMAJORANT,
                "Woodcock majorant violated in the kernel despite table headroom. The geometry "
                "contains material or density the majorant declaration did not cover.");
  return HMC_OK;
}

int hmc_fold(hmc_ctx* ctx, int n_slots_used, const double* nhist, int64_t* total_quanta) {
  if (!ctx || !nhist) return fail(ctx, HMC_EINVAL, "null");
  if (!ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  if (n_slots_used < 1 || n_slots_used > ctx->n_slots)
    return fail(ctx, HMC_EINVAL, "slot count out of range");
  const long long n_nh = ctx->n_columns * n_slots_used;
  for (long long c = 0; c < n_nh; ++c)
    if (!(nhist[c] > 0.0)) return fail(ctx, HMC_EINVAL, "empty batch for a column");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipMemcpyAsync(ctx->d_nhist, nhist, sizeof(double) * n_nh, hipMemcpyHostToDevice,
                              ctx->stream));
  unsigned long long* folded = ctx->d_ledger + L_COUNT;  // scratch slot past the ledger
  // The counter starts from zero unless earlier folds' quanta are still to be read.
  if (!ctx->fold_pending)
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, folded, 1ll, 0ull);
  const long long total = ctx->s.n_voxels * ctx->n_columns;
  hipLaunchKernelGGL(k_fold, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream,
                     ctx->d_edep, ctx->d_mass, ctx->d_nhist, ctx->s.n_voxels, ctx->n_columns,
                     n_slots_used, ctx->d_s1, ctx->d_s2, folded);
  HMC_HIP(ctx, hipGetLastError());
  if (n_slots_used >= ctx->dirty_slots) ctx->dirty_slots = 0;  // the fold zeroes what it reads
  if (!total_quanta) {  // no wait: the total comes back with the next hmc_finalize
    ctx->fold_pending = true;
    return HMC_OK;
  }
  HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage + kStageFolded, folded, sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->fold_pending = false;
  *total_quanta = static_cast<int64_t>(ctx->h_stage[kStageFolded]);
  return HMC_OK;
}

int hmc_folded_quanta(hmc_ctx* ctx, int64_t* total_quanta) {
  if (!ctx || !total_quanta) return fail(ctx, HMC_EINVAL, "null");
  if (ctx->fold_pending) {  // no finalize since those folds: read the counter now
    HMC_HIP(ctx, hipSetDevice(ctx->device));
    HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage + kStageFolded, ctx->d_ledger + L_COUNT, sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, ctx->stream));
    HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->folded_quanta = static_cast<long long>(ctx->h_stage[kStageFolded]);
    ctx->fold_pending = false;
  }
  *total_quanta = static_cast<int64_t>(ctx->folded_quanta);
  ctx->folded_quanta = 0;
  return HMC_OK;
}

int hmc_finalize(hmc_ctx* ctx, int n_batches, double* mean, double* sigma) {
  if (!ctx || n_batches < 1) return fail(ctx, HMC_EINVAL, "bad finalize arguments");
  if (!ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const long long total = ctx->s.n_voxels * ctx->n_columns;
  if (!ctx->finalized) {
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(total, 256, 65536)), dim3(256), 0,
                       ctx->stream, ctx->d_s1, ctx->d_s2, total, n_batches);
    HMC_HIP(ctx, hipGetLastError());
    ctx->finalized = true;
  }
  if (mean)
    HMC_HIP(ctx, hipMemcpyAsync(mean, ctx->d_s1, sizeof(double) * total, hipMemcpyDeviceToHost,
                                ctx->stream));
  if (sigma)
    HMC_HIP(ctx, hipMemcpyAsync(sigma, ctx->d_s2, sizeof(double) * total, hipMemcpyDeviceToHost,
                                ctx->stream));
  const bool take_folded = ctx->fold_pending;
  if (take_folded)
    HMC_HIP(ctx, hipMemcpyAsync(ctx->h_stage + kStageFolded, ctx->d_ledger + L_COUNT, sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (take_folded) {
    ctx->folded_quanta = static_cast<long long>(ctx->h_stage[kStageFolded]);
    ctx->fold_pending = false;
  }
  return HMC_OK;
}

int hmc_truncate(hmc_ctx* ctx, double truncation, int64_t* counts, int64_t* total_out) {
  if (!ctx || !counts || truncation < 0.0) return fail(ctx, HMC_EINVAL, "bad truncation arguments");
  if (!ctx->finalized) return fail(ctx, HMC_ESTATE, "finalize before truncation");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  const long long nv = ctx->s.n_voxels;
  const int chunks = static_cast<int>((nv + kTruncChunk - 1) / kTruncChunk);
  const long long n_blocks = static_cast<long long>(chunks) * ctx->n_columns;
  if (n_blocks > 0x7fffffffll) return fail(ctx, HMC_EINVAL, "too many truncation chunks");
  int rc;
  if ((rc = dev_alloc(ctx, &ctx->d_chunk_counts, static_cast<size_t>(n_blocks)))) return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_chunk_offsets, static_cast<size_t>(n_blocks)))) return rc;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, ctx->stream, ctx->d_colmax,
                     ctx->n_columns, 0ull);
  hipLaunchKernelGGL(k_column_max, dim3(static_cast<unsigned>(n_blocks)), dim3(kTruncBlock), 0,
                     ctx->stream, ctx->d_s1, nv, chunks, ctx->d_colmax);
  hipLaunchKernelGGL(k_count_kept, dim3(static_cast<unsigned>(n_blocks)), dim3(kTruncBlock), 0,
                     ctx->stream, ctx->d_s1, nv, chunks, ctx->d_colmax, truncation,
                     ctx->d_chunk_counts);
  HMC_HIP(ctx, hipGetLastError());
  std::vector<long long> chunk_counts(static_cast<size_t>(n_blocks));
  HMC_HIP(ctx, hipMemcpyAsync(chunk_counts.data(), ctx->d_chunk_counts,
                              sizeof(long long) * n_blocks, hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<long long> offsets(static_cast<size_t>(n_blocks));
  long long running = 0;
  for (long long c = 0; c < ctx->n_columns; ++c) {
    long long col_count = 0;
    for (int k = 0; k < chunks; ++k) {
      const long long b = c * chunks + k;
      offsets[b] = running;
      running += chunk_counts[b];
      col_count += chunk_counts[b];
    }
    counts[c] = col_count;
  }
  ctx->trunc_total = running;
  if (total_out) *total_out = running;
  if (static_cast<size_t>(running) > ctx->out_cap || !ctx->d_out_idx) {
    const size_t cap = static_cast<size_t>(std::max<long long>(running, 1));
    if ((rc = dev_alloc(ctx, &ctx->d_out_idx, cap))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_out_dose, cap))) return rc;
    if ((rc = dev_alloc(ctx, &ctx->d_out_sigma, cap))) return rc;
    ctx->out_cap = cap;
  }
  HMC_HIP(ctx, hipMemcpyAsync(ctx->d_chunk_offsets, offsets.data(), sizeof(long long) * n_blocks,
                              hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_compact, dim3(static_cast<unsigned>(n_blocks)), dim3(kTruncBlock), 0,
                     ctx->stream, ctx->d_s1, ctx->d_s2, nv, chunks, ctx->d_colmax, truncation,
                     ctx->d_chunk_offsets, ctx->d_out_idx, ctx->d_out_dose, ctx->d_out_sigma);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

static int check_windows(hmc_ctx* ctx, const int32_t* origins, int64_t n_cols, int wx, int wy) {
  if (!ctx || !origins) return fail(ctx, HMC_EINVAL, "null");
  if (!ctx->has_scoring) return fail(ctx, HMC_ESTATE, "scoring grid not set");
  if (n_cols != ctx->n_columns) return fail(ctx, HMC_EINVAL, "window count must equal the column count");
  if (wx < 1 || wy < 1 || wx > ctx->s.nx || wy > ctx->s.ny) return fail(ctx, HMC_EINVAL, "bad window size");
  for (int64_t c = 0; c < n_cols; ++c) {
    const int x0 = origins[2 * c], y0 = origins[2 * c + 1];
    if (x0 < 0 || y0 < 0 || x0 + wx > ctx->s.nx || y0 + wy > ctx->s.ny)
      return fail(ctx, HMC_EINVAL, "window outside the scoring grid");
  }
  return HMC_OK;
}

int hmc_gather_windows(hmc_ctx* ctx, const int32_t* origins, int64_t n_cols, int wx, int wy,
                       float* out) {
  int rc = check_windows(ctx, origins, n_cols, wx, wy);
  if (rc) return rc;
  if (!out) return fail(ctx, HMC_EINVAL, "null output");
  if (!ctx->finalized) return fail(ctx, HMC_ESTATE, "finalize before gathering windows");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  int* d_org = nullptr;
  if ((rc = upload(ctx, &ctx->d_origins, origins, static_cast<size_t>(2 * n_cols)))) return rc;
  d_org = ctx->d_origins;
  const long long total = n_cols * ctx->s.nz * wx * wy;
  hipLaunchKernelGGL(k_gather_windows, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream,
                     ctx->d_s1, d_org, static_cast<long long>(n_cols), wx, wy, ctx->s.ny, ctx->s.nz,
                     ctx->s.n_voxels, out);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

int hmc_scatter_windows(hmc_ctx* ctx, const int32_t* origins, int64_t n_cols, int wx, int wy,
                        const float* mean, const float* sigma) {
  int rc = check_windows(ctx, origins, n_cols, wx, wy);
  if (rc) return rc;
  if (!mean) return fail(ctx, HMC_EINVAL, "null input");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  if ((rc = upload(ctx, &ctx->d_origins, origins, static_cast<size_t>(2 * n_cols)))) return rc;
  const long long nvc = ctx->s.n_voxels * ctx->n_columns;
  const int blocks = blocks_for(nvc, 256, 65536);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(ctx->d_s1), nvc, 0ull);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(ctx->d_s2), nvc, 0ull);
  const long long total = n_cols * ctx->s.nz * wx * wy;
  hipLaunchKernelGGL(k_scatter_windows, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, ctx->stream,
                     ctx->d_s1, ctx->d_s2, ctx->d_origins, static_cast<long long>(n_cols), wx, wy,
                     ctx->s.ny, ctx->s.nz, ctx->s.n_voxels, mean, sigma);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->finalized = true;  // s1 / s2 now hold mean / sigma for hmc_truncate
  return HMC_OK;
}

int hmc_set_column_entries(hmc_ctx* ctx, int64_t column, int64_t n, const int64_t* indices, const double* dose,
                           const double* sigma) {
  if (!ctx) return fail(ctx, HMC_EINVAL, "null");
  if (!ctx->finalized) return fail(ctx, HMC_ESTATE, "finalize (or scatter windows) before setting entries");
  if (column < 0 || column >= ctx->n_columns) return fail(ctx, HMC_EINVAL, "column outside the scoring columns");
  if (n < 0 || (n > 0 && (!indices || !dose || !sigma))) return fail(ctx, HMC_EINVAL, "null entries");
  if (n == 0) return HMC_OK;
  for (int64_t i = 0; i < n; ++i)
    if (indices[i] < 0 || indices[i] >= ctx->s.n_voxels) return fail(ctx, HMC_EINVAL, "entry index outside the grid");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  if ((rc = upload(ctx, &ctx->d_entry_idx, reinterpret_cast<const long long*>(indices), static_cast<size_t>(n))))
    return rc;
  if ((rc = dev_alloc(ctx, &ctx->d_entry_val, static_cast<size_t>(2 * n)))) return rc;
  HMC_HIP(ctx, hipMemcpy(ctx->d_entry_val, dose, sizeof(double) * n, hipMemcpyHostToDevice));
  HMC_HIP(ctx, hipMemcpy(ctx->d_entry_val + n, sigma, sizeof(double) * n, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_set_entries, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, ctx->stream, ctx->d_s1, ctx->d_s2,
                     static_cast<long long>(column) * ctx->s.n_voxels, static_cast<long long>(n), ctx->d_entry_idx,
                     ctx->d_entry_val, ctx->d_entry_val + n);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

int hmc_host_alloc(size_t bytes, void** out) {
  if (!out) return HMC_EINVAL;
  *out = nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes == 0 ? 1 : bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return HMC_ENOMEM;
  }
  *out = p;
  return HMC_OK;
}

void hmc_host_free(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int hmc_fetch_csc_f32(hmc_ctx* ctx, int64_t* indices, float* dose, float* sigma) {
  if (!ctx) return fail(ctx, HMC_EINVAL, "null");
  const long long n = ctx->trunc_total;
  if (n == 0) return HMC_OK;
  if (!indices || !dose || !sigma) return fail(ctx, HMC_EINVAL, "null output");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  if ((rc = dev_alloc(ctx, &ctx->d_out32, static_cast<size_t>(2 * n)))) return rc;
  hipLaunchKernelGGL(k_narrow_f32, dim3(blocks_for(n, 256, 65536)), dim3(256), 0, ctx->stream, ctx->d_out_dose,
                     ctx->d_out_sigma, n, ctx->d_out32);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipMemcpyAsync(indices, ctx->d_out_idx, sizeof(long long) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  HMC_HIP(ctx, hipMemcpyAsync(dose, ctx->d_out32, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipMemcpyAsync(sigma, ctx->d_out32 + n, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

int hmc_fetch_csc(hmc_ctx* ctx, int64_t* indices, double* dose, double* sigma) {
  if (!ctx) return fail(ctx, HMC_EINVAL, "null");
  const long long n = ctx->trunc_total;
  if (n == 0) return HMC_OK;
  if (!indices || !dose || !sigma) return fail(ctx, HMC_EINVAL, "null output");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  HMC_HIP(ctx, hipMemcpyAsync(indices, ctx->d_out_idx, sizeof(long long) * n,
                              hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipMemcpyAsync(dose, ctx->d_out_dose, sizeof(double) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  HMC_HIP(ctx, hipMemcpyAsync(sigma, ctx->d_out_sigma, sizeof(double) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

int hmc_sample_physics(hmc_ctx* ctx, uint64_t seed, int64_t n, const hmc_sampler_args* args, float* out) {
  if (!ctx) return fail(ctx, HMC_EINVAL, "null");
  if (!args || !out) return fail(ctx, HMC_EINVAL, "null argument");
  if (n < 1 || n > (int64_t(1) << 24)) return fail(ctx, HMC_EINVAL, "sample count must be in [1, 2^24]");
  if (!(args->path_mu > 0.0) || !(args->compton_energy[0] > 0.0) || !(args->compton_energy[1] > 0.0) ||
      !(args->moller_energy > 2.0 * args->moller_cut) || !(args->moller_cut > 0.0) || args->hinge_ms < 0.0 ||
      !(args->brems_pcut > 0.0) || !(args->brems_energy > 0.0) || !(args->coherent_energy > 0.0))
    return fail(ctx, HMC_EINVAL, "sampler arguments out of range");
  const double mu_sum = args->channel_mu[0] + args->channel_mu[1] + args->channel_mu[2] + args->channel_mu[3];
  for (int k = 0; k < 4; ++k)
    if (args->channel_mu[k] < 0.0) return fail(ctx, HMC_EINVAL, "negative channel coefficient");
  if (!(mu_sum > 0.0)) return fail(ctx, HMC_EINVAL, "channel coefficients sum to zero");
  if (ctx->has_tables && (args->coherent_material < 0 || args->coherent_material >= ctx->t.n_mat))
    return fail(ctx, HMC_EINVAL, "coherent material has no table row");
  HMC_HIP(ctx, hipSetDevice(ctx->device));
  int rc;
  const size_t count = static_cast<size_t>(n) * HMC_SAMPLE_COLUMNS;
  if ((rc = dev_alloc(ctx, &ctx->d_samples, count))) return rc;
  hipLaunchKernelGGL(k_sample_physics, dim3(blocks_for(n, 256, 1LL << 20)), dim3(256), 0, ctx->stream, seed,
                     static_cast<long long>(n), *args, ctx->t, ctx->has_tables ? 1 : 0,
                     (ctx->has_tables && ctx->has_gs) ? 1 : 0, ctx->d_samples);
  HMC_HIP(ctx, hipGetLastError());
  HMC_HIP(ctx, hipMemcpyAsync(out, ctx->d_samples, count * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HMC_OK;
}

}  // extern "C"
