// Device-side building blocks of the transport kernels: counter-based RNG, table
// lookups, voxel geometry, and the scalar physics samplers.
//
// Every sampler restates one function of the reference's single-source physics
// (cited per function); the transport kernels in hipmc_kernels.hip sequence them in
// the order the reference step loops do (pyradmc/transport/photon.py:100-282,
// pyradmc/transport/electron.py:207-481). Arithmetic is float32 like the reference's
// GPU kernels; fixed-point tally conversion is float64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hmc {

constexpr float kElectronMass = 0.51099895069f;
constexpr float kTwoPi = 6.283185307179586f;
constexpr float kPi = 3.141592653589793f;
constexpr float kRouletteMeV = 0.5f;
constexpr float kRouletteSurvival = 0.5f;
constexpr float kRouletteWeightCap = 4.0f;
constexpr float kEntryNudge = 1.0e-4f;
constexpr float kBoundaryNudge = 1.0e-4f;
constexpr float kInfinityCm = 1.0e30f;
constexpr double kInvQuantum = 1.0e9;
constexpr float kGsBinsPerLog = 4.0f;
constexpr float kGsTheta2Min = 1.0e-12f;
constexpr float kHingeTheta2Max = 0.10f;
constexpr int kSubstepDepositCap = 1024;
constexpr float kPoleFloor = 1.0e-20f;

enum Kind : uint32_t { PHOTON = 0, ELECTRON = 1, POSITRON = 2 };
enum Process : int { COMPTON = 0, PHOTOELECTRIC = 1, PAIR = 2, RAYLEIGH = 3 };

// ----------------------------------------------------------------------------- RNG
//
// One 64-bit PCG state per particle (PCG-XSH-RR 64/32, O'Neill 2014). A history's
// stream is a pure function of (seed, key) through a SplitMix64 finalizer; a
// secondary gets a child stream seeded from two draws of its parent, so every
// trajectory depends only on its ancestry, never on which lane or queue slot ran
// it (the reproducibility scheme of pyradmc/rng/warp_shim.py:44-81). Uniforms carry
// 24 bits and lie in [0, 1 - 2^-24]: exactly 1.0 is impossible.

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t rng_init(uint64_t seed, uint64_t key) {
  uint64_t s = mix64(seed + 0x9E3779B97F4A7C15ULL);
  return mix64(s ^ (key * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
}

__host__ __device__ inline uint32_t rng_next(uint64_t& s) {
  const uint64_t old = s;
  s = old * 6364136223846793005ULL + 1442695040888963407ULL;
  const uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = static_cast<uint32_t>(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

__device__ inline float uniform(uint64_t& s) {
  return static_cast<float>(rng_next(s) >> 8) * 5.9604644775390625e-08f;
}

__device__ inline uint64_t spawn_stream(uint64_t& s) {
  const uint64_t a = rng_next(s);
  const uint64_t b = rng_next(s);
  return mix64((a << 32) ^ b ^ 0xA0761D6478BD642FULL);
}

// -------------------------------------------------------------------------- tables

struct DevTables {
  const float* majorant;  // [n]
  const float4* photon;   // [n_mat][n]: compton, photo, pair, rayleigh (cm^2/g)
  const float4* elec_a;   // [n_mat][n]: csda_range, moller, restricted_range, radiative
  const float4* elec_b;   // [n_mat][n]: scattering_power, log_eta, stopping_restricted, 0
  const float* inv_range; // [n_mat][n]: energy of restricted range (shared range grid)
  const float* coherent_x;          // [n_coherent]
  const float* coherent_cumulative; // [n_mat][n_coherent]
  const float* gs;        // [n_eta][n_theta2][n_u]
  float p_log_min, p_inv_dlog;
  float e_log_min, e_inv_dlog;
  float r_log_min, r_inv_dlog;
  int n_points, n_mat, n_coherent;
  int gs_ix0, gs_iy0, gs_n_eta, gs_n_theta2, gs_n_u;
};

// Linear-in-value interpolation over a log grid, clamped flat outside
// (pyradmc/data/tables.py:83-133). Returns the lower node and the weight.
__device__ inline void grid_locate(float log_value, float log_min, float inv_dlog, int n,
                                   int& i, float& w) {
  float t = (log_value - log_min) * inv_dlog;
  t = fminf(fmaxf(t, 0.0f), static_cast<float>(n - 1));
  i = min(static_cast<int>(t), n - 2);
  w = t - static_cast<float>(i);
}

__device__ inline float lerp1(const float* v, int i, float w) {
  return v[i] * (1.0f - w) + v[i + 1] * w;
}

__device__ inline float4 lerp4(const float4* v, int i, float w) {
  const float4 a = v[i];
  const float4 b = v[i + 1];
  const float wa = 1.0f - w;
  return make_float4(a.x * wa + b.x * w, a.y * wa + b.y * w, a.z * wa + b.z * w,
                     a.w * wa + b.w * w);
}

// ------------------------------------------------------------------------ geometry

struct DevGrid {
  float x_lo, y_lo, z_lo, x_hi, y_hi, z_hi;
  float sx, sy, sz;
  float inv_sx, inv_sy, inv_sz;
  int nx, ny, nz;
  const float* density;     // [nx][ny][nz]
  const uint8_t* material;  // [nx][ny][nz]
};

struct DevScoring {
  float x_lo, y_lo, z_lo, x_hi, y_hi, z_hi;
  float inv_sx, inv_sy, inv_sz;
  int nx, ny, nz;
  long long n_voxels;
};

// Half-open box containment (pyradmc/geometry/grid.py:35-55).
__device__ inline bool inside(float x, float y, float z, float xl, float yl, float zl, float xh,
                              float yh, float zh) {
  if (x < xl || x >= xh) return false;
  if (y < yl || y >= yh) return false;
  return !(z < zl || z >= zh);
}

__device__ inline bool inside_grid(const DevGrid& g, float x, float y, float z) {
  return inside(x, y, z, g.x_lo, g.y_lo, g.z_lo, g.x_hi, g.y_hi, g.z_hi);
}

// floor index, clamped into [0, n-1] (pyradmc/geometry/grid.py:58-80).
__device__ inline int axis_index(float p, float lo, float inv_s, int n) {
  int i = static_cast<int>(floorf((p - lo) * inv_s));
  return min(max(i, 0), n - 1);
}

// Slab entry distance from outside; kInfinityCm or more when the ray misses
// (pyradmc/geometry/grid.py:83-148).
__device__ inline float slab_entry(const DevGrid& g, float x, float y, float z, float ux, float uy,
                                   float uz) {
  float t_near = 0.0f;
  float t_far = 3.0e38f;
  const float pos[3] = {x, y, z};
  const float dir[3] = {ux, uy, uz};
  const float lo[3] = {g.x_lo, g.y_lo, g.z_lo};
  const float hi[3] = {g.x_hi, g.y_hi, g.z_hi};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (dir[a] == 0.0f) {
      if (pos[a] < lo[a] || pos[a] >= hi[a]) return 3.0e38f;
    } else {
      float t1 = (lo[a] - pos[a]) / dir[a];
      float t2 = (hi[a] - pos[a]) / dir[a];
      if (t1 > t2) {
        const float s = t1;
        t1 = t2;
        t2 = s;
      }
      t_near = fmaxf(t_near, t1);
      t_far = fminf(t_far, t2);
    }
  }
  if (t_near > t_far || t_far <= 0.0f) return 3.0e38f;
  return t_near;
}

// Distance to the next voxel face along the direction, clamped >= 0
// (pyradmc/geometry/grid.py:151-223; the clamp is what removes the float32
// boundary-trap fixed point).
__device__ inline float boundary_distance(const DevGrid& g, float x, float y, float z, float ux,
                                          float uy, float uz) {
  float r = 3.0e38f;
  const int ix = axis_index(x, g.x_lo, g.inv_sx, g.nx);
  if (ux > 0.0f) r = fminf(r, (g.x_lo + static_cast<float>(ix + 1) * g.sx - x) / ux);
  else if (ux < 0.0f) r = fminf(r, (g.x_lo + static_cast<float>(ix) * g.sx - x) / ux);
  const int iy = axis_index(y, g.y_lo, g.inv_sy, g.ny);
  if (uy > 0.0f) r = fminf(r, (g.y_lo + static_cast<float>(iy + 1) * g.sy - y) / uy);
  else if (uy < 0.0f) r = fminf(r, (g.y_lo + static_cast<float>(iy) * g.sy - y) / uy);
  const int iz = axis_index(z, g.z_lo, g.inv_sz, g.nz);
  if (uz > 0.0f) r = fminf(r, (g.z_lo + static_cast<float>(iz + 1) * g.sz - z) / uz);
  else if (uz < 0.0f) r = fminf(r, (g.z_lo + static_cast<float>(iz) * g.sz - z) / uz);
  return fmaxf(r, 0.0f);
}

// Flat scoring index (ix*ny + iy)*nz + iz of a position inside the scoring grid.
__device__ inline long long grid_index(const DevScoring& s, float x, float y, float z) {
  const int ix = axis_index(x, s.x_lo, s.inv_sx, s.nx);
  const int iy = axis_index(y, s.y_lo, s.inv_sy, s.ny);
  const int iz = axis_index(z, s.z_lo, s.inv_sz, s.nz);
  return (static_cast<long long>(ix) * s.ny + iy) * s.nz + iz;
}

// Flat scoring index, or -1 outside the scoring grid
// (pyradmc/backends/warp/kernels.py:492-532, rectilinear branch).
__device__ inline long long scoring_index(const DevScoring& s, float x, float y, float z) {
  if (!inside(x, y, z, s.x_lo, s.y_lo, s.z_lo, s.x_hi, s.y_hi, s.z_hi)) return -1;
  return grid_index(s, x, y, z);
}

// ------------------------------------------------------------------------- physics

// Direction-cosine rotation with the pole guard (pyradmc/physics/direction.py:37-81).
__device__ inline void rotate(float& ux, float& uy, float& uz, float cos_t, float phi) {
  const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
  float sin_p, cos_p;
  __sincosf(phi, &sin_p, &cos_p);
  const float t2 = ux * ux + uy * uy;
  if (t2 > kPoleFloor) {
    const float t = sqrtf(t2);
    const float inv = sin_t / t;
    const float vx = ux * cos_t + inv * (ux * uz * cos_p - uy * sin_p);
    const float vy = uy * cos_t + inv * (uy * uz * cos_p + ux * sin_p);
    const float vz = uz * cos_t - t * sin_t * cos_p;
    ux = vx;
    uy = vy;
    uz = vz;
  } else {
    const float sign = uz > 0.0f ? 1.0f : -1.0f;
    ux = sin_t * cos_p;
    uy = sign * sin_t * sin_p;
    uz = sign * cos_t;
  }
}

// Photon channel from the four attenuation coefficients with one uniform, in the order
// Compton, photoelectric, pair, Rayleigh (pyradmc/physics/channel.py:17-60).
__device__ inline int select_process(float mu_c, float mu_p, float mu_pp, float mu_r, uint64_t& rng) {
  const float threshold = uniform(rng) * (mu_c + mu_p + mu_pp + mu_r);
  if (threshold < mu_c) return COMPTON;
  if (threshold < mu_c + mu_p) return PHOTOELECTRIC;
  if (threshold < mu_c + mu_p + mu_pp) return PAIR;
  return RAYLEIGH;
}

// Isotropic unit direction from two uniforms (pyradmc/physics/direction.py:21-31).
__device__ inline void sample_isotropic(uint64_t& rng, float& ax, float& ay, float& az) {
  const float c = 1.0f - 2.0f * uniform(rng);
  const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
  const float phi = kTwoPi * uniform(rng);
  float sp, cp;
  __sincosf(phi, &sp, &cp);
  ax = s * cp;
  ay = s * sp;
  az = c;
}

// Bernoulli bremsstrahlung over a sub-step whose radiative loss is radiative_loss:
// the emission probability that matches the radiative stopping power with photons
// of log-uniform energy in (pcut, e), and the part below pcut deposited locally
// (pyradmc/physics/brems.py:30-66). log_e and log_pcut are logf(e) and logf(pcut).
__device__ inline void brems_step(float e, float radiative_loss, float pcut, float log_e, float log_pcut,
                                  float& emit_probability, float& local) {
  emit_probability = 0.0f;
  local = radiative_loss;
  if (e > pcut) {
    const float emitted_fraction = 1.0f - pcut / e;
    const float mean_k = (e - pcut) / (log_e - log_pcut);
    emit_probability = radiative_loss * emitted_fraction / mean_k;
    local = radiative_loss * (1.0f - emitted_fraction);
  }
}

// Log-uniform bremsstrahlung photon energy k = pcut (e / pcut)^u
// (pyradmc/physics/brems.py:69-84).
__device__ inline float sample_brems_energy(float pcut, float log_e, float log_pcut, uint64_t& rng) {
  return pcut * __expf(uniform(rng) * (log_e - log_pcut));
}

// Exponential free path -log(1-u)/mu (pyradmc/physics/path.py:15-35).
__device__ inline float sample_path(float mu, uint64_t& rng) {
  return -__logf(1.0f - uniform(rng)) / mu;
}

// Kahn rejection for the Klein-Nishina energy ratio (pyradmc/physics/compton.py:19-72).
__device__ inline float sample_compton_ratio(float e, uint64_t& rng) {
  const float alpha = e / kElectronMass;
  const float two_alpha = 2.0f * alpha;
  const float branch = (1.0f + two_alpha) / (9.0f + two_alpha);
  for (int guard = 0; guard < 100000; ++guard) {
    const float u1 = uniform(rng);
    const float u2 = uniform(rng);
    const float u3 = uniform(rng);
    if (u1 <= branch) {
      const float x = 1.0f + two_alpha * u2;
      if (u3 <= 4.0f * (x - 1.0f) / (x * x)) return 1.0f / x;
    } else {
      const float x = (1.0f + two_alpha) / (1.0f + two_alpha * u2);
      const float c = 1.0f - (x - 1.0f) / alpha;
      if (u3 <= 0.5f * (c * c + 1.0f / x)) return 1.0f / x;
    }
  }
  return 1.0f;
}

__device__ inline float compton_cos(float e, float ratio) {
  const float alpha = e / kElectronMass;
  const float c = 1.0f + 1.0f / alpha - 1.0f / (alpha * ratio);
  return fminf(1.0f, fmaxf(-1.0f, c));
}

__device__ inline float compton_electron_cos(float e, float ratio) {
  const float recoil = e * (1.0f - ratio);
  const float p = sqrtf(recoil * (recoil + 2.0f * kElectronMass));
  const float cg = compton_cos(e, ratio);
  const float c = (e - ratio * e * cg) / p;
  return fminf(1.0f, fmaxf(-1.0f, c));
}

// Restricted Moller delta energy by rejection from a 1/x^2 envelope
// (pyradmc/physics/moller.py:18-57).
__device__ inline float sample_moller(float e, float cut, uint64_t& rng) {
  const float tau = e / kElectronMass;
  const float tr = tau / (tau + 1.0f);
  const float tr2 = tr * tr;
  const float two_tau = (2.0f * tau + 1.0f) / ((tau + 1.0f) * (tau + 1.0f));
  const float w_min = cut / e;
  const float bound = 2.0f + 0.25f * tr2;
  for (int guard = 0; guard < 100000; ++guard) {
    const float u = uniform(rng);
    const float x = w_min / (1.0f - u * (1.0f - 2.0f * w_min));
    const float ratio = x / (1.0f - x);
    const float accept = 1.0f + ratio * ratio + x * x * tr2 - two_tau * ratio;
    if (uniform(rng) * bound <= accept) return x * e;
  }
  return w_min * e;
}

__device__ inline void moller_cosines(float e, float w, float& cos_delta, float& cos_primary) {
  const float two_m = 2.0f * kElectronMass;
  cos_delta = fminf(1.0f, sqrtf(w * (e + two_m) / (e * (w + two_m))));
  const float surv = e - w;
  cos_primary = fminf(1.0f, sqrtf(surv * (e + two_m) / (e * (surv + two_m))));
}

__device__ inline float expm1_small(float x) {
  // exp(x)-1 with a second-order branch where the subtraction would cancel
  // (pyradmc/backends/warp/physics.py:89-101).
  if (fabsf(x) < 1.0e-4f) return x * (1.0f + 0.5f * x);
  return __expf(x) - 1.0f;
}

// Gaussian (Fermi-Eyges) hinge cosine (pyradmc/physics/msc.py:37-54).
__device__ inline float sample_gaussian_hinge(float ms, uint64_t& rng) {
  const float theta = sqrtf(-ms * __logf(1.0f - uniform(rng)));
  return __cosf(fminf(theta, kPi));
}

// Goudsmit-Saunderson bilinear grid sample (pyradmc/physics/gs.py:39-166).
__device__ inline float sample_gs(const DevTables& t, float fx, float fy, float ms,
                                  uint64_t& rng) {
  int ix = static_cast<int>(floorf(fx)) - t.gs_ix0;
  int iy = static_cast<int>(floorf(fy)) - t.gs_iy0;
  ix = min(max(ix, 0), t.gs_n_eta - 2);
  iy = min(max(iy, 0), t.gs_n_theta2 - 2);
  const float tx = fminf(fmaxf(fx - static_cast<float>(t.gs_ix0 + ix), 0.0f), 1.0f);
  const float ty = fminf(fmaxf(fy - static_cast<float>(t.gs_iy0 + iy), 0.0f), 1.0f);
  if (ms <= 0.0f) return 1.0f;
  const int n_u = t.gs_n_u;
  int index = static_cast<int>(uniform(rng) * static_cast<float>(n_u));
  index = min(index, n_u - 1);
  const long long row = static_cast<long long>(n_u);
  const long long col = static_cast<long long>(t.gs_n_theta2) * row;
  const float* base = t.gs + ix * col + iy * row + index;
  const float w_lo = base[0] * (1.0f - tx) + base[col] * tx;
  const float w_hi = base[row] * (1.0f - tx) + base[col + row] * tx;
  const float w = w_lo * (1.0f - ty) + w_hi * ty;
  const float omc = -expm1_small(-0.5f * ms);
  return fmaxf(-1.0f, 1.0f - w * omc);
}

// Linear interpolation of ys(xs) for ascending xs, flat outside; bisection
// (pyradmc/physics/rayleigh.py:86-110).
__device__ inline float interp_increasing(float q, const float* xs, const float* ys, int n) {
  if (q <= xs[0]) return ys[0];
  if (q >= xs[n - 1]) return ys[n - 1];
  int lo = 0;
  int hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xs[mid] <= q) lo = mid;
    else hi = mid;
  }
  const float t = (q - xs[lo]) / (xs[hi] - xs[lo]);
  return ys[lo] * (1.0f - t) + ys[hi] * t;
}

// Coherent cosine from the form-factor cumulative with Thomson acceptance
// (pyradmc/physics/rayleigh.py:39-83). Unreachable with analytic (zero) coherent data.
__device__ inline float sample_coherent(const DevTables& t, int mat, float e, uint64_t& rng) {
  const float x_max = 80.65543f * e;
  const float* cum = t.coherent_cumulative + static_cast<long long>(mat) * t.n_coherent;
  const float a_max = interp_increasing(x_max, t.coherent_x, cum, t.n_coherent);
  for (int guard = 0; guard < 100000; ++guard) {
    const float target = uniform(rng) * a_max;
    const float x = interp_increasing(target, cum, t.coherent_x, t.n_coherent);
    const float ratio = x / x_max;
    const float mu = 1.0f - 2.0f * ratio * ratio;
    if (2.0f * uniform(rng) <= 1.0f + mu * mu) return mu;
  }
  return 1.0f;
}

// Dose-to-water weight: restricted stopping ratio water/medium at the
// cutoff-clamped energy; exactly 1 for water and under dose-to-medium
// (pyradmc/backends/warp/kernels.py:459-489).
__device__ inline float spr_factor(const DevTables& t, int dose_to_water, float e, int mat,
                                   float ecut) {
  if (dose_to_water == 0 || mat == 0) return 1.0f;
  int i;
  float w;
  grid_locate(__logf(fmaxf(e, ecut)), t.e_log_min, t.e_inv_dlog, t.n_points, i, w);
  const float s_water = lerp4(t.elec_b, i, w).z;
  const float s_medium = lerp4(t.elec_b + static_cast<long long>(mat) * t.n_points, i, w).z;
  return s_water / s_medium;
}

// An integer-valued double of magnitude below 2^51 as int64, exactly: adding 1.5 * 2^52
// places the integer in the low mantissa bits, which the bit pattern then holds (a few
// instructions where a general float64 -> int64 conversion takes a dozen).
__device__ inline long long whole_to_ll(double whole) {
  return __double_as_longlong(whole + 6755399441055744.0) - 0x4338000000000000LL;
}

// Energy to fixed-point quanta, floor(E / quantum + 1/2): round to nearest for either
// sign (pyradmc/backends/warp/kernels.py:584-597); |E| stays far below 2^51 quanta.
__device__ inline long long quanta_round(double energy) {
  return whole_to_ll(floor(energy * kInvQuantum + 0.5));
}

}  // namespace hmc
