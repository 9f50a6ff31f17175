// Goudsmit-Saunderson scaled-deflection node grid, built natively on the host.
//
// Each node (eta = exp(i/bins), <theta^2> = exp(j/bins)) is a 512-bin table of the
// conditional mean of 1 - mu over equal-probability bins of the multiple-scattering
// angular distribution after Lambda = <theta^2>/(2 G_1) elastic mean free paths,
// normalised to unit mean (pyradmc/data/goudsmit_saunderson.py:652-759).
//   Lambda >= 300: Legendre-series CDF with exp(-Lambda G_l) damping, the unscattered
//                  delta removed before summation, inverted at 2^19 midpoints
//                  (goudsmit_saunderson.py:456-516, 743-759).
//   Lambda <  300: explicit composition of a Poisson number of screened-Rutherford
//                  scatters, then log2(Lambda) self-convolutions by random matching
//                  (goudsmit_saunderson.py:567-604, 727-741). The matching here is a
//                  Fisher-Yates permutation from this library's own counter-based
//                  stream, so MC nodes agree with the reference's in distribution,
//                  not bit for bit (its nodes depend on NumPy's Generator).
// Nodes are independent; threads pull them from an atomic counter, so the result
// does not depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/hipmc.h"

namespace {

constexpr int kRawSamples = 1 << 19;
constexpr int kMuBins = 4096;
constexpr int kQuadrature = 1 << 16;
constexpr int kMomentLMax = 1024;
constexpr double kSeriesMinLambda = 300.0;
constexpr double kSeriesCutoff = 1.0e-12;
constexpr uint64_t kBuildSeed = 20260720ull;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t s;
  explicit Stream(uint64_t key) : s(mix64(mix64(kBuildSeed + 0x9E3779B97F4A7C15ULL) ^ key)) {}
  uint64_t next() {  // SplitMix64
    s += 0x9E3779B97F4A7C15ULL;
    return mix64(s);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

double first_transport_moment(double eta) {
  return 2.0 * eta * (1.0 + eta) * std::log((1.0 + eta) / eta) - 2.0 * eta;
}

// G_l, l = 0..l_max, by midpoint quadrature in t = ln((xi + eta)/eta)
// (goudsmit_saunderson.py:402-453).
std::vector<double> moments(double eta) {
  std::vector<double> mu(kQuadrature), weight(kQuadrature);
  const double t_max = std::log((1.0 + eta) / eta);
  const double dt = t_max / kQuadrature;
  double total = 0.0;
  for (int k = 0; k < kQuadrature; ++k) {
    const double t = (k + 0.5) * dt;
    const double et = std::exp(t);
    const double xi = eta * (et - 1.0);
    weight[k] = eta * et * dt;
    mu[k] = 1.0 - 2.0 * eta * (1.0 - xi) / (xi + eta);
    total += weight[k];
  }
  std::vector<double> g(kMomentLMax + 1, 0.0);
  std::vector<double> p_prev(kQuadrature, 1.0), p_curr(mu);
  for (int l = 1; l <= kMomentLMax; ++l) {
    double dot = 0.0;
    for (int k = 0; k < kQuadrature; ++k) dot += p_curr[k] * weight[k];
    g[l] = total - dot;
    const double a = (2.0 * l + 1.0) / (l + 1.0), b = static_cast<double>(l) / (l + 1.0);
    for (int k = 0; k < kQuadrature; ++k) {
      const double next = a * mu[k] * p_curr[k] - b * p_prev[k];
      p_prev[k] = p_curr[k];
      p_curr[k] = next;
    }
  }
  return g;
}

// Series CDF on the mu grid (goudsmit_saunderson.py:456-516).
void series_cdf(double lam, const std::vector<double>& g, const std::vector<double>& mu,
                std::vector<double>& cdf) {
  const double unscattered = std::exp(-lam);
  std::vector<double> coef(g.size());
  int l_max = 0;
  for (size_t l = 0; l < g.size(); ++l) {
    coef[l] = std::exp(-lam * g[l]) - unscattered;
    if (std::fabs(coef[l]) > kSeriesCutoff) l_max = static_cast<int>(l);
  }
  const int n = static_cast<int>(mu.size());
  for (int k = 0; k < n; ++k) {
    const double x = mu[k];
    double c = (1.0 - unscattered) * 0.5 * (1.0 + x);
    if (l_max >= 1) {
      double p_prev = 1.0, p_curr = x, p_next = 1.5 * x * x - 0.5;
      for (int l = 1; l <= l_max; ++l) {
        c += 0.5 * coef[l] * (p_next - p_prev);
        p_prev = p_curr;
        p_curr = p_next;
        const int m = l + 1;
        p_next = ((2.0 * m + 1.0) * x * p_curr - m * p_prev) / (m + 1.0);
      }
    }
    cdf[k] = c;
  }
  double running = 0.0;
  for (int k = 0; k < n; ++k) {
    double c = std::min(std::max(cdf[k], 0.0), 1.0);
    running = std::max(running, c);
    cdf[k] = running;
  }
  cdf[0] = 0.0;
  cdf[n - 1] = 1.0;
}

void series_raw(double lam, const std::vector<double>& g, std::vector<double>& raw) {
  std::vector<double> mu(kMuBins), cdf(kMuBins);
  const double a = std::log(1.0e-12), b = std::log(2.0);
  for (int k = 0; k < kMuBins; ++k) {
    // one_minus_mu = geomspace(1e-12, 2, n), mirrored into an ascending mu grid.
    const double omm = k == 0 ? 1.0e-12 : (k == kMuBins - 1 ? 2.0 : std::exp(a + (b - a) * k / (kMuBins - 1)));
    mu[kMuBins - 1 - k] = std::min(std::max(1.0 - omm, -1.0), 1.0);
  }
  mu[0] = -1.0;
  mu[kMuBins - 1] = 1.0;
  series_cdf(lam, g, mu, cdf);
  int keep_lo = 0, keep_hi = kMuBins - 1;
  for (int k = 0; k < kMuBins; ++k)
    if (cdf[k] <= 0.0) keep_lo = k;
  for (int k = kMuBins - 1; k >= 0; --k)
    if (cdf[k] >= 1.0) keep_hi = k;
  if (keep_hi <= keep_lo) {
    keep_lo = 0;
    keep_hi = kMuBins - 1;
  }
  std::vector<double> xp, fp;
  for (int k = keep_lo; k <= keep_hi; ++k) {
    if (k == keep_lo || cdf[k] > cdf[k - 1]) {
      xp.push_back(cdf[k]);
      fp.push_back(mu[k]);
    }
  }
  // np.interp at ascending midpoints u.
  size_t j = 0;
  const size_t m = xp.size();
  for (int i = 0; i < kRawSamples; ++i) {
    const double u = (i + 0.5) / kRawSamples;
    if (u <= xp[0]) {
      raw[i] = fp[0];
      continue;
    }
    if (u >= xp[m - 1]) {
      raw[i] = fp[m - 1];
      continue;
    }
    while (j + 1 < m && xp[j + 1] <= u) ++j;
    const double t = (u - xp[j]) / (xp[j + 1] - xp[j]);
    raw[i] = fp[j] + t * (fp[j + 1] - fp[j]);
  }
}

inline double compose(double ca, double cb, double phi) {
  const double sa = std::sqrt(std::max(0.0, 1.0 - ca * ca));
  const double sb = std::sqrt(std::max(0.0, 1.0 - cb * cb));
  return std::min(std::max(ca * cb + sa * sb * std::cos(phi), -1.0), 1.0);
}

// Poisson composition plus self-convolution doublings (goudsmit_saunderson.py:567-604).
void mc_raw(double eta, double lam, uint64_t key, std::vector<double>& raw,
            std::vector<double>& partner) {
  Stream rng(key);
  int doublings = lam > 1.0 ? static_cast<int>(std::ceil(std::log2(lam))) : 0;
  if (doublings < 0) doublings = 0;
  const double lam0 = lam / std::ldexp(1.0, doublings);
  const double p0 = std::exp(-lam0);
  const double two_pi = 6.283185307179586;
  for (int i = 0; i < kRawSamples; ++i) {
    // Poisson(lam0 <= 1) by sequential inversion.
    const double u = rng.uniform();
    int count = 0;
    double p = p0, cum = p0;
    while (u > cum && count < 64) {
      ++count;
      p *= lam0 / count;
      cum += p;
    }
    double c = 1.0;
    for (int s = 0; s < count; ++s) {
      const double xi = rng.uniform();
      const double inverse = xi / (2.0 * eta * (1.0 + eta)) + 1.0 / (2.0 + 2.0 * eta);
      const double single = 1.0 + 2.0 * eta - 1.0 / inverse;
      c = compose(c, single, two_pi * rng.uniform());
    }
    raw[i] = c;
  }
  for (int d = 0; d < doublings; ++d) {
    partner = raw;
    for (int i = kRawSamples - 1; i > 0; --i) {
      const int j = static_cast<int>(rng.next() % static_cast<uint64_t>(i + 1));
      std::swap(partner[i], partner[j]);
    }
    for (int i = 0; i < kRawSamples; ++i) raw[i] = compose(raw[i], partner[i], two_pi * rng.uniform());
  }
  std::sort(raw.begin(), raw.end());
}

void condense(const std::vector<double>& raw, int n_u, double* out) {
  const int per = kRawSamples / n_u;
  double total = 0.0;
  for (int b = 0; b < n_u; ++b) {
    double s = 0.0;
    for (int k = 0; k < per; ++k) s += 1.0 - raw[static_cast<size_t>(b) * per + k];
    out[b] = s / per;
    total += out[b];
  }
  const double mean = total / n_u;
  for (int b = 0; b < n_u; ++b) out[b] = mean > 0.0 ? out[b] / mean : 0.0;
}

}  // namespace

extern "C" int hmc_build_gs_grid(int ix0, int ix1, int iy0, int iy1, int n_u, double bins,
                                 int n_threads, double* out) {
  if (!out || ix1 < ix0 || iy1 < iy0 || n_u < 1 || bins <= 0.0) return HMC_EINVAL;
  if (kRawSamples % n_u != 0) return HMC_EINVAL;
  const int n_eta = ix1 - ix0 + 1, n_th = iy1 - iy0 + 1;
  int workers = n_threads > 0 ? n_threads : static_cast<int>(std::thread::hardware_concurrency());
  workers = std::max(1, std::min(workers, 64));

  // Which etas need series moments?
  std::vector<char> need(n_eta, 0);
  for (int a = 0; a < n_eta; ++a) {
    const double eta = std::exp((ix0 + a) / bins);
    const double g1 = first_transport_moment(eta);
    for (int b = 0; b < n_th; ++b) {
      const double lam = 0.5 * std::exp((iy0 + b) / bins) / g1;
      if (lam >= kSeriesMinLambda) need[a] = 1;
    }
  }
  std::vector<std::vector<double>> g(n_eta);
  {
    std::atomic<int> next{0};
    auto work = [&]() {
      for (int a; (a = next.fetch_add(1)) < n_eta;)
        if (need[a]) g[a] = moments(std::exp((ix0 + a) / bins));
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; ++w) pool.emplace_back(work);
    for (auto& th : pool) th.join();
  }
  std::atomic<long long> next{0};
  const long long total = static_cast<long long>(n_eta) * n_th;
  auto work = [&]() {
    std::vector<double> raw(kRawSamples), partner(kRawSamples);
    for (long long node; (node = next.fetch_add(1)) < total;) {
      const int a = static_cast<int>(node / n_th), b = static_cast<int>(node % n_th);
      const double eta = std::exp((ix0 + a) / bins);
      const double theta2 = std::exp((iy0 + b) / bins);
      const double lam = 0.5 * theta2 / first_transport_moment(eta);
      if (lam < kSeriesMinLambda) {
        const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(ix0 + a)) << 32) ^
                             static_cast<uint32_t>(iy0 + b);
        mc_raw(eta, lam, key, raw, partner);
      } else {
        series_raw(lam, g[a], raw);
      }
      condense(raw, n_u, out + node * n_u);
    }
  };
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  return HMC_OK;
}
