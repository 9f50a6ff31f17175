// Per-primary weight factors of collimated and mask-shaped sources, evaluated by the
// generator in float64: narrow-beam transmission through a stack of jaws and a
// rounded-tip MLC (pyradmc/geometry/collimation.py:280-489, 927-984), or a 0..1 mask
// read bilinearly where the ray crosses a plane rectangle (collimation.py:987-1068).
//
// Geometry in the beam frame (origin at the focal spot, basis rows u, v, w): every
// device is a slab z_top <= w <= z_bottom; a jaw pair holds q >= edge_pos and
// q <= edge_neg (q = u or v; a focused edge is the plane z_mid q = edge w); an MLC tiles
// v with leaf strips whose ends are the flat body u >= tip + R (mirrored) united with
// the disc of radius R about (tip + R, z_mid) in the (u, w) plane, so a leaf's chord
// is |S & H| + |S & D| - |S & H & D|. Chords are taken along the whole line. An MLC
// visits only the strips the ray's chord of the slab can reach (found by bisection,
// one strip of margin each side); every other strip's chord is zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hmc {

struct DevRayWeights {
  int kind;  // 0 none, 1 stack transmission, 2 mask
  int n_dev, n_table, nu, nv;
  double o[3], b[9];
  const int32_t* dkind;   // [n_dev] 0 jaw pair, 1 MLC
  const int32_t* flags;   // jaw: axis (0 u, 1 v) | focused << 1; MLC: focused sides << 1
  const int32_t* n_pairs;
  const int32_t* off;     // first strip edge of an MLC in edges
  const int32_t* poff;    // first pair of an MLC in tips
  const double* z_top;
  const double* z_bot;
  const double* e_neg;    // jaw edge; MLC tip radius
  const double* e_pos;
  const double* dens;
  const double* edges;
  const double* tn;
  const double* tp;
  const double* log_e;    // [n_table]
  const double* log_mu;   // [n_dev][n_table] log of mu/rho (cm^2/g)
  double le0, inv_dle;    // first log energy and inverse mean spacing (index guess)
  double c[3], u[3], v[3], nrm[3], wu, wv;
  const double* mask;     // [nu][nv]
};

__host__ __device__ inline double span(double lo, double hi) { return hi > lo ? hi - lo : 0.0; }

// |[s_lo, s_hi] & {a + t b >= 0}| (<= 0 unless above).
__host__ __device__ inline double clip_affine(double s_lo, double s_hi, double a, double b, bool above) {
  if (b == 0.0) return (above ? a >= 0.0 : a <= 0.0) ? span(s_lo, s_hi) : 0.0;
  const double root = -a / b;
  if ((b > 0.0) == above) return span(fmax(s_lo, root), s_hi);
  return span(s_lo, fmin(s_hi, root));
}

// Slab interval of the ray (the forward half-line only if from_origin); false when it
// misses the slab.
__host__ __device__ inline bool slab_interval(double p_w, double d_w, double zt, double zb, bool from_origin,
                                              double& lo, double& hi) {
  if (d_w == 0.0) {
    if (p_w < zt || p_w > zb) return false;
    lo = -INFINITY;
    hi = INFINITY;
  } else {
    const double t1 = (zt - p_w) / d_w, t2 = (zb - p_w) / d_w;
    lo = fmin(t1, t2);
    hi = fmax(t1, t2);
  }
  if (from_origin) lo = fmax(lo, 0.0);
  return hi > lo;
}

__host__ __device__ inline double jaw_chord(const double* p, const double* d, double zt, double zb,
                                            double en, double ep, int flags, bool from_origin = false) {
  double s_lo, s_hi;
  if (!slab_interval(p[2], d[2], zt, zb, from_origin, s_lo, s_hi)) return 0.0;
  const int k = flags & 1;
  double ap, bp, an, bn;
  if (flags & 2) {
    const double zm = 0.5 * (zt + zb);
    ap = zm * p[k] - ep * p[2];
    bp = zm * d[k] - ep * d[2];
    an = zm * p[k] - en * p[2];
    bn = zm * d[k] - en * d[2];
  } else {
    ap = p[k] - ep;
    bp = d[k];
    an = p[k] - en;
    bn = d[k];
  }
  return clip_affine(s_lo, s_hi, ap, bp, true) + clip_affine(s_lo, s_hi, an, bn, false);
}

// Chord of one leaf end (body half-space united with the tip disc) inside [i_lo, i_hi].
__host__ __device__ inline double leaf_end(double i_lo, double i_hi, const double* p, const double* d,
                                           double cu, double zm, double r, bool above) {
  double h_lo, h_hi;
  if (d[0] == 0.0) {
    const bool in = above ? p[0] >= cu : p[0] <= cu;
    h_lo = in ? -INFINITY : INFINITY;
    h_hi = in ? INFINITY : -INFINITY;
  } else {
    const double root = (cu - p[0]) / d[0];
    if ((d[0] > 0.0) == above) {
      h_lo = root;
      h_hi = INFINITY;
    } else {
      h_lo = -INFINITY;
      h_hi = root;
    }
  }
  const double ou = p[0] - cu, ow = p[2] - zm;
  const double a = d[0] * d[0] + d[2] * d[2];
  const double bq = 2.0 * (d[0] * ou + d[2] * ow);
  const double cq = ou * ou + ow * ow - r * r;
  double c_lo = INFINITY, c_hi = -INFINITY;
  if (a > 0.0) {
    const double disc = bq * bq - 4.0 * a * cq;
    if (disc > 0.0) {
      const double q = sqrt(disc);
      c_lo = (-bq - q) / (2.0 * a);
      c_hi = (-bq + q) / (2.0 * a);
    }
  } else if (cq <= 0.0) {
    c_lo = -INFINITY;
    c_hi = INFINITY;
  }
  const double body = span(fmax(i_lo, h_lo), fmin(i_hi, h_hi));
  const double tip = span(fmax(i_lo, c_lo), fmin(i_hi, c_hi));
  const double both = span(fmax(i_lo, fmax(h_lo, c_lo)), fmin(i_hi, fmin(h_hi, c_hi)));
  return body + tip - both;
}

// First index in [lo, hi) with x < e[index] (upper bound), by bisection.
__host__ __device__ inline int upper_bound(const double* e, int lo, int hi, double x) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__host__ __device__ inline double mlc_chord(const DevRayWeights& W, int j, const double* p, const double* d,
                                            bool from_origin = false) {
  double s_lo, s_hi;
  const double zt = W.z_top[j], zb = W.z_bot[j];
  if (!slab_interval(p[2], d[2], zt, zb, from_origin, s_lo, s_hi)) return 0.0;
  const double zm = 0.5 * (zt + zb), r = W.e_neg[j];
  const bool focused = (W.flags[j] & 2) != 0;
  const int n = W.n_pairs[j];
  const double* e = W.edges + W.off[j];
  const double* tn = W.tn + W.poff[j];
  const double* tp = W.tp + W.poff[j];
  // Strips the ray can reach inside the slab: the strip coordinate (v, or the
  // focal-spot projection zm v / w, monotonic where w > 0) at the two slab crossings.
  int k0 = 0, k1 = n - 1;
  if (isfinite(s_lo) && isfinite(s_hi)) {
    double v0 = p[1] + s_lo * d[1], v1 = p[1] + s_hi * d[1];
    if (focused) {
      v0 = zm * v0 / (p[2] + s_lo * d[2]);
      v1 = zm * v1 / (p[2] + s_hi * d[2]);
    }
    const double lo = fmin(v0, v1), hi = fmax(v0, v1);
    k0 = max(upper_bound(e, 0, n + 1, lo) - 2, 0);
    k1 = min(upper_bound(e, 0, n + 1, hi), n - 1);
  }
  double total = 0.0;
  for (int i = k0; i <= k1; ++i) {
    const double e0 = e[i], e1 = e[i + 1];
    double a0, b0, a1, b1;
    if (focused) {
      a0 = zm * p[1] - e0 * p[2];
      b0 = zm * d[1] - e0 * d[2];
      a1 = zm * p[1] - e1 * p[2];
      b1 = zm * d[1] - e1 * d[2];
    } else {
      a0 = p[1] - e0;
      b0 = d[1];
      a1 = p[1] - e1;
      b1 = d[1];
    }
    double lo = s_lo, hi = s_hi;
    if (b0 == 0.0) {
      if (a0 < 0.0) continue;
    } else if (b0 > 0.0) {
      lo = fmax(lo, -a0 / b0);
    } else {
      hi = fmin(hi, -a0 / b0);
    }
    if (b1 == 0.0) {
      if (a1 > 0.0) continue;
    } else if (b1 < 0.0) {
      lo = fmax(lo, -a1 / b1);
    } else {
      hi = fmin(hi, -a1 / b1);
    }
    if (!(hi > lo)) continue;
    total += leaf_end(lo, hi, p, d, tp[i] + r, zm, r, true);
    total += leaf_end(lo, hi, p, d, tn[i] - r, zm, r, false);
  }
  return total;
}

// exp(interp(le)) of a table row y over increasing nodes x, clamped at the ends
// (numpy.interp's rule); the node is guessed from the mean spacing and corrected.
__host__ __device__ inline double log_table(const double* x, const double* y, int n, double le0, double inv_dle,
                                            double le) {
  if (le <= x[0]) return exp(y[0]);
  if (le >= x[n - 1]) return exp(y[n - 1]);
  int k = static_cast<int>((le - le0) * inv_dle);
  k = min(max(k, 0), n - 2);
  while (k > 0 && x[k] > le) --k;
  while (k < n - 2 && x[k + 1] <= le) ++k;
  const double f = (le - x[k]) / (x[k + 1] - x[k]);
  return exp(y[k] + f * (y[k + 1] - y[k]));
}

__host__ __device__ inline double table_mu(const DevRayWeights& W, int j, double le) {
  return log_table(W.log_e, W.log_mu + static_cast<long long>(j) * W.n_table, W.n_table, W.le0, W.inv_dle, le);
}

// Chord of one ray (beam frame) through device j.
__host__ __device__ inline double device_chord(const DevRayWeights& W, int j, const double* p, const double* d,
                                               bool from_origin) {
  return W.dkind[j] == 0 ? jaw_chord(p, d, W.z_top[j], W.z_bot[j], W.e_neg[j], W.e_pos[j], W.flags[j], from_origin)
                         : mlc_chord(W, j, p, d, from_origin);
}

__host__ __device__ inline double stack_factor(const DevRayWeights& W, double energy, const double* pe,
                                               const double* de) {
  double p[3], d[3];
  const double q[3] = {pe[0] - W.o[0], pe[1] - W.o[1], pe[2] - W.o[2]};
  for (int i = 0; i < 3; ++i) {
    p[i] = W.b[3 * i] * q[0] + W.b[3 * i + 1] * q[1] + W.b[3 * i + 2] * q[2];
    d[i] = W.b[3 * i] * de[0] + W.b[3 * i + 1] * de[1] + W.b[3 * i + 2] * de[2];
  }
  const double le = log(energy);
  double tau = 0.0;
  for (int j = 0; j < W.n_dev; ++j) {
    const double t = device_chord(W, j, p, d, false);
    if (t > 0.0) tau += W.dens[j] * table_mu(W, j, le) * t;
  }
  return exp(-tau);
}

__host__ __device__ inline double mask_axis(double c, int n, double w, int& i0) {
  double x = (c + 0.5 * w) / (w / n) - 0.5;
  x = fmin(fmax(x, 0.0), static_cast<double>(n - 1));
  i0 = min(static_cast<int>(x), max(n - 2, 0));
  return x - i0;
}

__host__ __device__ inline double mask_factor(const DevRayWeights& W, const double* p, const double* d) {
  const double den = d[0] * W.nrm[0] + d[1] * W.nrm[1] + d[2] * W.nrm[2];
  if (den == 0.0) return 0.0;
  const double t = ((W.c[0] - p[0]) * W.nrm[0] + (W.c[1] - p[1]) * W.nrm[1] + (W.c[2] - p[2]) * W.nrm[2]) / den;
  double off[3];
  for (int k = 0; k < 3; ++k) off[k] = p[k] + t * d[k] - W.c[k];
  const double cu = off[0] * W.u[0] + off[1] * W.u[1] + off[2] * W.u[2];
  const double cv = off[0] * W.v[0] + off[1] * W.v[1] + off[2] * W.v[2];
  if (fabs(cu) > 0.5 * W.wu || fabs(cv) > 0.5 * W.wv) return 0.0;
  int i0, j0;
  const double fu = mask_axis(cu, W.nu, W.wu, i0);
  const double fv = mask_axis(cv, W.nv, W.wv, j0);
  const int i1 = min(i0 + 1, W.nu - 1), j1 = min(j0 + 1, W.nv - 1);
  const double a = W.mask[i0 * W.nv + j0], b = W.mask[i1 * W.nv + j0];
  const double c = W.mask[i0 * W.nv + j1], e = W.mask[i1 * W.nv + j1];
  return a + fu * (b - a) + fv * (c - a) + fu * fv * (e - b - c + a);
}

// The factor of one primary (float32 state, as the generator holds it).
__host__ __device__ inline double ray_weight(const DevRayWeights& W, float e, float x, float y, float z, float ux,
                                             float uy, float uz) {
  const double p[3] = {x, y, z}, d[3] = {ux, uy, uz};
  if (W.kind == 1) return stack_factor(W, static_cast<double>(e), p, d);
  if (W.kind == 2) return mask_factor(W, p, d);
  return 1.0;
}

}  // namespace hmc
