/* ORACLE (test infrastructure only): CPU restatement of pyRadMC's transport loops.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this
 * library, and only as the checker. It restates, in double precision:
 *   transport_history   pyradmc/transport/history.py:31-150  (LIFO stack per history)
 *   photon_steps        pyradmc/transport/photon.py:100-282  (Woodcock, channels, roulette)
 *   electron_steps      pyradmc/transport/electron.py:207-481 (Class II substep)
 *   annihilate_at_rest  pyradmc/transport/particles.py:75-102
 *   physics samplers    pyradmc/physics/{path,channel,compton,direction,roulette,
 *                       moller,brems,msc,gs}.py
 *   table lookups       pyradmc/data/tables.py:83-133 (the kernels' data access,
 *                       pyradmc/backends/warp/kernels.py:1232-1260, 1646-1776)
 *   batched scoring     pyradmc/scoring/dose.py:127-207; Dij: ref/engine.py:296-463
 *   given primaries     a phase space's sampled records, one per history key, each with
 *                       its own kind and weight (pyradmc/geometry/phasespace.py:547-560,
 *                       pyradmc/backends/warp/engine.py:2506-2597)
 * One sequential random stream per history serves its whole particle family, as in
 * ReferenceEngine (pyradmc/backends/ref/engine.py:241-278); the stream itself is
 * this file's own SplitMix64 (the reference's PCG64 streams are not reproduced).
 * Histories of a batch are split over threads; each thread tallies privately and the
 * batch is summed in thread order, so results depend only on (seed, threads).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ME 0.51099895069
#define TWO_PI 6.283185307179586
#define PI 3.141592653589793

typedef struct {
  /* transport grid */
  int nx, ny, nz;
  double ox, oy, oz, sx, sy, sz;
  const double *density;
  const int32_t *material;
  /* scoring grid */
  int snx, sny, snz;
  double sox, soy, soz, ssx, ssy, ssz;
  const double *mass;
  /* tables, per material rows of n */
  int n, n_mat;
  const double *majorant, *mu_c, *mu_p, *mu_pp, *mu_r;
  const double *csda, *moller, *rrange, *srad, *tpow, *log_eta, *inv_range;
  double p_log_min, p_inv, e_log_min, e_inv, r_log_min, r_inv;
  /* GS grid */
  const double *gs;
  int gs_ix0, gs_iy0, gs_neta, gs_nth, gs_nu;
  /* source */
  int src_kind; /* 0 pencil, 1 parallel, 2 parallel spectral, 3 fan, 4 lattice, 5 given, 6 spot fan */
  int particle; /* 0 photon, 1 electron */
  double energy;
  double pos[3], dir[3];
  double z0, x0, xe, y0, ye;
  double focal[3], ua[3], va[3], wu, wv;
  int n_bins;
  const double *edges, *cdf;
  int n_beamlets;
  const double *beamlet; /* lattice [n][4]; fan [n][3] */
  /* physics */
  double pcut, ecut, step_fraction, resolution;
  int transport_electrons, msc_gs;
  /* dose to water: restricted stopping rows and the switch (scoring/dose_to_water.py) */
  const double *srest;
  int dose_to_water;
  /* given primaries (src_kind 5), indexed by history key: kind 0 photon, 1 electron,
   * 2 positron; prim [8][n_prim] = energy, x, y, z, ux, uy, uz, weight */
  long long n_prim;
  const int32_t *prim_kind;
  const double *prim;
  /* Gaussian focal spot of the spot fan (src_kind 6; apertures in beamlet [n][3]) */
  double sigma_u, sigma_v;
} om_config;

typedef struct {
  double emitted, deposited, escaped, unscored;
  long long vox, deposits, pushes, photon_steps, electron_steps;
} om_books;

/* ------------------------------------------------------------------ stream */
typedef struct { uint64_t s; } rng_t;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static void rng_seed(rng_t *r, uint64_t seed, uint64_t key) {
  r->s = mix64(mix64(seed ^ 0x5851F42D4C957F2DULL) + key * 0x9E3779B97F4A7C15ULL);
}
static double uniform(rng_t *r) { /* [0, 1), 53 bits */
  r->s += 0x9E3779B97F4A7C15ULL;
  return (double)(mix64(r->s) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------ lookups */
static void locate(double v, double lo, double inv, int n, int *i, double *w) {
  double t = (v - lo) * inv;
  if (t < 0.0) t = 0.0;
  if (t > (double)(n - 1)) t = (double)(n - 1);
  int k = (int)t;
  if (k > n - 2) k = n - 2;
  *i = k;
  *w = t - (double)k;
}
static double lerp(const double *row, int i, double w) { return row[i] * (1.0 - w) + row[i + 1] * w; }
static double elook(const om_config *c, const double *t, int mat, double e) {
  int i; double w;
  locate(log(e), c->e_log_min, c->e_inv, c->n, &i, &w);
  return lerp(t + (size_t)mat * c->n, i, w);
}

/* ------------------------------------------------------------------ geometry */
static int inside(const om_config *c, double x, double y, double z) {
  if (x < c->ox || x >= c->ox + c->nx * c->sx) return 0;
  if (y < c->oy || y >= c->oy + c->ny * c->sy) return 0;
  return !(z < c->oz || z >= c->oz + c->nz * c->sz);
}
static int axis(double p, double o, double s, int n) {
  int i = (int)floor((p - o) / s);
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
static size_t voxel(const om_config *c, double x, double y, double z) {
  int ix = axis(x, c->ox, c->sx, c->nx), iy = axis(y, c->oy, c->sy, c->ny), iz = axis(z, c->oz, c->sz, c->nz);
  return ((size_t)ix * c->ny + iy) * c->nz + iz;
}
static double entry_distance(const om_config *c, const double p[3], const double u[3]) {
  const double lo[3] = {c->ox, c->oy, c->oz};
  const double hi[3] = {c->ox + c->nx * c->sx, c->oy + c->ny * c->sy, c->oz + c->nz * c->sz};
  double tn = 0.0, tf = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (u[a] == 0.0) {
      if (p[a] < lo[a] || p[a] >= hi[a]) return INFINITY;
    } else {
      double t1 = (lo[a] - p[a]) / u[a], t2 = (hi[a] - p[a]) / u[a];
      if (t1 > t2) { double s = t1; t1 = t2; t2 = s; }
      if (t1 > tn) tn = t1;
      if (t2 < tf) tf = t2;
    }
  }
  if (tn > tf || tf <= 0.0) return INFINITY;
  return tn;
}
static double face_distance(const om_config *c, const double p[3], const double u[3]) {
  const double o[3] = {c->ox, c->oy, c->oz}, s[3] = {c->sx, c->sy, c->sz};
  const int n[3] = {c->nx, c->ny, c->nz};
  double r = INFINITY;
  for (int a = 0; a < 3; ++a) {
    int i = axis(p[a], o[a], s[a], n[a]);
    if (u[a] > 0.0) { double d = (o[a] + (double)(i + 1) * s[a] - p[a]) / u[a]; if (d < r) r = d; }
    else if (u[a] < 0.0) { double d = (o[a] + (double)i * s[a] - p[a]) / u[a]; if (d < r) r = d; }
  }
  return r > 0.0 ? r : 0.0;
}

/* ------------------------------------------------------------------ physics */
static void rotate(double u[3], double ct, double phi) {
  double st = sqrt(fmax(0.0, 1.0 - ct * ct)), cp = cos(phi), sp = sin(phi);
  double t2 = u[0] * u[0] + u[1] * u[1];
  if (t2 > 1.0e-20) {
    double t = sqrt(t2), k = st / t;
    double vx = u[0] * ct + k * (u[0] * u[2] * cp - u[1] * sp);
    double vy = u[1] * ct + k * (u[1] * u[2] * cp + u[0] * sp);
    double vz = u[2] * ct - t * st * cp;
    u[0] = vx; u[1] = vy; u[2] = vz;
  } else {
    double sg = u[2] > 0.0 ? 1.0 : -1.0;
    u[0] = st * cp; u[1] = sg * st * sp; u[2] = sg * ct;
  }
}
static double path(double mu, rng_t *r) { return -log1p(-uniform(r)) / mu; }

static double kahn(double e, rng_t *r) {
  double a = e / ME, ta = 2.0 * a, pb = (1.0 + ta) / (9.0 + ta);
  for (;;) {
    double u1 = uniform(r), u2 = uniform(r), u3 = uniform(r);
    if (u1 <= pb) {
      double x = 1.0 + ta * u2;
      if (u3 <= 4.0 * (x - 1.0) / (x * x)) return 1.0 / x;
    } else {
      double x = (1.0 + ta) / (1.0 + ta * u2), ct = 1.0 - (x - 1.0) / a;
      if (u3 <= 0.5 * (ct * ct + 1.0 / x)) return 1.0 / x;
    }
  }
}
static double clamp1(double v) { return v > 1.0 ? 1.0 : (v < -1.0 ? -1.0 : v); }
static double compton_cos(double e, double r) { double a = e / ME; return clamp1(1.0 + 1.0 / a - 1.0 / (a * r)); }
static double compton_e_cos(double e, double r) {
  double rec = e * (1.0 - r), p = sqrt(rec * (rec + 2.0 * ME));
  return clamp1((e - r * e * compton_cos(e, r)) / p);
}
static double moller_w(double e, double cut, rng_t *r) {
  double tau = e / ME, tr = tau / (tau + 1.0), tr2 = tr * tr;
  double tt = (2.0 * tau + 1.0) / ((tau + 1.0) * (tau + 1.0)), wm = cut / e, bound = 2.0 + 0.25 * tr2;
  for (;;) {
    double u = uniform(r), x = wm / (1.0 - u * (1.0 - 2.0 * wm)), q = x / (1.0 - x);
    double acc = 1.0 + q * q + x * x * tr2 - tt * q;
    if (uniform(r) * bound <= acc) return x * e;
  }
}
static double gs_cos(const om_config *c, double fx, double fy, double ms, rng_t *r) {
  int ix = (int)floor(fx) - c->gs_ix0, iy = (int)floor(fy) - c->gs_iy0;
  if (ix < 0) ix = 0;
  if (ix > c->gs_neta - 2) ix = c->gs_neta - 2;
  if (iy < 0) iy = 0;
  if (iy > c->gs_nth - 2) iy = c->gs_nth - 2;
  double tx = fx - (double)(c->gs_ix0 + ix), ty = fy - (double)(c->gs_iy0 + iy);
  tx = tx < 0.0 ? 0.0 : (tx > 1.0 ? 1.0 : tx);
  ty = ty < 0.0 ? 0.0 : (ty > 1.0 ? 1.0 : ty);
  if (ms <= 0.0) return 1.0;
  int k = (int)(uniform(r) * (double)c->gs_nu);
  if (k > c->gs_nu - 1) k = c->gs_nu - 1;
  size_t row = (size_t)c->gs_nu, col = (size_t)c->gs_nth * row;
  const double *b = c->gs + ix * col + iy * row + k;
  double wlo = b[0] * (1.0 - tx) + b[col] * tx, whi = b[row] * (1.0 - tx) + b[col + row] * tx;
  double w = wlo * (1.0 - ty) + whi * ty;
  double v = 1.0 - w * (-expm1(-0.5 * ms));
  return v < -1.0 ? -1.0 : v;
}

/* ------------------------------------------------------------------ scoring */
typedef struct {
  double *energy; /* [columns][scoring voxels] */
  size_t base;
  om_books b;
} tally_t;

/* Dose-to-water weight of a deposit: the restricted collision stopping-power ratio
 * water / medium at the cutoff-clamped energy, exactly 1 for water or dose to medium
 * (water_spr, pyradmc/scoring/dose_to_water.py:67-85; table form kernels.py:462-488). */
static double spr(const om_config *c, double e, int m) {
  if (!c->dose_to_water || m == 0) return 1.0;
  double ec = e > c->ecut ? e : c->ecut;
  return elook(c, c->srest, 0, ec) / elook(c, c->srest, m, ec);
}
static int material_at(const om_config *c, const double p[3]) {
  return c->material[voxel(c, p[0], p[1], p[2])];
}

/* The tally takes factor * amount; the books take the physical amount. */
static void deposit(const om_config *c, tally_t *t, double x, double y, double z, double amount,
                    double factor) {
  double hx = c->sox + c->snx * c->ssx, hy = c->soy + c->sny * c->ssy, hz = c->soz + c->snz * c->ssz;
  if (x < c->sox || x >= hx || y < c->soy || y >= hy || z < c->soz || z >= hz) {
    t->b.unscored += amount;
    return;
  }
  int ix = axis(x, c->sox, c->ssx, c->snx), iy = axis(y, c->soy, c->ssy, c->sny), iz = axis(z, c->soz, c->ssz, c->snz);
  t->energy[t->base + ((size_t)ix * c->sny + iy) * c->snz + iz] += factor * amount;
  t->b.deposited += amount;
  t->b.deposits++;
}
static double deposit_or_escape(const om_config *c, tally_t *t, const double p[3], double amount,
                                double factor) {
  if (amount <= 0.0) return 0.0;
  if (inside(c, p[0], p[1], p[2])) { deposit(c, t, p[0], p[1], p[2], amount, factor); return 0.0; }
  return amount;
}
static double deposit_spread(const om_config *c, tally_t *t, const double p[3], const double u[3], double len,
                             double amount, double factor) {
  int n = 1;
  if (c->resolution > 0.0) {
    n = (int)ceil(len / c->resolution);
    if (n < 1) n = 1;
    if (n > 1024) n = 1024;
  }
  double share = amount / (double)n, piece = len / (double)n, out = 0.0;
  for (int k = 0; k < n; ++k) {
    double s = ((double)k + 0.5) * piece, q[3] = {p[0] + u[0] * s, p[1] + u[1] * s, p[2] + u[2] * s};
    out += deposit_or_escape(c, t, q, share, factor);
  }
  return out;
}

/* ------------------------------------------------------------------ stack */
typedef struct { int kind; double e, w, p[3], u[3]; } particle_t;
typedef struct { particle_t *a; int n, cap; } stack_t;
static void push(stack_t *s, int kind, double e, double w, const double p[3], const double u[3], tally_t *t) {
  if (s->n == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 64;
    s->a = (particle_t *)realloc(s->a, sizeof(particle_t) * (size_t)s->cap);
  }
  particle_t *q = &s->a[s->n++];
  q->kind = kind; q->e = e; q->w = w;
  memcpy(q->p, p, sizeof(q->p));
  memcpy(q->u, u, sizeof(q->u));
  t->b.pushes++;
}

static void annihilate(const om_config *c, tally_t *t, stack_t *s, const double p[3], double w, double factor,
                       rng_t *r) {
  if (c->pcut >= ME) { deposit(c, t, p[0], p[1], p[2], w * 2.0 * ME, factor); return; }
  double ct = 1.0 - 2.0 * uniform(r), st = sqrt(fmax(0.0, 1.0 - ct * ct)), phi = TWO_PI * uniform(r);
  double a[3] = {st * cos(phi), st * sin(phi), ct}, b[3] = {-a[0], -a[1], -a[2]};
  push(s, 0, ME, w, p, a, t);
  push(s, 0, ME, w, p, b, t);
}

static double vacuum_entry(const om_config *c, double p[3], const double u[3]) {
  if (inside(c, p[0], p[1], p[2])) return 0.0;
  double d = entry_distance(c, p, u);
  if (isinf(d)) return -1.0;
  d += 1.0e-9;
  for (int a = 0; a < 3; ++a) p[a] += d * u[a];
  return inside(c, p[0], p[1], p[2]) ? 0.0 : -1.0;
}

/* photon_steps: returns escaped weight-energy. */
static double photon(const om_config *c, tally_t *t, stack_t *s, particle_t q, rng_t *r) {
  double e = q.e, w = q.w, p[3], u[3], esc = 0.0;
  memcpy(p, q.p, sizeof(p)); memcpy(u, q.u, sizeof(u));
  if (vacuum_entry(c, p, u) < 0.0) return w * e;
  for (;;) {
    t->b.photon_steps++;
    int pi; double pw;
    locate(log(e), c->p_log_min, c->p_inv, c->n, &pi, &pw);
    double mu_maj = lerp(c->majorant, pi, pw);
    double st = path(mu_maj, r);
    for (int a = 0; a < 3; ++a) p[a] += st * u[a];
    if (!inside(c, p[0], p[1], p[2])) return esc + w * e;
    size_t v = voxel(c, p[0], p[1], p[2]);
    double rho = c->density[v];
    int m = c->material[v];
    t->b.vox++;
    size_t row = (size_t)m * c->n;
    double mc = rho * lerp(c->mu_c + row, pi, pw), mp = rho * lerp(c->mu_p + row, pi, pw);
    double mpp = rho * lerp(c->mu_pp + row, pi, pw), mr = rho * lerp(c->mu_r + row, pi, pw);
    double real = mc + mp + mpp + mr;
    if (uniform(r) * mu_maj >= real) continue;
    double th = uniform(r) * (mc + mp + mpp + mr);
    if (th < mc) { /* Compton; the scattered photon re-enters the stack */
      double ratio = kahn(e, r), recoil = e * (1.0 - ratio), phi = TWO_PI * uniform(r);
      if (c->transport_electrons && recoil > c->ecut) {
        double d[3] = {u[0], u[1], u[2]};
        rotate(d, compton_e_cos(e, ratio), phi + PI);
        push(s, 1, recoil, w, p, d, t);
      } else {
        deposit(c, t, p[0], p[1], p[2], w * recoil, spr(c, recoil, m));
      }
      double g[3] = {u[0], u[1], u[2]};
      rotate(g, compton_cos(e, ratio), phi);
      double es = e * ratio;
      if (es <= c->pcut) { deposit(c, t, p[0], p[1], p[2], w * es, spr(c, es, m)); return esc; }
      double cw = w;
      if (es < 0.5 && cw < 4.0) {
        double nw = uniform(r) < 0.5 ? cw / 0.5 : 0.0;
        esc += (cw - nw) * es;
        if (nw == 0.0) return esc;
        cw = nw;
      }
      push(s, 0, es, cw, p, g, t);
      return esc;
    } else if (th < mc + mp) { /* photoelectric */
      if (c->transport_electrons && e > c->ecut) push(s, 1, e, w, p, u, t);
      else deposit(c, t, p[0], p[1], p[2], w * e, spr(c, e, m));
      return esc;
    } else if (th < mc + mp + mpp) { /* pair */
      double kin = e - 2.0 * ME;
      if (c->transport_electrons) {
        double f = uniform(r), se = f * kin, sp = (1.0 - f) * kin;
        if (se > c->ecut) push(s, 1, se, w, p, u, t);
        else deposit(c, t, p[0], p[1], p[2], w * se, spr(c, se, m));
        if (sp > c->ecut) push(s, 2, sp, w, p, u, t);
        else { deposit(c, t, p[0], p[1], p[2], w * sp, spr(c, sp, m)); annihilate(c, t, s, p, w, spr(c, ME, m), r); }
      } else {
        deposit(c, t, p[0], p[1], p[2], w * kin, spr(c, kin, m));
        annihilate(c, t, s, p, w, spr(c, ME, m), r);
      }
      return esc;
    } else { /* Rayleigh: unreachable with zero coherent data; Thomson angle */
      double mu;
      for (;;) { mu = 1.0 - 2.0 * uniform(r); if (2.0 * uniform(r) <= 1.0 + mu * mu) break; }
      rotate(u, mu, TWO_PI * uniform(r));
    }
  }
}

/* electron_steps: returns escaped weight-energy (incl. a positron's rest mass). */
static double electron(const om_config *c, tally_t *t, stack_t *s, particle_t q, rng_t *r) {
  int pos = q.kind == 2;
  double e = q.e, w = q.w, p[3], u[3], esc = 0.0, latent = pos ? 2.0 * ME : 0.0;
  memcpy(p, q.p, sizeof(p)); memcpy(u, q.u, sizeof(u));
  if (vacuum_entry(c, p, u) < 0.0) return w * (e + latent);
  for (;;) {
    if (e <= c->ecut) {
      int tm = material_at(c, p);
      deposit(c, t, p[0], p[1], p[2], w * e, spr(c, e, tm));
      if (pos) annihilate(c, t, s, p, w, spr(c, ME, tm), r);
      return esc;
    }
    t->b.electron_steps++;
    size_t v = voxel(c, p[0], p[1], p[2]);
    double rho = c->density[v];
    int m = c->material[v];
    t->b.vox++;
    double smax = c->step_fraction * elook(c, c->csda, m, e) / rho;
    if (!c->msc_gs) {
      double sth = 0.10 / (elook(c, c->tpow, m, e) * rho);
      if (sth < smax) smax = sth;
    }
    double sb = face_distance(c, p, u) + 1.0e-4;
    double sg = smax < sb ? smax : sb;
    double sigm = rho * elook(c, c->moller, m, e);
    double sint = sigm > 0.0 ? path(sigm, r) : INFINITY;
    double sl = sint < sg ? sint : sg;
    int pending = sint <= sg;
    double avail = elook(c, c->rrange, m, e), de;
    if (rho * sl >= avail) {
      sl = avail / rho; de = e - c->ecut; pending = 0;
    } else {
      int ri; double rw;
      locate(log(avail - rho * sl), c->r_log_min, c->r_inv, c->n, &ri, &rw);
      double ee = lerp(c->inv_range + (size_t)m * c->n, ri, rw);
      ee = ee < c->ecut ? c->ecut : (ee > e ? e : ee);
      de = e - ee;
    }
    double rad = elook(c, c->srad, m, e) * rho * sl, prob = 0.0, local = rad;
    if (e > c->pcut) {
      double fr = 1.0 - c->pcut / e, mk = (e - c->pcut) / log(e / c->pcut);
      prob = rad * fr / mk;
      local = rad * (1.0 - fr);
    }
    double cont = de + local;
    int emit = uniform(r) < prob;
    double s1 = uniform(r) * sl, s2 = sl - s1, d1 = sl > 0.0 ? cont * (s1 / sl) : 0.0, d2 = cont - d1;
    /* one weight per substep, at the (E, material) that set its stopping power */
    double sf = spr(c, e, m);
    esc += deposit_spread(c, t, p, u, s1, w * d1, sf);
    e -= d1;
    for (int a = 0; a < 3; ++a) p[a] += u[a] * s1;
    if (!inside(c, p[0], p[1], p[2])) return esc + w * (e + latent);
    double ms = elook(c, c->tpow, m, e) * rho * sl, ch;
    if (c->msc_gs) {
      double fx = elook(c, c->log_eta, m, e) * 4.0, fy = log(fmax(ms, 1.0e-12)) * 4.0;
      ch = gs_cos(c, fx, fy, ms, r);
    } else {
      double th = sqrt(-ms * log1p(-uniform(r)));
      ch = cos(th < PI ? th : PI);
    }
    rotate(u, ch, TWO_PI * uniform(r));
    if (emit && e > c->pcut) {
      double k = c->pcut * exp(uniform(r) * log(e / c->pcut));
      if (k > e - d2) k = e - d2;
      if (k > c->pcut) {
        double pw = w;
        if (k < 0.5 && w < 4.0) { pw = uniform(r) < 0.5 ? w / 0.5 : 0.0; esc += (w - pw) * k; }
        if (pw > 0.0) push(s, 0, k, pw, p, u, t);
      } else {
        esc += deposit_or_escape(c, t, p, w * k, spr(c, k, m));
      }
      e -= k;
    }
    esc += deposit_spread(c, t, p, u, s2, w * d2, sf);
    e -= d2;
    for (int a = 0; a < 3; ++a) p[a] += u[a] * s2;
    if (!inside(c, p[0], p[1], p[2])) return esc + w * (e + latent);
    if (pending && e > 2.0 * c->ecut) {
      double dw = moller_w(e, c->ecut, r), two_m = 2.0 * ME;
      double cd = fmin(1.0, sqrt(dw * (e + two_m) / (e * (dw + two_m))));
      double sv = e - dw, cp = fmin(1.0, sqrt(sv * (e + two_m) / (e * (sv + two_m))));
      double phi = TWO_PI * uniform(r), d[3] = {u[0], u[1], u[2]};
      rotate(d, cd, phi);
      push(s, 1, dw, w, p, d, t);
      rotate(u, cp, phi + PI);
      e -= dw;
    }
  }
}

static double spectrum_energy(const om_config *c, rng_t *r) {
  double ub = uniform(r);
  int k = 0;
  while (k < c->n_bins && c->cdf[k] < ub) ++k; /* searchsorted, left */
  if (k > c->n_bins - 1) k = c->n_bins - 1;
  return c->edges[k] + (c->edges[k + 1] - c->edges[k]) * uniform(r);
}

/* Emit one primary for (beamlet) and transport its family; returns emitted energy. */
static double history(const om_config *c, tally_t *t, stack_t *s, int beamlet, long long key,
                      rng_t *r) {
  particle_t q; double emitted;
  q.w = 1.0; q.kind = c->particle;
  memset(q.u, 0, sizeof(q.u)); q.u[2] = 1.0;
  switch (c->src_kind) {
    case 0: q.e = c->energy; memcpy(q.p, c->pos, sizeof(q.p)); memcpy(q.u, c->dir, sizeof(q.u)); break;
    case 1: q.e = c->energy; q.p[0] = c->x0 + c->xe * uniform(r); q.p[1] = c->y0 + c->ye * uniform(r); q.p[2] = c->z0; break;
    case 2: q.kind = 0; q.e = spectrum_energy(c, r); q.p[0] = c->x0 + c->xe * uniform(r);
            q.p[1] = c->y0 + c->ye * uniform(r); q.p[2] = c->z0; break;
    case 5: { const size_t n = (size_t)c->n_prim, k = (size_t)key;
              q.kind = c->prim_kind[k]; q.e = c->prim[k];
              for (int a = 0; a < 3; ++a) { q.p[a] = c->prim[(1 + a) * n + k]; q.u[a] = c->prim[(4 + a) * n + k]; }
              q.w = c->prim[7 * n + k]; break; }
    case 4: { const double *b = c->beamlet + 4 * (size_t)beamlet;
              q.kind = 0; q.e = c->energy; q.p[0] = b[0] + b[1] * uniform(r); q.p[1] = b[2] + b[3] * uniform(r);
              q.p[2] = c->z0; break; }
    case 6: { /* spot fan: energy, rectangle offset, Box-Muller spot; starts on the rectangle
               * (pyradmc/geometry/source.py:760-792) */
              const double *cc = c->beamlet + 3 * (size_t)beamlet;
              q.kind = 0; q.e = spectrum_energy(c, r);
              double du = (uniform(r) - 0.5) * c->wu, dv = (uniform(r) - 0.5) * c->wv;
              double u4 = uniform(r); if (u4 < 1e-12) u4 = 1e-12;
              double rad = sqrt(-2.0 * log(u4)), ang = TWO_PI * uniform(r);
              double su = c->sigma_u * rad * cos(ang), sv = c->sigma_v * rad * sin(ang), d[3];
              for (int a = 0; a < 3; ++a) {
                q.p[a] = cc[a] + du * c->ua[a] + dv * c->va[a];
                d[a] = q.p[a] - (c->focal[a] + su * c->ua[a] + sv * c->va[a]);
              }
              double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
              for (int a = 0; a < 3; ++a) q.u[a] = d[a] / nrm;
              break; }
    default: { const double *cc = c->beamlet + 3 * (size_t)beamlet;
               q.kind = 0; q.e = spectrum_energy(c, r);
               double du = (uniform(r) - 0.5) * c->wu, dv = (uniform(r) - 0.5) * c->wv, d[3];
               for (int a = 0; a < 3; ++a) d[a] = cc[a] + du * c->ua[a] + dv * c->va[a] - c->focal[a];
               double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
               for (int a = 0; a < 3; ++a) { q.p[a] = c->focal[a]; q.u[a] = d[a] / nrm; }
               break; }
  }
  emitted = q.w * (q.e + (q.kind == 2 ? 2.0 * ME : 0.0));
  s->n = 0;
  if (s->cap == 0) { s->cap = 64; s->a = (particle_t *)malloc(sizeof(particle_t) * 64); }
  s->a[s->n++] = q;
  double esc = 0.0;
  while (s->n > 0) {
    particle_t x = s->a[--s->n];
    esc += x.kind == 0 ? photon(c, t, s, x, r) : electron(c, t, s, x, r);
  }
  t->b.escaped += esc;
  return emitted;
}

/* ------------------------------------------------------------------ driver */
typedef struct {
  const om_config *c;
  uint64_t seed;
  const long long *first;   /* per beamlet: first key of this batch slice */
  const long long *count;   /* per beamlet: histories in this batch */
  int n_cols;
  long long total, lo, hi;  /* flat history range of this worker */
  const long long *start;   /* per beamlet exclusive prefix of counts */
  tally_t t;
  stack_t s;
} work_t;

static void *worker(void *arg) {
  work_t *w = (work_t *)arg;
  int j = 0;
  for (long long g = w->lo; g < w->hi; ++g) {
    while (j + 1 < w->n_cols && w->start[j + 1] <= g) ++j;
    long long i = g - w->start[j];
    rng_t r;
    rng_seed(&r, w->seed, (uint64_t)(w->first[j] + i));
    w->t.base = (size_t)j * (size_t)w->c->snx * w->c->sny * w->c->snz;
    w->t.b.emitted += history(w->c, &w->t, &w->s, j, w->first[j] + i, &r);
  }
  return NULL;
}

/* One batch: for each column j, count[j] histories keyed first[j] + i. Adds the batch
 * energy into batch_energy [n_cols][V] (zeroed by the caller) and books. */
int om_batch(const om_config *c, uint64_t seed, int n_cols, const long long *first,
             const long long *count, int threads, double *batch_energy, om_books *books) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  if (c->src_kind == 5)
    for (int j = 0; j < n_cols; ++j)
      if (first[j] < 0 || count[j] < 0 || first[j] + count[j] > c->n_prim) return -1;
  size_t nv = (size_t)c->snx * c->sny * c->snz, cells = nv * (size_t)n_cols;
  long long *start = (long long *)malloc(sizeof(long long) * (size_t)(n_cols + 1));
  start[0] = 0;
  for (int j = 0; j < n_cols; ++j) start[j + 1] = start[j] + count[j];
  long long total = start[n_cols];
  work_t *w = (work_t *)calloc((size_t)threads, sizeof(work_t));
  pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  for (int k = 0; k < threads; ++k) {
    w[k].c = c; w[k].seed = seed; w[k].first = first; w[k].count = count; w[k].n_cols = n_cols;
    w[k].start = start; w[k].total = total;
    w[k].lo = total * k / threads; w[k].hi = total * (k + 1) / threads;
    w[k].t.energy = (double *)calloc(cells, sizeof(double));
    pthread_create(&th[k], NULL, worker, &w[k]);
  }
  for (int k = 0; k < threads; ++k) {
    pthread_join(th[k], NULL);
    for (size_t i = 0; i < cells; ++i) batch_energy[i] += w[k].t.energy[i];
    books->emitted += w[k].t.b.emitted; books->deposited += w[k].t.b.deposited;
    books->escaped += w[k].t.b.escaped; books->unscored += w[k].t.b.unscored;
    books->vox += w[k].t.b.vox; books->deposits += w[k].t.b.deposits; books->pushes += w[k].t.b.pushes;
    books->photon_steps += w[k].t.b.photon_steps; books->electron_steps += w[k].t.b.electron_steps;
    free(w[k].t.energy);
    free(w[k].s.a);
  }
  free(w); free(th); free(start);
  return 0;
}
