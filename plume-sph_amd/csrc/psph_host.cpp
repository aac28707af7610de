// psph_host.cpp -- host side of the MI355X SPH particle path: configuration
// with the reference defaults, the .data deck reader, meteo tables, particle
// keys and ordering, the inlet schedule and the synthetic input generators.
// Everything here is ordinary C++ on the CPU; the per-step particle work is
// in psph_device.hip.

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>

#include "psph_internal.h"

namespace psph {

static thread_local char g_err[1024];

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

// ---------------------------------------------------------------------------
// HostParticles
// ---------------------------------------------------------------------------
#define PSPH_FOR_D(F) F(x) F(y) F(z) F(vx) F(vy) F(vz) F(svx) F(svy) F(svz) F(rho) F(e) F(m) \
  F(h) F(p) F(c) F(T) F(gam) F(cp) F(xi)
#define PSPH_FOR_U(F) F(k0) F(k1) F(k2)
#define PSPH_FOR_I(F) F(bc) F(phase) F(inv)

void HostParticles::resize(size_t n) {
#define R(a) a.resize(n);
  PSPH_FOR_D(R) PSPH_FOR_U(R) PSPH_FOR_I(R)
#undef R
}
void HostParticles::reserve(size_t n) {
#define R(a) a.reserve(n);
  PSPH_FOR_D(R) PSPH_FOR_U(R) PSPH_FOR_I(R)
#undef R
}
size_t HostParticles::push() {
  size_t i = size();
#define R(a) a.push_back(0);
  PSPH_FOR_D(R) PSPH_FOR_U(R) PSPH_FOR_I(R)
#undef R
  return i;
}
void HostParticles::permute(const std::vector<uint32_t> &perm) {
  size_t n = perm.size();
#define R(a) { decltype(a) t(n); for (size_t i = 0; i < n; i++) t[i] = a[perm[i]]; a.swap(t); }
  PSPH_FOR_D(R) PSPH_FOR_U(R) PSPH_FOR_I(R)
#undef R
}
void HostParticles::to_soa(psph_soa *o, size_t first) const {
  size_t n = size();
  auto cp8 = [&](double *d, const std::vector<double> &s) { if (d) std::memcpy(d + first, s.data(), n * 8); };
  auto cp4u = [&](uint32_t *d, const std::vector<uint32_t> &s) { if (d) std::memcpy(d + first, s.data(), n * 4); };
  auto cp4i = [&](int32_t *d, const std::vector<int32_t> &s) { if (d) std::memcpy(d + first, s.data(), n * 4); };
  cp8(o->x, x); cp8(o->y, y); cp8(o->z, z); cp8(o->vx, vx); cp8(o->vy, vy); cp8(o->vz, vz);
  cp8(o->svx, svx); cp8(o->svy, svy); cp8(o->svz, svz); cp8(o->rho, rho); cp8(o->e, e);
  cp8(o->m, m); cp8(o->h, h); cp8(o->p, p); cp8(o->c, c); cp8(o->T, T); cp8(o->gam, gam);
  cp8(o->cp, cp); cp8(o->xi, xi);
  cp4u(o->key0, k0); cp4u(o->key1, k1); cp4u(o->key2, k2);
  cp4i(o->bc, bc); cp4i(o->phase, phase); cp4i(o->involved, inv);
}
void HostParticles::from_soa(const psph_soa *s, size_t n) {
  resize(n);
  auto cp8 = [&](std::vector<double> &d, const double *src) { if (src) std::memcpy(d.data(), src, n * 8); else std::fill(d.begin(), d.end(), 0.0); };
  auto cp4u = [&](std::vector<uint32_t> &d, const uint32_t *src) { if (src) std::memcpy(d.data(), src, n * 4); else std::fill(d.begin(), d.end(), 0u); };
  auto cp4i = [&](std::vector<int32_t> &d, const int32_t *src) { if (src) std::memcpy(d.data(), src, n * 4); else std::fill(d.begin(), d.end(), 0); };
  cp8(x, s->x); cp8(y, s->y); cp8(z, s->z); cp8(vx, s->vx); cp8(vy, s->vy); cp8(vz, s->vz);
  cp8(svx, s->svx); cp8(svy, s->svy); cp8(svz, s->svz); cp8(rho, s->rho); cp8(e, s->e);
  cp8(m, s->m); cp8(h, s->h); cp8(p, s->p); cp8(c, s->c); cp8(T, s->T); cp8(gam, s->gam);
  cp8(cp, s->cp); cp8(xi, s->xi);
  cp4u(k0, s->key0); cp4u(k1, s->key1); cp4u(k2, s->key2);
  cp4i(bc, s->bc); cp4i(phase, s->phase); cp4i(inv, s->involved);
}

// ---------------------------------------------------------------------------
// Hilbert keys.  The curve is the Zoltan BSFC 3D Hilbert map the reference
// ships (hilbert.c:156-271): Gray-coded octant digits, 21 levels from the top
// bits of 32-bit coordinates, with the per-octant axis permutation/reflection
// rules of hilbert.c:219-266.  Here the 48 possible (axis order, reflection)
// states are enumerated once into a transition table.
// ---------------------------------------------------------------------------
namespace {
struct HilbertTable {
  // state: three axis words (axis<<1 | reflect), packed 3 bits each
  uint16_t next[512][8];
  uint8_t digit[512][8];   // octant bit pattern -> Hilbert digit
  HilbertTable() {
    static const uint8_t gray_inv[8] = {0, 1, 3, 2, 7, 6, 4, 5};
    static const uint8_t src[8][3] = {{2, 1, 0}, {0, 2, 1}, {0, 1, 2}, {2, 0, 1},
                                      {2, 0, 1}, {0, 1, 2}, {0, 2, 1}, {2, 1, 0}};
    static const uint8_t flip[8][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0},
                                       {0, 1, 1}, {0, 0, 0}, {0, 1, 1}, {1, 0, 1}};
    for (int s = 0; s < 512; s++)
      for (int b = 0; b < 8; b++) {
        uint8_t ax[3] = {uint8_t(s & 7), uint8_t((s >> 3) & 7), uint8_t((s >> 6) & 7)};
        // b holds the raw coordinate bits of the three (permuted) axes
        unsigned idx = 0;
        for (int n = 0; n < 3; n++) idx |= (((b >> n) & 1u) ^ (ax[n] & 1u)) << n;
        uint8_t c = gray_inv[idx];
        digit[s][b] = c;
        uint8_t nx[3];
        for (int n = 0; n < 3; n++) nx[n] = uint8_t(ax[src[c][n]] ^ flip[c][n]);
        next[s][b] = uint16_t(nx[0] | (nx[1] << 3) | (nx[2] << 6));
      }
  }
};
const HilbertTable &htab() {
  static const HilbertTable t;
  return t;
}
}  // namespace

static uint64_t hilbert63(const uint32_t c[3]) {
  const HilbertTable &T = htab();
  unsigned s = 0 | (2u << 3) | (4u << 6);   // axes x,y,z unreflected
  uint64_t h = 0;
  for (int i = 1; i <= 21; i++) {
    unsigned sh = 32 - i;
    unsigned ax0 = (s & 7) >> 1, ax1 = ((s >> 3) & 7) >> 1, ax2 = ((s >> 6) & 7) >> 1;
    unsigned b = ((c[ax0] >> sh) & 1u) | (((c[ax1] >> sh) & 1u) << 1) | (((c[ax2] >> sh) & 1u) << 2);
    h = (h << 3) | T.digit[s][b];
    s = T.next[s][b];
  }
  return h;
}

void thsfc3d(const double normc[3], uint32_t add_step, uint32_t key[3]) {
  const double imax = 4294967295.0;
  uint32_t c[3];
  for (int k = 0; k < 3; k++) {
    double ck = (normc[k] <= 0) ? 0 : imax * normc[k];
    c[k] = (uint32_t)((imax < ck) ? imax : ck);
  }
  // The reference's two-word key holds Hilbert bits 1..62 of 63: the last
  // digit bit lands past the second word and is dropped (hilbert.c:199-212).
  uint64_t k64 = hilbert63(c) >> 1;
  key[0] = uint32_t(k64 >> 32);
  key[1] = uint32_t(k64);
  key[2] = add_step;
}

int64_t table_slot(uint32_t k0, uint32_t k1, int table_size, int dim) {
  if (dim == 1) return (int64_t)k0;                   // 1D hash = key[0]
  const double umax = 4294967295.0;
  int S02 = table_size / 10;
  int S01 = table_size - S02;
  int i1 = (int)((double)k0 / umax * S01);
  int i2 = (int)((double)k1 / umax * S02);
  return (int64_t)i1 + i2;
}

std::vector<uint32_t> canonical_order(const psph_config &cfg, const uint32_t *k0, const uint32_t *k1,
                                      const uint32_t *k2, size_t n) {
  std::vector<int64_t> slot(n);
  for (size_t i = 0; i < n; i++) slot[i] = table_slot(k0[i], k1[i], cfg.p_table_size, cfg.dim);
  std::vector<uint32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0u);
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
    if (slot[a] != slot[b]) return slot[a] < slot[b];
    if (k0[a] != k0[b]) return k0[a] < k0[b];
    if (k1[a] != k1[b]) return k1[a] < k1[b];
    return k2[a] < k2[b];
  });
  return perm;
}

void eos(const psph_config &cfg, double rho, double e, double xi, double &p, double &c, double &T,
         double &gam, double &cp) {
  double na = 1 - xi, ng = cfg.ng0 * xi, ns = xi - ng;
  double Cvm = ns * cfg.cvs + ng * cfg.cvg + na * cfg.cva;
  double Rm = ng * cfg.rg + na * cfg.ra;
  double Rmg = Rm / (ng + na);
  gam = 1 + Rm / Cvm;
  double desmg = rho * (na + ng);
  p = Rmg * desmg * e / Cvm;
  c = std::pow(gam * p / desmg, 0.5);
  T = e / Cvm;
  if (cfg.lans) cp = Cvm;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2,
// 3", SC'11): counter-based, so each lattice point's jitter depends only on
// its index and the seed.
// ---------------------------------------------------------------------------
static void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t a = uint64_t(0xD2511F53u) * c[0], b = uint64_t(0xCD9E8D57u) * c[2];
    uint32_t n0 = uint32_t(b >> 32) ^ c[1] ^ k0, n2 = uint32_t(a >> 32) ^ c[3] ^ k1;
    c[0] = n0; c[1] = uint32_t(b); c[2] = n2; c[3] = uint32_t(a);
  }
}

// ---------------------------------------------------------------------------
// Inlet (add_new_erupt, setup_erupt.cc:389-670)
// ---------------------------------------------------------------------------
static double vel_prof(const psph_config &cfg, double R, double rsq, double u) {
  return cfg.vent_profile == 0 ? u : (u * 2) * (1 - (rsq / (R * R)));   // sph_lib.cc:1223-1243
}

// Vent column centres in the reference's scan order: x+ then x-, each with
// y+ then y- (setup_erupt.cc:105-205 and 458-660).
template <class F>
static void vent_columns(const psph_config &cfg, F &&fn) {
  double sml = cfg.vent_sml, sml2 = 0.5 * sml;
  for (int sx = 0; sx < 2; sx++) {
    double x = sx == 0 ? cfg.vent_x + sml2 : cfg.vent_x - sml2;
    while (sx == 0 ? x <= cfg.vent_r : x >= -cfg.vent_r) {
      for (int sy = 0; sy < 2; sy++) {
        double y = sy == 0 ? cfg.vent_y + sml2 : cfg.vent_y - sml2;
        while (sy == 0 ? y <= cfg.vent_r : y >= -cfg.vent_r) {
          fn(x, y);
          y = sy == 0 ? y + sml : y - sml;
        }
      }
      x = sx == 0 ? x + sml : x - sml;
    }
  }
}

static void new_erupt(const psph_config &cfg, HostParticles &P, double x, double y, double z,
                      uint32_t add_step, double vel) {
  size_t i = P.push();
  P.x[i] = x; P.y[i] = y; P.z[i] = z;
  double normc[3] = {(x - cfg.ext_lo[0]) / (cfg.ext_hi[0] - cfg.ext_lo[0]),
                     (y - cfg.ext_lo[1]) / (cfg.ext_hi[1] - cfg.ext_lo[1]),
                     (z - cfg.ext_lo[2]) / (cfg.ext_hi[2] - cfg.ext_lo[2])};
  uint32_t key[3];
  thsfc3d(normc, add_step, key);
  P.k0[i] = key[0]; P.k1[i] = key[1]; P.k2[i] = key[2];
  // erupt-ghost constructor (particle.cc:210-313)
  P.m[i] = cfg.vent_mass;
  P.h[i] = cfg.ratio_sml_dx * cfg.vent_sml;
  P.bc[i] = PSPH_BC_ERUPT_GHOST; P.inv[i] = 0; P.phase[i] = 2; P.xi[i] = 1.0;
  P.vz[i] = vel;
  P.e[i] = cfg.vent_e;
  P.rho[i] = cfg.vent_rho;
  double pd, cd, gd;
  eos(cfg, P.rho[i], P.e[i], 1.0, pd, cd, P.T[i], gd, P.cp[i]);
  P.p[i] = cfg.vent_p;
  P.c[i] = cfg.vent_c;
  P.gam[i] = cfg.vent_gamma;
}

int inject_host(const psph_config &cfg, double time, double dt, int64_t step, HostParticles &out,
                std::vector<int32_t> *rec, const int32_t *rep, size_t rep_len) {
  double t_total = time - dt;   // time of the previous step (setup_erupt.cc:451)
  double rvsq = cfg.vent_r * cfg.vent_r;
  size_t col = 0;
  int rc = 0;
  vent_columns(cfg, [&](double x, double y) {
    double dist = 0;
    dist += x * x;
    dist += y * y;
    if (!(dist <= rvsq)) return;
    double vel = vel_prof(cfg, cfg.vent_r, dist, cfg.vent_v);
    double t_each = cfg.vent_sml / vel;
    double t_add;
    int n;
    if (rep) {
      if (2 * col + 1 >= rep_len) { rc = PSPH_EINVAL; return; }
      double q = rep[2 * col];
      n = rep[2 * col + 1];
      t_add = (t_total - q * t_each) + dt;
      t_add = t_add - n * t_each;
    } else {
      double r = std::fmod(t_total, t_each);
      t_add = r + dt;
      n = (int)std::floor(t_add / t_each);
      t_add = t_add - n * t_each;
      if (rec) {
        rec->push_back((int32_t)std::llround((t_total - r) / t_each));
        rec->push_back(n);
      }
    }
    col++;
    for (int i = 0; i < n; i++)
      new_erupt(cfg, out, x, y, cfg.vent_bot + (t_add + i * t_each) * vel, (uint32_t)step, vel);
  });
  if (rc) return fail(rc, "inject: recorded inlet schedule has fewer columns than the vent");
  return 0;
}

void lattice_point(const psph_config &cfg, uint32_t seed, double jitter, const int64_t l[3],
                   const int64_t nl[3], double exact[3], double pos[3]) {
  const double dx = cfg.h_air, dx2 = 0.5 * dx, del = cfg.bucket_edge;
  const int per = (int)std::llround(del / dx);
  for (int k = 0; k < 3; k++) {
    double mincrd = cfg.ext_lo[k] + (double)(l[k] / per) * del;   // preprocess.cc:147-159
    exact[k] = pos[k] = mincrd + dx2 + (double)(l[k] % per) * dx;   // add_air.cc:575-577
  }
  if (jitter > 0) {
    uint64_t L = (uint64_t)((l[2] * nl[1] + l[1]) * nl[0] + l[0]);
    uint32_t r[4] = {uint32_t(L), uint32_t(L >> 32), 0, 0};
    philox(r, seed, 0x50535048u);
    for (int k = 0; k < 3; k++) {
      double u = ((double)r[k] + 0.5) * (1.0 / 4294967296.0);
      pos[k] += (2.0 * u - 1.0) * jitter * dx;
    }
  }
}

size_t push_air(const psph_config &cfg, const psph_meteo *met, HostParticles &P, const double pc[3],
                double h, int bc, int inv, uint32_t add_step) {
  size_t i = P.push();
  P.x[i] = pc[0]; P.y[i] = pc[1]; P.z[i] = pc[2];
  double normc[3];
  for (int k = 0; k < 3; k++) normc[k] = (pc[k] - cfg.ext_lo[k]) / (cfg.ext_hi[k] - cfg.ext_lo[k]);
  uint32_t key[3];
  thsfc3d(normc, add_step, key);
  P.k0[i] = key[0]; P.k1[i] = key[1]; P.k2[i] = key[2];
  P.h[i] = h;
  P.bc[i] = bc; P.phase[i] = 1; P.inv[i] = inv; P.xi[i] = 0;
  // initial_air + air_prop_meteo_based (sph_lib.cc:724-766, 831-872)
  double sml2 = 0.5 * P.h[i];
  double range[6] = {pc[0] - sml2, pc[0] + sml2, pc[1] - sml2, pc[1] + sml2, pc[2] - sml2, pc[2] + sml2};
  double pressure = meteo_interp(met, pc[2] * 0.001, 1) * 100;
  double density = meteo_interp(met, pc[2] * 0.001, 0);
  double energy = pressure / (density * (cfg.gamma_p - 1));
  double d1 = meteo_interp(met, range[4] * 0.001, 0);
  double d2 = meteo_interp(met, range[5] * 0.001, 0);
  double mass = (range[1] - range[0]) * (range[3] - range[2]) * (range[5] - range[4]) *
                (0.1666667 * d1 + 0.6666666 * density + 0.1666667 * d2);
  P.rho[i] = density; P.e[i] = energy; P.p[i] = pressure; P.m[i] = mass;
  eos(cfg, density, energy, 0.0, P.p[i], P.c[i], P.T[i], P.gam[i], P.cp[i]);
  return i;
}

void setup_vent(const psph_config &cfg, HostParticles &P) {
  // every non-erupt particle inside the vent duct goes (setup_erupt.cc:302-320)
  const double rvsq = cfg.vent_r * cfg.vent_r;
  std::vector<uint32_t> keep;
  keep.reserve(P.size());
  for (size_t i = 0; i < P.size(); i++) {
    double dist = 0;
    dist += P.x[i] * P.x[i];
    dist += P.y[i] * P.y[i];
    if (P.bc[i] != PSPH_BC_ERUPT_GHOST && dist <= rvsq && P.z[i] <= cfg.vent_z && P.z[i] >= cfg.vent_bot) continue;
    keep.push_back((uint32_t)i);
  }
  if (keep.size() != P.size()) P.permute(keep);
  // the duct's erupt ghosts (setup_erupt.cc:105-205)
  double z = cfg.vent_z - 0.5 * cfg.vent_sml;
  while (z >= cfg.vent_bot) {
    vent_columns(cfg, [&](double x, double y) {
      double dist = 0;
      dist += x * x;
      dist += y * y;
      if (dist <= rvsq) new_erupt(cfg, P, x, y, z, 0, vel_prof(cfg, cfg.vent_r, dist, cfg.vent_v));
    });
    z -= cfg.vent_sml;
  }
}

}  // namespace psph

using namespace psph;

#define PSPH_TRY_H(expr)          \
  do {                            \
    int rc_ = (expr);             \
    if (rc_) return rc_;          \
  } while (0)

// ===========================================================================
// C ABI: configuration
// ===========================================================================
extern "C" int psph_config_default(psph_config *c, int dim) {
  if (!c) return fail(PSPH_EINVAL, "null config");
  if (dim != 1 && dim != 3) return fail(PSPH_EINVAL, "dim must be 1 or 3 (got %d)", dim);
  std::memset(c, 0, sizeof *c);
  c->dim = dim;
  c->precision = PSPH_FP64;
  c->kernel = PSPH_KERNEL_GAUSS_BOX;
  c->cutoff = 3.0;                       // constant.h:42,137
  c->list_factor = 4.25;                 // neighs.cc:84,130
  c->ata = 0.01;                         // parameters.h:998, :1403
  c->eps_half = 0.4;                     // constant.h:175
  c->prandtl = 0.85;                     // constant.h:176
  c->e_smooth_ratio = 4.0;               // constant.h:187
  c->heat_scale = 1.0;                   // constant.h:186
  c->thresh = 1e-5;                      // parameters.h:1002, :1407
  c->c_smooth = 2.0;                     // parameters.h:1003, :1408
  c->ratio_sml_dx = 1.0;                 // constant.h:190
  c->msfrc_thresh = 1.0e-6;              // constant.h:167
  c->cvs = 1100; c->cvg = 1348; c->cva = 717; c->rg = 462; c->ra = 287;  // parameters.h:906-958
  c->gamma_p = 1.4;                      // parameters.h:917
  c->vent_c = 340.;                      // setup_erupt.cc:64 (sndspd)
  c->max_steps = 100000000;
  c->t_end = 100000.0;
  c->t_output = 1.0;
  c->erupt_start = 0.0;
  c->erupt_end = 7200.0;
  if (dim == 3) {                        // options.h:452-837
    c->lans = 1; c->energy_smooth = 1; c->pghost_contrib = 1; c->list_all = 0;
    c->num_loop = 2; c->sml_update_int = 50; c->adapt_at_setup = 0;  // parameters.h:1001, constant.h:189
    c->vent_profile = 1;                 // ERUPT_VELOCITY_PROF 1
    c->num_erupt_particles = 1080;       // parameters.h:983
    c->p_table_size = 4000000;           // dataread.cc:254
    c->cutoff2 = 8.0;                    // constant.h:138
    c->alpha = 0.3; c->beta = 0.6;       // parameters.h:996-997
    c->g = 9.80665;                      // parameters.h:894
    c->cfl = 0.80;                       // parameters.h:1026
    c->eta = 1.5;                        // parameters.h:1004
    c->ng0 = 0.026;                      // parameters.h:944
    const double Ll[3] = {-20800, -20800, 1666}, Lu[3] = {20800, 20800, 24000};  // parameters.h:889-890
    for (int k = 0; k < 3; k++) { c->dom_lo[k] = Ll[k]; c->dom_hi[k] = Lu[k]; }
    c->h_air = 100.0;
    c->vent_r = 120; c->vent_v = 75.0; c->vent_T = 1273.15; c->vent_p = 0.831e5;  // parameters.h:946-981
    c->vent_x = 0; c->vent_y = 0; c->vent_z = Ll[2];
  } else {                               // options.h:16-450, parameters.h:1225-1446
    c->lans = 0; c->energy_smooth = 0; c->pghost_contrib = 2; c->list_all = 1;
    c->num_loop = 5; c->sml_update_int = 1; c->adapt_at_setup = 1;
    c->vent_profile = 1;
    c->num_erupt_particles = 0;
    c->p_table_size = 10000;             // dataread.cc:506
    c->cutoff2 = 12.0;                   // constant.h:43
    c->alpha = 1.0; c->beta = 2.0;       // parameters.h:1401-1402
    c->g = 0.0;                          // ATMOSPHERE_TYPE 2 -> g_P = 0
    c->cfl = 0.1;                        // parameters.h:1429
    c->eta = 1.0;                        // parameters.h:1409
    c->ng0 = 0.05;                       // parameters.h:1330
    c->dom_lo[0] = -0.4; c->dom_hi[0] = 0.4;   // SHOCK_TUBE_TESTS 1
    c->h_air = 0.003;
  }
  return PSPH_OK;
}

extern "C" int psph_config_vent(psph_config *c) {
  if (!c) return fail(PSPH_EINVAL, "null config");
  if (c->dim != 3) return PSPH_OK;
  if (!(c->ng0 > 0 && c->vent_T > 0 && c->vent_p > 0)) return fail(PSPH_EINVAL, "vent needs ng0, T, p > 0");
  // parameters.h:958-975
  double Rv = c->ng0 * c->rg;
  double Cvv = c->ng0 * c->cvg + (1 - c->ng0) * c->cvs;
  c->vent_gamma = 1 + Rv / Cvv;
  double rhov_g = c->vent_p / (c->rg * c->vent_T);
  c->vent_rho = rhov_g / c->ng0;
  c->vent_e = c->ng0 * c->cvg * c->vent_T + (1 - c->ng0) * c->cvs * c->vent_T;
  // setup_erupt.cc:86-104 with Compute_mass (sph_lib.cc:1024-1040)
  double rz1 = c->vent_z;
  double rz0 = rz1 - (c->h_air) * 1.5 * 8;   // EXT_DOM_COF * PARTICLE_DENSITY
  double dom[3] = {c->vent_r - (-c->vent_r), c->vent_r - (-c->vent_r), rz1 - rz0};
  double volume = 1.;
  for (int i = 0; i < 3; i++) volume *= dom[i];
  int np = c->num_erupt_particles > 0 ? c->num_erupt_particles : 1;
  c->vent_sml = std::cbrt(volume / np);
  c->vent_mass = c->vent_rho * volume / np;
  c->vent_bot = rz0;
  return PSPH_OK;
}

extern "C" int psph_config_domain(psph_config *c) {
  if (!c) return fail(PSPH_EINVAL, "null config");
  if (!(c->h_air > 0)) return fail(PSPH_EINVAL, "h_air must be > 0");
  if (c->dim == 1) {
    c->bucket_edge = 6 * c->h_air;       // PARTICLE_DENSITY 6 (1D); unused: lists are all-pairs
    for (int k = 0; k < 3; k++) { c->ext_lo[k] = c->dom_lo[k]; c->ext_hi[k] = c->dom_hi[k]; }
    return PSPH_OK;
  }
  double del = 8 * c->h_air;             // PARTICLE_DENSITY * smlen (preprocess.cc:88)
  c->bucket_edge = del;
  for (int k = 0; k < 3; k++) {          // preprocess.cc:95-118
    double lo = c->dom_lo[k] - 1.5 * del;
    double hi = c->dom_hi[k] + 1.5 * del;
    int nb = (int)std::ceil((hi - lo) / del);
    c->ext_lo[k] = lo;
    c->ext_hi[k] = lo + nb * del;
  }
  return PSPH_OK;
}

// ===========================================================================
// C ABI: deck and meteo
// ===========================================================================
extern "C" int psph_read_deck(const char *path, psph_config *cfg, char *meteo_path_out,
                              size_t meteo_cap, int32_t *meteo_rows, int32_t *meteo_cols) {
  if (!path || !cfg) return fail(PSPH_EINVAL, "null argument");
  std::ifstream in(path);
  if (!in.good()) return fail(PSPH_EIO, "cannot open deck %s", path);
  std::vector<std::string> lines;
  std::string line;
  while (std::getline(in, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == '\n')) line.pop_back();
    lines.push_back(line);
  }
  if (lines.size() < 15) return fail(PSPH_EIO, "deck %s has %zu lines, needs at least 15", path, lines.size());
  // value = text after the last ':' converted with stof (dataread.cc:107-240)
  auto val = [&](size_t i, double &out) -> int {
    const std::string &l = lines[i];
    size_t pos = l.find_last_of(':');
    std::string t = l.substr(pos == std::string::npos ? 0 : pos + 1);
    char *end = nullptr;
    float f = std::strtof(t.c_str(), &end);
    if (end == t.c_str()) return fail(PSPH_EIO, "deck line %zu has no number: '%s'", i + 1, l.c_str());
    out = (double)f;
    return 0;
  };
  double v[17];
  for (size_t i = 0; i < 15; i++)
    if (int rc = val(i, v[i])) return rc;
  cfg->t_end = v[0];
  cfg->max_steps = (int32_t)v[1];
  cfg->t_output = v[2];
  cfg->erupt_start = v[3];
  cfg->erupt_end = v[4];
  cfg->h_air = v[8];
  cfg->dom_lo[0] = v[9];  cfg->dom_hi[0] = v[10];
  cfg->dom_lo[1] = v[11]; cfg->dom_hi[1] = v[12];
  cfg->dom_lo[2] = v[13]; cfg->dom_hi[2] = v[14];
  if (cfg->dim == 3) cfg->vent_z = cfg->dom_lo[2];
  int32_t rows = 0, cols = 0;
  if (lines.size() >= 18) {
    double r, c;
    if (int rc = val(15, r)) return rc;
    if (int rc = val(16, c)) return rc;
    rows = (int32_t)r; cols = (int32_t)c;
    const std::string &l = lines[17];
    size_t pos = l.find_last_of(':');
    std::string fn = l.substr(pos == std::string::npos ? 0 : pos + 1);
    if (meteo_path_out && meteo_cap) {
      std::snprintf(meteo_path_out, meteo_cap, "%s", fn.c_str());
    }
  } else if (meteo_path_out && meteo_cap) {
    meteo_path_out[0] = 0;
  }
  if (meteo_rows) *meteo_rows = rows;
  if (meteo_cols) *meteo_cols = cols;
  return PSPH_OK;
}

extern "C" int psph_meteo_read(const char *path, int32_t rows, int32_t cols, double *z, double *data) {
  if (!path || !z || !data || rows <= 0 || cols < 2) return fail(PSPH_EINVAL, "bad meteo arguments");
  std::ifstream in(path);
  if (!in.good()) return fail(PSPH_EIO, "cannot open meteo file %s", path);
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) {
      double v;
      if (!(in >> v)) return fail(PSPH_EIO, "meteo file %s: fewer than %d x %d numbers", path, rows, cols);
      if (c == 0) z[r] = v; else data[(size_t)(c - 1) * rows + r] = v;
    }
  return PSPH_OK;
}

extern "C" int psph_meteo_standard(double t0, double p0, double z_top_km, double dz_km, double *z,
                                   double *data, int32_t cap, int32_t *rows) {
  if (!(dz_km > 0) || !(z_top_km > 0) || !rows) return fail(PSPH_EINVAL, "bad standard-atmosphere arguments");
  int n = (int)std::floor(z_top_km / dz_km + 1e-9) + 1;
  *rows = n;
  if (!z || !data) return PSPH_OK;
  if (n > cap) return fail(PSPH_ECAPACITY, "standard atmosphere needs %d rows (cap %d)", n, cap);
  const double R = 287.0, g = 9.80665, L = 6.5e-3, z11 = 11000.0;
  const double expo = g / (R * L);   // 5.2559
  double T11 = t0 - L * z11;
  double p11 = p0 * std::pow(T11 / t0, expo);
  for (int r = 0; r < n; r++) {
    double zm = r * dz_km * 1000.0;
    double T, p;
    if (zm <= z11) { T = t0 - L * zm; p = p0 * std::pow(T / t0, expo); }
    else { T = T11; p = p11 * std::exp(-g * (zm - z11) / (R * T11)); }
    z[r] = zm / 1000.0;
    data[0 * (size_t)n + r] = p / (R * T);  // rho
    data[1 * (size_t)n + r] = p / 100.0;    // hPa
    data[2 * (size_t)n + r] = T;
    data[3 * (size_t)n + r] = 0.0;          // q
    data[4 * (size_t)n + r] = 0.0;          // u
    data[5 * (size_t)n + r] = 0.0;          // v
  }
  return PSPH_OK;
}

// Meteo::interpolate (meteo.cc:64-143)
double psph::meteo_interp(const psph_meteo *m, double z0, int prop) {
  const double *z = m->z, *d = m->data + (size_t)prop * m->rows;
  int R = m->rows;
  for (int r = 0; r < R; r++)
    if (z[r] == z0) return d[r];
  if (z0 < z[0]) return d[0];
  if (z0 > z[R - 1]) return d[R - 1];
  int ui = int(std::upper_bound(z, z + R, z0) - z);
  int li = ui - 1;
  return d[li] + (z0 - z[li]) / (z[ui] - z[li]) * (d[ui] - d[li]);
}

extern "C" int psph_keys(const psph_config *cfg, const double *x, const double *y, const double *z,
                         size_t n, uint32_t add_step, uint32_t *k0, uint32_t *k1, uint32_t *k2) {
  if (!cfg || !x || !y || !z || !k0 || !k1 || !k2) return fail(PSPH_EINVAL, "null argument");
  for (size_t i = 0; i < n; i++) {
    double normc[3] = {(x[i] - cfg->ext_lo[0]) / (cfg->ext_hi[0] - cfg->ext_lo[0]),
                       (y[i] - cfg->ext_lo[1]) / (cfg->ext_hi[1] - cfg->ext_lo[1]),
                       (z[i] - cfg->ext_lo[2]) / (cfg->ext_hi[2] - cfg->ext_lo[2])};
    uint32_t key[3];
    thsfc3d(normc, add_step, key);
    k0[i] = key[0]; k1[i] = key[1]; k2[i] = key[2];
  }
  return PSPH_OK;
}

// ===========================================================================
// C ABI: synthetic inputs
// ===========================================================================
extern "C" int psph_gen_plume(const psph_config *cfg, const psph_meteo *met, uint32_t seed,
                              double jitter, psph_soa *out, size_t cap, size_t *n_out) {
  if (!cfg || !met || !n_out) return fail(PSPH_EINVAL, "null argument");
  if (cfg->dim != 3) return fail(PSPH_EINVAL, "psph_gen_plume needs dim 3");
  if (!(cfg->bucket_edge > 0) || !(cfg->vent_sml > 0)) return fail(PSPH_EINVAL, "call psph_config_domain and psph_config_vent first");
  const double dx = cfg->h_air, del = cfg->bucket_edge;
  const int per = (int)std::llround(del / dx);
  int64_t nl[3];
  for (int k = 0; k < 3; k++) nl[k] = std::llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / del) * per;
  // vent column count bound
  double rr = cfg->vent_r / cfg->vent_sml + 2;
  size_t ncol = (size_t)(4 * rr * rr) + 16;
  size_t nz = (size_t)((cfg->vent_z - cfg->vent_bot) / cfg->vent_sml) + 2;
  size_t bound = (size_t)(nl[0] * nl[1] * nl[2]) + ncol * nz;
  if (!out) { *n_out = bound; return PSPH_OK; }

  HostParticles P;
  if (cfg->adapt_domain) {
    // domain adaptation: the initial domain's buckets and their ghost layers
    PSPH_TRY_H(gen_plume_adaptive(*cfg, met, seed, jitter, P));
  } else {
    P.reserve(bound);
    for (int64_t lz = 0; lz < nl[2]; lz++)
      for (int64_t ly = 0; ly < nl[1]; ly++)
        for (int64_t lx = 0; lx < nl[0]; lx++) {
          const int64_t l[3] = {lx, ly, lz};
          double exact[3], pc[3];
          lattice_point(*cfg, seed, jitter, l, nl, exact, pc);
          int bc;
          if (!(pc[2] > cfg->dom_lo[2])) bc = PSPH_BC_WALL_GHOST;
          else if (pc[0] < cfg->dom_lo[0] || pc[0] > cfg->dom_hi[0] || pc[1] < cfg->dom_lo[1] ||
                   pc[1] > cfg->dom_hi[1] || pc[2] > cfg->dom_hi[2]) bc = PSPH_BC_PRESSURE_GHOST;
          else bc = PSPH_BC_REAL;
          push_air(*cfg, met, P, pc, cfg->ratio_sml_dx * cfg->h_air, bc, bc == PSPH_BC_REAL ? 1 : 0, 0);
        }
    setup_vent(*cfg, P);
  }
  auto perm = canonical_order(*cfg, P.k0.data(), P.k1.data(), P.k2.data(), P.size());
  P.permute(perm);
  if (P.size() > cap) return fail(PSPH_ECAPACITY, "gen_plume: %zu particles exceed cap %zu", P.size(), cap);
  P.to_soa(out);
  *n_out = P.size();
  return PSPH_OK;
}

extern "C" int psph_gen_shock_tube(const psph_config *cfg, const double left[3], const double right[3],
                                   double mid, int32_t n_nosave_l, int32_t n_nosave_r, int32_t n_ghost,
                                   int32_t equal_mass, psph_soa *out, size_t cap, size_t *n_out) {
  if (!cfg || !left || !right || !n_out) return fail(PSPH_EINVAL, "null argument");
  const double des_l = left[0], prss_l = left[1], vel_l = left[2];
  const double des_r = right[0], prss_r = right[1], vel_r = right[2];
  double dx_l = cfg->h_air;
  double dx_r = equal_mass ? dx_l * des_l / des_r : cfg->h_air;
  double smlen = std::max(dx_r, dx_l);
  double dx2_l = 0.5 * dx_l;
  double bnd0 = cfg->dom_lo[0], bnd1 = cfg->dom_hi[0];
  HostParticles P;
  uint32_t num = 0;
  auto add = [&](double x, double m, double h, double des, double vel, double prss, int bc, int inv) {
    size_t i = P.push();
    P.k0[i] = num++;
    P.x[i] = x; P.m[i] = m; P.h[i] = cfg->ratio_sml_dx * h;
    P.bc[i] = bc; P.inv[i] = inv; P.phase[i] = 1;
    P.rho[i] = des; P.vx[i] = vel; P.e[i] = prss / ((cfg->gamma_p - 1) * des);
    P.p[i] = prss; P.gam[i] = cfg->gamma_p; P.c[i] = 340.; P.xi[i] = 0.;
  };
  auto rho = [&](double x) { return x > mid ? des_r : des_l; };
  auto prs = [&](double x) { return x > mid ? prss_r : prss_l; };
  auto vel = [&](double x) { return x > mid ? vel_r : vel_l; };
  double pcrd = 0;
  for (int k = 0; k < n_nosave_l; k++) {     // add_air.cc:289-311
    pcrd = bnd0 - (dx2_l + k * dx_l);
    double des = rho(pcrd);
    add(pcrd, des * dx_l, smlen, des, vel_l, prss_l, PSPH_BC_NOSAVE_REAL, 0);
  }
  for (int k = 0; k < n_ghost; k++) {        // add_air.cc:314-340 (ADAPTIVE_SML_GHOST 11)
    pcrd -= dx_l;
    double des = rho(pcrd);
    add(pcrd, des * dx_l, dx_l * 1.1, des, vel_l, prss_l, PSPH_BC_ERUPT_GHOST, 0);
  }
  pcrd = bnd0 + dx2_l;                       // add_air.cc:342-367
  while (pcrd <= mid) {
    double des = rho(pcrd);
    add(pcrd, des * dx_l, smlen, des, vel(pcrd), prs(pcrd), PSPH_BC_REAL, 2);
    pcrd += dx_l;
  }
  pcrd = pcrd - dx_l + dx_r;                 // add_air.cc:369-394
  while (pcrd <= bnd1) {
    double des = rho(pcrd);
    add(pcrd, des * dx_r, smlen, des, vel(pcrd), prs(pcrd), PSPH_BC_REAL, 2);
    pcrd += dx_r;
  }
  for (int k = 0; k < n_nosave_r; k++) {     // add_air.cc:396-419
    double des = rho(pcrd);
    add(pcrd, des * dx_r, smlen, des, vel_r, prss_r, PSPH_BC_NOSAVE_REAL, 2);
    pcrd += dx_r;
  }
  for (int k = 0; k < n_ghost; k++) {        // add_air.cc:421-451
    double des = rho(pcrd);
    add(pcrd, des * dx_r, dx_r * 1.1, des, vel_r, prss_r, PSPH_BC_ERUPT_GHOST, 2);
    pcrd += dx_r;
  }
  if (!out) { *n_out = P.size(); return PSPH_OK; }
  if (P.size() > cap) return fail(PSPH_ECAPACITY, "shock tube: %zu particles exceed cap %zu", P.size(), cap);
  P.to_soa(out);
  *n_out = P.size();
  return PSPH_OK;
}

extern "C" const char *psph_last_error(void) { return psph::g_err; }
extern "C" const char *psph_version(void) { return "psph 0.1.0 (gfx950)"; }
extern "C" size_t psph_abi_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(psph_config);
    case 1: return sizeof(psph_soa);
    case 2: return sizeof(psph_step_out);
    case 3: return sizeof(psph_adjoint_out);
    default: return 0;
  }
}
