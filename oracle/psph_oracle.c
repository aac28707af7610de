/*
 * psph_oracle.c -- CPU FP64 restatement of the Plume-SPH per-step particle
 * path.  TEST INFRASTRUCTURE ONLY (the parity checker): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference file:line it restates.  Floating-point
 * expressions keep the reference's evaluation order so that the restatement
 * reproduces the reference's arithmetic, not just its mathematics.
 * Parallelism: OpenMP over target particles; every particle's result depends
 * only on its own loop, so the output is independent of the thread count.
 */
#define _GNU_SOURCE
#include "psph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
#define ORC_FAIL(code, ...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return (code); } while (0)
const char *orc_last_error(void) { return g_err; }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n; return 1;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ===================================================================== */
/* Kernel: Gaussian truncated per axis at |s_k| > CUTOFF (sph_lib.cc:207-253) */
/* ===================================================================== */
static const double ORC_INV_SQRT_PI = 0.564189583547756;   /* sph_lib.cc:207 */
#define ORC_CUTOFF 3.0                                    /* constant.h:137 */

double orc_weight(const double *s, double h, int dim) {   /* sph_lib.cc:211-231 */
  for (int i = 0; i < dim; i++)
    if (fabs(s[i]) > ORC_CUTOFF) return 0;
  double norm = 1, expt = 0;
  for (int i = 0; i < dim; i++) {
    norm *= ORC_INV_SQRT_PI / h;
    expt += s[i] * s[i];
  }
  return norm * exp(-expt);
}

double orc_d_weight(const double *s, double h, int dir, int dim) { /* :241-253 */
  for (int i = 0; i < dim; i++)
    if (fabs(s[i]) > ORC_CUTOFF) return 0;
  double tmp = -2 * s[dir] / h;
  return tmp * orc_weight(s, h, dim);
}

/* in_support (sph_header.h:49-58): |dx_k| <= h for every axis */
static inline int in_support(const double *dx, double h, int dim) {
  for (int i = 0; i < dim; i++)
    if (fabs(dx[i]) > h) return 0;
  return 1;
}

/* ===================================================================== */
/* EOS: Particle::update_second_var (particle.cc:528-604), ideal-gas      */
/* mixture of erupted gas + air carrying solids.  Closed form.            */
/* ===================================================================== */
void orc_eos(const psph_config *cfg, double rho, double e, double xi,
             double *p, double *c, double *T, double *gam, double *cp) {
  double na = 1 - xi;
  double ng = cfg->ng0 * xi;
  double ns = xi - ng;
  double Cvm = ns * cfg->cvs + ng * cfg->cvg + na * cfg->cva;
  double Rm = ng * cfg->rg + na * cfg->ra;
  double Rmg = Rm / (ng + na);
  double g = 1 + Rm / Cvm;
  double desmg = rho * (na + ng);
  double pr = Rmg * desmg * e / Cvm;
  *gam = g;
  *p = pr;
  *c = pow(g * pr / desmg, 0.5);
  *T = e / Cvm;
  if (cfg->lans) *cp = Cvm;
}

/* ===================================================================== */
/* Hilbert keys: Zoltan-derived 3D Hilbert curve (hilbert.c:156-271) and  */
/* the particle key THSFC3d (hilbert.c:329-359).                          */
/* ===================================================================== */
/* inverse of the 3-bit reflected Gray code g(k) = k ^ (k >> 1) */
static const unsigned char GRAY_INV[8] = {0, 1, 3, 2, 7, 6, 4, 5};
/* Per Hilbert digit c: the next level's (axis<<1 | reflect) words are the
 * current ones permuted (SRC) and reflected (FLIP) -- hilbert.c:219-266. */
static const unsigned char HSRC[8][3] = {
  {2, 1, 0}, {0, 2, 1}, {0, 1, 2}, {2, 0, 1}, {2, 0, 1}, {0, 1, 2}, {0, 2, 1}, {2, 1, 0}};
static const unsigned char HFLIP[8][3] = {
  {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {0, 1, 1}, {0, 0, 0}, {0, 1, 1}, {1, 0, 1}};

/* 21 digits of 3 bits from the top of each 32-bit coordinate.  The reference
 * writes digit bit `off` (3..65, MSB first) into a two-word key; the bit at
 * offset 65 falls outside the two words and is lost, so the stored 64-bit
 * key is the 63-bit Hilbert index shifted right by one. */
void orc_hsfc3d_bits(const uint32_t coord[3], uint32_t key[2]) {
  unsigned char ax[3] = {0 << 1, 1 << 1, 2 << 1};
  uint64_t hidx = 0;
  for (int i = 1; i <= 21; i++) {
    unsigned s = 32 - i;
    unsigned idx = 0;
    for (int n = 0; n < 3; n++)
      idx |= (((coord[ax[n] >> 1] >> s) ^ ax[n]) & 1u) << n;
    unsigned c = GRAY_INV[idx];
    hidx = (hidx << 3) | c;
    unsigned char nx[3];
    for (int n = 0; n < 3; n++) nx[n] = ax[HSRC[c][n]] ^ HFLIP[c][n];
    ax[0] = nx[0]; ax[1] = nx[1]; ax[2] = nx[2];
  }
  uint64_t k64 = hidx >> 1;
  key[0] = (uint32_t)(k64 >> 32);
  key[1] = (uint32_t)(k64 & 0xffffffffu);
}

void orc_thsfc3d(const double normc[3], uint32_t add_step, uint32_t key[3]) {
  const double imax = 4294967295.0;   /* IScale, thashtab.cc:24 */
  uint32_t c[3];
  for (int k = 0; k < 3; k++) {
    double ck = (normc[k] <= 0) ? 0 : imax * normc[k];
    c[k] = (uint32_t)((imax < ck) ? imax : ck);
  }
  uint32_t kk[2];
  orc_hsfc3d_bits(c, kk);
  key[0] = kk[0];
  key[1] = kk[1];
  key[2] = add_step;
}

/* THashTable::hash (thashtab.cc:209-222; 1D: key[0]). */
int64_t orc_slot(const uint32_t *key, int table_size, int dim) {
  if (dim == 1) return (int64_t)key[0];
  const double umax = 4294967295.0;
  int S02 = table_size / 10;
  int S01 = table_size - S02;
  int i1 = (int)((double)key[0] / umax * S01);
  int i2 = (int)((double)key[1] / umax * S02);
  return (int64_t)i1 + i2;
}

/* ===================================================================== */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), used for the seeded  */
/* lattice jitter of the synthetic plume inputs.                          */
/* ===================================================================== */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ===================================================================== */
/* Meteo table: linear interpolation in km (meteo.cc:64-143).             */
/* ===================================================================== */
int orc_meteo_load(const char *path, int rows, int cols, orc_meteo *met) {
  FILE *f = fopen(path, "r");
  if (!f) ORC_FAIL(PSPH_EIO, "cannot open meteo file %s", path);
  double *z = malloc(sizeof(double) * rows);
  double *d = malloc(sizeof(double) * rows * (cols - 1));
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) {
      double v;
      if (fscanf(f, "%lf", &v) != 1) { fclose(f); free(z); free(d); ORC_FAIL(PSPH_EIO, "short meteo file"); }
      if (c == 0) z[r] = v; else d[(c - 1) * rows + r] = v;
    }
  fclose(f);
  met->rows = rows; met->nprop = cols - 1; met->z = z; met->data = d;
  return 0;
}

void orc_meteo_from_arrays(orc_meteo *met, int rows, int nprop, const double *z, const double *data) {
  met->rows = rows; met->nprop = nprop;
  met->z = malloc(sizeof(double) * rows);
  met->data = malloc(sizeof(double) * rows * nprop);
  memcpy(met->z, z, sizeof(double) * rows);
  memcpy(met->data, data, sizeof(double) * rows * nprop);
}

void orc_meteo_free(orc_meteo *met) { free(met->z); free(met->data); met->z = met->data = NULL; }

double orc_meteo_interp(const orc_meteo *met, double z0, int prop) {
  const double *z = met->z, *d = met->data + (size_t)prop * met->rows;
  int R = met->rows;
  for (int r = 0; r < R; r++)
    if (z[r] == z0) return d[r];                      /* sample hit */
  if (z0 < z[0]) return d[0];
  if (z0 > z[R - 1]) return d[R - 1];
  int ui = 0;                                         /* first z > z0 */
  while (ui < R && !(z[ui] > z0)) ui++;
  int li = ui - 1;
  return d[li] + (z0 - z[li]) / (z[ui] - z[li]) * (d[ui] - d[li]);
}

void orc_meteo_interp_all(const orc_meteo *met, double z0, double *out) {
  for (int p = 0; p < met->nprop; p++) out[p] = orc_meteo_interp(met, z0, p);
}

/* ===================================================================== */
/* State                                                                  */
/* ===================================================================== */
#define ORC_DARRAYS(st) \
  X(st->x[0]) X(st->x[1]) X(st->x[2]) X(st->v[0]) X(st->v[1]) X(st->v[2]) \
  X(st->sv[0]) X(st->sv[1]) X(st->sv[2]) X(st->rho) X(st->e) X(st->m) X(st->h) \
  X(st->p) X(st->c) X(st->T) X(st->gam) X(st->cp) X(st->xi)
#define ORC_UARRAYS(st) Y(st->key[0]) Y(st->key[1]) Y(st->key[2])
#define ORC_IARRAYS(st) Z(st->bc) Z(st->phase) Z(st->inv)

int orc_state_reserve(orc_state *st, size_t cap) {
  if (cap <= st->cap) return 0;
#define X(a) { void *q = realloc(a, cap * sizeof(double)); if (!q) ORC_FAIL(PSPH_ENOMEM, "oom"); a = q; }
#define Y(a) { void *q = realloc(a, cap * sizeof(uint32_t)); if (!q) ORC_FAIL(PSPH_ENOMEM, "oom"); a = q; }
#define Z(a) { void *q = realloc(a, cap * sizeof(int32_t)); if (!q) ORC_FAIL(PSPH_ENOMEM, "oom"); a = q; }
  ORC_DARRAYS(st) ORC_UARRAYS(st) ORC_IARRAYS(st)
#undef X
#undef Y
#undef Z
  st->cap = cap;
  return 0;
}

int orc_state_init(orc_state *st, size_t cap) {
  memset(st, 0, sizeof *st);
  return orc_state_reserve(st, cap < 16 ? 16 : cap);
}

void orc_state_free(orc_state *st) {
#define X(a) free(a);
#define Y(a) free(a);
#define Z(a) free(a);
  ORC_DARRAYS(st) ORC_UARRAYS(st) ORC_IARRAYS(st)
#undef X
#undef Y
#undef Z
  free(st->nb_off); free(st->nb_idx);
  free(st->bval); free(st->berupt);
  orc_meteo_free(&st->met);
  memset(st, 0, sizeof *st);
}

static void orc_zero_particle(orc_state *st, size_t i) {
  for (int k = 0; k < 3; k++) { st->x[k][i] = st->v[k][i] = st->sv[k][i] = 0; st->key[k][i] = 0; }
  st->rho[i] = st->e[i] = st->m[i] = st->h[i] = st->p[i] = st->c[i] = 0;
  st->T[i] = st->gam[i] = st->cp[i] = st->xi[i] = 0;
  st->bc[i] = st->phase[i] = st->inv[i] = 0;
}

static size_t orc_push(orc_state *st) {
  if (st->n == st->cap) orc_state_reserve(st, st->cap * 2);
  size_t i = st->n++;
  orc_zero_particle(st, i);
  return i;
}

int orc_state_load(orc_state *st, const psph_soa *s, size_t n) {
  int rc = orc_state_reserve(st, n < 16 ? 16 : n);
  if (rc) return rc;
  st->n = n;
  memcpy(st->x[0], s->x, n * 8); memcpy(st->x[1], s->y, n * 8); memcpy(st->x[2], s->z, n * 8);
  memcpy(st->v[0], s->vx, n * 8); memcpy(st->v[1], s->vy, n * 8); memcpy(st->v[2], s->vz, n * 8);
  memcpy(st->sv[0], s->svx, n * 8); memcpy(st->sv[1], s->svy, n * 8); memcpy(st->sv[2], s->svz, n * 8);
  memcpy(st->rho, s->rho, n * 8); memcpy(st->e, s->e, n * 8); memcpy(st->m, s->m, n * 8);
  memcpy(st->h, s->h, n * 8); memcpy(st->p, s->p, n * 8); memcpy(st->c, s->c, n * 8);
  memcpy(st->T, s->T, n * 8); memcpy(st->gam, s->gam, n * 8); memcpy(st->cp, s->cp, n * 8);
  memcpy(st->xi, s->xi, n * 8);
  memcpy(st->key[0], s->key0, n * 4); memcpy(st->key[1], s->key1, n * 4); memcpy(st->key[2], s->key2, n * 4);
  memcpy(st->bc, s->bc, n * 4); memcpy(st->phase, s->phase, n * 4); memcpy(st->inv, s->involved, n * 4);
  return 0;
}

int orc_state_store(const orc_state *st, psph_soa *s) {
  size_t n = st->n;
  memcpy(s->x, st->x[0], n * 8); memcpy(s->y, st->x[1], n * 8); memcpy(s->z, st->x[2], n * 8);
  memcpy(s->vx, st->v[0], n * 8); memcpy(s->vy, st->v[1], n * 8); memcpy(s->vz, st->v[2], n * 8);
  memcpy(s->svx, st->sv[0], n * 8); memcpy(s->svy, st->sv[1], n * 8); memcpy(s->svz, st->sv[2], n * 8);
  memcpy(s->rho, st->rho, n * 8); memcpy(s->e, st->e, n * 8); memcpy(s->m, st->m, n * 8);
  memcpy(s->h, st->h, n * 8); memcpy(s->p, st->p, n * 8); memcpy(s->c, st->c, n * 8);
  memcpy(s->T, st->T, n * 8); memcpy(s->gam, st->gam, n * 8); memcpy(s->cp, st->cp, n * 8);
  memcpy(s->xi, st->xi, n * 8);
  memcpy(s->key0, st->key[0], n * 4); memcpy(s->key1, st->key[1], n * 4); memcpy(s->key2, st->key[2], n * 4);
  memcpy(s->bc, st->bc, n * 4); memcpy(s->phase, st->phase, n * 4); memcpy(s->involved, st->inv, n * 4);
  return 0;
}

/* Canonical order = THTIterator walk (thashtab.cc:229-262): ascending slot,
 * chains kept sorted by (k0, k1, k2) on insert (thashtab.cc:86-134). */
typedef struct { int64_t slot; uint32_t k0, k1, k2; uint32_t idx; } orc_sortrec;
static int orc_cmp_rec(const void *a, const void *b) {
  const orc_sortrec *x = a, *y = b;
  if (x->slot != y->slot) return x->slot < y->slot ? -1 : 1;
  if (x->k0 != y->k0) return x->k0 < y->k0 ? -1 : 1;
  if (x->k1 != y->k1) return x->k1 < y->k1 ? -1 : 1;
  if (x->k2 != y->k2) return x->k2 < y->k2 ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static void orc_permute(orc_state *st, const uint32_t *perm) {
  size_t n = st->n;
  double *td = malloc(n * sizeof(double));
  uint32_t *tu = malloc(n * sizeof(uint32_t));
#define X(a) { for (size_t i = 0; i < n; i++) td[i] = a[perm[i]]; memcpy(a, td, n * sizeof(double)); }
#define Y(a) { for (size_t i = 0; i < n; i++) tu[i] = a[perm[i]]; memcpy(a, tu, n * sizeof(uint32_t)); }
#define Z(a) { for (size_t i = 0; i < n; i++) tu[i] = (uint32_t)a[perm[i]]; memcpy(a, tu, n * sizeof(int32_t)); }
  ORC_DARRAYS(st) ORC_UARRAYS(st) ORC_IARRAYS(st)
#undef X
#undef Y
#undef Z
  free(td); free(tu);
}

void orc_sort_canonical(const psph_config *cfg, orc_state *st) {
  size_t n = st->n;
  orc_sortrec *r = malloc(n * sizeof *r);
  for (size_t i = 0; i < n; i++) {
    uint32_t k[3] = {st->key[0][i], st->key[1][i], st->key[2][i]};
    r[i].slot = orc_slot(k, cfg->p_table_size, cfg->dim);
    r[i].k0 = k[0]; r[i].k1 = k[1]; r[i].k2 = k[2]; r[i].idx = (uint32_t)i;
  }
  qsort(r, n, sizeof *r, orc_cmp_rec);
  uint32_t *perm = malloc(n * sizeof(uint32_t));
  for (size_t i = 0; i < n; i++) perm[i] = r[i].idx;
  orc_permute(st, perm);
  free(perm); free(r);
}

/* ===================================================================== */
/* Particle masks (particle.h:555-612)                                    */
/* ===================================================================== */
static inline int need_neigh(int bc) { return bc == 100 || bc == 99; }   /* no guests at one rank */
static inline int is_real(int bc) { return bc == 100 || bc == 99; }
static inline int contr_dens(const psph_config *cfg, int bc) {
  if (cfg->pghost_contrib == 1) return bc == 100 || bc == 99 || bc == 1;
  if (cfg->pghost_contrib == 2) return bc == 100 || bc == 99 || bc == 1 || bc == 0;
  return bc == 100 || bc == 99;
}
static inline int contr_vel_smooth(int bc) { return bc == 100 || bc == 99 || bc == 2; }

/* ===================================================================== */
/* Neighbour lists: search_neighs_consth (neighs.cc:37-211).              */
/* list(i) = every particle of i's bucket (i included), then, bucket by    */
/* bucket in neighbour-index order, each particle j of the 26 adjacent     */
/* buckets with |x_i - x_j|_k <= 4.25 h_i on every axis.  Buckets are the  */
/* preprocessor's cells of edge PARTICLE_DENSITY*h_air anchored at the     */
/* extended-domain minimum (preprocess.cc:88-133).  1D: list = all         */
/* particles (add_air.cc:471-479).                                         */
/* ===================================================================== */
static int orc_bucket_of(const psph_config *cfg, const orc_state *st, size_t i, int b[3]) {
  int nb[3];
  for (int k = 0; k < 3; k++) {
    nb[k] = (int)llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / cfg->bucket_edge);
    double f = floor((st->x[k][i] - cfg->ext_lo[k]) / cfg->bucket_edge);
    if (!(f >= 0 && f < nb[k])) return -1;
    b[k] = (int)f;
  }
  return 0;
}

static int orc_ensure_nb(orc_state *st, size_t total) {
  if (total > st->nb_cap) {
    void *q = realloc(st->nb_idx, total * sizeof(uint32_t));
    if (!q) ORC_FAIL(PSPH_ENOMEM, "oom (neighbour lists: %zu entries)", total);
    st->nb_idx = q; st->nb_cap = total;
  }
  return 0;
}

int orc_neighbors(const psph_config *cfg, orc_state *st) {
  size_t n = st->n;
  free(st->nb_off);
  st->nb_off = calloc(n + 1, sizeof(int64_t));
  if (cfg->list_all || cfg->dim == 1) {
    size_t nn = 0;
    for (size_t i = 0; i < n; i++) nn += need_neigh(st->bc[i]) ? n : 0;
    if (orc_ensure_nb(st, nn ? nn : 1)) return PSPH_ENOMEM;
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
      st->nb_off[i] = (int64_t)o;
      if (need_neigh(st->bc[i]))
        for (size_t j = 0; j < n; j++) st->nb_idx[o++] = (uint32_t)j;
    }
    st->nb_off[n] = (int64_t)o;
    return 0;
  }
  int nb[3];
  for (int k = 0; k < 3; k++) nb[k] = (int)llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / cfg->bucket_edge);
  size_t nbk = (size_t)nb[0] * nb[1] * nb[2];
  int64_t *bstart = calloc(nbk + 1, sizeof(int64_t));
  uint32_t *bmem = malloc((n ? n : 1) * sizeof(uint32_t));
  int64_t *blin = malloc((n ? n : 1) * sizeof(int64_t));
  for (size_t i = 0; i < n; i++) {
    int b[3];
    if (orc_bucket_of(cfg, st, i, b)) {
      free(bstart); free(bmem); free(blin);
      ORC_FAIL(PSPH_EINVARIANT, "particle %zu outside the bucket grid", i);
    }
    blin[i] = ((int64_t)b[2] * nb[1] + b[1]) * nb[0] + b[0];
    bstart[blin[i] + 1]++;
  }
  for (size_t b = 0; b < nbk; b++) bstart[b + 1] += bstart[b];
  {
    int64_t *fill = malloc((nbk + 1) * sizeof(int64_t));
    memcpy(fill, bstart, (nbk + 1) * sizeof(int64_t));
    for (size_t i = 0; i < n; i++) bmem[fill[blin[i]]++] = (uint32_t)i;   /* canonical order */
    free(fill);
  }
  /* pass 1: counts */
  int64_t *cnt = calloc(n + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    int b[3]; orc_bucket_of(cfg, st, i, b);
    double hi = cfg->list_factor * st->h[i];
    int64_t c = 0;
    for (int dx = -1; dx <= 1; dx++) for (int dy = -1; dy <= 1; dy++) for (int dz = -1; dz <= 1; dz++) {
      int q[3] = {b[0] + dx, b[1] + dy, b[2] + dz};
      if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= nb[0] || q[1] >= nb[1] || q[2] >= nb[2]) continue;
      int64_t L = ((int64_t)q[2] * nb[1] + q[1]) * nb[0] + q[0];
      int self = (dx == 0 && dy == 0 && dz == 0);
      for (int64_t t = bstart[L]; t < bstart[L + 1]; t++) {
        uint32_t j = bmem[t];
        if (self) { c++; continue; }
        double ds[3] = {st->x[0][i] - st->x[0][j], st->x[1][i] - st->x[1][j], st->x[2][i] - st->x[2][j]};
        if (in_support(ds, hi, 3)) c++;
      }
    }
    cnt[i + 1] = c;
  }
  for (size_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  if (orc_ensure_nb(st, cnt[n] ? (size_t)cnt[n] : 1)) { free(cnt); free(bstart); free(bmem); free(blin); return PSPH_ENOMEM; }
  /* pass 2: fill (own bucket first, then neighbour buckets in index order) */
#pragma omp parallel for schedule(dynamic, 256)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    int b[3]; orc_bucket_of(cfg, st, i, b);
    double hi = cfg->list_factor * st->h[i];
    int64_t o = cnt[i];
    int64_t Lself = ((int64_t)b[2] * nb[1] + b[1]) * nb[0] + b[0];
    for (int64_t t = bstart[Lself]; t < bstart[Lself + 1]; t++) st->nb_idx[o++] = bmem[t];
    for (int dx = -1; dx <= 1; dx++) for (int dy = -1; dy <= 1; dy++) for (int dz = -1; dz <= 1; dz++) {
      if (dx == 0 && dy == 0 && dz == 0) continue;
      int q[3] = {b[0] + dx, b[1] + dy, b[2] + dz};
      if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= nb[0] || q[1] >= nb[1] || q[2] >= nb[2]) continue;
      int64_t L = ((int64_t)q[2] * nb[1] + q[1]) * nb[0] + q[0];
      for (int64_t t = bstart[L]; t < bstart[L + 1]; t++) {
        uint32_t j = bmem[t];
        double ds[3] = {st->x[0][i] - st->x[0][j], st->x[1][i] - st->x[1][j], st->x[2][i] - st->x[2][j]};
        if (in_support(ds, hi, 3)) st->nb_idx[o++] = j;
      }
    }
  }
  memcpy(st->nb_off, cnt, (n + 1) * sizeof(int64_t));
  free(cnt); free(bstart); free(bmem); free(blin);
  return 0;
}

/* The same membership rule evaluated pair by pair over all particles
 * (O(N^2)); list order is ascending index.  For tests on small inputs. */
int orc_neighbors_bruteforce(const psph_config *cfg, orc_state *st) {
  if (cfg->list_all || cfg->dim == 1) return orc_neighbors(cfg, st);
  size_t n = st->n;
  int *bb = malloc((n ? n : 1) * 3 * sizeof(int));
  for (size_t i = 0; i < n; i++)
    if (orc_bucket_of(cfg, st, i, bb + 3 * i)) { free(bb); ORC_FAIL(PSPH_EINVARIANT, "outside grid"); }
  free(st->nb_off);
  st->nb_off = calloc(n + 1, sizeof(int64_t));
  size_t o = 0;
  for (size_t i = 0; i < n; i++) {
    st->nb_off[i] = (int64_t)o;
    if (!need_neigh(st->bc[i])) continue;
    for (size_t j = 0; j < n; j++) {
      int same = 1, adj = 1;
      for (int k = 0; k < 3; k++) {
        int d = bb[3 * j + k] - bb[3 * i + k];
        if (d != 0) same = 0;
        if (d < -1 || d > 1) adj = 0;
      }
      int in = same;
      if (!in && adj) {
        double ds[3] = {st->x[0][i] - st->x[0][j], st->x[1][i] - st->x[1][j], st->x[2][i] - st->x[2][j]};
        in = in_support(ds, cfg->list_factor * st->h[i], 3);
      }
      if (in) {
        if (orc_ensure_nb(st, o + 1 > st->nb_cap ? 2 * (o + 1) : o + 1)) { free(bb); return PSPH_ENOMEM; }
        st->nb_idx[o++] = (uint32_t)j;
      }
    }
  }
  st->nb_off[n] = (int64_t)o;
  free(bb);
  return 0;
}

/* ===================================================================== */
/* Adaptive smoothing length: adaptive_sml, ADAPTIVE_SML==1               */
/* (neighs.cc:479-693).  Uses the list of the current step.               */
/* ===================================================================== */
int orc_adaptive_sml(const psph_config *cfg, orc_state *st) {
  const int D = cfg->dim;
  size_t n = st->n;
#pragma omp parallel for schedule(dynamic, 128)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    double xi[3] = {st->x[0][i], st->x[1][i], st->x[2][i]};
    for (int loop = 0; loop < cfg->num_loop; loop++) {
      double H_old = st->h[i];
      double H_star = H_old * cfg->c_smooth;
      double supp = cfg->cutoff2 * H_star;
      double rho_star = 0.;
      for (int64_t t = st->nb_off[i]; t < st->nb_off[i + 1]; t++) {
        uint32_t j = st->nb_idx[t];
        if (!contr_dens(cfg, st->bc[j])) continue;
        double dx[3], si[3];
        for (int k = 0; k < D; k++) { dx[k] = xi[k] - st->x[k][j]; si[k] = dx[k] / H_star; }
        if (in_support(dx, supp, D)) rho_star += orc_weight(si, H_star, D) * st->m[j];
      }
      double H_new = (D == 3) ? cfg->eta * cbrt(st->m[i] / rho_star) : cfg->eta * (st->m[i] / rho_star);
      st->h[i] = cfg->ratio_sml_dx * H_new;
      double err = fabs((H_new - H_old) / H_old);
      if (err < cfg->thresh) break;
    }
  }
  return 0;
}

/* ===================================================================== */
/* Momentum / energy RHS, standard SPH (mom_engr_update.cc:30-439) with    */
/* art_vis (sph_lib.cc:402-424), SPH_epsilon_mom (:1331-1338),             */
/* SPH_epsilon_heat_conductivity 6-arg (:1399-1439), compute_F (:257-281). */
/* ===================================================================== */
int orc_rhs(const psph_config *cfg, orc_state *st, double dt) {
  const int D = cfg->dim;
  size_t n = st->n;
  double *nv[3], *ne;
  for (int k = 0; k < 3; k++) nv[k] = malloc((n ? n : 1) * sizeof(double));
  ne = malloc((n ? n : 1) * sizeof(double));
  int bad = 0;
  int64_t pairs = 0;
  const double gz = -cfg->g;
#pragma omp parallel for schedule(dynamic, 128) reduction(|:bad) reduction(+:pairs)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    double xi[3] = {st->x[0][i], st->x[1][i], st->x[2][i]};
    double hi = st->h[i];
    double supp = cfg->cutoff2 * hi;
    double pressi = st->p[i], tempi = st->T[i], Cp_i = st->cp[i];
    if (!(st->rho[i] > 0)) { bad = 1; continue; }
    double Vi = 1.0 / st->rho[i];
    double pvsqi = pressi * Vi * Vi;
    double veli[3];
    for (int k = 0; k < 3; k++) veli[k] = cfg->lans ? st->sv[k][i] : st->v[k][i];
    double rhs_v[3] = {0, 0, 0}, rhs_e = 0;
    for (int64_t t = st->nb_off[i]; t < st->nb_off[i + 1]; t++) {
      uint32_t j = st->nb_idx[t];
      if (j == i) continue;                              /* dW(0)=0 */
      double h = 0.5 * (hi + st->h[j]);
      double dx[3], si[3], dist_sq = 0;
      for (int k = 0; k < D; k++) { dx[k] = xi[k] - st->x[k][j]; si[k] = dx[k] / h; dist_sq += dx[k] * dx[k]; }
      if (!in_support(dx, supp, D)) continue;
      double mj = st->m[j], pressj = st->p[j];
      if (!(st->rho[j] > 0)) { bad = 1; continue; }
      double Vj = 1.0 / st->rho[j];
      double velij[3];
      for (int k = 0; k < D; k++) velij[k] = veli[k] - (cfg->lans ? st->sv[k][j] : st->v[k][j]);
      double rhoab = 0.5 * (st->rho[i] + st->rho[j]);
      double sndspdab = 0.5 * (st->c[i] + st->c[j]);
      /* art_vis (sph_lib.cc:402-424) */
      double vrab = 0.;
      for (int k = 0; k < D; k++) vrab += dx[k] * velij[k];
      double miuab = h * vrab / (dist_sq + cfg->ata * h * h);
      double vis = 0.;
      if (vrab > 0) vis = 0.;
      else if (vrab < 0) vis = (cfg->beta * miuab - cfg->alpha * sndspdab) * miuab / rhoab;
      if (cfg->lans) {                                   /* SPH_epsilon_mom */
        double dotv = 0.;
        for (int k = 0; k < D; k++) dotv += velij[k] * velij[k];
        vis -= cfg->eps_half * dotv * Vj;
      }
      double mpvsqij = mj * (pressj * Vj * Vj + pvsqi + vis);
      double s_heat[3], dwdx[3], dwdx_heat[3];
      for (int k = 0; k < D; k++) s_heat[k] = si[k] * cfg->heat_scale;
      int nz = 0;
      for (int k = 0; k < D; k++) {
        dwdx[k] = orc_d_weight(si, h, k, D);
        dwdx_heat[k] = orc_d_weight(s_heat, h / cfg->heat_scale, k, D);
        nz |= (dwdx[k] != 0);
      }
      {
        int insup = 1;
        for (int k = 0; k < D; k++) if (fabs(si[k]) > ORC_CUTOFF) insup = 0;
        pairs += insup;
      }
      (void)nz;
      double heat_tran = 0.;
      if (cfg->lans) {
        double Cp_ij = 0.5 * (st->cp[j] + Cp_i);
        double hab = 0.5 * (hi + st->h[j]);
        /* SPH_epsilon_heat_conductivity, 6-argument form */
        double dotvv = 0., dotvr = 0., dotrr = 0., kij;
        for (int k = 0; k < D; k++) dotvr += fabs(velij[k] * dx[k]);
        if (dotvr == 0.) kij = 0.;
        else {
          for (int k = 0; k < D; k++) dotrr += dx[k] * dx[k];
          for (int k = 0; k < D; k++) dotvv += velij[k] * velij[k];
          double kab;
          if (D == 3) kab = 10.0 * cfg->eps_half * Cp_ij * dotrr * dotvv / (cfg->prandtl * dotvr);
          else kab = 0.0;
          kij = kab / (hab * sndspdab);
        }
        /* compute_F */
        double W_r = 0., r_sq = 0., Fij;
        for (int k = 0; k < D; k++) W_r += dwdx_heat[k] * dx[k];
        if (W_r == 0) Fij = 0.0;
        else {
          for (int k = 0; k < D; k++) r_sq += dx[k] * dx[k];
          Fij = W_r / r_sq;
        }
        heat_tran = mj * Vj * Vi * kij * (tempi - st->T[j]) * Fij;
      }
      for (int k = 0; k < D; k++) rhs_v[k] -= mpvsqij * dwdx[k];
      double deltae = 0.;
      for (int k = 0; k < D; k++) deltae += 0.5 * (mpvsqij * dwdx[k]) * velij[k];
      rhs_e += (deltae + heat_tran);
    }
    for (int k = 0; k < D; k++) nv[k][i] = st->v[k][i] + dt * (rhs_v[k]);
    for (int k = D; k < 3; k++) nv[k][i] = st->v[k][i];
    nv[D - 1][i] += dt * gz;
    ne[i] = st->e[i] + dt * (rhs_e + veli[D - 1] * gz);
  }
  if (!bad) {
    /* second sweep: update_state_vars + EOS (mom_engr_update.cc:412-432) */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
      if (!need_neigh(st->bc[i])) continue;
      for (int k = 0; k < 3; k++) st->v[k][i] = nv[k][i];
      st->e[i] = ne[i];
      orc_eos(cfg, st->rho[i], st->e[i], st->xi[i], &st->p[i], &st->c[i], &st->T[i], &st->gam[i], &st->cp[i]);
    }
  }
  for (int k = 0; k < 3; k++) free(nv[k]);
  free(ne);
  st->last_pairs = pairs;
  if (bad) ORC_FAIL(PSPH_EINVARIANT, "rhs: non-positive density");
  return 0;
}

/* ===================================================================== */
/* Density / mass fraction / energy smoothing (density_massfrac.cc:25-359) */
/* DENSITY_UPDATE_SML 11, DENSITY_UPDATE_SPH 11, HAVE_ENERGY_SMOOTH 0|1.   */
/* ===================================================================== */
int orc_density(const psph_config *cfg, orc_state *st) {
  const int D = cfg->dim;
  size_t n = st->n;
  double *nrho = malloc((n ? n : 1) * sizeof(double));
  double *nxi = malloc((n ? n : 1) * sizeof(double));
  double *ne = malloc((n ? n : 1) * sizeof(double));
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 128) reduction(|:bad)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    double xi[3] = {st->x[0][i], st->x[1][i], st->x[2][i]};
    double hi0 = st->h[i];
    double supp = cfg->cutoff2 * hi0;
    double tmprho[2] = {0., 0.};
    double wnorm_e = 0.0, engr = 0.0;
    for (int64_t t = st->nb_off[i]; t < st->nb_off[i + 1]; t++) {
      uint32_t j = st->nb_idx[t];
      if (!contr_dens(cfg, st->bc[j])) continue;
      double mj = st->m[j], rhoj = st->rho[j];
      int ph = st->phase[j];
      if (ph != 1 && ph != 2) continue;
      double ds[3], s[3];
      for (int k = 0; k < D; k++) ds[k] = xi[k] - st->x[k][j];
      if (!in_support(ds, supp, D)) continue;
      double hij = 0.5 * (st->h[i] + st->h[j]);
      for (int k = 0; k < D; k++) s[k] = ds[k] / hij;
      if (cfg->energy_smooth) {
        double s_e[3];
        for (int k = 0; k < D; k++) s_e[k] = cfg->e_smooth_ratio * ds[k] / hij;
        double wm_e = orc_weight(s_e, hij / cfg->e_smooth_ratio, D) * mj;
        wnorm_e += wm_e / rhoj;
        engr += wm_e * st->e[j] / rhoj;
      }
      double wght = orc_weight(s, hij, D);
      tmprho[ph - 1] += wght * mj;
    }
    if (cfg->energy_smooth) {
      if (!(wnorm_e > 0)) { bad = 1; continue; }
      ne[i] = engr / wnorm_e;
    }
    double density = 0.0;
    density += tmprho[0];
    density += tmprho[1];
    double mssfrc = tmprho[1] / density;
    if (!(density > 0) || !(mssfrc <= 1)) { bad = 1; continue; }
    nrho[i] = density;
    nxi[i] = mssfrc;
  }
  if (!bad) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
      if (!need_neigh(st->bc[i])) continue;
      st->xi[i] = nxi[i];
      if (st->xi[i] >= cfg->msfrc_thresh && st->inv[i] != 2) st->inv[i] = 2;
      st->rho[i] = nrho[i];
      if (cfg->energy_smooth) st->e[i] = ne[i];
      orc_eos(cfg, st->rho[i], st->e[i], st->xi[i], &st->p[i], &st->c[i], &st->T[i], &st->gam[i], &st->cp[i]);
    }
  }
  free(nrho); free(nxi); free(ne);
  if (bad) ORC_FAIL(PSPH_EINVARIANT, "density: density<=0, mass fraction>1 or wnorm_e<=0");
  return 0;
}

/* ===================================================================== */
/* LANS velocity filter (smooth_velocity.cc:25-165)                       */
/* ===================================================================== */
int orc_vsmooth(const psph_config *cfg, orc_state *st) {
  const int D = cfg->dim;
  size_t n = st->n;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 128) reduction(|:bad)
  for (size_t i = 0; i < n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    double xi[3] = {st->x[0][i], st->x[1][i], st->x[2][i]};
    double vi[3] = {st->v[0][i], st->v[1][i], st->v[2][i]};
    double hi = st->h[i];
    double supp = cfg->cutoff2 * hi;
    double dv[3] = {0., 0., 0.}, wnorm = 0.;
    for (int64_t t = st->nb_off[i]; t < st->nb_off[i + 1]; t++) {
      uint32_t j = st->nb_idx[t];
      if (!contr_vel_smooth(st->bc[j])) continue;
      double ds[3], s[3];
      for (int k = 0; k < D; k++) ds[k] = xi[k] - st->x[k][j];
      if (!in_support(ds, supp, D)) continue;
      double rhoj = st->rho[j], mj = st->m[j];
      for (int k = 0; k < D; k++) s[k] = ds[k] / hi;
      double wght = orc_weight(s, hi, D);
      double mvw = wght * mj / rhoj;
      for (int k = 0; k < D; k++) dv[k] += mvw * (st->v[k][j] - vi[k]);
      wnorm += mvw;
    }
    if (!(wnorm > 0)) { bad = 1; continue; }
    for (int k = 0; k < D; k++) dv[k] /= wnorm;
    for (int k = 0; k < D; k++) st->sv[k][i] = vi[k] + dv[k];
  }
  if (bad) ORC_FAIL(PSPH_EINVARIANT, "vsmooth: wnorm<=0");
  return 0;
}

/* ===================================================================== */
/* Time step and clock (timestep.cc:94-183, properties.h:71-135)          */
/* ===================================================================== */
double orc_timestep(const psph_config *cfg, const orc_state *st) {
  double dt = 1.0E+10;
  for (size_t i = 0; i < st->n; i++) {
    if (!need_neigh(st->bc[i])) continue;
    double temp = st->h[i] / st->c[i];
    if (temp < dt) dt = temp;
  }
  return cfg->cfl * dt;
}

void orc_incrtime(const psph_config *cfg, orc_state *st, double *dt) {
  if ((st->time + (*dt)) > st->next_output) (*dt) = st->next_output - st->time;
  if ((st->time + (*dt)) > cfg->t_end) (*dt) = cfg->t_end - st->time;
  st->time += (*dt);
  st->step++;
}

static void orc_ifoutput(const psph_config *cfg, orc_state *st) {   /* properties.h:104-120 */
  if (st->time >= st->next_output || st->time == cfg->t_end) {
    st->ioutput++;
    st->next_output = ((st->ioutput + 1) * cfg->t_output) / 1.0;
  }
}

int orc_iferupt(const psph_config *cfg, const orc_state *st) {
  if (st->time < cfg->erupt_start) return 0;
  if (st->time > cfg->erupt_end) return 0;
  return 1;
}

/* ===================================================================== */
/* Position update (update_pos.cc:34-168) and escape removal of real       */
/* particles that leave the domain through a boundary bucket               */
/* (Bucket::determine_escape, bucket.cc:435-452; update_pos.cc:171-463).  */
/* ===================================================================== */
int orc_update_pos(const psph_config *cfg, orc_state *st, double dt) {
  const int D = cfg->dim;
  int erupting = orc_iferupt(cfg, st);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < st->n; i++) {
    int bc = st->bc[i];
    if (is_real(bc)) {
      for (int k = 0; k < D; k++) {
        double vel = cfg->lans ? st->sv[k][i] : st->v[k][i];
        st->x[k][i] = st->x[k][i] + dt * vel;
      }
    } else if (bc == 0 && erupting) {
      for (int k = 0; k < D; k++) st->x[k][i] = st->x[k][i] + dt * st->v[k][i];
      if (D == 3 && st->x[2][i] >= cfg->dom_lo[2]) {
        st->bc[i] = 100;                                 /* erupt_turn_real */
        st->h[i] = cfg->ratio_sml_dx * cfg->h_air;       /* put_smlen(sml_of_phase2) */
        st->inv[i] = 2;
      }
    }
  }
  return 0;
}

int orc_remove_escaped(const psph_config *cfg, orc_state *st) {
  const double TINY = 1.0E-08;
  size_t w = 0, removed = 0;
  for (size_t i = 0; i < st->n; i++) {
    int esc = 0;
    if (is_real(st->bc[i]))
      for (int k = 0; k < 3; k++)
        if ((st->x[k][i] - cfg->dom_lo[k]) < -TINY || (st->x[k][i] - cfg->dom_hi[k]) > TINY) esc = 1;
    if (esc) { removed++; continue; }
    if (w != i) {
#define X(a) a[w] = a[i];
#define Y(a) a[w] = a[i];
#define Z(a) a[w] = a[i];
      ORC_DARRAYS(st) ORC_UARRAYS(st) ORC_IARRAYS(st)
#undef X
#undef Y
#undef Z
    }
    w++;
  }
  st->n = w;
  st->last_removed = (int64_t)removed;
  return 0;
}

/* ===================================================================== */
/* Inlet: erupt-ghost particles (Particle ctor particle.cc:210-313)        */
/* ===================================================================== */
static double orc_vel_prof(const psph_config *cfg, double R, double rsq, double u_avg) {
  if (cfg->vent_profile == 0) return u_avg;                     /* uniform_vel */
  double umax = u_avg * 2;                                       /* parabolic_vel */
  return umax * (1 - (rsq / (R * R)));
}

static void orc_new_erupt(const psph_config *cfg, orc_state *st, const double crd[3],
                          const uint32_t key[3], double vel) {
  size_t i = orc_push(st);
  for (int k = 0; k < 3; k++) { st->x[k][i] = crd[k]; st->key[k][i] = key[k]; }
  st->m[i] = cfg->vent_mass;
  st->h[i] = cfg->ratio_sml_dx * cfg->vent_sml;
  st->bc[i] = 0; st->inv[i] = 0; st->phase[i] = 2; st->xi[i] = 1.;
  st->v[2][i] = vel;
  st->e[i] = cfg->vent_e;
  st->rho[i] = cfg->vent_rho;
  double pdummy, cdummy, gdummy;
  st->cp[i] = 0.;
  orc_eos(cfg, st->rho[i], st->e[i], 1., &pdummy, &cdummy, &st->T[i], &gdummy, &st->cp[i]);
  st->p[i] = cfg->vent_p;
  st->c[i] = cfg->vent_c;
  st->gam[i] = cfg->vent_gamma;
}

/* Scan order of the vent columns: x+ (y+, y-) then x- (y+, y-)
 * (setup_erupt.cc:105-205, 458-660). */
typedef void (*orc_col_fn)(const psph_config *, orc_state *, double x, double y, void *);
static void orc_vent_columns(const psph_config *cfg, orc_state *st, orc_col_fn fn, void *ud) {
  double sml = cfg->vent_sml, sml2 = 0.5 * sml;
  double rx0 = -cfg->vent_r, rx1 = cfg->vent_r, ry0 = -cfg->vent_r, ry1 = cfg->vent_r;
  for (int sx = 0; sx < 2; sx++) {
    double x = sx == 0 ? cfg->vent_x + sml2 : cfg->vent_x - sml2;
    while (sx == 0 ? x <= rx1 : x >= rx0) {
      for (int sy = 0; sy < 2; sy++) {
        double y = sy == 0 ? cfg->vent_y + sml2 : cfg->vent_y - sml2;
        while (sy == 0 ? y <= ry1 : y >= ry0) {
          fn(cfg, st, x, y, ud);
          y = sy == 0 ? y + sml : y - sml;
        }
      }
      x = sx == 0 ? x + sml : x - sml;
    }
  }
}

typedef struct { double dt, t_total, bot; uint32_t add_step; int64_t added; int dup; } orc_inject_ud;

static void orc_inject_col(const psph_config *cfg, orc_state *st, double x, double y, void *vud) {
  orc_inject_ud *u = vud;
  double rvsq = cfg->vent_r * cfg->vent_r;
  double dist = 0;
  dist += x * x;
  dist += y * y;
  if (!(dist <= rvsq)) return;
  double vel = orc_vel_prof(cfg, cfg->vent_r, dist, cfg->vent_v);
  double t_each = cfg->vent_sml / vel;
  double t_add = fmod(u->t_total, t_each) + u->dt;
  int n = (int)floor(t_add / t_each);
  t_add = t_add - n * t_each;
  for (int i = 0; i < n; i++) {
    double crd[3] = {x, y, u->bot + (t_add + i * t_each) * vel};
    double normc[3];
    for (int k = 0; k < 3; k++) normc[k] = (crd[k] - cfg->ext_lo[k]) / (cfg->ext_hi[k] - cfg->ext_lo[k]);
    uint32_t key[3];
    orc_thsfc3d(normc, u->add_step, key);
    orc_new_erupt(cfg, st, crd, key, vel);
    u->added++;
  }
}

int orc_inject(const psph_config *cfg, orc_state *st, double dt) {  /* setup_erupt.cc:389-670 */
  orc_inject_ud u = {dt, st->time - dt, cfg->vent_bot, (uint32_t)st->step, 0, 0};
  size_t n0 = st->n;
  orc_vent_columns(cfg, st, orc_inject_col, &u);
  st->last_injected = u.added;
  /* duplicate-key check (setup_erupt.cc:486-491 exits) */
  for (size_t i = n0; i < st->n; i++)
    for (size_t j = 0; j < n0; j++)
      if (st->key[0][i] == st->key[0][j] && st->key[1][i] == st->key[1][j] && st->key[2][i] == st->key[2][j])
        ORC_FAIL(PSPH_EINVARIANT, "inject: particle added twice on the same key");
  return 0;
}

typedef struct { double z; } orc_setup_ud;
static void orc_setup_col(const psph_config *cfg, orc_state *st, double x, double y, void *vud) {
  orc_setup_ud *u = vud;
  double rvsq = cfg->vent_r * cfg->vent_r;
  double dist = 0;
  dist += x * x;
  dist += y * y;
  if (!(dist <= rvsq)) return;
  double crd[3] = {x, y, u->z};
  double normc[3];
  for (int k = 0; k < 3; k++) normc[k] = (crd[k] - cfg->ext_lo[k]) / (cfg->ext_hi[k] - cfg->ext_lo[k]);
  uint32_t key[3];
  orc_thsfc3d(normc, 0, key);
  orc_new_erupt(cfg, st, crd, key, orc_vel_prof(cfg, cfg->vent_r, dist, cfg->vent_v));
}

int orc_setup_erupt(const psph_config *cfg, orc_state *st) {       /* setup_erupt.cc:43-386 */
  double rz1 = cfg->vent_z, rz0 = cfg->vent_bot;
  double rvsq = cfg->vent_r * cfg->vent_r;
  /* remove non-erupt particles inside the duct (setup_erupt.cc:302-320) */
  size_t w = 0;
  for (size_t i = 0; i < st->n; i++) {
    double dist = 0;
    dist += st->x[0][i] * st->x[0][i];
    dist += st->x[1][i] * st->x[1][i];
    int inside = dist <= rvsq && st->x[2][i] <= rz1 && st->x[2][i] >= rz0 && st->bc[i] != 0;
    if (inside) continue;
    if (w != i) {
#define X(a) a[w] = a[i];
#define Y(a) a[w] = a[i];
#define Z(a) a[w] = a[i];
      ORC_DARRAYS(st) ORC_UARRAYS(st) ORC_IARRAYS(st)
#undef X
#undef Y
#undef Z
    }
    w++;
  }
  st->n = w;
  orc_setup_ud u;
  u.z = rz1 - 0.5 * cfg->vent_sml;
  while (u.z >= rz0) {
    orc_vent_columns(cfg, st, orc_setup_col, &u);
    u.z -= cfg->vent_sml;
  }
  orc_sort_canonical(cfg, st);
  return 0;
}

/* ===================================================================== */
/* Synthetic plume input: air lattice (add_air.cc:490-813, initial_air     */
/* sph_lib.cc:831-872, air_prop_meteo_based :724-766) over the extended    */
/* domain: real air inside [dom_lo, dom_hi] above ground, pressure ghosts  */
/* (bc 1) in the outer shell above ground, wall ghosts (bc 2) below.       */
/* Optional seeded jitter of each lattice point (Philox4x32-10, keyed by   */
/* the global lattice index).                                              */
/* ===================================================================== */
static void orc_air_state(const psph_config *cfg, const orc_meteo *met, orc_state *st, size_t i) {
  double xi[3] = {st->x[0][i], st->x[1][i], st->x[2][i]};
  double sml2 = 0.5 * st->h[i];
  double range[6] = {xi[0] - sml2, xi[0] + sml2, xi[1] - sml2, xi[1] + sml2, xi[2] - sml2, xi[2] + sml2};
  double prop[16];
  orc_meteo_interp_all(met, xi[2] * 0.001, prop);
  double pressure = prop[1] * 100;
  double density = prop[0];
  double energy = pressure / (density * (cfg->gamma_p - 1));
  double d1 = orc_meteo_interp(met, range[4] * 0.001, 0);
  double d2 = orc_meteo_interp(met, range[5] * 0.001, 0);
  double d = density;
  double mass = (range[1] - range[0]) * (range[3] - range[2]) * (range[5] - range[4]) *
                (0.1666667 * d1 + 0.6666666 * d + 0.1666667 * d2);
  st->rho[i] = density;
  st->e[i] = energy;
  st->p[i] = pressure;
  for (int k = 0; k < 3; k++) st->v[k][i] = 0.;
  st->m[i] = mass;
  orc_eos(cfg, st->rho[i], st->e[i], st->xi[i], &st->p[i], &st->c[i], &st->T[i], &st->gam[i], &st->cp[i]);
}

/* Lattice point l of the air (spacing h_air, offset h_air/2 inside each
 * bucket): pc0 exact, pc with the seeded jitter of the synthetic inputs. */
static void orc_lattice_point(const psph_config *cfg, uint32_t seed, double jitter, const int64_t l[3],
                              const int64_t nl[3], double pc0[3], double pc[3]) {
  const double dx = cfg->h_air, dx2 = 0.5 * dx, del = cfg->bucket_edge;
  const int per = (int)llround(del / dx);
  for (int k = 0; k < 3; k++) {
    double mincrd = cfg->ext_lo[k] + (double)(l[k] / per) * del;   /* preprocess.cc:147-159 */
    pc0[k] = pc[k] = mincrd + dx2 + (double)(l[k] % per) * dx;      /* add_air.cc:575-577 */
  }
  if (jitter > 0) {
    uint64_t L = (uint64_t)((l[2] * nl[1] + l[1]) * nl[0] + l[0]);
    uint32_t ctr[4] = {(uint32_t)L, (uint32_t)(L >> 32), 0, 0}, key[2] = {seed, 0x50535048u}, r[4];
    orc_philox4x32_10(ctr, key, r);
    for (int k = 0; k < 3; k++) {
      double u = ((double)r[k] + 0.5) * (1.0 / 4294967296.0);
      pc[k] += (2.0 * u - 1.0) * jitter * dx;
    }
  }
}

int orc_air_lattice(const psph_config *cfg, const orc_meteo *met, uint32_t seed,
                    double jitter, orc_state *st) {
  double dx = cfg->h_air, del = cfg->bucket_edge;
  int nb[3], per = (int)llround(del / dx);
  for (int k = 0; k < 3; k++) nb[k] = (int)llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / del);
  int64_t nl[3] = {(int64_t)nb[0] * per, (int64_t)nb[1] * per, (int64_t)nb[2] * per};
  size_t total = (size_t)(nl[0] * nl[1] * nl[2]);
  if (orc_state_reserve(st, st->n + total)) return PSPH_ENOMEM;
  for (int64_t lz = 0; lz < nl[2]; lz++)
    for (int64_t ly = 0; ly < nl[1]; ly++)
      for (int64_t lx = 0; lx < nl[0]; lx++) {
        int64_t l[3] = {lx, ly, lz};
        double pc[3], pc0[3];
        orc_lattice_point(cfg, seed, jitter, l, nl, pc0, pc);
        int bc;
        if (!(pc[2] > cfg->dom_lo[2])) bc = 2;
        else if (pc[0] < cfg->dom_lo[0] || pc[0] > cfg->dom_hi[0] || pc[1] < cfg->dom_lo[1] ||
                 pc[1] > cfg->dom_hi[1] || pc[2] > cfg->dom_hi[2]) bc = 1;
        else bc = 100;
        size_t i = orc_push(st);
        for (int k = 0; k < 3; k++) st->x[k][i] = pc[k];
        double normc[3];
        for (int k = 0; k < 3; k++) normc[k] = (pc[k] - cfg->ext_lo[k]) / (cfg->ext_hi[k] - cfg->ext_lo[k]);
        uint32_t key[3];
        orc_thsfc3d(normc, 0, key);
        for (int k = 0; k < 3; k++) st->key[k][i] = key[k];
        st->h[i] = cfg->ratio_sml_dx * cfg->h_air;
        st->bc[i] = bc;
        st->phase[i] = 1;
        st->inv[i] = bc == 100 ? 1 : 0;
        st->xi[i] = 0.;
        st->cp[i] = 0.;
        orc_air_state(cfg, met, st, i);
      }
  orc_sort_canonical(cfg, st);
  return 0;
}

/* ===================================================================== */
/* 1D shock tube (add_air.cc:36-485, SHOCK_TUBE_SMOOTH 0)                  */
/* ===================================================================== */
static void orc_new_1d(const psph_config *cfg, orc_state *st, uint32_t id, double x, double m,
                       double h, double des, double vel, double prss, int bc, int inv) {
  size_t i = orc_push(st);
  st->key[0][i] = id; st->key[1][i] = 0; st->key[2][i] = 0;
  st->x[0][i] = x;
  st->m[i] = m;
  st->h[i] = cfg->ratio_sml_dx * h;
  st->bc[i] = bc; st->inv[i] = inv; st->phase[i] = 1;
  st->rho[i] = des;
  st->v[0][i] = vel;
  st->e[i] = prss / ((cfg->gamma_p - 1) * des);
  st->p[i] = prss;
  st->gam[i] = cfg->gamma_p;
  st->c[i] = 340.;
  st->xi[i] = 0.;
}

int orc_shock_tube(const psph_config *cfg, double des_l, double prss_l, double vel_l,
                   double des_r, double prss_r, double vel_r, double mid,
                   int n_nosave_l, int n_nosave_r, int n_ghost, int equal_mass, orc_state *st) {
  double dx_l = cfg->h_air;
  double dx_r = equal_mass ? dx_l * des_l / des_r : cfg->h_air;
  double smlen = dx_r > dx_l ? dx_r : dx_l;
  double dx2_l = 0.5 * dx_l;
  double bnd0 = cfg->dom_lo[0], bnd1 = cfg->dom_hi[0];
  uint32_t num = 0;
  double pcrd = 0;
#define RHO(x) ((x) > mid ? des_r : des_l)
#define PRS(x) ((x) > mid ? prss_r : prss_l)
#define VEL(x) ((x) > mid ? vel_r : vel_l)
  for (int k = 0; k < n_nosave_l; k++) {
    pcrd = bnd0 - (dx2_l + k * dx_l);
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_l, smlen, des, vel_l, prss_l, 99, 0);
  }
  for (int k = 0; k < n_ghost; k++) {
    pcrd -= dx_l;
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_l, dx_l * 1.1, des, vel_l, prss_l, 0, 0);
  }
  pcrd = bnd0 + dx2_l;
  while (pcrd <= mid) {
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_l, smlen, des, VEL(pcrd), PRS(pcrd), 100, 2);
    pcrd += dx_l;
  }
  pcrd = pcrd - dx_l + dx_r;
  while (pcrd <= bnd1) {
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_r, smlen, des, VEL(pcrd), PRS(pcrd), 100, 2);
    pcrd += dx_r;
  }
  for (int k = 0; k < n_nosave_r; k++) {
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_r, smlen, des, vel_r, prss_r, 99, 2);
    pcrd += dx_r;
  }
  for (int k = 0; k < n_ghost; k++) {
    double des = RHO(pcrd);
    orc_new_1d(cfg, st, num++, pcrd, des * dx_r, dx_r * 1.1, des, vel_r, prss_r, 0, 2);
    pcrd += dx_r;
  }
#undef RHO
#undef PRS
#undef VEL
  orc_sort_canonical(cfg, st);
  return 0;
}

/* ===================================================================== */
/* Conservation totals and the plume diagnostic (no reference counterpart) */
/* ===================================================================== */
void orc_conservation(const orc_state *st, double out[7]) {
  for (int k = 0; k < 7; k++) out[k] = 0;
  for (size_t i = 0; i < st->n; i++) {
    if (!is_real(st->bc[i])) continue;
    double m = st->m[i];
    double v2 = st->v[0][i] * st->v[0][i] + st->v[1][i] * st->v[1][i] + st->v[2][i] * st->v[2][i];
    out[0] += m;
    out[1] += m * st->e[i];
    out[2] += 0.5 * m * v2;
    out[3] += m * st->v[0][i];
    out[4] += m * st->v[1][i];
    out[5] += m * st->v[2][i];
    out[6] += m * st->xi[i];
  }
}

double orc_diag_height(const orc_state *st) {
  double num = 0, den = 0;
  for (size_t i = 0; i < st->n; i++) {
    if (!is_real(st->bc[i])) continue;
    double w = st->m[i] * st->xi[i];
    num += w * st->x[2][i];
    den += w;
  }
  return den > 0 ? num / den : 0.;
}

/* ===================================================================== */
/* One step in the reference order (main.cc:217-486 3D, 676-751 1D)        */
/* ===================================================================== */
/* ===================================================================== */
/* apply_bcond (bcond.cc:24-249) with the images of search_bnd_images     */
/* (bnd_image.cc:33-254) for a flat ground at z = dom_lo[2]: every wall   */
/* ghost of a bucket that straddles the ground (bucket type MIXED) takes  */
/* the state interpolated at its mirror point refc = (x, y, 2 z_g - z)    */
/* from the real and pressure-ghost particles j of the image's bucket and */
/* its 26 neighbours: Shepard sums of m_j W(s, h_j) (s = (refc - x_j)/h_j,*/
/* in_support box CUTOFF2 h_j) over rho, v / rho, e / rho, normalised by  */
/* sum m_j W / rho_j; the mass fraction sums only the neighbour buckets   */
/* (bcond.cc:140-146, BC_FOR_KX 1); the velocity is reflected in the      */
/* ground plane (sph_lib.cc:75-83) and the EOS re-evaluated.              */
/* ===================================================================== */
static int orc_contr_image(int bc) { return bc == 100 || bc == 99 || bc == 1; }   /* particle.h:562-565 */

int orc_apply_bcond(const psph_config *cfg, orc_state *st) {
  if (cfg->dim != 3) return 0;
  const double zg = cfg->dom_lo[2];
  const double fz = (zg - cfg->ext_lo[2]) / cfg->bucket_edge;
  if (fz == floor(fz)) return 0;                  /* ground on a bucket face: no MIXED bucket */
  const int bzg = (int)floor(fz);
  const size_t n = st->n;
  /* the new ghost states are computed from the old ones of the contributors only
     (bc 100, 99, 1), and only wall ghosts (bc 2) are written: ghosts are independent */
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t g = 0; g < n; g++) {
    if (st->bc[g] != 2) continue;
    int bg[3];
    if (orc_bucket_of(cfg, st, g, bg) || bg[2] != bzg) continue;
    double coord[3], refc[3], normal[3], d = 0;
    for (int k = 0; k < 3; k++) coord[k] = st->x[k][g];
    refc[0] = coord[0]; refc[1] = coord[1]; refc[2] = 2 * zg - coord[2];
    for (int k = 0; k < 3; k++) { normal[k] = refc[k] - coord[k]; d += normal[k] * normal[k]; }
    if (!(d > 0)) { normal[0] = normal[1] = 0; normal[2] = 1; d = 1; }
    for (int k = 0; k < 3; k++) normal[k] /= sqrt(d);
    /* bucket of the image */
    int bi[3], nb[3], ok = 1;
    for (int k = 0; k < 3; k++) {
      nb[k] = (int)llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / cfg->bucket_edge);
      double f = floor((refc[k] - cfg->ext_lo[k]) / cfg->bucket_edge);
      if (!(f >= 0 && f < nb[k])) ok = 0; else bi[k] = (int)f;
    }
    if (!ok) continue;
    double uvec[5] = {0, 0, 0, 0, 0}, wnorm = 0, ks = 0;
    for (size_t j = 0; j < n; j++) {
      if (!orc_contr_image(st->bc[j])) continue;
      int bj[3];
      if (orc_bucket_of(cfg, st, j, bj)) continue;
      if (abs(bj[0] - bi[0]) > 1 || abs(bj[1] - bi[1]) > 1 || abs(bj[2] - bi[2]) > 1) continue;
      const double h = st->h[j], supp = cfg->cutoff2 * h;
      double dx[3], sv[3];
      int in = 1;
      for (int k = 0; k < 3; k++) { dx[k] = refc[k] - st->x[k][j]; if (fabs(dx[k]) > supp) in = 0; }
      if (!in) continue;
      for (int k = 0; k < 3; k++) sv[k] = dx[k] / h;
      const double w = orc_weight(sv, h, 3), mj = st->m[j], rj = st->rho[j];
      uvec[0] += mj * w;
      uvec[1] += mj * w * st->v[0][j] / rj;
      uvec[2] += mj * w * st->v[1][j] / rj;
      uvec[3] += mj * w * st->v[2][j] / rj;
      uvec[4] += mj * w * st->e[j] / rj;
      wnorm += mj * w / rj;
      if (bj[0] != bi[0] || bj[1] != bi[1] || bj[2] != bi[2]) ks += mj * w * st->xi[j] / rj;
    }
    if (wnorm > 0) {
      for (int k = 0; k < 5; k++) uvec[k] /= wnorm;
      ks = ks / wnorm;
    } else {
      uvec[0] = 1; uvec[1] = uvec[2] = uvec[3] = uvec[4] = 0; ks = 0;
    }
    const double vdotn = uvec[1] * normal[0] + uvec[2] * normal[1] + uvec[3] * normal[2];
    st->rho[g] = uvec[0];
    for (int k = 0; k < 3; k++) st->v[k][g] = uvec[1 + k] - 2 * vdotn * normal[k];
    st->e[g] = uvec[4];
    st->xi[g] = ks;
    orc_eos(cfg, st->rho[g], st->e[g], st->xi[g], &st->p[g], &st->c[g], &st->T[g], &st->gam[g], &st->cp[g]);
  }
  return 0;
}

/* ===================================================================== */
/* Domain adaptation (ADJUST_DOMAIN, main.cc:256-304 and 433-436).         */
/* Buckets are the preprocessor's cells of edge del = PARTICLE_DENSITY      */
/* h_air over the extended domain; a bucket's type comes from its faces     */
/* against the largest domain dom (preprocess.cc:164-199, 380-440).  Each   */
/* bucket carries has_involved (bit 0 potentially involved, bit 1 involved, */
/* bucket.h:400-414, 579-589) and erupt_flag.  At the end of a step          */
/* scan_outside_layer marks the buckets of value 1 that hold an involved     */
/* particle with both bits and counts those next to a non-empty, uninvolved */
/* bucket that is not underground; the next step, after its time step,      */
/* adapt_domain turns the particles of the uninvolved buckets next to an    */
/* involved one real (potentially involved), add_pressure_ghost fills the   */
/* empty uninvolved buckets next to any involved or potentially involved one */
/* with pressure ghosts, add_wall_ghost the ground layer and the layer below */
/* with wall ghosts, and the wall ghosts are refreshed.  Bucket bookkeeping */
/* of the reference's hash tables (brief buckets, active marks, guests) has */
/* no effect on the particles and is not restated.                          */
/* ===================================================================== */
enum { ORC_UNDERGROUND = 1, ORC_MIXED = 2, ORC_OVERGROUND = 3, ORC_PRESS_BC = 4 };

static int orc_face(double crd, double max, double min) {    /* preprocess.cc:386-399 */
  if (crd < min) return -1;
  if (crd > max) return 1;
  return 0;
}

static void orc_bucket_box(const psph_config *cfg, int b, int k, double *lo, double *hi) {
  *lo = cfg->ext_lo[k] + b * cfg->bucket_edge;                /* preprocess.cc:166-179 */
  *hi = cfg->ext_lo[k] + (b + 1) * cfg->bucket_edge;
}

void orc_bucket_type(const psph_config *cfg, int bx, int by, int bz, int bt[6], int *type) {
  const int b[3] = {bx, by, bz};
  int flag[3], sum = 0;
  for (int k = 0; k < 3; k++) {
    double lo, hi;
    orc_bucket_box(cfg, b[k], k, &lo, &hi);
    bt[2 * k] = orc_face(lo, cfg->dom_hi[k], cfg->dom_lo[k]);
    bt[2 * k + 1] = orc_face(hi, cfg->dom_hi[k], cfg->dom_lo[k]);
    flag[k] = abs(bt[2 * k] + bt[2 * k + 1]);
    sum += flag[k];
  }
  if (flag[0] == 2 || flag[1] == 2 || flag[2] == 2)          /* preprocess.cc:419-433 */
    *type = (bt[4] == -1 && bt[5] == -1) ? ORC_UNDERGROUND : ORC_PRESS_BC;
  else if (flag[0] == 1 || flag[1] == 1 || flag[2] == 1)
    *type = ORC_MIXED;
  else
    *type = sum == 0 ? ORC_OVERGROUND : 0;
}

/* a box [lo, hi] meets the bucket unless it lies wholly on one side of it
 * on some axis (determine_bucket_involved_flag, dataread.cc:52-75, and
 * determine_erupt_buket, sph_lib.cc:1597-1620) */
static int orc_box_meets_bucket(const psph_config *cfg, const int b[3], const double *lo, const double *hi,
                                int box_vs_bucket) {
  for (int k = 0; k < 3; k++) {
    double blo, bhi;
    orc_bucket_box(cfg, b[k], k, &blo, &bhi);
    int f0, f1;
    if (box_vs_bucket) { f0 = orc_face(lo[k], bhi, blo); f1 = orc_face(hi[k], bhi, blo); }
    else { f0 = orc_face(blo, hi[k], lo[k]); f1 = orc_face(bhi, hi[k], lo[k]); }
    if (abs(f0 + f1) == 2) return 0;
  }
  return 1;
}

static size_t orc_nbuckets(const orc_state *st) { return (size_t)st->nbk[0] * st->nbk[1] * st->nbk[2]; }
static size_t orc_bidx(const orc_state *st, int bx, int by, int bz) {
  return ((size_t)bz * st->nbk[1] + by) * st->nbk[0] + bx;
}

int orc_domain_init(const psph_config *cfg, const orc_meteo *met, uint32_t seed, double jitter,
                    orc_state *st) {
  for (int k = 0; k < 3; k++) st->nbk[k] = (int)llround((cfg->ext_hi[k] - cfg->ext_lo[k]) / cfg->bucket_edge);
  const size_t nb = orc_nbuckets(st);
  free(st->bval); free(st->berupt);
  st->bval = calloc(nb ? nb : 1, 1);
  st->berupt = calloc(nb ? nb : 1, 1);
  if (!st->bval || !st->berupt) ORC_FAIL(PSPH_ENOMEM, "oom (bucket flags)");
  orc_meteo_free(&st->met);
  orc_meteo_from_arrays(&st->met, met->rows, met->nprop, met->z, met->data);
  st->lat_seed = seed;
  st->lat_jitter = jitter;
  st->adapt_pending = 0;
  /* the vent duct (setup_erupt.cc:86-91): the vent disc over one extended-  */
  /* domain bottom layer below the ground                                    */
  const double dlo[3] = {cfg->vent_x - cfg->vent_r, cfg->vent_y - cfg->vent_r, cfg->vent_z - 1.5 * cfg->bucket_edge};
  const double dhi[3] = {cfg->vent_x + cfg->vent_r, cfg->vent_y + cfg->vent_r, cfg->vent_z};
  for (int bz = 0; bz < st->nbk[2]; bz++)
    for (int by = 0; by < st->nbk[1]; by++)
      for (int bx = 0; bx < st->nbk[0]; bx++) {
        const int b[3] = {bx, by, bz};
        const size_t q = orc_bidx(st, bx, by, bz);
        /* potentially involved: the bucket meets the initial domain */
        st->bval[q] = (uint8_t)orc_box_meets_bucket(cfg, b, cfg->idom_lo, cfg->idom_hi, 0);
        st->berupt[q] = (uint8_t)orc_box_meets_bucket(cfg, b, dlo, dhi, 1);
      }
  /* a state loaded mid-run: buckets holding real particles are potentially */
  /* involved as well                                                        */
  for (size_t i = 0; i < st->n; i++) {
    int b[3];
    if (is_real(st->bc[i]) && orc_bucket_of(cfg, st, i, b) == 0) st->bval[orc_bidx(st, b[0], b[1], b[2])] |= 1;
  }
  return 0;
}

/* per-bucket occupancy: bit 0 any particle, bit 1 an involved one (2),     */
/* bit 2 a pressure ghost, bit 3 a wall ghost                               */
static uint8_t *orc_bucket_occupancy(const psph_config *cfg, const orc_state *st) {
  uint8_t *occ = calloc(orc_nbuckets(st) ? orc_nbuckets(st) : 1, 1);
  if (!occ) return NULL;
  for (size_t i = 0; i < st->n; i++) {
    int b[3];
    if (orc_bucket_of(cfg, st, i, b)) continue;
    uint8_t f = 1;
    if (st->inv[i] == 2) f |= 2;
    if (st->bc[i] == 1) f |= 4;
    if (st->bc[i] == 2) f |= 8;
    occ[orc_bidx(st, b[0], b[1], b[2])] |= f;
  }
  return occ;
}

static int orc_any_neighbour(const orc_state *st, int bx, int by, int bz, uint8_t mask) {
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (!dx && !dy && !dz) continue;
        const int x = bx + dx, y = by + dy, z = bz + dz;
        if (x < 0 || y < 0 || z < 0 || x >= st->nbk[0] || y >= st->nbk[1] || z >= st->nbk[2]) continue;
        if (st->bval[orc_bidx(st, x, y, z)] & mask) return 1;
      }
  return 0;
}

int orc_domain_scan(const psph_config *cfg, orc_state *st) {      /* scan_outside_layer.cc:44-121 */
  if (!st->bval) ORC_FAIL(PSPH_EINVAL, "domain adaptation: call orc_domain_init first");
  uint8_t *occ = orc_bucket_occupancy(cfg, st);
  if (!occ) ORC_FAIL(PSPH_ENOMEM, "oom (bucket occupancy)");
  int adapt = 0;
  for (int bz = 0; bz < st->nbk[2]; bz++)
    for (int by = 0; by < st->nbk[1]; by++)
      for (int bx = 0; bx < st->nbk[0]; bx++) {
        const size_t q = orc_bidx(st, bx, by, bz);
        if (st->bval[q] != 1 || !(occ[q] & 2)) continue;
        st->bval[q] = 3;                                          /* HAS_BOTH */
        for (int dz = -1, done = 0; dz <= 1 && !done; dz++)
          for (int dy = -1; dy <= 1 && !done; dy++)
            for (int dx = -1; dx <= 1 && !done; dx++) {
              if (!dx && !dy && !dz) continue;
              const int x = bx + dx, y = by + dy, z = bz + dz;
              if (x < 0 || y < 0 || z < 0 || x >= st->nbk[0] || y >= st->nbk[1] || z >= st->nbk[2]) continue;
              const size_t r = orc_bidx(st, x, y, z);
              int bt[6], type;
              orc_bucket_type(cfg, x, y, z, bt, &type);
              if ((occ[r] & 1) && st->bval[r] == 0 && type != ORC_UNDERGROUND) { adapt++; done = 1; }
            }
      }
  free(occ);
  st->adapt_pending = adapt;
  return 0;
}

/* a new air particle at lattice point l (add_pressure_ghost.cc:164-232,    */
/* add_wall_ghost.cc:150-205): initial_air state, keys with this step       */
static int orc_add_lattice_particle(const psph_config *cfg, orc_state *st, const double pc[3], int bc) {
  if (orc_state_reserve(st, st->n + 1 > st->cap ? 2 * st->cap + 1024 : st->cap)) return PSPH_ENOMEM;
  size_t i = orc_push(st);
  for (int k = 0; k < 3; k++) st->x[k][i] = pc[k];
  double normc[3];
  for (int k = 0; k < 3; k++) normc[k] = (pc[k] - cfg->ext_lo[k]) / (cfg->ext_hi[k] - cfg->ext_lo[k]);
  uint32_t key[3];
  orc_thsfc3d(normc, (uint32_t)st->step, key);
  for (int k = 0; k < 3; k++) st->key[k][i] = key[k];
  st->h[i] = cfg->h_air;
  st->bc[i] = bc;
  st->phase[i] = 1;
  st->inv[i] = 0;
  st->xi[i] = 0.;
  st->cp[i] = 0.;
  orc_air_state(cfg, &st->met, st, i);
  return 0;
}

/* calls fn on the lattice points of bucket (bx, by, bz) */
typedef int (*orc_lat_fn)(const psph_config *, orc_state *, const double pc0[3], const double pc[3], void *);
static int orc_bucket_lattice(const psph_config *cfg, orc_state *st, int bx, int by, int bz, orc_lat_fn fn, void *ud) {
  const int per = (int)llround(cfg->bucket_edge / cfg->h_air);
  const int64_t nl[3] = {(int64_t)st->nbk[0] * per, (int64_t)st->nbk[1] * per, (int64_t)st->nbk[2] * per};
  const int b[3] = {bx, by, bz};
  for (int i = 0; i < per; i++)
    for (int j = 0; j < per; j++)
      for (int k = 0; k < per; k++) {
        const int64_t l[3] = {(int64_t)b[0] * per + i, (int64_t)b[1] * per + j, (int64_t)b[2] * per + k};
        double pc0[3], pc[3];
        orc_lattice_point(cfg, st->lat_seed, st->lat_jitter, l, nl, pc0, pc);
        int rc = fn(cfg, st, pc0, pc, ud);
        if (rc) return rc;
      }
  return 0;
}

static int orc_lat_pghost_above(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud;
  return pc0[2] > cfg->dom_lo[2] ? orc_add_lattice_particle(cfg, st, pc, 1) : 0;
}
static int orc_lat_pghost(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud; (void)pc0;
  return orc_add_lattice_particle(cfg, st, pc, 1);
}
static int orc_lat_wghost_below(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud;
  return pc0[2] < cfg->dom_lo[2] ? orc_add_lattice_particle(cfg, st, pc, 2) : 0;
}
static int orc_lat_wghost(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud; (void)pc0;
  return orc_add_lattice_particle(cfg, st, pc, 2);
}
static int orc_lat_real_above(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud;
  if (!(pc0[2] > cfg->dom_lo[2])) return 0;
  int rc = orc_add_lattice_particle(cfg, st, pc, 100);
  st->inv[st->n - 1] = 1;
  return rc;
}
static int orc_lat_real(const psph_config *cfg, orc_state *st, const double pc0[3], const double pc[3], void *ud) {
  (void)ud; (void)pc0;
  int rc = orc_add_lattice_particle(cfg, st, pc, 100);
  st->inv[st->n - 1] = 1;
  return rc;
}

static int orc_add_pressure_ghost(const psph_config *cfg, orc_state *st) {  /* add_pressure_ghost.cc:37-316 */
  uint8_t *occ = orc_bucket_occupancy(cfg, st);
  if (!occ) ORC_FAIL(PSPH_ENOMEM, "oom (bucket occupancy)");
  int rc = 0;
  for (int bz = 0; bz < st->nbk[2] && !rc; bz++)
    for (int by = 0; by < st->nbk[1] && !rc; by++)
      for (int bx = 0; bx < st->nbk[0] && !rc; bx++) {
        const size_t q = orc_bidx(st, bx, by, bz);
        if (st->bval[q] || (occ[q] & 4) || st->berupt[q]) continue;
        if (!orc_any_neighbour(st, bx, by, bz, 3)) continue;
        int bt[6], type;
        orc_bucket_type(cfg, bx, by, bz, bt, &type);
        if (type == ORC_MIXED)
          rc = orc_bucket_lattice(cfg, st, bx, by, bz, bt[4] == -1 ? orc_lat_pghost_above : orc_lat_pghost, NULL);
        else if (type == ORC_OVERGROUND || type == ORC_PRESS_BC)
          rc = orc_bucket_lattice(cfg, st, bx, by, bz, orc_lat_pghost, NULL);
      }
  free(occ);
  return rc;
}

static int orc_add_wall_ghost(const psph_config *cfg, orc_state *st) {      /* add_wall_ghost.cc:45-293 */
  uint8_t *occ = orc_bucket_occupancy(cfg, st);
  if (!occ) ORC_FAIL(PSPH_ENOMEM, "oom (bucket occupancy)");
  int rc = 0;
  for (int bz = 0; bz < st->nbk[2] && !rc; bz++)
    for (int by = 0; by < st->nbk[1] && !rc; by++)
      for (int bx = 0; bx < st->nbk[0] && !rc; bx++) {
        const size_t q = orc_bidx(st, bx, by, bz);
        int bt[6], type;
        orc_bucket_type(cfg, bx, by, bz, bt, &type);
        if (!(type == ORC_MIXED && bt[4] == -1) || !(occ[q] & 1) || (occ[q] & 8) || st->berupt[q]) continue;
        rc = orc_bucket_lattice(cfg, st, bx, by, bz, orc_lat_wghost_below, NULL);
        if (!rc && bz > 0) rc = orc_bucket_lattice(cfg, st, bx, by, bz - 1, orc_lat_wghost, NULL);
      }
  free(occ);
  return rc;
}

int orc_domain_adapt(const psph_config *cfg, orc_state *st) {      /* main.cc:256-304 */
  if (!st->bval) ORC_FAIL(PSPH_EINVAL, "domain adaptation: call orc_domain_init first");
  const size_t n0 = st->n;
  uint8_t *occ = orc_bucket_occupancy(cfg, st);
  if (!occ) ORC_FAIL(PSPH_ENOMEM, "oom (bucket occupancy)");
  /* adapt_domain (adapt_domain.cc:41-194) */
  const size_t nb = orc_nbuckets(st);
  uint8_t *conv = calloc(nb ? nb : 1, 1);
  if (!conv) { free(occ); ORC_FAIL(PSPH_ENOMEM, "oom"); }
  for (int bz = 0; bz < st->nbk[2]; bz++)
    for (int by = 0; by < st->nbk[1]; by++)
      for (int bx = 0; bx < st->nbk[0]; bx++) {
        const size_t q = orc_bidx(st, bx, by, bz);
        int bt[6], type;
        orc_bucket_type(cfg, bx, by, bz, bt, &type);
        if (st->bval[q] || !(occ[q] & 1) || type == ORC_UNDERGROUND || type == ORC_PRESS_BC) continue;
        if (!orc_any_neighbour(st, bx, by, bz, 2)) continue;
        conv[q] = (uint8_t)(type == ORC_MIXED ? (bt[4] == -1 ? 2 : 3) : 1);
      }
  const double TINY = 1.0E-08;
  for (size_t i = 0; i < st->n; i++) {
    int b[3];
    if (orc_bucket_of(cfg, st, i, b)) continue;
    const size_t q = orc_bidx(st, b[0], b[1], b[2]);
    if (!conv[q]) continue;
    int turn = 1;
    if (conv[q] == 2) turn = st->bc[i] == 1;              /* ground MIXED: pressure ghosts only */
    else if (conv[q] == 3) {                               /* MIXED elsewhere: not outside dom   */
      int bt[6], type;                                     /* (Bucket::determine_escape,        */
      orc_bucket_type(cfg, b[0], b[1], b[2], bt, &type);   /*  bucket.cc:435-452)               */
      for (int k = 0; k < 3 && turn; k++) {
        if (bt[2 * k] == -1) { if ((st->x[k][i] - cfg->dom_lo[k]) < -TINY) turn = 0; }
        else if (bt[2 * k + 1] == 1) { if ((st->x[k][i] - cfg->dom_hi[k]) > TINY) turn = 0; }
      }
    }
    if (turn) { st->bc[i] = 100; st->inv[i] = 1; }
  }
  for (size_t q = 0; q < nb; q++) if (conv[q]) st->bval[q] = 1;
  free(conv);
  free(occ);
  int rc = orc_add_pressure_ghost(cfg, st);
  if (!rc) rc = orc_add_wall_ghost(cfg, st);
  st->last_adapted = (int64_t)(st->n - n0);
  if (!rc) orc_sort_canonical(cfg, st);
  return rc;
}

int orc_domain_setup(const psph_config *cfg, const orc_meteo *met, uint32_t seed, double jitter,
                     orc_state *st) {
  int rc = orc_domain_init(cfg, met, seed, jitter, st);
  if (rc) return rc;
  /* add_air (add_air.cc:491-813): the potentially involved buckets, real  */
  for (int bz = 0; bz < st->nbk[2] && !rc; bz++)
    for (int by = 0; by < st->nbk[1] && !rc; by++)
      for (int bx = 0; bx < st->nbk[0] && !rc; bx++) {
        if (!st->bval[orc_bidx(st, bx, by, bz)]) continue;
        int bt[6], type;
        orc_bucket_type(cfg, bx, by, bz, bt, &type);
        if (type == ORC_MIXED)
          rc = orc_bucket_lattice(cfg, st, bx, by, bz, bt[4] == -1 ? orc_lat_real_above : orc_lat_real, NULL);
        else if (type == ORC_OVERGROUND)
          rc = orc_bucket_lattice(cfg, st, bx, by, bz, orc_lat_real, NULL);
      }
  if (!rc) rc = orc_add_pressure_ghost(cfg, st);          /* main.cc:128-132 */
  if (!rc) rc = orc_add_wall_ghost(cfg, st);
  if (!rc) rc = orc_setup_erupt(cfg, st);
  if (!rc) orc_sort_canonical(cfg, st);
  return rc;
}

int orc_step(const psph_config *cfg, orc_state *st) {
  int rc;
  double dt = orc_timestep(cfg, st);
  orc_incrtime(cfg, st, &dt);
  st->last_injected = st->last_removed = 0;
  st->last_adapted = 0;
  if (cfg->dim == 3 && cfg->adapt_domain && st->adapt_pending) {
    st->adapt_pending = 0;
    if ((rc = orc_domain_adapt(cfg, st))) return rc;
    if (cfg->wall_bc && (rc = orc_apply_bcond(cfg, st))) return rc;
  }
  if ((rc = orc_neighbors(cfg, st))) return rc;
  if (cfg->sml_update_int > 0 && st->step % cfg->sml_update_int == 0)
    if ((rc = orc_adaptive_sml(cfg, st))) return rc;
  if ((rc = orc_rhs(cfg, st, dt))) return rc;
  if ((rc = orc_density(cfg, st))) return rc;
  if (cfg->lans && (rc = orc_vsmooth(cfg, st))) return rc;
  if ((rc = orc_update_pos(cfg, st, dt))) return rc;
  if (cfg->dim == 3) {
    if ((rc = orc_remove_escaped(cfg, st))) return rc;
    if (orc_iferupt(cfg, st) && (rc = orc_inject(cfg, st, dt))) return rc;
    if (cfg->adapt_domain && (rc = orc_domain_scan(cfg, st))) return rc;
    if (cfg->wall_bc && (rc = orc_apply_bcond(cfg, st))) return rc;
    orc_sort_canonical(cfg, st);
  }
  orc_ifoutput(cfg, st);
  return 0;
}
