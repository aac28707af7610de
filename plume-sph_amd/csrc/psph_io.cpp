// On-disk output of the particle state: H5Part steps (h5part_writer.cc:40-444,
// hdf5calls.cc:288-346) and the 1D CSV table (h5part_writer.cc:448-489).
// Host code only; links HDF5 and libpsph.so (psph_read for a context's
// particles, fail / psph_last_error for errors).
#include <hdf5.h>

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psph_internal.h"
#include "../../include/psph_io.h"

#define PSPH_TRY(expr)   \
  do {                   \
    int rc_ = (expr);    \
    if (rc_) return rc_; \
  } while (0)

using psph::fail;

namespace {

// Closes an HDF5 handle on scope exit (id < 0: nothing to close).
struct H5Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H5Id() {
    if (id >= 0 && close) close(id);
  }
  H5Id(const H5Id &) = delete;
  H5Id &operator=(const H5Id &) = delete;
  operator hid_t() const { return id; }
};

// HDF5 prints its error stack to stderr by default; errors here are
// reported through the return code and psph_last_error instead.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

std::string group_name(int64_t group) {
  char g[32];
  std::snprintf(g, sizeof g, "Step#%" PRId64, group);
  return g;
}

// One-dimensional dataset of n elements (GH5_WriteS, hdf5calls.cc:288-346:
// dims {n, 0} gives a rank-1 simple dataspace, default creation properties,
// so a contiguous layout).
int write_1d(hid_t gid, const char *name, hid_t type, const void *buf, size_t n) {
  hsize_t dims[1] = {(hsize_t)n};
  H5Id space(H5Screate_simple(1, dims, nullptr), H5Sclose);
  if (space < 0) return fail(PSPH_EIO, "h5part: dataspace for %s", name);
  H5Id ds(H5Dcreate2(gid, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (ds < 0) return fail(PSPH_EIO, "h5part: cannot create dataset %s", name);
  if (n && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
    return fail(PSPH_EIO, "h5part: cannot write dataset %s", name);
  return PSPH_OK;
}

bool selected(int32_t which, int32_t bc) {
  switch (which) {
    case PSPH_H5_REAL: return bc == 100;               // h5part_writer.cc:87
    case PSPH_H5_SHOW: return bc == 100 || bc == 99;   // need_neigh (particle.h:555-558)
    default: return bc >= 0;                           // every particle held
  }
}

}  // namespace

extern "C" int psph_h5part_write_soa(const char *path, int create, int64_t group, const psph_soa *p, size_t n,
                                     int32_t dim, int32_t which, int32_t extras, int32_t rank) {
  if (!path || (!p && n)) return fail(PSPH_EINVAL, "h5part: null argument");
  if (which < PSPH_H5_REAL || which > PSPH_H5_ALL) return fail(PSPH_EINVAL, "h5part: which = %d", which);
  if (dim != 1 && dim != 3) return fail(PSPH_EINVAL, "h5part: dim = %d", dim);
  if (group < 0) return fail(PSPH_EINVAL, "h5part: group = %" PRId64, group);
  std::vector<size_t> sel;
  sel.reserve(n);
  for (size_t i = 0; i < n; i++)
    if (selected(which, p->bc[i])) sel.push_back(i);
  const size_t m = sel.size();
  if (m > (size_t)INT32_MAX) return fail(PSPH_ECAPACITY, "h5part: %zu particles exceed the int32 ID range", m);

  quiet();
  H5Id fp(create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                 : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT),
          H5Fclose);
  if (fp < 0) return fail(PSPH_EIO, "h5part: cannot %s %s", create ? "create" : "open", path);
  const std::string gname = group_name(group);
  H5Id gid(H5Gcreate2(fp, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (gid < 0) return fail(PSPH_EIO, "h5part: cannot create group %s in %s", gname.c_str(), path);

  std::vector<int32_t> ib(m);
  std::vector<double> db(m);
  auto put_i = [&](const char *name) { return write_1d(gid, name, H5T_NATIVE_INT32, ib.data(), m); };
  auto put_d = [&](const char *name, const double *src) {
    for (size_t r = 0; r < m; r++) db[r] = src ? src[sel[r]] : 0.0;
    return write_1d(gid, name, H5T_NATIVE_DOUBLE, db.data(), m);
  };
  const bool d3 = dim == 3;
  // dataset order of h5part_writer.cc:179-252
  for (size_t r = 0; r < m; r++) ib[r] = (int32_t)r;
  PSPH_TRY(put_i("ID"));
  if (extras & PSPH_H5_INVOLVED) {
    for (size_t r = 0; r < m; r++) ib[r] = p->involved[sel[r]];
    PSPH_TRY(put_i("Involved"));
  }
  if (extras & PSPH_H5_PID) {
    for (size_t r = 0; r < m; r++) ib[r] = rank;
    PSPH_TRY(put_i("myprocess"));
  }
  if (extras & PSPH_H5_MASS) PSPH_TRY(put_d("mymass", p->m));
  if (extras & PSPH_H5_SML) PSPH_TRY(put_d("sml", p->h));
  PSPH_TRY(put_d("x", p->x));
  PSPH_TRY(put_d("y", d3 ? p->y : nullptr));
  PSPH_TRY(put_d("z", d3 ? p->z : nullptr));
  PSPH_TRY(put_d("Rho", p->rho));
  PSPH_TRY(put_d("Vx", p->vx));
  PSPH_TRY(put_d("Vy", d3 ? p->vy : nullptr));
  PSPH_TRY(put_d("Vz", d3 ? p->vz : nullptr));
  PSPH_TRY(put_d("engr", p->e));
  PSPH_TRY(put_d("mssfrc", p->xi));
  PSPH_TRY(put_d("prss", p->p));
  for (size_t r = 0; r < m; r++) ib[r] = p->phase[sel[r]];
  PSPH_TRY(put_i("phase"));
  for (size_t r = 0; r < m; r++) ib[r] = p->bc[sel[r]];
  PSPH_TRY(put_i("bctp"));
  for (size_t r = 0; r < m; r++) ib[r] = 1;            // get_guest: 1 for an owned particle
  PSPH_TRY(put_i("guest"));
  if (H5Fflush(fp, H5F_SCOPE_LOCAL) < 0) return fail(PSPH_EIO, "h5part: flush of %s", path);
  return PSPH_OK;
}

namespace {

// Host copy of every particle of a context, in psph_read order.
struct Snapshot {
  psph::HostParticles P;
  psph_soa soa{};
  int fill(psph_ctx *ctx) {
    size_t n = 0;
    PSPH_TRY(psph_read(ctx, nullptr, 0, &n));
    P.resize(n);
#define PSPH_VIEW(f) soa.f = P.f.data();
    PSPH_VIEW(x) PSPH_VIEW(y) PSPH_VIEW(z) PSPH_VIEW(vx) PSPH_VIEW(vy) PSPH_VIEW(vz) PSPH_VIEW(svx)
    PSPH_VIEW(svy) PSPH_VIEW(svz) PSPH_VIEW(rho) PSPH_VIEW(e) PSPH_VIEW(m) PSPH_VIEW(h) PSPH_VIEW(p)
    PSPH_VIEW(c) PSPH_VIEW(T) PSPH_VIEW(gam) PSPH_VIEW(cp) PSPH_VIEW(xi) PSPH_VIEW(bc) PSPH_VIEW(phase)
#undef PSPH_VIEW
    soa.key0 = P.k0.data(); soa.key1 = P.k1.data(); soa.key2 = P.k2.data();
    soa.involved = P.inv.data();
    size_t got = 0;
    PSPH_TRY(psph_read(ctx, &soa, n, &got));
    if (got != n) return fail(PSPH_EINVARIANT, "h5part: particle count changed (%zu, %zu)", n, got);
    return PSPH_OK;
  }
};

}  // namespace

extern "C" int psph_write_h5part(psph_ctx *ctx, const char *path, int create, int64_t group, int32_t which,
                                 int32_t extras) {
  if (!ctx || !path) return fail(PSPH_EINVAL, "h5part: null argument");
  psph_config cfg;
  PSPH_TRY(psph_get_config(ctx, &cfg));
  Snapshot s;
  PSPH_TRY(s.fill(ctx));
  return psph_h5part_write_soa(path, create, group, &s.soa, s.P.size(), cfg.dim, which, extras, 0);
}

extern "C" int psph_h5part_steps(const char *path, int64_t *nsteps) {
  if (!path || !nsteps) return fail(PSPH_EINVAL, "h5part: null argument");
  quiet();
  H5Id fp(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (fp < 0) return fail(PSPH_EIO, "h5part: cannot open %s", path);
  int64_t k = 0;
  while (H5Lexists(fp, group_name(k).c_str(), H5P_DEFAULT) > 0) k++;
  *nsteps = k;
  return PSPH_OK;
}

extern "C" int psph_h5part_read(const char *path, int64_t group, const char *name, int32_t is_int, void *buf,
                                size_t cap, size_t *n_out) {
  if (!path || !name || !n_out) return fail(PSPH_EINVAL, "h5part: null argument");
  quiet();
  H5Id fp(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (fp < 0) return fail(PSPH_EIO, "h5part: cannot open %s", path);
  const std::string full = group_name(group) + "/" + name;
  if (H5Lexists(fp, group_name(group).c_str(), H5P_DEFAULT) <= 0 || H5Lexists(fp, full.c_str(), H5P_DEFAULT) <= 0)
    return fail(PSPH_EIO, "h5part: %s has no dataset %s", path, full.c_str());
  H5Id ds(H5Dopen2(fp, full.c_str(), H5P_DEFAULT), H5Dclose);
  if (ds < 0) return fail(PSPH_EIO, "h5part: cannot open %s", full.c_str());
  H5Id space(H5Dget_space(ds), H5Sclose);
  if (H5Sget_simple_extent_ndims(space) != 1) return fail(PSPH_EIO, "h5part: %s is not one-dimensional", full.c_str());
  hsize_t dims[1] = {0};
  H5Sget_simple_extent_dims(space, dims, nullptr);
  *n_out = (size_t)dims[0];
  if (!buf) return PSPH_OK;
  if (*n_out > cap) return fail(PSPH_ECAPACITY, "h5part: %s holds %zu values, buffer %zu", full.c_str(), *n_out, cap);
  if (*n_out && H5Dread(ds, is_int ? H5T_NATIVE_INT32 : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
    return fail(PSPH_EIO, "h5part: cannot read %s", full.c_str());
  return PSPH_OK;
}

extern "C" int psph_csv_write_soa(const char *path, const psph_soa *p, size_t n) {
  if (!path || (!p && n)) return fail(PSPH_EINVAL, "csv: null argument");
  FILE *f = std::fopen(path, "a");
  if (!f) return fail(PSPH_EIO, "csv: cannot open %s", path);
  // header as h5part_writer.cc:462 writes it ("smlen" and "pmass" run together)
  std::fputs("RHO,X,U,E,KS,P,PHASE,BCTP,GUEST_FLAG,INVOLVE_FLAG,smlenpmass\n", f);
  // operator<< of a double with the default stream state prints %g with six
  // significant digits; the guest flag is a double (1 for an owned particle)
  for (size_t i = 0; i < n; i++) {
    if (p->bc[i] != 100) continue;                      // h5part_writer.cc:471
    std::fprintf(f, "%g,%g,%g,%g,%g,%g,%d,%d,%g,%d,%g,%g\n", p->rho[i], p->x[i], p->vx[i], p->e[i], p->xi[i],
                 p->p[i], p->phase[i], p->bc[i], 1.0, p->involved[i], p->h[i], p->m[i]);
  }
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) return fail(PSPH_EIO, "csv: write to %s failed", path);
  return PSPH_OK;
}

extern "C" int psph_write_csv(psph_ctx *ctx, const char *path) {
  if (!ctx || !path) return fail(PSPH_EINVAL, "csv: null argument");
  Snapshot s;
  PSPH_TRY(s.fill(ctx));
  return psph_csv_write_soa(path, &s.soa, s.P.size());
}
