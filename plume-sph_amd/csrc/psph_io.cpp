// On-disk output of the particle state: H5Part steps (h5part_writer.cc:40-444,
// hdf5calls.cc:288-346) and the 1D CSV table (h5part_writer.cc:448-489).
// Host code only; links HDF5 and libpsph.so (psph_read for a context's
// particles, fail / psph_last_error for errors).
#include <hdf5.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstddef>
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

// ---------------------------------------------------------------------------
// Grid files (funky%04d.h5): createfunky (createfunky.cc:32-71) and Read_Grid
// (dataread.cc:248-499).
// ---------------------------------------------------------------------------
namespace {

constexpr int kKeyLen = 2;    // bucket keys (the preprocessor's KEYLENGTH)
constexpr int kNeigh = 27;    // NEIGH_SIZE (constant.h:117)

// The preprocessor's record (buckstr.h:14-26): the file's compound type takes
// its size and member offsets from it.
struct BucketRecord {
  uint32_t key[kKeyLen];
  uint32_t neighs[kNeigh * kKeyLen];
  int32_t buckettype;
  int32_t myproc;
  int32_t neigh_proc[kNeigh];
  int32_t bucket_index[6];
  double xcoord[2], ycoord[2], zcoord[2];
  double elev[4];
};
static_assert(sizeof(BucketRecord) == 448 && offsetof(BucketRecord, buckettype) == 224 &&
                  offsetof(BucketRecord, neigh_proc) == 232 && offsetof(BucketRecord, xcoord) == 368 &&
                  offsetof(BucketRecord, elev) == 416,
              "the record must keep BucketStruct's layout");

// GH5_bucketstruct (hdf5calls.cc:175-203): one scalar per member, at the
// record's offsets
hid_t bucket_file_type() {
  const hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(BucketRecord));
  H5Tinsert(t, "Key", offsetof(BucketRecord, key), H5T_NATIVE_UINT);
  H5Tinsert(t, "neighs", offsetof(BucketRecord, neighs), H5T_NATIVE_UINT);
  H5Tinsert(t, "buckettype", offsetof(BucketRecord, buckettype), H5T_NATIVE_INT);
  H5Tinsert(t, "myproc", offsetof(BucketRecord, myproc), H5T_NATIVE_INT);
  H5Tinsert(t, "neigh_proc", offsetof(BucketRecord, neigh_proc), H5T_NATIVE_INT);
  H5Tinsert(t, "xcoord", offsetof(BucketRecord, xcoord), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "ycoord", offsetof(BucketRecord, ycoord), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "zcoord", offsetof(BucketRecord, zcoord), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "elev", offsetof(BucketRecord, elev), H5T_NATIVE_DOUBLE);
  return t;
}

// what a record holds, read by member name
struct BucketRead {
  uint32_t key, neighs;
  int32_t buckettype, myproc, neigh_proc;
  double x, y, z, elev;
};
hid_t bucket_read_type() {
  const hid_t t = H5Tcreate(H5T_COMPOUND, sizeof(BucketRead));
  H5Tinsert(t, "Key", HOFFSET(BucketRead, key), H5T_NATIVE_UINT32);
  H5Tinsert(t, "neighs", HOFFSET(BucketRead, neighs), H5T_NATIVE_UINT32);
  H5Tinsert(t, "buckettype", HOFFSET(BucketRead, buckettype), H5T_NATIVE_INT32);
  H5Tinsert(t, "myproc", HOFFSET(BucketRead, myproc), H5T_NATIVE_INT32);
  H5Tinsert(t, "neigh_proc", HOFFSET(BucketRead, neigh_proc), H5T_NATIVE_INT32);
  H5Tinsert(t, "xcoord", HOFFSET(BucketRead, x), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "ycoord", HOFFSET(BucketRead, y), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "zcoord", HOFFSET(BucketRead, z), H5T_NATIVE_DOUBLE);
  H5Tinsert(t, "elev", HOFFSET(BucketRead, elev), H5T_NATIVE_DOUBLE);
  return t;
}

// determine_face_type (preprocess.cc:386-399)
int face_of(double crd, double hi, double lo) { return crd < lo ? -1 : (crd > hi ? 1 : 0); }

int dataset_len(hid_t fp, const char *name, hsize_t *n) {
  if (H5Lexists(fp, name, H5P_DEFAULT) <= 0) return fail(PSPH_EIO, "grid file: no dataset %s", name);
  H5Id ds(H5Dopen2(fp, name, H5P_DEFAULT), H5Dclose);
  if (ds < 0) return fail(PSPH_EIO, "grid file: cannot open %s", name);
  H5Id space(H5Dget_space(ds), H5Sclose);
  if (H5Sget_simple_extent_ndims(space) != 1) return fail(PSPH_EIO, "grid file: %s is not one-dimensional", name);
  H5Sget_simple_extent_dims(space, n, nullptr);
  return PSPH_OK;
}

int grid_load(const char *path, psph_grid_info *info, std::vector<BucketRead> &rec) {
  if (!path || !info) return fail(PSPH_EINVAL, "grid file: null argument");
  quiet();
  H5Id fp(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (fp < 0) return fail(PSPH_EIO, "grid file: cannot open %s", path);
  std::memset(info, 0, sizeof *info);
  hsize_t nh = 0, nb = 0, np = 0;
  PSPH_TRY(dataset_len(fp, "/hashtable_constants", &nh));
  if (nh != 6) return fail(PSPH_EIO, "grid file: /hashtable_constants holds %llu values, not 6", (unsigned long long)nh);
  double ht[6];
  {
    H5Id ds(H5Dopen2(fp, "/hashtable_constants", H5P_DEFAULT), H5Dclose);
    if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, ht) < 0)
      return fail(PSPH_EIO, "grid file: cannot read /hashtable_constants");
  }
  for (int d = 0; d < 3; d++) { info->ext_lo[d] = ht[2 * d]; info->ext_hi[d] = ht[2 * d + 1]; }
  PSPH_TRY(dataset_len(fp, "/Buckets", &nb));
  rec.resize(nb);
  if (nb) {
    H5Id ds(H5Dopen2(fp, "/Buckets", H5P_DEFAULT), H5Dclose);
    H5Id t(bucket_read_type(), H5Tclose);
    if (H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, rec.data()) < 0)
      return fail(PSPH_EIO, "grid file: cannot read /Buckets");
  }
  PSPH_TRY(dataset_len(fp, "/partition_table", &np));
  info->n_buckets = (int64_t)nb;
  info->n_partition = (int64_t)np;
  info->flat = 1;
  info->ground = nb ? rec[0].elev : 0.0;
  int32_t maxproc = -1;
  double edge = 0;
  for (const BucketRead &r : rec) {
    if (r.buckettype >= 0 && r.buckettype <= 4) info->n_type[r.buckettype]++;
    maxproc = std::max(maxproc, r.myproc);
    if (r.elev != info->ground) info->flat = 0;
    const double dx = r.x - info->ext_lo[0];
    if (dx > 0 && (edge == 0 || dx < edge)) edge = dx;
  }
  info->nproc = maxproc + 1;
  // the lower corners lie on ext_lo + i del: the smallest positive offset
  // from the origin on x is del when the file holds a bucket at i = 1
  info->bucket_edge = edge;
  for (int d = 0; d < 3; d++)
    info->nb[d] = edge > 0 ? (int32_t)std::llround((info->ext_hi[d] - info->ext_lo[d]) / edge) : 0;
  return PSPH_OK;
}

}  // namespace

extern "C" int psph_grid_read(const char *path, psph_grid_info *out) {
  std::vector<BucketRead> rec;
  return grid_load(path, out, rec);
}

extern "C" int psph_grid_check(const char *path, const psph_config *cfg) {
  if (!cfg) return fail(PSPH_EINVAL, "grid file: null argument");
  if (cfg->dim != 3) return fail(PSPH_EINVAL, "grid file: the grid describes a 3D domain");
  psph_grid_info g;
  std::vector<BucketRead> rec;
  PSPH_TRY(grid_load(path, &g, rec));
  if (g.nproc > 1)
    return fail(PSPH_ENOTSUP, "grid file %s: buckets of %d processes; one GPU runs the whole domain (preprocess "
                "with one process)", path, g.nproc);
  if (!g.flat) return fail(PSPH_ENOTSUP, "grid file %s: the ground elevation varies (terrain), not supported", path);
  const double del = cfg->bucket_edge;
  for (int d = 0; d < 3; d++) {
    const double tol = 1e-9 * (cfg->ext_hi[d] - cfg->ext_lo[d]);
    if (std::fabs(g.ext_lo[d] - cfg->ext_lo[d]) > tol || std::fabs(g.ext_hi[d] - cfg->ext_hi[d]) > tol)
      return fail(PSPH_EINVAL, "grid file %s: extended domain [%.17g, %.17g] on axis %d, the config's [%.17g, %.17g]",
                  path, g.ext_lo[d], g.ext_hi[d], d, cfg->ext_lo[d], cfg->ext_hi[d]);
  }
  if (std::fabs(g.bucket_edge - del) > 1e-9 * del)
    return fail(PSPH_EINVAL, "grid file %s: bucket edge %.17g, the config's %.17g", path, g.bucket_edge, del);
  if (std::fabs(g.ground - cfg->dom_lo[2]) > 1e-9 * del)
    return fail(PSPH_EINVAL, "grid file %s: ground at %.17g, the config's dom_lo[2] %.17g", path, g.ground, cfg->dom_lo[2]);
  psph::DomainBuckets B;
  psph::domain_buckets(*cfg, B);
  if ((int64_t)B.size() != g.n_buckets)
    return fail(PSPH_EINVAL, "grid file %s: %lld buckets, the config's grid %zu", path, (long long)g.n_buckets, B.size());
  std::vector<uint8_t> seen(B.size(), 0);
  for (const BucketRead &r : rec) {
    const double c[3] = {r.x, r.y, r.z};
    int b[3];
    for (int d = 0; d < 3; d++) {
      const double f = (c[d] - cfg->ext_lo[d]) / del;
      b[d] = (int)std::llround(f);
      if (b[d] < 0 || b[d] >= B.nb[d] || std::fabs(f - b[d]) > 1e-6)
        return fail(PSPH_EINVAL, "grid file %s: a bucket corner (%.17g, %.17g, %.17g) off the config's grid", path, c[0],
                    c[1], c[2]);
    }
    const size_t q = B.idx(b[0], b[1], b[2]);
    if (seen[q]++) return fail(PSPH_EINVAL, "grid file %s: bucket (%d, %d, %d) twice", path, b[0], b[1], b[2]);
    if (r.buckettype != B.type[q])
      return fail(PSPH_EINVAL, "grid file %s: bucket (%d, %d, %d) of type %d, this build derives %d", path, b[0], b[1],
                  b[2], r.buckettype, (int)B.type[q]);
  }
  return PSPH_OK;
}

extern "C" int psph_grid_write(const char *path, const psph_config *cfg) {
  if (!path || !cfg) return fail(PSPH_EINVAL, "grid file: null argument");
  if (cfg->dim != 3) return fail(PSPH_EINVAL, "grid file: the grid describes a 3D domain");
  quiet();
  psph::DomainBuckets B;
  psph::domain_buckets(*cfg, B);
  const int nx = B.nb[0], ny = B.nb[1], nz = B.nb[2];
  const double del = cfg->bucket_edge;
  const double *lo = cfg->ext_lo, *hi = cfg->ext_hi;
  std::vector<BucketRecord> bg(B.size());
  std::memset(bg.data(), 0, bg.size() * sizeof(BucketRecord));
  auto key_at = [&](int i, int j, int k, uint32_t key[kKeyLen]) {
    // bucket centre normalised to the hash-table domain, HSFC3d
    // (preprocess.cc:196-203; the two-word key of THSFC3d's words 0-1)
    const double cr[3] = {((lo[0] + i * del) + (lo[0] + (i + 1) * del)) * 0.5,
                          ((lo[1] + j * del) + (lo[1] + (j + 1) * del)) * 0.5,
                          ((lo[2] + k * del) + (lo[2] + (k + 1) * del)) * 0.5};
    double normc[3];
    for (int l = 0; l < 3; l++) normc[l] = (cr[l] - lo[l]) / (hi[l] - lo[l]);
    uint32_t k3[3];
    psph::thsfc3d(normc, 0, k3);
    key[0] = k3[0];
    key[1] = k3[1];
  };
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        BucketRecord &r = bg[B.idx(i, j, k)];
        const int b[3] = {i, j, k};
        for (int l = 0; l < 2; l++) {
          r.xcoord[l] = lo[0] + (i + l) * del;
          r.ycoord[l] = lo[1] + (j + l) * del;
          r.zcoord[l] = lo[2] + (k + l) * del;
        }
        for (int l = 0; l < 4; l++) r.elev[l] = cfg->dom_lo[2];      // flat ground (preprocess.cc:160-162)
        for (int d = 0; d < 3; d++) {
          const double c0 = lo[d] + b[d] * del, c1 = lo[d] + (b[d] + 1) * del;
          r.bucket_index[2 * d] = face_of(c0, cfg->dom_hi[d], cfg->dom_lo[d]);
          r.bucket_index[2 * d + 1] = face_of(c1, cfg->dom_hi[d], cfg->dom_lo[d]);
        }
        r.buckettype = B.type[B.idx(i, j, k)];
        r.myproc = 0;
        key_at(i, j, k, r.key);
      }
  // neighbours (preprocess.cc:280-315)
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        BucketRecord &r = bg[B.idx(i, j, k)];
        if (!r.buckettype) continue;
        int nc = 0, npc = 0;
        for (int ii = i - 1; ii < i + 2; ii++)
          for (int jj = j - 1; jj < j + 2; jj++)
            for (int kk = k - 1; kk < k + 2; kk++) {
              if (ii > -1 && ii < nx && jj > -1 && jj < ny && kk > -1 && kk < nz) {
                const BucketRecord &o = bg[B.idx(ii, jj, kk)];
                for (int l = 0; l < kKeyLen; l++) r.neighs[nc++] = o.key[l];
                r.neigh_proc[npc++] = (ii == i && jj == j && kk == k) ? -2 : o.myproc;
              } else {
                r.neighs[nc++] = 0;
                r.neighs[nc++] = 0;
                r.neigh_proc[npc++] = -1;
              }
            }
      }
  // partition order (PartiHead::operator<, preprocess.h:59-110): a bucket of
  // the lowest layer sorts right after the bucket above it
  struct Part { uint32_t k0, k1, u; size_t q; };
  std::vector<Part> part;
  part.reserve(B.size());
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        const size_t q = B.idx(i, j, k);
        if (k == 0 && nz > 1) {
          uint32_t top[kKeyLen];
          key_at(i, j, 1, top);
          part.push_back({top[0], top[1], 1u, q});
        } else {
          part.push_back({bg[q].key[0], bg[q].key[1], 0u, q});
        }
      }
  std::stable_sort(part.begin(), part.end(), [](const Part &a, const Part &b) {
    if (a.k0 != b.k0) return a.k0 < b.k0;
    if (a.k1 != b.k1) return a.k1 < b.k1;
    return a.u < b.u;
  });
  std::vector<BucketRecord> out;
  std::vector<uint32_t> ptab;
  for (const Part &p : part) {
    const BucketRecord &r = bg[p.q];
    if (r.buckettype > 0) out.push_back(r);
    // the partition key is the bucket's own key (createfunky's table,
    // preprocess.cc:341-349), then the bucket's key again
    for (int l = 0; l < kKeyLen; l++) ptab.push_back(r.key[l]);
    for (int l = 0; l < kKeyLen; l++) ptab.push_back(r.key[l]);
  }
  H5Id fp(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (fp < 0) return fail(PSPH_EIO, "grid file: cannot create %s", path);
  const double ht[6] = {lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
  PSPH_TRY(write_1d(fp, "/hashtable_constants", H5T_NATIVE_DOUBLE, ht, 6));
  {
    H5Id t(bucket_file_type(), H5Tclose);
    PSPH_TRY(write_1d(fp, "/Buckets", t, out.data(), out.size()));
  }
  PSPH_TRY(write_1d(fp, "/partition_table", H5T_NATIVE_UINT, ptab.data(), ptab.size()));
  return PSPH_OK;
}
