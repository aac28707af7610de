// psph_domain.cpp -- domain adaptation (ADJUST_DOMAIN) on the bucket grid:
// the host side of the grow-the-domain step (main.cc:256-304, 433-436).
//
// The reference keeps, per bucket of its background mesh, a has_involved bit
// field (bit 0 potentially involved, bit 1 involved, bucket.h:400-414) and
// grows the region of real particles when involved material reaches the
// outer layer of that region.  Here the buckets are a small dense array (a
// few thousand to tens of thousands of entries: cells of 8 h_air), so their
// decisions run on the host from a per-bucket occupancy that the GPU reduces
// from the particles each step; the particle updates (conversion to real,
// new ghost layers) run on the GPU or are appended to its arrays.

#include <cmath>
#include <cstring>
#include <vector>

#include "psph_internal.h"

namespace psph {

namespace {

// face of a coordinate against [lo, hi]: -1 below, 1 above, 0 on or inside
// (determine_face_type, preprocess.cc:386-399)
int face(double crd, double hi, double lo) { return crd < lo ? -1 : (crd > hi ? 1 : 0); }

// 26 neighbour offsets of a bucket
struct Offsets {
  int d[26][3];
  Offsets() {
    int n = 0;
    for (int z = -1; z <= 1; z++)
      for (int y = -1; y <= 1; y++)
        for (int x = -1; x <= 1; x++)
          if (x || y || z) { d[n][0] = x; d[n][1] = y; d[n][2] = z; n++; }
  }
};
const Offsets kNbr;

template <class F>
bool any_neighbour(const DomainBuckets &B, int x, int y, int z, F &&pred) {
  for (const auto &o : kNbr.d) {
    const int a = x + o[0], b = y + o[1], c = z + o[2];
    if (a < 0 || b < 0 || c < 0 || a >= B.nb[0] || b >= B.nb[1] || c >= B.nb[2]) continue;
    if (pred(B.idx(a, b, c))) return true;
  }
  return false;
}

bool ground_mixed(const DomainBuckets &B, size_t q) {
  // MIXED with its lower z face below the ground (is_onground_or_underground)
  return B.type[q] == BK_MIXED && (B.sides[q] & 16);
}

}  // namespace

void domain_buckets(const psph_config &cfg, DomainBuckets &B) {
  const double del = cfg.bucket_edge;
  for (int k = 0; k < 3; k++) B.nb[k] = (int)std::llround((cfg.ext_hi[k] - cfg.ext_lo[k]) / del);
  const size_t n = B.size();
  B.val.assign(n, 0);
  B.erupt.assign(n, 0);
  B.type.assign(n, 0);
  B.sides.assign(n, 0);
  // the vent duct's box (setup_erupt.cc:86-91): the vent disc over one
  // extended-domain bottom layer below the ground
  const double dlo[3] = {cfg.vent_x - cfg.vent_r, cfg.vent_y - cfg.vent_r, cfg.vent_z - 1.5 * del};
  const double dhi[3] = {cfg.vent_x + cfg.vent_r, cfg.vent_y + cfg.vent_r, cfg.vent_z};
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const int b[3] = {x, y, z};
        const size_t q = B.idx(x, y, z);
        int bt[6], flag[3], sum = 0;
        bool meets_idom = true, meets_duct = true;
        for (int k = 0; k < 3; k++) {
          const double lo = cfg.ext_lo[k] + b[k] * del, hi = cfg.ext_lo[k] + (b[k] + 1) * del;   // preprocess.cc:166-179
          bt[2 * k] = face(lo, cfg.dom_hi[k], cfg.dom_lo[k]);
          bt[2 * k + 1] = face(hi, cfg.dom_hi[k], cfg.dom_lo[k]);
          flag[k] = std::abs(bt[2 * k] + bt[2 * k + 1]);
          sum += flag[k];
          // the bucket against the initial domain (dataread.cc:52-75, 455-457)
          if (std::abs(face(lo, cfg.idom_hi[k], cfg.idom_lo[k]) + face(hi, cfg.idom_hi[k], cfg.idom_lo[k])) == 2)
            meets_idom = false;
          // the duct's box against the bucket (determine_erupt_buket, sph_lib.cc:1597-1620)
          if (std::abs(face(dlo[k], hi, lo) + face(dhi[k], hi, lo)) == 2) meets_duct = false;
        }
        int type;
        if (flag[0] == 2 || flag[1] == 2 || flag[2] == 2)                    // preprocess.cc:419-433
          type = (bt[4] == -1 && bt[5] == -1) ? BK_UNDERGROUND : BK_PRESS_BC;
        else if (flag[0] == 1 || flag[1] == 1 || flag[2] == 1)
          type = BK_MIXED;
        else
          type = sum == 0 ? BK_OVERGROUND : 0;
        B.type[q] = (int8_t)type;
        uint8_t sd = 0;
        for (int k = 0; k < 3; k++) {
          if (bt[2 * k] == -1) sd |= (uint8_t)(1u << (2 * k));
          if (bt[2 * k + 1] == 1) sd |= (uint8_t)(1u << (2 * k + 1));
        }
        B.sides[q] = sd;
        B.val[q] = meets_idom ? 1 : 0;
        B.erupt[q] = meets_duct ? 1 : 0;
      }
}

int64_t bucket_of(const psph_config &cfg, const DomainBuckets &B, double x, double y, double z) {
  const double p[3] = {x, y, z};
  int b[3];
  for (int k = 0; k < 3; k++) {
    const double f = std::floor((p[k] - cfg.ext_lo[k]) / cfg.bucket_edge);
    if (!(f >= 0 && f < B.nb[k])) return -1;
    b[k] = (int)f;
  }
  return (int64_t)B.idx(b[0], b[1], b[2]);
}

int domain_scan(DomainBuckets &B, const std::vector<uint8_t> &occ) {
  int adapt = 0;
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const size_t q = B.idx(x, y, z);
        if (B.val[q] != 1 || !(occ[q] & OCC_INVOLVED)) continue;
        B.val[q] = 3;
        // the outermost layer: a non-empty, uninvolved neighbour above ground
        if (any_neighbour(B, x, y, z, [&](size_t r) {
              return (occ[r] & OCC_ANY) && B.val[r] == 0 && B.type[r] != BK_UNDERGROUND;
            }))
          adapt++;
      }
  return adapt;
}

int domain_convert_plan(DomainBuckets &B, const std::vector<uint8_t> &occ, std::vector<uint8_t> &conv) {
  conv.assign(B.size(), 0);
  int n = 0;
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const size_t q = B.idx(x, y, z);
        if (B.val[q] || !(occ[q] & OCC_ANY) || B.type[q] == BK_UNDERGROUND || B.type[q] == BK_PRESS_BC) continue;
        if (!any_neighbour(B, x, y, z, [&](size_t r) { return (B.val[r] & 2) != 0; })) continue;
        conv[q] = B.type[q] == BK_MIXED ? (ground_mixed(B, q) ? 2 : 3) : 1;
        n++;
      }
  // the converted buckets are potentially involved from now on; the test above
  // reads only the involved bit of neighbours, which this does not change
  for (size_t q = 0; q < conv.size(); q++)
    if (conv[q]) B.val[q] = 1;
  return n;
}

namespace {

// visits the air lattice points of bucket q: fn(exact, jittered)
template <class F>
void bucket_points(const psph_config &cfg, const DomainBuckets &B, int x, int y, int z, uint32_t seed,
                   double jitter, F &&fn) {
  const int per = (int)std::llround(cfg.bucket_edge / cfg.h_air);
  const int64_t nl[3] = {(int64_t)B.nb[0] * per, (int64_t)B.nb[1] * per, (int64_t)B.nb[2] * per};
  for (int i = 0; i < per; i++)
    for (int j = 0; j < per; j++)
      for (int k = 0; k < per; k++) {
        const int64_t l[3] = {(int64_t)x * per + i, (int64_t)y * per + j, (int64_t)z * per + k};
        double exact[3], pos[3];
        lattice_point(cfg, seed, jitter, l, nl, exact, pos);
        fn(exact, pos);
      }
}

}  // namespace

void domain_add_ghosts(const psph_config &cfg, const DomainBuckets &B, std::vector<uint8_t> &occ,
                       const psph_meteo *met, uint32_t seed, double jitter, uint32_t add_step,
                       HostParticles &out) {
  const double ground = cfg.dom_lo[2], h = cfg.h_air;
  // add_pressure_ghost: uninvolved buckets without pressure ghosts, not in the
  // duct, next to a bucket with any involvement (add_pressure_ghost.cc:100-290)
  std::vector<size_t> filled;
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const size_t q = B.idx(x, y, z);
        if (B.val[q] || (occ[q] & OCC_PGHOST) || B.erupt[q]) continue;
        const int t = B.type[q];
        if (t != BK_MIXED && t != BK_OVERGROUND && t != BK_PRESS_BC) continue;
        if (!any_neighbour(B, x, y, z, [&](size_t r) { return B.val[r] != 0; })) continue;
        const bool above_only = ground_mixed(B, q);
        size_t before = out.size();
        bucket_points(cfg, B, x, y, z, seed, jitter, [&](const double *exact, const double *pos) {
          if (above_only && !(exact[2] > ground)) return;
          push_air(cfg, met, out, pos, h, PSPH_BC_PRESSURE_GHOST, 0, add_step);
        });
        if (out.size() > before) filled.push_back(q);
      }
  for (size_t q : filled) occ[q] |= OCC_ANY | OCC_PGHOST;
  // add_wall_ghost: the non-empty ground-layer buckets without wall ghosts,
  // not in the duct, below the ground, and the whole bucket under each
  // (add_wall_ghost.cc:131-285)
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const size_t q = B.idx(x, y, z);
        if (!ground_mixed(B, q) || !(occ[q] & OCC_ANY) || (occ[q] & OCC_WGHOST) || B.erupt[q]) continue;
        bucket_points(cfg, B, x, y, z, seed, jitter, [&](const double *exact, const double *pos) {
          if (exact[2] < ground) push_air(cfg, met, out, pos, h, PSPH_BC_WALL_GHOST, 0, add_step);
        });
        if (z > 0)
          bucket_points(cfg, B, x, y, z - 1, seed, jitter, [&](const double *, const double *pos) {
            push_air(cfg, met, out, pos, h, PSPH_BC_WALL_GHOST, 0, add_step);
          });
      }
}

int gen_plume_adaptive(const psph_config &cfg, const psph_meteo *met, uint32_t seed, double jitter,
                       HostParticles &P) {
  DomainBuckets B;
  domain_buckets(cfg, B);
  const double ground = cfg.dom_lo[2];
  // add_air on the buckets that meet the initial domain: real air, above the
  // ground in the ground layer (add_air.cc:566-800)
  for (int z = 0; z < B.nb[2]; z++)
    for (int y = 0; y < B.nb[1]; y++)
      for (int x = 0; x < B.nb[0]; x++) {
        const size_t q = B.idx(x, y, z);
        if (!B.val[q]) continue;
        const int t = B.type[q];
        if (t != BK_MIXED && t != BK_OVERGROUND) continue;
        const bool above_only = ground_mixed(B, q);
        bucket_points(cfg, B, x, y, z, seed, jitter, [&](const double *exact, const double *pos) {
          if (above_only && !(exact[2] > ground)) return;
          push_air(cfg, met, P, pos, cfg.h_air, PSPH_BC_REAL, 1, 0);
        });
      }
  // the ghost layers around them (main.cc:128-132), then the vent
  std::vector<uint8_t> occ(B.size(), 0);
  for (size_t i = 0; i < P.size(); i++) {
    const int64_t q = bucket_of(cfg, B, P.x[i], P.y[i], P.z[i]);
    if (q >= 0) occ[q] |= OCC_ANY;
  }
  domain_add_ghosts(cfg, B, occ, met, seed, jitter, 0, P);
  setup_vent(cfg, P);
  return PSPH_OK;
}

}  // namespace psph
