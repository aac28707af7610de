// psph_internal.h -- shared declarations of the host-side library (not ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/psph.h"

namespace psph {

// Thread-local last error; returns `code` so callers can `return fail(...)`.
int fail(int code, const char *fmt, ...);

// Host copy of a particle set in SoA form (owning vectors).
struct HostParticles {
  std::vector<double> x, y, z, vx, vy, vz, svx, svy, svz, rho, e, m, h, p, c, T, gam, cp, xi;
  std::vector<uint32_t> k0, k1, k2;
  std::vector<int32_t> bc, phase, inv;
  size_t size() const { return x.size(); }
  void resize(size_t n);
  void reserve(size_t n);
  size_t push();                       // appends a zeroed particle, returns its index
  void permute(const std::vector<uint32_t> &perm);
  void to_soa(psph_soa *out, size_t first = 0) const;
  void from_soa(const psph_soa *in, size_t n);
};

// Hilbert particle key THSFC3d of a position normalised to [0,1]^3.
void thsfc3d(const double normc[3], uint32_t add_step, uint32_t key[3]);
// THashTable slot (thashtab.cc:209-222).
int64_t table_slot(uint32_t k0, uint32_t k1, int table_size, int dim);
// Permutation that puts particles in THashTable iteration order.
std::vector<uint32_t> canonical_order(const psph_config &cfg, const uint32_t *k0,
                                      const uint32_t *k1, const uint32_t *k2, size_t n);

// Closed-form EOS (particle.cc:528-604); cp is written only when lans != 0.
void eos(const psph_config &cfg, double rho, double e, double xi, double &p, double &c,
         double &T, double &gam, double &cp);

// Inlet for one step (setup_erupt.cc:389-670): appends new erupt ghosts.
// With rec, the per-column whole-period count q of fmod(t, t_each) and the
// spawn count n are appended (q, n) in scan order; with rep, they are taken
// from there instead (the schedule of a recorded run, for the adjoint and its
// finite-difference checks), and the spawn time is t - q t_each + dt - n t_each.
int inject_host(const psph_config &cfg, double time, double dt, int64_t step,
                HostParticles &out, std::vector<int32_t> *rec = nullptr,
                const int32_t *rep = nullptr, size_t rep_len = 0);

// Meteo::interpolate (meteo.cc:64-143) of property prop at height z0 (km).
double meteo_interp(const psph_meteo *m, double z0, int prop);
// Air lattice point l (spacing h_air, offset h_air / 2 inside each bucket of
// the extended domain, nl points per axis): exact position and the position
// with the seeded jitter of the synthetic inputs.
void lattice_point(const psph_config &cfg, uint32_t seed, double jitter, const int64_t l[3],
                   const int64_t nl[3], double exact[3], double pos[3]);
// Appends an air particle at pos in its initial state (initial_air and
// air_prop_meteo_based, sph_lib.cc:724-766, 831-872): smoothing length h,
// keys of add_step.
size_t push_air(const psph_config &cfg, const psph_meteo *met, HostParticles &P, const double pos[3],
                double h, int bc, int inv, uint32_t add_step);
// setup_erupt (setup_erupt.cc:43-386): drops every non-erupt particle inside
// the vent duct and adds the duct's erupt ghosts.
void setup_vent(const psph_config &cfg, HostParticles &P);

// ---- domain adaptation (ADJUST_DOMAIN, main.cc:256-304) -----------------
// Buckets: the grid preprocessor's cells of edge bucket_edge over the
// extended domain.  Occupancy bits of a bucket, from its particles: any
// particle, an involved one (involved == 2), a pressure ghost, a wall ghost.
enum : uint8_t { OCC_ANY = 1, OCC_INVOLVED = 2, OCC_PGHOST = 4, OCC_WGHOST = 8 };
struct DomainBuckets {
  int nb[3] = {0, 0, 0};
  std::vector<uint8_t> val;     // has_involved: bit 0 potentially involved, bit 1 involved
  std::vector<uint8_t> erupt;   // erupt_flag (setup_erupt.cc:221-253)
  std::vector<int8_t> type;     // bucket type (preprocess.cc:401-440)
  std::vector<uint8_t> sides;   // bit 2k / 2k+1: the bucket straddles dom's lower / upper face on axis k
  size_t size() const { return (size_t)nb[0] * nb[1] * nb[2]; }
  size_t idx(int x, int y, int z) const { return ((size_t)z * nb[1] + y) * nb[0] + x; }
};
enum { BK_UNDERGROUND = 1, BK_MIXED = 2, BK_OVERGROUND = 3, BK_PRESS_BC = 4 };
// Geometry, types and the initial has_involved (the bucket meets the initial
// domain idom, dataread.cc:326-457) of the buckets.
void domain_buckets(const psph_config &cfg, DomainBuckets &B);
// Bucket of a position, or -1 outside the extended domain.
int64_t bucket_of(const psph_config &cfg, const DomainBuckets &B, double x, double y, double z);
// scan_outside_layer (scan_outside_layer.cc:44-121): marks the potentially
// involved buckets that hold an involved particle as involved and returns how
// many of them border a non-empty, uninvolved bucket that is not underground.
int domain_scan(DomainBuckets &B, const std::vector<uint8_t> &occ);
// adapt_domain (adapt_domain.cc:41-194): for every non-empty uninvolved
// bucket that is neither underground nor on the outer boundary and borders an
// involved one, the conversion code of its particles (1 all, 2 the pressure
// ghosts, 3 those inside dom) in conv, and the bucket becomes potentially
// involved.  Returns the number of buckets converted.
int domain_convert_plan(DomainBuckets &B, const std::vector<uint8_t> &occ, std::vector<uint8_t> &conv);
// add_pressure_ghost (add_pressure_ghost.cc:37-316) then add_wall_ghost
// (add_wall_ghost.cc:45-293): the particles of the empty-of-ghost buckets the
// grown region now borders, appended to out (keys of add_step); occ is updated
// with them.
void domain_add_ghosts(const psph_config &cfg, const DomainBuckets &B, std::vector<uint8_t> &occ,
                       const psph_meteo *met, uint32_t seed, double jitter, uint32_t add_step,
                       HostParticles &out);
// The initial state of a run with domain adaptation (add_air.cc:491-813 on
// the initial domain's buckets, then the ghost layers and the vent).
int gen_plume_adaptive(const psph_config &cfg, const psph_meteo *met, uint32_t seed, double jitter,
                       HostParticles &P);

}  // namespace psph
