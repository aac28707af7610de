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

}  // namespace psph
