// psph_device.hip -- the per-step SPH particle path on one MI355X (gfx950).
//
// Data layout: structure-of-arrays in HBM, one array per particle field, all
// in "cell order" (particles sorted by the linear index of a uniform cell
// grid whose edge covers the largest kernel support).  Every step:
//   1. cell keys -> radix sort (key, index) -> gather every field (neighbours
//      are then contiguous: cell rows along x are contiguous index ranges);
//   2. targets (bc 100/99) are compacted and cut into "chunks" of <= 64
//      consecutive targets inside one (y,z) cell row;
//   3. pair passes: one 64-lane wavefront per chunk.  The chunk's candidate
//      range for each neighbouring (y,z) row is one contiguous slice; it is
//      staged 64 candidates at a time into LDS with coalesced loads (derived
//      per-candidate terms such as 1/rho and p/rho^2 are computed once, by the
//      staging lane), then every lane walks the tile with broadcast LDS reads.
//      No list is materialised; the reference's list membership rule is
//      evaluated per pair (same bucket, or |dx| <= 4.25 h_i per axis).
// No MFMA: there is no dense contraction.  Reductions are deterministic
// (min/max via integer atomics on positive doubles; sums via fixed-order
// block partials).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <vector>

#include "psph_internal.h"

using namespace psph;

// A failed HIP call also leaves the thread's last-error state set; it is
// cleared here so that the error is reported once, by the return code, and not
// again by the next unrelated check of another user of the runtime (torch).
#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      (void)hipGetLastError();                                                               \
      return fail(PSPH_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    }                                                                                        \
  } while (0)

#define PSPH_TRY(expr)            \
  do {                            \
    int rc_ = (expr);             \
    if (rc_) return rc_;          \
  } while (0)

namespace {

constexpr int kWave = 64;
constexpr uint32_t kXcds = 8;      // MI355X: 8 XCDs, each with its own L2
constexpr uint32_t kDead = 0xffffffffu;

// ---------------------------------------------------------------------------
// device-side error word: [0] = PSPH_E* code (0 = none), [1] = particle index
// ---------------------------------------------------------------------------
__device__ inline void raise_err(int *err, int code, uint32_t idx) {
  if (atomicCAS(err, 0, code) == 0) err[1] = (int)idx;
}

__device__ inline bool need_neigh(int bc) { return bc == 100 || bc == 99; }
__device__ inline bool is_real(int bc) { return bc == 100 || bc == 99; }
__device__ inline bool contr_dens(int pg, int bc) {
  return bc == 100 || bc == 99 || (pg >= 1 && bc == 1) || (pg >= 2 && bc == 0);
}
__device__ inline bool contr_vel(int bc) { return bc == 100 || bc == 99 || bc == 2; }

// The reference tests abs(s_k) > CUTOFF with s_k = dx_k / h rounded to
// nearest.  Exactly equivalent and division-free: with u = half an ulp of cut,
// fl(|dx| / h) > cut  <=>  |dx| / h > cut + u (ties go to cut, whose last
// mantissa bit is even for cut = 3)  <=>  |dx| - cut h > u h.  The FMA gives
// round(|dx| - cut h); that difference is a multiple of ~h 2^-53, so it cannot
// lie within half an ulp of u h without equalling it, and the comparison of
// the rounded value with the representable u h is exact.  Applied to
// max_k |dx_k| it decides the whole box.
struct CutTest {
  double cut, u;
  __device__ inline bool outside(double amax, double h) const { return fma(-cut, h, amax) > u * h; }
};

// Closed-form EOS (particle.cc:528-604).
struct EosConst { double ng0, cvs, cvg, cva, rg, ra; int lans; };
__device__ inline void eos_dev(const EosConst &k, double rho, double e, double xi, double &p,
                               double &c, double &T, double &gam, double *cp) {
  double na = 1 - xi, ng = k.ng0 * xi, ns = xi - ng;
  double Cvm = ns * k.cvs + ng * k.cvg + na * k.cva;
  double Rm = ng * k.rg + na * k.ra;
  double Rmg = Rm / (ng + na);
  gam = 1 + Rm / Cvm;
  double desmg = rho * (na + ng);
  p = Rmg * desmg * e / Cvm;
  c = sqrt(gam * p / desmg);
  T = e / Cvm;
  if (k.lans) *cp = Cvm;
}

// ---------------------------------------------------------------------------
// simple kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int *p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Block-wide min / max of one value per thread (blockDim = 256); the
// result is valid in thread 0.  One atomic per block keeps the global
// update off a single contended address.
template <bool MAX>
__device__ inline double block_ext(double v) {
  __shared__ double part[4];
  v = MAX ? wave_max(v) : wave_min(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < 4; w++) v = MAX ? fmax(v, part[w]) : fmin(v, part[w]);
  return v;
}

// dt candidate min h/c over need_neigh particles (timestep.cc:94-135);
// grid-stride, 256 threads per block.
__global__ __launch_bounds__(256) void k_dt_min(uint32_t n, const int32_t *bc, const double *h, const double *c,
                                                unsigned long long *out, int *err) {
  double v = 1.0e300;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    if (!need_neigh(bc[i])) continue;
    const double ci = c[i];
    if (!(ci > 0)) raise_err(err, PSPH_EINVARIANT, i);
    else v = fmin(v, h[i] / ci);
  }
  v = block_ext<false>(v);
  if (threadIdx.x == 0 && v < 1.0e300) atomicMin(out, (unsigned long long)__double_as_longlong(v));
}

// max h over live particles (the cell edge must cover every pair support)
__global__ __launch_bounds__(256) void k_h_max(uint32_t n, const int32_t *bc, const double *h, unsigned long long *out) {
  double v = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (bc[i] >= 0) v = fmax(v, h[i]);
  v = block_ext<true>(v);
  if (threadIdx.x == 0 && v > 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// Cell grid of the neighbour search.  Particles are sorted by
// (cell, Morton code of their position inside the cell at fbits bits per
// axis), so every cell is a contiguous index range and every run of 64
// consecutive particles of a cell (a "tile") is a compact cluster.
struct GridDesc {
  double lo[3];
  double inv_e;      // 1 / cell edge; 0: one cell holds everything
  int nx, ny, nz;
  uint32_t ncells;
  int fbits;
};

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// floor((v - lo) * inv_e) clamped to [0, n - 1] without overflowing the int conversion
__device__ inline int cell_of(double v, double lo, double inv_e, int n) {
  const double f = floor((v - lo) * inv_e);
  return (int)fmin(fmax(f, 0.0), (double)(n - 1));
}

__global__ void k_cellkey(uint32_t n, const double *x, const double *y, const double *z,
                          const int32_t *bc, GridDesc g, uint32_t *key, uint32_t *idx) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  idx[i] = i;
  const int sh = 3 * g.fbits;
  if (bc[i] < 0) { key[i] = g.ncells << sh; return; }   // removed: sorts past every cell
  const double p[3] = {x[i], y[i], z[i]};
  const int nn[3] = {g.nx, g.ny, g.nz};
  int cc[3];
  uint32_t m = 0;
  const double nf = (double)(1 << g.fbits);
  for (int d = 0; d < 3; d++) {
    const double f = (p[d] - g.lo[d]) * g.inv_e;
    cc[d] = (int)fmin(fmax(floor(f), 0.0), (double)(nn[d] - 1));
    const int u = (int)fmin(fmax(floor((f - cc[d]) * nf), 0.0), nf - 1);
    for (int b = 0; b < g.fbits; b++) m |= (uint32_t)((u >> b) & 1) << (3 * b + d);
  }
  key[i] = ((((uint32_t)cc[2] * g.ny + cc[1]) * g.nx + cc[0]) << sh) | m;
}

template <class T>
__global__ void k_gather(uint32_t n, const uint32_t *perm, const T *__restrict__ src, T *__restrict__ dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// cell_start[c] = first sorted index whose cell (key >> sh) >= c, c in [0, ncells]
__global__ void k_cell_start(uint32_t n, const uint32_t *key, int sh, uint32_t ncells, uint32_t *cs) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  uint32_t prev = (i == 0) ? 0u : min(key[i - 1] >> sh, ncells) + 1u;
  uint32_t cur = (i == n) ? ncells + 1u : min(key[i] >> sh, ncells) + 1u;
  for (uint32_t c = prev; c < cur; c++) cs[c] = i;
}

// tiles per cell (64 consecutive particles each); tcnt[ncells] = 0
__global__ void k_tile_counts(uint32_t ncells, const uint32_t *cs, uint32_t *tcnt) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncells) return;
  tcnt[c] = c == ncells ? 0u : (cs[c + 1] - cs[c] + kWave - 1) / kWave;
}

// Bounding box of every tile, rounded outwards to float: [lo, hi] per tile.
// One wavefront per tile; the tile's cell is found by bisection of toff.
__global__ __launch_bounds__(kWave) void k_tile_boxes(uint32_t ncells, const uint32_t *cs, const uint32_t *toff,
                                                      const double *x, const double *y, const double *z,
                                                      float4 *tbox) {
  const uint32_t t = blockIdx.x;
  if (t >= toff[ncells]) return;
  uint32_t lo = 0, hi = ncells;              // toff[lo] <= t < toff[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (toff[mid] <= t) lo = mid; else hi = mid;
  }
  const uint32_t base = cs[lo] + kWave * (t - toff[lo]);
  const uint32_t end = min(cs[lo + 1], base + kWave);
  const uint32_t j = base + threadIdx.x < end ? base + threadIdx.x : base;
  const double px = x[j], py = y[j], pz = z[j];
  const double ax = wave_min(px), ay = wave_min(py), az = wave_min(pz);
  const double bx = wave_max(px), by = wave_max(py), bz = wave_max(pz);
  if (threadIdx.x == 0) {
    tbox[2 * t] = make_float4(__double2float_rd(ax), __double2float_rd(ay), __double2float_rd(az), 0.f);
    tbox[2 * t + 1] = make_float4(__double2float_ru(bx), __double2float_ru(by), __double2float_ru(bz), 0.f);
  }
}

// Reference bucket (preprocessor cells of edge 8 h_air, preprocess.cc:88-133)
// as packed coordinates: x in bits 0-9, y in bits 11-20, z in bits 22-31 (at
// most 1023 buckets per axis; kDead outside the grid).  A list holds the own
// bucket's particles and, from the 26 adjacent buckets, those within the
// list box (neighs.cc:37-211).
constexpr int kRefbMaxPerAxis = 1023;
__device__ inline uint32_t refb_pack(uint32_t x, uint32_t y, uint32_t z) { return x | (y << 11) | (z << 22); }
// buckets at most one apart on every axis: per field 1024 + a - b + 1 (the
// guard bits 10 and 21 stop borrows between fields), (a - b + 1) mod 1024 for z
__device__ inline bool refb_adjacent(uint32_t a, uint32_t b) {
  if (a == kDead || b == kDead) return false;
  const uint32_t d = ((a | 0x00200400u) - b) + 0x00400801u;
  return ((d & 0x7ffu) - 1024u <= 2u) & (((d >> 11) & 0x7ffu) - 1024u <= 2u) & ((d >> 22) <= 2u);
}

// the smoothing length at list time, and the reference bucket
__global__ void k_refb(uint32_t n, const double *x, const double *y, const double *z, const double *h,
                       double lx, double ly, double lz, double del, int bx, int by, int bz,
                       uint32_t *refb, double *h_list) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double fx = floor((x[i] - lx) / del), fy = floor((y[i] - ly) / del), fz = floor((z[i] - lz) / del);
  uint32_t r = kDead;
  if (fx >= 0 && fx < bx && fy >= 0 && fy < by && fz >= 0 && fz < bz)
    r = refb_pack((uint32_t)fx, (uint32_t)fy, (uint32_t)fz);
  refb[i] = r;
  h_list[i] = h[i];
}

__global__ void k_target_flags(uint32_t n, const int32_t *bc, uint32_t *flags) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) flags[i] = (i < n && need_neigh(bc[i])) ? 1u : 0u;
}

__global__ void k_targets(uint32_t n, const uint32_t *flags, const uint32_t *tpos, uint32_t *targets) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) targets[tpos[i]] = i;
}

// chunks: up to csz consecutive targets of one cell
__global__ void k_cell_chunks(uint32_t ncells, const uint32_t *cs, const uint32_t *tpos, uint32_t csz,
                              uint32_t *cell_chunks) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > ncells) return;
  if (r == ncells) { cell_chunks[r] = 0; return; }
  uint32_t a = tpos[cs[r]], b = tpos[cs[r + 1]];
  cell_chunks[r] = (b - a + csz - 1) / csz;
}

__global__ void k_chunks(uint32_t ncells, const uint32_t *cs, const uint32_t *tpos, const uint32_t *chunk_off,
                         uint32_t csz, uint32_t *cfirst, uint32_t *ccount, uint32_t *nchunks) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ncells) {
    if (r == ncells) *nchunks = chunk_off[ncells];
    return;
  }
  uint32_t a = tpos[cs[r]], b = tpos[cs[r + 1]];
  uint32_t o = chunk_off[r];
  for (uint32_t t = a; t < b; t += csz) {
    cfirst[o] = t;
    ccount[o] = min(csz, b - t);
    o++;
  }
}

// ---------------------------------------------------------------------------
// Pair-pass skeleton
// ---------------------------------------------------------------------------
// Cells and tiles a wavefront's chunk of particles must visit: the chunk's
// bounding box grown by `reach` (the largest per-axis pair distance any
// predicate of the pass can accept) selects a block of cells, and of their
// tiles only those whose box meets the grown box.  Every candidate within
// reach of a chunk particle lies in such a tile.
struct Traversal {
  const uint32_t *cs, *toff;
  const float4 *tbox;
  GridDesc g;
  double reach;
};

struct ChunkBox {
  int c0[3], c1[3];
  float lo[3], hi[3];
};

__device__ inline ChunkBox chunk_box(const Traversal &tr, const double *p) {
  ChunkBox b;
  const int nn[3] = {tr.g.nx, tr.g.ny, tr.g.nz};
  for (int d = 0; d < 3; d++) {
    const double lo = wave_min(p[d]) - tr.reach, hi = wave_max(p[d]) + tr.reach;
    b.lo[d] = __double2float_rd(lo);
    b.hi[d] = __double2float_ru(hi);
    if (tr.g.inv_e > 0) {
      b.c0[d] = cell_of(lo, tr.g.lo[d], tr.g.inv_e, nn[d]);
      b.c1[d] = cell_of(hi, tr.g.lo[d], tr.g.inv_e, nn[d]);
    } else {
      b.c0[d] = b.c1[d] = 0;
    }
  }
  return b;
}

__device__ inline bool tile_meets(const ChunkBox &b, const float4 *tbox, uint32_t t) {
  const float4 lo = tbox[2 * t], hi = tbox[2 * t + 1];
  return !(lo.x > b.hi[0] || hi.x < b.lo[0] || lo.y > b.hi[1] || hi.y < b.lo[1] || lo.z > b.hi[2] || hi.z < b.lo[2]);
}

struct PairCommon {
  const double *x, *y, *z;     // positions of the predicates (cells, boxes, tests)
  const double *ax, *ay, *az;  // positions of the pair arithmetic (the same arrays
                               // except in a frozen replay, Tape::xrec)
  const uint32_t *refb;
  const double *h_list;
  const int32_t *bc;
  uint32_t *gflag;            // optional: marks boundary ghosts (bc 1, 2) met as neighbours
  Traversal tr;
  const uint32_t *targets;
  const uint32_t *cfirst, *ccount, *nchunks;
  int list_all;
  int list_adj;               // some list box may reach past the adjacent buckets
  double list_factor;
  double del;                 // reference bucket edge
  double ext_lo[3];           // origin of the reference buckets
  CutTest cut;
  int *err;
};

// The region a chunk's neighbour lists can reach (neighs.cc:37-211): per
// target its own reference bucket and, within the adjacent buckets, its list
// box; the chunk's box is their union, with a margin for the rounding of the
// bucket index (1e-9 of a bucket).  It is often much smaller than the
// traversal's reach, which has to bound the largest list box of all.
__device__ inline ChunkBox chunk_list_box(const PairCommon &pc, const double *p, double lf, uint32_t rb) {
  ChunkBox b;
  const int nn[3] = {pc.tr.g.nx, pc.tr.g.ny, pc.tr.g.nz};
  const uint32_t bk[3] = {rb & 1023u, (rb >> 11) & 1023u, rb >> 22};
  const double m = 1e-9 * pc.del;
  for (int d = 0; d < 3; d++) {
    double lo, hi;
    if (rb == kDead || pc.list_all) {
      lo = p[d] - pc.tr.reach;
      hi = p[d] + pc.tr.reach;
    } else {
      const double blo = pc.ext_lo[d] + bk[d] * pc.del, bhi = blo + pc.del;
      lo = fmin(fmax(p[d] - lf, blo - pc.del), blo) - m;
      hi = fmax(fmin(p[d] + lf, bhi + pc.del), bhi) + m;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    b.lo[d] = __double2float_rd(lo);
    b.hi[d] = __double2float_ru(hi);
    if (pc.tr.g.inv_e > 0) {
      b.c0[d] = cell_of(lo, pc.tr.g.lo[d], pc.tr.g.inv_e, nn[d]);
      b.c1[d] = cell_of(hi, pc.tr.g.lo[d], pc.tr.g.inv_e, nn[d]);
    } else {
      b.c0[d] = b.c1[d] = 0;
    }
  }
  return b;
}

// Work split of a wavefront: kChunk targets (consecutive in the cell's fine
// Morton order, so they sit within ~2 particle spacings of each other), each
// served by kSub lanes that take every kSub-th candidate of a staged tile.
// The wave streams the union of its targets' neighbourhoods; with so compact
// a chunk that union is little more than one neighbourhood, and a lane's hit
// count per tile is bounded by 64 / kSub, so lanes rarely idle in phase 2.
// The kSub partial sums of a target are combined by a fixed shuffle tree.
#ifndef PSPH_SUB_LANES
#define PSPH_SUB_LANES 2
#endif
constexpr int kSub = PSPH_SUB_LANES;
constexpr double kFar = 1.0e300;
constexpr int kChunk = kWave / kSub;

// Each pass supplies: Tile (LDS SoA), Tgt (registers), stage(), pair(),
// combine() (sums a target's sub-lane partials), finish().
template <int D, class P>
__global__ __launch_bounds__(kWave) void k_pairs(PairCommon pc, typename P::Args a) {
  __shared__ double sx[kWave], sy[kWave], sz[kWave];
  __shared__ uint32_t srefb[kWave];
  __shared__ uint8_t sgh[kWave];   // boundary ghost (bc 1, 2), one byte each
  __shared__ typename P::Tile tile;
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so
  // XCD x runs workgroups x, x+8, ...; give it the contiguous chunk range
  // [x*per, (x+1)*per) so neighbouring chunks share that XCD's L2.
  const uint32_t nch = *pc.nchunks;
  const uint32_t per = (nch + kXcds - 1) / kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  if (local >= per) return;
  const uint32_t chunk = (blockIdx.x % kXcds) * per + local;
  if (chunk >= nch) return;
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const int lane = threadIdx.x;
  const uint32_t slot = (uint32_t)lane / P::kSub, sub = (uint32_t)lane % P::kSub;
  const bool active = slot < cc;
  using Mask = typename std::conditional<(kWave / P::kSub > 32), uint64_t, uint32_t>::type;
  const uint32_t i = pc.targets[cf + (active ? slot : 0)];

  double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const uint32_t rb_i = pc.refb[i];
  const double lf_i = pc.list_factor * pc.h_list[i];
  typename P::Tgt t;
  P::load_target(a, i, t);
  // a chunk none of whose targets wants the pass (converged in an earlier
  // iteration of the adaptive h) skips the traversal
  if (!__any(active && P::want(t))) {
    if (active && sub == 0) P::finish(a, t, i, pc.err);
    return;
  }
  // the passes run here (adaptive h, list materialisation) accept list
  // members only: the chunk visits the region its lists can reach
  const ChunkBox box = chunk_list_box(pc, xi, lf_i, rb_i);
  const GridDesc &g = pc.tr.g;
  P::init(tile, lane);
  __syncthreads();

  for (int pass = 0; pass < P::npasses(a); pass++) {
    P::begin_pass(a, t, pass);
    for (int cz = box.c0[2]; cz <= box.c1[2]; cz++)
      for (int cy = box.c0[1]; cy <= box.c1[1]; cy++)
        for (int cx = box.c0[0]; cx <= box.c1[0]; cx++) {
          const uint32_t cell = ((uint32_t)cz * g.ny + cy) * g.nx + cx;
          const uint32_t lo = pc.tr.cs[cell], hi = pc.tr.cs[cell + 1];
          uint32_t tl = pc.tr.toff[cell];
          for (uint32_t base = lo; base < hi; base += kWave, tl++) {
            if (!tile_meets(box, pc.tr.tbox, tl)) continue;
            const uint32_t j = base + lane;
            if (j < hi) {
              sx[lane] = pc.x[j]; sy[lane] = pc.y[j]; sz[lane] = pc.z[j];
              srefb[lane] = pc.refb[j];
              if (pc.gflag) { const int b = pc.bc[j]; sgh[lane] = b == 1 || b == 2; }
              P::stage(a, j, tile, lane);
            } else {
              // empty slot: a position no predicate accepts (|dx| > every support)
              sx[lane] = sy[lane] = sz[lane] = kFar;
              srefb[lane] = ~0u;
              sgh[lane] = 0;
            }
            __syncthreads();
            const uint32_t cnt = min((uint32_t)kWave, hi - base);
            if (active && P::want(t)) {
              // phase 1: exact FP64 predicate on this lane's share of the tile,
              // without branches (the same trip count on every lane, q < 64)
              Mask mask = 0;
              const uint32_t nb = (cnt + P::kSub - 1) / P::kSub;
#pragma unroll 4
              for (uint32_t b = 0; b < nb; b++) {
                const uint32_t q = sub + b * P::kSub;
                double amax = fabs(xi[0] - sx[q]);
                if (D == 3) amax = fmax(amax, fmax(fabs(xi[1] - sy[q]), fabs(xi[2] - sz[q])));
                // inside the kernel box (exact, see CutTest), list membership
                // (neighs.cc:37-211: same bucket or 4.25 h_i box), in_support
                // and the pass's own mask
                bool in = !pc.cut.outside(amax, P::hker(a, t, tile, q));
                in &= (pc.list_all != 0) | (srefb[q] == rb_i) |
                      ((amax <= lf_i) & ((pc.list_adj == 0) || refb_adjacent(srefb[q], rb_i)));
                in &= (q < cnt) & (amax <= t.supp) & P::extra(a, t, tile, q, base + q, i);
                mask |= (Mask)in << b;
              }
              // phase 2: each lane walks its own hits
              while (mask) {
                const uint32_t b = (uint32_t)__builtin_ctzll((uint64_t)mask);
                mask &= mask - 1;
                const uint32_t q = sub + b * P::kSub;
                double dx[3];
                dx[0] = xi[0] - sx[q];
                double amax = fabs(dx[0]);
                if (D == 3) {
                  dx[1] = xi[1] - sy[q]; dx[2] = xi[2] - sz[q];
                  amax = fmax(amax, fmax(fabs(dx[1]), fabs(dx[2])));
                }
                P::pair(a, t, tile, q, base + q, i, dx, amax);
                if (pc.gflag && sgh[q]) pc.gflag[base + q] = 1u;
              }
            }
            __syncthreads();
          }
        }
    if constexpr (P::kSub > 1) P::combine(t);
    P::end_pass(a, t, pass, i, active);
  }
  if (active && sub == 0) P::finish(a, t, i, pc.err);
}

// ---------------------------------------------------------------------------
// Visit lists.  The RHS, density and filter passes of a step see the same
// positions, smoothing lengths and list-time data, so the exact candidate
// scan is done once per step: for every chunk, the tiles that meet its grown
// box and hold a pair of any of the three passes, each with one 32-bit hit
// mask per lane (bit b <-> candidate sub + b kSub of the lane's target).  The
// mask is a superset: list membership and in_support are exact, the kernel
// box is widened to cut * max(h_ij, h_i) (the union of the three passes'
// boxes) with a relative margin of 1e-12; each pass re-tests its own exact box
// and masks on the set bits only.
// ---------------------------------------------------------------------------
static_assert(kWave / kSub <= 32, "visit masks hold one bit per candidate of a lane's share");

struct VisitList {
  const uint32_t *off;      // first visit of each chunk
  const uint32_t *num;      // visits written for each chunk
  const uint2 *ent;         // (first particle, particle count) of the visited tile
  const uint32_t *mask;     // per visit, kWave lane masks
};

// tiles that meet each chunk's grown box (an upper bound on its visits);
// one wavefront per chunk, chunk = blockIdx.x, zero past the last chunk; the
// lanes take the box's cells in turn
__device__ inline uint32_t count_meeting_tiles(const Traversal &tr, const ChunkBox &box) {
  const GridDesc &g = tr.g;
  const int bx = box.c1[0] - box.c0[0] + 1, by = box.c1[1] - box.c0[1] + 1, bz = box.c1[2] - box.c0[2] + 1;
  const int nbox = bx * by * bz;
  uint32_t cnt = 0;
  for (int k = threadIdx.x; k < nbox; k += kWave) {
    const int cx = box.c0[0] + k % bx, cy = box.c0[1] + (k / bx) % by, cz = box.c0[2] + k / (bx * by);
    const uint32_t cell = ((uint32_t)cz * g.ny + cy) * g.nx + cx;
    const uint32_t t0 = tr.toff[cell], t1 = tr.toff[cell + 1];
    for (uint32_t t = t0; t < t1; t++) cnt += tile_meets(box, tr.tbox, t) ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += (uint32_t)__shfl_xor((int)cnt, o);
  return cnt;
}

// The chunk's grown box for the visit lists: per target the widest kernel
// box of the three passes, cutx max((h_i + h_max) / 2, h_i), with a relative
// margin of 1e-9 (the traversal's reach bounds it with h_i = h_max).
__device__ inline ChunkBox chunk_pass_box(const Traversal &tr, const double *p, double h, double cutx,
                                          double h_max) {
  ChunkBox b;
  const int nn[3] = {tr.g.nx, tr.g.ny, tr.g.nz};
  const double r = cutx * fmax(0.5 * h + 0.5 * h_max, h) * (1 + 1e-9);
  for (int d = 0; d < 3; d++) {
    const double lo = wave_min(p[d] - r), hi = wave_max(p[d] + r);
    b.lo[d] = __double2float_rd(lo);
    b.hi[d] = __double2float_ru(hi);
    if (tr.g.inv_e > 0) {
      b.c0[d] = cell_of(lo, tr.g.lo[d], tr.g.inv_e, nn[d]);
      b.c1[d] = cell_of(hi, tr.g.lo[d], tr.g.inv_e, nn[d]);
    } else {
      b.c0[d] = b.c1[d] = 0;
    }
  }
  return b;
}

__global__ __launch_bounds__(kWave) void k_visit_count(PairCommon pc, const double *h, double cutx, double h_max,
                                                       uint32_t *vcount) {
  const uint32_t chunk = blockIdx.x;
  const uint32_t nch = *pc.nchunks;
  if (chunk >= nch) {
    if (threadIdx.x == 0) vcount[chunk] = 0;
    return;
  }
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const uint32_t slot = (uint32_t)threadIdx.x / kSub;
  const uint32_t i = pc.targets[cf + (slot < cc ? slot : 0)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const ChunkBox box = chunk_pass_box(pc.tr, xi, h[i], cutx, h_max);
  const uint32_t cnt = count_meeting_tiles(pc.tr, box);
  if (threadIdx.x == 0) vcount[chunk] = cnt;
}

// The scan is transposed: lane l holds candidate c(l) = 2 (l mod 32) + l / 32
// of the tile (so the ballot of a test over the 64 lanes is, in its low and
// high halves, exactly the masks of a target's sub-lanes 0 and 1), and the
// wave loops over the chunk's targets.  A target is skipped for a tile when
// its widened box misses the tile's box (one vector test for all 32 targets,
// then a scalar loop over the survivors), which drops about half of the
// (target, tile) combinations of a chunk's grown box.
struct VisitTgt {
  double x, y, z, chh;       // position, cutx h / 2
  double ch, lf, supp, wide; // cutx h, list box, in_support box, widest union box
};

// m &= ~(1 << b) for a wave-uniform 64-bit mask in scalar registers: one
// s_bitset0_b64 (the and-not of a shifted one takes two scalar instructions)
__device__ inline void clear_bit_u(uint64_t &m, uint32_t b) {
  asm volatile("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(b));
}

// v = value in lane `lane` only (value and lane wave-uniform): one
// v_writelane_b32, which ignores EXEC.  With the value in a scalar register
// the lane select goes through M0 (one scalar operand per VALU instruction).
// M0 is written by a scalar move, so the lane select has no VALU-write
// hazard; the asm touches registers only (M0 declared clobbered).  lane < 64.
__device__ inline void set_lane(uint32_t &v, uint32_t value, uint32_t lane) {
  asm volatile("s_mov_b32 m0, %2\n\tv_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(value), "s"(lane) : "m0");
}

// FAST: for every target of the chunk the widened box lies inside both its
// list box (4.25 h at list time) and its in_support box (CUTOFF2 h), and is
// narrower than a bucket (so it reaches the adjacent buckets only): those
// tests cannot fail for a candidate inside the widened box.
template <int D>
__global__ __launch_bounds__(kWave) void k_visit_masks(PairCommon pc, const double *h, double cutoff2, double cutx,
                                                       double h_max, const uint32_t *voff, uint32_t *vnum,
                                                       uint2 *vent, uint32_t *vmask) {
  static_assert(kSub == 2, "the transposed scan maps the ballot halves onto two sub-lanes");
  __shared__ VisitTgt tg[kChunk];
  __shared__ uint32_t trb[kChunk];
  const uint32_t nch = *pc.nchunks;
  const uint32_t per = (nch + kXcds - 1) / kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  if (local >= per) return;
  const uint32_t chunk = (blockIdx.x % kXcds) * per + local;
  if (chunk >= nch) return;
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const int lane = threadIdx.x;
  // lanes 0..cc-1 own one target each: its record in LDS, its float box in registers
  const bool owner = (uint32_t)lane < cc;
  const uint32_t i = pc.targets[cf + (owner ? (uint32_t)lane : 0u)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const double hi = h[i], hh_i = 0.5 * hi;
  const double lf_i = pc.list_factor * pc.h_list[i], supp_i = cutoff2 * hi;
  const double widest = cutx * fmax(hh_i + 0.5 * h_max, hi);
  if (owner) {
    tg[lane] = VisitTgt{xi[0], xi[1], xi[2], cutx * hh_i, cutx * hi, lf_i, supp_i, widest};
    trb[lane] = pc.refb[i];
  }
  float blo[3], bhi[3];
  for (int d = 0; d < 3; d++) {
    blo[d] = D == 3 || d == 0 ? __double2float_rd(xi[d] - widest) : -INFINITY;
    bhi[d] = D == 3 || d == 0 ? __double2float_ru(xi[d] + widest) : INFINITY;
  }
  const bool fast = __all(!owner || (widest <= supp_i &&
                                     (pc.list_all != 0 || (widest <= lf_i && widest * (1 + 1e-12) < pc.del))));
  // the chunk's grown box (same as the passes': every lane's target counts)
  const uint32_t slot = (uint32_t)lane / kSub;
  const uint32_t ib = pc.targets[cf + (slot < cc ? slot : 0u)];
  const double xb[3] = {pc.x[ib], pc.y[ib], pc.z[ib]};
  const ChunkBox box = chunk_pass_box(pc.tr, xb, h[ib], cutx, h_max);
  __syncthreads();
  const GridDesc &g = pc.tr.g;
  const uint32_t cq = 2u * ((uint32_t)lane & 31u) + ((uint32_t)lane >> 5);   // this lane's candidate
  const uint32_t w0 = voff[chunk];
  uint32_t w = w0;
  for (int cz = box.c0[2]; cz <= box.c1[2]; cz++)
    for (int cy = box.c0[1]; cy <= box.c1[1]; cy++)
      for (int cx = box.c0[0]; cx <= box.c1[0]; cx++) {
        const uint32_t cell = ((uint32_t)cz * g.ny + cy) * g.nx + cx;
        const uint32_t lo = pc.tr.cs[cell], hi_ = pc.tr.cs[cell + 1];
        uint32_t tl = pc.tr.toff[cell];
        for (uint32_t base = lo; base < hi_; base += kWave, tl++) {
          if (!tile_meets(box, pc.tr.tbox, tl)) continue;
          const uint32_t cnt = min((uint32_t)kWave, hi_ - base);
          // targets whose widened box meets the tile's box
          const float4 tlo = pc.tr.tbox[2 * tl], thi = pc.tr.tbox[2 * tl + 1];
          bool meets = owner && !(tlo.x > bhi[0] || thi.x < blo[0]);
          if (D == 3) meets = meets && !(tlo.y > bhi[1] || thi.y < blo[1] || tlo.z > bhi[2] || thi.z < blo[2]);
          uint64_t tm = __ballot(meets);
          if (!tm) continue;
          const bool valid = cq < cnt;
          const uint32_t j = base + (valid ? cq : 0u);
          const double cxp = valid ? pc.x[j] : kFar;
          const double cyp = valid ? pc.y[j] : kFar;
          const double czp = valid ? pc.z[j] : kFar;
          const double chh = valid ? cutx * (0.5 * h[j]) : 0.0;
          const uint32_t crb = valid ? pc.refb[j] : ~0u;
          uint32_t mask = 0;
          // the same loop with and without the list and support tests (the
          // wave-uniform choice taken outside it)
          auto scan = [&](auto fast_c) {
            while (tm) {
              const uint32_t t = (uint32_t)__builtin_ctzll(tm);
              clear_bit_u(tm, t);
              const VisitTgt T = tg[t];
              double amax = fabs(T.x - cxp);
              if (D == 3) amax = fmax(amax, fmax(fabs(T.y - cyp), fabs(T.z - czp)));
              // amax <= cutx max(h_ij, h_t), with cutx h_ij ~ cutx hh_t + cutx hh_j
              // (the 1e-12 margin of cutx covers the rounding of the split sum,
              // taken as amax - cutx hh_j <= cutx hh_t)
              uint64_t bal;
              if constexpr (decltype(fast_c)::value) {
                // one ballot per compare: each compare writes its lane mask directly
                bal = __ballot(amax - chh <= T.chh) | __ballot(amax <= T.ch);
              } else {
                bool in = (amax - chh <= T.chh) | (amax <= T.ch);
                in &= (pc.list_all != 0) | (crb == trb[t]) |
                      ((amax <= T.lf) & ((pc.list_adj == 0) || refb_adjacent(crb, trb[t])));
                in &= amax <= T.supp;
                bal = __ballot(in);
              }
              // lanes 2t and 2t+1 (target t's sub-lanes) take the two halves:
              // t and the ballot are wave-uniform (scalar registers), so each
              // half is written into its one lane
              set_lane(mask, (uint32_t)bal, 2u * t);
              set_lane(mask, (uint32_t)(bal >> 32), 2u * t + 1u);
            }
          };
          if (fast) scan(std::true_type{}); else scan(std::false_type{});
          if (__ballot(mask != 0)) {
            if (lane == 0) vent[w] = make_uint2(base, cnt);
            __builtin_nontemporal_store(mask, vmask + (size_t)w * kWave + lane);
            w++;
          }
        }
      }
  if (lane == 0) vnum[chunk] = w - w0;
}

// A pair pass over the visit lists: stage each visited tile, test the pass's
// exact kernel box and masks on the set bits, and run the pair arithmetic.
struct NoArith {};
struct ArithPos { double x[kWave], y[kWave], z[kWave]; };

// FROZEN (a frozen replay): the predicates read pc.x (the recorded run's
// positions) and the pair arithmetic pc.ax; otherwise both are pc.x.
template <int D, class P, bool FROZEN = false>
__global__ __launch_bounds__(kWave) void k_pairs_v(PairCommon pc, VisitList vl, typename P::Args a) {
  __shared__ double sx[kWave], sy[kWave], sz[kWave];
  __shared__ uint8_t sgh[kWave];   // boundary ghost (bc 1, 2), one byte each
  __shared__ typename P::Tile tile;
  __shared__ typename std::conditional<FROZEN, ArithPos, NoArith>::type sa;
  const uint32_t nch = *pc.nchunks;
  const uint32_t per = (nch + kXcds - 1) / kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  if (local >= per) return;
  const uint32_t chunk = (blockIdx.x % kXcds) * per + local;
  if (chunk >= nch) return;
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const int lane = threadIdx.x;
  const uint32_t slot = (uint32_t)lane / P::kSub, sub = (uint32_t)lane % P::kSub;
  const bool active = slot < cc;
  const uint32_t i = pc.targets[cf + (active ? slot : 0)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  double xa[3] = {xi[0], xi[1], xi[2]};
  if constexpr (FROZEN) { xa[0] = pc.ax[i]; xa[1] = pc.ay[i]; xa[2] = pc.az[i]; }
  typename P::Tgt t;
  P::load_target(a, i, t);
  P::init(tile, lane);
  __syncthreads();
  P::begin_pass(a, t, 0);
  const uint32_t v0 = vl.off[chunk], nv = vl.num[chunk];
  for (uint32_t v = 0; v < nv; v++) {
    const uint2 e = vl.ent[v0 + v];
    const uint32_t base = e.x, cnt = e.y;
    if ((uint32_t)lane < cnt) {
      const uint32_t j = base + lane;
      sx[lane] = pc.x[j]; sy[lane] = pc.y[j]; sz[lane] = pc.z[j];
      if constexpr (FROZEN) { sa.x[lane] = pc.ax[j]; sa.y[lane] = pc.ay[j]; sa.z[lane] = pc.az[j]; }
      if (pc.gflag) { const int b = pc.bc[j]; sgh[lane] = b == 1 || b == 2; }
      P::stage(a, j, tile, lane);
    }
    // streamed once per pass: keep it out of the caches that hold the tiles
    uint32_t mask = __builtin_nontemporal_load(vl.mask + (size_t)(v0 + v) * kWave + lane);
    __syncthreads();
    if (active) {
      while (mask) {
        const uint32_t b = (uint32_t)__builtin_ctz(mask);
        mask &= mask - 1;
        const uint32_t q = sub + b * P::kSub;
        double dx[3];
        dx[0] = xi[0] - sx[q];
        double amax = fabs(dx[0]);
        if (D == 3) {
          dx[1] = xi[1] - sy[q]; dx[2] = xi[2] - sz[q];
          amax = fmax(amax, fmax(fabs(dx[1]), fabs(dx[2])));
        }
        if (pc.cut.outside(amax, P::hker(a, t, tile, q)) || !P::extra(a, t, tile, q, base + q, i)) continue;
        if constexpr (FROZEN) {
          dx[0] = xa[0] - sa.x[q];
          if (D == 3) { dx[1] = xa[1] - sa.y[q]; dx[2] = xa[2] - sa.z[q]; }
        }
        P::pair(a, t, tile, q, base + q, i, dx, amax);
        if (pc.gflag && sgh[q]) pc.gflag[base + q] = 1u;
      }
    }
    __syncthreads();
  }
  if constexpr (P::kSub > 1) P::combine(t);
  P::end_pass(a, t, 0, i, active);
  if (active && sub == 0) P::finish(a, t, i, pc.err);
}

// sum of v over the kSub lanes of a target (lanes slot*kSub .. +kSub-1)
template <int S>
__device__ inline double sub_sum(double v) {
  for (int o = 1; o < S; o <<= 1) v += __shfl_xor(v, o);
  return v;
}
template <int S>
__device__ inline uint32_t sub_sum_u(uint32_t v) {
  for (int o = 1; o < S; o <<= 1) v += (uint32_t)__shfl_xor((int)v, o);
  return v;
}

__device__ inline bool in_support(const double *dx, double h, int D) {
  for (int k = 0; k < D; k++)
    if (fabs(dx[k]) > h) return false;
  return true;
}

constexpr double kInvSqrtPi = 0.564189583547756;   // sph_lib.cc:207

// |dx_k| > 3 h (1 + 1e-12) on any axis implies fl(dx_k / h) > 3: the pair is
// outside the kernel box and contributes exactly zero.
template <int D>
__device__ inline bool beyond(const double *dx, double bound) {
  bool out = fabs(dx[0]) > bound;
  if (D == 3) out = out || fabs(dx[1]) > bound || fabs(dx[2]) > bound;
  return out;
}
constexpr double kCutMargin = 1.0 + 1e-12;

// Each pass supplies, besides load_target/stage/pair/finish:
//   hker(q):  smoothing length of the kernel-box test, extra(q): masks.
// The pair arithmetic type R is double (FP64 form) or float (FP32-pair form):
// the pair predicate (phase 1) is FP64 in both, so both forms visit exactly
// the same pairs; accumulators, integration and the EOS stay FP64.
// ---------------------------------------------------------------------------
template <class R> __device__ inline R rexp(R x);
template <> __device__ inline double rexp<double>(double x) { return exp(x); }
template <> __device__ inline float rexp<float>(float x) { return __expf(x); }
template <class R> __device__ inline R rrcp(R x);
template <> __device__ inline double rrcp<double>(double x) { return 1.0 / x; }
template <> __device__ inline float rrcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// |x| as the source modifier of the consuming instruction (x < 0 ? -x : x
// would cost a compare and a select)
template <class R> __device__ inline R rabs(R x);
template <> __device__ inline double rabs<double>(double x) { return __builtin_fabs(x); }
template <> __device__ inline float rabs<float>(float x) { return __builtin_fabsf(x); }

// exp(-e) in FP64 for 0 <= e < kETab / 8 from a table of exp(-k/8) and the
// Taylor series of exp(-r), 0 <= r < 1/8, to r^10 (truncation < 6e-18).  The
// table lives in each workgroup's LDS; 224 entries (1.75 KB) keep the density
// pass at 6 waves per SIMD.
constexpr int kETab = 224;            // covers the kernel box: e = |s|^2 <= 27
__device__ inline void etab_fill(double *tab, int lane) {
  for (int k = lane; k < kETab; k += kWave) tab[k] = exp(-0.125 * k);
}
__device__ inline double exp_neg_tab(double e, const double *tab) {
  const double k = fmin(floor(e * 8.0), (double)(kETab - 1));
  const double r = e - 0.125 * k;
  double p = 1.0 / 3628800;
  p = fma(p, -r, 1.0 / 362880); p = fma(p, -r, 1.0 / 40320); p = fma(p, -r, 1.0 / 5040);
  p = fma(p, -r, 1.0 / 720); p = fma(p, -r, 1.0 / 120); p = fma(p, -r, 1.0 / 24);
  p = fma(p, -r, 1.0 / 6); p = fma(p, -r, 0.5); p = fma(p, -r, 1.0); p = fma(p, -r, 1.0);
  return tab[(int)k] * p;
}

// The exponential of the FP32-pair forms' FP64 weights.  FX false
// (PSPH_FP32_PAIR): exp_neg_tab, FP64 throughout.  FX true (PSPH_FP32_EXP):
// the FP32 hardware exponential 2^-tf of the argument t = e log2(e), formed
// in FP64 and rounded once to tf: a relative error of the instruction's own
// (about one FP32 ulp) plus t 2^-24 ln 2 from the rounding of t (largest at
// the kernel box's corner, where the weight is e^-27 of the centre's), while
// h, 1/h and the norm stay FP64.  Taking the rounding of t back out (times
// 1 - (t - tf) ln 2) halves the form's difference to FP64 over the bench's
// steps and costs 4 % of the step: not kept.
template <bool FX> constexpr int wtab_size() { return FX ? 1 : kETab; }
template <bool FX> __device__ inline void wtab_fill(double *tab, int lane) {
  if constexpr (!FX) etab_fill(tab, lane);
}
constexpr double kLog2e = 1.4426950408889634;
// 2^-t of the FX form (t >= 0 in FP64)
__device__ inline double exp2_neg_fx(double t) { return (double)__builtin_amdgcn_exp2f(-(float)t); }
template <bool FX> __device__ inline double exp_neg_w(double e, const double *tab) {
  if constexpr (!FX) return exp_neg_tab(e, tab);
  else return exp2_neg_fx(e * kLog2e);
}

template <class R> __device__ inline R pow16(R x) {
  x *= x; x *= x; x *= x;
  return x * x;
}

// ---------------------------------------------------------------------------
// RHS pass: mom_engr_update, USE_GSPH==0 (mom_engr_update.cc:30-439)
// ---------------------------------------------------------------------------
struct RhsArgs {
  const double *h, *m, *rho, *p, *c, *T, *cp, *vx, *vy, *vz, *svx, *svy, *svz, *e, *xi;
  const int32_t *bc;
  double *o_vx, *o_vy, *o_vz, *o_e, *o_p, *o_c, *o_T, *o_gam, *o_cp;
  double dt, g, cutoff, cutoff2, alpha, beta, ata, eps_half, prandtl, heat_scale;
  // per-pair constants, folded on the host: 2 beta; 20 eps / Pr (the heat
  // term's 10 eps / Pr, doubled because the energy sum holds 2 de/dt)
  double beta2, kheat;
  int lans, dim;
  EosConst eos;
  unsigned long long *pairs;
};

// Pair terms, written with the halves folded: c2 = c_i + c_j = 2 c_ab,
// rs = rho_i + rho_j = 2 rho_ab, cp2 = 2 cp_ab.  Pi = (beta mu - alpha c_ab)
// mu / rho_ab = (2 beta mu - alpha c2) mu / rs with mu = h v.r / den (one
// reciprocal of den rs gives both 1/den and 1/rs); the heat term's cp_ab /
// c_ab = cp2 / c2.  The energy sum holds 2 de/dt (finish takes half).
//
// The FP32-pair form keeps the pressure term m_j (p_j/rho_j^2 + p_i/rho_i^2)
// grad W in FP64: in air near hydrostatic balance its pair terms cancel to
// about 1 % of their size, and FP32 rounding there is what the velocities
// feel.  The kernel value comes from the exp table (FP64 at FP32-like cost);
// viscosity, heat conduction and the energy work terms stay FP32.
template <bool FX> struct RhsMixTile { double m64[kWave], mpv64[kWave], etab[wtab_size<FX>()]; float mV[kWave]; };
template <class R> struct RhsTileR { R PV2[kWave]; };

// -2 pi^(-D/2): grad W = G dx with G = -2 W / h^2 = kGradNorm h^-(D+2) exp(-r^2/h^2)
template <int D> constexpr double kGradNorm = -2.0 * kInvSqrtPi * kInvSqrtPi * kInvSqrtPi;
template <> constexpr double kGradNorm<1> = -2.0 * kInvSqrtPi;

template <class R, bool MIXED = std::is_same<R, float>::value, bool FX = false>
struct RhsPass {
  using Args = RhsArgs;
  static constexpr bool kMixed = MIXED;
  struct Tile : std::conditional<kMixed, RhsMixTile<FX>, RhsTileR<R>>::type {
    double hh[kWave];       // h / 2
    R m[kWave], V[kWave], rho[kWave], c[kWave], T[kWave], cp[kWave];
    R vx[kWave], vy[kWave], vz[kWave];
  };
  struct Tgt {
    double hi, hh, supp, rv[3], re, pv64;
    R rhoi, Vi, pvsqi, ci, Ti, cpi, vel[3];
    unsigned npairs;
    int bad;
  };
  __device__ static void init(Tile &s, int lane) {
    if constexpr (kMixed) wtab_fill<FX>(s.etab, lane);
  }
  static constexpr int kSub = ::kSub;
  __device__ static void combine(Tgt &t) {
    for (int k = 0; k < 3; k++) t.rv[k] = sub_sum<kSub>(t.rv[k]);
    t.re = sub_sum<kSub>(t.re);
    t.npairs = sub_sum_u<kSub>(t.npairs);
    t.bad = sub_sum_u<kSub>((uint32_t)t.bad) ? 1 : 0;
  }
  __device__ static int npasses(const Args &) { return 1; }
  __device__ static void begin_pass(const Args &, Tgt &, int) {}
  __device__ static void end_pass(const Args &, Tgt &, int, uint32_t, bool) {}
  __device__ static bool want(const Tgt &) { return true; }
  __device__ static void load_target(const Args &a, uint32_t i, Tgt &t) {
    t.hi = a.h[i];
    t.hh = 0.5 * t.hi;
    t.supp = a.cutoff2 * t.hi;
    const double rho = a.rho[i];
    t.bad = !(rho > 0) ? 1 : 0;
    const double Vi = 1.0 / rho;
    t.pv64 = a.p[i] * Vi * Vi;
    t.rhoi = (R)rho; t.Vi = (R)Vi; t.pvsqi = (R)t.pv64;
    t.ci = (R)a.c[i]; t.Ti = (R)a.T[i]; t.cpi = (R)a.cp[i];
    if (a.lans) { t.vel[0] = (R)a.svx[i]; t.vel[1] = (R)a.svy[i]; t.vel[2] = (R)a.svz[i]; }
    else { t.vel[0] = (R)a.vx[i]; t.vel[1] = (R)a.vy[i]; t.vel[2] = (R)a.vz[i]; }
    t.rv[0] = t.rv[1] = t.rv[2] = 0; t.re = 0; t.npairs = 0;
  }
  __device__ static void stage(const Args &a, uint32_t j, Tile &s, int l) {
    const double rj = a.rho[j];
    const double V = 1.0 / rj;
    const double pv = a.p[j] * V * V;
    const double mj = a.m[j];
    s.hh[l] = 0.5 * a.h[j]; s.m[l] = (R)mj; s.V[l] = (R)V; s.rho[l] = (R)rj;
    if constexpr (kMixed) { s.m64[l] = mj; s.mpv64[l] = mj * pv; s.mV[l] = (float)(mj * V); }
    else s.PV2[l] = (R)pv;
    s.c[l] = (R)a.c[j]; s.T[l] = (R)a.T[j]; s.cp[l] = (R)a.cp[j];
    if (a.lans) { s.vx[l] = (R)a.svx[j]; s.vy[l] = (R)a.svy[j]; s.vz[l] = (R)a.svz[j]; }
    else { s.vx[l] = (R)a.vx[j]; s.vy[l] = (R)a.vy[j]; s.vz[l] = (R)a.vz[j]; }
  }
  // h_ij = (h_i + h_j)/2 (ME_UPDATE_SML 2).  The reference skips j == i; here
  // the target meets itself once, with dx = 0 and v_ij = 0, and every term it
  // adds is an exact zero (f dx, f v.r, no viscosity for v.r = 0, no heat for
  // |v_ij . dx| = 0), so no test is spent on it; finish leaves it out of the
  // pair count.
  __device__ static double hker(const Args &, const Tgt &t, const Tile &s, uint32_t q) { return t.hh + s.hh[q]; }
  __device__ static bool extra(const Args &, const Tgt &, const Tile &, uint32_t, uint32_t, uint32_t) { return true; }
  template <int D>
  __device__ static void pair_d(const Args &a, Tgt &t, const Tile &s, uint32_t q, uint32_t, uint32_t,
                                const double *dxd, double) {
    t.npairs++;
    if (!(s.rho[q] > R(0))) { t.bad = 1; return; }      // assert(uvecj[0] > 0)
    if constexpr (kMixed) pair_mixed<D>(a, t, s, q, dxd);
    else pair_plain<D>(a, t, s, q, dxd);
  }
  template <int D>
  __device__ static void pair_plain(const Args &a, Tgt &t, const Tile &s, uint32_t q, const double *dxd) {
    const R rhoj = s.rho[q];
    const R h = (R)(t.hh + s.hh[q]);
    const R ih = rrcp<R>(h);
    R dx[3], dist_sq = 0, expt = 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      dx[k] = (R)dxd[k];
      const R sk = dx[k] * ih;
      dist_sq += dx[k] * dx[k];
      expt += sk * sk;
    }
    const R ih2 = ih * ih;
    const R ihD = (D == 3) ? ih2 * ih : ih;
    // G = -2 W / h^2: grad W = G dx
    const R G = (R(kGradNorm<D>) * ih2) * (ihD * rexp<R>(-expt));
    const R Vj = s.V[q];
    R vij[3];
    vij[0] = t.vel[0] - s.vx[q];
    if (D == 3) { vij[1] = t.vel[1] - s.vy[q]; vij[2] = t.vel[2] - s.vz[q]; }
    const R c2 = t.ci + s.c[q];
    R vrab = 0, dotvv = 0;
#pragma unroll
    for (int k = 0; k < D; k++) { vrab += dx[k] * vij[k]; dotvv += vij[k] * vij[k]; }
    // art_vis (sph_lib.cc:402-424): Pi = (beta mu - alpha c) mu / rho_ab for
    // approaching pairs, mu = h v.r / (r^2 + ata h^2)
    R vis = 0;
    if (vrab < R(0)) {
      const R rs = t.rhoi + rhoj;
      const R den = dist_sq + R(a.ata) * h * h;
      const R r = rrcp<R>(den * rs);
      const R miu = h * vrab * (rs * r);
      vis = (R(a.beta2) * miu - R(a.alpha) * c2) * miu * (den * r);
    }
    if (a.lans) vis -= R(a.eps_half) * dotvv * Vj;        // SPH_epsilon_mom
    const R mpv = s.m[q] * (s.PV2[q] + t.pvsqi + vis);
    R heat = 0;
    if (a.lans && D == 3) {
      // k_ij F_ij of SPH_epsilon_heat_conductivity (6-arg, sph_lib.cc:1399-1439)
      // and compute_F (sph_lib.cc:257-281): 10 eps cp r^2 |v|^2 / (Pr sum|v_k dx_k|)
      // / (h c) * (dW.dx / r^2); the r^2 factors cancel: dW.dx = G r^2.
      R dotvr = 0;
#pragma unroll
      for (int k = 0; k < D; k++) dotvr += rabs<R>(vij[k] * dx[k]);
      const R wr = G * dist_sq;
      if (dotvr != R(0) && wr != R(0)) {
        const R cp2 = s.cp[q] + t.cpi;
        const R kF = R(a.kheat) * cp2 * dotvv * wr * rrcp<R>(dotvr * h * c2);
        heat = s.m[q] * Vj * t.Vi * (t.Ti - s.T[q]) * kF;
      }
    }
    R de = 0;
    const R Gm = mpv * G;
#pragma unroll
    for (int k = 0; k < D; k++) {
      const R f = Gm * dx[k];
      t.rv[k] -= (double)f;
      de += f * vij[k];
    }
    t.re += (double)(de + heat);
  }
  template <int D>
  __device__ static void pair_mixed(const Args &a, Tgt &t, const Tile &s, uint32_t q, const double *dxd) {
    const double hd = t.hh + s.hh[q];
    const float h = (float)hd;
    const float ih = rrcp<float>(h);
    const double ih64 = (double)ih * fma(-hd, (double)ih, 2.0);     // one Newton step: ~1e-14
    const double ih2 = ih64 * ih64;
    const double ihD = (D == 3) ? ih2 * ih64 : ih64;
    double r2 = 0;
    float dx[3];
#pragma unroll
    for (int k = 0; k < D; k++) { r2 = fma(dxd[k], dxd[k], r2); dx[k] = (float)dxd[k]; }
    // G = -2 W / h^2: grad W = G dx
    const double G64 = (kGradNorm<D> * ih2) * (ihD * exp_neg_w<FX>(r2 * ih2, s.etab));
    const float G = (float)G64, dist_sq = (float)r2;
    const float Vj = s.V[q], rhoj = s.rho[q];
    float vij[3];
    vij[0] = t.vel[0] - s.vx[q];
    if (D == 3) { vij[1] = t.vel[1] - s.vy[q]; vij[2] = t.vel[2] - s.vz[q]; }
    const float c2 = t.ci + s.c[q];
    float vrab = 0, dotvv = 0;
#pragma unroll
    for (int k = 0; k < D; k++) { vrab += dx[k] * vij[k]; dotvv += vij[k] * vij[k]; }
    float vis = 0;
    if (vrab < 0.f) {
      const float rs = t.rhoi + rhoj;
      const float den = dist_sq + (float)a.ata * h * h;
      const float r = rrcp<float>(den * rs);
      const float miu = h * vrab * (rs * r);
      vis = ((float)a.beta2 * miu - (float)a.alpha * c2) * miu * (den * r);
    }
    if (a.lans) vis -= (float)a.eps_half * dotvv * Vj;
    // pressure part in FP64, viscous part in FP32; f = (M_p + M_v) G dx
    const double f = fma(s.m64[q], t.pv64, s.mpv64[q]) * G64 + (double)(s.m[q] * vis * G);
#pragma unroll
    for (int k = 0; k < D; k++) t.rv[k] = fma(-f, dxd[k], t.rv[k]);
    float heat = 0;
    if (a.lans && D == 3) {
      float dotvr = 0;
#pragma unroll
      for (int k = 0; k < D; k++) dotvr += rabs<float>(vij[k] * dx[k]);
      const float wr = G * dist_sq;
      if (dotvr != 0.f && wr != 0.f) {
        const float cp2 = s.cp[q] + t.cpi;
        const float kF = (float)a.kheat * cp2 * dotvv * wr * rrcp<float>(dotvr * h * c2);
        heat = s.mV[q] * t.Vi * (t.Ti - s.T[q]) * kF;
      }
    }
    t.re += fma(f, (double)vrab, (double)heat);
  }
  __device__ static void finish(const Args &a, Tgt &t, uint32_t i, int *err) {
    if (t.bad) { raise_err(err, PSPH_EINVARIANT, i); return; }
    const double gz = -a.g, dt = a.dt;
    // v += dt rhs_v; gravity on the last axis; e += dt (rhs_e + v_last g)
    // (mom_engr_update.cc:372-388); the velocity in the energy term is the
    // FP64 one of the target; t.re holds 2 rhs_e
    const double vl = a.dim == 3 ? (a.lans ? a.svz[i] : a.vz[i]) : (a.lans ? a.svx[i] : a.vx[i]);
    double nv0 = a.vx[i] + dt * (t.rv[0]);
    double nv1 = a.vy[i] + dt * (t.rv[1]);
    double nv2 = a.vz[i] + dt * (t.rv[2]);
    if (a.dim == 3) nv2 += dt * gz; else nv0 += dt * gz;
    const double ne = a.e[i] + dt * (0.5 * t.re + vl * gz);
    a.o_vx[i] = nv0; a.o_vy[i] = nv1; a.o_vz[i] = nv2;
    a.o_e[i] = ne;
    double p, c, T, gam, cp = a.cp[i];
    eos_dev(a.eos, a.rho[i], ne, a.xi[i], p, c, T, gam, &cp);
    a.o_p[i] = p; a.o_c[i] = c; a.o_T[i] = T; a.o_gam[i] = gam; a.o_cp[i] = cp;
    if (t.npairs > 1) atomicAdd(a.pairs, (unsigned long long)(t.npairs - 1));   // without the target itself
  }
};

// ---------------------------------------------------------------------------
// Density pass: smooth_density (density_massfrac.cc:25-359)
// ---------------------------------------------------------------------------
struct DensArgs {
  const double *h, *m, *rho, *e;
  const int32_t *bc, *phase;
  double *o_rho, *o_e, *xi, *p, *c, *T, *gam, *cp;
  int32_t *inv;
  double cutoff, cutoff2, e_ratio, msfrc_thresh;
  int pghost, energy_smooth;
  EosConst eos;
  CutTest cut;
  double *o_wn;            // optional: energy-smoothing normalisation (adjoint recompute)
  int e_sq16;              // e_ratio^2 == 16: exp(-16 x) = exp(-x)^16 by four squarings
  double e_norm;           // the energy-smoothing kernel's norm times h^D: (e_ratio / sqrt(pi))^D
};

// pi^(-D/2): W = kNorm h^-D exp(-s^2)
template <int D> constexpr double kNorm = kInvSqrtPi * kInvSqrtPi * kInvSqrtPi;
template <> constexpr double kNorm<1> = kInvSqrtPi;

// mr = m / rho of the candidate (the energy-smoothing weight's factor).  m is
// staged twice, as m1 (phase 1, else 0) and m2 (phase 2, else 0), so that a
// pair adds to both sums without selecting between them; a candidate that
// does not contribute (contr_dens false or another phase) stages zero masses
// and mr = 0, so its terms are exact zeros and no pair tests for it.
template <bool FX> struct DensMixTile { double m1[kWave], m2[kWave], mr64[kWave], e64[kWave], etab[wtab_size<FX>()]; };
template <class R> struct DensTileR { R m1[kWave], m2[kWave], mr[kWave], e[kWave]; };

// The FP32-pair form evaluates the density and energy-smoothing weights in
// FP64 with the exp table: the pressure gradient of the next step amplifies
// relative errors of rho and e by its cancellation.
template <class R, bool MIXED = std::is_same<R, float>::value, bool FX = false>
struct DensPass {
  using Args = DensArgs;
  static constexpr bool kMixed = MIXED;
  // the mixed form stages only the FP64 copies it reads
  struct Tile : std::conditional<kMixed, DensMixTile<FX>, DensTileR<R>>::type {
    double hh[kWave];       // h / 2
  };
  __device__ static void init(Tile &s, int lane) {
    if constexpr (kMixed) wtab_fill<FX>(s.etab, lane);
  }
  struct Tgt { double hi, hh, supp, r1, r2, wn, en; };
  static constexpr int kSub = ::kSub;
  __device__ static void combine(Tgt &t) {
    t.r1 = sub_sum<kSub>(t.r1); t.r2 = sub_sum<kSub>(t.r2);
    t.wn = sub_sum<kSub>(t.wn); t.en = sub_sum<kSub>(t.en);
  }
  __device__ static int npasses(const Args &) { return 1; }
  __device__ static void begin_pass(const Args &, Tgt &, int) {}
  __device__ static void end_pass(const Args &, Tgt &, int, uint32_t, bool) {}
  __device__ static bool want(const Tgt &) { return true; }
  __device__ static void load_target(const Args &a, uint32_t i, Tgt &t) {
    t.hi = a.h[i]; t.hh = 0.5 * t.hi; t.supp = a.cutoff2 * t.hi;
    t.r1 = t.r2 = t.wn = t.en = 0;
  }
  __device__ static void stage(const Args &a, uint32_t j, Tile &s, int l) {
    s.hh[l] = 0.5 * a.h[j];
    const int ph = a.phase[j];
    const int ok = (contr_dens(a.pghost, a.bc[j]) && (ph == 1 || ph == 2)) ? ph : 0;
    const double mj = a.m[j];
    const double m1 = ok == 1 ? mj : 0.0, m2 = ok == 2 ? mj : 0.0;
    const double mr = ok ? mj / a.rho[j] : 0.0, e = ok ? a.e[j] : 0.0;
    if constexpr (kMixed) {
      s.m1[l] = m1; s.m2[l] = m2; s.mr64[l] = mr; s.e64[l] = e;
    } else {
      s.m1[l] = (R)m1; s.m2[l] = (R)m2; s.mr[l] = (R)mr; s.e[l] = (R)e;
    }
  }
  // kernel at h_ij (DENSITY_UPDATE_SML 11); contr_dens and the phase are in
  // the staged masses.  The energy-smoothing kernel (support 0.75 h_ij) lies
  // inside that box.
  __device__ static double hker(const Args &, const Tgt &t, const Tile &s, uint32_t q) { return t.hh + s.hh[q]; }
  __device__ static bool extra(const Args &, const Tgt &, const Tile &, uint32_t, uint32_t, uint32_t) { return true; }
  template <int D>
  __device__ static void pair_d(const Args &a, Tgt &t, const Tile &s, uint32_t q, uint32_t, uint32_t,
                                const double *ds, double amax) {
    const double hijd = t.hh + s.hh[q];
    if constexpr (kMixed && FX) {
      // The FP32-exp form: the weights in FP32 (reciprocal of h_ij, s^2, the
      // exponential, the norm), the sums in FP64.  Over the 4M bench window
      // this moves rho by 2e-8 rel L1 against FP64 (3e-9 with FP64 weights
      // and the FP32 exponential) and the velocities not measurably (3.9e-7
      // against 3.6e-7), for 10 % of the pass.
      const float ih = rrcp<float>((float)hijd);
      float r2 = 0;
#pragma unroll
      for (int k = 0; k < D; k++) { const float d = (float)ds[k]; r2 = fmaf(d, d, r2); }
      const float ih2 = ih * ih;
      const float ex = __builtin_amdgcn_exp2f(-(r2 * (ih2 * (float)kLog2e)));
      const float ihD = (D == 3) ? ih2 * ih : ih;
      const double wk = (double)(((float)kNorm<D> * ihD) * ex);
      t.r1 = fma(wk, s.m1[q], t.r1);
      t.r2 = fma(wk, s.m2[q], t.r2);
      if (a.energy_smooth && !a.cut.outside(a.e_ratio * amax, hijd)) {
        const float ee = a.e_sq16 ? pow16(ex) : __builtin_amdgcn_exp2f(-(float)(a.e_ratio * a.e_ratio) * (r2 * (ih2 * (float)kLog2e)));
        const double vr = (double)(((float)a.e_norm * ihD) * ee) * s.mr64[q];
        t.wn += vr;
        t.en += vr * s.e64[q];
      }
    } else if constexpr (kMixed) {
      const float ih32 = rrcp<float>((float)hijd);
      const double ih = (double)ih32 * fma(-hijd, (double)ih32, 2.0);
      double r2 = 0;
#pragma unroll
      for (int k = 0; k < D; k++) r2 = fma(ds[k], ds[k], r2);
      const double ih2 = ih * ih;
      const double expt = r2 * ih2;
      const double ihD = (D == 3) ? ih2 * ih : ih;
      const double ex = exp_neg_w<FX>(expt, s.etab);
      // one of m1, m2 is 0: that sum gains an exact zero
      const double wk = (kNorm<D> * ihD) * ex;
      t.r1 = fma(wk, s.m1[q], t.r1);
      t.r2 = fma(wk, s.m2[q], t.r2);
      if (a.energy_smooth && !a.cut.outside(a.e_ratio * amax, hijd)) {
        const double ee = a.e_sq16 ? pow16(ex) : exp_neg_w<FX>(a.e_ratio * a.e_ratio * expt, s.etab);
        const double vr = (a.e_norm * ihD) * (ee * s.mr64[q]);
        t.wn += vr;
        t.en += vr * s.e64[q];
      }
    } else {
    const R hij = (R)hijd;
    const R ih = rrcp<R>(hij);
    R expt = 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      R sk = (R)ds[k] * ih;
      expt += sk * sk;
    }
    const R ihD = (D == 3) ? ih * ih * ih : ih;
    const R ex = rexp<R>(-expt);
    const R wk = (R(kNorm<D>) * ihD) * ex;
    t.r1 += (double)(wk * s.m1[q]);
    t.r2 += (double)(wk * s.m2[q]);
    if (a.energy_smooth) {
      // weight(4 ds / h_ij, h_ij / 4) m_j / rho_j (HAVE_ENERGY_SMOOTH 1); exact FP64 box test
      if (!a.cut.outside(a.e_ratio * amax, hijd)) {
        const R e2 = R(a.e_ratio * a.e_ratio);
        const R vr = (R(a.e_norm) * ihD) * ((a.e_sq16 ? pow16(ex) : rexp<R>(-(e2 * expt))) * s.mr[q]);
        t.wn += (double)vr;
        t.en += (double)(vr * s.e[q]);
      }
    }
    }
  }
  __device__ static void finish(const Args &a, Tgt &t, uint32_t i, int *err) {
    double density = 0.0;
    density += t.r1;
    density += t.r2;
    const double xi = t.r2 / density;
    double e = a.e[i];
    if (a.energy_smooth) {
      if (!(t.wn > 0)) { raise_err(err, PSPH_EINVARIANT, i); return; }
      e = t.en / t.wn;
      if (a.o_wn) a.o_wn[i] = t.wn;
    }
    if (!(density > 0) || !(xi <= 1)) { raise_err(err, PSPH_EINVARIANT, i); return; }
    a.o_rho[i] = density;
    a.o_e[i] = e;
    a.xi[i] = xi;
    if (xi >= a.msfrc_thresh && a.inv[i] != 2) a.inv[i] = 2;
    double p, c, T, gam, cp = a.cp[i];
    eos_dev(a.eos, density, e, xi, p, c, T, gam, &cp);
    a.p[i] = p; a.c[i] = c; a.T[i] = T; a.gam[i] = gam; a.cp[i] = cp;
  }
};

// ---------------------------------------------------------------------------
// LANS velocity filter: smooth_velocity (smooth_velocity.cc:25-165)
// ---------------------------------------------------------------------------
struct VsmArgs {
  const double *h, *m, *rho, *vx, *vy, *vz;
  const int32_t *bc;
  double *svx, *svy, *svz;
  double cutoff, cutoff2;
  double *o_wn;            // optional: the filter's normalisation sum (adjoint recompute)
  // the per-pair contr_vel test is skipped (a non-contributing candidate
  // stages m / rho = 0, so its terms are exact zeros) unless the step is
  // recorded, where the boundary-ghost marks follow the tested pairs
  int ok_test;
};

// The FP32-pair form evaluates the filter weights in FP64 from the exp table
// (kernel argument, exponential, m_j / rho_j): sv moves the particles, and
// FP32 weights (relative error ~3e-7, dominated by the rounding of s^2 and of
// h) made the velocity filter the largest source of FP32-vs-FP64 difference,
// ~15x the RHS's FP32 terms.
template <bool FX> struct VsmMixTile { double mr64[kWave], v64[3][kWave], etab[wtab_size<FX>()]; };
template <class R> struct VsmTileR { R mr[kWave], vx[kWave], vy[kWave], vz[kWave]; };

template <class R, bool FX = false>
struct VsmPass {
  using Args = VsmArgs;
  static constexpr bool kMixed = std::is_same<R, float>::value;
  struct Tile : std::conditional<kMixed, VsmMixTile<FX>, VsmTileR<R>>::type {
    int ok[kWave];
  };
  __device__ static void init(Tile &s, int lane) {
    if constexpr (kMixed) wtab_fill<FX>(s.etab, lane);
  }
  struct Tgt { double hi, supp, ih2, nrm, v64[3], dv[3], wn; R ih, v[3]; };
  static constexpr int kSub = ::kSub;
  __device__ static void combine(Tgt &t) {
    for (int k = 0; k < 3; k++) t.dv[k] = sub_sum<kSub>(t.dv[k]);
    t.wn = sub_sum<kSub>(t.wn);
  }
  __device__ static int npasses(const Args &) { return 1; }
  __device__ static void begin_pass(const Args &, Tgt &, int) {}
  __device__ static void end_pass(const Args &, Tgt &, int, uint32_t, bool) {}
  __device__ static bool want(const Tgt &) { return true; }
  __device__ static void load_target(const Args &a, uint32_t i, Tgt &t) {
    t.hi = a.h[i]; t.supp = a.cutoff2 * t.hi;
    const double ih = 1.0 / t.hi;
    const double nk = kInvSqrtPi * ih;
    t.ih = (R)ih;
    t.ih2 = FX ? ih * ih * kLog2e : ih * ih;    // the FX form's exponential takes s^2 log2(e)
    t.nrm = a.vy ? nk * nk * nk : nk;
    t.v64[0] = a.vx[i]; t.v64[1] = a.vy ? a.vy[i] : 0; t.v64[2] = a.vz ? a.vz[i] : 0;
    for (int k = 0; k < 3; k++) t.v[k] = (R)t.v64[k];
    t.dv[0] = t.dv[1] = t.dv[2] = 0; t.wn = 0;
  }
  __device__ static void stage(const Args &a, uint32_t j, Tile &s, int l) {
    const double v[3] = {a.vx[j], a.vy ? a.vy[j] : 0.0, a.vz ? a.vz[j] : 0.0};
    const int ok = contr_vel(a.bc[j]);
    if constexpr (kMixed) {
      s.mr64[l] = ok ? a.m[j] / a.rho[j] : 0.0;
      for (int k = 0; k < 3; k++) s.v64[k][l] = v[k];
    } else {
      s.mr[l] = ok ? (R)(a.m[j] / a.rho[j]) : R(0);
      s.vx[l] = (R)v[0]; s.vy[l] = (R)v[1]; s.vz[l] = (R)v[2];
    }
    s.ok[l] = ok;
  }
  __device__ static double hker(const Args &, const Tgt &t, const Tile &, uint32_t) { return t.hi; }
  __device__ static bool extra(const Args &a, const Tgt &, const Tile &s, uint32_t q, uint32_t, uint32_t) {
    return !a.ok_test || s.ok[q] != 0;
  }
  template <int D>
  __device__ static void pair_d(const Args &a, Tgt &t, const Tile &s, uint32_t q, uint32_t, uint32_t,
                                const double *ds, double) {
    if constexpr (kMixed) {
      double r2 = 0;
#pragma unroll
      for (int k = 0; k < D; k++) r2 = fma(ds[k], ds[k], r2);
      // the weight without the kernel's norm, which cancels in dv / wn
      double ew;
      if constexpr (FX) ew = exp2_neg_fx(r2 * t.ih2);
      else ew = exp_neg_tab(r2 * t.ih2, s.etab);
      const double w = ew * s.mr64[q];
#pragma unroll
      for (int k = 0; k < D; k++) t.dv[k] = fma(w, s.v64[k][q] - t.v64[k], t.dv[k]);
      t.wn += w;
    }
    else {
      R expt = 0;
#pragma unroll
      for (int k = 0; k < D; k++) {
        R sk = (R)ds[k] * t.ih;
        expt += sk * sk;
      }
      const R mvw = rexp<R>(-expt) * s.mr[q];        // the norm cancels in dv / wn
      t.dv[0] += (double)(mvw * (s.vx[q] - t.v[0]));
      if (D == 3) {
        t.dv[1] += (double)(mvw * (s.vy[q] - t.v[1]));
        t.dv[2] += (double)(mvw * (s.vz[q] - t.v[2]));
      }
      t.wn += (double)mvw;
    }
  }
  __device__ static void finish(const Args &a, Tgt &t, uint32_t i, int *err) {
    if (!(t.wn > 0)) { raise_err(err, PSPH_EINVARIANT, i); return; }
    const double iw = 1.0 / t.wn;
    if (a.o_wn) a.o_wn[i] = t.wn * t.nrm;          // the normalisation sum with the kernel's norm
    a.svx[i] = t.v64[0] + t.dv[0] * iw;
    if (a.svy) { a.svy[i] = t.v64[1] + t.dv[1] * iw; a.svz[i] = t.v64[2] + t.dv[2] * iw; }
  }
};

// ---------------------------------------------------------------------------
// Adaptive smoothing length: adaptive_sml (neighs.cc:479-693).  One launch
// per iteration of the reference's loop; h is updated in place (no pass reads
// another particle's h) and done[] marks particles whose |dh/h| < thresh.
// ---------------------------------------------------------------------------
struct AdaptArgs {
  const double *m;
  const int32_t *bc;
  double *h;
  uint32_t *done;
  double cutoff, cutoff2, c_smooth, eta, ratio, thresh;
  int pghost, dim;
};

// The kernel value comes from the FP64 exp table (exp_neg_tab, about an ulp),
// or in the FX form (PSPH_FP32_EXP) from the FP32 exponential of the FP64
// argument: its rounding is independent from pair to pair, so rho* and h
// move by ~1e-9 relative over a particle's ~2000 list pairs.
template <class R, bool FX = false>
struct AdaptPass {
  using Args = AdaptArgs;
  struct Tile {
    R m[kWave];
    int ok[kWave];
    double etab[wtab_size<FX>()];
  };
  __device__ static void init(Tile &s, int lane) { wtab_fill<FX>(s.etab, lane); }
  struct Tgt { double mi, hcur, Hs, supp, rho_star; R iHs, norm; bool done; };
  static constexpr int kSub = ::kSub;
  __device__ static void combine(Tgt &t) { t.rho_star = sub_sum<kSub>(t.rho_star); }
  __device__ static int npasses(const Args &) { return 1; }
  __device__ static void begin_pass(const Args &a, Tgt &t, int) {
    t.Hs = t.hcur * a.c_smooth;
    const double iHs = 1.0 / t.Hs;
    t.iHs = (R)iHs;
    t.supp = a.cutoff2 * t.Hs;
    const double nk = kInvSqrtPi * iHs;
    t.norm = (R)(a.dim == 3 ? nk * nk * nk : nk);
    t.rho_star = 0;
  }
  __device__ static bool want(const Tgt &t) { return !t.done; }
  __device__ static void load_target(const Args &a, uint32_t i, Tgt &t) {
    t.mi = a.m[i]; t.hcur = a.h[i]; t.done = a.done[i] != 0;
  }
  __device__ static void stage(const Args &a, uint32_t j, Tile &s, int l) {
    s.m[l] = (R)a.m[j];
    s.ok[l] = contr_dens(a.pghost, a.bc[j]);
  }
  __device__ static double hker(const Args &, const Tgt &t, const Tile &, uint32_t) { return t.Hs; }
  __device__ static bool extra(const Args &, const Tgt &, const Tile &s, uint32_t q, uint32_t, uint32_t) { return s.ok[q] != 0; }
  template <int D>
  __device__ static void pair_d(const Args &a, Tgt &t, const Tile &s, uint32_t q, uint32_t, uint32_t,
                                const double *dx, double) {
    R expt = 0;
#pragma unroll
    for (int k = 0; k < D; k++) {
      R sk = (R)dx[k] * t.iHs;
      expt += sk * sk;
    }
    t.rho_star += (double)(t.norm * (R)exp_neg_w<FX>((double)expt, s.etab) * s.m[q]);
  }
  __device__ static void end_pass(const Args &a, Tgt &t, int, uint32_t, bool active) {
    if (!active || t.done) return;
    double Hold = t.hcur;
    double Hnew = (a.dim == 3) ? a.eta * cbrt(t.mi / t.rho_star) : a.eta * (t.mi / t.rho_star);
    t.hcur = a.ratio * Hnew;
    if (fabs((Hnew - Hold) / Hold) < a.thresh) t.done = true;
  }
  __device__ static void finish(const Args &a, Tgt &t, uint32_t i, int *) {
    a.h[i] = t.hcur;
    a.done[i] = t.done ? 1u : 0u;
  }
};

// ---------------------------------------------------------------------------
// List materialisation (parity checks only): the candidates the skeleton's
// membership rule accepts are exactly the reference's list (neighs.cc:37-211).
// mode 0 counts, mode 1 writes device indices at offsets[i].
// ---------------------------------------------------------------------------
struct ListPass {
  template <class T> __device__ static void init(T &, int) {}
  struct Args {
    int mode;
    uint32_t *count;
    const uint32_t *offset;
    uint32_t *out;
    double cutoff;
  };
  struct Tile { int dummy; };
  struct Tgt { uint32_t n, o; double supp; };
  static constexpr int kSub = 1;
  __device__ static void combine(Tgt &) {}
  __device__ static int npasses(const Args &) { return 1; }
  __device__ static void begin_pass(const Args &, Tgt &, int) {}
  __device__ static void end_pass(const Args &, Tgt &, int, uint32_t, bool) {}
  __device__ static bool want(const Tgt &) { return true; }
  __device__ static void load_target(const Args &a, uint32_t i, Tgt &t) {
    t.n = 0;
    t.o = a.mode ? a.offset[i] : 0;
    t.supp = HUGE_VAL;
  }
  __device__ static void stage(const Args &, uint32_t, Tile &, int) {}
  __device__ static double bound(const Args &, const Tgt &, const Tile &, uint32_t) { return HUGE_VAL; }
  __device__ static double hker(const Args &, const Tgt &, const Tile &, uint32_t) { return HUGE_VAL; }
  __device__ static bool extra(const Args &, const Tgt &, const Tile &, uint32_t, uint32_t, uint32_t) { return true; }
  template <int D>
  __device__ static void pair_d(const Args &a, Tgt &t, const Tile &, uint32_t, uint32_t j, uint32_t,
                                const double *, double) {
    if (a.mode) a.out[t.o + t.n] = j;
    t.n++;
  }
  __device__ static void finish(const Args &a, Tgt &t, uint32_t i, int *) {
    if (!a.mode) a.count[i] = t.n;
  }
};

// Binds the dimension into the pass's pair routine for the skeleton.
template <int D, class P>
struct Bound : P {
  // dx: the pair's position difference; amax: max_k |dx_k| of the positions
  // the predicates read (the same dx except in a frozen replay)
  __device__ static void pair(const typename P::Args &a, typename P::Tgt &t, const typename P::Tile &s,
                             uint32_t q, uint32_t j, uint32_t i, const double *dx, double amax) {
    P::template pair_d<D>(a, t, s, q, j, i, dx, amax);
  }
};

// ---------------------------------------------------------------------------
// Wall-ghost refresh: apply_bcond (bcond.cc:24-249) with the mirror images of
// search_bnd_images (bnd_image.cc:33-254) for a flat ground at z = zg.  Each
// wall ghost of the bucket layer that straddles the ground (bucket type
// MIXED) takes the state at its image (x, y, 2 zg - z): Shepard sums over the
// real and pressure-ghost particles j of the image's bucket and its 26
// neighbours, m_j W(s, h_j) with s = (image - x_j) / h_j inside the in_support
// box CUTOFF2 h_j, of rho, v / rho and e / rho over sum m_j W / rho_j; the
// mass fraction sums the neighbour buckets only (bcond.cc:140-146); the
// velocity is reflected in the ground plane and the EOS re-evaluated.
// Candidates come from the cell grid of the step's neighbour search, with the
// reach widened by the largest displacement since (positions have moved, the
// grid has not).  Only bc 2 entries are written and only bc 100/99/1 entries
// read, so ghosts are independent.
// ---------------------------------------------------------------------------
struct BcondArgs {
  const double *x, *y, *z, *h, *m;
  double *rho, *vx, *vy, *vz, *e, *xi, *p, *c, *T, *gam, *cp;
  const int32_t *bc;
  const uint32_t *cs, *toff;
  const float4 *tbox;        // tile boxes of the step's grid (positions at grid time)
  GridDesc g;
  double reach, zg, del, cutoff2;
  double ext_lo[3];
  int nb[3], bzg;
  CutTest cut;
  EosConst eos;
  double *wrec;              // recording for the adjoint: Shepard normaliser of each refreshed ghost
  uint32_t wrec_n;
  const double *px, *py, *pz;   // positions the support tests read (x, y, z except in a frozen replay)
};

// unit normal of the mirror plane, from a ghost towards its image
__device__ inline void mirror_normal(const double *coord, const double *refc, double *nrm) {
  double d = 0;
  for (int k = 0; k < 3; k++) { nrm[k] = refc[k] - coord[k]; d += nrm[k] * nrm[k]; }
  if (!(d > 0)) { nrm[0] = nrm[1] = 0; nrm[2] = 1; d = 1; }
  const double sd = sqrt(d);
  for (int k = 0; k < 3; k++) nrm[k] /= sd;
}

__device__ inline bool bucket_of(const BcondArgs &a, const double *p, int *b) {
  bool ok = true;
  for (int k = 0; k < 3; k++) {
    const double f = floor((p[k] - a.ext_lo[k]) / a.del);
    ok = ok && f >= 0 && f < a.nb[k];
    b[k] = ok ? (int)f : 0;
  }
  return ok;
}

// One wavefront per kBcondSpan consecutive particles (in cell order, so a
// wave's ghosts and their images are compact): the wave stages the candidate cells'
// particles 64 at a time into LDS, with their bucket indices and 1/rho; each
// lane holding an imaged ghost marks the staged candidates that reach its
// image, then walks only those, in index order.
// consecutive particles per wavefront: 32 (lanes 32-63 only stage and test
// candidates) gives twice the waves to hide the staging latency behind;
// 1.74 ms per refresh at 4M against 1.83 ms with 64 and 2.19 ms with 16
constexpr uint32_t kBcondSpan = 32;
__global__ __launch_bounds__(kWave) void k_bcond(uint32_t n, BcondArgs a) {
  __shared__ double sx[kWave], sy[kWave], sz[kWave], sh[kWave], smw[kWave], sir[kWave];
  __shared__ double svx[kWave], svy[kWave], svz[kWave], se[kWave], sxi[kWave];
  __shared__ double sqx[kWave], sqy[kWave], sqz[kWave];
  __shared__ int sbk[3][kWave];
  const int lane = threadIdx.x;
  const uint32_t i = blockIdx.x * kBcondSpan + lane;
  double coord[3] = {0, 0, 0}, refc[3] = {0, 0, 0};
  int bi[3] = {0, 0, 0};
  bool act = false;
  if ((uint32_t)lane < kBcondSpan && i < n && a.bc[i] == 2) {
    coord[0] = a.x[i]; coord[1] = a.y[i]; coord[2] = a.z[i];
    int bg[3];
    if (bucket_of(a, coord, bg) && bg[2] == a.bzg) {
      refc[0] = coord[0]; refc[1] = coord[1]; refc[2] = 2 * a.zg - coord[2];
      act = bucket_of(a, refc, bi);
    }
  }
  if (!__any(act)) return;
  // Each active lane's cells within reach of its image form a box, walked
  // in increasing cell index with a cursor; the wave visits the union of the
  // boxes in that order (the smallest pending cell of any lane next).  A
  // lane's hits therefore come cell by cell in increasing index, as in the
  // bounding-box walk, without the cells between far-apart images (a wave's
  // ghosts can straddle two rows of cells).
  const int nn[3] = {a.g.nx, a.g.ny, a.g.nz};
  int c0[3], c1[3], cur[3];
  for (int k = 0; k < 3; k++) {
    if (a.g.inv_e > 0) {
      c0[k] = cell_of(refc[k] - a.reach, a.g.lo[k], a.g.inv_e, nn[k]);
      c1[k] = cell_of(refc[k] + a.reach, a.g.lo[k], a.g.inv_e, nn[k]);
    } else {
      c0[k] = c1[k] = 0;
    }
    cur[k] = c0[k];
  }
  constexpr uint32_t kDone = 0xffffffffu;
  uint32_t mine = act ? ((uint32_t)cur[2] * a.g.ny + cur[1]) * a.g.nx + cur[0] : kDone;
  double u0 = 0, u1 = 0, u2 = 0, u3 = 0, u4 = 0, wnorm = 0, ks = 0;
  for (;;) {
    uint32_t cell = mine;
    for (int o = 32; o > 0; o >>= 1) cell = min(cell, (uint32_t)__shfl_xor((int)cell, o));
    if (cell == kDone) break;
    const bool here = mine == cell;
    if (here) {
      if (++cur[0] > c1[0]) {
        cur[0] = c0[0];
        if (++cur[1] > c1[1]) {
          cur[1] = c0[1];
          ++cur[2];
        }
      }
      mine = cur[2] > c1[2] ? kDone : ((uint32_t)cur[2] * a.g.ny + cur[1]) * a.g.nx + cur[0];
    }
    const uint32_t lo = a.cs[cell], hi = a.cs[cell + 1];
    uint32_t tl = a.toff[cell];
    for (uint32_t base = lo; base < hi; base += kWave, tl++) {
      // A tile is staged only if some lane of this cell's box can reach it:
      // its particles have moved by at most the displacement that the reach
      // includes, so the tile box at grid time within reach of the image is
      // necessary for any of them to be within the kernel box now.
      const float4 tlo = a.tbox[2 * tl], thi = a.tbox[2 * tl + 1];
      const bool near_t = here && fmax((double)tlo.x - refc[0], refc[0] - (double)thi.x) <= a.reach &&
                          fmax((double)tlo.y - refc[1], refc[1] - (double)thi.y) <= a.reach &&
                          fmax((double)tlo.z - refc[2], refc[2] - (double)thi.z) <= a.reach;
      if (!__any(near_t)) continue;
      const uint32_t j = base + lane;
      bool ok = false;
      if (j < hi) {
        const int b = a.bc[j];
        const double qj[3] = {a.px[j], a.py[j], a.pz[j]};
        int bj[3];
        ok = (b == 100 || b == 99 || b == 1) && bucket_of(a, qj, bj);   // contr_image (particle.h:562-565)
        if (ok) {
          sx[lane] = a.x[j]; sy[lane] = a.y[j]; sz[lane] = a.z[j];
          sqx[lane] = qj[0]; sqy[lane] = qj[1]; sqz[lane] = qj[2];
          sh[lane] = a.h[j]; smw[lane] = a.m[j]; sir[lane] = 1.0 / a.rho[j];
          svx[lane] = a.vx[j]; svy[lane] = a.vy[j]; svz[lane] = a.vz[j];
          se[lane] = a.e[j]; sxi[lane] = a.xi[j];
          for (int k = 0; k < 3; k++) sbk[k][lane] = bj[k];
        }
      }
      if (!ok) sbk[0][lane] = -1000;            // never within one bucket of an image
      const bool any_contr = __any(ok);          // e.g. tiles of wall ghosts only
      __syncthreads();
      const uint32_t cnt = min((uint32_t)kWave, hi - base);
      if (any_contr) {
        // phase 1, transposed: lane l tests staged candidate l against the
        // image of each lane within reach of the tile, one lane g at a time;
        // lane g keeps the ballot (bit q: candidate q reaches g's image)
        uint64_t mask = 0;
        uint64_t need = __ballot(near_t);
        const bool cand = (uint32_t)lane < cnt;
        const double ox = sqx[lane], oy = sqy[lane], oz = sqz[lane], oh = sh[lane];
        const int o0 = sbk[0][lane], o1 = sbk[1][lane], o2 = sbk[2][lane];
        while (need) {
          const int g = __builtin_ctzll(need);
          need &= need - 1;
          const double r0 = __shfl(refc[0], g), r1 = __shfl(refc[1], g), r2 = __shfl(refc[2], g);
          const int b0 = __shfl(bi[0], g), b1 = __shfl(bi[1], g), b2 = __shfl(bi[2], g);
          const bool near = abs(o0 - b0) <= 1 && abs(o1 - b1) <= 1 && abs(o2 - b2) <= 1;
          const double amax = fmax(fabs(r0 - ox), fmax(fabs(r1 - oy), fabs(r2 - oz)));
          // in_support; W = 0 outside the kernel box
          const bool in = cand & near & (amax <= a.cutoff2 * oh) & !a.cut.outside(amax, oh);
          const uint64_t bal = __ballot(in);
          mask = lane == g ? bal : mask;
        }
        // phase 2: the lane's own hits, in index order
        while (mask) {
          const uint32_t q = (uint32_t)__builtin_ctzll(mask);
          mask &= mask - 1;
          const double hj = sh[q];
          const double dx[3] = {refc[0] - sx[q], refc[1] - sy[q], refc[2] - sz[q]};
          const double ih = 1.0 / hj;
          double expt = 0;
          for (int k = 0; k < 3; k++) { const double sk = dx[k] * ih; expt += sk * sk; }
          const double nk = kInvSqrtPi * ih;
          const double w = nk * nk * nk * exp(-expt);
          const double ir = sir[q];
          const double mw = smw[q] * w;
          u0 += mw;
          u1 += mw * svx[q] * ir;
          u2 += mw * svy[q] * ir;
          u3 += mw * svz[q] * ir;
          u4 += mw * se[q] * ir;
          wnorm += mw * ir;
          if (sbk[0][q] != bi[0] || sbk[1][q] != bi[1] || sbk[2][q] != bi[2]) ks += mw * sxi[q] * ir;
        }
      }
      __syncthreads();
    }
  }
  if (!act) return;
  if (a.wrec && i < a.wrec_n) a.wrec[i] = wnorm;
  double nrm[3];
  mirror_normal(coord, refc, nrm);
  if (wnorm > 0) {
    u0 /= wnorm; u1 /= wnorm; u2 /= wnorm; u3 /= wnorm; u4 /= wnorm;
    ks = ks / wnorm;
  } else {
    u0 = 1; u1 = u2 = u3 = u4 = 0; ks = 0;
  }
  const double vdotn = u1 * nrm[0] + u2 * nrm[1] + u3 * nrm[2];
  a.rho[i] = u0;
  a.vx[i] = u1 - 2 * vdotn * nrm[0];
  a.vy[i] = u2 - 2 * vdotn * nrm[1];
  a.vz[i] = u3 - 2 * vdotn * nrm[2];
  a.e[i] = u4;
  a.xi[i] = ks;
  double p, c, T, gam, cp = a.cp[i];
  eos_dev(a.eos, u0, u4, ks, p, c, T, gam, &cp);
  a.p[i] = p; a.c[i] = c; a.T[i] = T; a.gam[i] = gam; a.cp[i] = cp;
}

// largest per-axis speed of the particles update_pos moves (reals with the
// filtered velocity under LANS, erupt ghosts with v)
__global__ __launch_bounds__(256) void k_speed_max(uint32_t n, const int32_t *bc, int lans, const double *vx,
                                                   const double *vy, const double *vz, const double *svx,
                                                   const double *svy, const double *svz, unsigned long long *out) {
  double v = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int b = bc[i];
    if (b == 100 || b == 99) {
      const bool s = lans != 0;
      v = fmax(v, fmax(fabs(s ? svx[i] : vx[i]), fmax(fabs(s ? svy[i] : vy[i]), fabs(s ? svz[i] : vz[i]))));
    } else if (b == 0) {
      v = fmax(v, fmax(fabs(vx[i]), fmax(fabs(vy[i]), fabs(vz[i]))));
    }
  }
  v = block_ext<true>(v);
  if (threadIdx.x == 0 && v > 0) atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// Position update + escape (update_pos.cc:34-168, bucket.cc:435-452)
// ---------------------------------------------------------------------------
// Schedule of a recorded run: tmode 1 records, by particle id, the tape step
// at which an erupt ghost turns real and a real particle escapes; tmode 2
// replays those events instead of testing positions.
struct PosTape {
  int mode;
  int32_t k;
  const uint32_t *pid;
  int32_t *tr_step, *esc_step;
};

__global__ void k_update_pos(uint32_t n, int dim, int lans, int erupting, double dt, double *x,
                             double *y, double *z, const double *vx, const double *vy,
                             const double *vz, const double *svx, const double *svy,
                             const double *svz, int32_t *bc, double *h, int32_t *inv,
                             double ground, double h_new, double lo0, double lo1, double lo2,
                             double hi0, double hi1, double hi2, unsigned long long *dead, PosTape tp) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int b = bc[i];
  if (b == 100 || b == 99) {
    if (dim == 3) {
      double ux = lans ? svx[i] : vx[i], uy = lans ? svy[i] : vy[i], uz = lans ? svz[i] : vz[i];
      x[i] = x[i] + dt * ux; y[i] = y[i] + dt * uy; z[i] = z[i] + dt * uz;
    } else {
      x[i] = x[i] + dt * (lans ? svx[i] : vx[i]);
    }
  } else if (b == 0 && erupting) {
    x[i] = x[i] + dt * vx[i];
    if (dim == 3) {
      y[i] = y[i] + dt * vy[i];
      z[i] = z[i] + dt * vz[i];
      const bool turn = tp.mode == 2 ? tp.tr_step[tp.pid[i]] == tp.k : z[i] >= ground;
      if (turn) {
        b = 100; bc[i] = 100; h[i] = h_new; inv[i] = 2;
        if (tp.mode == 1) tp.tr_step[tp.pid[i]] = tp.k;
      }
    }
  }
  if (dim == 3 && (b == 100 || b == 99)) {
    const double TINY = 1.0E-08;
    bool esc = (x[i] - lo0) < -TINY || (y[i] - lo1) < -TINY || (z[i] - lo2) < -TINY ||
               (x[i] - hi0) > TINY || (y[i] - hi1) > TINY || (z[i] - hi2) > TINY;
    if (tp.mode == 2) esc = tp.esc_step[tp.pid[i]] == tp.k;
    if (esc) {
      bc[i] = -1; atomicAdd(dead, 1ull);
      if (tp.mode == 1) tp.esc_step[tp.pid[i]] = tp.k;
    }
  }
}

// real particles (bc 100/99): grid-stride count, one atomic per block
__global__ __launch_bounds__(256) void k_count_real(uint32_t n, const int32_t *bc, unsigned long long *out) {
  __shared__ unsigned int part[256 / kWave];
  unsigned int cnt = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int b = bc[i];
    cnt += (b == 100 || b == 99) ? 1u : 0u;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) cnt += (unsigned int)__shfl_xor((int)cnt, o);
  if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sum = 0;
    for (int w = 0; w < 256 / kWave; w++) sum += part[w];
    if (sum) atomicAdd(out, sum);
  }
}

__global__ void k_iota_u32(uint32_t *p, uint32_t first, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = first + i;
}
// tab[id] = h (record) or h = tab[id] (replay) for the particles present
__global__ void k_h_by_id(uint32_t n, const uint32_t *pid, double *h, double *tab, int replay) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (replay) h[i] = tab[pid[i]];
  else tab[pid[i]] = h[i];
}

// positions of a step by particle id (rec: x block, y block, z block of nrec)
__global__ void k_pos_to_rec(uint32_t n, const uint32_t *pid, const double *x, const double *y, const double *z,
                             double *rec, uint32_t nrec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = pid[i];
  if (id >= nrec) return;
  rec[id] = x[i]; rec[(size_t)nrec + id] = y[i]; rec[2 * (size_t)nrec + id] = z[i];
}
// and back, into the current order; a live particle without a record is an
// error (the replay no longer follows the recorded run)
__global__ void k_pos_from_rec(uint32_t n, const uint32_t *pid, const int32_t *bc, const double *rec,
                               uint32_t nrec, const double *x, const double *y, const double *z, double *px,
                               double *py, double *pz, int *err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = pid[i];
  if (bc[i] < 0 || id >= nrec) {
    if (bc[i] >= 0) raise_err(err, PSPH_EINVARIANT, i);
    px[i] = x[i]; py[i] = y[i]; pz[i] = z[i];
    return;
  }
  px[i] = rec[id]; py[i] = rec[(size_t)nrec + id]; pz[i] = rec[2 * (size_t)nrec + id];
}

// fixed-order block partial sums (deterministic)
constexpr int kRedBlocks = 512;
__global__ void k_conservation(uint32_t n, const int32_t *bc, const double *m, const double *e,
                               const double *vx, const double *vy, const double *vz,
                               const double *xi, const double *z, double *part) {
  __shared__ double sh[8][256];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int b = bc[i];
    if (!(b == 100 || b == 99)) continue;
    double mi = m[i], v0 = vx[i], v1 = vy ? vy[i] : 0, v2 = vz ? vz[i] : 0;
    acc[0] += mi;
    acc[1] += mi * e[i];
    acc[2] += 0.5 * mi * (v0 * v0 + v1 * v1 + v2 * v2);
    acc[3] += mi * v0; acc[4] += mi * v1; acc[5] += mi * v2;
    acc[6] += mi * xi[i];
    acc[7] += mi * xi[i] * (z ? z[i] : 0);
  }
  for (int k = 0; k < 8; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 8; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 8; k++) part[blockIdx.x * 8 + k] = sh[k][0];
}

inline unsigned blocks(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }
// grid of a grid-stride reduction: enough blocks to fill the chip, no more
inline unsigned red_blocks(size_t n) { return std::min(blocks(n), 2048u); }

}  // namespace

// ===========================================================================
// Context
// ===========================================================================
#define PSPH_DFIELDS(F) F(x) F(y) F(z) F(vx) F(vy) F(vz) F(svx) F(svy) F(svz) F(rho) F(e) F(m) \
  F(h) F(p) F(c) F(T) F(gam) F(cp) F(xi)
#define PSPH_UFIELDS(F) F(k0) F(k1) F(k2)
#define PSPH_IFIELDS(F) F(bc) F(phase) F(inv)

struct DevFields {
#define D_(a) double *a = nullptr;
#define U_(a) uint32_t *a = nullptr;
#define I_(a) int32_t *a = nullptr;
  PSPH_DFIELDS(D_) PSPH_UFIELDS(U_) PSPH_IFIELDS(I_)
#undef D_
#undef U_
#undef I_
};

// Frozen schedule of a recorded run (the adjoint and its finite-difference
// checks differentiate the step map with these discrete choices held fixed).
struct Tape {
  int mode = 0;                             // 0 off, 1 record, 2 replay
  int64_t k = 0;                            // steps taken since the start of the tape
  double t0 = 0;
  int64_t step0 = 0;
  size_t n0 = 0;
  std::vector<double> dt;                   // per step, after TimeProps::incrtime
  std::vector<int> hslot;                   // per step: adapted-h table, or -1
  std::vector<double *> htab;               // smoothing lengths after adaptation, by particle id
  std::vector<std::vector<int32_t>> inlet;  // per step: (q, n) per vent column
  int32_t *tr_step = nullptr, *esc_step = nullptr;   // by id, -1 = never
  size_t id_cap = 0;
  // Positions at each step's neighbour search, by particle id (x block, y
  // block, z block of xrec_n[k] entries).  A replay with `frozen` set builds
  // its cell order, tiles, visit lists and every pair predicate (kernel
  // boxes, list membership, support) from these, and only the pair
  // arithmetic from its own positions: the pair set and the summation order
  // are the recorded run's, so a finite difference through the replay sees
  // no kernel-box face crossings, as the adjoint does not.
  std::vector<double *> xrec;
  std::vector<uint32_t> xrec_n;
  // with cfg.wall_bc, the same at each step's wall-ghost refresh (after
  // update_pos): a frozen replay's refresh takes its contributor sets from them
  std::vector<double *> xrecb;
  std::vector<uint32_t> xrecb_n;
  bool frozen = false;
};

// What the reverse sweep needs of one step, in that step's cell order.
enum RecField {
  RF_X, RF_Y, RF_Z, RF_H, RF_M, RF_RHO, RF_P, RF_C, RF_T, RF_CP, RF_SVX, RF_SVY, RF_SVZ,
  RF_VX, RF_VY, RF_VZ, RF_E, RF_HLIST,                  // state after sort and h update
  RF_V1X, RF_V1Y, RF_V1Z, RF_E1,                         // after the RHS update
  RF_RHO2, RF_E2, RF_XI2, RF_WNE,                        // after the density pass
  RF_SV2X, RF_SV2Y, RF_SV2Z, RF_WNV,                     // after the velocity filter
  // wall-ghost refresh (only with cfg.wall_bc): positions after update_pos,
  // the refreshed ghost fields and each ghost's Shepard normaliser
  RF_XN, RF_YN, RF_ZN, RF_BRHO, RF_BVX, RF_BVY, RF_BVZ, RF_BE, RF_BXI, RF_BW,
  RF_N
};
constexpr int RF_NBASE = RF_XN;                          // fields recorded in every step
struct StepRec {
  size_t n = 0, cap = 0;
  double *d[RF_N] = {};
  uint32_t *pid = nullptr, *refb = nullptr, *ckey = nullptr;
  uint32_t *gflag = nullptr;                  // boundary ghosts met as neighbours in this step
  int32_t *bc = nullptr, *phase = nullptr;
  int32_t *bcn = nullptr;                     // bc after update_pos (wall-ghost refresh only)
  uint32_t *cs = nullptr, *toff = nullptr;   // cell starts and first tile of each cell
  float4 *tbox = nullptr;                     // tile boxes
  size_t cs_cap = 0, tile_cap = 0;
  GridDesc grid{};
  double reach = 0;
  double hmax = 0;                            // bounds h at list time
  double dt = 0, time = 0;
  int erupting = 0;
  uint32_t inj_lo = 0, inj_hi = 0;          // ids of the particles this step injected
  std::vector<double> inj_vel;              // their column exit velocity
};
struct Adjoint;

struct psph_ctx {
  psph_config cfg{};
  hipStream_t stream = nullptr;
  size_t cap = 0, n = 0;
  DevFields f;
  double *dsp[10] = {};        // spare double arrays (gather / double buffers)
  uint32_t *usp = nullptr;     // spare u32 array
  int32_t *isp = nullptr;      // spare i32 array
  double *h_list = nullptr;
  uint32_t *refb = nullptr, *key_a = nullptr, *key_b = nullptr, *idx_a = nullptr, *idx_b = nullptr;
  uint32_t *flags = nullptr, *tpos = nullptr, *targets = nullptr;
  uint32_t *cfirst = nullptr, *ccount = nullptr;
  uint32_t *cs = nullptr, *tcnt = nullptr, *toff = nullptr, *cchunks = nullptr, *coff = nullptr, *nchunks = nullptr;
  float4 *tbox = nullptr;
  size_t cs_cap = 0, tile_cap = 0, chunk_cap = 0;
  void *cub = nullptr;
  size_t cub_bytes = 0;
  unsigned long long *d_u64 = nullptr;   // [0] dt bits [1] hmax bits [2] pairs [3] dead
  int *d_err = nullptr;
  double *d_part = nullptr;
  // grid of the last neighbour search
  bool grid_valid = false;
  GridDesc grid{};
  double reach_main = 0;       // per-axis reach of the RHS / density / filter predicates
  // visit lists of the RHS / density / filter passes (built once per step)
  bool vis_valid = false;
  uint32_t *vcount = nullptr, *voff = nullptr, *vnum = nullptr, *vmask = nullptr;
  uint2 *vent = nullptr;
  size_t vchunk_cap = 0, vis_cap = 0;
  double h_max = 0;
  // clock (TimeProps)
  double time = 0, next_output = 0;
  int64_t step = 0;
  int ioutput = 0;
  int64_t last_injected = 0, last_removed = 0, last_real = -1;
  unsigned long long last_pairs = 0;
  HostParticles inj;
  // persistent particle ids: load order, then injection order
  uint32_t *pid = nullptr, *pid_sp = nullptr;
  uint32_t next_pid = 0;
  // positions the pair predicates read in a frozen replay (Tape::xrec)
  double *px = nullptr, *py = nullptr, *pz = nullptr;
  // domain adaptation (psph_domain_init, psph_domain.inl)
  DomainBuckets dom;
  bool dom_ready = false;
  int dom_pending = 0;                 // scan_outside_layer's count of the last step
  std::vector<double> met_z, met_data;
  psph_meteo met{};
  uint32_t dom_seed = 0;
  double dom_jitter = 0;
  uint32_t *d_occ = nullptr;
  uint8_t *d_conv = nullptr, *d_sides = nullptr;
  int64_t last_domain = 0;
  Tape tape;
  StepRec *rec = nullptr;      // set: psph_step captures the step for the reverse sweep
  Adjoint *adj = nullptr;
  // per-stage HIP-event timing
  bool prof = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double prof_ms[PSPH_NSTAGES] = {};
  int64_t prof_n[PSPH_NSTAGES] = {};
};

// Brackets the main kernel(s) of a stage with events on the context stream.
struct StageTimer {
  psph_ctx *c;
  int id;
  StageTimer(psph_ctx *ctx, int stage) : c(ctx), id(stage) {
    if (c->prof) (void)hipEventRecord(c->ev0, c->stream);
  }
  void stop() {
    if (!c->prof || id < 0) return;
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
      c->prof_ms[id] += ms;
      c->prof_n[id] += 1;
    }
    id = -1;
  }
  ~StageTimer() { stop(); }
};

static int dmalloc(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    (void)hipGetLastError();             // reported here, not again by the next check
    *p = nullptr;
    return fail(PSPH_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return 0;
}

static int check_err(psph_ctx *c, const char *stage) {
  int he[2];
  HIP_TRY(hipMemcpyAsync(he, c->d_err, sizeof he, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (he[0]) {
    int code = he[0];
    hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, c->stream, c->d_err, 0, 2);
    return fail(code, "%s: invariant violated at device particle %d (density<=0, mass fraction>1, "
                      "wnorm<=0 or sound speed<=0)", stage, he[1]);
  }
  return 0;
}

static int ensure_cub(psph_ctx *c, size_t bytes) {
  if (bytes <= c->cub_bytes) return 0;
  if (c->cub) hipFree(c->cub);
  c->cub = nullptr;
  PSPH_TRY(dmalloc(&c->cub, bytes));
  c->cub_bytes = bytes;
  return 0;
}

extern "C" int psph_create(const psph_config *cfg, size_t capacity, psph_ctx **out) {
  if (!cfg || !out) return fail(PSPH_EINVAL, "null argument");
  if (cfg->dim != 1 && cfg->dim != 3) return fail(PSPH_EINVAL, "dim must be 1 or 3");
  if (cfg->precision != PSPH_FP64 && cfg->precision != PSPH_FP32_PAIR && cfg->precision != PSPH_FP32_PURE &&
      cfg->precision != PSPH_FP32_EXP)
    return fail(PSPH_EINVAL, "unknown precision form %d", cfg->precision);
  if (cfg->kernel != PSPH_KERNEL_GAUSS_BOX)
    return fail(PSPH_EINVAL, "kernel %d: the reference's one kernel, the Gaussian box (PSPH_KERNEL_GAUSS_BOX), "
                             "is the only one", cfg->kernel);
  if (!(cfg->h_air > 0)) return fail(PSPH_EINVAL, "h_air must be > 0");
  {
    uint64_t bits;
    std::memcpy(&bits, &cfg->cutoff, 8);
    if (!(cfg->cutoff > 0) || (bits & 1u))
      return fail(PSPH_ENOTSUP, "cutoff %g: the exact kernel-box test needs a positive cut-off with an even "
                                "last significand bit", cfg->cutoff);
  }
  if (cfg->dim == 3 && !(cfg->bucket_edge > 0)) return fail(PSPH_EINVAL, "call psph_config_domain first");
  if (capacity < 64) capacity = 64;
  if (capacity > 0xffffff00u) return fail(PSPH_EINVAL, "capacity above 2^32 particles");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PSPH_EHIP, "no HIP device");
  psph_ctx *c = new psph_ctx();
  c->cfg = *cfg;
  c->cap = capacity;
  int rc = 0;
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail(PSPH_EHIP, "hipStreamCreate"); }
#define A_(a) if (!rc) rc = dmalloc((void **)&c->f.a, capacity * sizeof(*c->f.a));
  PSPH_DFIELDS(A_) PSPH_UFIELDS(A_) PSPH_IFIELDS(A_)
#undef A_
  for (int k = 0; k < 10 && !rc; k++) rc = dmalloc((void **)&c->dsp[k], capacity * 8);
  if (!rc) rc = dmalloc((void **)&c->usp, capacity * 4);
  if (!rc) rc = dmalloc((void **)&c->isp, capacity * 4);
  if (!rc) rc = dmalloc((void **)&c->h_list, capacity * 8);
  if (!rc) rc = dmalloc((void **)&c->pid, capacity * 4);
  if (!rc) rc = dmalloc((void **)&c->pid_sp, capacity * 4);
  uint32_t **u32s[] = {&c->refb, &c->key_a, &c->key_b, &c->idx_a, &c->idx_b, &c->targets};
  for (auto pp : u32s)
    if (!rc) rc = dmalloc((void **)pp, capacity * 4);
  if (!rc) rc = dmalloc((void **)&c->flags, (capacity + 1) * 4);
  if (!rc) rc = dmalloc((void **)&c->tpos, (capacity + 1) * 4);
  if (!rc) rc = dmalloc((void **)&c->nchunks, 4);
  if (!rc) rc = dmalloc((void **)&c->d_u64, 8 * 8);
  if (!rc) rc = dmalloc((void **)&c->d_err, 2 * sizeof(int));
  if (!rc) rc = dmalloc((void **)&c->d_part, kRedBlocks * 8 * 8);
  if (rc) { psph_destroy(c); return rc; }
  hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, c->stream, c->d_err, 0, 2);
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64, 0ull, 8);
  c->next_output = cfg->t_output;
  *out = c;
  return 0;
}

static void adjoint_free(psph_ctx *c);
static void tape_clear(Tape &tp);

extern "C" int psph_destroy(psph_ctx *c) {
  if (!c) return 0;
  if (c->stream) hipStreamSynchronize(c->stream);
  adjoint_free(c);
  tape_clear(c->tape);
  if (c->tape.tr_step) hipFree(c->tape.tr_step);
  if (c->tape.esc_step) hipFree(c->tape.esc_step);
#define F_(a) if (c->f.a) hipFree(c->f.a);
  PSPH_DFIELDS(F_) PSPH_UFIELDS(F_) PSPH_IFIELDS(F_)
#undef F_
  for (auto p : c->dsp) if (p) hipFree(p);
  void *ptrs[] = {c->usp, c->isp, c->h_list, c->pid, c->pid_sp, c->refb, c->key_a, c->key_b, c->idx_a, c->idx_b,
                  c->flags, c->tpos, c->targets, c->cfirst, c->ccount, c->cs, c->tcnt, c->toff,
                  c->cchunks, c->coff, c->tbox, c->nchunks, c->cub, c->d_u64, c->d_err, c->d_part,
                  c->vcount, c->voff, c->vnum, c->vmask, c->vent, c->px, c->py, c->pz,
                  c->d_occ, c->d_conv, c->d_sides};
  for (auto p : ptrs) if (p) hipFree(p);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

static int upload(psph_ctx *c, const HostParticles &P, size_t at) {
  size_t n = P.size();
  if (!n) return 0;
#define U_(a) HIP_TRY(hipMemcpyAsync(c->f.a + at, P.a.data(), n * sizeof(P.a[0]), hipMemcpyHostToDevice, c->stream));
  PSPH_DFIELDS(U_) PSPH_UFIELDS(U_)
  U_(bc) U_(phase)
#undef U_
  HIP_TRY(hipMemcpyAsync(c->f.inv + at, P.inv.data(), n * 4, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int psph_load(psph_ctx *c, const psph_soa *host, size_t n, double t0, int64_t step0) {
  if (!c || !host) return fail(PSPH_EINVAL, "null argument");
  if (n > c->cap) return fail(PSPH_ECAPACITY, "load: %zu particles exceed capacity %zu", n, c->cap);
  if (!host->x || !host->rho || !host->m || !host->h || !host->bc) return fail(PSPH_EINVAL, "load: x, rho, m, h and bc are required");
  HostParticles P;
  P.from_soa(host, n);
  PSPH_TRY(upload(c, P, 0));
  c->n = n;
  if (n) hipLaunchKernelGGL(k_iota_u32, blocks(n), 256, 0, c->stream, c->pid, 0u, (uint32_t)n);
  c->next_pid = (uint32_t)n;
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 3, 0ull, 1);
  c->time = t0;
  c->step = step0;
  c->ioutput = 0;
  c->next_output = c->cfg.t_output;
  while (c->next_output <= t0 && c->cfg.t_output > 0) { c->ioutput++; c->next_output = (c->ioutput + 1) * c->cfg.t_output; }
  c->grid_valid = false;
  return 0;
}

extern "C" int psph_count(psph_ctx *c, size_t *n_total, size_t *n_real) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  std::vector<int32_t> bc(c->n);
  if (c->n) HIP_TRY(hipMemcpy(bc.data(), c->f.bc, c->n * 4, hipMemcpyDeviceToHost));
  size_t nt = 0, nr = 0;
  for (auto b : bc) { if (b >= 0) nt++; if (b == 100 || b == 99) nr++; }
  if (n_total) *n_total = nt;
  if (n_real) *n_real = nr;
  return 0;
}

extern "C" int psph_device_view(psph_ctx *c, psph_soa *dev, size_t *n, void **stream) {
  if (!c || !dev) return fail(PSPH_EINVAL, "null argument");
  // finish pending work so that the caller may use the arrays on any stream
  HIP_TRY(hipStreamSynchronize(c->stream));
#define V_(a) dev->a = c->f.a;
  PSPH_DFIELDS(V_)
  V_(bc) V_(phase)
#undef V_
  dev->key0 = c->f.k0; dev->key1 = c->f.k1; dev->key2 = c->f.k2;
  dev->involved = c->f.inv;
  if (n) *n = c->n;
  if (stream) *stream = (void *)c->stream;
  return 0;
}

extern "C" int psph_get_config(psph_ctx *c, psph_config *cfg) {
  if (!c || !cfg) return fail(PSPH_EINVAL, "null argument");
  *cfg = c->cfg;
  return 0;
}

extern "C" int psph_time(psph_ctx *c, double *time, int64_t *step) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (time) *time = c->time;
  if (step) *step = c->step;
  return 0;
}

static int download(psph_ctx *c, HostParticles &P) {
  size_t n = c->n;
  P.resize(n);
  if (!n) return 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
#define D_(a) HIP_TRY(hipMemcpy(P.a.data(), c->f.a, n * sizeof(P.a[0]), hipMemcpyDeviceToHost));
  PSPH_DFIELDS(D_) PSPH_UFIELDS(D_) PSPH_IFIELDS(D_)
#undef D_
  return 0;
}

extern "C" int psph_read(psph_ctx *c, psph_soa *host, size_t cap, size_t *n_out) {
  if (!c || !n_out) return fail(PSPH_EINVAL, "null argument");
  HostParticles P;
  PSPH_TRY(download(c, P));
  // drop removed particles, then THashTable order
  std::vector<uint32_t> keep;
  keep.reserve(P.size());
  for (size_t i = 0; i < P.size(); i++) if (P.bc[i] >= 0) keep.push_back((uint32_t)i);
  if (keep.size() != P.size()) { P.permute(keep); }
  auto perm = canonical_order(c->cfg, P.k0.data(), P.k1.data(), P.k2.data(), P.size());
  P.permute(perm);
  *n_out = P.size();
  if (!host) return 0;
  if (P.size() > cap) return fail(PSPH_ECAPACITY, "read: %zu particles, buffer holds %zu", P.size(), cap);
  P.to_soa(host);
  return 0;
}

// ---------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------
extern "C" int psph_timestep(psph_ctx *c, double *dt_out) {
  if (!c || !dt_out) return fail(PSPH_EINVAL, "null argument");
  unsigned long long init = (unsigned long long)0x7FF0000000000000ull;   // +inf bits
  StageTimer tm(c, 0);
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64, init, 1);
  if (c->n) hipLaunchKernelGGL(k_dt_min, red_blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.bc, c->f.h, c->f.c, c->d_u64, c->d_err);
  tm.stop();
  unsigned long long bits;
  HIP_TRY(hipMemcpyAsync(&bits, c->d_u64, 8, hipMemcpyDeviceToHost, c->stream));
  PSPH_TRY(check_err(c, "timestep"));
  double mn;
  std::memcpy(&mn, &bits, 8);
  double dt = 1.0E+10;                    // timestep.cc:99
  if (mn < dt) dt = mn;
  *dt_out = c->cfg.cfl * dt;
  return 0;
}

extern "C" int psph_advance_time(psph_ctx *c, double *dt) {   // TimeProps::incrtime
  if (!c || !dt) return fail(PSPH_EINVAL, "null argument");
  if ((c->time + (*dt)) > c->next_output) (*dt) = c->next_output - c->time;
  if ((c->time + (*dt)) > c->cfg.t_end) (*dt) = c->cfg.t_end - c->time;
  c->time += (*dt);
  c->step++;
  return 0;
}

static int build_grid(psph_ctx *c) {
  const psph_config &k = c->cfg;
  // h_max over live particles
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 1, 0ull, 1);
  if (c->n) hipLaunchKernelGGL(k_h_max, red_blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.bc, c->f.h, c->d_u64 + 1);
  unsigned long long bits;
  HIP_TRY(hipMemcpyAsync(&bits, c->d_u64 + 1, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double hmax;
  std::memcpy(&hmax, &bits, 8);
  if (!(hmax > 0) || !std::isfinite(hmax)) hmax = k.h_air;
  c->h_max = hmax;
  c->reach_main = k.cutoff * hmax * (1 + 1e-9);
  GridDesc g{};
  if (k.list_all || k.dim == 1) {
    g.lo[0] = g.lo[1] = g.lo[2] = 0;
    g.inv_e = 0;
    g.nx = g.ny = g.nz = 1;
  } else {
    // cell edge: at least the reach of the RHS / density / velocity passes
    // (|dx_k| <= 3 (h_i + h_j) / 2 <= 3 h_max) and about 125 particles per
    // cell, so that a cell holds whole 64-particle chunks
    double vol = 1;
    for (int d = 0; d < 3; d++) vol *= k.ext_hi[d] - k.ext_lo[d];
    const double spacing = std::cbrt(vol / std::max<double>((double)c->n, 1.0));
    const double E = std::max(c->reach_main, 5.0 * spacing);
    int n3[3];
    for (int d = 0; d < 3; d++) {
      g.lo[d] = k.ext_lo[d] - E;
      n3[d] = (int)std::ceil((k.ext_hi[d] - k.ext_lo[d] + 2 * E) / E);
      if (n3[d] < 1) n3[d] = 1;
    }
    double ncell = (double)n3[0] * n3[1] * n3[2];
    if (ncell > 2.0e8) return fail(PSPH_ECAPACITY, "cell grid of %.3g cells (h_max %.4g) too large", ncell, hmax);
    g.inv_e = 1.0 / E;
    g.nx = n3[0]; g.ny = n3[1]; g.nz = n3[2];
  }
  g.ncells = (uint32_t)g.nx * g.ny * g.nz;
  int cbits = 0;
  while (cbits < 32 && ((uint64_t)1 << cbits) <= (uint64_t)g.ncells) cbits++;
  g.fbits = g.inv_e > 0 ? std::min(4, (32 - cbits) / 3) : 0;
  c->grid = g;
  size_t need = (size_t)g.ncells + 1;
  if (need > c->cs_cap) {
    uint32_t **arrs[] = {&c->cs, &c->tcnt, &c->toff, &c->cchunks, &c->coff};
    for (auto pp : arrs) { if (*pp) hipFree(*pp); *pp = nullptr; }
    for (auto pp : arrs) PSPH_TRY(dmalloc((void **)pp, need * 4));
    c->cs_cap = need;
  }
  const size_t ub = c->n / kWave + (size_t)g.ncells + 1;
  const size_t cub_ = c->n / kChunk + (size_t)g.ncells + 1;
  if (ub > c->tile_cap) {
    if (c->tbox) hipFree(c->tbox);
    c->tbox = nullptr;
    PSPH_TRY(dmalloc((void **)&c->tbox, ub * 2 * sizeof(float4)));
    c->tile_cap = ub;
  }
  if (cub_ > c->chunk_cap) {
    if (c->cfirst) hipFree(c->cfirst);
    if (c->ccount) hipFree(c->ccount);
    c->cfirst = c->ccount = nullptr;
    PSPH_TRY(dmalloc((void **)&c->cfirst, cub_ * 4));
    PSPH_TRY(dmalloc((void **)&c->ccount, cub_ * 4));
    c->chunk_cap = cub_;
  }
  return 0;
}

// exclusive scan of n + 1 values (the last input is 0) with hipcub
static int excl_scan(psph_ctx *c, const uint32_t *in, uint32_t *out, size_t n1) {
  size_t bytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n1, c->stream));
  PSPH_TRY(ensure_cub(c, bytes));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->cub, bytes, in, out, (int)n1, c->stream));
  return 0;
}

extern "C" int psph_neighbors(psph_ctx *c) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  PSPH_TRY(build_grid(c));
  StageTimer tm(c, 1);
  const GridDesc &g = c->grid;
  uint32_t n = (uint32_t)c->n;
  hipStream_t s = c->stream;
  Tape &tp = c->tape;
  const bool frozen = tp.mode == 2 && tp.frozen;
  if (frozen) {
    if (tp.k >= (int64_t)tp.xrec.size()) return fail(PSPH_EINVAL, "replay: no recorded positions for step %lld", (long long)tp.k);
    if (!c->px) {
      PSPH_TRY(dmalloc((void **)&c->px, c->cap * 8));
      PSPH_TRY(dmalloc((void **)&c->py, c->cap * 8));
      PSPH_TRY(dmalloc((void **)&c->pz, c->cap * 8));
    }
    if (n) hipLaunchKernelGGL(k_pos_from_rec, blocks(n), 256, 0, s, n, c->pid, c->f.bc, tp.xrec[tp.k], tp.xrec_n[tp.k],
                              c->f.x, c->f.y, c->f.z, c->px, c->py, c->pz, c->d_err);
  }
  if (n) {
    hipLaunchKernelGGL(k_cellkey, blocks(n), 256, 0, s, n, frozen ? c->px : c->f.x, frozen ? c->py : c->f.y,
                       frozen ? c->pz : c->f.z, c->f.bc, g, c->key_a, c->idx_a);
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= ((uint64_t)g.ncells << (3 * g.fbits))) end_bit++;
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, c->key_a, c->key_b, c->idx_a, c->idx_b, (int)n, 0, end_bit, s));
    PSPH_TRY(ensure_cub(c, bytes));
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(c->cub, bytes, c->key_a, c->key_b, c->idx_a, c->idx_b, (int)n, 0, end_bit, s));
    // gather every field into cell order
#define G_(a) { hipLaunchKernelGGL(k_gather<double>, blocks(n), 256, 0, s, n, c->idx_b, c->f.a, c->dsp[0]); std::swap(c->f.a, c->dsp[0]); }
    PSPH_DFIELDS(G_)
#undef G_
#define G_(a) { hipLaunchKernelGGL(k_gather<uint32_t>, blocks(n), 256, 0, s, n, c->idx_b, c->f.a, c->usp); std::swap(c->f.a, c->usp); }
    PSPH_UFIELDS(G_)
#undef G_
#define G_(a) { hipLaunchKernelGGL(k_gather<int32_t>, blocks(n), 256, 0, s, n, c->idx_b, c->f.a, c->isp); std::swap(c->f.a, c->isp); }
    PSPH_IFIELDS(G_)
#undef G_
    hipLaunchKernelGGL(k_gather<uint32_t>, blocks(n), 256, 0, s, n, c->idx_b, c->pid, c->pid_sp);
    std::swap(c->pid, c->pid_sp);
    if (frozen) {
      double **pp[3] = {&c->px, &c->py, &c->pz};
      for (auto q : pp) {
        hipLaunchKernelGGL(k_gather<double>, blocks(n), 256, 0, s, n, c->idx_b, *q, c->dsp[0]);
        std::swap(*q, c->dsp[0]);
      }
    }
    std::swap(c->key_a, c->key_b);   // key_a: sorted keys
  }
  // drop removed particles (their key == ncells sorts them to the tail)
  unsigned long long dead = 0;
  HIP_TRY(hipMemcpyAsync(&dead, c->d_u64 + 3, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (dead) {
    if (dead > n) return fail(PSPH_EINVARIANT, "removed-particle count %llu exceeds %u", dead, n);
    n -= (uint32_t)dead;
    c->n = n;
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, s, c->d_u64 + 3, 0ull, 1);
  }
  hipLaunchKernelGGL(k_cell_start, blocks(n + 1), 256, 0, s, n, c->key_a, 3 * g.fbits, g.ncells, c->cs);
  // tiles of each cell and their boxes
  hipLaunchKernelGGL(k_tile_counts, blocks(g.ncells + 1), 256, 0, s, g.ncells, c->cs, c->tcnt);
  PSPH_TRY(excl_scan(c, c->tcnt, c->toff, (size_t)g.ncells + 1));
  // the positions every predicate of this step reads
  const double *qx = frozen ? c->px : c->f.x, *qy = frozen ? c->py : c->f.y, *qz = frozen ? c->pz : c->f.z;
  if (tp.mode == 1 && (int64_t)tp.xrec.size() == tp.k) {
    double *rec = nullptr;
    const uint32_t nrec = std::max<uint32_t>(c->next_pid, 1);
    PSPH_TRY(dmalloc((void **)&rec, (size_t)nrec * 3 * 8));
    tp.xrec.push_back(rec);
    tp.xrec_n.push_back(nrec);
    if (n) hipLaunchKernelGGL(k_pos_to_rec, blocks(n), 256, 0, s, n, c->pid, c->f.x, c->f.y, c->f.z, rec, nrec);
  }
  hipLaunchKernelGGL(k_tile_boxes, (unsigned)(n / kWave + g.ncells + 1), kWave, 0, s, g.ncells, c->cs, c->toff,
                     qx, qy, qz, c->tbox);
  // reference buckets (list membership) and h at list time
  int nb[3] = {1, 1, 1};
  double del = c->cfg.bucket_edge > 0 ? c->cfg.bucket_edge : 1.0;
  if (c->cfg.dim == 3)
    for (int d = 0; d < 3; d++) nb[d] = (int)std::llround((c->cfg.ext_hi[d] - c->cfg.ext_lo[d]) / del);
  for (int d = 0; d < 3; d++)
    if (nb[d] > kRefbMaxPerAxis)
      return fail(PSPH_ECAPACITY, "%d reference buckets on axis %d (at most %d)", nb[d], d, kRefbMaxPerAxis);
  if (n) hipLaunchKernelGGL(k_refb, blocks(n), 256, 0, s, n, qx, qy, qz, c->f.h,
                            c->cfg.ext_lo[0], c->cfg.ext_lo[1], c->cfg.ext_lo[2], del, nb[0], nb[1], nb[2],
                            c->refb, c->h_list);
  // targets and their chunks (<= kChunk consecutive targets of one cell)
  hipLaunchKernelGGL(k_target_flags, blocks(n + 1), 256, 0, s, n, c->f.bc, c->flags);
  PSPH_TRY(excl_scan(c, c->flags, c->tpos, (size_t)n + 1));
  if (n) hipLaunchKernelGGL(k_targets, blocks(n), 256, 0, s, n, c->flags, c->tpos, c->targets);
  hipLaunchKernelGGL(k_cell_chunks, blocks(g.ncells + 1), 256, 0, s, g.ncells, c->cs, c->tpos, (uint32_t)kChunk,
                     c->cchunks);
  PSPH_TRY(excl_scan(c, c->cchunks, c->coff, (size_t)g.ncells + 1));
  hipLaunchKernelGGL(k_chunks, blocks(g.ncells + 1), 256, 0, s, g.ncells, c->cs, c->tpos, c->coff,
                     (uint32_t)kChunk, c->cfirst, c->ccount, c->nchunks);
  HIP_TRY(hipGetLastError());
  tm.stop();
  c->grid_valid = true;
  c->vis_valid = false;
  return 0;
}

// half an ulp of the cut-off; the test above assumes its mantissa is even
static CutTest cut_test(double cut) {
  int e;
  std::frexp(cut, &e);                       // cut = f 2^e, f in [0.5, 1)
  return CutTest{cut, std::ldexp(1.0, e - 54)};
}

static Traversal traversal(psph_ctx *c, double reach) {
  return Traversal{c->cs, c->toff, c->tbox, c->grid, reach};
}

// a list box 4.25 h can reach past the adjacent buckets: the list test needs
// the bucket adjacency as well
static int list_adj(const psph_config &k, double hmax) {
  return (!k.list_all && k.dim == 3 && k.list_factor * hmax >= k.bucket_edge * (1 - 1e-12)) ? 1 : 0;
}

static PairCommon make_common(psph_ctx *c, double reach) {
  PairCommon pc;
  const bool frozen = c->tape.mode == 2 && c->tape.frozen;
  pc.x = frozen ? c->px : c->f.x; pc.y = frozen ? c->py : c->f.y; pc.z = frozen ? c->pz : c->f.z;
  pc.ax = c->f.x; pc.ay = c->f.y; pc.az = c->f.z;
  pc.refb = c->refb; pc.h_list = c->h_list;
  pc.bc = c->f.bc;
  pc.gflag = c->rec ? c->rec->gflag : nullptr;
  pc.tr = traversal(c, reach);
  pc.targets = c->targets;
  pc.cfirst = c->cfirst; pc.ccount = c->ccount; pc.nchunks = c->nchunks;
  pc.list_all = (c->cfg.list_all || c->cfg.dim == 1) ? 1 : 0;
  pc.list_factor = c->cfg.list_factor;
  pc.del = c->cfg.bucket_edge;
  for (int d = 0; d < 3; d++) pc.ext_lo[d] = c->cfg.ext_lo[d];
  // h_max bounds h at list time (it only grows with an adaptation)
  pc.list_adj = list_adj(c->cfg, c->h_max);
  pc.cut = cut_test(c->cfg.cutoff);
  pc.err = c->d_err;
  return pc;
}

// workgroups for a pair pass: an upper bound on the chunks of csz targets
// (each cell may add a partial chunk), padded so that kXcds * ceil(nchunks /
// kXcds) exist
static unsigned chunk_grid(size_t n, size_t ncells, size_t csz) {
  size_t ub = n / csz + ncells + 1 + kXcds;
  return (unsigned)((ub + kXcds - 1) / kXcds * kXcds);
}

template <class P>
static void launch_pairs(psph_ctx *c, const PairCommon &pc, const typename P::Args &a) {
  unsigned nb = chunk_grid(c->n, c->grid.ncells, kChunk);
  if (c->cfg.dim == 3)
    hipLaunchKernelGGL((k_pairs<3, Bound<3, P>>), nb, kWave, 0, c->stream, pc, a);
  else
    hipLaunchKernelGGL((k_pairs<1, Bound<1, P>>), nb, kWave, 0, c->stream, pc, a);
}

// Visit lists of the current step (positions, h and targets fixed since the
// last psph_neighbors / adaptation): count the tiles each chunk meets, scan,
// then write the tiles with hits and their lane masks.
static int ensure_visits(psph_ctx *c) {
  if (c->vis_valid) return 0;
  StageTimer tm(c, 1);
  const PairCommon pc = make_common(c, c->reach_main);
  const unsigned G = chunk_grid(c->n, c->grid.ncells, kChunk);
  if ((size_t)G + 1 > c->vchunk_cap) {
    uint32_t **arrs[] = {&c->vcount, &c->voff, &c->vnum};
    for (auto pp : arrs) { if (*pp) hipFree(*pp); *pp = nullptr; }
    for (auto pp : arrs) PSPH_TRY(dmalloc((void **)pp, ((size_t)G + 1) * 4));
    c->vchunk_cap = (size_t)G + 1;
  }
  // a superset of the three passes' kernel boxes with a relative margin
  const double cutx = c->cfg.cutoff * (1.0 + 1e-12);
  hipLaunchKernelGGL(k_visit_count, G, kWave, 0, c->stream, pc, c->f.h, cutx, c->h_max, c->vcount);
  hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, c->stream, (int *)c->vcount + G, 0, 1);
  HIP_TRY(hipGetLastError());
  PSPH_TRY(excl_scan(c, c->vcount, c->voff, (size_t)G + 1));
  uint32_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, c->voff + G, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if ((size_t)total > c->vis_cap) {
    if (c->vmask) hipFree(c->vmask);
    if (c->vent) hipFree(c->vent);
    c->vmask = nullptr; c->vent = nullptr;
    const size_t cap = (size_t)total + total / 8 + 1024;
    PSPH_TRY(dmalloc((void **)&c->vmask, cap * kWave * 4));
    PSPH_TRY(dmalloc((void **)&c->vent, cap * sizeof(uint2)));
    c->vis_cap = cap;
  }
  if (c->cfg.dim == 3)
    hipLaunchKernelGGL(k_visit_masks<3>, G, kWave, 0, c->stream, pc, c->f.h, c->cfg.cutoff2, cutx, c->h_max,
                       c->voff, c->vnum, c->vent, c->vmask);
  else
    hipLaunchKernelGGL(k_visit_masks<1>, G, kWave, 0, c->stream, pc, c->f.h, c->cfg.cutoff2, cutx, c->h_max,
                       c->voff, c->vnum, c->vent, c->vmask);
  HIP_TRY(hipGetLastError());
  c->vis_valid = true;
  return 0;
}

// the RHS, density and filter passes run over the visit lists; PM, PP, PD:
// the pass in the FP32-pair, FP32-pure and FP64 forms
template <int D, class P>
static void launch_visit_pass(psph_ctx *c, const PairCommon &pc, const VisitList &vl, const typename P::Args &a) {
  const unsigned nb = chunk_grid(c->n, c->grid.ncells, kChunk);
  if (c->tape.mode == 2 && c->tape.frozen)
    hipLaunchKernelGGL((k_pairs_v<D, Bound<D, P>, true>), nb, kWave, 0, c->stream, pc, vl, a);
  else
    hipLaunchKernelGGL((k_pairs_v<D, Bound<D, P>>), nb, kWave, 0, c->stream, pc, vl, a);
}
template <class PM, class PX, class PP, class PD>
static int launch_visits(psph_ctx *c, const typename PD::Args &a) {
  PSPH_TRY(ensure_visits(c));
  const PairCommon pc = make_common(c, c->reach_main);
  const VisitList vl{c->voff, c->vnum, c->vent, c->vmask};
  const int form = c->cfg.precision;
  if (c->cfg.dim == 3) {
    if (form == PSPH_FP32_PAIR) launch_visit_pass<3, PM>(c, pc, vl, a);
    else if (form == PSPH_FP32_EXP) launch_visit_pass<3, PX>(c, pc, vl, a);
    else if (form == PSPH_FP32_PURE) launch_visit_pass<3, PP>(c, pc, vl, a);
    else launch_visit_pass<3, PD>(c, pc, vl, a);
  } else {
    if (form == PSPH_FP32_PAIR) launch_visit_pass<1, PM>(c, pc, vl, a);
    else if (form == PSPH_FP32_EXP) launch_visit_pass<1, PX>(c, pc, vl, a);
    else if (form == PSPH_FP32_PURE) launch_visit_pass<1, PP>(c, pc, vl, a);
    else launch_visit_pass<1, PD>(c, pc, vl, a);
  }
  return 0;
}

static EosConst eos_const(const psph_config &k) {
  return EosConst{k.ng0, k.cvs, k.cvg, k.cva, k.rg, k.ra, k.lans};
}

static int reduce_hmax(psph_ctx *c, double *out) {
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 1, 0ull, 1);
  if (c->n) hipLaunchKernelGGL(k_h_max, red_blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.bc, c->f.h, c->d_u64 + 1);
  unsigned long long bits;
  HIP_TRY(hipMemcpyAsync(&bits, c->d_u64 + 1, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::memcpy(out, &bits, 8);
  return 0;
}

// the later passes of a step use the new h: widen the reach
static void set_reach(psph_ctx *c, double hmax) {
  c->vis_valid = false;                // h changed
  c->h_max = hmax;
  c->reach_main = c->cfg.cutoff * hmax * (1 + 1e-9);
}

extern "C" int psph_adapt_h(psph_ctx *c) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (!c->grid_valid) return fail(PSPH_EINVAL, "adapt_h: call psph_neighbors first");
  const psph_config &k = c->cfg;
  const bool all = k.list_all || k.dim == 1;
  // the list never reaches past the own bucket or the 4.25 h_i box of list time
  const double list_reach = all ? HUGE_VAL : std::max(k.list_factor * c->h_max, k.bucket_edge);
  AdaptArgs a{};
  a.m = c->f.m; a.bc = c->f.bc; a.h = c->f.h; a.done = c->usp;
  a.cutoff = k.cutoff; a.cutoff2 = k.cutoff2; a.c_smooth = k.c_smooth; a.eta = k.eta;
  a.ratio = k.ratio_sml_dx; a.thresh = k.thresh; a.pghost = k.pghost_contrib; a.dim = k.dim;
  if (c->n) HIP_TRY(hipMemsetAsync(c->usp, 0, c->n * 4, c->stream));
  double hmax = c->h_max;
  for (int it = 0; it < k.num_loop; it++) {
    const double reach = std::min(k.cutoff * k.c_smooth * hmax, list_reach) * (1 + 1e-9);
    {
      StageTimer tm(c, 2);
      // FP64 arithmetic in every form (h enters every later pass, and a
      // relative error of h moves every kernel value by ~|s|^2 times as
      // much); the FX form takes its exponentials from the FP32 instruction
      if (c->n) {
        if (k.precision == PSPH_FP32_EXP) launch_pairs<AdaptPass<double, true>>(c, make_common(c, reach), a);
        else launch_pairs<AdaptPass<double>>(c, make_common(c, reach), a);
      }
    }
    HIP_TRY(hipGetLastError());
    PSPH_TRY(reduce_hmax(c, &hmax));
  }
  PSPH_TRY(check_err(c, "adapt_h"));
  set_reach(c, hmax);
  return 0;
}

// replays (or records) the smoothing lengths of an adaptation step by id
static int h_by_id(psph_ctx *c, double *tab, int replay) {
  if (c->n) hipLaunchKernelGGL(k_h_by_id, blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->pid, c->f.h, tab, replay);
  HIP_TRY(hipGetLastError());
  if (!replay) return 0;
  double hmax;
  PSPH_TRY(reduce_hmax(c, &hmax));
  set_reach(c, hmax);
  return 0;
}

// A pass writes its outputs for its targets (need_neigh) into spare arrays;
// the other particles keep their values: one launch copies those.
struct FieldCopy {
  const double *src[9];
  double *dst[9];
  int nf;
};
__global__ void k_copy_nontargets(uint32_t n, const int32_t *bc, FieldCopy fc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || need_neigh(bc[i])) return;
  for (int q = 0; q < fc.nf; q++) fc.dst[q][i] = fc.src[q][i];
}

extern "C" int psph_rhs(psph_ctx *c, double dt) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (!c->grid_valid) return fail(PSPH_EINVAL, "rhs: call psph_neighbors first");
  const psph_config &k = c->cfg;
  uint32_t n = (uint32_t)c->n;
  // outputs go to spare arrays; non-targets keep their values
  double *ov[9] = {c->dsp[1], c->dsp[2], c->dsp[3], c->dsp[4], c->dsp[5], c->dsp[6], c->dsp[7], c->dsp[8], c->dsp[9]};
  double *src[9] = {c->f.vx, c->f.vy, c->f.vz, c->f.e, c->f.p, c->f.c, c->f.T, c->f.gam, c->f.cp};
  if (n) {
    FieldCopy fc{};
    for (int q = 0; q < 9; q++) { fc.src[q] = src[q]; fc.dst[q] = ov[q]; }
    fc.nf = 9;
    hipLaunchKernelGGL(k_copy_nontargets, blocks(n), 256, 0, c->stream, n, c->f.bc, fc);
  }
  RhsArgs a{};
  a.h = c->f.h; a.m = c->f.m; a.rho = c->f.rho; a.p = c->f.p; a.c = c->f.c; a.T = c->f.T; a.cp = c->f.cp;
  a.vx = c->f.vx; a.vy = c->f.vy; a.vz = c->f.vz; a.svx = c->f.svx; a.svy = c->f.svy; a.svz = c->f.svz;
  a.e = c->f.e; a.xi = c->f.xi; a.bc = c->f.bc;
  a.o_vx = ov[0]; a.o_vy = ov[1]; a.o_vz = ov[2];
  a.o_e = ov[3]; a.o_p = ov[4]; a.o_c = ov[5]; a.o_T = ov[6]; a.o_gam = ov[7]; a.o_cp = ov[8];
  a.dt = dt; a.g = k.g; a.cutoff = k.cutoff; a.cutoff2 = k.cutoff2; a.alpha = k.alpha; a.beta = k.beta;
  a.ata = k.ata; a.eps_half = k.eps_half; a.prandtl = k.prandtl; a.heat_scale = k.heat_scale;
  a.beta2 = 2.0 * k.beta; a.kheat = 20.0 * k.eps_half / k.prandtl;
  a.lans = k.lans; a.dim = k.dim; a.eos = eos_const(k); a.pairs = c->d_u64 + 2;
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 2, 0ull, 1);
  if (n) PSPH_TRY(ensure_visits(c));
  {
    StageTimer tm(c, 3);
    if (n) PSPH_TRY((launch_visits<RhsPass<float>, RhsPass<float, true, true>, RhsPass<float, false>, RhsPass<double>>(c, a)));
  }
  HIP_TRY(hipGetLastError());
  std::swap(c->f.vx, c->dsp[1]); std::swap(c->f.vy, c->dsp[2]); std::swap(c->f.vz, c->dsp[3]);
  std::swap(c->f.e, c->dsp[4]); std::swap(c->f.p, c->dsp[5]); std::swap(c->f.c, c->dsp[6]);
  std::swap(c->f.T, c->dsp[7]); std::swap(c->f.gam, c->dsp[8]); std::swap(c->f.cp, c->dsp[9]);
  HIP_TRY(hipMemcpyAsync(&c->last_pairs, c->d_u64 + 2, 8, hipMemcpyDeviceToHost, c->stream));
  return check_err(c, "rhs");
}

extern "C" int psph_density(psph_ctx *c) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (!c->grid_valid) return fail(PSPH_EINVAL, "density: call psph_neighbors first");
  const psph_config &k = c->cfg;
  uint32_t n = (uint32_t)c->n;
  if (n) {
    FieldCopy fc{};
    fc.src[0] = c->f.rho; fc.dst[0] = c->dsp[1];
    fc.src[1] = c->f.e; fc.dst[1] = c->dsp[2];
    fc.nf = 2;
    hipLaunchKernelGGL(k_copy_nontargets, blocks(n), 256, 0, c->stream, n, c->f.bc, fc);
  }
  DensArgs a{};
  a.h = c->f.h; a.m = c->f.m; a.rho = c->f.rho; a.e = c->f.e; a.bc = c->f.bc; a.phase = c->f.phase;
  a.o_rho = c->dsp[1]; a.o_e = c->dsp[2]; a.xi = c->f.xi; a.p = c->f.p; a.c = c->f.c; a.T = c->f.T;
  a.gam = c->f.gam; a.cp = c->f.cp; a.inv = c->f.inv;
  a.cutoff = k.cutoff; a.cutoff2 = k.cutoff2; a.e_ratio = k.e_smooth_ratio; a.msfrc_thresh = k.msfrc_thresh;
  a.pghost = k.pghost_contrib; a.energy_smooth = k.energy_smooth; a.eos = eos_const(k);
  a.cut = cut_test(k.cutoff);
  a.o_wn = c->rec ? c->rec->d[RF_WNE] : nullptr;
  a.e_sq16 = k.e_smooth_ratio * k.e_smooth_ratio == 16.0;
  {
    const double en = kInvSqrtPi * k.e_smooth_ratio;
    a.e_norm = k.dim == 3 ? en * en * en : en;
  }
  if (n) PSPH_TRY(ensure_visits(c));
  {
    StageTimer tm(c, 4);
    if (n) PSPH_TRY((launch_visits<DensPass<float>, DensPass<float, true, true>, DensPass<float, false>, DensPass<double>>(c, a)));
  }
  HIP_TRY(hipGetLastError());
  std::swap(c->f.rho, c->dsp[1]);
  std::swap(c->f.e, c->dsp[2]);
  return check_err(c, "density");
}

extern "C" int psph_vsmooth(psph_ctx *c) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (!c->grid_valid) return fail(PSPH_EINVAL, "vsmooth: call psph_neighbors first");
  const psph_config &k = c->cfg;
  VsmArgs a{};
  a.h = c->f.h; a.m = c->f.m; a.rho = c->f.rho; a.vx = c->f.vx;
  a.vy = k.dim == 3 ? c->f.vy : nullptr; a.vz = k.dim == 3 ? c->f.vz : nullptr;
  a.bc = c->f.bc; a.svx = c->f.svx; a.svy = k.dim == 3 ? c->f.svy : nullptr; a.svz = k.dim == 3 ? c->f.svz : nullptr;
  a.cutoff = k.cutoff; a.cutoff2 = k.cutoff2;
  a.o_wn = c->rec ? c->rec->d[RF_WNV] : nullptr;
  a.ok_test = c->rec ? 1 : 0;
  if (c->n) PSPH_TRY(ensure_visits(c));
  {
    StageTimer tm(c, 5);
    if (c->n) PSPH_TRY((launch_visits<VsmPass<float>, VsmPass<float, true>, VsmPass<float>, VsmPass<double>>(c, a)));
  }
  HIP_TRY(hipGetLastError());
  return check_err(c, "vsmooth");
}

static int iferupt(const psph_ctx *c) {     // TimeProps::iferupt (properties.h:126-135)
  if (c->time < c->cfg.erupt_start) return 0;
  if (c->time > c->cfg.erupt_end) return 0;
  return 1;
}

extern "C" int psph_update_pos(psph_ctx *c, double dt) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  const psph_config &k = c->cfg;
  uint32_t n = (uint32_t)c->n;
  PosTape tp{c->tape.mode, (int32_t)c->tape.k, c->pid, c->tape.tr_step, c->tape.esc_step};
  if (tp.mode && (!tp.tr_step || c->next_pid > c->tape.id_cap)) return fail(PSPH_EINVAL, "update_pos: tape has no id table");
  StageTimer tm(c, 6);
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 4, 0ull, 1);
  if (n)
    hipLaunchKernelGGL(k_update_pos, blocks(n), 256, 0, c->stream, n, k.dim, k.lans, iferupt(c), dt,
                       c->f.x, c->f.y, c->f.z, c->f.vx, c->f.vy, c->f.vz, c->f.svx, c->f.svy, c->f.svz,
                       c->f.bc, c->f.h, c->f.inv, k.dom_lo[2], k.ratio_sml_dx * k.h_air,
                       k.dom_lo[0], k.dom_lo[1], k.dom_lo[2], k.dom_hi[0], k.dom_hi[1], k.dom_hi[2],
                       c->d_u64 + 3, tp);
  if (n) hipLaunchKernelGGL(k_count_real, red_blocks(n), 256, 0, c->stream, n, c->f.bc, c->d_u64 + 4);
  tm.stop();
  HIP_TRY(hipGetLastError());
  unsigned long long cnt[2] = {0, 0};      // escaped, real after the update
  HIP_TRY(hipMemcpyAsync(cnt, c->d_u64 + 3, 16, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->last_removed = (int64_t)cnt[0];
  c->last_real = (int64_t)cnt[1];
  c->grid_valid = false;
  return 0;
}

// id-indexed event tables of the tape, grown to hold `need` ids (new: -1)
static int tape_ids(psph_ctx *c, size_t need) {
  Tape &tp = c->tape;
  if (need <= tp.id_cap && tp.tr_step) return 0;
  size_t cap = std::max(need, std::max(2 * tp.id_cap, c->cap));
  int32_t *a = nullptr, *b = nullptr;
  PSPH_TRY(dmalloc((void **)&a, cap * 4));
  if (dmalloc((void **)&b, cap * 4)) { hipFree(a); return PSPH_ENOMEM; }
  hipLaunchKernelGGL(k_fill_i32, blocks(cap), 256, 0, c->stream, a, -1, (int)cap);
  hipLaunchKernelGGL(k_fill_i32, blocks(cap), 256, 0, c->stream, b, -1, (int)cap);
  if (tp.tr_step) {
    HIP_TRY(hipMemcpyAsync(a, tp.tr_step, tp.id_cap * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(b, tp.esc_step, tp.id_cap * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    hipFree(tp.tr_step);
    hipFree(tp.esc_step);
  }
  tp.tr_step = a;
  tp.esc_step = b;
  tp.id_cap = cap;
  return 0;
}

static void tape_clear(Tape &tp) {
  for (double *p : tp.htab) if (p) hipFree(p);
  tp.htab.clear();
  tp.dt.clear();
  tp.hslot.clear();
  tp.inlet.clear();
  for (double *p : tp.xrec) if (p) hipFree(p);
  tp.xrec.clear();
  tp.xrec_n.clear();
  for (double *p : tp.xrecb) if (p) hipFree(p);
  tp.xrecb.clear();
  tp.xrecb_n.clear();
  tp.frozen = false;
  tp.k = 0;
  tp.mode = 0;
}

// captures the step's intermediates into c->rec (defined with the adjoint)
static int rec_capture(psph_ctx *c, int stage, double dt);

extern "C" int psph_inject(psph_ctx *c, double dt, int64_t *added) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (added) *added = 0;
  c->last_injected = 0;
  Tape &tp = c->tape;
  std::vector<int32_t> *rec = nullptr;
  const std::vector<int32_t> *rep = nullptr;
  if (tp.mode == 1) { tp.inlet.emplace_back(); rec = &tp.inlet.back(); }
  if (tp.mode == 2) {
    if (tp.k >= (int64_t)tp.inlet.size()) return fail(PSPH_EINVAL, "inject: replay past the recorded run");
    rep = &tp.inlet[tp.k];
  }
  c->inj.resize(0);
  if (c->cfg.dim != 3 || !iferupt(c)) {
    if (rep && !rep->empty()) return fail(PSPH_EINVAL, "inject: recorded run erupted at this step");
    return 0;
  }
  PSPH_TRY(inject_host(c->cfg, c->time, dt, c->step, c->inj, rec, rep ? rep->data() : nullptr,
                       rep ? rep->size() : 0));
  size_t m = c->inj.size();
  if (c->n + m > c->cap)
    return fail(PSPH_ECAPACITY, "inject: %zu + %zu particles exceed capacity %zu", c->n, m, c->cap);
  if (tp.mode && c->next_pid + m > tp.id_cap) PSPH_TRY(tape_ids(c, c->next_pid + m));
  {
    StageTimer tm(c, 7);
    PSPH_TRY(upload(c, c->inj, c->n));
  }
  if (m) hipLaunchKernelGGL(k_iota_u32, blocks(m), 256, 0, c->stream, c->pid + c->n, c->next_pid, (uint32_t)m);
  c->next_pid += (uint32_t)m;
  c->n += m;
  c->last_injected = (int64_t)m;
  if (added) *added = (int64_t)m;
  c->grid_valid = false;
  return 0;
}

// wall-ghost refresh over the cell grid of the last neighbour search; disp:
// the largest per-axis displacement of any particle since that search
// ground on a bucket face: no MIXED bucket, nothing to refresh
static bool bcond_trivial(const psph_config &k) {
  const double fz = (k.dom_lo[2] - k.ext_lo[2]) / k.bucket_edge;
  return fz == std::floor(fz);
}

// wrec (optional, wrec_n entries): receives each refreshed ghost's Shepard
// normaliser, for the adjoint
// the bucket geometry, ground, support tests and EOS of the refresh
static BcondArgs bcond_geom(const psph_config &k) {
  BcondArgs a{};
  a.zg = k.dom_lo[2]; a.del = k.bucket_edge; a.cutoff2 = k.cutoff2;
  for (int d = 0; d < 3; d++) {
    a.ext_lo[d] = k.ext_lo[d];
    a.nb[d] = (int)std::llround((k.ext_hi[d] - k.ext_lo[d]) / k.bucket_edge);
  }
  a.bzg = (int)std::floor((k.dom_lo[2] - k.ext_lo[2]) / k.bucket_edge);
  a.cut = cut_test(k.cutoff);
  a.eos = eos_const(k);
  return a;
}

// qx, qy, qz (optional): the positions the support tests read
static int apply_bcond_grid(psph_ctx *c, double disp, double *wrec = nullptr, size_t wrec_n = 0,
                            const double *qx = nullptr, const double *qy = nullptr, const double *qz = nullptr) {
  const psph_config &k = c->cfg;
  if (k.dim != 3 || !c->n || bcond_trivial(k)) return 0;
  BcondArgs a = bcond_geom(k);
  a.wrec = wrec;
  a.wrec_n = (uint32_t)wrec_n;
  a.x = c->f.x; a.y = c->f.y; a.z = c->f.z; a.h = c->f.h; a.m = c->f.m;
  a.rho = c->f.rho; a.vx = c->f.vx; a.vy = c->f.vy; a.vz = c->f.vz; a.e = c->f.e; a.xi = c->f.xi;
  a.p = c->f.p; a.c = c->f.c; a.T = c->f.T; a.gam = c->f.gam; a.cp = c->f.cp;
  a.bc = c->f.bc; a.cs = c->cs; a.toff = c->toff; a.tbox = c->tbox; a.g = c->grid;
  a.px = qx ? qx : c->f.x; a.py = qx ? qy : c->f.y; a.pz = qx ? qz : c->f.z;
  const double hm = std::max(c->h_max, k.ratio_sml_dx * k.h_air);
  a.reach = k.cutoff * hm * (1 + 1e-9) + disp * (1 + 1e-9);
  hipLaunchKernelGGL(k_bcond, blocks(c->n, kBcondSpan), kWave, 0, c->stream, (uint32_t)c->n, a);
  HIP_TRY(hipGetLastError());
  return 0;
}

extern "C" int psph_apply_bcond(psph_ctx *c) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (!c->grid_valid) PSPH_TRY(psph_neighbors(c));
  return apply_bcond_grid(c, 0.0);
}

// domain adaptation (psph_domain.inl)
namespace {
int domain_adapt(psph_ctx *c);
int domain_scan_step(psph_ctx *c);
}  // namespace

extern "C" int psph_step(psph_ctx *c, psph_step_out *out) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  const psph_config &k = c->cfg;
  Tape &tp = c->tape;
  double dt;
  if (k.adapt_domain && k.dim == 3 && !c->dom_ready)
    return fail(PSPH_EINVAL, "adapt_domain: call psph_domain_init after psph_load");
  if (tp.mode == 2) {
    if (tp.k >= (int64_t)tp.dt.size()) return fail(PSPH_EINVAL, "replay: past the end of the recorded run");
    dt = tp.dt[tp.k];
    c->time += dt;
    c->step++;
  } else {
    PSPH_TRY(psph_timestep(c, &dt));
    PSPH_TRY(psph_advance_time(c, &dt));
    if (tp.mode == 1) tp.dt.push_back(dt);
  }
  c->last_domain = 0;
  if (k.adapt_domain && k.dim == 3 && c->dom_pending) {
    // the region of real particles grows before this step's search
    // (main.cc:256-304); the wall ghosts are refreshed after it
    c->dom_pending = 0;
    PSPH_TRY(domain_adapt(c));
    if (k.wall_bc && c->n) {
      PSPH_TRY(psph_neighbors(c));
      PSPH_TRY(apply_bcond_grid(c, 0.0));
    }
  }
  PSPH_TRY(psph_neighbors(c));
  if (tp.mode == 2) {
    const int slot = tp.hslot[tp.k];
    if (slot >= 0) PSPH_TRY(h_by_id(c, tp.htab[slot], 1));
  } else if (k.sml_update_int > 0 && c->step % k.sml_update_int == 0) {
    PSPH_TRY(psph_adapt_h(c));
    if (tp.mode == 1) {
      double *tab = nullptr;
      PSPH_TRY(dmalloc((void **)&tab, (size_t)c->next_pid * 8));
      tp.htab.push_back(tab);
      tp.hslot.push_back((int)tp.htab.size() - 1);
      PSPH_TRY(h_by_id(c, tab, 0));
    }
  } else if (tp.mode == 1) {
    tp.hslot.push_back(-1);
  }
  if (c->rec) PSPH_TRY(rec_capture(c, 0, dt));
  PSPH_TRY(psph_rhs(c, dt));
  if (c->rec) PSPH_TRY(rec_capture(c, 1, dt));
  PSPH_TRY(psph_density(c));
  if (c->rec) PSPH_TRY(rec_capture(c, 2, dt));
  if (k.lans) PSPH_TRY(psph_vsmooth(c));
  if (c->rec) PSPH_TRY(rec_capture(c, 3, dt));
  PSPH_TRY(psph_update_pos(c, dt));
  int64_t added = 0;
  const uint32_t id0 = c->next_pid;
  PSPH_TRY(psph_inject(c, dt, &added));
  if (c->rec) {
    c->rec->inj_lo = id0;
    c->rec->inj_hi = c->next_pid;
    c->rec->inj_vel.assign(c->inj.vz.begin(), c->inj.vz.end());
  }
  if (k.adapt_domain && k.dim == 3) PSPH_TRY(domain_scan_step(c));   // main.cc:433-436
  if (k.wall_bc && k.dim == 3 && c->n) {
    // apply_bcond at the end of the step (main.cc:468): the grid is the one of
    // this step's neighbour search; particles have moved by at most dt vmax
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, c->stream, c->d_u64 + 1, 0ull, 1);
    hipLaunchKernelGGL(k_speed_max, red_blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.bc, k.lans, c->f.vx,
                       c->f.vy, c->f.vz, c->f.svx, c->f.svy, c->f.svz, c->d_u64 + 1);
    unsigned long long bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, c->d_u64 + 1, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double vmax;
    std::memcpy(&vmax, &bits, 8);
    double disp = std::fabs(dt) * vmax;
    const uint32_t nb = (uint32_t)c->n;
    if (tp.mode == 1 && (int64_t)tp.xrecb.size() == tp.k) {
      // positions at the refresh, by id, for a frozen replay
      double *rec = nullptr;
      const uint32_t nrec = std::max<uint32_t>(c->next_pid, 1);
      PSPH_TRY(dmalloc((void **)&rec, (size_t)nrec * 3 * 8));
      tp.xrecb.push_back(rec);
      tp.xrecb_n.push_back(nrec);
      hipLaunchKernelGGL(k_pos_to_rec, blocks(nb), 256, 0, c->stream, nb, c->pid, c->f.x, c->f.y, c->f.z, rec, nrec);
    }
    const bool frozen = tp.mode == 2 && tp.frozen;
    if (frozen) {
      // the contributor sets of the recorded run: support tests on its
      // positions, whose displacement since the search differs from this
      // run's by the perturbation (the reach takes 1% more)
      if (tp.k >= (int64_t)tp.xrecb.size() || !c->px)
        return fail(PSPH_EINVAL, "replay: no recorded refresh positions for step %lld", (long long)tp.k);
      hipLaunchKernelGGL(k_pos_from_rec, blocks(nb), 256, 0, c->stream, nb, c->pid, c->f.bc, tp.xrecb[tp.k],
                         tp.xrecb_n[tp.k], c->f.x, c->f.y, c->f.z, c->px, c->py, c->pz, c->d_err);
      disp *= 1.01;
    }
    if (c->rec) PSPH_TRY(rec_capture(c, 4, dt));
    PSPH_TRY(apply_bcond_grid(c, disp, c->rec ? c->rec->d[RF_BW] : nullptr, c->rec ? c->rec->n : 0,
                              frozen ? c->px : nullptr, c->py, c->pz));
    if (c->rec) PSPH_TRY(rec_capture(c, 5, dt));
    if (frozen) PSPH_TRY(check_err(c, "replay refresh"));
  }
  if (tp.mode) tp.k++;
  // TimeProps::ifoutput (properties.h:104-120)
  bool output = false;
  if (c->time >= c->next_output || c->time == k.t_end) {
    c->ioutput++;
    c->next_output = ((c->ioutput + 1) * k.t_output) / 1.0;
    output = true;
  }
  if (out) {
    out->dt = dt; out->time = c->time; out->step = c->step;
    out->n_total = (int64_t)c->n - c->last_removed;
    out->n_real = c->last_real;            // the inlet adds erupt ghosts only
    out->n_injected = added; out->n_removed = c->last_removed;
    out->n_pairs = (int64_t)c->last_pairs;
    out->output = output;
    out->n_domain = c->last_domain;
  }
  return 0;
}

extern "C" int psph_run(psph_ctx *c, int64_t nsteps, psph_step_out *last) {
  for (int64_t s = 0; s < nsteps; s++) PSPH_TRY(psph_step(c, last));
  return 0;
}

extern "C" int psph_conservation(psph_ctx *c, double out[7]) {
  if (!c || !out) return fail(PSPH_EINVAL, "null argument");
  uint32_t n = (uint32_t)c->n;
  bool d3 = c->cfg.dim == 3;
  hipLaunchKernelGGL(k_conservation, kRedBlocks, 256, 0, c->stream, n, c->f.bc, c->f.m, c->f.e, c->f.vx,
                     d3 ? c->f.vy : nullptr, d3 ? c->f.vz : nullptr, c->f.xi, d3 ? c->f.z : nullptr, c->d_part);
  std::vector<double> part(kRedBlocks * 8);
  HIP_TRY(hipMemcpyAsync(part.data(), c->d_part, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double acc[8] = {0};
  for (int b = 0; b < kRedBlocks; b++)
    for (int k = 0; k < 8; k++) acc[k] += part[b * 8 + k];
  for (int k = 0; k < 7; k++) out[k] = acc[k];
  return 0;
}

extern "C" int psph_diag_height(psph_ctx *c, double *H) {
  if (!c || !H) return fail(PSPH_EINVAL, "null argument");
  double cons[7];
  uint32_t n = (uint32_t)c->n;
  bool d3 = c->cfg.dim == 3;
  hipLaunchKernelGGL(k_conservation, kRedBlocks, 256, 0, c->stream, n, c->f.bc, c->f.m, c->f.e, c->f.vx,
                     d3 ? c->f.vy : nullptr, d3 ? c->f.vz : nullptr, c->f.xi, d3 ? c->f.z : nullptr, c->d_part);
  std::vector<double> part(kRedBlocks * 8);
  HIP_TRY(hipMemcpyAsync(part.data(), c->d_part, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double num = 0, den = 0;
  for (int b = 0; b < kRedBlocks; b++) { den += part[b * 8 + 6]; num += part[b * 8 + 7]; }
  (void)cons;
  *H = den > 0 ? num / den : 0.0;
  return 0;
}

// Neighbour lists for parity checks, produced on the GPU by the same
// skeleton and membership rule as the physics passes, then mapped to read
// order and sorted (the reference's list order is bucket-insertion history;
// the set is what every pass consumes).
extern "C" int psph_get_neighbors(psph_ctx *c, int64_t *offsets, size_t n_cap, uint32_t *idx, size_t idx_cap) {
  if (!c || !offsets) return fail(PSPH_EINVAL, "null argument");
  if (!c->grid_valid) return fail(PSPH_EINVAL, "get_neighbors: call psph_neighbors first");
  const psph_config &k = c->cfg;
  uint32_t n = (uint32_t)c->n;
  // the list rule reaches the own bucket or the 4.25 h_i box (all pairs in 1D)
  const double r = (k.list_all || k.dim == 1) ? HUGE_VAL
                                              : std::max(k.list_factor * c->h_max, k.bucket_edge) * (1 + 1e-9);
  uint32_t *cnt = nullptr, *off = nullptr, *out = nullptr;
  PSPH_TRY(dmalloc((void **)&cnt, ((size_t)n + 1) * 4));
  PSPH_TRY(dmalloc((void **)&off, ((size_t)n + 1) * 4));
  HIP_TRY(hipMemsetAsync(cnt, 0, ((size_t)n + 1) * 4, c->stream));
  ListPass::Args a{0, cnt, nullptr, nullptr, k.cutoff};
  if (n) launch_pairs<ListPass>(c, make_common(c, r), a);
  size_t bytes = 0;
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, off, (int)n + 1, c->stream));
  PSPH_TRY(ensure_cub(c, bytes));
  HIP_TRY(hipcub::DeviceScan::ExclusiveSum(c->cub, bytes, cnt, off, (int)n + 1, c->stream));
  uint32_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, off + n, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  PSPH_TRY(dmalloc((void **)&out, ((size_t)total + 1) * 4));
  a = ListPass::Args{1, cnt, off, out, k.cutoff};
  if (n) launch_pairs<ListPass>(c, make_common(c, r), a);
  std::vector<uint32_t> hoff(n + 1), hout(total);
  HIP_TRY(hipMemcpyAsync(hoff.data(), off, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, c->stream));
  if (total) HIP_TRY(hipMemcpyAsync(hout.data(), out, (size_t)total * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  hipFree(cnt); hipFree(off); hipFree(out);
  // device order -> read order
  HostParticles P;
  PSPH_TRY(download(c, P));
  auto perm = canonical_order(c->cfg, P.k0.data(), P.k1.data(), P.k2.data(), n);   // read pos -> dev idx
  std::vector<uint32_t> rank(n);
  for (uint32_t q = 0; q < n; q++) rank[perm[q]] = q;
  if ((size_t)n + 1 > n_cap) return fail(PSPH_ECAPACITY, "get_neighbors: offsets need %u entries", n + 1);
  size_t o = 0;
  std::vector<uint32_t> tmp;
  for (uint32_t q = 0; q < n; q++) {
    offsets[q] = (int64_t)o;
    uint32_t d = perm[q];
    uint32_t b = hoff[d], e = hoff[d + 1];
    if (P.bc[d] != 100 && P.bc[d] != 99) continue;
    tmp.assign(hout.begin() + b, hout.begin() + e);
    for (auto &v : tmp) v = rank[v];
    std::sort(tmp.begin(), tmp.end());
    if (idx) {
      if (o + tmp.size() > idx_cap) return fail(PSPH_ECAPACITY, "get_neighbors: idx buffer holds %zu", idx_cap);
      std::memcpy(idx + o, tmp.data(), tmp.size() * 4);
    }
    o += tmp.size();
  }
  offsets[n] = (int64_t)o;
  return 0;
}

extern "C" int psph_profile(psph_ctx *c, int enable) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (enable && !c->ev0) {
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
  }
  c->prof = enable != 0;
  for (int k = 0; k < PSPH_NSTAGES; k++) { c->prof_ms[k] = 0; c->prof_n[k] = 0; }
  return 0;
}

extern "C" int psph_profile_read(psph_ctx *c, double ms[PSPH_NSTAGES], int64_t launches[PSPH_NSTAGES]) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  for (int k = 0; k < PSPH_NSTAGES; k++) {
    if (ms) ms[k] = c->prof_ms[k];
    if (launches) launches[k] = c->prof_n[k];
  }
  return 0;
}

#include "psph_adjoint.inl"
#include "psph_domain.inl"
#include "psph_noalloc.inl"
