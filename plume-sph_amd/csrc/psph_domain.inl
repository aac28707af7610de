// psph_domain.inl -- domain adaptation on the device (part of psph_device.hip).
//
// Each step of a run with cfg.adapt_domain ends with a per-bucket occupancy
// of the particles (any particle, an involved one, a pressure ghost, a wall
// ghost, a real one) reduced on the GPU, from which the host runs
// scan_outside_layer (scan_outside_layer.cc:44-121) on the bucket array.
// When it reports the outer layer reached, the next step, after its time
// step (main.cc:256-304), turns the particles of the bordering buckets real
// on the GPU (adapt_domain, adapt_domain.cc:41-194), and appends the new
// pressure- and wall-ghost layers that the host builds on the air lattice
// (add_pressure_ghost.cc, add_wall_ghost.cc), then refreshes the wall ghosts
// when cfg.wall_bc is set.

namespace {

// Bucket of a position (floor((x - ext_lo) / del) per axis), -1 outside.
struct BucketGeom {
  double lo[3], del;
  int nb[3];
  __device__ int64_t of(double x, double y, double z) const {
    const double p[3] = {x, y, z};
    int b[3];
    for (int k = 0; k < 3; k++) {
      const double f = floor((p[k] - lo[k]) / del);
      if (!(f >= 0 && f < nb[k])) return -1;
      b[k] = (int)f;
    }
    return ((int64_t)b[2] * nb[1] + b[1]) * nb[0] + b[0];
  }
};

constexpr uint32_t kOccReal = 16;   // beyond the host's OCC_* bits: a real particle

// OR of the occupancy bits of every live particle into its bucket.  The
// particles are in cell order, so a wave's lanes usually share a bucket:
// they combine their bits first and one lane does the atomic.
__global__ void k_bucket_occ(uint32_t n, const double *x, const double *y, const double *z, const int32_t *bc,
                             const int32_t *inv, BucketGeom g, uint32_t *occ) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  int64_t q = -1;
  uint32_t f = 0;
  if (i < n) {
    const int b = bc[i];
    if (b >= 0) {
      q = g.of(x[i], y[i], z[i]);
      f = 1u | (inv[i] == 2 ? 2u : 0u) | (b == 1 ? 4u : 0u) | (b == 2 ? 8u : 0u) |
          ((b == 100 || b == 99) ? kOccReal : 0u);
    }
  }
  const int64_t q0 = __shfl(q, 0);
  if (__all(q == q0)) {
    for (int o = 32; o > 0; o >>= 1) f |= (uint32_t)__shfl_xor((int)f, o);
    if ((threadIdx.x & (kWave - 1)) == 0 && q0 >= 0 && f) atomicOr(occ + q0, f);
  } else if (q >= 0) {
    atomicOr(occ + q, f);
  }
}

// adapt_domain's particle updates: conv[q] 1 every particle of the bucket,
// 2 its pressure ghosts, 3 those not outside dom on the faces the bucket
// straddles (Bucket::determine_escape, bucket.cc:435-452) turn real and
// potentially involved
__global__ void k_domain_convert(uint32_t n, const double *x, const double *y, const double *z, int32_t *bc,
                                 int32_t *inv, BucketGeom g, const uint8_t *conv, const uint8_t *sides,
                                 double lo0, double lo1, double lo2, double hi0, double hi1, double hi2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || bc[i] < 0) return;
  const double p[3] = {x[i], y[i], z[i]};
  const int64_t q = g.of(p[0], p[1], p[2]);
  if (q < 0) return;
  const uint8_t code = conv[q];
  if (!code) return;
  bool turn = true;
  if (code == 2) {
    turn = bc[i] == 1;
  } else if (code == 3) {
    constexpr double kTiny = 1.0E-08;
    const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
    const uint8_t sd = sides[q];
    for (int k = 0; k < 3; k++) {
      if (sd & (1u << (2 * k))) { if ((p[k] - lo[k]) < -kTiny) turn = false; }
      else if (sd & (1u << (2 * k + 1))) { if ((p[k] - hi[k]) > kTiny) turn = false; }
    }
  }
  if (turn) { bc[i] = 100; inv[i] = 1; }
}

BucketGeom bucket_geom(const psph_ctx *c) {
  BucketGeom g;
  for (int k = 0; k < 3; k++) { g.lo[k] = c->cfg.ext_lo[k]; g.nb[k] = c->dom.nb[k]; }
  g.del = c->cfg.bucket_edge;
  return g;
}

// occupancy bits of every bucket (OCC_* and kOccReal)
int domain_occupancy(psph_ctx *c, std::vector<uint8_t> &occ) {
  const size_t nb = c->dom.size();
  hipLaunchKernelGGL(k_fill_i32, blocks(nb), 256, 0, c->stream, (int *)c->d_occ, 0, (int)nb);
  if (c->n)
    hipLaunchKernelGGL(k_bucket_occ, blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.x, c->f.y, c->f.z,
                       c->f.bc, c->f.inv, bucket_geom(c), c->d_occ);
  std::vector<uint32_t> o(nb);
  HIP_TRY(hipMemcpyAsync(o.data(), c->d_occ, nb * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  occ.resize(nb);
  for (size_t q = 0; q < nb; q++) occ[q] = (uint8_t)o[q];
  return 0;
}

// the grow step: conversions on the device, new ghost layers appended
int domain_adapt(psph_ctx *c) {
  const psph_config &k = c->cfg;
  std::vector<uint8_t> occ, conv;
  PSPH_TRY(domain_occupancy(c, occ));
  const int nconv = domain_convert_plan(c->dom, occ, conv);
  if (nconv && c->n) {
    HIP_TRY(hipMemcpyAsync(c->d_conv, conv.data(), conv.size(), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_domain_convert, blocks(c->n), 256, 0, c->stream, (uint32_t)c->n, c->f.x, c->f.y, c->f.z,
                       c->f.bc, c->f.inv, bucket_geom(c), c->d_conv, c->d_sides, k.dom_lo[0], k.dom_lo[1],
                       k.dom_lo[2], k.dom_hi[0], k.dom_hi[1], k.dom_hi[2]);
    HIP_TRY(hipGetLastError());
    PSPH_TRY(domain_occupancy(c, occ));
  }
  HostParticles G;
  domain_add_ghosts(k, c->dom, occ, &c->met, c->dom_seed, c->dom_jitter, (uint32_t)c->step, G);
  const size_t m = G.size();
  if (c->n + m > c->cap)
    return fail(PSPH_ECAPACITY, "domain adaptation: %zu + %zu particles exceed capacity %zu", c->n, m, c->cap);
  PSPH_TRY(upload(c, G, c->n));
  if (m) hipLaunchKernelGGL(k_iota_u32, blocks(m), 256, 0, c->stream, c->pid + c->n, c->next_pid, (uint32_t)m);
  c->next_pid += (uint32_t)m;
  c->n += m;
  c->last_domain = (int64_t)m;
  c->grid_valid = false;
  return 0;
}

// scan_outside_layer at the end of a step
int domain_scan_step(psph_ctx *c) {
  std::vector<uint8_t> occ;
  PSPH_TRY(domain_occupancy(c, occ));
  c->dom_pending = domain_scan(c->dom, occ);
  return 0;
}

}  // namespace

extern "C" int psph_domain_init(psph_ctx *c, const psph_meteo *met, uint32_t seed, double jitter) {
  if (!c || !met || !met->z || !met->data || met->rows < 1) return fail(PSPH_EINVAL, "null argument");
  const psph_config &k = c->cfg;
  if (!k.adapt_domain || k.dim != 3) return fail(PSPH_EINVAL, "domain_init: needs dim 3 and cfg.adapt_domain");
  c->met_z.assign(met->z, met->z + met->rows);
  c->met_data.assign(met->data, met->data + (size_t)met->rows * met->nprop);
  c->met.rows = met->rows;
  c->met.nprop = met->nprop;
  c->met.z = c->met_z.data();
  c->met.data = c->met_data.data();
  c->dom_seed = seed;
  c->dom_jitter = jitter;
  domain_buckets(k, c->dom);
  const size_t nb = c->dom.size();
  if (c->d_occ) hipFree(c->d_occ);
  if (c->d_conv) hipFree(c->d_conv);
  if (c->d_sides) hipFree(c->d_sides);
  c->d_occ = nullptr; c->d_conv = nullptr; c->d_sides = nullptr;
  PSPH_TRY(dmalloc((void **)&c->d_occ, nb * 4));
  PSPH_TRY(dmalloc((void **)&c->d_conv, nb));
  PSPH_TRY(dmalloc((void **)&c->d_sides, nb));
  HIP_TRY(hipMemcpyAsync(c->d_sides, c->dom.sides.data(), nb, hipMemcpyHostToDevice, c->stream));
  // a state loaded mid-run: buckets holding real particles are potentially
  // involved as well
  std::vector<uint8_t> occ;
  PSPH_TRY(domain_occupancy(c, occ));
  for (size_t q = 0; q < nb; q++)
    if (occ[q] & kOccReal) c->dom.val[q] |= 1;
  c->dom_pending = 0;
  c->dom_ready = true;
  return 0;
}
