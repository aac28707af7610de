// psph_adjoint.inl -- reverse-mode adjoint of the SPH step (part of psph_device.hip).
//
// Differentiates the plume diagnostic H = sum m xi z / sum m xi over real
// particles after N steps with respect to the vent exit velocity, the vent
// temperature, the gas mass fraction ng0 of the erupted material and every
// particle's initial velocity.  The step map is differentiated with its
// discrete schedule frozen: the time steps, the adapted smoothing lengths,
// the inlet spawn counts and the erupt-ghost -> real and escape events of the
// recorded run (Tape).  The finite-difference checks run the same frozen map
// (psph_replay_run).
//
// Memory: a checkpoint of the whole state every K steps; a reverse segment
// recomputes its K steps from the checkpoint, keeping each step's
// intermediates (StepRec, ~256 B/particle/step), then sweeps them backwards.
// The adjoint variables live in arrays indexed by persistent particle id, so
// each step's cell-order permutation is a gather on entry and a scatter on
// exit.  Pair terms are reversed by a gather over chunks of rows, built like
// the forward pass's target chunks: a row particle collects both its own-role
// terms (it is the particle whose sum holds the pair) and its neighbour-role
// terms (it is the neighbour in a candidate's sum), so no atomics are needed
// and sums are deterministic.  All adjoint arithmetic is FP64.

namespace {

enum LamField { L_X, L_Y, L_Z, L_VX, L_VY, L_VZ, L_SVX, L_SVY, L_SVZ, L_RHO, L_E, L_XI, L_P, L_C, L_T, L_CP, L_M, L_N };
struct LamPtrs { double *f[L_N]; };

__global__ void k_lam_gather(uint32_t n, const uint32_t *pid, LamPtrs src, LamPtrs dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = pid[i];
#pragma unroll
  for (int k = 0; k < L_N; k++) dst.f[k][i] = src.f[k][id];
}
__global__ void k_lam_scatter(uint32_t n, const uint32_t *pid, LamPtrs src, LamPtrs dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t id = pid[i];
#pragma unroll
  for (int k = 0; k < L_N; k++) dst.f[k][id] = src.f[k][i];
}
__global__ void k_fill_f64(double *p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Seeds of H = sum m xi z / D, D = sum m xi (real particles).
__global__ void k_lam_seed(uint32_t n, const int32_t *bc, const uint32_t *pid, const double *m,
                           const double *xi, const double *z, double H, double D, LamPtrs L) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !is_real(bc[i])) return;
  const uint32_t id = pid[i];
  L.f[L_Z][id] = m[i] * xi[i] / D;
  L.f[L_XI][id] = m[i] * (z[i] - H) / D;
  L.f[L_M][id] = xi[i] * (z[i] - H) / D;
}

// update_pos reverse (update_pos.cc:34-168): x += dt sv (real) or dt v
// (erupting ghost); x is carried, so only the velocity adjoint moves.
__global__ void k_adj_pos(uint32_t n, const int32_t *bc, int erupting, int lans, double dt, LamPtrs l) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = bc[i];
  int v0;
  if (is_real(b)) v0 = lans ? L_SVX : L_VX;
  else if (b == 0 && erupting) v0 = L_VX;
  else return;
#pragma unroll
  for (int k = 0; k < 3; k++) l.f[v0 + k][i] += dt * l.f[L_X + k][i];
}

// Velocity filter reverse, per target: sv'' = sum_j w_ij v'_j / Wn, so with
// a = lambda_sv'' / Wn the pair weight gets a.v'_j - a.sv''.  sv is
// overwritten for targets: its adjoint restarts at zero.
__global__ void k_adj_vsm_pre(uint32_t n, const int32_t *bc, const double *wn, const double *s0,
                              const double *s1, const double *s2, LamPtrs l, double *ax, double *ay,
                              double *az, double *ac) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!need_neigh(bc[i])) { ax[i] = ay[i] = az[i] = ac[i] = 0; return; }
  const double iw = 1.0 / wn[i];
  const double a0 = l.f[L_SVX][i] * iw, a1 = l.f[L_SVY][i] * iw, a2 = l.f[L_SVZ][i] * iw;
  ax[i] = a0; ay[i] = a1; az[i] = a2;
  ac[i] = a0 * s0[i] + a1 * s1[i] + a2 * s2[i];
  l.f[L_SVX][i] = l.f[L_SVY][i] = l.f[L_SVZ][i] = 0;
}

// d/d(rho, e, xi, ng0) of lp p + lc c + lT T + lcp Cp for the closed-form EOS
// (particle.cc:528-604): p = rho e u, c^2 = (u + u^2) e / G, T = e / Cvm,
// Cp = Cvm (LANS), u = Rm / Cvm, G = ng + na.
__device__ inline void eos_adj(const EosConst &k, double rho, double e, double xi, double lp, double lc,
                               double lT, double lcp, double &grho, double &ge, double &gxi, double &gng0) {
  const double ng0 = k.ng0;
  const double na = 1 - xi, ng = ng0 * xi, ns = xi - ng;
  const double Cvm = ns * k.cvs + ng * k.cvg + na * k.cva;
  const double Rm = ng * k.rg + na * k.ra;
  const double G = ng + na;
  const double u = Rm / Cvm;
  const double c = sqrt((u + u * u) * e / G);
  const double T = e / Cvm;
  double gu = lp * rho * e;
  double gCvm = -lT * T / Cvm + (k.lans ? lcp : 0.0);
  double gG = 0;
  ge = lp * rho * u + lT / Cvm;
  grho = lp * e * u;
  if (c > 0) {
    gu += lc * c * (1 + 2 * u) / (2 * (u + u * u));
    ge += lc * c / (2 * e);
    gG -= lc * c / (2 * G);
  }
  gCvm -= gu * u / Cvm;
  const double gRm = gu / Cvm;
  gxi = gCvm * ((1 - ng0) * k.cvs + ng0 * k.cvg - k.cva) + gRm * (ng0 * k.rg - k.ra) - gG * (1 - ng0);
  gng0 = gCvm * xi * (k.cvg - k.cvs) + gRm * xi * k.rg + gG * xi;
}

// Density reverse, per target: EOS, then rho = r1 + r2, xi = r2 / rho,
// e = En / Wn (energy smoothing).  The pass overwrites rho, xi, e and the EOS
// outputs of targets, so their adjoints restart at zero (e keeps the direct
// term when energy smoothing is off).
__global__ void k_adj_dens_pre(uint32_t n, const int32_t *bc, const double *rho2, const double *e2,
                               const double *xi2, const double *wne, EosConst eos, int energy_smooth,
                               LamPtrs l, double *s1, double *s2, double *sen, double *swn, double *gng0) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!need_neigh(bc[i])) { s1[i] = s2[i] = sen[i] = swn[i] = 0; gng0[i] = 0; return; }
  const double rho = rho2[i], e = e2[i], xi = xi2[i];
  double grho, ge, gxi, gn;
  eos_adj(eos, rho, e, xi, l.f[L_P][i], l.f[L_C][i], l.f[L_T][i], l.f[L_CP][i], grho, ge, gxi, gn);
  grho += l.f[L_RHO][i];
  ge += l.f[L_E][i];
  gxi += l.f[L_XI][i];
  gng0[i] = gn;
  s1[i] = grho - gxi * xi / rho;
  s2[i] = grho + gxi * (1 - xi) / rho;
  if (energy_smooth) {
    const double iw = 1.0 / wne[i];
    sen[i] = ge * iw;
    swn[i] = -ge * e * iw;
    l.f[L_E][i] = 0;
  } else {
    sen[i] = swn[i] = 0;
    l.f[L_E][i] = ge;
  }
  l.f[L_RHO][i] = 0; l.f[L_XI][i] = 0; l.f[L_P][i] = 0; l.f[L_C][i] = 0; l.f[L_T][i] = 0;
  if (eos.lans) l.f[L_CP][i] = 0;
}

// RHS reverse, per target: v' = v + dt (rhs_v + g), e' = e + dt (rhs_e + vl g):
// the pair seeds are dt lambda_v', dt lambda_e'; v and e carry lambda through.
__global__ void k_adj_rhs_pre(uint32_t n, const int32_t *bc, double dt, double gz, int lans, LamPtrs l,
                              double *a0, double *a1, double *a2, double *ae) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!need_neigh(bc[i])) { a0[i] = a1[i] = a2[i] = ae[i] = 0; return; }
  const double le = l.f[L_E][i];
  a0[i] = dt * l.f[L_VX][i]; a1[i] = dt * l.f[L_VY][i]; a2[i] = dt * l.f[L_VZ][i];
  ae[i] = dt * le;
  l.f[lans ? L_SVZ : L_VZ][i] += dt * gz * le;
}

// ---------------------------------------------------------------------------
// Adjoint pair passes.  Rows: the particles whose state can depend on the
// inputs (real particles, erupt ghosts, boundary ghosts next to targets),
// in chunks of kChunk consecutive rows of one cell.
// ---------------------------------------------------------------------------
struct AdjCommon {
  const double *x, *y, *z, *h, *h_list;
  const uint32_t *refb, *rows, *cfirst, *ccount, *nchunks;
  const int32_t *bc;
  Traversal tr;
  int list_all;
  int list_adj;              // some list box may reach past the adjacent buckets
  double list_factor, cutoff2;
  CutTest cut;
};

// Visit lists of a reversed step, shared by its three adjoint passes: for
// every chunk of rows, the tiles that meet its grown box and hold a candidate
// of either role of any pass, with two 32-bit masks per lane (own role,
// neighbour role; bit b <-> candidate sub + b kSub).  List membership,
// in_support and the row/candidate target tests are exact; the kernel box is
// widened to the union of the passes' boxes (cut * max(h_s, h_ij), h_s the
// summing particle's h) with a relative margin of 1e-12.  Each pass re-tests
// its own box and masks on the set bits.
static_assert(kWave / kSub <= 32, "the adjoint's role masks hold one bit per candidate of a lane's share");

struct AdjVisits {
  const uint32_t *off, *num;
  const uint2 *ent;          // (first particle, particle count) of the visited tile
  const uint64_t *mask;      // per visit and lane: own-role mask | neighbour-role mask << 32
};

__global__ __launch_bounds__(kWave) void k_adj_visit_count(AdjCommon pc, uint32_t *vcount) {
  const uint32_t chunk = blockIdx.x;
  const uint32_t nch = *pc.nchunks;
  if (chunk >= nch) {
    if (threadIdx.x == 0) vcount[chunk] = 0;
    return;
  }
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const uint32_t slot = (uint32_t)threadIdx.x / kSub;
  const uint32_t i = pc.rows[cf + (slot < cc ? slot : 0)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const ChunkBox box = chunk_box(pc.tr, xi);
  const uint32_t cnt = count_meeting_tiles(pc.tr, box);
  if (threadIdx.x == 0) vcount[chunk] = cnt;
}

// The scan is transposed as the forward one (k_visit_masks): lane l holds
// candidate 2 (l mod 32) + l / 32 of the tile, the wave loops over the
// chunk's rows, and the two ballots of a row's role tests are, in their low
// and high halves, the masks of the row's two sub-lanes.  A row is skipped
// for a tile when its distance to the tile's box exceeds cut * max(h_i,
// h_max of the tile): both roles' kernel boxes, cut * max(h_s, h_ij) with h_s
// the summing particle's h, lie within that reach for every candidate of the
// tile, and the rounded pair quantities stay on the same side of it (a
// rounded mean lies between its operands; rounding is monotonic), so the
// masks are exactly those of testing every row.
struct AdjVisitRow {
  double x, y, z, h, lf, supp;   // position, h, list box, in_support box
  uint32_t refb;
  int tgt;                       // need_neigh: the row sums (own role)
};

__global__ __launch_bounds__(kWave) void k_adj_visit_masks(AdjCommon pc, double cutx, const uint32_t *voff,
                                                           uint32_t *vnum, uint2 *vent, uint64_t *vmask) {
  static_assert(kSub == 2, "the transposed scan maps the ballot halves onto two sub-lanes");
  __shared__ AdjVisitRow rw[kChunk];
  const uint32_t nch = *pc.nchunks;
  const uint32_t per = (nch + kXcds - 1) / kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  if (local >= per) return;
  const uint32_t chunk = (blockIdx.x % kXcds) * per + local;
  if (chunk >= nch) return;
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const int lane = threadIdx.x;
  // lanes 0..cc-1 own one row each: its record in LDS, its position in registers
  const bool owner = (uint32_t)lane < cc;
  const uint32_t i = pc.rows[cf + (owner ? (uint32_t)lane : 0u)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const double hi = pc.h[i];
  if (owner)
    rw[lane] = AdjVisitRow{xi[0], xi[1], xi[2], hi, pc.list_factor * pc.h_list[i], pc.cutoff2 * hi,
                           pc.refb[i], (int)need_neigh(pc.bc[i])};
  // the chunk's grown box (every lane's row counts, as in the passes)
  const uint32_t slot = (uint32_t)lane / kSub;
  const uint32_t ib = pc.rows[cf + (slot < cc ? slot : 0u)];
  const double xb[3] = {pc.x[ib], pc.y[ib], pc.z[ib]};
  const ChunkBox box = chunk_box(pc.tr, xb);
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
          const bool valid = cq < cnt;
          const uint32_t j = base + (valid ? cq : 0u);
          const double cxp = valid ? pc.x[j] : kFar;
          const double cyp = valid ? pc.y[j] : kFar;
          const double czp = valid ? pc.z[j] : kFar;
          const double hj = valid ? pc.h[j] : 0.0;
          const double lfj = valid ? pc.list_factor * pc.h_list[j] : 0.0;
          const uint32_t crb = valid ? pc.refb[j] : ~0u;
          const bool nn_j = valid && need_neigh(pc.bc[j]);
          const double c2hj = pc.cutoff2 * hj;
          // rows within reach of the tile in either role
          const double hmax_t = wave_max(hj);
          const float4 tlo = pc.tr.tbox[2 * tl], thi = pc.tr.tbox[2 * tl + 1];
          const double r_i = cutx * fmax(hi, hmax_t);
          const double dxl = fmax((double)tlo.x - xi[0], xi[0] - (double)thi.x);
          const double dyl = fmax((double)tlo.y - xi[1], xi[1] - (double)thi.y);
          const double dzl = fmax((double)tlo.z - xi[2], xi[2] - (double)thi.z);
          const bool meets = owner && dxl <= r_i && dyl <= r_i && dzl <= r_i;
          uint64_t tm = __ballot(meets);
          if (!tm) continue;
          uint32_t mown = 0, mnbr = 0;   // this lane's halves of the two role ballots
          while (tm) {
            const uint32_t t = (uint32_t)__builtin_ctzll(tm);
            clear_bit_u(tm, t);
            const AdjVisitRow R = rw[t];
            const double amax = fmax(fabs(R.x - cxp), fmax(fabs(R.y - cyp), fabs(R.z - czp)));
            const bool same = (pc.list_all != 0) | (crb == R.refb);
            const bool adj = (pc.list_adj == 0) || refb_adjacent(crb, R.refb);
            const double hij = 0.5 * (R.h + hj);
            // row t sums, this lane's candidate is its neighbour
            const bool own = (R.tgt != 0) & (same | ((amax <= R.lf) & adj)) & (amax <= R.supp) &
                             (amax <= cutx * fmax(R.h, hij));
            // this lane's candidate sums, row t is its neighbour
            const bool nbr = nn_j & (same | ((amax <= lfj) & adj)) & (amax <= c2hj) & (amax <= cutx * fmax(hj, hij));
            const uint64_t bo = __ballot(own), bn = __ballot(nbr);
            // lanes 2t and 2t+1 (row t's sub-lanes) take the two halves of
            // each ballot (one v_writelane each)
            set_lane(mown, (uint32_t)bo, 2u * t);
            set_lane(mown, (uint32_t)(bo >> 32), 2u * t + 1u);
            set_lane(mnbr, (uint32_t)bn, 2u * t);
            set_lane(mnbr, (uint32_t)(bn >> 32), 2u * t + 1u);
          }
          const uint64_t mask = (uint64_t)mnbr << 32 | mown;
          if (__ballot(mask != 0)) {
            if (lane == 0) vent[w] = make_uint2(base, cnt);
            __builtin_nontemporal_store(mask, vmask + (size_t)w * kWave + lane);
            w++;
          }
        }
      }
  if (lane == 0) vnum[chunk] = w - w0;
}

// An adjoint pair pass over the visit lists.  For row i and candidate j it
// evaluates the forward pass's exact pair predicate with i as the summing
// particle (own role, i a target) and with j as the summing particle
// (neighbour role, j a target), and calls pair() once with both flags and
// dx = x_i - x_j.  Within a pass, dx given to an own()/nbr() role routine is
// always summing minus neighbour.
template <class P>
__global__ __launch_bounds__(kWave) void k_adj_pairs(AdjCommon pc, AdjVisits vl, typename P::Args a) {
  __shared__ double sx[kWave], sy[kWave], sz[kWave], sh[kWave];
  __shared__ typename P::Tile tile;
  const uint32_t nch = *pc.nchunks;
  const uint32_t per = (nch + kXcds - 1) / kXcds;
  const uint32_t local = blockIdx.x / kXcds;
  if (local >= per) return;
  const uint32_t chunk = (blockIdx.x % kXcds) * per + local;
  if (chunk >= nch) return;
  const uint32_t cf = pc.cfirst[chunk], cc = pc.ccount[chunk];
  const int lane = threadIdx.x;
  // kSub lanes per row, taking alternate candidates of each tile
  const uint32_t slot = (uint32_t)lane / kSub, sub = (uint32_t)lane % kSub;
  const bool active = slot < cc;
  const uint32_t i = pc.rows[cf + (active ? slot : 0)];
  const double xi[3] = {pc.x[i], pc.y[i], pc.z[i]};
  const double hi = pc.h[i];
  typename P::Row t;
  P::load_row(a, i, t);
  P::init(tile, lane);
  __syncthreads();
  const uint32_t v0 = vl.off[chunk], nv = vl.num[chunk];
  for (uint32_t v = 0; v < nv; v++) {
    const uint2 e = vl.ent[v0 + v];
    const uint32_t base = e.x, cnt = e.y;
    if ((uint32_t)lane < cnt) {
      const uint32_t j = base + lane;
      sx[lane] = pc.x[j]; sy[lane] = pc.y[j]; sz[lane] = pc.z[j];
      sh[lane] = pc.h[j];
      P::stage(a, j, tile, lane);
    }
    const uint64_t m64 = __builtin_nontemporal_load(vl.mask + (size_t)(v0 + v) * kWave + lane);
    const uint2 m = make_uint2((uint32_t)m64, (uint32_t)(m64 >> 32));
    __syncthreads();
    if (active) {
      uint32_t mb = m.x | m.y;
      while (mb) {
        const uint32_t b = (uint32_t)__builtin_ctz(mb);
        mb &= mb - 1;
        const uint32_t q = sub + b * kSub, jq = base + q;
        const double dx[3] = {xi[0] - sx[q], xi[1] - sy[q], xi[2] - sz[q]};
        const double amax = fmax(fabs(dx[0]), fmax(fabs(dx[1]), fabs(dx[2])));
        // all operands evaluated (no short-circuit branches), combined as ints
        const bool fo = ((int)((m.x >> b) & 1u) & (int)P::own_ok(t, tile, q, jq, i) &
                         (int)!pc.cut.outside(amax, P::hker_own(hi, sh[q]))) != 0;
        const bool fn = ((int)((m.y >> b) & 1u) & (int)P::nbr_ok(t, tile, q, jq, i) &
                         (int)!pc.cut.outside(amax, P::hker_nbr(hi, sh[q]))) != 0;
        if (fo | fn) P::pair(a, t, tile, q, dx, amax, hi, sh[q], fo, fn);
      }
    }
    __syncthreads();
  }
  P::combine(t);
  if (active && sub == 0) P::finish(a, t, i);
}

__device__ inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// 1 / x in FP64: hardware reciprocal refined by two Newton steps
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * fma(-x, r, 2.0);
  return r * fma(-x, r, 2.0);
}

// --- velocity filter (smooth_velocity.cc:25-165) ----------------------------
struct VsmAdj {
  struct Args {
    const double *m, *rho, *h, *v0, *v1, *v2;   // rho'' and v' of the step
    const double *ax, *ay, *az, *ac;             // per-target seeds
    const int32_t *bc;
    LamPtrs l;
  };
  struct Tile {
    double mr[kWave], ih2[kWave], nrm[kWave], v[3][kWave], a[3][kWave], c[kWave], etab[kETab];
    int ok[kWave];
  };
  struct Row { double imi, irho, mr, ih2, nrm, v[3], a[3], c, g[3], gv[3], gm, grho; bool ok; };
  __device__ static void init(Tile &s, int lane) { etab_fill(s.etab, lane); }
  __device__ static void load_row(const Args &A, uint32_t i, Row &t) {
    const double mi = A.m[i], rhoi = A.rho[i], ih = 1.0 / A.h[i], nk = kInvSqrtPi * ih;
    t.imi = 1.0 / mi; t.irho = 1.0 / rhoi; t.mr = mi * t.irho;
    t.ih2 = ih * ih; t.nrm = nk * nk * nk;
    t.v[0] = A.v0[i]; t.v[1] = A.v1[i]; t.v[2] = A.v2[i];
    t.a[0] = A.ax[i]; t.a[1] = A.ay[i]; t.a[2] = A.az[i]; t.c = A.ac[i];
    for (int k = 0; k < 3; k++) t.g[k] = t.gv[k] = 0;
    t.gm = t.grho = 0;
    t.ok = contr_vel(A.bc[i]);
  }
  __device__ static void stage(const Args &A, uint32_t j, Tile &s, int l) {
    const double ih = 1.0 / A.h[j], nk = kInvSqrtPi * ih;
    s.mr[l] = A.m[j] / A.rho[j];
    s.ih2[l] = ih * ih; s.nrm[l] = nk * nk * nk;
    s.v[0][l] = A.v0[j]; s.v[1][l] = A.v1[j]; s.v[2][l] = A.v2[j];
    s.a[0][l] = A.ax[j]; s.a[1][l] = A.ay[j]; s.a[2][l] = A.az[j]; s.c[l] = A.ac[j];
    s.ok[l] = contr_vel(A.bc[j]);
  }
  __device__ static double hker_own(double hi, double) { return hi; }
  __device__ static double hker_nbr(double, double hj) { return hj; }
  __device__ static bool own_ok(const Row &, const Tile &s, uint32_t q, uint32_t, uint32_t) { return s.ok[q] != 0; }
  __device__ static bool nbr_ok(const Row &t, const Tile &, uint32_t, uint32_t, uint32_t) { return t.ok; }
  __device__ static void own(const Args &, Row &t, const Tile &s, uint32_t q, const double *dx, double,
                             double, double) {
    const double w = t.nrm * exp_neg_tab(dot3(dx, dx) * t.ih2, s.etab) * s.mr[q];
    const double gw = t.a[0] * s.v[0][q] + t.a[1] * s.v[1][q] + t.a[2] * s.v[2][q] - t.c;
    const double f = -2.0 * gw * w * t.ih2;
    for (int k = 0; k < 3; k++) t.g[k] += f * dx[k];
  }
  __device__ static void nbr(const Args &, Row &t, const Tile &s, uint32_t q, const double *dx, double,
                             double, double) {
    const double ih2 = s.ih2[q];
    const double w = s.nrm[q] * exp_neg_tab(dot3(dx, dx) * ih2, s.etab) * t.mr;
    const double a[3] = {s.a[0][q], s.a[1][q], s.a[2][q]};
    const double gw = dot3(a, t.v) - s.c[q];
    for (int k = 0; k < 3; k++) t.gv[k] += a[k] * w;
    t.gm += gw * w * t.imi;
    t.grho -= gw * w * t.irho;
    const double f = -2.0 * gw * w * ih2;
    for (int k = 0; k < 3; k++) t.g[k] -= f * dx[k];
  }
  // the two roles use different kernels (h_i and h_j): nothing to share
  __device__ static void pair(const Args &A, Row &t, const Tile &s, uint32_t q, const double *dx, double amax,
                              double hi, double hj, bool fo, bool fn) {
    if (fo) own(A, t, s, q, dx, amax, hi, hj);
    if (fn) {
      const double dn[3] = {-dx[0], -dx[1], -dx[2]};
      nbr(A, t, s, q, dn, amax, hi, hj);
    }
  }
  __device__ static void combine(Row &t) {
    for (int k = 0; k < 3; k++) { t.g[k] = sub_sum<kSub>(t.g[k]); t.gv[k] = sub_sum<kSub>(t.gv[k]); }
    t.gm = sub_sum<kSub>(t.gm); t.grho = sub_sum<kSub>(t.grho);
  }
  __device__ static void finish(const Args &A, Row &t, uint32_t i) {
    for (int k = 0; k < 3; k++) { A.l.f[L_X + k][i] += t.g[k]; A.l.f[L_VX + k][i] += t.gv[k]; }
    A.l.f[L_M][i] += t.gm;
    A.l.f[L_RHO][i] += t.grho;
  }
};

// --- density + energy smoothing (density_massfrac.cc:25-359) ----------------
struct DensAdj {
  struct Args {
    const double *m, *rho, *e1;                  // rho of the step's start (energy weights), e'
    const int32_t *bc, *phase;
    const double *s1, *s2, *sen, *swn;           // per-target seeds
    double e_ratio;
    int pghost, energy_smooth, e_sq16;
    CutTest cut;
    LamPtrs l;
  };
  struct Tile {
    double m[kWave], irho[kWave], e1[kWave], s1[kWave], s2[kWave], sen[kWave], swn[kWave], etab[kETab];
    int ok[kWave];
  };
  struct Row { double mi, imi, irho, e1i, s1, s2, sen, swn, g[3], gm, grho, ge; int ok; };
  __device__ static void init(Tile &s, int lane) { etab_fill(s.etab, lane); }
  __device__ static void load_row(const Args &A, uint32_t i, Row &t) {
    t.mi = A.m[i]; t.imi = 1.0 / t.mi; t.irho = 1.0 / A.rho[i]; t.e1i = A.e1[i];
    t.s1 = A.s1[i]; t.s2 = A.s2[i]; t.sen = A.sen[i]; t.swn = A.swn[i];
    const int ph = A.phase[i];
    t.ok = (contr_dens(A.pghost, A.bc[i]) && (ph == 1 || ph == 2)) ? ph : 0;
    t.g[0] = t.g[1] = t.g[2] = 0;
    t.gm = t.grho = t.ge = 0;
  }
  __device__ static void stage(const Args &A, uint32_t j, Tile &s, int l) {
    s.m[l] = A.m[j]; s.irho[l] = 1.0 / A.rho[j]; s.e1[l] = A.e1[j];
    s.s1[l] = A.s1[j]; s.s2[l] = A.s2[j]; s.sen[l] = A.sen[j]; s.swn[l] = A.swn[j];
    const int ph = A.phase[j];
    s.ok[l] = (contr_dens(A.pghost, A.bc[j]) && (ph == 1 || ph == 2)) ? ph : 0;
  }
  __device__ static double hker_own(double hi, double hj) { return 0.5 * (hi + hj); }
  __device__ static double hker_nbr(double hi, double hj) { return 0.5 * (hj + hi); }
  __device__ static bool own_ok(const Row &, const Tile &s, uint32_t q, uint32_t, uint32_t) { return s.ok[q] != 0; }
  __device__ static bool nbr_ok(const Row &t, const Tile &, uint32_t, uint32_t, uint32_t) { return t.ok != 0; }
  // kernel values of the pair: wm / m_j (density) and the energy weight / (m_j / rho_j)
  __device__ static void weights(const Args &A, const Tile &s, double r2, double amax, double h, double &w,
                                 double &we, double &ih2, bool &ebox) {
    const double ih = rcp64(h);
    ih2 = ih * ih;
    const double nk = kInvSqrtPi * ih;
    const double ex = exp_neg_tab(r2 * ih2, s.etab);
    w = nk * nk * nk * ex;
    ebox = A.energy_smooth && !A.cut.outside(A.e_ratio * amax, h);
    we = 0;
    if (ebox) {
      const double nke = nk * A.e_ratio;
      we = nke * nke * nke * (A.e_sq16 ? pow16(ex) : exp_neg_tab(A.e_ratio * A.e_ratio * r2 * ih2, s.etab));
    }
  }
  // h_ij is symmetric: both roles share the pair's two kernel weights.
  // dx = x_i - x_j; the neighbour role's position term, 2 f (x_j - x_i) added,
  // is the same subtraction as the own role's.
  __device__ static void pair(const Args &A, Row &t, const Tile &s, uint32_t q, const double *dx, double amax,
                              double hi, double hj, bool fo, bool fn) {
    double w, we, ih2;
    bool ebox;
    weights(A, s, dot3(dx, dx), amax, 0.5 * (hi + hj), w, we, ih2, ebox);
    double f = 0;
    if (fo) {
      const double lr = s.ok[q] == 1 ? t.s1 : t.s2;
      f = lr * (w * s.m[q]) * ih2;
      if (ebox) {
        const double vr = we * s.m[q] * s.irho[q];
        const double gvr = t.swn + t.sen * s.e1[q];
        f += gvr * vr * (A.e_ratio * A.e_ratio) * ih2;
      }
      for (int k = 0; k < 3; k++) t.g[k] -= 2.0 * f * dx[k];
    }
    if (fn) {
      const double lr = t.ok == 1 ? s.s1[q] : s.s2[q];
      f = lr * (w * t.mi) * ih2;
      t.gm += lr * w;
      if (ebox) {
        const double vr = we * t.mi * t.irho;
        const double gvr = s.swn[q] + s.sen[q] * t.e1i;
        t.ge += s.sen[q] * vr;
        t.gm += gvr * we * t.irho;
        t.grho -= gvr * vr * t.irho;
        f += gvr * vr * (A.e_ratio * A.e_ratio) * ih2;
      }
      for (int k = 0; k < 3; k++) t.g[k] -= 2.0 * f * dx[k];
    }
  }
  __device__ static void combine(Row &t) {
    for (int k = 0; k < 3; k++) t.g[k] = sub_sum<kSub>(t.g[k]);
    t.gm = sub_sum<kSub>(t.gm); t.grho = sub_sum<kSub>(t.grho); t.ge = sub_sum<kSub>(t.ge);
  }
  __device__ static void finish(const Args &A, Row &t, uint32_t i) {
    for (int k = 0; k < 3; k++) A.l.f[L_X + k][i] += t.g[k];
    A.l.f[L_M][i] += t.gm;
    A.l.f[L_RHO][i] += t.grho;
    A.l.f[L_E][i] += t.ge;
  }
};

// --- momentum / energy RHS (mom_engr_update.cc:30-439) ----------------------
struct RhsAdjConst { double alpha, beta, ata, eps_half, prandtl; int lans; };
struct RhsSide { double m, im, rho, V, p, c, T, cp, vel[3]; };
struct RhsGrad { double dx[3], vij[3], m_n, rho_s, rho_n, p_s, p_n, c_s, c_n, T_s, T_n, cp_s, cp_n; };

// Reverse of one pair term of the summing particle s with neighbour n,
// dx = x_s - x_n, seeds av = dt lambda_v'_s, ae = dt lambda_e'_s:
//   L = av . (-M dW) + ae (M dW.v_sn / 2 + heat),  dW = G dx,
//   G = -2 W / h^2, M = m_n (p_n / rho_n^2 + p_s / rho_s^2 + Pi).
// The pair's geometry (kernel, viscosity switch and mu, heat denominator) is
// symmetric in s <-> n (dx and v_sn both change sign), so it is computed once
// (RhsGeo) and serves both roles.  Quotients share reciprocals: 1/h,
// 1/(den rho_ab) for the viscosity and 1/(sum|v_k dx_k| h c_ab) for the heat.
struct RhsGeo {
  double h, ih2, r2, G, vrab, dotvv, cab, mu, irhoab, iden, vis_a, dotvr, Gd, cpab, hinv;
  bool avis, heat;
};

__device__ inline RhsGeo rhs_geo(const RhsAdjConst &k, const double *etab, const double *dx, const double *vij,
                                 double h, const RhsSide &S, const RhsSide &N) {
  RhsGeo q;
  q.h = h;
  const double ih = rcp64(h);
  q.ih2 = ih * ih;
  q.r2 = dot3(dx, dx);
  const double nk = kInvSqrtPi * ih;
  const double W = nk * nk * nk * exp_neg_tab(q.r2 * q.ih2, etab);
  q.G = -2.0 * q.ih2 * W;
  q.vrab = dot3(dx, vij);
  q.dotvv = dot3(vij, vij);
  q.cab = 0.5 * (S.c + N.c);
  q.avis = q.vrab < 0;
  q.vis_a = q.mu = q.irhoab = q.iden = 0;
  if (q.avis) {
    const double rhoab = 0.5 * (S.rho + N.rho);
    const double den = q.r2 + k.ata * h * h;
    const double inv = rcp64(den * rhoab);
    q.irhoab = den * inv;
    q.iden = rhoab * inv;
    q.mu = h * q.vrab * q.iden;
    q.vis_a = (k.beta * q.mu - k.alpha * q.cab) * q.mu * q.irhoab;
  }
  q.dotvr = 0;
  for (int d = 0; d < 3; d++) q.dotvr += fabs(vij[d] * dx[d]);
  q.Gd = q.G * q.r2;
  q.heat = k.lans && q.dotvr != 0 && q.Gd != 0;
  q.cpab = q.hinv = 0;
  if (q.heat) {
    q.cpab = 0.5 * (S.cp + N.cp);
    q.hinv = rcp64(q.dotvr * h * q.cab);
  }
  return q;
}

// The role-dependent part, with dx and vij oriented summing minus neighbour.
__device__ inline void rhs_role(const RhsAdjConst &k, const RhsGeo &q, const double *dx, const double *vij,
                                const RhsSide &S, const RhsSide &N, const double *av, double ae, RhsGrad &g) {
  const double Vn = N.V, Vs = S.V;
  double vis = q.vis_a;
  if (k.lans) vis -= k.eps_half * q.dotvv * Vn;
  const double PV2n = N.p * Vn * Vn, PV2s = S.p * Vs * Vs;
  const double Msum = PV2n + PV2s + vis;
  const double mpv = N.m * Msum;
  const double Q = -dot3(av, dx) + 0.5 * ae * q.vrab;
  const double g_mpv = q.G * Q, g_Q = mpv * q.G;
  double g_G = mpv * Q;
  double g_r2 = 0, g_dotvv = 0, g_cab = 0, g_vrab = 0.5 * ae * g_Q;
  double g_dx[3], g_vij[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) g_dx[d] = -g_Q * av[d];
  g.m_n = g_mpv * Msum;
  const double g_vis = g_mpv * N.m;
  g.p_n = g_vis * Vn * Vn;
  g.rho_n = -2.0 * g_vis * PV2n * Vn;
  g.p_s = g_vis * Vs * Vs;
  g.rho_s = -2.0 * g_vis * PV2s * Vs;
  if (k.lans) {
    g_dotvv -= g_vis * k.eps_half * Vn;
    g.rho_n += g_vis * k.eps_half * q.dotvv * Vn * Vn;
  }
  if (q.avis) {
    const double g_mu = g_vis * (2.0 * k.beta * q.mu - k.alpha * q.cab) * q.irhoab;
    g_cab -= g_vis * k.alpha * q.mu * q.irhoab;
    const double g_rhoab = -g_vis * q.vis_a * q.irhoab;
    g.rho_s += 0.5 * g_rhoab;
    g.rho_n += 0.5 * g_rhoab;
    g_vrab += g_mu * q.h * q.iden;
    g_r2 -= g_mu * q.mu * q.iden;
  }
  g.T_s = g.T_n = g.cp_s = g.cp_n = 0;
  if (q.heat) {
    // heat = base cp_ab |v|^2 Gd (T_s - T_n), Gd = G r^2 = dW.dx,
    // base = 10 eps / Pr m_n V_n V_s / (sum|v_k dx_k| h c_ab)
    const double base = (10.0 * k.eps_half / k.prandtl) * N.m * Vn * Vs * q.hinv;
    const double B = base * q.cpab * q.dotvv * q.Gd;
    const double dT = S.T - N.T;
    const double hh = ae * B * dT;
    const double aeb = ae * dT * base;
    g.T_s = ae * B;
    g.T_n = -ae * B;
    g.m_n += hh * N.im;
    g.rho_n -= hh * Vn;
    g.rho_s -= hh * Vs;
    g.cp_s = g.cp_n = 0.5 * aeb * q.dotvv * q.Gd;
    g_dotvv += aeb * q.cpab * q.Gd;
    const double g_Gd = aeb * q.cpab * q.dotvv;
    g_G += g_Gd * q.r2;
    g_r2 += g_Gd * q.G;
    const double g_dotvr = -hh * (q.hinv * q.h * q.cab);
    for (int d = 0; d < 3; d++) {
      const double pr = vij[d] * dx[d];
      const double sg = pr > 0 ? 1.0 : (pr < 0 ? -1.0 : 0.0);
      g_vij[d] += g_dotvr * sg * dx[d];
      g_dx[d] += g_dotvr * sg * vij[d];
    }
    g_cab -= hh * (q.hinv * q.dotvr * q.h);
  }
  g_r2 -= g_G * q.G * q.ih2;
  for (int d = 0; d < 3; d++) {
    g.dx[d] = g_dx[d] + 2.0 * dx[d] * g_r2 + g_vrab * vij[d];
    g.vij[d] = g_vij[d] + g_vrab * dx[d] + 2.0 * vij[d] * g_dotvv;
  }
  g.c_s = g.c_n = 0.5 * g_cab;
}

struct RhsAdj {
  struct Args {
    const double *m, *rho, *p, *c, *T, *cp, *u0, *u1, *u2;   // u: sv (LANS) or v
    const double *a0, *a1, *a2, *ae;                          // per-target seeds
    RhsAdjConst k;
    LamPtrs l;
  };
  struct Tile {
    double m[kWave], im[kWave], rho[kWave], V[kWave], p[kWave], c[kWave], T[kWave], cp[kWave], u[3][kWave];
    double a[3][kWave], ae[kWave], etab[kETab];
  };
  struct Row { RhsSide me; double a[3], ae, g[3], gu[3], gm, grho, gp, gc, gT, gcp; };
  __device__ static void init(Tile &s, int lane) { etab_fill(s.etab, lane); }
  __device__ static void load_row(const Args &A, uint32_t i, Row &t) {
    const double m = A.m[i], rho = A.rho[i];
    t.me = RhsSide{m, 1.0 / m, rho, 1.0 / rho, A.p[i], A.c[i], A.T[i], A.cp[i], {A.u0[i], A.u1[i], A.u2[i]}};
    t.a[0] = A.a0[i]; t.a[1] = A.a1[i]; t.a[2] = A.a2[i]; t.ae = A.ae[i];
    for (int k = 0; k < 3; k++) t.g[k] = t.gu[k] = 0;
    t.gm = t.grho = t.gp = t.gc = t.gT = t.gcp = 0;
  }
  __device__ static void stage(const Args &A, uint32_t j, Tile &s, int l) {
    const double m = A.m[j], rho = A.rho[j];
    s.m[l] = m; s.im[l] = 1.0 / m; s.rho[l] = rho; s.V[l] = 1.0 / rho;
    s.p[l] = A.p[j]; s.c[l] = A.c[j]; s.T[l] = A.T[j]; s.cp[l] = A.cp[j];
    s.u[0][l] = A.u0[j]; s.u[1][l] = A.u1[j]; s.u[2][l] = A.u2[j];
    s.a[0][l] = A.a0[j]; s.a[1][l] = A.a1[j]; s.a[2][l] = A.a2[j]; s.ae[l] = A.ae[j];
  }
  __device__ static RhsSide side(const Tile &s, uint32_t q) {
    return RhsSide{s.m[q], s.im[q], s.rho[q], s.V[q], s.p[q], s.c[q], s.T[q], s.cp[q], {s.u[0][q], s.u[1][q], s.u[2][q]}};
  }
  __device__ static double hker_own(double hi, double hj) { return 0.5 * (hi + hj); }
  __device__ static double hker_nbr(double hi, double hj) { return 0.5 * (hj + hi); }
  __device__ static bool own_ok(const Row &, const Tile &, uint32_t, uint32_t j, uint32_t i) { return j != i; }
  __device__ static bool nbr_ok(const Row &, const Tile &, uint32_t, uint32_t j, uint32_t i) { return j != i; }
  // dx = x_i - x_j (row minus candidate); the neighbour role sees -dx, -v_ij
  __device__ static void pair(const Args &A, Row &t, const Tile &s, uint32_t q, const double *dx, double,
                              double hi, double hj, bool fo, bool fn) {
    const RhsSide J = side(s, q);
    const double vij[3] = {t.me.vel[0] - J.vel[0], t.me.vel[1] - J.vel[1], t.me.vel[2] - J.vel[2]};
    const RhsGeo G = rhs_geo(A.k, s.etab, dx, vij, 0.5 * (hi + hj), t.me, J);
    RhsGrad g;
    if (fo) {
      rhs_role(A.k, G, dx, vij, t.me, J, t.a, t.ae, g);
      for (int k = 0; k < 3; k++) { t.g[k] += g.dx[k]; t.gu[k] += g.vij[k]; }
      t.grho += g.rho_s; t.gp += g.p_s; t.gc += g.c_s; t.gT += g.T_s; t.gcp += g.cp_s;
    }
    if (fn) {
      const double dn[3] = {-dx[0], -dx[1], -dx[2]}, vn[3] = {-vij[0], -vij[1], -vij[2]};
      const double a[3] = {s.a[0][q], s.a[1][q], s.a[2][q]};
      rhs_role(A.k, G, dn, vn, J, t.me, a, s.ae[q], g);
      for (int k = 0; k < 3; k++) { t.g[k] -= g.dx[k]; t.gu[k] -= g.vij[k]; }
      t.gm += g.m_n; t.grho += g.rho_n; t.gp += g.p_n; t.gc += g.c_n; t.gT += g.T_n; t.gcp += g.cp_n;
    }
  }
  __device__ static void combine(Row &t) {
    for (int k = 0; k < 3; k++) { t.g[k] = sub_sum<kSub>(t.g[k]); t.gu[k] = sub_sum<kSub>(t.gu[k]); }
    t.gm = sub_sum<kSub>(t.gm); t.grho = sub_sum<kSub>(t.grho); t.gp = sub_sum<kSub>(t.gp);
    t.gc = sub_sum<kSub>(t.gc); t.gT = sub_sum<kSub>(t.gT); t.gcp = sub_sum<kSub>(t.gcp);
  }
  __device__ static void finish(const Args &A, Row &t, uint32_t i) {
    const int u0 = A.k.lans ? L_SVX : L_VX;
    for (int k = 0; k < 3; k++) { A.l.f[L_X + k][i] += t.g[k]; A.l.f[u0 + k][i] += t.gu[k]; }
    A.l.f[L_M][i] += t.gm; A.l.f[L_RHO][i] += t.grho; A.l.f[L_P][i] += t.gp;
    A.l.f[L_C][i] += t.gc; A.l.f[L_T][i] += t.gT; A.l.f[L_CP][i] += t.gcp;
  }
};

// Adjoint rows: real particles, erupt ghosts, and the boundary ghosts (bc 1,
// 2: constant in time, but their velocity and state enter their neighbours'
// sums) that some target's sum met in the recorded step (gflag).
__global__ void k_adj_row_flags(uint32_t n, const int32_t *bc, const uint32_t *gflag, uint32_t *flags) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int b = i < n ? bc[i] : -1;
  flags[i] = (is_real(b) || b == 0 || ((b == 1 || b == 2) && gflag[i])) ? 1u : 0u;
}

// fixed-order block sums of one array
__global__ void k_sum_blocks(uint32_t n, const double *v, double *part) {
  __shared__ double sh[256];
  double acc = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += v[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace

// ===========================================================================
// Checkpoints, step records and the adjoint context
// ===========================================================================
#define PSPH_NDF 19
#define PSPH_NUF 3
#define PSPH_NIF 3

struct Snapshot {
  size_t n = 0, cap = 0;
  double *d[PSPH_NDF] = {};
  uint32_t *u[PSPH_NUF] = {};
  int32_t *i[PSPH_NIF] = {};
  uint32_t *pid = nullptr;
  double time = 0, next_output = 0;
  int64_t step = 0, tk = 0;
  int ioutput = 0;
  uint32_t next_pid = 0;
  unsigned long long dead = 0;
};

struct Adjoint {
  std::vector<Snapshot> ck;
  Snapshot fin;
  std::vector<StepRec> recs;
  size_t lam_cap = 0, l_cap = 0;
  double *L[L_N] = {};        // by particle id
  double *l[L_N] = {};        // cell order of the step being reversed
  double *aux[8] = {};        // per-particle seeds of one pass
  double *part = nullptr;     // reduction partials
  // chunks of adjoint rows
  uint32_t *flags = nullptr, *tpos = nullptr, *rows = nullptr, *cchunks = nullptr, *coff = nullptr;
  uint32_t *cfirst = nullptr, *ccount = nullptr, *nchunks = nullptr;
  uint32_t *tcell = nullptr;
  size_t n_cap = 0, rows_cap = 0, chunk_cap = 0, tcell_cap = 0;
  // visit lists of the step being reversed
  uint32_t *vcount = nullptr, *voff = nullptr, *vnum = nullptr;
  uint2 *vent = nullptr;
  uint64_t *vmask = nullptr;
  size_t vchunk_cap = 0, vis_cap = 0;
};

static void dev_fields(psph_ctx *c, double **d, uint32_t **u, int32_t **i) {
  int k = 0;
#define D_(a) d[k++] = c->f.a;
  PSPH_DFIELDS(D_)
#undef D_
  k = 0;
#define U_(a) u[k++] = c->f.a;
  PSPH_UFIELDS(U_)
#undef U_
  k = 0;
#define I_(a) i[k++] = c->f.a;
  PSPH_IFIELDS(I_)
#undef I_
}

static void snap_free(Snapshot &s) {
  for (auto &p : s.d) { if (p) hipFree(p); p = nullptr; }
  for (auto &p : s.u) { if (p) hipFree(p); p = nullptr; }
  for (auto &p : s.i) { if (p) hipFree(p); p = nullptr; }
  if (s.pid) hipFree(s.pid);
  s.pid = nullptr;
  s.cap = s.n = 0;
}

static void rec_free(StepRec &r) {
  for (auto &p : r.d) { if (p) hipFree(p); p = nullptr; }
  void *ptrs[] = {r.pid, r.refb, r.ckey, r.gflag, r.bc, r.phase, r.bcn, r.cs, r.toff, r.tbox};
  for (auto p : ptrs) if (p) hipFree(p);
  r.pid = r.refb = r.ckey = r.gflag = nullptr;
  r.bc = r.phase = r.bcn = nullptr;
  r.cs = r.toff = nullptr;
  r.tbox = nullptr;
  r.cap = r.cs_cap = r.tile_cap = r.n = 0;
}

static void adjoint_free(psph_ctx *c) {
  Adjoint *A = c->adj;
  if (!A) return;
  for (auto &s : A->ck) snap_free(s);
  snap_free(A->fin);
  for (auto &r : A->recs) rec_free(r);
  for (auto p : A->L) if (p) hipFree(p);
  for (auto p : A->l) if (p) hipFree(p);
  for (auto p : A->aux) if (p) hipFree(p);
  void *ptrs[] = {A->part, A->flags, A->tpos, A->rows, A->cchunks, A->coff, A->cfirst, A->ccount,
                  A->nchunks, A->tcell, A->vcount, A->voff, A->vnum, A->vent, A->vmask};
  for (auto p : ptrs) if (p) hipFree(p);
  delete A;
  c->adj = nullptr;
}

static int snap_save(psph_ctx *c, Snapshot &s) {
  const size_t n = c->n;
  if (n > s.cap) {
    snap_free(s);
    const size_t cap = std::max<size_t>(n, 64);
    for (auto &p : s.d) PSPH_TRY(dmalloc((void **)&p, cap * 8));
    for (auto &p : s.u) PSPH_TRY(dmalloc((void **)&p, cap * 4));
    for (auto &p : s.i) PSPH_TRY(dmalloc((void **)&p, cap * 4));
    PSPH_TRY(dmalloc((void **)&s.pid, cap * 4));
    s.cap = cap;
  }
  double *d[PSPH_NDF]; uint32_t *u[PSPH_NUF]; int32_t *iv[PSPH_NIF];
  dev_fields(c, d, u, iv);
  hipStream_t st = c->stream;
  if (n) {
    for (int k = 0; k < PSPH_NDF; k++) HIP_TRY(hipMemcpyAsync(s.d[k], d[k], n * 8, hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < PSPH_NUF; k++) HIP_TRY(hipMemcpyAsync(s.u[k], u[k], n * 4, hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < PSPH_NIF; k++) HIP_TRY(hipMemcpyAsync(s.i[k], iv[k], n * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(s.pid, c->pid, n * 4, hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(&s.dead, c->d_u64 + 3, 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  s.n = n;
  s.time = c->time; s.next_output = c->next_output; s.step = c->step; s.ioutput = c->ioutput;
  s.next_pid = c->next_pid; s.tk = c->tape.k;
  return 0;
}

static int snap_load(psph_ctx *c, const Snapshot &s) {
  double *d[PSPH_NDF]; uint32_t *u[PSPH_NUF]; int32_t *iv[PSPH_NIF];
  dev_fields(c, d, u, iv);
  hipStream_t st = c->stream;
  const size_t n = s.n;
  if (n) {
    for (int k = 0; k < PSPH_NDF; k++) HIP_TRY(hipMemcpyAsync(d[k], s.d[k], n * 8, hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < PSPH_NUF; k++) HIP_TRY(hipMemcpyAsync(u[k], s.u[k], n * 4, hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < PSPH_NIF; k++) HIP_TRY(hipMemcpyAsync(iv[k], s.i[k], n * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->pid, s.pid, n * 4, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, st, c->d_u64 + 3, (unsigned long long)s.dead, 1);
  HIP_TRY(hipStreamSynchronize(st));
  c->n = n;
  c->time = s.time; c->next_output = s.next_output; c->step = s.step; c->ioutput = s.ioutput;
  c->next_pid = s.next_pid; c->tape.k = s.tk;
  c->grid_valid = false;
  return 0;
}

// stage 0: state after sort and h update; 1: after RHS; 2: after density;
// 3: after the velocity filter
static int rec_capture(psph_ctx *c, int stage, double dt) {
  StepRec &r = *c->rec;
  hipStream_t st = c->stream;
  const size_t n = c->n;
  auto cp = [&](int f, const double *src) -> int {
    if (n) HIP_TRY(hipMemcpyAsync(r.d[f], src, n * 8, hipMemcpyDeviceToDevice, st));
    return 0;
  };
  if (stage == 0) {
    if (n > r.cap) {
      const size_t cap = std::max(c->cap, n);
      for (auto &p : r.d) { if (p) hipFree(p); p = nullptr; }
      uint32_t **u32s[] = {&r.pid, &r.refb, &r.ckey, &r.gflag};
      for (auto pp : u32s) { if (*pp) hipFree(*pp); *pp = nullptr; }
      if (r.bc) hipFree(r.bc);
      if (r.phase) hipFree(r.phase);
      if (r.bcn) hipFree(r.bcn);
      r.bc = r.phase = r.bcn = nullptr;
      const bool wall = c->cfg.wall_bc && c->cfg.dim == 3;
      for (int f = 0; f < (wall ? RF_N : RF_NBASE); f++) PSPH_TRY(dmalloc((void **)&r.d[f], cap * 8));
      for (auto pp : u32s) PSPH_TRY(dmalloc((void **)pp, cap * 4));
      PSPH_TRY(dmalloc((void **)&r.bc, cap * 4));
      PSPH_TRY(dmalloc((void **)&r.phase, cap * 4));
      if (wall) PSPH_TRY(dmalloc((void **)&r.bcn, cap * 4));
      r.cap = cap;
    }
    const size_t ncs = (size_t)c->grid.ncells + 1;
    if (ncs > r.cs_cap) {
      if (r.cs) hipFree(r.cs);
      if (r.toff) hipFree(r.toff);
      r.cs = r.toff = nullptr;
      PSPH_TRY(dmalloc((void **)&r.cs, ncs * 4));
      PSPH_TRY(dmalloc((void **)&r.toff, ncs * 4));
      r.cs_cap = ncs;
    }
    const size_t ntb = n / kWave + (size_t)c->grid.ncells + 1;
    if (ntb > r.tile_cap) {
      if (r.tbox) hipFree(r.tbox);
      r.tbox = nullptr;
      PSPH_TRY(dmalloc((void **)&r.tbox, ntb * 2 * sizeof(float4)));
      r.tile_cap = ntb;
    }
    r.n = n;
    r.grid = c->grid;
    r.reach = c->reach_main;
    r.hmax = c->h_max;
    r.dt = dt;
    r.time = c->time;
    const double *src[] = {c->f.x, c->f.y, c->f.z, c->f.h, c->f.m, c->f.rho, c->f.p, c->f.c, c->f.T, c->f.cp,
                           c->f.svx, c->f.svy, c->f.svz, c->f.vx, c->f.vy, c->f.vz, c->f.e, c->h_list};
    for (int f = RF_X; f <= RF_HLIST; f++) PSPH_TRY(cp(f, src[f]));
    if (n) {
      HIP_TRY(hipMemcpyAsync(r.pid, c->pid, n * 4, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(r.refb, c->refb, n * 4, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(r.ckey, c->key_a, n * 4, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(r.bc, c->f.bc, n * 4, hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(r.phase, c->f.phase, n * 4, hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_fill_i32, blocks(n), 256, 0, st, (int *)r.gflag, 0, (int)n);
    }
    HIP_TRY(hipMemcpyAsync(r.cs, c->cs, ncs * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(r.toff, c->toff, ncs * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(r.tbox, c->tbox, ntb * 2 * sizeof(float4), hipMemcpyDeviceToDevice, st));
  } else if (stage == 1) {
    PSPH_TRY(cp(RF_V1X, c->f.vx)); PSPH_TRY(cp(RF_V1Y, c->f.vy)); PSPH_TRY(cp(RF_V1Z, c->f.vz));
    PSPH_TRY(cp(RF_E1, c->f.e));
  } else if (stage == 2) {
    PSPH_TRY(cp(RF_RHO2, c->f.rho)); PSPH_TRY(cp(RF_E2, c->f.e)); PSPH_TRY(cp(RF_XI2, c->f.xi));
  } else if (stage == 3) {
    PSPH_TRY(cp(RF_SV2X, c->f.svx)); PSPH_TRY(cp(RF_SV2Y, c->f.svy)); PSPH_TRY(cp(RF_SV2Z, c->f.svz));
    r.erupting = iferupt(c);
  } else {
    // 4: before the wall-ghost refresh (after update_pos and the inlet, whose
    // particles come after the first r.n); 5: after it
    if (!r.d[RF_BW] || !r.bcn) return fail(PSPH_EINVARIANT, "record: no wall-ghost fields");
    const size_t m = r.n;
    auto cpm = [&](int f, const double *src) -> int {
      if (m) HIP_TRY(hipMemcpyAsync(r.d[f], src, m * 8, hipMemcpyDeviceToDevice, st));
      return 0;
    };
    if (stage == 4) {
      PSPH_TRY(cpm(RF_XN, c->f.x)); PSPH_TRY(cpm(RF_YN, c->f.y)); PSPH_TRY(cpm(RF_ZN, c->f.z));
      if (m) {
        HIP_TRY(hipMemcpyAsync(r.bcn, c->f.bc, m * 4, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_fill_f64, blocks(m), 256, 0, st, r.d[RF_BW], -1.0, m);
      }
    } else {
      PSPH_TRY(cpm(RF_BRHO, c->f.rho)); PSPH_TRY(cpm(RF_BE, c->f.e)); PSPH_TRY(cpm(RF_BXI, c->f.xi));
      PSPH_TRY(cpm(RF_BVX, c->f.vx)); PSPH_TRY(cpm(RF_BVY, c->f.vy)); PSPH_TRY(cpm(RF_BVZ, c->f.vz));
    }
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

static int adj_arrays(psph_ctx *c, Adjoint &A, size_t nid) {
  if (nid > A.lam_cap) {
    for (auto &p : A.L) { if (p) hipFree(p); p = nullptr; }
    const size_t cap = std::max<size_t>(nid, 64);
    for (auto &p : A.L) PSPH_TRY(dmalloc((void **)&p, cap * 8));
    A.lam_cap = cap;
  }
  if (c->cap > A.l_cap) {
    for (auto &p : A.l) { if (p) hipFree(p); p = nullptr; }
    for (auto &p : A.aux) { if (p) hipFree(p); p = nullptr; }
    for (auto &p : A.l) PSPH_TRY(dmalloc((void **)&p, c->cap * 8));
    for (auto &p : A.aux) PSPH_TRY(dmalloc((void **)&p, c->cap * 8));
    A.l_cap = c->cap;
  }
  if (!A.part) PSPH_TRY(dmalloc((void **)&A.part, kRedBlocks * 8));
  return 0;
}

static LamPtrs lam_ptrs(double *const *p) {
  LamPtrs r;
  for (int k = 0; k < L_N; k++) r.f[k] = p[k];
  return r;
}

// deterministic sum of v[0..n)
static int dev_sum(psph_ctx *c, Adjoint &A, const double *v, size_t n, double *out) {
  *out = 0;
  if (!n) return 0;
  hipLaunchKernelGGL(k_sum_blocks, kRedBlocks, 256, 0, c->stream, (uint32_t)n, v, A.part);
  std::vector<double> part(kRedBlocks);
  HIP_TRY(hipMemcpyAsync(part.data(), A.part, kRedBlocks * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  double s = 0;
  for (double x : part) s += x;
  *out = s;
  return 0;
}

// chunks of adjoint rows of one step (<= 64 consecutive rows of one cell)
static int adj_rows(psph_ctx *c, Adjoint &A, const StepRec &R) {
  const uint32_t n = (uint32_t)R.n;
  const GridDesc &g = R.grid;
  const uint32_t ncells = g.ncells;
  hipStream_t s = c->stream;
  if (n + 1 > A.n_cap) {
    void *ptrs[] = {A.flags, A.tpos, A.rows};
    for (auto p : ptrs) if (p) hipFree(p);
    const size_t cap = std::max<size_t>(c->cap, n) + 1;
    PSPH_TRY(dmalloc((void **)&A.flags, cap * 4));
    PSPH_TRY(dmalloc((void **)&A.tpos, cap * 4));
    PSPH_TRY(dmalloc((void **)&A.rows, cap * 4));
    A.n_cap = cap;
  }
  if (ncells + 1 > A.tcell_cap) {
    void *ptrs[] = {A.cchunks, A.coff, A.tcell};
    for (auto p : ptrs) if (p) hipFree(p);
    PSPH_TRY(dmalloc((void **)&A.cchunks, ((size_t)ncells + 1) * 4));
    PSPH_TRY(dmalloc((void **)&A.coff, ((size_t)ncells + 1) * 4));
    PSPH_TRY(dmalloc((void **)&A.tcell, ((size_t)ncells + 1) * 4));
    A.tcell_cap = (size_t)ncells + 1;
  }
  const size_t chunks_ub = n / kChunk + (size_t)ncells + 1;
  if (chunks_ub > A.chunk_cap) {
    void *ptrs[] = {A.cfirst, A.ccount};
    for (auto p : ptrs) if (p) hipFree(p);
    PSPH_TRY(dmalloc((void **)&A.cfirst, chunks_ub * 4));
    PSPH_TRY(dmalloc((void **)&A.ccount, chunks_ub * 4));
    A.chunk_cap = chunks_ub;
  }
  if (!A.nchunks) PSPH_TRY(dmalloc((void **)&A.nchunks, 4));
  hipLaunchKernelGGL(k_adj_row_flags, blocks(n + 1), 256, 0, s, n, R.bc, R.gflag, A.flags);
  PSPH_TRY(excl_scan(c, A.flags, A.tpos, (size_t)n + 1));
  if (n) hipLaunchKernelGGL(k_targets, blocks(n), 256, 0, s, n, A.flags, A.tpos, A.rows);
  hipLaunchKernelGGL(k_cell_chunks, blocks(ncells + 1), 256, 0, s, ncells, R.cs, A.tpos, (uint32_t)kChunk, A.cchunks);
  PSPH_TRY(excl_scan(c, A.cchunks, A.coff, (size_t)ncells + 1));
  hipLaunchKernelGGL(k_chunks, blocks(ncells + 1), 256, 0, s, ncells, R.cs, A.tpos, A.coff,
                     (uint32_t)kChunk, A.cfirst, A.ccount, A.nchunks);
  HIP_TRY(hipGetLastError());
  return 0;
}

// visit lists of the step's row chunks (count, scan, masks)
static int adj_visits(psph_ctx *c, Adjoint &A, const StepRec &R, const AdjCommon &pc, AdjVisits &vl) {
  const unsigned G = chunk_grid(R.n, R.grid.ncells, kChunk);
  if ((size_t)G + 1 > A.vchunk_cap) {
    uint32_t **arrs[] = {&A.vcount, &A.voff, &A.vnum};
    for (auto pp : arrs) { if (*pp) hipFree(*pp); *pp = nullptr; }
    for (auto pp : arrs) PSPH_TRY(dmalloc((void **)pp, ((size_t)G + 1) * 4));
    A.vchunk_cap = (size_t)G + 1;
  }
  hipLaunchKernelGGL(k_adj_visit_count, G, kWave, 0, c->stream, pc, A.vcount);
  hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, c->stream, (int *)A.vcount + G, 0, 1);
  HIP_TRY(hipGetLastError());
  PSPH_TRY(excl_scan(c, A.vcount, A.voff, (size_t)G + 1));
  uint32_t total = 0;
  HIP_TRY(hipMemcpyAsync(&total, A.voff + G, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if ((size_t)total > A.vis_cap) {
    if (A.vmask) hipFree(A.vmask);
    if (A.vent) hipFree(A.vent);
    A.vmask = nullptr;
    A.vent = nullptr;
    const size_t cap = (size_t)total + total / 8 + 1024;
    PSPH_TRY(dmalloc((void **)&A.vmask, cap * kWave * sizeof(uint64_t)));
    PSPH_TRY(dmalloc((void **)&A.vent, cap * sizeof(uint2)));
    A.vis_cap = cap;
  }
  hipLaunchKernelGGL(k_adj_visit_masks, G, kWave, 0, c->stream, pc, c->cfg.cutoff * (1.0 + 1e-12), A.voff, A.vnum,
                     A.vent, A.vmask);
  HIP_TRY(hipGetLastError());
  vl = AdjVisits{A.voff, A.vnum, A.vent, A.vmask};
  return 0;
}

template <class P>
static void launch_adj(psph_ctx *c, const AdjCommon &pc, const AdjVisits &vl, const StepRec &R,
                       const typename P::Args &a) {
  hipLaunchKernelGGL((k_adj_pairs<P>), chunk_grid(R.n, R.grid.ncells, kChunk), kWave, 0, c->stream, pc, vl, a);
}

struct ThetaGrad { double v = 0, T = 0, ng0 = 0; };

// vent state derivatives (parameters.h:958-975): rho_v = p / (Rg T) / ng0,
// e_v = Cvm(1) T, m_v = rho_v V / np, T = e_v / Cvm(1) = T_v, Cp = Cvm(1)
// Wall-ghost refresh reverse (bcond.cc:24-249, as k_bcond computes it).  A
// refreshed ghost i holds u_k = N_k / Wn: N_rho = sum w_j, N_v = sum w_j v_j /
// rho_j, N_e = sum w_j e_j / rho_j, N_xi = the same sum of xi_j over the
// neighbour buckets only, Wn = sum w_j / rho_j, with w_j = m_j W(r_i - x_j,
// h_j) at its image r_i; then rho = u_rho, v = R u_v (R the mirror), e = u_e,
// xi = u_xi and the EOS.  Per ghost, Lambda_k (the adjoint of u_k, the EOS
// folded in) gives a_k = Lambda_k / Wn and b = -sum_k Lambda_k u_k / Wn (the
// adjoint of Wn); per contributor j, lambda_w = a_rho + (a_v.v_j + a_e e_j +
// [nbr] a_xi xi_j + b) / rho_j, and v_j, e_j, xi_j, rho_j, m_j and x_j
// (through W) take their shares.  The refreshed fields are overwritten, so
// the adjoints of the ghost's values before the refresh restart at zero.
// Contributors gather their ghosts (no atomics): a contributor's candidates
// are the wall ghosts around its mirror point, found in the step's cell grid
// (wall ghosts do not move).  Pressure ghosts contribute as well: their
// state is never updated, so the adjoint of their initial state gathers the
// refresh's terms of every step.
enum { BA_ARHO, BA_AVX, BA_AVY, BA_AVZ, BA_AE, BA_AXI, BA_B, BA_GNG0, BA_N };
struct BcondAdj { double *f[BA_N]; };

__global__ void k_adj_bcond_ghost(uint32_t n, BcondArgs g, const double *gx, const double *gy, const double *gz,
                                  const double *wrec, const double *brho, const double *bvx, const double *bvy,
                                  const double *bvz, const double *be, const double *bxi, LamPtrs l, BcondAdj o) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  o.f[BA_GNG0][i] = 0;
  const double wn = wrec[i];
  if (!(wn >= 0)) return;                   // not refreshed: the adjoints pass through
  const double lr = l.f[L_RHO][i], le = l.f[L_E][i], lxi = l.f[L_XI][i];
  const double lv[3] = {l.f[L_VX][i], l.f[L_VY][i], l.f[L_VZ][i]};
  const double lp = l.f[L_P][i], lc = l.f[L_C][i], lT = l.f[L_T][i], lcp = l.f[L_CP][i];
  l.f[L_RHO][i] = l.f[L_E][i] = l.f[L_XI][i] = 0;
  l.f[L_VX][i] = l.f[L_VY][i] = l.f[L_VZ][i] = 0;
  l.f[L_P][i] = l.f[L_C][i] = l.f[L_T][i] = 0;
  if (g.eos.lans) l.f[L_CP][i] = 0;          // Cp is an EOS output under LANS only
  if (!(wn > 0)) return;                    // no contributor: a constant state
  const double coord[3] = {gx[i], gy[i], gz[i]};
  const double refc[3] = {coord[0], coord[1], 2 * g.zg - coord[2]};
  double nrm[3];
  mirror_normal(coord, refc, nrm);
  const double v[3] = {bvx[i], bvy[i], bvz[i]};
  const double vn = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
  const double ln = lv[0] * nrm[0] + lv[1] * nrm[1] + lv[2] * nrm[2];
  double u[3], lu[3];
  for (int k = 0; k < 3; k++) { u[k] = v[k] - 2 * vn * nrm[k]; lu[k] = lv[k] - 2 * ln * nrm[k]; }
  const double rho = brho[i], e = be[i], xi = bxi[i];
  double grho, ge, gxi, gn;
  eos_adj(g.eos, rho, e, xi, lp, lc, lT, lcp, grho, ge, gxi, gn);
  const double Lr = lr + grho, Le = le + ge, Lxi = lxi + gxi;
  const double iw = 1.0 / wn;
  o.f[BA_ARHO][i] = Lr * iw;
  o.f[BA_AVX][i] = lu[0] * iw; o.f[BA_AVY][i] = lu[1] * iw; o.f[BA_AVZ][i] = lu[2] * iw;
  o.f[BA_AE][i] = Le * iw;
  o.f[BA_AXI][i] = Lxi * iw;
  o.f[BA_B][i] = -(Lr * rho + lu[0] * u[0] + lu[1] * u[1] + lu[2] * u[2] + Le * e + Lxi * xi) * iw;
  o.f[BA_GNG0][i] = gn;
}

struct BcondAdjIn {
  const double *xn, *yn, *zn;       // positions at the refresh (after update_pos)
  const int32_t *bcn, *bc0;         // bc at the refresh and at the step's search
  const double *gx, *gy, *gz;       // positions at the search (wall ghosts do not move)
  const double *h, *m, *rho, *vx, *vy, *vz, *e, *xi;   // the values the refresh read
  const double *wrec;
  const uint32_t *cs;
  GridDesc grid;
  double h_new;                      // h of an erupt ghost that turned real in update_pos
  double rfac;                       // per-axis reach over h (both support tests, with margin)
};

__global__ void k_adj_bcond_contr(uint32_t n, BcondArgs g, BcondAdjIn in, BcondAdj cf, LamPtrs l) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = in.bcn[j];
  if (b != 100 && b != 99 && b != 1) return;        // contr_image (particle.h:562-565)
  const double pj[3] = {in.xn[j], in.yn[j], in.zn[j]};
  int bj[3];
  if (!bucket_of(g, pj, bj)) return;
  // images lie in the ground bucket layer or the one above it
  if (bj[2] > g.bzg + 2) return;
  const double hj = in.bc0[j] == 0 ? in.h_new : in.h[j];
  const double mj[3] = {pj[0], pj[1], 2 * g.zg - pj[2]};
  const double reach = in.rfac * hj;
  const int nn[3] = {in.grid.nx, in.grid.ny, in.grid.nz};
  int c0[3], c1[3];
  for (int k = 0; k < 3; k++) {
    if (in.grid.inv_e > 0) {
      c0[k] = cell_of(mj[k] - reach, in.grid.lo[k], in.grid.inv_e, nn[k]);
      c1[k] = cell_of(mj[k] + reach, in.grid.lo[k], in.grid.inv_e, nn[k]);
    } else {
      c0[k] = c1[k] = 0;
    }
  }
  const double ir = 1.0 / in.rho[j], mj_ = in.m[j];
  const double vj[3] = {in.vx[j], in.vy[j], in.vz[j]}, ej = in.e[j], xij = in.xi[j];
  const double ih = 1.0 / hj, nk = kInvSqrtPi * ih, nk3 = nk * nk * nk;
  double lvx = 0, lvy = 0, lvz = 0, le = 0, lxi = 0, lrho = 0, lm = 0, lx[3] = {0, 0, 0};
  for (int cz = c0[2]; cz <= c1[2]; cz++)
    for (int cy = c0[1]; cy <= c1[1]; cy++)
      for (int cx = c0[0]; cx <= c1[0]; cx++) {
        const uint32_t cell = ((uint32_t)cz * in.grid.ny + cy) * in.grid.nx + cx;
        for (uint32_t q = in.cs[cell], qe = in.cs[cell + 1]; q < qe; q++) {
          if (in.bc0[q] != 2 || !(in.wrec[q] > 0)) continue;
          const double refc[3] = {in.gx[q], in.gy[q], 2 * g.zg - in.gz[q]};
          int bi[3];
          if (!bucket_of(g, refc, bi)) continue;
          if (abs(bj[0] - bi[0]) > 1 || abs(bj[1] - bi[1]) > 1 || abs(bj[2] - bi[2]) > 1) continue;
          const double amax = fmax(fabs(refc[0] - pj[0]), fmax(fabs(refc[1] - pj[1]), fabs(refc[2] - pj[2])));
          if (!(amax <= g.cutoff2 * hj) || g.cut.outside(amax, hj)) continue;
          const double dx[3] = {refc[0] - pj[0], refc[1] - pj[1], refc[2] - pj[2]};
          double expt = 0;
          for (int k = 0; k < 3; k++) { const double sk = dx[k] * ih; expt += sk * sk; }
          const double w = nk3 * exp(-expt);
          const double mw = mj_ * w, mwr = mw * ir;
          const bool nbr = bj[0] != bi[0] || bj[1] != bi[1] || bj[2] != bi[2];
          const double av[3] = {cf.f[BA_AVX][q], cf.f[BA_AVY][q], cf.f[BA_AVZ][q]};
          const double ae = cf.f[BA_AE][q], axi = nbr ? cf.f[BA_AXI][q] : 0.0;
          const double s = av[0] * vj[0] + av[1] * vj[1] + av[2] * vj[2] + ae * ej + axi * xij + cf.f[BA_B][q];
          const double lw = cf.f[BA_ARHO][q] + s * ir;
          lvx += av[0] * mwr; lvy += av[1] * mwr; lvz += av[2] * mwr;
          le += ae * mwr;
          lxi += axi * mwr;
          lrho -= s * mwr * ir;
          lm += lw * w;
          const double gw = lw * mw * 2 * ih * ih;          // dW/dx_j = W 2 (r_i - x_j) / h^2
          for (int k = 0; k < 3; k++) lx[k] += gw * dx[k];
        }
      }
  l.f[L_VX][j] += lvx; l.f[L_VY][j] += lvy; l.f[L_VZ][j] += lvz;
  l.f[L_E][j] += le; l.f[L_XI][j] += lxi; l.f[L_RHO][j] += lrho; l.f[L_M][j] += lm;
  l.f[L_X][j] += lx[0]; l.f[L_Y][j] += lx[1]; l.f[L_Z][j] += lx[2];
}

static int adj_bcond(psph_ctx *c, Adjoint &A, const StepRec &R, ThetaGrad &th) {
  const psph_config &k = c->cfg;
  hipStream_t st = c->stream;
  const uint32_t n = (uint32_t)R.n;
  const BcondArgs g = bcond_geom(k);
  BcondAdj cf;
  for (int q = 0; q < BA_N; q++) cf.f[q] = A.aux[q];
  const LamPtrs l = lam_ptrs(A.l);
  hipLaunchKernelGGL(k_adj_bcond_ghost, blocks(n), 256, 0, st, n, g, R.d[RF_X], R.d[RF_Y], R.d[RF_Z], R.d[RF_BW],
                     R.d[RF_BRHO], R.d[RF_BVX], R.d[RF_BVY], R.d[RF_BVZ], R.d[RF_BE], R.d[RF_BXI], l, cf);
  double gn = 0;
  PSPH_TRY(dev_sum(c, A, A.aux[BA_GNG0], n, &gn));
  th.ng0 += gn;
  BcondAdjIn in;
  in.xn = R.d[RF_XN]; in.yn = R.d[RF_YN]; in.zn = R.d[RF_ZN];
  in.bcn = R.bcn; in.bc0 = R.bc;
  in.gx = R.d[RF_X]; in.gy = R.d[RF_Y]; in.gz = R.d[RF_Z];
  in.h = R.d[RF_H]; in.m = R.d[RF_M]; in.rho = R.d[RF_RHO2];
  in.vx = R.d[RF_V1X]; in.vy = R.d[RF_V1Y]; in.vz = R.d[RF_V1Z];
  in.e = R.d[RF_E2]; in.xi = R.d[RF_XI2];
  in.wrec = R.d[RF_BW];
  in.cs = R.cs;
  in.grid = R.grid;
  in.h_new = k.ratio_sml_dx * k.h_air;
  in.rfac = std::min(k.cutoff, k.cutoff2) * (1 + 1e-9) + 1e-12;
  hipLaunchKernelGGL(k_adj_bcond_contr, blocks(n), 256, 0, st, n, g, in, cf, l);
  HIP_TRY(hipGetLastError());
  return 0;
}

struct VentDeriv { double drho_dT, drho_dn, de_dT, de_dn, dm_dT, dm_dn, dcp_dn; };
static VentDeriv vent_deriv(const psph_config &k) {
  VentDeriv d;
  d.drho_dT = -k.vent_rho / k.vent_T;
  d.drho_dn = -k.vent_rho / k.ng0;
  d.de_dT = k.ng0 * k.cvg + (1 - k.ng0) * k.cvs;
  d.de_dn = (k.cvg - k.cvs) * k.vent_T;
  d.dm_dT = -k.vent_mass / k.vent_T;
  d.dm_dn = -k.vent_mass / k.ng0;
  d.dcp_dn = k.cvg - k.cvs;
  return d;
}

// theta-gradient of erupt ghosts created with column velocity vel (z of
// injected ghosts moves with vel: dz/dvel = injection time; setup ghosts sit
// at fixed z).  Mass terms are added once at the end.
static void ghost_theta(const psph_config &k, const VentDeriv &d, double vel, double tz, const double *lam,
                        ThetaGrad &g) {
  // lam: z, vz, rho, e, T, cp
  const double dvel = k.vent_v != 0 ? vel / k.vent_v : 0.0;
  g.v += lam[1] * dvel + lam[0] * tz * dvel;
  g.T += lam[2] * d.drho_dT + lam[3] * d.de_dT + lam[4];
  g.ng0 += lam[2] * d.drho_dn + lam[3] * d.de_dn + lam[5] * d.dcp_dn;
}

static int reverse_step(psph_ctx *c, Adjoint &A, StepRec &R, ThetaGrad &th) {
  const psph_config &k = c->cfg;
  hipStream_t st = c->stream;
  const uint32_t n = (uint32_t)R.n;
  const VentDeriv vd = vent_deriv(k);
  // 1. inlet: the ghosts this step injected
  if (R.inj_hi > R.inj_lo) {
    const size_t m = R.inj_hi - R.inj_lo;
    const int fl[6] = {L_Z, L_VZ, L_RHO, L_E, L_T, L_CP};
    std::vector<double> lam(6 * m);
    for (int f = 0; f < 6; f++)
      HIP_TRY(hipMemcpyAsync(lam.data() + f * m, A.L[fl[f]] + R.inj_lo, m * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (R.inj_vel.size() != m) return fail(PSPH_EINVARIANT, "adjoint: inlet record of %zu ghosts, %zu ids", R.inj_vel.size(), m);
    for (size_t q = 0; q < m; q++) {
      double one[6];
      for (int f = 0; f < 6; f++) one[f] = lam[f * m + q];
      ghost_theta(k, vd, R.inj_vel[q], R.time, one, th);
    }
  }
  if (!n) return 0;
  const LamPtrs Lid = lam_ptrs(A.L), l = lam_ptrs(A.l);
  hipLaunchKernelGGL(k_lam_gather, blocks(n), 256, 0, st, n, R.pid, Lid, l);
  // 2. wall-ghost refresh at the end of the step
  if (k.wall_bc && k.dim == 3 && !bcond_trivial(k)) PSPH_TRY(adj_bcond(c, A, R, th));
  // 3. position update
  hipLaunchKernelGGL(k_adj_pos, blocks(n), 256, 0, st, n, R.bc, R.erupting, k.lans, R.dt, l);
  PSPH_TRY(adj_rows(c, A, R));
  AdjCommon pc;
  pc.x = R.d[RF_X]; pc.y = R.d[RF_Y]; pc.z = R.d[RF_Z]; pc.h = R.d[RF_H]; pc.h_list = R.d[RF_HLIST];
  pc.refb = R.refb; pc.rows = A.rows;
  pc.cfirst = A.cfirst; pc.ccount = A.ccount; pc.nchunks = A.nchunks;
  pc.bc = R.bc;
  pc.tr = Traversal{R.cs, R.toff, R.tbox, R.grid, R.reach};
  pc.list_all = k.list_all ? 1 : 0;
  pc.list_adj = list_adj(k, R.hmax);
  pc.list_factor = k.list_factor; pc.cutoff2 = k.cutoff2;
  pc.cut = cut_test(k.cutoff);
  AdjVisits vl;
  PSPH_TRY(adj_visits(c, A, R, pc, vl));
  // 4. velocity filter
  if (k.lans) {
    hipLaunchKernelGGL(k_adj_vsm_pre, blocks(n), 256, 0, st, n, R.bc, R.d[RF_WNV], R.d[RF_SV2X], R.d[RF_SV2Y],
                       R.d[RF_SV2Z], l, A.aux[0], A.aux[1], A.aux[2], A.aux[3]);
    VsmAdj::Args a{R.d[RF_M], R.d[RF_RHO2], R.d[RF_H], R.d[RF_V1X], R.d[RF_V1Y], R.d[RF_V1Z], A.aux[0], A.aux[1],
                   A.aux[2], A.aux[3], R.bc, l};
    launch_adj<VsmAdj>(c, pc, vl, R, a);
  }
  // 5. density + EOS
  {
    hipLaunchKernelGGL(k_adj_dens_pre, blocks(n), 256, 0, st, n, R.bc, R.d[RF_RHO2], R.d[RF_E2], R.d[RF_XI2],
                       R.d[RF_WNE], eos_const(k), k.energy_smooth, l, A.aux[0], A.aux[1], A.aux[2], A.aux[3], A.aux[4]);
    double gn = 0;
    PSPH_TRY(dev_sum(c, A, A.aux[4], n, &gn));
    th.ng0 += gn;
    DensAdj::Args a{R.d[RF_M], R.d[RF_RHO], R.d[RF_E1], R.bc, R.phase, A.aux[0], A.aux[1], A.aux[2], A.aux[3],
                    k.e_smooth_ratio, k.pghost_contrib, k.energy_smooth,
                    k.e_smooth_ratio * k.e_smooth_ratio == 16.0 ? 1 : 0, cut_test(k.cutoff), l};
    launch_adj<DensAdj>(c, pc, vl, R, a);
  }
  // 6. momentum / energy
  {
    hipLaunchKernelGGL(k_adj_rhs_pre, blocks(n), 256, 0, st, n, R.bc, R.dt, -k.g, k.lans, l,
                       A.aux[0], A.aux[1], A.aux[2], A.aux[3]);
    RhsAdj::Args a;
    a.m = R.d[RF_M]; a.rho = R.d[RF_RHO]; a.p = R.d[RF_P]; a.c = R.d[RF_C]; a.T = R.d[RF_T]; a.cp = R.d[RF_CP];
    a.u0 = k.lans ? R.d[RF_SVX] : R.d[RF_VX];
    a.u1 = k.lans ? R.d[RF_SVY] : R.d[RF_VY];
    a.u2 = k.lans ? R.d[RF_SVZ] : R.d[RF_VZ];
    a.a0 = A.aux[0]; a.a1 = A.aux[1]; a.a2 = A.aux[2]; a.ae = A.aux[3];
    a.k = RhsAdjConst{k.alpha, k.beta, k.ata, k.eps_half, k.prandtl, k.lans};
    a.l = l;
    launch_adj<RhsAdj>(c, pc, vl, R, a);
  }
  hipLaunchKernelGGL(k_lam_scatter, blocks(n), 256, 0, st, n, R.pid, l, Lid);
  HIP_TRY(hipGetLastError());
  return 0;
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

extern "C" int psph_adjoint_run(psph_ctx *c, int64_t nsteps, int64_t ckpt_every, psph_adjoint_out *out,
                                double *grad_v, double *grad_sv, size_t n_cap) {
  if (!c || !out) return fail(PSPH_EINVAL, "null argument");
  const psph_config &k = c->cfg;
  if (k.adapt_domain) return fail(PSPH_ENOTSUP, "adjoint: domain adaptation (cfg.adapt_domain) has no reverse; set it to 0");
  if (k.dim != 3) return fail(PSPH_ENOTSUP, "adjoint: built for the 3D plume only");
  if (nsteps < 1 || ckpt_every < 1) return fail(PSPH_EINVAL, "adjoint: nsteps and ckpt_every must be >= 1");
  if (c->tape.mode) return fail(PSPH_EINVAL, "adjoint: a recorded run is already in progress");
  if (!c->adj) c->adj = new Adjoint();
  Adjoint &A = *c->adj;
  std::memset(out, 0, sizeof *out);
  hipStream_t st = c->stream;
  HIP_TRY(hipStreamSynchronize(st));
  const double t_start = now_ms();

  // ---- forward, recording the schedule ----
  Tape &tp = c->tape;
  tape_clear(tp);
  const size_t n0 = c->n;
  if (n0) hipLaunchKernelGGL(k_iota_u32, blocks(n0), 256, 0, st, c->pid, 0u, (uint32_t)n0);
  c->next_pid = (uint32_t)n0;
  PSPH_TRY(tape_ids(c, std::max<size_t>(n0, 1)));
  hipLaunchKernelGGL(k_fill_i32, blocks(tp.id_cap), 256, 0, st, tp.tr_step, -1, (int)tp.id_cap);
  hipLaunchKernelGGL(k_fill_i32, blocks(tp.id_cap), 256, 0, st, tp.esc_step, -1, (int)tp.id_cap);
  tp.t0 = c->time; tp.step0 = c->step; tp.n0 = n0;
  tp.mode = 1;
  const int64_t K = std::min(ckpt_every, nsteps);
  const int64_t nseg = (nsteps + K - 1) / K;
  if ((int64_t)A.ck.size() < nseg) A.ck.resize(nseg);
  if ((int64_t)A.recs.size() < K) A.recs.resize(K);
  // Segments of K steps, the shorter one (if any) first: the last segment is
  // reversed straight from the recorded run, the others are recomputed from
  // their checkpoints, so a full last segment recomputes the fewest steps.
  const int64_t first_len = nsteps - (nseg - 1) * K;
  auto seg_begin = [&](int64_t seg) { return seg == 0 ? (int64_t)0 : first_len + (seg - 1) * K; };
  const int64_t last0 = seg_begin(nseg - 1);
  int rc = 0;
  for (int64_t s = 0, seg = 0; s < nsteps && !rc; s++) {
    if (seg < nseg && s == seg_begin(seg)) rc = snap_save(c, A.ck[seg++]);
    if (rc) break;
    c->rec = s >= last0 ? &A.recs[s - last0] : nullptr;
    rc = psph_step(c, nullptr);
  }
  c->rec = nullptr;
  if (rc) { tp.mode = 0; return rc; }
  double H = 0;
  PSPH_TRY(psph_diag_height(c, &H));
  PSPH_TRY(snap_save(c, A.fin));
  // ids of the whole run (the recompute of a segment restores the id counter
  // of its checkpoint, so it is read here, after the recorded run)
  const uint32_t pid_end = c->next_pid;
  HIP_TRY(hipStreamSynchronize(st));
  const double t_fwd = now_ms();

  // ---- seeds ----
  PSPH_TRY(adj_arrays(c, A, c->next_pid));
  for (int f = 0; f < L_N; f++)
    hipLaunchKernelGGL(k_fill_f64, blocks(A.lam_cap), 256, 0, st, A.L[f], 0.0, A.lam_cap);
  {
    double cons[8];
    uint32_t n = (uint32_t)c->n;
    hipLaunchKernelGGL(k_conservation, kRedBlocks, 256, 0, st, n, c->f.bc, c->f.m, c->f.e, c->f.vx,
                       c->f.vy, c->f.vz, c->f.xi, c->f.z, c->d_part);
    std::vector<double> part(kRedBlocks * 8);
    HIP_TRY(hipMemcpyAsync(part.data(), c->d_part, part.size() * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double D = 0;
    for (int b = 0; b < kRedBlocks; b++) D += part[b * 8 + 6];
    (void)cons;
    if (!(D > 0)) { tp.mode = 0; return fail(PSPH_EINVARIANT, "adjoint: no erupted material among real particles"); }
    if (n) hipLaunchKernelGGL(k_lam_seed, blocks(n), 256, 0, st, n, c->f.bc, c->pid, c->f.m, c->f.xi, c->f.z, H, D, lam_ptrs(A.L));
  }

  // ---- reverse sweep, segment by segment ----
  ThetaGrad th;
  tp.mode = 2;
  for (int64_t seg = nseg - 1; seg >= 0 && !rc; seg--) {
    const int64_t s0 = seg_begin(seg), s1 = seg + 1 < nseg ? seg_begin(seg + 1) : nsteps;
    if (seg != nseg - 1) {
      rc = snap_load(c, A.ck[seg]);
      for (int64_t s = s0; s < s1 && !rc; s++) {
        c->rec = &A.recs[s - s0];
        rc = psph_step(c, nullptr);
      }
      c->rec = nullptr;
    }
    for (int64_t s = s1 - 1; s >= s0 && !rc; s--) rc = reverse_step(c, A, A.recs[s - s0], th);
  }
  c->rec = nullptr;
  if (!rc) rc = check_err(c, "adjoint");
  if (rc) { tp.mode = 0; snap_load(c, A.fin); return rc; }

  // ---- initial state: setup ghosts and the per-particle output ----
  const Snapshot &S0 = A.ck[0];
  std::vector<int32_t> bc0(n0);
  std::vector<double> x0(n0), y0(n0), lam0(6 * std::max<size_t>(n0, 1));
  std::vector<double> lm(pid_end);
  if (n0) {
    HIP_TRY(hipMemcpyAsync(bc0.data(), S0.i[0], n0 * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(x0.data(), S0.d[0], n0 * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(y0.data(), S0.d[1], n0 * 8, hipMemcpyDeviceToHost, st));
    const int fl[6] = {L_Z, L_VZ, L_RHO, L_E, L_T, L_CP};
    for (int f = 0; f < 6; f++) HIP_TRY(hipMemcpyAsync(lam0.data() + f * n0, A.L[fl[f]], n0 * 8, hipMemcpyDeviceToHost, st));
  }
  if (pid_end) HIP_TRY(hipMemcpyAsync(lm.data(), A.L[L_M], (size_t)pid_end * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const VentDeriv vd = vent_deriv(k);
  double lm_erupt = 0;
  for (size_t q = 0; q < n0; q++) {
    if (bc0[q] != 0) continue;
    double one[6];
    for (int f = 0; f < 6; f++) one[f] = lam0[f * n0 + q];
    one[0] = 0;                                   // setup ghosts: z independent of the vent velocity
    // column exit velocity (vel_prof, sph_lib.cc:1223-1243) at the ghost's column centre
    const double rsq = x0[q] * x0[q] + y0[q] * y0[q];
    const double vel = k.vent_profile == 0 ? k.vent_v : 2 * k.vent_v * (1 - rsq / (k.vent_r * k.vent_r));
    ghost_theta(k, vd, vel, 0.0, one, th);
    lm_erupt += lm[q];
  }
  for (size_t q = n0; q < pid_end; q++) lm_erupt += lm[q];   // every ghost the inlet added
  th.T += lm_erupt * vd.dm_dT;
  th.ng0 += lm_erupt * vd.dm_dn;

  if (grad_v || grad_sv) {
    // psph_read order of the state at the start: live particles, THashTable order
    std::vector<uint32_t> k0(n0), k1(n0), k2(n0);
    if (n0) {
      HIP_TRY(hipMemcpy(k0.data(), S0.u[0], n0 * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(k1.data(), S0.u[1], n0 * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(k2.data(), S0.u[2], n0 * 4, hipMemcpyDeviceToHost));
    }
    std::vector<uint32_t> live;
    for (size_t q = 0; q < n0; q++) if (bc0[q] >= 0) live.push_back((uint32_t)q);
    std::vector<uint32_t> a0(live.size()), a1(live.size()), a2(live.size());
    for (size_t q = 0; q < live.size(); q++) { a0[q] = k0[live[q]]; a1[q] = k1[live[q]]; a2[q] = k2[live[q]]; }
    auto perm = canonical_order(k, a0.data(), a1.data(), a2.data(), live.size());
    if (live.size() > n_cap) { tp.mode = 0; snap_load(c, A.fin); return fail(PSPH_ECAPACITY, "adjoint: %zu particles, gradient buffers hold %zu", live.size(), n_cap); }
    for (int which = 0; which < 2; which++) {
      double *dst = which == 0 ? grad_v : grad_sv;
      if (!dst) continue;
      std::vector<double> g3(3 * std::max<size_t>(n0, 1));
      for (int d = 0; d < 3; d++)
        if (n0) HIP_TRY(hipMemcpy(g3.data() + d * n0, A.L[(which == 0 ? L_VX : L_SVX) + d], n0 * 8, hipMemcpyDeviceToHost));
      for (size_t r = 0; r < perm.size(); r++) {
        const uint32_t id = live[perm[r]];
        for (int d = 0; d < 3; d++) dst[3 * r + d] = g3[d * n0 + id];
      }
    }
    out->n_particles = (int64_t)live.size();
  }
  tp.mode = 0;
  PSPH_TRY(snap_load(c, A.fin));
  HIP_TRY(hipStreamSynchronize(st));
  const double t_end = now_ms();
  out->H = H;
  out->dH_dvent_v = th.v;
  out->dH_dvent_T = th.T;
  out->dH_dng0 = th.ng0;
  out->dH_dmass_frac = -th.ng0;
  out->forward_ms = t_fwd - t_start;
  out->backward_ms = t_end - t_fwd;
  out->steps = nsteps;
  out->checkpoints = nseg;
  return 0;
}

extern "C" int psph_replay_run(psph_ctx *c, int64_t nsteps) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  if (c->cfg.adapt_domain) return fail(PSPH_ENOTSUP, "replay: domain adaptation is not part of the recorded map");
  Tape &tp = c->tape;
  if (tp.mode) return fail(PSPH_EINVAL, "replay: a recorded run is in progress");
  if (nsteps < 0 || nsteps > (int64_t)tp.dt.size()) return fail(PSPH_EINVAL, "replay: %lld steps, %zu recorded", (long long)nsteps, tp.dt.size());
  if (c->time != tp.t0 || c->step != tp.step0 || c->n != tp.n0)
    return fail(PSPH_EINVAL, "replay: load the state the recorded run started from (t %g, step %lld, %zu particles)",
                tp.t0, (long long)tp.step0, tp.n0);
  if ((int64_t)tp.xrec.size() < nsteps)
    return fail(PSPH_EINVAL, "replay: positions of %zu steps recorded, %lld asked", tp.xrec.size(), (long long)nsteps);
  tp.mode = 2;
  tp.frozen = true;
  tp.k = 0;
  int rc = 0;
  for (int64_t s = 0; s < nsteps && !rc; s++) rc = psph_step(c, nullptr);
  if (!rc) rc = check_err(c, "replay");
  tp.mode = 0;
  tp.frozen = false;
  return rc;
}

extern "C" int psph_set_vent(psph_ctx *c, double vent_v, double vent_T, double ng0) {
  if (!c) return fail(PSPH_EINVAL, "null ctx");
  psph_config k = c->cfg;
  k.vent_v = vent_v; k.vent_T = vent_T; k.ng0 = ng0;
  PSPH_TRY(psph_config_vent(&k));
  c->cfg = k;
  return 0;
}
