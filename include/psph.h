/*
 * psph.h -- C ABI of the MI355X per-step SPH particle path (Plume-SPH drop-in).
 *
 * The reference has no plugin or FFI layer: its hot-path stages are C++ free
 * functions that mutate Particle objects through THashTable and HashTable pointers
 * (src/plume/particler.h:22-357), called only from the time loop in
 * src/plume/main.cc:217-486 (3D) and main.cc:676-751 (1D).  Each entry point
 * below replaces one of those functions and cites it.  Plain pointers and
 * sizes only; no torch or HIP types cross this boundary.
 *
 * Ownership: the caller owns every host buffer (copied in/out).  A psph_ctx
 * owns all device memory and one HIP stream.  Calls on one context are
 * serialised and stream-ordered; contexts are independent.
 *
 * Errors: every function returns 0 on success or a negative PSPH_E* code; the
 * reference aborts (assert/exit/MPI_Abort) where this returns a code.  The
 * message of the last failure on the calling thread is psph_last_error().
 */
#ifndef PSPH_H
#define PSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PSPH_OK           0
#define PSPH_EINVAL      -1   /* bad argument or configuration                */
#define PSPH_ENOMEM      -2   /* device or host allocation failed             */
#define PSPH_EINVARIANT  -3   /* rho<=0, c<=0, xi>1, wnorm<=0 ... (asserts)   */
#define PSPH_ECAPACITY   -4   /* particle capacity / inject buffer overflow   */
#define PSPH_EHIP        -5   /* HIP runtime error                            */
#define PSPH_EIO         -6   /* deck / meteo file cannot be read             */
#define PSPH_ENOTSUP     -7   /* option not implemented                       */

/* ---- particle boundary-condition types (particle.h:157-166) ------------- */
#define PSPH_BC_ERUPT_GHOST     0
#define PSPH_BC_PRESSURE_GHOST  1
#define PSPH_BC_WALL_GHOST      2
#define PSPH_BC_NOSAVE_REAL    99
#define PSPH_BC_REAL          100

/* ---- precision forms --------------------------------------------------- */
#define PSPH_FP64        0   /* plain FP64 (drop-in parity form)              */
#define PSPH_FP32_PAIR   1   /* FP32 pair arithmetic where the rounding does
                                not reach the velocities: the RHS pressure term
                                and the density / filter weights stay FP64
                                (every field within ~4e-7 rel L1 of FP64 over
                                95 steps of the 4M plume)                      */
#define PSPH_FP32_PURE   2   /* RHS and density pair arithmetic all FP32 (the
                                filter weights and the h update stay FP64):
                                faster, but velocities drift ~1e-4 rel L1 from
                                FP64 over the same run                         */
#define PSPH_FP32_EXP    3   /* PSPH_FP32_PAIR with the pair weights'
                                exponential in FP32 hardware (argument formed
                                in FP64 and rounded once; h, 1/h, norms and
                                sums FP64), and the density weights in FP32
                                throughout, summed in FP64                    */

/* ---- smoothing kernel -------------------------------------------------- */
/* The reference has one kernel, a Gaussian truncated per axis at |s_k| = 3
 * (USE_GAUSSIAN, sph_lib.cc:207-253, options.h:18-20,454-456); it has no
 * cubic spline.  psph_create rejects any other value of cfg.kernel. */
#define PSPH_KERNEL_GAUSS_BOX   0

/*
 * Solver configuration: every compile-time switch (options.h), constant
 * (constant.h) and physical parameter (parameters.h) the hot path reads,
 * as runtime values.  psph_config_default() fills the reference defaults.
 */
typedef struct psph_config {
  int32_t dim;              /* CODE_DIMENSION: 1 or 3                          */
  int32_t precision;        /* PSPH_FP64 / PSPH_FP32_PAIR / PSPH_FP32_PURE     */
  int32_t kernel;           /* PSPH_KERNEL_*                                   */
  int32_t lans;             /* HAVE_TURBULENCE_LANS (0 or 1)                   */
  int32_t energy_smooth;    /* HAVE_ENERGY_SMOOTH (0 or 1)                     */
  int32_t pghost_contrib;   /* PGHOST_CONTRIBUTE_DES (0, 1 or 2)               */
  int32_t list_all;         /* 1: every particle lists every particle (1D)     */
  int32_t num_loop;         /* num_loop_P: adaptive-h iterations               */
  int32_t sml_update_int;   /* SML_UPDATE_INT: adaptive-h cadence (0 = off)    */
  int32_t adapt_at_setup;   /* 1D main.cc:641: adaptive_sml once before loop   */
  int32_t vent_profile;     /* ERUPT_VELOCITY_PROF: 0 uniform, 1 parabolic     */
  int32_t max_steps;        /* deck line 2                                     */
  int32_t num_erupt_particles; /* parameters.h:983                            */
  int32_t p_table_size;     /* P_TABLE_SIZE: slots of the particle hash table  */
  int32_t wall_bc;          /* 1: refresh the wall ghosts next to the ground
                               from their mirror images after every step
                               (apply_bcond, main.cc:468); 0: keep their state */
  int32_t adapt_domain;     /* ADJUST_DOMAIN (main.cc:256-304): 1 grows the
                               region of real particles bucket by bucket as
                               involved material reaches its outer layer,
                               starting from idom (psph_domain_init)          */

  double cutoff;            /* CUTOFF = 3: kernel box half-width in h         */
  double cutoff2;           /* CUTOFF2: pair pre-filter box in h_i (8 / 12)   */
  double list_factor;       /* 4.25: neighbour-list box in h_i (neighs.cc:84) */
  double alpha, beta, ata;  /* alf_P, beta_P, ata_P (Monaghan viscosity)      */
  double eps_half;          /* EPSILON_HALF (SPH-epsilon turbulence)          */
  double prandtl;           /* PRANDTL_NUM                                     */
  double e_smooth_ratio;    /* E_SMOOTH_RATIO                                  */
  double heat_scale;        /* HEAT_TRANS_SCALE_RATIO                          */
  double g;                 /* g_P (gravity along the last axis, downward)     */
  double cfl;               /* CFL_P                                           */
  double thresh;            /* thresh_P (adaptive-h tolerance)                 */
  double c_smooth;          /* C_smooth_P                                      */
  double eta;               /* eta_smooth_P                                    */
  double ratio_sml_dx;      /* RATIO_SML_DX                                    */
  double msfrc_thresh;      /* MSFRC_THRESH                                    */
  double ng0, cvs, cvg, cva, rg, ra;   /* EOS constants (parameters.h:905-958) */
  double gamma_p;           /* gamma_P: initial-energy gamma (1.4)             */

  double h_air;             /* deck line 9: air smoothing length               */
  double bucket_edge;       /* PARTICLE_DENSITY * h_air                        */
  double dom_lo[3], dom_hi[3];   /* Ll_P, Lu_P: the simulated domain           */
  double ext_lo[3], ext_hi[3];   /* extended domain (key normalisation, cells) */

  double t_end;             /* deck line 1                                     */
  double t_output;          /* deck line 3 (dt is clipped to hit it)           */
  double erupt_start, erupt_end;  /* deck lines 4-5                            */

  /* vent (parameters.h:944-982); the derived ones are set by psph_config_vent */
  double vent_r, vent_v, vent_T, vent_p, vent_x, vent_y, vent_z;
  double vent_rho, vent_e, vent_gamma, vent_c;
  double vent_sml, vent_mass, vent_bot;
  double idom_lo[3], idom_hi[3];  /* initial domain (deck lines 10-15,
                               dataread.cc:326-332); dom is then the largest
                               one (Ll_P, Lu_P of the grid preprocessor)      */
  double reserved_d[2];
} psph_config;

/*
 * Structure-of-arrays particle view (particle.h:34-175).  Every pointer
 * addresses `n` elements (the caller sizes them).  key = TKey {hsfc[0],
 * hsfc[1], add_step}.  v = state_vars[1..3], rho = state_vars[0],
 * e = state_vars[NO_OF_EQNS-1], sv = smoothed_v, xi = mass_frac,
 * cp = specific_heat_p, gam = gamma, c = sound_speed, T = temperature.
 * For dim 1 only x, v[0], sv[0] are meaningful.
 */
typedef struct psph_soa {
  double *x, *y, *z;
  double *vx, *vy, *vz;
  double *svx, *svy, *svz;
  double *rho, *e, *m, *h, *p, *c, *T, *gam, *cp, *xi;
  uint32_t *key0, *key1, *key2;
  int32_t *bc, *phase, *involved;
} psph_soa;

typedef struct psph_step_out {
  double dt;                /* step size actually taken (after clipping)       */
  double time;              /* simulated time after the step                   */
  int64_t step;             /* step counter after the step                     */
  int64_t n_total;          /* particles held (all bc types)                   */
  int64_t n_real;           /* bc in {100, 99}                                 */
  int64_t n_injected;       /* particles added by the inlet this step          */
  int64_t n_removed;        /* real particles that escaped the domain          */
  int64_t n_pairs;          /* kernel-support pairs in the RHS pass            */
  int64_t output;           /* 1: the step reached an output time, the point
                               where the reference's loop writes its files
                               (TimeProps::ifoutput, main.cc:473-477)         */
  int64_t n_domain;         /* particles the domain adaptation added this step
                               (pressure and wall ghosts of the new layer)    */
} psph_step_out;

typedef struct psph_ctx psph_ctx;

/* ---- configuration ----------------------------------------------------- */
/* Reference defaults for dim 3 (options.h:452-837, constant.h:109-203,
 * parameters.h:889-1043) or dim 1 (options.h:16-450, parameters.h:1225-1446). */
int psph_config_default(psph_config *cfg, int dim);
/* Derived vent state: rho_v, e_v, gamma_v (parameters.h:958-975), the vent
 * column spacing and mass (Compute_mass, sph_lib.cc:1024-1040, called from
 * setup_erupt.cc:94-104) and the column bottom (setup_erupt.cc:93). */
int psph_config_vent(psph_config *cfg);
/* Extended domain and cell lattice from dom_lo/dom_hi and h_air exactly as
 * the grid preprocessor builds it (src/preproc/preprocess.cc:88-133). */
int psph_config_domain(psph_config *cfg);
/* Parse a reference simulation.data deck (dataread.cc:77-245): values go
 * through stof as in the reference.  meteo_path_out receives line 18. */
int psph_read_deck(const char *path, psph_config *cfg, char *meteo_path_out,
                   size_t meteo_path_cap, int32_t *meteo_rows, int32_t *meteo_cols);

/* ---- inputs ------------------------------------------------------------ */
/* Meteo table in the reference layout (meteo.cc:22-40): z[rows] in km and
 * data[prop*rows + r] for prop = rho(kg/m3), p(hPa), T(K), q, u, v. */
typedef struct psph_meteo {
  int32_t rows, nprop;
  const double *z;
  const double *data;
} psph_meteo;

/* Read a meteo file of rows x cols numbers (readFile, sph_lib.cc:1559;
 * Meteo::Meteo, meteo.cc:22-40).  z and data are caller buffers. */
int psph_meteo_read(const char *path, int32_t rows, int32_t cols, double *z, double *data);
/* Standard atmosphere at rest in that layout: T = T0 - 6.5 K/km up to 11 km,
 * isothermal above; hydrostatic p; rho = p / (R_air T).  Rows every dz_km
 * from 0 to z_top_km.  Returns the row count in *rows. */
int psph_meteo_standard(double t0, double p0, double z_top_km, double dz_km,
                        double *z, double *data, int32_t cap, int32_t *rows);
/* Particle keys THSFC3d (hilbert.c:329-359) of positions normalised to the
 * extended domain (add_air.cc:604-607). */
int psph_keys(const psph_config *cfg, const double *x, const double *y, const double *z,
              size_t n, uint32_t add_step, uint32_t *k0, uint32_t *k1, uint32_t *k2);
/* Synthetic plume input (add_air.cc:490-813 lattice, initial_air
 * sph_lib.cc:831-872, setup_erupt.cc:43-386 vent column): air lattice of
 * spacing h_air over the extended domain -- real air inside dom, pressure
 * ghosts (bc 1) in the shell above ground, wall ghosts (bc 2) below --
 * plus the vent's erupt-ghost column.  Seeded jitter (fraction of h_air)
 * from Philox4x32-10 keyed by lattice index.  Canonical order.
 * With out == NULL only *n_out (an upper bound) is returned. */
int psph_gen_plume(const psph_config *cfg, const psph_meteo *met, uint32_t seed,
                   double jitter, psph_soa *out, size_t cap, size_t *n_out);
/* 1D shock tube (add_air.cc:36-485): left/right state {rho, p, u}. */
int psph_gen_shock_tube(const psph_config *cfg, const double left[3], const double right[3],
                        double mid, int32_t n_nosave_l, int32_t n_nosave_r, int32_t n_ghost,
                        int32_t equal_mass, psph_soa *out, size_t cap, size_t *n_out);

/* ---- context ----------------------------------------------------------- */
int psph_create(const psph_config *cfg, size_t capacity, psph_ctx **out);
int psph_destroy(psph_ctx *ctx);
/* Upload n particles (any order) and reset time to t0 / step counter. */
int psph_load(psph_ctx *ctx, const psph_soa *host, size_t n, double t0, int64_t step0);
/* Download all particles in the reference's THashTable iteration order
 * (thashtab.cc:209-262): ascending slot, then (key0,key1,key2). */
int psph_read(psph_ctx *ctx, psph_soa *host, size_t cap, size_t *n_out);
int psph_count(psph_ctx *ctx, size_t *n_total, size_t *n_real);
int psph_time(psph_ctx *ctx, double *time, int64_t *step);
/* The configuration the context runs (psph_create's, as completed there). */
int psph_get_config(psph_ctx *ctx, psph_config *cfg);
/* Zero-copy view of the context's particle arrays: every pointer of *dev
 * receives the DEVICE address of that field's array, *n the particle count,
 * *stream (if not NULL) the context's hipStream_t.  Entries [0, n) are the
 * particles in the context's current order: the order given to psph_load right
 * after it, the cell order after a neighbour search.  The call waits for the
 * context's pending work.  The addresses stay valid until the next call that
 * loads, sorts (psph_neighbors, psph_step, psph_run) or runs the adjoint: a
 * sort gathers every field into a spare array and swaps.  Writes through the
 * view change the state the next stage reads.  The reference hands its
 * particles out as Particle pointers through THashTable iteration
 * (thashtab.h THTIterator); this is the device-resident counterpart, used by
 * the torch wrapper. */
int psph_device_view(psph_ctx *ctx, psph_soa *dev, size_t *n, void **stream);

/* ---- stages (reference granularity) ------------------------------------ */
/* dt = CFL * min h/c over need_neigh particles (timestep.cc:94-135, 1D
 * :139-183), then TimeProps::incrtime clipping (properties.h:71-82).  Only
 * computes; psph_advance_time commits. */
int psph_timestep(psph_ctx *ctx, double *dt_out);
int psph_advance_time(psph_ctx *ctx, double *dt_inout);
/* search_neighs_consth (neighs.cc:37-211): cell-key radix sort + cell
 * ranges.  Lists are not materialised (see psph_get_neighbors). */
int psph_neighbors(psph_ctx *ctx);
/* adaptive_sml (neighs.cc:479-693). */
int psph_adapt_h(psph_ctx *ctx);
/* mom_engr_update, USE_GSPH==0 (mom_engr_update.cc:30-439) + EOS. */
int psph_rhs(psph_ctx *ctx, double dt);
/* smooth_density (density_massfrac.cc:25-359) + EOS. */
int psph_density(psph_ctx *ctx);
/* smooth_velocity (smooth_velocity.cc:25-165). */
int psph_vsmooth(psph_ctx *ctx);
/* update_pos (update_pos.cc:34-168) + escape removal
 * (update_particles_in_bucket, update_pos.cc:171-463). */
int psph_update_pos(psph_ctx *ctx, double dt);
/* add_new_erupt (setup_erupt.cc:389-670); *added receives the count. */
int psph_inject(psph_ctx *ctx, double dt, int64_t *added);
/* apply_bcond (bcond.cc:24-249): refresh the wall ghosts of the bucket layer
 * that straddles the ground (z = dom_lo[2]) from their mirror images; runs a
 * neighbour search first when none is current.  psph_step calls it at the end
 * of every step when cfg.wall_bc is set (as main.cc:468 does); call it once
 * after psph_load for the refresh the reference makes before its loop
 * (main.cc:196). */
int psph_apply_bcond(psph_ctx *ctx);
/* Domain adaptation (ADJUST_DOMAIN; scan_outside_layer.cc:44-121,
 * adapt_domain.cc:41-194, add_pressure_ghost.cc:37-316, add_wall_ghost.cc:
 * 45-293, called from main.cc:256-304 and 433-436).  With cfg.adapt_domain,
 * call this once after psph_load: it sets up the buckets of the grid
 * preprocessor (edge cfg.bucket_edge over the extended domain), whose
 * has_involved starts from the initial domain cfg.idom_lo/idom_hi and the
 * buckets that hold real particles, and keeps the atmosphere, seed and jitter
 * of the air lattice the new ghost layers are cut from (psph_gen_plume's).
 * psph_step then ends with scan_outside_layer and, when involved material
 * reached the outer layer of the region of real particles, starts the next
 * step by growing that region by the bordering buckets: their particles turn
 * real (potentially involved), the empty buckets around get pressure ghosts,
 * the ground layer under them wall ghosts, and the wall ghosts are refreshed
 * when cfg.wall_bc is set.  psph_gen_plume with cfg.adapt_domain makes the
 * matching initial state: real air in the initial domain's buckets and the
 * ghost layers around them. */
int psph_domain_init(psph_ctx *ctx, const psph_meteo *met, uint32_t seed, double jitter);
/* One whole step in the order of main.cc:217-486 (3D) / 676-751 (1D). */
int psph_step(psph_ctx *ctx, psph_step_out *out);
int psph_run(psph_ctx *ctx, int64_t nsteps, psph_step_out *last);

/* ---- parity and diagnostics -------------------------------------------- */
/* Neighbour lists of the need_neigh particles (neighs.cc:37-211) as CSR in
 * read order: offsets[n+1], idx[] (ascending read-order indices).  Pass
 * idx=NULL to get only offsets (offsets[n] = total). */
int psph_get_neighbors(psph_ctx *ctx, int64_t *offsets, size_t n_cap,
                       uint32_t *idx, size_t idx_cap);
/* Totals over real particles: out[0]=sum m, out[1]=sum m e,
 * out[2]=sum m |v|^2/2, out[3..5]=sum m v, out[6]=sum m xi. FP64. */
int psph_conservation(psph_ctx *ctx, double out[7]);
/* Plume diagnostic: mass-weighted mean height of erupted material over real
 * particles, H = sum m xi z / sum m xi.  No reference counterpart. */
int psph_diag_height(psph_ctx *ctx, double *H);
/* Allocation-free variant of psph_conservation + psph_diag_height, with no
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  It needs no context: dev holds DEVICE pointers of n
 * particles (bc, m, e, vx, xi; also vy, vz, z when dim is 3; e.g. from
 * psph_device_view), out_dev is a DEVICE array of 8 doubles that receives the
 * 7 sums of psph_conservation and sum m xi z (H = out[7] / out[6]), and
 * stream is a hipStream_t (NULL: the null stream).  The work is queued on the
 * stream and not waited for.  It allocates nothing: its block partials live in
 * a 32 KiB array in the library's device memory, one per GPU, held until the
 * process ends, so calls that overlap on one GPU race.  n above 4294967040
 * (2^32 - 256) is rejected with PSPH_EINVAL.  Results are bit-identical to
 * psph_conservation's on the same particles in the same order. */
int psph_conservation_noalloc(const psph_soa *dev, size_t n, int32_t dim, double *out_dev, void *stream);

/* ---- adjoint ------------------------------------------------------------ */
/* Reverse-mode gradient of the plume diagnostic H (psph_diag_height) after
 * nsteps steps (psph_step) from the context's current state.  No reference
 * counterpart: the reference has no adjoint.  Differentiated with respect to
 * the vent exit velocity (cfg.vent_v), the vent temperature (cfg.vent_T), the
 * gas mass fraction of the erupted material (cfg.ng0; its complement is the
 * erupted solid mass fraction) and every particle's initial velocity, with
 * the step's discrete schedule frozen at the recorded run: time steps,
 * adapted smoothing lengths, inlet spawn counts, erupt-ghost -> real and
 * escape events, and the pair set of every pass (which pairs lie inside the
 * kernel boxes, the list and support boxes; a box indicator has zero
 * derivative wherever it is differentiable).  The recorded run keeps every
 * step's particle positions by id (24 B per particle and step) for
 * psph_replay_run.  With cfg.wall_bc the wall-ghost refresh that ends every
 * step is differentiated too (its contributor sets frozen like the pair set;
 * the step records take 84 B and the recorded run 24 B more per particle and
 * step); a context with cfg.adapt_domain is refused (PSPH_ENOTSUP).  Erupt
 * ghosts present at the start (bc 0) are taken to hold the vent state of the
 * config, as psph_gen_plume makes them.
 * The forward run stores a checkpoint every ckpt_every steps; the reverse
 * sweep recomputes one segment at a time from its checkpoint.  grad_v and
 * grad_sv (either may be NULL) receive dH/dv and dH/dsv of the starting
 * state, 3 per particle, in psph_read order of that state (n_cap particles
 * at most).  On return the context holds the state after nsteps steps, and
 * the recorded schedule stays available to psph_replay_run. */
typedef struct psph_adjoint_out {
  double H;                  /* diagnostic at the end of the run */
  double dH_dvent_v;         /* per m/s */
  double dH_dvent_T;         /* per K */
  double dH_dng0;            /* per unit gas mass fraction */
  double dH_dmass_frac;      /* per unit erupted solid mass fraction (= -dH_dng0) */
  double forward_ms;         /* recorded forward run incl. checkpoints */
  double backward_ms;        /* recompute + reverse sweep */
  int64_t steps, checkpoints, n_particles;
  double reserved[4];
} psph_adjoint_out;
int psph_adjoint_run(psph_ctx *ctx, int64_t nsteps, int64_t ckpt_every, psph_adjoint_out *out,
                     double *grad_v, double *grad_sv, size_t n_cap);
/* Runs nsteps steps with the frozen schedule of the last psph_adjoint_run
 * (the map the adjoint differentiates; finite-difference checks use it):
 * besides the schedule, the cell order, tiles, visit lists and every pair
 * predicate of a step (and, with cfg.wall_bc, the wall refresh's support
 * tests) are evaluated on the recorded run's positions of that step, and
 * only the arithmetic on the replay's own, so the replay visits the recorded
 * pairs in the recorded summation order.  The context
 * must hold a state loaded at that run's start time and step, with the same
 * particle count. */
int psph_replay_run(psph_ctx *ctx, int64_t nsteps);
/* Sets the vent exit velocity, temperature and gas mass fraction of the
 * context's config and recomputes the derived vent state (psph_config_vent). */
int psph_set_vent(psph_ctx *ctx, double vent_v, double vent_T, double ng0);

/* Per-stage GPU time measured with HIP events on the context's stream.
 * Stage ids: 0 timestep, 1 neighbours (keys+sort+gather+ranges), 2 adapt_h,
 * 3 rhs, 4 density, 5 vsmooth, 6 update_pos, 7 inject.  psph_profile(ctx, 1)
 * enables and clears; psph_profile_read returns accumulated ms and the
 * number of launches of each stage's main kernel. */
#define PSPH_NSTAGES 8
int psph_profile(psph_ctx *ctx, int enable);
int psph_profile_read(psph_ctx *ctx, double ms[PSPH_NSTAGES], int64_t launches[PSPH_NSTAGES]);

const char *psph_last_error(void);
const char *psph_version(void);
/* sizeof(psph_config), sizeof(psph_soa), sizeof(psph_step_out),
 * sizeof(psph_adjoint_out) for FFI bindings to check their struct layout:
 * which = 0, 1, 2, 3. */
size_t psph_abi_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif /* PSPH_H */
