/*
 * psph_oracle.h -- CPU FP64 restatement of the Plume-SPH per-step particle
 * path.  TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Particles are
 * held in the reference's THashTable iteration order (thashtab.cc:209-262)
 * after every call that adds or removes particles.
 */
#ifndef PSPH_ORACLE_H
#define PSPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/psph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_meteo {
  int rows, nprop;          /* nprop = columns - 1 (meteo.cc:22-40)          */
  double *z;                /* km, ascending                                 */
  double *data;             /* data[prop*rows + r]: rho, p(hPa), T, ...      */
} orc_meteo;

typedef struct orc_state {
  size_t n, cap;
  double *x[3], *v[3], *sv[3];
  double *rho, *e, *m, *h, *p, *c, *T, *gam, *cp, *xi;
  uint32_t *key[3];
  int32_t *bc, *phase, *inv;
  /* neighbour lists (CSR, indices into the current order) */
  int64_t *nb_off;          /* n+1                                           */
  uint32_t *nb_idx;
  size_t nb_cap;
  /* time (TimeProps, properties.h:20-150) */
  double time, next_output;
  int64_t step;
  int ioutput;
  /* counters of the last step */
  int64_t last_injected, last_removed, last_pairs;
  /* domain adaptation (ADJUST_DOMAIN), set up by orc_domain_init */
  uint8_t *bval;            /* has_involved of each bucket: bit 0 potentially
                               involved, bit 1 involved (bucket.h:400-414)  */
  uint8_t *berupt;          /* erupt_flag of each bucket (setup_erupt.cc:221-253) */
  int32_t nbk[3];           /* buckets per axis                              */
  int32_t adapt_pending;    /* scan_outside_layer's count of the last step   */
  uint32_t lat_seed;        /* jitter of the lattice the adaptation fills     */
  double lat_jitter;
  orc_meteo met;            /* atmosphere of the particles it adds           */
  int64_t last_adapted;     /* particles the last adaptation added           */
} orc_state;

/* kernel (sph_lib.cc:207-253) */
double orc_weight(const double *s, double h, int dim);
double orc_d_weight(const double *s, double h, int dir, int dim);
/* EOS update_second_var (particle.cc:528-604) */
void orc_eos(const psph_config *cfg, double rho, double e, double xi,
             double *p, double *c, double *T, double *gam, double *cp);
/* Hilbert keys (hilbert.c:156-271, :329-359) */
void orc_hsfc3d_bits(const uint32_t coord[3], uint32_t key[2]);
void orc_thsfc3d(const double normc[3], uint32_t add_step, uint32_t key[3]);
int64_t orc_slot(const uint32_t *key, int table_size, int dim);
/* Philox4x32-10 (Salmon et al., SC'11) */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* meteo (meteo.cc:64-143) */
int orc_meteo_load(const char *path, int rows, int cols, orc_meteo *met);
void orc_meteo_from_arrays(orc_meteo *met, int rows, int nprop, const double *z, const double *data);
void orc_meteo_free(orc_meteo *met);
void orc_meteo_interp_all(const orc_meteo *met, double zkm, double *out);
double orc_meteo_interp(const orc_meteo *met, double zkm, int prop);

/* state */
int orc_state_init(orc_state *st, size_t cap);
void orc_state_free(orc_state *st);
int orc_state_reserve(orc_state *st, size_t cap);
int orc_state_load(orc_state *st, const psph_soa *soa, size_t n);
int orc_state_store(const orc_state *st, psph_soa *soa);
void orc_sort_canonical(const psph_config *cfg, orc_state *st);

/* set-up */
int orc_air_lattice(const psph_config *cfg, const orc_meteo *met, uint32_t seed,
                    double jitter, orc_state *st);                 /* add_air.cc:490-813 */
int orc_setup_erupt(const psph_config *cfg, orc_state *st);        /* setup_erupt.cc:43-386 */
int orc_shock_tube(const psph_config *cfg, double rho_l, double p_l, double u_l,
                   double rho_r, double p_r, double u_r, double mid,
                   int n_nosave_l, int n_nosave_r, int n_ghost, int equal_mass,
                   orc_state *st);                                  /* add_air.cc:36-485 */

/* stages */
double orc_timestep(const psph_config *cfg, const orc_state *st);   /* timestep.cc:94-183 */
void orc_incrtime(const psph_config *cfg, orc_state *st, double *dt);  /* properties.h:71-82 */
int orc_neighbors(const psph_config *cfg, orc_state *st);           /* neighs.cc:37-211 */
int orc_adaptive_sml(const psph_config *cfg, orc_state *st);        /* neighs.cc:479-693 */
int orc_rhs(const psph_config *cfg, orc_state *st, double dt);      /* mom_engr_update.cc:30-439 */
int orc_density(const psph_config *cfg, orc_state *st);             /* density_massfrac.cc:25-359 */
int orc_vsmooth(const psph_config *cfg, orc_state *st);             /* smooth_velocity.cc:25-165 */
int orc_update_pos(const psph_config *cfg, orc_state *st, double dt); /* update_pos.cc:34-168 */
int orc_remove_escaped(const psph_config *cfg, orc_state *st);      /* update_pos.cc:171-463 */
int orc_inject(const psph_config *cfg, orc_state *st, double dt);   /* setup_erupt.cc:389-670 */
int orc_apply_bcond(const psph_config *cfg, orc_state *st);         /* bcond.cc:24-249 */
int orc_step(const psph_config *cfg, orc_state *st);                /* main.cc:217-486, 676-751 */
/* domain adaptation: bucket flags of the state (dataread.cc:326-457,
 * setup_erupt.cc:221-253); the initial state of an adaptive run (add_air.cc:
 * 491-813, add_pressure_ghost.cc, add_wall_ghost.cc, setup_erupt.cc); one
 * scan (scan_outside_layer.cc:44-121) and one adaptation (main.cc:256-304) */
int orc_domain_init(const psph_config *cfg, const orc_meteo *met, uint32_t seed, double jitter,
                    orc_state *st);
int orc_domain_setup(const psph_config *cfg, const orc_meteo *met, uint32_t seed, double jitter,
                     orc_state *st);
int orc_domain_scan(const psph_config *cfg, orc_state *st);
int orc_domain_adapt(const psph_config *cfg, orc_state *st);
void orc_bucket_type(const psph_config *cfg, int bx, int by, int bz, int bt[6], int *type);
int orc_iferupt(const psph_config *cfg, const orc_state *st);       /* properties.h:126-135 */
void orc_conservation(const orc_state *st, double out[7]);
double orc_diag_height(const orc_state *st);

/* brute-force O(N^2) list with the same predicate, for pinning (tests) */
int orc_neighbors_bruteforce(const psph_config *cfg, orc_state *st);

int orc_set_threads(int n);
int orc_get_threads(void);
const char *orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
