/*
 * psph_io.h -- on-disk output of the particle state (host library
 * libpsph_io.so, linked against HDF5 and libpsph.so).
 *
 * H5Part files as the reference writes them (h5part_writer.cc:40-271,
 * hdf5calls.cc:288-346): one HDF5 group "Step#k" per output, k counting the
 * outputs of the run from 0, each holding one-dimensional datasets of the
 * selected particles in THashTable order --
 *   ID (int32, 0..n-1), x, y, z, Rho, Vx, Vy, Vz, engr, mssfrc, prss
 *   (float64), phase, bctp, guest (int32; guest = 1 for a particle this
 *   process owns, Particle::get_guest, particle.h:623-626)
 * -- contiguous layout, native little-endian types.  The file is created
 * (truncated) for the first output of a run and opened read-write after
 * (TimeProps::ifstart, GH5_fopen_serial modes 'w' / 'a').  The one-process
 * file name is pvplot%03d.h5part of the rank (pvplot000.h5part here); the
 * caller passes the path.
 *
 * The 1D driver also appends CSV tables (write_csv, h5part_writer.cc:
 * 448-489) when the output time is within 1.5 % of the comparison time
 * (main.cc:743-750).
 *
 * Return codes and psph_last_error() as in psph.h.
 */
#ifndef PSPH_IO_H
#define PSPH_IO_H

#include <stddef.h>
#include <stdint.h>

#include "psph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which particles a step holds */
#define PSPH_H5_REAL  0   /* bc 100 and not guest: write_h5part (h5part_writer.cc:83-121)          */
#define PSPH_H5_SHOW  1   /* need_neigh, bc 100/99: write_h5part_show (h5part_writer.cc:302-320)  */
#define PSPH_H5_ALL   2   /* every particle: the WRITE_GHOSTS build (h5part_writer.cc:82-121)     */

/* extra datasets of the reference's DEBUG build (h5part_writer.cc:181-205,
 * options.h:430-445), written after ID in this order */
#define PSPH_H5_INVOLVED 1   /* "Involved"  int32                         */
#define PSPH_H5_PID      2   /* "myprocess" int32 (the rank argument)    */
#define PSPH_H5_MASS     4   /* "mymass"    float64                      */
#define PSPH_H5_SML      8   /* "sml"       float64                      */

/* Write group "Step#<group>" of the particles p[0..n) (already in the order
 * the file should hold) to path.  create != 0 truncates or creates the file,
 * otherwise it must exist.  dim 1 writes y = z = Vy = Vz = 0.  Fails with
 * PSPH_EIO when the file cannot be opened or the group already exists.
 * Replaces write_h5part / write_h5part_show (h5part_writer.cc:40-271,
 * 274-444) for a caller that holds the particles on the host. */
int psph_h5part_write_soa(const char *path, int create, int64_t group, const psph_soa *p, size_t n,
                          int32_t dim, int32_t which, int32_t extras, int32_t rank);

/* The same for the particles of a context (psph_read order, which is the
 * reference's THashTable iteration order).  The caller counts the outputs
 * (the writer's static step counter, h5part_writer.cc:65, 128-129) and
 * creates the file on the first (timeprops->ifstart(), h5part_writer.cc:
 * 75-78); psph_step_out.output marks the steps after which the reference's
 * loop writes (main.cc:473-477). */
int psph_write_h5part(psph_ctx *ctx, const char *path, int create, int64_t group, int32_t which,
                      int32_t extras);

/* Number of Step#k groups in an H5Part file (k = 0.. consecutive). */
int psph_h5part_steps(const char *path, int64_t *nsteps);

/* Read dataset `name` of group Step#<group>: *n_out = its length; with
 * buf != NULL and n_out <= cap the values are converted to float64
 * (is_int == 0) or int32 (is_int != 0) into buf. */
int psph_h5part_read(const char *path, int64_t group, const char *name, int32_t is_int, void *buf,
                     size_t cap, size_t *n_out);

/* 1D CSV table (write_csv, h5part_writer.cc:448-489): appends the header
 * line and one row per real particle (bc 100), columns
 * rho, x, u, e, xi, p, phase, bc, guest flag, involved, h, m, in the
 * default C++ stream format (six significant digits). */
int psph_csv_write_soa(const char *path, const psph_soa *p, size_t n);
int psph_write_csv(psph_ctx *ctx, const char *path);

/* ---- the preprocessor's grid file funky%04d.h5 --------------------------
 * The grid preprocessor (preprocess.cc, createfunky.cc:32-71) writes per
 * process "/hashtable_constants" (6 float64: x min, x max, y min, y max,
 * z min, z max of the extended domain, preprocess.cc:136-141), "/Buckets"
 * (one compound record per bucket, in partition order) and
 * "/partition_table" (uint32, 4 per bucket: partition key, bucket key), and
 * Read_Grid (dataread.cc:248-499) reads them back.  The record's compound
 * type (GH5_bucketstruct, hdf5calls.cc:175-203) has the size and member
 * offsets of BucketStruct (buckstr.h:14-26) but declares each array member
 * as a scalar, so a file holds, per bucket, Key (key[0]), neighs (the first
 * neighbour's key[0]), buckettype (1 underground, 2 mixed, 3 overground, 4
 * pressure boundary), myproc, neigh_proc (the first neighbour's process),
 * xcoord, ycoord, zcoord (the lower corner) and elev (the ground elevation,
 * elev[0]).  This build derives its buckets from the deck (psph_config_domain
 * and the preprocessor's rules); these functions read a grid file, check a
 * configuration against it, and write one. */
typedef struct psph_grid_info {
  double ext_lo[3], ext_hi[3];   /* /hashtable_constants                         */
  double bucket_edge;            /* spacing of the buckets' lower corners        */
  int32_t nb[3];                 /* buckets per axis                             */
  int64_t n_buckets;             /* records in /Buckets                          */
  int64_t n_type[5];             /* records per buckettype 0..4                  */
  int32_t nproc;                 /* 1 + the largest myproc                       */
  int32_t flat;                  /* every record holds the same elev             */
  double ground;                 /* that elevation (the first record's if not)   */
  int64_t n_partition;           /* uint32 words in /partition_table             */
} psph_grid_info;
int psph_grid_read(const char *path, psph_grid_info *out);
/* Checks cfg (after psph_config_domain) against a grid file: the same
 * extended domain and bucket edge (1e-9 relative), every bucket of the file
 * at its place with the type this build derives, and a flat ground at
 * cfg->dom_lo[2].  PSPH_EINVAL on a mismatch; PSPH_ENOTSUP for a grid split
 * over several processes or with varying ground elevations (terrain), which
 * this build does not run. */
int psph_grid_check(const char *path, const psph_config *cfg);
/* Writes the one-process grid file of cfg's domain (the preprocessor with
 * np = 1 and a flat ground at cfg->dom_lo[2]) in the layout above. */
int psph_grid_write(const char *path, const psph_config *cfg);

#ifdef __cplusplus
}
#endif
#endif
