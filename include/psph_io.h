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

#ifdef __cplusplus
}
#endif
#endif
