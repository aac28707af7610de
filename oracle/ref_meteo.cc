// C entry points over the reference's Meteo class (src/datstr/meteo.cc, built
// from where it lies by `make -C oracle ref`), for the tests' checks of the
// atmosphere interpolation.  Test infrastructure: never part of the product.
//
// A Meteo is created per table and never destroyed: its destructor releases
// the per-property vectors allocated with new[] by a plain delete
// (meteo.cc:59-62), which is undefined behaviour.
#include "meteo.h"

extern "C" void *ref_meteo_new(double *mat, int nr, int np) { return new Meteo(mat, nr, np); }

extern "C" double ref_meteo_interp(void *m, double z0, int prop) {
  return static_cast<Meteo *>(m)->interpolate(z0, prop);
}

extern "C" void ref_meteo_interp_all(void *m, double z0, double *out) {
  static_cast<Meteo *>(m)->interpolate(z0, out);
}
