// Test infrastructure only: a C entry layer over the reference's own CRng
// (compiled from its rng.cpp, rng.h:30-65) so that the oracle's and the
// library's random-number restatements can be checked against it.  Built by
// oracle/Makefile (target `ref`) into oracle/_ref/ when the reference sources
// are present; nothing outside tests/ loads it.
#include "rng.h"

extern "C" {
void* ref_rng_create(int seed) { return new CRng(seed); }
void ref_rng_destroy(void* r) { delete static_cast<CRng*>(r); }
void ref_rng_seed(void* r, int seed) { static_cast<CRng*>(r)->seed(seed); }
int ref_rng_rand(void* r) { return static_cast<CRng*>(r)->rand(); }
double ref_rng_unirand(void* r, double a, double b) { return static_cast<CRng*>(r)->unirand(a, b); }
double ref_rng_unirand_oub(void* r, double a, double b) { return static_cast<CRng*>(r)->unirand_oub(a, b); }
double ref_rng_normrand(void* r, double mu, double s) { return static_cast<CRng*>(r)->normrand(mu, s); }
int ref_rng_poisrand(void* r, double mu) { return static_cast<CRng*>(r)->poisrand(mu); }
// n consecutive draws of unirand(a, b) (oub != 0: unirand_oub) into out
void ref_rng_unirand_n(void* r, double a, double b, int n, int oub, double* out) {
  CRng* g = static_cast<CRng*>(r);
  for (int i = 0; i < n; ++i) out[i] = oub ? g->unirand_oub(a, b) : g->unirand(a, b);
}
}
