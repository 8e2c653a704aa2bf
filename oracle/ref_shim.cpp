// C-ABI shim around a g++ build of the reference CPU solver (TEST INFRASTRUCTURE ONLY).
//
// oracle/Makefile compiles this file together with CPU_Solver/*.cpp and Auxiliary/*.cpp taken in place
// from the reference tree, once per (POT_SCHEME, POT_GHOST_SIZE, GRA_GHOST_SIZE, FLOAT8) combination,
// into oracle/_ref/.  Tests use those libraries to pin the oracle restatement bit for bit; nothing in the
// product links them.  The entry point forwards to CPU_PoissonGravitySolver (CPU_Solver/
// CPU_PoissonGravitySolver.cpp:52-87) with the argument list of Header/Prototype.h:6-14.
#include "GAMER.h"

#include <fcntl.h>
#include <unistd.h>

extern "C" {

// Sizes the library was built for: {RHO_NXT, POT_NXT, GRA_NXT, POT_GHOST_SIZE, GRA_GHOST_SIZE,
// POT_SCHEME, sizeof(real)}.
void ref_geometry(int* out) {
  out[0] = RHO_NXT;
  out[1] = POT_NXT;
  out[2] = GRA_NXT;
  out[3] = POT_GHOST_SIZE;
  out[4] = GRA_GHOST_SIZE;
  out[5] = POT_SCHEME;
  out[6] = int(sizeof(real));
}

// The reference SOR prints one line per patch (CPU_PoissonSolver_SOR.cpp:288); stdout is sent to
// /dev/null for the duration of the call so that callers see only their own output.
int ref_poisson_gravity(const real* rho, real* cpot, real* pot_out, real* flu, int npatch_group, double dt,
                        double dh, int sor_min_iter, int sor_max_iter, double sor_omega, int mg_max_iter,
                        int mg_npre, int mg_npost, double mg_tol, double poi_coeff, int int_scheme, int p5,
                        int poisson, int graacc, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  fflush(stdout);
  const int saved = dup(1);
  const int devnull = open("/dev/null", O_WRONLY);
  if (devnull >= 0) dup2(devnull, 1);
  CPU_PoissonGravitySolver((const real(*)[RHO_NXT][RHO_NXT][RHO_NXT])rho, (real(*)[POT_NXT][POT_NXT][POT_NXT])cpot,
                           (real(*)[GRA_NXT][GRA_NXT][GRA_NXT])pot_out,
                           (real(*)[5][PATCH_SIZE][PATCH_SIZE][PATCH_SIZE])flu, npatch_group, (real)dt, (real)dh,
                           sor_min_iter, sor_max_iter, (real)sor_omega, mg_max_iter, mg_npre, mg_npost,
                           (real)mg_tol, (real)poi_coeff, (IntScheme_t)int_scheme, p5 != 0, poisson != 0,
                           graacc != 0);
  fflush(stdout);
  if (saved >= 0) {
    dup2(saved, 1);
    close(saved);
  }
  if (devnull >= 0) close(devnull);
  return 0;
}

}  // extern "C"
