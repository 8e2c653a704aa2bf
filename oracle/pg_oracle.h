/*
 * CPU oracle for the batched per-patch Poisson-gravity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C++, the CPU solvers of the reference
 * testbed (CPU_Solver/CPU_PoissonGravitySolver.cpp, CPU_PoissonSolver_SOR.cpp, CPU_PoissonSolver_MG.cpp,
 * CPU_GravitySolver.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker; the product library never links or calls it.
 */
#ifndef PG_ORACLE_H
#define PG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of one batched call.  Integer enums use the reference values:
 *   scheme      1 = SOR, 2 = MG                       (Header/Macro.h:47-48)
 *   int_scheme  4 = CQUAD, 5 = QUAD                   (Header/Typedef.h:7-8)
 * New in this project (no reference counterpart):
 *   init_mode   0 = interpolated cube (reference), 1 = zero interior, 2 = caller-supplied interior guess
 *   term_mode   SOR only: 0 = reference rule (stop when the L1 residual grows), 1 = relative residual */
typedef struct {
  int ghost_size;          /* POT_GHOST_SIZE, 1..5 */
  int gra_ghost_size;      /* GRA_GHOST_SIZE, 1..2, <= ghost_size */
  int scheme;
  int int_scheme;
  int init_mode;
  int term_mode;
  int check_every;         /* relative-residual check cadence in iterations (>= 1) */
  int sor_min_iter;
  int sor_max_iter;
  double sor_omega;
  int mg_max_iter;
  int mg_npre_smooth;
  int mg_npost_smooth;
  double mg_tolerated_error;
  double rel_tol;          /* relative-residual tolerance for term_mode 1 */
  double poi_coeff;        /* 4*pi*G */
  double dh;
  double dt;
  int p5_gradient;
  int poisson;
  int graacc;
} OrcParams;

/* Batched solve.  Layouts (x fastest):
 *   rho     [npatch][R][R][R],  R = 6 + 2*ghost_size
 *   cpot    [npatch][C][C][C],  C = 4 + 2*((ghost_size+3)/2)
 *   pot_out [npatch][G][G][G],  G = 8 + 2*gra_ghost_size
 *   flu     [npatch][5][8][8][8]  (rho, px, py, pz, E), updated in place when graacc
 *   guess   [npatch][R][R][R]   interior initial guess (init_mode 2 only, else may be NULL)
 * iters (may be NULL) receives the iteration count per patch; relres (may be NULL) receives the FP64
 * relative residual sum|R|/sum|phi| of the returned field.  Returns 0 or a negative error code. */
int orc_poisson_gravity_f32(const float* rho, const float* cpot, float* pot_out, float* flu, const float* guess,
                            int npatch, const OrcParams* p, int* iters, double* relres, int nthreads);
int orc_poisson_gravity_f64(const double* rho, const double* cpot, double* pot_out, double* flu,
                            const double* guess, int npatch, const OrcParams* p, int* iters, double* relres,
                            int nthreads);

/* Building blocks, exposed so that tests can pin each one on its own. */
int orc_interp_f64(const double* cpot, double* cube, int ghost_size, int int_scheme, int npatch);
int orc_interp_f32(const float* cpot, float* cube, int ghost_size, int int_scheme, int npatch);
int orc_restrict_f64(const double* fine, double* coarse, int nf, int nc);
int orc_prolongate_add_f64(const double* coarse, double* fine, int nc, int nf);
int orc_mg_levels(int n0, int* ngrid, int* nbottom_smooth);  /* returns the number of levels */
/* 7-point residual R = sum6 - 6 phi - Const*rho on the interior of an N^3 cube; returns sum|R|/sum|phi| */
double orc_relres_f64(const double* cube, const double* rho, int ghost_size, double poi_coeff, double dh);
double orc_relres_f32(const float* cube, const float* rho, int ghost_size, double poi_coeff, double dh);

#ifdef __cplusplus
}
#endif
#endif
