/*
 * thirring_hmc.h -- the reference's HMC interface (hmc.c:73-746) with its
 * own function signatures, implemented over libth2d (libthirring_hmc.so).
 *
 * A program written against hmc.c links this library instead of compiling
 * those function bodies.  Its globals keep their meaning: m, g, mu
 * (hmc.c:38-40) and the gauge field A (hmc.c:47, used by fermion_matrix)
 * are read at every call; the library carries weak definitions of them,
 * so the program's own definitions take precedence when it has them.  The
 * random stream is the reference's: mersenne_i, mersenne_array,
 * seed_mersenne and mersenne_generate (mersenne.h, mersenne_inline.c) are
 * weak here as well and are replaced by the program's, if it links
 * mersenne_inline.c, so both draw from one stream.
 *
 * Arrays are the reference's jagged ones: vectors v[t][x], fields
 * A[t][x][dir] (rows may be longer than NX, as alloc_vector makes them).
 * The lattice extents that hmc.c takes from #define NT/NX are set once
 * with thirring_hmc_init().  Boundary conditions are antiperiodic in t and
 * x, as hmc.c:143-180.
 *
 * Differences from the reference, by design:
 *  - fm_conjugate_mul is the true adjoint M^dag (the reference body,
 *    hmc.c:188-249, repeats fm_mul);
 *  - the solver stop is configurable (thirring_hmc_set_cg); the default is
 *    the reference's absolute r.r < 1e-30 with at most 100000 iterations
 *    (hmc.c:34-35);
 *  - errors that hmc.c ends with exit(1) (hmc.c:383-388) are returned by
 *    thirring_hmc_last_status() instead; the functions still fill their
 *    outputs.
 */
#ifndef THIRRING_HMC_H
#define THIRRING_HMC_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

/* globals of hmc.c (weak in the library) */
extern double m, g, mu;
extern double ***A;

/* reference RNG interface (mersenne.h) */
extern int mersenne_i;
extern double mersenne_array[624];
void seed_mersenne(long seed);
double mersenne_generate(void);

/* ---- set-up replacing the compile-time constants ---------------------- */
int thirring_hmc_init(int nt, int nx);          /* NT, NX (hmc.c:14-15)     */
void thirring_hmc_finalize(void);
/* Solver of every inversion: relative true residual tol_rel (0 = off),
 * absolute r.r stop tol_abs_rr, max_iter, precision (0 FP64, 2 mixed). */
int thirring_hmc_set_cg(double tol_rel, double tol_abs_rr, int max_iter, int precision);
/* Leapfrog steps and trajectory length of update_gauge (hmc.c:708-709). */
int thirring_hmc_set_md(int nsteps, double traj_len);
/* Status of the last call (TH2D_* code; 0 = OK). */
int thirring_hmc_last_status(void);
/* Acceptance of the last update_gauge, its dH and CG iterations. */
int thirring_hmc_last_trajectory(int *accepted, double *dH, int *cg_iterations);

/* ---- checkpoint (no counterpart in the reference, whose HMC runs cannot
 * resume: hmc.c:932-939) ---------------------------------------------------
 * Write the field A (FP64), m, g, mu, the caller's trajectory counter and
 * the whole MT19937 state to `path`; reading it back (same lattice extents)
 * restores A, m, g, mu, the counter and the random stream, so the next
 * update_gauge repeats the trajectory the uninterrupted run would have
 * made, bit for bit.  Returns 0, TH2D_EARG (I/O error, other lattice, not a
 * checkpoint) or, on reading, THIRRING_HMC_RNG_PARTIAL when the program
 * supplies its own generator (mersenne_inline.c): its state words are
 * file-static, so only the current block of 624 draws is restored. */
#define THIRRING_HMC_RNG_PARTIAL 1
int thirring_hmc_write_checkpoint(const char *path, double ***A, long trajectory);
int thirring_hmc_read_checkpoint(const char *path, double ***A, long *trajectory);

/* ---- hmc.c ------------------------------------------------------------ */
double calc_gauge_action(double ***A);                                  /* :73-79   */
double gauge_step(double ***A, double ***mom, double eps);              /* :97-101  */
void fm_mul(_Complex double **v_in, _Complex double **v_out, double ***A);          /* :123-184 */
void fm_conjugate_mul(_Complex double **v_in, _Complex double **v_out, double ***A); /* :188-249 */
void fmdm_mul(_Complex double **v_in, _Complex double **v_out, double ***A);        /* :259-264 */
void fermion_matrix(_Complex double *M);                                /* :269-310 */
double complex determinant(void);                                       /* :314-336 */
void fmdm_invert_cg(_Complex double **v_in, _Complex double **v_out, double ***A);  /* :341-404 */
void fm_invert_cg(_Complex double **v_in, _Complex double **v_out, double ***A);    /* :408-414 */
double random_pseudofermion(_Complex double **v, double ***A);          /* :418-436 */
void stochastic_vector(_Complex double **v);                            /* :439-447 */
double pseudofermion_action(_Complex double **v, double ***A);          /* :451-463 */
double stochastic_md_action(_Complex double **v, double ***A);          /* :467-479 */
double random_momentum(double ***mom);                                  /* :483-499 */
double momentum_step(double ***mom, double ***A, _Complex double **psi,
                     _Complex double **st, double eps);                 /* :504-665 */
void update_gauge(double ***A);                                         /* :671-746 */
void update_puregauge_hb(double ***A);                                  /* :82-93   */

#ifdef __cplusplus
}
#endif
#endif
