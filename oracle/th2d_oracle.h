/*
 * th2d_oracle.h -- CPU restatement of the Thirring2D fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in thirring2d_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Conventions (shared with the reference hmc.c / vec_ops.c):
 *   site index      s = t*nx + x          (reference v[t][x])
 *   complex vector  double[2*V], interleaved (re, im)
 *   link field      double[2*V], A[2*s + dir], dir 0 = t, 1 = x (A[t][x][dir])
 *   eo index        (nx*t + x)/2          (fluctuation_determinant.c:181-192)
 *
 * Every function cites the reference file:line it restates.  The "parity
 * pinned by" line says which independent check in tests/ pins it.
 */
#ifndef TH2D_ORACLE_H
#define TH2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Lattice for the U(1) HMC family (hmc.c)                             */
/* ------------------------------------------------------------------ */
typedef struct {
  int nt, nx;        /* lattice extent                                  */
  int anti_t;        /* 1: antiperiodic in t (hmc.c:143-158)            */
  int anti_x;        /* 1: antiperiodic in x (hmc.c:165-180)        */
  double m, g, mu;   /* mass, coupling, chemical potential              */
  int nf;            /* Nf in the gauge action (hmc.c:28)               */
} orc_lattice;

/* MT19937, 1998 sgenrand seeding, blocks of 624 doubles consumed in
 * reverse (mersenne_inline.c:59-114, mersenne.h:11).  Written from
 * Matsumoto & Nishimura, ACM TOMACS 8 (1998) 3-30.
 * Parity pinned by: bitwise comparison with oracle/_ref (the reference's
 * own mersenne_inline.c compiled unmodified) and the golden draws. */
typedef struct {
  uint32_t mt[624];
  double out[624];
  int i;             /* number of unread doubles left in out[]          */
} orc_mt;

void   orc_mt_seed(orc_mt *st, long seed);
double orc_mt_next(orc_mt *st);
void   orc_mt_skip(orc_mt *st, long n);

/* Gaussian pair from two draws, x1 first (hmc.c:423-426).
 * Parity pinned by: moments and a KS test of N(0,1) (test_gauss_pair_moments). */
void orc_gauss_pair(orc_mt *st, double *c, double *s);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): the counter-based
 * generator behind the device noise fields (no reference counterpart; the
 * reference draws MT on the host).  Parity pinned by: the published
 * Random123 known-answer vectors (test_philox_known_answers). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The device's Gaussian pair for (seed, stream, index, counter): one
 * Philox block, two 53-bit uniforms in (0,1), Box-Muller.
 * Parity pinned by: moments and a KS test of N(0,1)
 * (test_philox_gauss_moments), the Philox known answers. */
void orc_philox_gauss2(uint64_t seed, uint32_t stream, uint32_t idx, uint32_t ctr, double *re,
                       double *im);

/* The device's uniform pair for (seed, stream, index, counter): the two
 * 53-bit uniforms in (0,1) of one Philox block (the draws of the device
 * pure-gauge sweeps).  Parity pinned by: the Philox known answers, and the
 * uniform moments in test_philox_uniform_moments. */
void orc_philox_uniform2(uint64_t seed, uint32_t stream, uint32_t idx, uint32_t ctr, double *u0,
                         double *u1);

/* Staggered phase eta_dir(t,x): eta_t = (-1)^x, eta_x = 1 (hmc.c:914-921) */
int orc_eta(int t, int x, int dir);

/* chi = M psi  (hmc.c:123-184), literal per-site form with sin/cos.
 * Parity pinned by: dense element test vs orc_fermion_matrix, free-field
 * plane-wave eigenvalues, gauge covariance. */
void orc_fm_mul(const orc_lattice *L, const double *A, const double *in, double *out);

/* chi = M^dagger psi, the true conjugate transpose that fm_conjugate_mul
 * (hmc.c:188-249) is meant to be (the reference body equals fm_mul).  Parity pinned by: <w,Mv> = <M^dag w, v> and M^dag = 2m - M at
 * mu = 0. */
void orc_fm_dag_mul(const orc_lattice *L, const double *A, const double *in, double *out);

/* Dense V x V complex M, column-major M[row + col*V] (hmc.c:269-310). */
void orc_fermion_matrix(const orc_lattice *L, const double *A, double *M);

/* Complex partial-pivot LU; log|det| and arg(det) (hmc.c:314-336, but
 * accumulated as logs, so that large orders neither overflow nor underflow).  Returns 0, or k>0 for a zero pivot.
 * Parity pinned by: free-field closed form, 2x2/3x3 explicit formulas. */
int orc_zlogdet(int n, double *M /* col-major, overwritten */, double *logabs, double *arg);

/* CG on M^dag M, x0 = 0 (hmc.c:341-404) with the true adjoint.
 * stop: rr < tol_abs_rr (reference CG_ACCURACY, hmc.c:34,381) or
 *       rr < tol_rel^2 * rr0 when tol_rel > 0.
 * Returns iterations; *status = 0 ok, 1 max_iter, 2 divergence (hmc.c:383).
 * Parity pinned by: true residual check, A = 0 Fourier solution. */
int orc_cg_mdm(const orc_lattice *L, const double *A, const double *b, double *x,
               double tol_abs_rr, double tol_rel, int max_iter, int *status);

/* Noise (hmc.c:418-499).  phys_norm = 0 reproduces the reference
 * normalisation (Re,Im ~ N(0,1)); 1 scales by 1/sqrt(2).
 * random_pseudofermion: returns sum |xi|^2, writes xi and psi = M^dag xi. */
double orc_random_pseudofermion(const orc_lattice *L, orc_mt *st, const double *A,
                                double *xi, double *psi, int phys_norm);
void   orc_stochastic_vector(const orc_lattice *L, orc_mt *st, double *v, int phys_norm);
double orc_random_momentum(const orc_lattice *L, orc_mt *st, double *mom, int phys_norm);

/* Actions (hmc.c:73-79, 451-479). */
double orc_gauge_action(const orc_lattice *L, const double *A);
double orc_pseudofermion_action(const orc_lattice *L, const double *A, const double *psi,
                                double tol_abs_rr, double tol_rel, int max_iter, int *iters);
double orc_stochastic_md_action(const orc_lattice *L, const double *A, const double *st);

/* A += 2 eps mom (hmc.c:97-101). */
void orc_gauge_step(const orc_lattice *L, double *A, const double *mom, double eps);

/* Fermion + gauge + stochastic force, mom -= eps dS/dA (hmc.c:504-665).
 * If force_out != NULL it receives dS/dA (without eps).
 * Parity pinned by: central finite differences of the actions. */
int orc_momentum_step(const orc_lattice *L, double *mom, const double *A, const double *psi,
                      const double *st, double eps, double tol_abs_rr, double tol_rel,
                      int max_iter, double *force_out);

/* One HMC trajectory (hmc.c:671-746).  Draw order: pseudofermion,
 * momentum, stochastic vector, then one accept draw (hmc.c:695-699,738).
 * Writes the action parts to acts[0..7] = {Sg0,Spf0,Smd0,Smom0,Sg1,Spf1,Smd1,Smom1}.
 * Returns 1 if accepted (A updated), 0 otherwise; *dS receives dH. */
int orc_update_gauge(const orc_lattice *L, orc_mt *st, double *A, int nsteps, double traj_len,
                     double tol_abs_rr, double tol_rel, int max_iter, int phys_norm,
                     double *dS, double *acts, int *cg_iters_total);

/* One pure-gauge Metropolis sweep (hmc.c:82-93). */
void orc_puregauge_sweep(const orc_lattice *L, orc_mt *st, double *A);

/* ------------------------------------------------------------------ */
/* Fermion-bag family (Thirring.h, vec_ops.c, fermion_matrix.c,        */
/* full_determinant.c, fluctuation_determinant.c, tests/thirring_brute.c)*/
/* ------------------------------------------------------------------ */
#define ORC_BC_ANTISYMMETRIC 0   /* Thirring.h:27 */
#define ORC_BC_OPENX         1   /* Thirring.h:29, tests/thirring_brute.h:28 */

typedef struct {
  int nt, nx;
  int bc;                 /* ORC_BC_*                                          */
  int with_mass_monomers; /* 1: no m on the diagonal of fM_index (Thirring.h:49) */
  double m, mu;
} orc_bag_lattice;

/* Matrix element from (t1,x1) to (t2,x2) (fermion_matrix.c:28-64 for
 * ANTISYMMETRIC, :152-182 for OPENX, including the NT==2/NX==2 cases). */
double orc_fM_index(const orc_bag_lattice *B, int t1, int x1, int t2, int x2, double mu);

/* Masked real apply, argument order (out, in) as the reference
 * (vec_ops.c:96-172).  field[s] != 0 marks an occupied site. */
void orc_fM(const orc_bag_lattice *B, const int *field, double *chi, const double *psi);
void orc_fM_transpose(const orc_bag_lattice *B, const int *field, double *chi, const double *psi);

/* CG on M^T M after b = M^T s: solves M x = s (vec_ops.c:261-321).
 * tol_abs_rr = CG_ACCURACY.  Returns iterations; *status as orc_cg_mdm
 * (on divergence x is filled with 1e50, vec_ops.c:292-296). */
int orc_cg_propagator(const orc_bag_lattice *B, const int *field, double *x, const double *src,
                      double tol_abs_rr, int max_iter, int *status);

/* Real partial-pivot LU: log|det| and sign (full_determinant.c:171-175
 * semantics, accumulated as logs).  M row-major or col-major (det is the same).
 * Returns 0 or k>0 at an exactly zero pivot (then sign = 0). */
int orc_dlogdet(int n, double *M, double *logabs, int *sign);

/* In-place inverse of a real n x n matrix (row-major) via LU. */
int orc_dinverse(int n, double *M);

/* The dense reduced matrix over the unoccupied sites in natural order,
 * M[i*n + j] = fM_index(free_i, free_j, mu) (full_determinant.c:157-167);
 * M holds n_free^2 doubles.  Returns n_free.
 * Parity pinned by: equal to the rows and columns of orc_fM applied to unit
 * vectors (test_reduced_matrix_matches_fM). */
int orc_reduced_matrix(const orc_bag_lattice *B, const int *field, double mu, double *M);

/* Dense reduced matrix over unoccupied sites, its log|det| and sign
 * (full_determinant.c:144-186).  Returns n_free. */
int orc_determinant_mu(const orc_bag_lattice *B, const int *field, double mu,
                       double *logabs, int *sign);

/* Free propagator blocks in the reference layout (fermion_matrix.c:191-351,
 * tests/thirring_brute.c:190-264): Dinv[e*V/2 + o] = (D^-1)_{e,o} and
 * Dinv[V^2/4 + o*V/2 + e] = (D^-1)_{o,e} (fermion_matrix.c:237-238,
 * 290-291), with D the unmasked matrix of
 * fM_index.  Computed by dense LU inversion of the two eo blocks. */
int orc_calc_Dinv(const orc_bag_lattice *B, double *Dinv);

/* det_e * det_o of the occupied-site propagator blocks
 * (fluctuation_determinant.c:203-237, check_det :880-970,
 * tests/thirring_brute.c:312-337).  Returns the product as log|.| and sign;
 * *n_out receives n. Returns 0 ok, 1 if the site counts do not match
 * (brute sets det = 0), 2 if an LU hits a zero pivot. */
int orc_bag_weight(const orc_bag_lattice *B, const int *field, const double *Dinv,
                   double *logabs, int *sign, int *n_out);

/* Exact enumeration (tests/thirring_brute.c:456-494): accumulates
 * Z[nl][nm] += det and Zabs[nl][nm] += |det| over every link/monomer
 * configuration.  Z, Zabs have (V/2+1)*(V+1) entries, row nl. Returns the
 * number of configurations visited. */
long orc_brute_enumerate(const orc_bag_lattice *B, double *Z, double *Zabs);

/* The worm estimate of the susceptibility (measure_susceptibility,
 * fermionbag.c:450-590) on the antisymmetric lattice: n_attempts worms,
 * each opening a random link into two source monomers and moving one of
 * them until it closes; field (natural order, int32) is updated in place,
 * n_links / n_monomers / sign carry the reference's globals (sign is
 * current_sign: the sign of det(D_FF) det(D)), and the draws come from the
 * MT stream in the reference's order.  With mass monomers the live moves
 * are the link flip, the monomer swap, a new link (det_added_link, weight U)
 * and a link removal (det_removed_link / U); without, the hop branch only
 * draws (its ratio vanishes: the neighbour has the other parity).  Ratios
 * come from orc_determinant_mu.  Returns the summed signs (`steps`);
 * *value = steps * n_links / (4 n_attempts V), the printed number.
 * Parity pinned by: field invariants (every link has its two ends, the
 * sources close, n_links counts the links) and the ratios' full
 * recomputes, test_worm_invariants. */
long orc_worm(const orc_bag_lattice *B, int *field, orc_mt *st, double U, int n_attempts,
              int *n_links, int *n_monomers, int *sign, double *value);

#ifdef __cplusplus
}
#endif
#endif
