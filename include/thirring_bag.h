/*
 * thirring_bag.h -- the reference's fermion-bag interface (Thirring.h:84-107,
 * fermionbag.c:49-62) with its own function signatures, implemented over
 * libth2d (libthirring_bag.so).
 *
 * A program written against the reference's fermion-bag modules (vec_ops.c,
 * fermion_matrix.c, full_determinant.c / fluctuation_determinant.c) links
 * this library instead of those files.  Its globals keep their meaning and
 * are read or written at every call: the configuration `field` (int
 * field[t][x]: 0 empty, 1 monomer, 2+dir link, SOURCE_MONOMER = 6;
 * Thirring.h:33-39), m and mu, the propagator Dinv (written by calc_Dinv),
 * n_monomers, n_links, current_sign and max_changes.  The library carries
 * weak definitions of all of them (and a no-op update_linklists, which the
 * reference's hooks call), so the program's own definitions take precedence.
 *
 * The compile-time choices of Thirring.h (NT, NX, ANTISYMMETRIC or OPENX,
 * WITH_MASS_MONOMERS) are set once with thirring_bag_init().
 *
 * Determinant ratios come from the device-resident inverse B = (D_FF)^-1
 * of the accepted configuration with rank-k updates (th2d_bag_*); the
 * hooks commit a change and refresh B from scratch every max_changes
 * accepted changes (as update_background does, fluctuation_determinant.c:
 * 305-545).  det_moved_monomer between sites of opposite parity returns 0
 * (fluctuation_determinant.c:785-789).
 */
#ifndef THIRRING_BAG_H
#define THIRRING_BAG_H

#ifdef __cplusplus
extern "C" {
#endif

/* globals of fermionbag.c (weak in the library) */
extern int **field;
extern double m, mu, U;  /* U: the coupling, named in the checkpoint file */
extern double *Dinv;
extern int n_monomers, n_links, current_sign, max_changes;
void update_linklists(void);

/* ---- set-up replacing the compile-time constants ---------------------- */
/* bc: 0 ANTISYMMETRIC, 1 OPENX; with_mass_monomers as WITH_MASS_MONOMERS */
int thirring_bag_init(int nt, int nx, int bc, int with_mass_monomers);
void thirring_bag_finalize(void);
/* Status of the last call (TH2D_* code; 0 = OK). */
int thirring_bag_last_status(void);

/* ---- fermion_matrix.c ------------------------------------------------- */
double fM_index(int t1, int x1, int t2, int x2, double mu);             /* :28-64     */
void calc_Dinv(void);  /* :191-351; the on-disk cache of :206-220,342-346 */

/* ---- fermionbag.c: configuration checkpoint ------------------------------ */
void write_config(void);                                                /* :120-138   */
void read_config(char *filename);                                       /* :141-161   */

/* ---- vec_ops.c (argument order out, in as the reference) -------------- */
void fM(double **chi, double **psi);                                    /* :96-131    */
void fM_transpose(double **chi, double **psi);                          /* :135-172   */
void cg_propagator(double **propagator, double **source);               /* :311-321   */

/* ---- full_determinant.c / fluctuation_determinant.c ------------------- */
double determinant_mu(double mu);                                       /* :144-186   */
double determinant(void);                                               /* :188-190   */
void update_background(void);                                           /* :126-301   */
double det_added_link(int t, int x, int t2, int x2);                    /* :664-708   */
double det_removed_link(int t, int x, int t2, int x2);                  /* :715-760   */
double det_moved_monomer(int t, int x, int t2, int x2);                 /* :764-871   */
void new_link(int t, int x, int nu);                                    /* :305-345   */
void new_monomer(int t1, int x1, int t2, int x2);                       /* :346-387   */
void removed_link(int t, int x, int nu);                                /* :388-434   */
void removed_monomer(int t1, int x1, int nu);                           /* :435-479   */
void moved_source_monomer(int t2, int x2, int t4, int x4);              /* :484-524   */
void not_moved_source_monomer(int t2, int x2, int t4, int x4);          /* :526-545   */

/* ---- fermionbag.c measurement over these ratios and hooks -------------- */
/* The worm estimate of the susceptibility (fermionbag.c:450-590) on the
 * program's field, drawing from the MT stream (mersenne.h) in the
 * reference's order: returns the summed signs (`steps`) and, in *value,
 * steps * n_links / (4 n_attempts V), the number measure_susceptibility
 * prints.  Antisymmetric lattices only. */
long thirring_bag_worm(int n_attempts, double *value);
/* 20 worms, printing "Susceptibility %g" (weak: a program's own
 * measure_susceptibility takes precedence). */
void measure_susceptibility(void);
/* reference RNG interface (mersenne.h), weak here as in libthirring_hmc */
extern int mersenne_i;
extern double mersenne_array[624];
void seed_mersenne(long seed);
double mersenne_generate(void);

#ifdef __cplusplus
}
#endif
#endif
