/*
 * th2d_bag.h -- C ABI of the fermion-bag part of libth2d (gfx950):
 * masked real staggered operator and its CG (vec_ops.c), dense batched LU,
 * log-determinants and inverses (full_determinant.c, fermion_matrix.c),
 * and the determinant-ratio state of the fermion-bag update
 * (fluctuation_determinant.c).
 *
 * Conventions: the bag field is int32[V] in natural order s = t*nx + x with
 * 0 = empty, 1 = monomer, 2+dir = link (Thirring.h:34-39); real vectors are
 * double[V] in natural order.  Dense matrices are row-major.
 */
#ifndef TH2D_BAG_H
#define TH2D_BAG_H

#include "th2d.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TH2D_BC_ANTISYMMETRIC 0 /* Thirring.h:27 */
#define TH2D_BC_OPENX 1         /* Thirring.h:29, tests/thirring_brute.h:28 */

/* Lattice of the fermion-bag family (Thirring.h:14-50, runtime values). */
typedef struct {
  int nt, nx;              /* even extents, >= 4                             */
  int bc;                  /* TH2D_BC_*                                       */
  int with_mass_monomers;  /* 1: no mass on the diagonal (Thirring.h:49)     */
  double m, mu;
} th2d_bag_geom;

/* ---- masked operator and solver --------------------------------------- */
/* out = fM in (transpose = 0) or fM^T in (vec_ops.c:96-172): rows of
 * occupied sites are identity rows, hops into occupied sites are dropped.
 * nrhs vectors are stored one after another (stride V). */
int th2d_bag_apply(th2d_ctx *ctx, const th2d_bag_geom *g, const int *field, double *out,
                   const double *in, int transpose, int nrhs);

/* Solve fM x = src for nrhs right-hand sides by CG on fM^T fM after
 * b = fM^T src (cg_propagator / cg_MdM, vec_ops.c:261-321), every RHS
 * with its own alpha, beta and stop r.r < tol_abs_rr.  On divergence
 * (r.r / r0.r0 > 1e10) that RHS is filled with 1e50 as the reference does.
 * iters_host / status_host (nrhs each, may be NULL) receive per-RHS
 * iteration counts and TH2D_OK / TH2D_ENOCONV / TH2D_EBREAKDOWN. */
int th2d_bag_cg(th2d_ctx *ctx, const th2d_bag_geom *g, const int *field, double *x,
                const double *src, int nrhs, double tol_abs_rr, int max_iter, int *iters_host,
                int *status_host);

/* ---- dense batched linear algebra ------------------------------------- */
/* A batch of row-major square matrices in one device buffer: matrix b has
 * order n[b], starts at element off[b]; its pivots start at poff[b].  The
 * three arrays are device arrays of length count; max_n is host-side. */
typedef struct {
  int count;
  int max_n;
  const long *off;
  const int *n;
  const long *poff;
} th2d_batch;

/* In-place LU with partial pivoting (dgetrf semantics: row i swapped with
 * piv[i] >= i, 0-based), FP64 or FP32 (TH2D_PREC_F64 / TH2D_PREC_F32); the
 * trailing updates run on MFMA.  info[b] = 0 or j+1 at the first exactly
 * zero pivot (device array). */
int th2d_lu_batched(th2d_ctx *ctx, const th2d_batch *b, void *A, int precision, int *piv,
                    int *info);
/* log|det| and sign(det) from the factors (device arrays of length count);
 * a zero pivot gives -inf and sign 0 (determinant_mu, full_determinant.c:
 * 171-175, accumulated as logs). */
int th2d_lu_logdet(th2d_ctx *ctx, const th2d_batch *b, const void *LU, int precision,
                   const int *piv, const int *info, double *logabs, int *sign);
/* log|det A| of the FP32 factors of A (th2d_lu_batched, TH2D_PREC_F32),
 * corrected in FP64 (determinant_mu, full_determinant.c:144-186, with the
 * factorisation in FP32 on the matrix cores): with B = P L U the product of
 * the FP32 factors and E = A - B formed in FP64,
 *   log|det A| = log|det B| + tr(B^-1 E) + O(|B^-1 E|^2),
 * where tr(B^-1 E) takes B^-1 from the FP32 factors.  A holds the FP64
 * matrices (device, batch layout, unchanged); sign as th2d_lu_logdet. */
int th2d_lu_logdet_refined(th2d_ctx *ctx, const th2d_batch *b, const double *A, const float *LU32,
                           const int *piv, const int *info, double *logabs, int *sign);
/* X = A^{-1} from the factors (dgetri semantics); X uses the batch layout. */
int th2d_lu_inverse(th2d_ctx *ctx, const th2d_batch *b, const void *LU, int precision,
                    const int *piv, void *X);

/* The same inverse without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * permutation scratch is a static device array of 4,194,304 ints (16 MiB,
 * one copy per GPU, held until the process ends): count * max_n must not
 * exceed that (else TH2D_EARG), and a context must have been created on
 * the device.  Every call on a GPU shares that array, so calls that
 * overlap on one GPU race.  Same results as th2d_lu_inverse. */
int th2d_lu_inverse_noalloc(th2d_ctx *ctx, const th2d_batch *b, const void *LU, int precision,
                            const int *piv, void *X);

/* C += alpha A B for a batch of jobs (element offsets into one buffer). */
typedef struct {
  long a, b, c;
  int lda, ldb, ldc;
  int M, N, K;
} th2d_gemm_job;
int th2d_gemm_batched(th2d_ctx *ctx, void *base, int precision, const th2d_gemm_job *jobs_dev,
                      int count, int max_m, int max_n, double alpha);
/* Dense kernels on MFMA (1, default) or on the vector ALUs (0). */
int th2d_set_dense_mfma(int on);
/* LU panel factorisation (process-wide): 1 (default), when every order of
 * the batch is at most 3072, factors blocks of 128 columns as 16-column
 * panels held in the registers of one block per matrix (256 to 1024
 * threads), with the row interchanges applied to the other columns
 * afterwards and one trailing update per block; 0 always uses the
 * 32-column panel that swaps whole rows as it goes, with a trailing update
 * per panel.  Both give the factors of partial pivoting with whole-row
 * interchanges (LAPACK getrf); the switch is there to compare them.
 * Replaces no reference function. */
int th2d_set_lu_panel(int reg);
/* Look-ahead of the blocked register-panel LU (process-wide; 1 default,
 * 0 off): a block's panels run on a stream confined to 64 CUs (8 per XCD)
 * while the rest of the previous block's trailing update runs on the other
 * CUs; the block's row interchanges outside its own columns wait until that
 * update is done.  The factors and pivots are the same either way.
 * Replaces no reference function. */
int th2d_set_lu_lookahead(int on);
int th2d_copy_f64_to_f32(th2d_ctx *ctx, float *dst, const double *src, long n);

/* ---- matrices of bag configurations ---------------------------------- */
/* Dense V x V masked matrix of `field` (identity rows and columns at
 * occupied sites, fM_index elements between free sites; vec_ops.c:107-131,
 * fermion_matrix.c:28-182) for a batch of configurations: config b has
 * geometry geoms_dev[b], field at field_off[b], matrix at the batch's off[b]
 * (FP64 or FP32). */
int th2d_bag_build_masked(th2d_ctx *ctx, const th2d_bag_geom *geoms_dev, const int *fields,
                          const long *field_off, const th2d_batch *b, void *A, int precision);

/* ---- determinant-ratio state of the fermion-bag update ---------------- */
/* Keeps B = (D_FF)^{-1}, the inverse of the fM_index matrix restricted to
 * the free sites F of the accepted configuration (stored V x V with
 * identity at occupied sites), and log|det D_FF| with its sign.  The
 * reference's bag weight det_e det_o (fluctuation_determinant.c:203-237)
 * equals det(D_FF)/det(D) by Jacobi's complementary-minor identity, so all
 * of its determinant ratios follow from B:
 *   occupy sites S:  ratio = det(B_SS)
 *   free sites S:    ratio = det(D_SS - D_SF B D_FS)
 * and an accepted change is a rank-|S| (Woodbury / Schur) update of B. */
typedef struct th2d_bag th2d_bag;

int th2d_bag_create(th2d_ctx *ctx, const th2d_bag_geom *g, th2d_bag **out);
int th2d_bag_destroy(th2d_bag *bag);
/* Load a configuration (host int32[V]) and refactor B from scratch by LU
 * (update_background, fluctuation_determinant.c:126-301). */
int th2d_bag_set_field(th2d_bag *bag, const int *field_host);
int th2d_bag_refresh(th2d_bag *bag);
/* log|det D_FF|, its sign, and log|det D| of the empty configuration. */
int th2d_bag_logdet(th2d_bag *bag, double *logabs_ff, int *sign_ff, double *logabs_d, int *sign_d);
/* Ratio det(D_F'F') / det(D_FF) for F' = F minus `occupy` plus `free_`
 * (site indices, host arrays, at most 4 each); returns it in *ratio
 * (signed). */
int th2d_bag_ratio(th2d_bag *bag, const int *occupy, int n_occupy, const int *free_, int n_free,
                   double *ratio);
/* Apply the change of th2d_bag_ratio to B and to the log-determinant and
 * set the field of the changed sites (values: field_vals for `occupy`,
 * 0 for `free_`). */
int th2d_bag_accept(th2d_bag *bag, const int *occupy, int n_occupy, const int *field_vals,
                    const int *free_, int n_free);
/* Device pointer of B and of the field (for tests and batched use). */
int th2d_bag_views(th2d_bag *bag, double **B, int **field, int *V);

/* ---- many configurations on the device: the update loop of fermionbag.c
 * (add_link / remove_link, :284-379) with the ratios and accept hooks of
 * fluctuation_determinant.c:305-871 for a batch of independent bag
 * matrices of any sizes.  Each matrix keeps B = (D_FF)^{-1}, its field and
 * log|det D_FF| on the device.  One step proposes one change per matrix
 * from device (Philox) draws -- with probability 0.3 a monomer on an empty
 * site or off a monomer (rank 1), otherwise, in direction t or x, a link
 * on two empty sites, off a link, or a monomer hop onto an empty
 * neighbour (rank 2) -- forms its ratio det(D_F'F')/det(D_FF) from B and
 * accepts it when a second draw u < |ratio|; accepted changes are applied
 * to every matrix in the same launches.  No host synchronisation happens
 * between steps. */
typedef struct th2d_bag_batch th2d_bag_batch;

/* One proposal of th2d_bag_batch_sweep: the sites occupied (with their new
 * field values) and freed (-1 for unused slots), the ratio (0 when the
 * site had no legal change or the ratio vanished) and the decision. */
typedef struct {
  int no, nr;
  int occ[2], fr[2], vals[2];
  int accepted, pad;
  double ratio;
} th2d_bag_move;

/* count configurations (1..65535): geoms (host) and their fields
 * concatenated (host int32, natural order, V_b values each); B of each by
 * batched LU. */
int th2d_bag_batch_create(th2d_ctx *ctx, int count, const th2d_bag_geom *geoms, const int *fields,
                          th2d_bag_batch **out);
int th2d_bag_batch_destroy(th2d_bag_batch *b);
/* nsteps proposals per matrix; step k draws with counter `counter + k`.
 * trace: NULL or a device array of nsteps * count records (step-major). */
int th2d_bag_batch_sweep(th2d_bag_batch *b, int nsteps, unsigned long long seed,
                         unsigned long long counter, th2d_bag_move *trace);
/* B, log|det D_FF| and the sign of every matrix again from scratch. */
int th2d_bag_batch_refresh(th2d_bag_batch *b);
/* Host copies (any may be NULL): log|det D_FF| and sign per matrix, the
 * fields (concatenated), the accepted changes per matrix.  Synchronises. */
int th2d_bag_batch_state(th2d_bag_batch *b, double *logabs, int *sign, int *fields, long *accepted);

/* Free propagator D^{-1} in the reference's eo block layout
 * (calc_Dinv, fermion_matrix.c:191-351): Dinv[e*V/2 + o] = (D^{-1})_{e,o},
 * Dinv[V^2/4 + o*V/2 + e] = (D^{-1})_{o,e} (device, 2 (V/2)^2 doubles),
 * computed by dense LU inversion on the device. */
int th2d_bag_calc_dinv(th2d_ctx *ctx, const th2d_bag_geom *g, double *Dinv);

/* Bag weights det_e det_o (fluctuation_determinant.c:203-237,
 * tests/thirring_brute.c:273-345) of `count` configurations given as uint8
 * fields (natural order, device) from Dinv; w[c] (device) is 0 for
 * configurations the reference gives weight 0.  V/2 <= 16. */
int th2d_bag_small_weights(th2d_ctx *ctx, const th2d_bag_geom *g, const double *Dinv,
                           const unsigned char *cfg, long count, double *w);

/* Matrix element fM_index(t1,x1,t2,x2,mu) (fermion_matrix.c:28-64,
 * 152-182) of the lattice g at the given mu (host, no device work). */
int th2d_bag_matrix_element(const th2d_bag_geom *g, int t1, int x1, int t2, int x2, double mu,
                            double *out);

/* log|det| and sign of the matrix restricted to the free sites of a
 * configuration (determinant_mu, full_determinant.c:144-186; field host
 * int32[V]) by LU on the device; sign 0 and -inf at an exact zero pivot. */
int th2d_bag_determinant(th2d_ctx *ctx, const th2d_bag_geom *g, const int *field_host,
                         double *logabs, int *sign);

/* Exact enumeration of all link/monomer configurations (tests/
 * thirring_brute.c:402-494): Z[nl*(V+1) + nm] += w and Zabs[...] += |w|
 * over every configuration with nl links and nm monomers, w = det_e det_o
 * evaluated on the device (host arrays of (V/2+1)(V+1) doubles; V/2 <= 16).
 * *visited receives the number of configurations. */
int th2d_bag_brute_enumerate(th2d_ctx *ctx, const th2d_bag_geom *g, double *Z, double *Zabs,
                             long *visited);

#ifdef __cplusplus
}
#endif
#endif
