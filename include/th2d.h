/*
 * th2d.h -- C ABI of libth2d, the MI355X (gfx950) implementation of the
 * Thirring2D fermion hot path.
 *
 * Plain C types only: device pointers are `void *` / `double *`, sizes are
 * ints, status codes are ints.  No torch, no C++ in the signatures.
 *
 * Device storage (all entry points below):
 *   - lattice extents nt, nx are even; V = nt*nx, Vh = V/2;
 *   - "eo order": site (t,x) lives at index  p*Vh + (t*nx + x)/2  with
 *     p = (t+x)&1, i.e. all even sites in the reference's eo order
 *     (fluctuation_determinant.c:181-192), then all odd sites;
 *   - complex vectors are interleaved (re,im) pairs: double[2V] (FP64) or
 *     float[2V] (FP32);
 *   - the link angles A (hmc.c:47) are two double[V] arrays in eo order,
 *     direction t first then x, packed as double[2V];
 *   - "folded links" K are the per-site forward hopping factors
 *     K_mu(x) = 1/2 eta_mu(x) s_mu(x) e^{i A_mu(x)} (s = -1 on a hop through
 *     an antiperiodic boundary), complex, packed [dir][V] in eo order, kept
 *     in FP64 (double[4V]) and optionally FP32 (float[4V]).
 *
 * Every call is asynchronous on the context's HIP stream unless it
 * returns a host value; a non-zero return is a TH2D_E* code.
 */
#ifndef TH2D_H
#define TH2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TH2D_OK 0
#define TH2D_EARG (-1)       /* bad argument (size, null, odd extent, ...)   */
#define TH2D_EHIP (-2)       /* a HIP runtime call failed                    */
#define TH2D_ENOCONV (-3)    /* solver hit max_iter before the tolerance     */
#define TH2D_EBREAKDOWN (-4) /* solver breakdown: p.Ap <= 0, NaN, divergence */
#define TH2D_ESINGULAR (-5)  /* LU met an exactly zero pivot                 */
#define TH2D_ENOMEM (-6)     /* device allocation failed                     */
#define TH2D_ELAUNCH (-7)    /* cooperative launch does not fit the device   */
#define TH2D_ESYNC (-8)      /* a bounded wait between workgroups timed out  */

#define TH2D_PREC_F64 0
#define TH2D_PREC_F32 1
#define TH2D_PREC_MIXED 2    /* FP32 inner solve, FP64 reliable updates      */

typedef struct th2d_ctx th2d_ctx;

/* Lattice descriptor of the U(1) staggered operator (hmc.c:13-40; the
 * reference fixes these at compile time, here they are runtime values). */
typedef struct {
  int nt, nx;      /* even extents                                          */
  int anti_t;      /* 1: antiperiodic in t (hmc.c:143-158)                   */
  int anti_x;      /* 1: antiperiodic in x (hmc.c:165-180)                   */
  double m;        /* mass                                                  */
  double mu;       /* chemical potential, weights e^{+-mu} on t hops         */
} th2d_geom;

/* Solver controls and diagnostics (hmc.c:341-404, vec_ops.c:261-307). */
typedef struct {
  double tol_rel;     /* stop at ||b - A x|| <= tol_rel ||b|| (true residual)  */
  double tol_abs_rr;  /* or at r.r < tol_abs_rr (reference CG_ACCURACY); 0 off */
  int max_iter;       /* cap on inner iterations (all restarts together)       */
  int precision;      /* TH2D_PREC_F64 or TH2D_PREC_MIXED                       */
  double delta;       /* mixed: inner FP32 solve reduces the residual by this   */
} th2d_cg_params;

typedef struct {
  int iterations;        /* inner iterations, all restarts                     */
  int restarts;          /* FP64 residual recomputations after the first       */
  int status;            /* TH2D_OK / TH2D_ENOCONV / TH2D_EBREAKDOWN            */
  double true_rel_resid; /* ||b - A x|| / ||b|| recomputed in FP64 at the end   */
  double rr;             /* final recursive r.r (reference stopping quantity)   */
} th2d_cg_info;

/* ---- context, memory, timing ------------------------------------------ */
int th2d_ctx_create(th2d_ctx **ctx, void *hip_stream /* NULL: default stream */);
int th2d_ctx_destroy(th2d_ctx *ctx);
const char *th2d_strerror(int code);
int th2d_device_count(int *n);
int th2d_device_props(int *num_cu, int *num_xcd, size_t *lds_per_cu, char *name, int name_len);

int th2d_malloc(void **dptr, size_t bytes);
int th2d_free(void *dptr);
int th2d_memset(th2d_ctx *ctx, void *dptr, int value, size_t bytes);
int th2d_h2d(th2d_ctx *ctx, void *dst, const void *src, size_t bytes);
int th2d_d2h(th2d_ctx *ctx, void *dst, const void *src, size_t bytes);
int th2d_d2d(th2d_ctx *ctx, void *dst, const void *src, size_t bytes);
int th2d_sync(th2d_ctx *ctx);

/* Event timing on the context stream: start, stop, elapsed ms. */
int th2d_timer_start(th2d_ctx *ctx);
int th2d_timer_stop(th2d_ctx *ctx, float *ms);

/* Kernel-level timing: when enabled, every launch of the dominant kernel
 * family (the CG solver kernel) is bracketed by events on the stream it is
 * launched on and accumulated; read with th2d_kernel_time. */
int th2d_kernel_timing(th2d_ctx *ctx, int enable);
int th2d_kernel_time(th2d_ctx *ctx, double *total_ms, long *launches, double *total_bytes);

/* ---- layout ------------------------------------------------------------ */
/* complex vector: natural order (s = t*nx + x) <-> eo order, FP64 both. */
int th2d_vec_to_eo(th2d_ctx *ctx, const th2d_geom *g, double *eo, const double *nat);
int th2d_vec_from_eo(th2d_ctx *ctx, const th2d_geom *g, double *nat, const double *eo);
/* links/momenta: natural A[V][2] <-> eo packed [dir][V]. */
int th2d_links_to_eo(th2d_ctx *ctx, const th2d_geom *g, double *eo, const double *nat);
int th2d_links_from_eo(th2d_ctx *ctx, const th2d_geom *g, double *nat, const double *eo);
/* FP64 <-> FP32 copies of n complex numbers. */
int th2d_vec_f64_to_f32(th2d_ctx *ctx, float *dst, const double *src, long n);
int th2d_vec_f32_to_f64(th2d_ctx *ctx, double *dst, const float *src, long n);

/* ---- gauge field -------------------------------------------------------- */
/* K = 1/2 eta s e^{iA}; K32 may be NULL (hmc.c:140-180 recompute sin/cos
 * per apply; this build folds them once per field update). */
int th2d_fold_links(th2d_ctx *ctx, const th2d_geom *g, double *K, float *K32, const double *A);
/* A += 2 eps mom, then refold K (hmc.c:97-101). */
int th2d_gauge_step(th2d_ctx *ctx, const th2d_geom *g, double *A, const double *mom, double eps,
                    double *K, float *K32);

/* ---- operator ----------------------------------------------------------- */
/* out = M in (dagger = 0, hmc.c:123-184) or out = M^dag in (dagger = 1, the
 * true adjoint of fm_mul; the reference's fm_conjugate_mul hmc.c:188-249
 * repeats fm_mul, defect D1).  precision TH2D_PREC_F64 (double vectors,
 * K) or TH2D_PREC_F32 (float vectors, K32).  nrhs vectors are stored one
 * after the other (stride 2V) and share the link reads. */
int th2d_dslash(th2d_ctx *ctx, const th2d_geom *g, void *out, const void *in, const void *K,
                int dagger, int precision, int nrhs);
/* The FP64 operator on every lattice of a batch in one launch: out[c] =
 * M in[c] (dagger = 0) or M^dag in[c] (dagger = 1) with links K[c], for
 * nch lattices of any sizes (geoms on the host, the pointer arrays on the
 * host, the vectors and links on the device, eo order as th2d_dslash).  The
 * reference applies fm_mul (hmc.c:123-184) to one lattice at a time; this
 * is that apply for many independent chains. */
int th2d_dslash_batched(th2d_ctx *ctx, int nch, const th2d_geom *geoms, double *const *out,
                        const double *const *in, const double *const *K, int dagger);

/* out = M^dag M in (FP64). */
int th2d_mdagm(th2d_ctx *ctx, const th2d_geom *g, double *out, const double *in, const double *K,
               double *tmp);

/* ---- solver ------------------------------------------------------------- */
/* x = (M^dag M)^{-1} b (fmdmm_invert_cg, hmc.c:341-404).  K must be folded;
 * K32 is required for TH2D_PREC_MIXED.  x and b are FP64 eo vectors.
 * info may be NULL. Returns info.status. */
int th2d_cg_mdagm(th2d_ctx *ctx, const th2d_geom *g, double *x, const double *b,
                  const double *K, const float *K32, const th2d_cg_params *p,
                  th2d_cg_info *info);

/* The same solve for nch independent lattices of any sizes in one batch:
 * every launch of the solver kernel serves all of them (per-lattice
 * scalars and stop decisions).  geoms, x, b, K, K32 and infos are host
 * arrays of length nch (x, b, K, K32 hold device pointers).  Returns the
 * first non-OK status, if any. */
int th2d_cg_mdagm_batched(th2d_ctx *ctx, int nch, const th2d_geom *geoms, double **x,
                          const double *const *b, const double *const *K,
                          const float *const *K32, const th2d_cg_params *p,
                          th2d_cg_info *infos);

/* Tuning of the solver and of the operator kernel: tile shape (tt x tx
 * sites per workgroup tile, one of 16x64 8x64 16x32 8x32 16x16 8x16 8x8
 * 4x8 4x4 2x4 2x2; 0 0 = the modelled default; a shape larger than a
 * lattice of the batch falls back to the default), a cap on the number of
 * workgroups of a solver launch (0 = one per tile), and the Dslash
 * variant (0 direct, 1 LDS-tiled). */
int th2d_set_cg_tile(th2d_ctx *ctx, int tt, int tx);
int th2d_set_cg_max_blocks(th2d_ctx *ctx, int n);
/* Measuring autotuner (0 off, default; 1 on): the first solve of a batch
 * signature (its distinct lattice shapes and size class) times a few
 * iterations with every tile shape that fits and keeps the fastest for
 * this context; a tile set with th2d_set_cg_tile takes precedence. */
int th2d_set_cg_autotune(th2d_ctx *ctx, int on);
/* Tile shape used by the last solve on this context. */
int th2d_cg_last_tile(th2d_ctx *ctx, int *tt, int *tx);
/* Resident FP32 solver (1 on, default; 0 off): an FP32 inner solve runs as
 * cooperative launches that keep every 64x64 region's vectors in registers,
 * one launch per group of chains that fits the device's resident
 * workgroups (one per CU on MI355X: 256 regions, e.g. 16 lattices of
 * 256^2).  A lattice the resident kernel does not take -- larger than one
 * XCD holds (32 regions), or cut into several regions with a clipped edge
 * region (a side above 64 that is not a multiple of 64) -- is solved by the
 * step kernel; a batch that mixes such lattices with others is split
 * between the two kernels.  The whole batch goes to the step kernel when a
 * cost model prices it faster, or when a tile is set with
 * th2d_set_cg_tile.  th2d_cg_last_resident tells which kernel the last
 * inner solve used: 0 the step kernel, 1 the resident kernel, 2 both (a
 * split batch), 3 the small-lattice kernel (FP64, th2d_set_cg_small). */
int th2d_set_cg_resident(th2d_ctx *ctx, int on);
/* Step-kernel launches as captured graphs (0 off, default; 1 on): the
 * step kernel's iterations are issued as replays of a hipGraph holding 8
 * launches, captured once per batch layout on a private stream and cached
 * on the context (at most 16).  A solve then runs a whole number of
 * replays; the launches after every chain has stopped find nothing to do.
 * Off by default: on MI355X the replays measured 1-6% slower than eager
 * launches (C4, batched and single-chain FP64 HMC), whose gaps the
 * kernels' own duration already hides.  Replaces no reference function. */
int th2d_set_cg_graph(th2d_ctx *ctx, int on);
/* Reliable updates of the mixed solver (0 off, default; 1 on).  Each FP64
 * pass replaces the FP32 inner solver's residual with the true residual.
 * Off, every pass restarts CG from p = 0 (defect correction).  On, a pass
 * run by the resident solver keeps its search direction p and
 * s = M^dag M p, and the next pass continues the same Krylov sequence
 * (residual replacement, with p.Ap of the first step recomputed), as long as
 * the FP32 recursive residual agrees with the true one at the end of the
 * pass (|r_true|^2 <= 1.1 |r_rec|^2).  With delta = 1e-4 this saves
 * iterations on well-conditioned systems (m = 0.2: 3-15%) and costs them on
 * ill-conditioned ones (m = 0.05: +13%), where the FP32 residual has drifted
 * by the end of a pass; hence off by default. */
int th2d_set_cg_reliable(th2d_ctx *ctx, int on);
int th2d_cg_last_resident(th2d_ctx *ctx, int *on);
/* Small lattices (1, default; 0 off): an FP64 inner solve of a batch whose
 * lattices have at most 2048 sites runs as one launch, one workgroup per
 * lattice, the whole CG loop inside it (vectors in registers, the operator
 * through LDS); th2d_cg_last_resident then reports 3.  Off, or with a tile
 * set by th2d_set_cg_tile, the step kernel solves them.  Replaces no
 * reference function. */
int th2d_set_cg_small(th2d_ctx *ctx, int on);
/* Regions of the resident kernel: 64 x 64 sites, or 32 x 64 for batches
 * whose lattices all cut so and whose twice as many regions still fit the
 * device in one launch (a single 256^2 chain then runs on 32 CUs instead
 * of 16).  rows = 0 (default) chooses per batch; 64 or 32 forces the shape
 * where the batch allows it.  Replaces no reference function. */
int th2d_set_cg_resident_rows(th2d_ctx *ctx, int rows);
/* Bounded waits of the resident solver.  Every wait between the regions
 * of a chain gives up after `polls` polls (polls < 0: the default, 2^20,
 * at most about 1 s) and ends every workgroup of the launch normally; the solve is
 * then redone by the step kernel from the same right-hand sides, and
 * th2d_cg_sync_fallbacks counts how often that happened on this context.
 * A small limit forces the fallback from the host (used by the tests).
 * Replaces no reference function (the reference's CG is serial). */
int th2d_set_cg_spin_limit(th2d_ctx *ctx, long polls);
int th2d_cg_sync_fallbacks(th2d_ctx *ctx, long *n);
/* Solver launches enqueued between two reads of the stop flags
 * (0 = the iterations the same restart pass took in the previous solve,
 * then 8 at a time). */
int th2d_set_cg_chunk(th2d_ctx *ctx, int n);
int th2d_set_dslash_variant(int variant);

/* ---- force -------------------------------------------------------------- */
/* mom -= eps * dS/dA for S = Nf/g sum(1-cos A) + psi^dag (M^dag M)^{-1} psi
 *                             + Re st^dag M^dag st
 * given chi = (M^dag M)^{-1} psi (momentum_step, hmc.c:504-665).  st may
 * be NULL (its force vanishes at mu = 0).  If force_out != NULL it receives
 * dS/dA in eo packed [dir][V] order and mom is left unchanged. */
int th2d_force(th2d_ctx *ctx, const th2d_geom *g, double *mom, const double *A, const double *K,
               const double *chi, const double *st, double eps, double gcoup, int nf,
               double *force_out);

/* ---- reductions (results on the host; these synchronise) --------------- */
int th2d_dot(th2d_ctx *ctx, long n_complex, const double *a, const double *b, double *re,
             double *im);
int th2d_norm2(th2d_ctx *ctx, long n_real, const double *a, double *out);
/* Nf/g sum(1 - cos A) (calc_gauge_action, hmc.c:73-79). */
int th2d_gauge_action(th2d_ctx *ctx, const th2d_geom *g, const double *A, double gcoup, int nf,
                      double *out);

/* ---- HMC ---------------------------------------------------------------- */
/* Parameters of one trajectory (update_gauge, hmc.c:671-746). */
typedef struct {
  double g;          /* coupling                                        */
  int nf;            /* Nf of the gauge action (hmc.c:28)                */
  int nsteps;        /* leapfrog steps (hmc.c:708)                       */
  double traj_len;   /* trajectory length (hmc.c:709)                     */
  int phys_norm;     /* 0: reference noise normalisation (Re, Im ~ N(0,1) each, hmc.c:418-436); 1: scaled by 1/sqrt(2) */
  th2d_cg_params cg; /* solver used for every inversion                 */
} th2d_hmc_params;

typedef struct {
  double dS;          /* H_end - H_start                              */
  double acts[8];     /* Sg0, Spf0, Smd0, Smom0, Sg1, Spf1, Smd1, Smom1  */
  int accepted;
  int cg_iterations;  /* inner iterations summed over the trajectory   */
  int cg_solves;
  double max_true_rel_resid;
  int status;
} th2d_hmc_result;

/* MD steps as graph replays (1, default; 0 off).  When every chain of a
 * trajectory has at most 2048 sites and the solves are FP64 (so that each
 * solve is one launch of the small-lattice kernel, th2d_set_cg_small), one
 * leapfrog step -- gauge step, solve, M chi, force, gauge step -- has no host
 * decision in it: it is captured once per trajectory as a hipGraph and
 * launched nsteps times, and the solver statistics are read once at the
 * end.  The results are those of the step-by-step launches.  Replaces no
 * reference function (the loop of hmc.c:708-716). */
int th2d_set_hmc_graph(th2d_ctx *ctx, int on);

/* One trajectory on device-resident fields.  A (eo, FP64) is the current
 * gauge field and is replaced on accept.  The noise comes either from the
 * host (xi, mom, st: natural-order host arrays drawn from the reference's
 * MT stream; u = the accept draw) or, when xi == NULL, from the device
 * counter-based generator keyed by (seed, counter). */
int th2d_hmc_trajectory(th2d_ctx *ctx, const th2d_geom *geo, const th2d_hmc_params *hp,
                        double *A, const double *xi_host, const double *mom_host,
                        const double *st_host, double u_accept, unsigned long long seed,
                        unsigned long long counter, th2d_hmc_result *res);

/* nch independent chains, one trajectory each, in lockstep: every
 * inversion of all chains is one batched solve.  geos, hps (one per chain:
 * g and nf may differ, nsteps/traj_len/phys_norm/cg must agree), A, u_accept
 * and res are host arrays of length nch (A holds device pointers).  xi_h,
 * mom_h, st_h are NULL (device generator: chain k uses streams 4k+1..4k+3 of
 * `seed` at counters[k]) or host arrays of host noise pointers. */
int th2d_hmc_trajectory_batched(th2d_ctx *ctx, int nch, const th2d_geom *geos,
                                const th2d_hmc_params *hps, double **A,
                                const double *const *xi_h, const double *const *mom_h,
                                const double *const *st_h, const double *u_accept,
                                unsigned long long seed, const unsigned long long *counters,
                                th2d_hmc_result *res);

/* ---- dense complex matrix ---------------------------------------------- */
/* The V x V complex fermion matrix, column-major M[row + col*V] in natural
 * site order (fermion_matrix, hmc.c:269-310), from folded links K (device
 * output of 2 V^2 doubles). */
int th2d_fermion_matrix(th2d_ctx *ctx, const th2d_geom *g, const double *K, double *M);
/* det of a column-major complex n x n device matrix (destroyed) by
 * partial-pivot LU (determinant, hmc.c:314-336): log|det| and arg(det);
 * TH2D_ESINGULAR (log|det| = -inf) at an exactly zero pivot. */
int th2d_zdet(th2d_ctx *ctx, int n, double *M, double *logabs, double *phase);

/* The device generator's raw blocks: Philox4x32-10 (Salmon et al., SC'11)
 * of n (counter, key) pairs, ctr_key = device uint32[6n] (4 counter words,
 * then 2 key words, per item), out = device uint32[4n].  Exposed so the
 * generator behind every device noise field can be checked against the
 * published known-answer vectors; replaces no reference function (the
 * reference draws MT19937 on the host, mersenne_inline.c:59-114). */
int th2d_philox4x32_10(th2d_ctx *ctx, long n, const unsigned *ctr_key, unsigned *out);

/* Gaussian N(0,1) link angles from the device generator (configs: A~N(0,1)). */
int th2d_random_links(th2d_ctx *ctx, const th2d_geom *g, double *A, unsigned long long seed);

/* One pure-gauge Metropolis sweep with host-supplied uniform draws, two
 * per link in natural link order (update_puregauge_hb, hmc.c:82-93). */
int th2d_puregauge_sweep(th2d_ctx *ctx, const th2d_geom *g, double *A, const double *draws_dev,
                         double gcoup, int nf);
/* nsweeps of the same Metropolis update with draws from the device
 * generator (sweep k of link l: one Philox block at index l, counter
 * counter + k, stream 0x50475357; proposal first, accept draw second), all
 * sweeps of a link in one launch: the thermalisation of hmc.c:927-929
 * (100 sweeps from A = 0) without host draws.  Every link is updated
 * independently, as the action Nf/g sum(1 - cos A) is a sum over links. */
int th2d_puregauge_sweeps(th2d_ctx *ctx, const th2d_geom *g, double *A, double gcoup, int nf,
                          int nsweeps, unsigned long long seed, unsigned long long counter);

#ifdef __cplusplus
}
#endif
#endif
