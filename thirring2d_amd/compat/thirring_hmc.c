/*
 * thirring_hmc.c -- the reference's HMC functions (hmc.c:73-746) with their
 * own signatures over libth2d.  Every numerical step runs on the GPU; this
 * file only marshals the reference's jagged arrays to and from contiguous
 * natural-order buffers, draws the reference's random numbers (host MT
 * stream, reverse consumption) and calls the C ABI of include/th2d.h.
 */
#include "../../include/thirring_hmc.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/th2d.h"

/* ---- globals of hmc.c; the program's own definitions take precedence -- */
__attribute__((weak)) double m = 0.1;
__attribute__((weak)) double g = 0.5;
__attribute__((weak)) double mu = 0.0;
__attribute__((weak)) double ***A = NULL;

/* ---- MT19937 stream of mersenne.h (weak: mersenne_inline.c replaces it) -
 * Matsumoto-Nishimura MT19937 with the Knuth LCG (69069) seeding of the
 * 1998 sgenrand; even seeds are made odd; blocks of 624 doubles y / 2^32
 * are consumed from the top down (mersenne.h:11). */
#define MT_N 624
#define MT_M 397
__attribute__((weak)) int mersenne_i = -1;
__attribute__((weak)) double mersenne_array[MT_N];
static unsigned int th_mt[MT_N];

static void th_seed(long seed)
{
  if (seed % 2 == 0) seed++;
  th_mt[0] = (unsigned int)(seed & 0xffffffffL);
  for (int k = 1; k < MT_N; k++) th_mt[k] = 69069u * th_mt[k - 1];
  mersenne_i = 0;
}

static double th_generate(void)
{
  if (mersenne_i < 0) th_seed(4357);
  for (int k = 0; k < MT_N; k++) {
    unsigned int y = (th_mt[k] & 0x80000000u) | (th_mt[(k + 1) % MT_N] & 0x7fffffffu);
    unsigned int v = th_mt[(k + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    th_mt[k] = v;
  }
  for (int k = 0; k < MT_N; k++) {
    unsigned int y = th_mt[k];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    mersenne_array[k] = (double)y * 2.3283064365386963e-10;
  }
  mersenne_i = MT_N;
  return mersenne_array[--mersenne_i];
}

/* weak, so that a program linking mersenne_inline.c uses its own */
void seed_mersenne(long seed) __attribute__((weak, alias("th_seed")));
double mersenne_generate(void) __attribute__((weak, alias("th_generate")));

/* Whether the stream is this library's generator (its state words are
 * then th_mt) or the program's (mersenne_inline.c keeps its words in a
 * file-static array). */
static int own_generator(void) { return (void *)mersenne_generate == (void *)th_generate; }

/* The state word behind one output of the last block: the MT tempering is
 * a bijection of 32-bit words, so x = temper^-1(y). */
static unsigned int untemper(unsigned int y)
{
  y ^= y >> 18;
  y ^= (y << 15) & 0xefc60000u;
  unsigned int x = y;
  for (int k = 0; k < 5; k++) x = y ^ ((x << 7) & 0x9d2c5680u);
  y = x;
  for (int k = 0; k < 3; k++) x = y ^ (x >> 11);
  return x;
}

static inline double draw(void)
{
  return mersenne_i > 0 ? mersenne_array[--mersenne_i] : mersenne_generate();
}

/* Box-Muller pair from two draws, x1 then x2 (hmc.c:424-428). */
static void gauss_pair(double *c, double *s)
{
  double x1 = draw(), x2 = draw();
  double r = sqrt(-2.0 * log(x1));
  *c = r * cos(2.0 * M_PI * x2);
  *s = r * sin(2.0 * M_PI * x2);
}

/* ---- state -------------------------------------------------------------- */
static struct {
  th2d_ctx *ctx;
  th2d_geom geo;
  int nt, nx, V;
  double *dA, *dK, *dnat, *dv0, *dv1, *dv2, *dmom, *dM, *ddraw;
  float *dK32;
  double *h0, *h1, *h2, *h3;
  th2d_cg_params cg;
  int nsteps;
  double traj_len;
  int status, last_acc, last_its;
  double last_dH;
} S;

static int fail(int rc)
{
  if (rc != TH2D_OK && S.status == TH2D_OK) S.status = rc;
  return rc;
}

int thirring_hmc_init(int nt, int nx)
{
  thirring_hmc_finalize();
  if (nt < 2 || nx < 2 || (nt & 1) || (nx & 1)) return TH2D_EARG;
  int rc = th2d_ctx_create(&S.ctx, NULL);
  if (rc) return rc;
  S.nt = nt;
  S.nx = nx;
  S.V = nt * nx;
  S.geo = (th2d_geom){nt, nx, 1, 1, m, mu};
  size_t vb = sizeof(double) * 2 * (size_t)S.V;
  void **dev[] = {(void **)&S.dA, (void **)&S.dK, (void **)&S.dnat, (void **)&S.dv0,
                  (void **)&S.dv1, (void **)&S.dv2, (void **)&S.dmom};
  size_t sz[] = {vb, 2 * vb, vb, vb, vb, vb, vb};
  for (int k = 0; k < 7; k++)
    if ((rc = th2d_malloc(dev[k], sz[k]))) return rc;
  if ((rc = th2d_malloc((void **)&S.dK32, vb))) return rc;
  if ((rc = th2d_malloc((void **)&S.ddraw, 2 * vb))) return rc;
  S.h0 = malloc(2 * vb);
  S.h1 = malloc(vb);
  S.h2 = malloc(vb);
  S.h3 = malloc(vb);
  if (!S.h0 || !S.h1 || !S.h2 || !S.h3) return TH2D_ENOMEM;
  S.cg = (th2d_cg_params){0.0, 1e-30, 100000, TH2D_PREC_F64, 1e-4};
  S.nsteps = 10;
  S.traj_len = 1.0;
  S.status = TH2D_OK;
  return TH2D_OK;
}

void thirring_hmc_finalize(void)
{
  if (!S.ctx) return;
  void *dev[] = {S.dA, S.dK, S.dnat, S.dv0, S.dv1, S.dv2, S.dmom, S.dM, S.ddraw, S.dK32};
  for (int k = 0; k < 10; k++)
    if (dev[k]) th2d_free(dev[k]);
  free(S.h0);
  free(S.h1);
  free(S.h2);
  free(S.h3);
  th2d_ctx_destroy(S.ctx);
  memset(&S, 0, sizeof(S));
}

int thirring_hmc_set_cg(double tol_rel, double tol_abs_rr, int max_iter, int precision)
{
  if (max_iter < 1 || (precision != TH2D_PREC_F64 && precision != TH2D_PREC_MIXED) ||
      !(tol_rel > 0 || tol_abs_rr > 0))
    return TH2D_EARG;
  S.cg = (th2d_cg_params){tol_rel, tol_abs_rr, max_iter, precision, 1e-4};
  return TH2D_OK;
}

int thirring_hmc_set_md(int nsteps, double traj_len)
{
  if (nsteps < 1 || !(traj_len > 0)) return TH2D_EARG;
  S.nsteps = nsteps;
  S.traj_len = traj_len;
  return TH2D_OK;
}

int thirring_hmc_last_status(void) { return S.status; }

int thirring_hmc_last_trajectory(int *accepted, double *dH, int *cg_iterations)
{
  if (accepted) *accepted = S.last_acc;
  if (dH) *dH = S.last_dH;
  if (cg_iterations) *cg_iterations = S.last_its;
  return TH2D_OK;
}

/* ---- marshalling -------------------------------------------------------- */
static int ready(void)
{
  if (!S.ctx) return 0;
  S.geo.m = m;
  S.geo.mu = mu;
  S.status = TH2D_OK;
  return 1;
}

static void pack_links(double ***F, double *h)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) {
      h[2 * (t * S.nx + x)] = F[t][x][0];
      h[2 * (t * S.nx + x) + 1] = F[t][x][1];
    }
}

static void unpack_links(const double *h, double ***F)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) {
      F[t][x][0] = h[2 * (t * S.nx + x)];
      F[t][x][1] = h[2 * (t * S.nx + x) + 1];
    }
}

static void pack_vec(_Complex double **v, double *h)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) {
      h[2 * (t * S.nx + x)] = creal(v[t][x]);
      h[2 * (t * S.nx + x) + 1] = cimag(v[t][x]);
    }
}

static void unpack_vec(const double *h, _Complex double **v)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++)
      v[t][x] = h[2 * (t * S.nx + x)] + I * h[2 * (t * S.nx + x) + 1];
}

static size_t vbytes(void) { return sizeof(double) * 2 * (size_t)S.V; }

/* device field dA (eo) and folded links from a jagged field */
static int load_links(double ***F)
{
  pack_links(F, S.h1);
  int rc = th2d_h2d(S.ctx, S.dnat, S.h1, vbytes());
  if (!rc) rc = th2d_links_to_eo(S.ctx, &S.geo, S.dA, S.dnat);
  if (!rc) rc = th2d_fold_links(S.ctx, &S.geo, S.dK, S.dK32, S.dA);
  return fail(rc);
}

static int store_links(double ***F)
{
  int rc = th2d_links_from_eo(S.ctx, &S.geo, S.dnat, S.dA);
  if (!rc) rc = th2d_d2h(S.ctx, S.h1, S.dnat, vbytes());
  if (!rc) unpack_links(S.h1, F);
  return fail(rc);
}

static int load_mom(double ***mom)
{
  pack_links(mom, S.h1);
  int rc = th2d_h2d(S.ctx, S.dnat, S.h1, vbytes());
  if (!rc) rc = th2d_links_to_eo(S.ctx, &S.geo, S.dmom, S.dnat);
  return fail(rc);
}

static int store_mom(double ***mom)
{
  int rc = th2d_links_from_eo(S.ctx, &S.geo, S.dnat, S.dmom);
  if (!rc) rc = th2d_d2h(S.ctx, S.h1, S.dnat, vbytes());
  if (!rc) unpack_links(S.h1, mom);
  return fail(rc);
}

static int load_vec(_Complex double **v, double *d_eo)
{
  pack_vec(v, S.h1);
  int rc = th2d_h2d(S.ctx, S.dnat, S.h1, vbytes());
  if (!rc) rc = th2d_vec_to_eo(S.ctx, &S.geo, d_eo, S.dnat);
  return fail(rc);
}

static int store_vec(const double *d_eo, _Complex double **v)
{
  int rc = th2d_vec_from_eo(S.ctx, &S.geo, S.dnat, d_eo);
  if (!rc) rc = th2d_d2h(S.ctx, S.h1, S.dnat, vbytes());
  if (!rc) unpack_vec(S.h1, v);
  return fail(rc);
}

static double re_dot(const double *a, const double *b)
{
  double re = 0.0, im = 0.0;
  fail(th2d_dot(S.ctx, S.V, a, b, &re, &im));
  return re;
}

/* ---- hmc.c -------------------------------------------------------------- */
double calc_gauge_action(double ***F)
{
  if (!ready() || load_links(F)) return NAN;
  double s = NAN;
  fail(th2d_gauge_action(S.ctx, &S.geo, S.dA, g, 2, &s));
  return s;
}

double gauge_step(double ***F, double ***mom, double eps)
{
  if (!ready() || load_links(F) || load_mom(mom)) return 0.0;
  if (!fail(th2d_gauge_step(S.ctx, &S.geo, S.dA, S.dmom, eps, S.dK, S.dK32))) store_links(F);
  return 0.0;
}

static void apply(_Complex double **in, _Complex double **out, double ***F, int dagger)
{
  if (!ready() || load_links(F) || load_vec(in, S.dv0)) return;
  if (!fail(th2d_dslash(S.ctx, &S.geo, S.dv1, S.dv0, S.dK, dagger, TH2D_PREC_F64, 1)))
    store_vec(S.dv1, out);
}

void fm_mul(_Complex double **v_in, _Complex double **v_out, double ***F) { apply(v_in, v_out, F, 0); }

void fm_conjugate_mul(_Complex double **v_in, _Complex double **v_out, double ***F)
{
  apply(v_in, v_out, F, 1);
}

void fmdm_mul(_Complex double **v_in, _Complex double **v_out, double ***F)
{
  if (!ready() || load_links(F) || load_vec(v_in, S.dv0)) return;
  if (!fail(th2d_mdagm(S.ctx, &S.geo, S.dv1, S.dv0, S.dK, S.dv2))) store_vec(S.dv1, v_out);
}

static int ensure_dense(void)
{
  if (S.dM) return TH2D_OK;
  return fail(th2d_malloc((void **)&S.dM, sizeof(double) * 2 * (size_t)S.V * S.V));
}

void fermion_matrix(_Complex double *M)
{
  if (!ready() || !A || load_links(A) || ensure_dense()) return;
  if (!fail(th2d_fermion_matrix(S.ctx, &S.geo, S.dK, S.dM)))
    fail(th2d_d2h(S.ctx, (double *)M, S.dM, sizeof(double) * 2 * (size_t)S.V * S.V));
}

double complex determinant(void)
{
  if (!ready() || !A || load_links(A) || ensure_dense()) return NAN;
  if (fail(th2d_fermion_matrix(S.ctx, &S.geo, S.dK, S.dM))) return NAN;
  double la = 0.0, ph = 0.0;
  int rc = th2d_zdet(S.ctx, S.V, S.dM, &la, &ph);
  if (rc == TH2D_ESINGULAR) return 0.0;
  if (fail(rc)) return NAN;
  return exp(la) * cexp(I * ph);
}

static int solve(const double *b, double *x)
{
  th2d_cg_info info;
  int rc = th2d_cg_mdagm(S.ctx, &S.geo, x, b, S.dK, S.dK32, &S.cg, &info);
  return fail(rc);
}

void fmdm_invert_cg(_Complex double **v_in, _Complex double **v_out, double ***F)
{
  if (!ready() || load_links(F) || load_vec(v_in, S.dv0)) return;
  solve(S.dv0, S.dv1);
  store_vec(S.dv1, v_out);
}

void fm_invert_cg(_Complex double **v_in, _Complex double **v_out, double ***F)
{
  if (!ready() || load_links(F) || load_vec(v_in, S.dv0)) return;
  if (fail(th2d_dslash(S.ctx, &S.geo, S.dv2, S.dv0, S.dK, 1, TH2D_PREC_F64, 1))) return;
  solve(S.dv2, S.dv1);
  store_vec(S.dv1, v_out);
}

/* V complex Gaussian numbers from V Box-Muller pairs, natural order */
static double gauss_field(double *h)
{
  double n2 = 0.0;
  for (int s = 0; s < S.V; s++) {
    gauss_pair(&h[2 * s], &h[2 * s + 1]);
    n2 += h[2 * s] * h[2 * s] + h[2 * s + 1] * h[2 * s + 1];
  }
  return n2;
}

double random_pseudofermion(_Complex double **v, double ***F)
{
  if (!ready()) return NAN;
  double n2 = gauss_field(S.h2);
  if (load_links(F)) return n2;
  int rc = th2d_h2d(S.ctx, S.dnat, S.h2, vbytes());
  if (!rc) rc = th2d_vec_to_eo(S.ctx, &S.geo, S.dv0, S.dnat);
  if (!rc) rc = th2d_dslash(S.ctx, &S.geo, S.dv1, S.dv0, S.dK, 1, TH2D_PREC_F64, 1);
  if (!fail(rc)) store_vec(S.dv1, v);
  return n2;
}

void stochastic_vector(_Complex double **v)
{
  if (!ready()) return;
  gauss_field(S.h2);
  unpack_vec(S.h2, v);
}

double pseudofermion_action(_Complex double **v, double ***F)
{
  if (!ready() || load_links(F) || load_vec(v, S.dv0)) return NAN;
  if (solve(S.dv0, S.dv1)) return NAN;
  return re_dot(S.dv0, S.dv1);
}

double stochastic_md_action(_Complex double **v, double ***F)
{
  if (!ready() || load_links(F) || load_vec(v, S.dv0)) return NAN;
  if (fail(th2d_dslash(S.ctx, &S.geo, S.dv1, S.dv0, S.dK, 1, TH2D_PREC_F64, 1))) return NAN;
  return re_dot(S.dv0, S.dv1);
}

double random_momentum(double ***mom)
{
  if (!ready()) return NAN;
  double n2 = gauss_field(S.h2); /* cos -> dir 0, sin -> dir 1 (hmc.c:490-491) */
  unpack_links(S.h2, mom);
  return n2;
}

double momentum_step(double ***mom, double ***F, _Complex double **psi, _Complex double **st,
                     double eps)
{
  if (!ready() || load_links(F) || load_vec(psi, S.dv0) || load_mom(mom)) return 0.0;
  if (solve(S.dv0, S.dv1)) return 0.0;
  if (st && load_vec(st, S.dv2)) return 0.0;
  if (!fail(th2d_force(S.ctx, &S.geo, S.dmom, S.dA, S.dK, S.dv1, st ? S.dv2 : NULL, eps, g, 2,
                       NULL)))
    store_mom(mom);
  return 0.0;
}

void update_gauge(double ***F)
{
  if (!ready()) return;
  /* draws in the reference's order: pseudofermion, momentum, stochastic
   * vector, accept (hmc.c:695-699, 738) */
  double *xi = S.h0, *pm = S.h0 + 2 * S.V, *stv = S.h2;
  gauss_field(xi);
  gauss_field(pm);
  gauss_field(stv);
  double u = draw();
  if (load_links(F)) return;
  th2d_hmc_params hp = {g, 2, S.nsteps, S.traj_len, 0, S.cg};
  th2d_hmc_result res;
  memset(&res, 0, sizeof(res));
  int rc = th2d_hmc_trajectory(S.ctx, &S.geo, &hp, S.dA, xi, pm, stv, u, 0, 0, &res);
  fail(rc);
  S.last_acc = res.accepted;
  S.last_dH = res.dS;
  S.last_its = res.cg_iterations;
  if (res.accepted) store_links(F);
}

void update_puregauge_hb(double ***F)
{
  if (!ready() || load_links(F)) return;
  for (int k = 0; k < 4 * S.V; k++) S.h0[k] = draw(); /* two per link, natural order */
  int rc = th2d_h2d(S.ctx, S.ddraw, S.h0, 2 * vbytes());
  if (!rc) rc = th2d_puregauge_sweep(S.ctx, &S.geo, S.dA, S.ddraw, g, 2);
  if (!fail(rc)) store_links(F);
}

/* ---- checkpoint of an HMC run --------------------------------------------
 * The reference cannot resume an HMC run (hmc.c:932-939 loops from a fresh
 * field).  A checkpoint holds what a resumed run needs to continue the same
 * Markov chain draw for draw: the field A in FP64, the trajectory counter
 * and the whole MT19937 state.  Layout (host byte order):
 *   char magic[8] "TH2DHMC1"; int32 nt, nx, has_words, reserved;
 *   double m, g, mu; int64 trajectory; double A[nt][nx][2];
 *   int32 mersenne_i, reserved; double mersenne_array[624];
 *   uint32 words[624]
 * The state words are this library's generator state when it supplies the
 * stream; when the program links its own mersenne_inline.c (whose words
 * are file-static), they are recovered from the last block by inverting
 * the tempering (has_words = 2).  Reading restores A, the counter and the
 * exported block, and the words into this library's generator; with the
 * program's generator the words cannot be installed, and the read returns
 * THIRRING_HMC_RNG_PARTIAL. */
static const char ckpt_magic[8] = {'T', 'H', '2', 'D', 'H', 'M', 'C', '1'};

int thirring_hmc_write_checkpoint(const char *path, double ***F, long trajectory)
{
  if (!S.ctx || !path || !F) return TH2D_EARG;
  FILE *f = fopen(path, "wb");
  if (!f) return TH2D_EARG;
  int own = own_generator();
  int hdr[4] = {S.nt, S.nx, own ? 1 : 2, 0};
  double par[3] = {m, g, mu};
  long long traj = trajectory;
  int st[2] = {mersenne_i, 0};
  unsigned int words[MT_N];
  for (int k = 0; k < MT_N; k++)
    words[k] = own ? th_mt[k] : untemper((unsigned int)(mersenne_array[k] * 4294967296.0));
  pack_links(F, S.h1);
  int ok = fwrite(ckpt_magic, 1, 8, f) == 8 && fwrite(hdr, sizeof(int), 4, f) == 4 &&
           fwrite(par, sizeof(double), 3, f) == 3 && fwrite(&traj, sizeof(traj), 1, f) == 1 &&
           fwrite(S.h1, sizeof(double), 2 * (size_t)S.V, f) == 2 * (size_t)S.V &&
           fwrite(st, sizeof(int), 2, f) == 2 &&
           fwrite(mersenne_array, sizeof(double), MT_N, f) == MT_N &&
           fwrite(words, sizeof(unsigned int), MT_N, f) == MT_N;
  ok = (fclose(f) == 0) && ok;
  return ok ? TH2D_OK : TH2D_EARG;
}

int thirring_hmc_read_checkpoint(const char *path, double ***F, long *trajectory)
{
  if (!S.ctx || !path || !F) return TH2D_EARG;
  FILE *f = fopen(path, "rb");
  if (!f) return TH2D_EARG;
  char magic[8];
  int hdr[4], st[2];
  double par[3];
  long long traj = 0;
  unsigned int words[MT_N];
  double arr[MT_N];
  int ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, ckpt_magic, 8) == 0 &&
           fread(hdr, sizeof(int), 4, f) == 4 && hdr[0] == S.nt && hdr[1] == S.nx &&
           fread(par, sizeof(double), 3, f) == 3 && fread(&traj, sizeof(traj), 1, f) == 1 &&
           fread(S.h1, sizeof(double), 2 * (size_t)S.V, f) == 2 * (size_t)S.V &&
           fread(st, sizeof(int), 2, f) == 2 && fread(arr, sizeof(double), MT_N, f) == MT_N &&
           fread(words, sizeof(unsigned int), MT_N, f) == MT_N;
  fclose(f);
  if (!ok || st[0] > MT_N) return TH2D_EARG;
  unpack_links(S.h1, F);
  m = par[0];
  g = par[1];
  mu = par[2];
  if (trajectory) *trajectory = (long)traj;
  memcpy(mersenne_array, arr, sizeof(arr));
  mersenne_i = st[0];
  if (!own_generator()) return THIRRING_HMC_RNG_PARTIAL;
  memcpy(th_mt, words, sizeof(words));
  return TH2D_OK;
}
