/*
 * thirring_bag.c -- the reference's fermion-bag functions with their own
 * signatures over libth2d.  Every matrix operation, solve, LU and rank-k
 * update runs on the GPU; this file marshals the jagged field and vectors
 * and keeps the reference's bookkeeping (field values, counters, sign).
 */
#include "../../include/thirring_bag.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/th2d.h"
#include "../../include/th2d_bag.h"

#define SOURCE_MONOMER 6 /* 2 + NDIRS, Thirring.h:39 */
#define CG_ACCURACY 1e-30 /* Thirring.h:42 */
#define CG_MAX_ITER 10000 /* Thirring.h:43 */

/* ---- globals of fermionbag.c; the program's own definitions win ------- */
__attribute__((weak)) int **field = NULL;
__attribute__((weak)) double m = 0.0;
__attribute__((weak)) double mu = 0.0;
__attribute__((weak)) double U = 0.0;
__attribute__((weak)) double *Dinv = NULL;
__attribute__((weak)) int n_monomers = 0;
__attribute__((weak)) int n_links = 0;
__attribute__((weak)) int current_sign = 1;
__attribute__((weak)) int max_changes = 0;
__attribute__((weak)) void update_linklists(void) {}

/* ---- MT19937 stream of mersenne.h (weak: a program's mersenne_inline.c,
 * or libthirring_hmc's copy when both libraries are loaded, takes over) --
 * Matsumoto-Nishimura MT19937 with the Knuth LCG (69069) seeding of the
 * 1998 sgenrand; blocks of 624 doubles y / 2^32 consumed from the top. */
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

void seed_mersenne(long seed) __attribute__((weak, alias("th_seed")));
double mersenne_generate(void) __attribute__((weak, alias("th_generate")));

static inline double draw(void)
{
  return mersenne_i > 0 ? mersenne_array[--mersenne_i] : mersenne_generate();
}

static struct {
  th2d_ctx *ctx;
  th2d_bag_geom geo;
  int nt, nx, V;
  th2d_bag *bag;
  double bag_m, bag_mu;
  int *dfield;
  double *dv0, *dv1, *dD;
  int *hf;
  double *hv;
  int changes;
  int status;
} S;

static int fail(int rc)
{
  if (rc != TH2D_OK && S.status == TH2D_OK) S.status = rc;
  return rc;
}

int thirring_bag_init(int nt, int nx, int bc, int with_mass_monomers)
{
  thirring_bag_finalize();
  if (nt < 4 || nx < 4 || (nt & 1) || (nx & 1) || (bc != 0 && bc != 1)) return TH2D_EARG;
  int rc = th2d_ctx_create(&S.ctx, NULL);
  if (rc) return rc;
  S.nt = nt;
  S.nx = nx;
  S.V = nt * nx;
  S.geo = (th2d_bag_geom){nt, nx, bc, with_mass_monomers, m, mu};
  size_t vb = sizeof(double) * (size_t)S.V;
  if ((rc = th2d_malloc((void **)&S.dfield, sizeof(int) * S.V))) return rc;
  if ((rc = th2d_malloc((void **)&S.dv0, vb))) return rc;
  if ((rc = th2d_malloc((void **)&S.dv1, vb))) return rc;
  if ((rc = th2d_malloc((void **)&S.dD, sizeof(double) * (size_t)S.V * S.V / 2))) return rc;
  S.hf = malloc(sizeof(int) * S.V);
  S.hv = malloc(vb);
  if (!S.hf || !S.hv) return TH2D_ENOMEM;
  S.status = TH2D_OK;
  return TH2D_OK;
}

void thirring_bag_finalize(void)
{
  if (!S.ctx) return;
  if (S.bag) th2d_bag_destroy(S.bag);
  void *dev[] = {S.dfield, S.dv0, S.dv1, S.dD};
  for (int k = 0; k < 4; k++)
    if (dev[k]) th2d_free(dev[k]);
  free(S.hf);
  free(S.hv);
  th2d_ctx_destroy(S.ctx);
  memset(&S, 0, sizeof(S));
}

int thirring_bag_last_status(void) { return S.status; }

static int ready(void)
{
  if (!S.ctx) return 0;
  S.geo.m = m;
  S.geo.mu = mu;
  S.status = TH2D_OK;
  return 1;
}

static void pack_field(void)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) S.hf[t * S.nx + x] = field ? field[t][x] : 0;
}

static int upload_field(void)
{
  pack_field();
  return fail(th2d_h2d(S.ctx, S.dfield, S.hf, sizeof(int) * S.V));
}

static int load_vec(double **v, double *d)
{
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) S.hv[t * S.nx + x] = v[t][x];
  return fail(th2d_h2d(S.ctx, d, S.hv, sizeof(double) * S.V));
}

static int store_vec(const double *d, double **v)
{
  int rc = fail(th2d_d2h(S.ctx, S.hv, d, sizeof(double) * S.V));
  if (!rc)
    for (int t = 0; t < S.nt; t++)
      for (int x = 0; x < S.nx; x++) v[t][x] = S.hv[t * S.nx + x];
  return rc;
}

/* neighbour in direction dir (0 t up, 1 x up, 2 t down, 3 x down) */
static void nb(int t, int x, int dir, int *t2, int *x2)
{
  *t2 = dir == 0 ? (t + 1) % S.nt : dir == 2 ? (t - 1 + S.nt) % S.nt : t;
  *x2 = dir == 1 ? (x + 1) % S.nx : dir == 3 ? (x - 1 + S.nx) % S.nx : x;
}

/* ---- fermion_matrix.c -------------------------------------------------- */
double fM_index(int t1, int x1, int t2, int x2, double mu_)
{
  if (!ready()) return NAN;
  double v = NAN;
  fail(th2d_bag_matrix_element(&S.geo, t1, x1, t2, x2, mu_, &v));
  return v;
}

/* The free propagator is cached on disk as fermion_matrix.c:206-220,
 * 342-346 and 376-384, 418-422 do, in the working directory, host byte
 * order.  A file that exists with the expected length is read instead of
 * computing; a short file is ignored here (the reference reads it
 * unchecked).  Otherwise Dinv is computed on the GPU and the file written.
 *
 * With mass monomers (fermion_matrix.c:191-351) Dinv holds the eo blocks,
 * V*V/2 doubles, in a file named for the lattice, mu and the boundary.
 * Without them (:355-425) Dinv is the full V x V inverse in natural site
 * order, Dinv[(NX*t1 + x1)*V + NX*t2 + x2] = (D^-1)_{1,2}, and the file
 * name also holds m.  The reference writes only the first V*V/2 doubles of
 * that inverse (and reads V*V back); this library writes all V*V, so its
 * files read back whole, and it ignores the reference's half-length ones. */
void calc_Dinv(void)
{
  if (!ready()) return;
  const int eo = S.geo.with_mass_monomers != 0;
  const size_t count = eo ? (size_t)S.V * S.V / 2 : (size_t)S.V * S.V;
  if (!Dinv) Dinv = malloc(sizeof(double) * (size_t)S.V * S.V);
  if (!Dinv) {
    S.status = TH2D_ENOMEM;
    return;
  }
  char filename[128];
  const char *open = S.geo.bc == 1 ? "_open" : "";
  if (eo)
    snprintf(filename, sizeof(filename), "free_propagator_eo_T%dX%d_mu%0.6f%s", S.nt, S.nx, mu,
             open);
  else
    snprintf(filename, sizeof(filename), "free_propagator_T%dX%d_m%.6f_mu%0.6f%s", S.nt, S.nx, m,
             mu, open);
  FILE *f = fopen(filename, "rb");
  if (f) {
    const size_t got = fread(Dinv, sizeof(double), count, f);
    fclose(f);
    if (got == count) return;
  }
  struct timeval start, end;
  gettimeofday(&start, NULL);
  if (eo) {
    if (fail(th2d_bag_calc_dinv(S.ctx, &S.geo, S.dD))) return;
    if (fail(th2d_d2h(S.ctx, Dinv, S.dD, sizeof(double) * count))) return;
  } else {
    /* B = D^-1 of the empty configuration, V x V row-major, natural order */
    th2d_bag *b = NULL;
    double *B = NULL;
    int *fd = NULL, V = 0;
    if (fail(th2d_bag_create(S.ctx, &S.geo, &b))) return;
    int rc = th2d_bag_views(b, &B, &fd, &V);
    if (!rc) rc = th2d_d2h(S.ctx, Dinv, B, sizeof(double) * count);
    th2d_bag_destroy(b);
    if (fail(rc)) return;
  }
  gettimeofday(&end, NULL);
  const long diff = 1000000L * (end.tv_sec - start.tv_sec) + (end.tv_usec - start.tv_usec);
  printf("Inverted fermion matrix in %.3g seconds\n", 1e-6 * diff);
  f = fopen(filename, "wb");
  if (f) {
    fwrite(Dinv, sizeof(double), count, f);
    fclose(f);
  }
}

/* ---- fermionbag.c: configuration checkpoint (:120-161) -------------------
 * NT*NX int32 values field[t][x] at t*NX + x, host byte order; a file that
 * cannot be opened ends the program with the reference's message. */
void write_config(void)
{
  if (!S.nt || !field) {
    S.status = TH2D_EARG;
    return;
  }
  char filename[100];
  snprintf(filename, sizeof(filename), "config_checkpoint_T%dX%d_U%.6gm%.6gmu%.6g", S.nt, S.nx, U,
           m, mu);
  int *buffer = malloc(sizeof(int) * (size_t)S.V);
  if (!buffer) {
    S.status = TH2D_ENOMEM;
    return;
  }
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) buffer[t * S.nx + x] = field[t][x];
  FILE *f = fopen(filename, "wb");
  if (f) {
    fwrite(buffer, sizeof(int), (size_t)S.V, f);
    fclose(f);
  } else {
    printf("Could not write configuration\n");
    exit(1);
  }
  free(buffer);
}

void read_config(char *filename)
{
  if (!S.nt || !field || !filename) {
    S.status = TH2D_EARG;
    return;
  }
  int *buffer = calloc((size_t)S.V, sizeof(int));
  if (!buffer) {
    S.status = TH2D_ENOMEM;
    return;
  }
  FILE *f = fopen(filename, "rb");
  if (f) {
    if (fread(buffer, sizeof(int), (size_t)S.V, f) != (size_t)S.V) S.status = TH2D_EARG;
    fclose(f);
  } else {
    printf("Could not read configuration\n");
    exit(1);
  }
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++) field[t][x] = buffer[t * S.nx + x];
  free(buffer);
}

/* ---- vec_ops.c ---------------------------------------------------------- */
static void apply(double **chi, double **psi, int transpose)
{
  if (!ready() || upload_field() || load_vec(psi, S.dv0)) return;
  if (!fail(th2d_bag_apply(S.ctx, &S.geo, S.dfield, S.dv1, S.dv0, transpose, 1))) store_vec(S.dv1, chi);
}

void fM(double **chi, double **psi) { apply(chi, psi, 0); }

void fM_transpose(double **chi, double **psi) { apply(chi, psi, 1); }

void cg_propagator(double **propagator, double **source)
{
  if (!ready() || upload_field() || load_vec(source, S.dv0)) return;
  int it = 0, st = 0;
  int rc = th2d_bag_cg(S.ctx, &S.geo, S.dfield, S.dv1, S.dv0, 1, CG_ACCURACY, CG_MAX_ITER, &it, &st);
  /* divergence fills the result with 1e50 as vec_ops.c:292-296 does */
  if (rc != TH2D_OK && rc != TH2D_EBREAKDOWN && rc != TH2D_ENOCONV) fail(rc);
  else if (rc != TH2D_OK) S.status = rc;
  store_vec(S.dv1, propagator);
}

/* ---- determinants --------------------------------------------------------- */
double determinant_mu(double mu_)
{
  if (!ready()) return NAN;
  pack_field();
  th2d_bag_geom g = S.geo;
  g.mu = mu_;
  double la = 0.0;
  int sg = 0;
  if (fail(th2d_bag_determinant(S.ctx, &g, S.hf, &la, &sg))) return NAN;
  return sg == 0 ? 0.0 : sg * exp(la);
}

double determinant(void) { return determinant_mu(mu); }

static int set_sign(void)
{
  double la, lad;
  int sg, sgd;
  int rc = th2d_bag_logdet(S.bag, &la, &sg, &lad, &sgd);
  if (!rc) current_sign = sg * sgd; /* sign of det_e det_o = det(D_FF) / det(D) */
  return fail(rc);
}

void update_background(void)
{
  if (!ready()) return;
  if (S.bag && (S.bag_m != m || S.bag_mu != mu)) {
    th2d_bag_destroy(S.bag);
    S.bag = NULL;
  }
  if (!S.bag) {
    if (fail(th2d_bag_create(S.ctx, &S.geo, &S.bag))) return;
    S.bag_m = m;
    S.bag_mu = mu;
  }
  pack_field();
  if (fail(th2d_bag_set_field(S.bag, S.hf))) return;
  S.changes = 0;
  set_sign();
}

static int state(void)
{
  if (!ready()) return 0;
  if (!S.bag) update_background();
  return S.bag != NULL;
}

static double ratio(const int *occ, int no, const int *fr, int nf)
{
  double r = 0.0;
  int rc = th2d_bag_ratio(S.bag, occ, no, fr, nf, &r);
  if (rc == TH2D_ESINGULAR) return 0.0;
  if (fail(rc)) return NAN;
  return fabs(r);
}

double det_added_link(int t, int x, int t2, int x2)
{
  if (!state()) return NAN;
  int occ[2] = {t * S.nx + x, t2 * S.nx + x2};
  return ratio(occ, 2, NULL, 0);
}

double det_removed_link(int t, int x, int t2, int x2)
{
  if (!state()) return NAN;
  int fr[2] = {t * S.nx + x, t2 * S.nx + x2};
  return ratio(NULL, 0, fr, 2);
}

double det_moved_monomer(int t, int x, int t2, int x2)
{
  if (!state()) return NAN;
  if (((t + x) & 1) != ((t2 + x2) & 1)) return 0.0;
  int fr[1] = {t * S.nx + x}, occ[1] = {t2 * S.nx + x2};
  return ratio(occ, 1, fr, 1);
}

/* commit a change to the device state, then the reference's bookkeeping */
static void commit(const int *occ, const int *vals, int no, const int *fr, int nf)
{
  if (fail(th2d_bag_accept(S.bag, occ, no, vals, fr, nf))) return;
  set_sign();
  update_linklists();
  if (max_changes > 0 && ++S.changes >= max_changes) update_background();
}

void new_link(int t, int x, int nu)
{
  if (!state()) return;
  int t2, x2;
  nb(t, x, nu, &t2, &x2);
  field[t][x] = nu + 2;
  field[t2][x2] = (nu + 2) % 4 + 2; /* opp_dir(nu) + 2 */
  int occ[2] = {t * S.nx + x, t2 * S.nx + x2}, vals[2] = {field[t][x], field[t2][x2]};
  commit(occ, vals, 2, NULL, 0);
}

void new_monomer(int t1, int x1, int t2, int x2)
{
  if (!state()) return;
  n_monomers += 2;
  field[t1][x1] = 1;
  field[t2][x2] = 1;
  int occ[2] = {t1 * S.nx + x1, t2 * S.nx + x2}, vals[2] = {1, 1};
  commit(occ, vals, 2, NULL, 0);
}

void removed_link(int t, int x, int nu)
{
  if (!state()) return;
  n_links -= 1;
  int t2, x2;
  nb(t, x, nu, &t2, &x2);
  field[t][x] = 0;
  field[t2][x2] = 0;
  int fr[2] = {t * S.nx + x, t2 * S.nx + x2};
  commit(NULL, NULL, 0, fr, 2);
}

void removed_monomer(int t1, int x1, int nu)
{
  if (!state()) return;
  n_monomers -= 2;
  int t2, x2;
  nb(t1, x1, nu, &t2, &x2);
  field[t1][x1] = 0;
  field[t2][x2] = 0;
  int fr[2] = {t1 * S.nx + x1, t2 * S.nx + x2};
  commit(NULL, NULL, 0, fr, 2);
}

void moved_source_monomer(int t2, int x2, int t4, int x4)
{
  if (!state()) return;
  field[t2][x2] = 0;
  field[t4][x4] = SOURCE_MONOMER;
  int fr[1] = {t2 * S.nx + x2}, occ[1] = {t4 * S.nx + x4}, vals[1] = {SOURCE_MONOMER};
  commit(occ, vals, 1, fr, 1);
}

void not_moved_source_monomer(int t2, int x2, int t4, int x4)
{
  (void)t2;
  (void)x2;
  (void)t4;
  (void)x4;
}

/* ---- fermionbag.c: the worm estimate of the susceptibility (:450-590) ----
 * n_attempts worms on the program's field: a random link (in the order
 * update_linklists lists them) becomes two source monomers, one of which
 * moves until it meets the other and the link closes again.  The moves,
 * the draws and their order are the reference's; the determinant ratios
 * (det_added_link, det_removed_link, det_moved_monomer) and the accept
 * hooks (new_link, removed_link) are this library's.  Antisymmetric
 * lattices only (the open-x neighbour arrays point at a column that does
 * not exist). */
#define LINK_TUP 2

static int worm_links(int *links)
{
  int n = 0;
  for (int t = 0; t < S.nt; t++)
    for (int x = 0; x < S.nx; x++)
      if (field[t][x] >= LINK_TUP && field[t][x] < LINK_TUP + 2)
        links[n++] = (field[t][x] - LINK_TUP) * S.V + t * S.nx + x;
  return n;
}

static void worm_link_on(int t, int x, int dir)
{
  int t2, x2;
  nb(t, x, dir, &t2, &x2);
  field[t][x] = dir + 2;
  field[t2][x2] = (dir + 2) % 4 + 2;
}

static void worm_link_off(int t, int x, int dir)
{
  int t2, x2;
  nb(t, x, dir, &t2, &x2);
  field[t][x] = 0;
  field[t2][x2] = 0;
}

long thirring_bag_worm(int n_attempts, double *value)
{
  if (value) *value = 0.0;
  if (n_attempts < 1 || !field || !state() || S.geo.bc != TH2D_BC_ANTISYMMETRIC) {
    S.status = TH2D_EARG;
    return 0;
  }
  int *links = malloc(sizeof(int) * 2 * (size_t)S.V);
  if (!links) {
    S.status = TH2D_ENOMEM;
    return 0;
  }
  long steps = 0;
  const double scale = (double)n_links / (4.0 * n_attempts * S.V);
  update_linklists();
  n_links = worm_links(links);
  if (n_links > 0)
    for (int attempt = 0; attempt < n_attempts && S.status == TH2D_OK; attempt++) {
      const int l = links[(int)(draw() * n_links)];
      int x1 = l % S.nx, t1 = (l / S.nx) % S.nt, nu = l / S.V, t2, x2;
      if (draw() > 0.5) {
        int ta, xa;
        nb(t1, x1, nu, &ta, &xa);
        t1 = ta;
        x1 = xa;
      }
      const int dir0 = field[t1][x1] - LINK_TUP;
      worm_link_off(t1, x1, dir0);
      nb(t1, x1, dir0, &t2, &x2);
      field[t1][x1] = SOURCE_MONOMER;
      field[t2][x2] = SOURCE_MONOMER;
      n_monomers += 2;
      n_links -= 1;
      steps += current_sign;
      for (;;) {
        const int dir = (int)(4 * draw());
        int t3, x3;
        nb(t2, x2, dir, &t3, &x3);
        if (draw() > 0.5) {
          if (t3 == t1 && x3 == x1) {  /* back at the start: the link again */
            field[t1][x1] = 0;
            field[t2][x2] = 0;
            worm_link_on(t2, x2, dir);
            n_monomers -= 2;
            n_links += 1;
            break;
          } else if (field[t3][x3] >= LINK_TUP) {  /* flip the link */
            const int ld = field[t3][x3] - LINK_TUP;
            int t4, x4;
            nb(t3, x3, ld, &t4, &x4);
            field[t2][x2] = 0;
            worm_link_off(t3, x3, ld);
            worm_link_on(t2, x2, dir);
            field[t4][x4] = SOURCE_MONOMER;
            t2 = t4;
            x2 = x4;
          } else if (field[t3][x3] == 1) {  /* swap with the monomer */
            field[t2][x2] = 1;
            field[t3][x3] = SOURCE_MONOMER;
            t2 = t3;
            x2 = x3;
          }
        } else if (S.geo.with_mass_monomers) {
          int dir2;
          do dir2 = (int)(4 * draw());
          while (dir2 == (dir + 2) % 4);
          int t4, x4;
          nb(t3, x3, dir2, &t4, &x4);
          (void)draw(); /* the hop of :532-544 is never taken (draw < 0) */
          if (field[t3][x3] == 0 && field[t4][x4] == 0) {
            const double p = U * det_added_link(t3, x3, t4, x4);
            if (draw() < p) {
              new_link(t3, x3, dir2);
              field[t2][x2] = 0;
              worm_link_off(t3, x3, dir2);
              worm_link_on(t2, x2, dir);
              field[t4][x4] = SOURCE_MONOMER;
              t2 = t4;
              x2 = x4;
            }
          } else if (field[t3][x3] == LINK_TUP + dir2) {
            const double p = det_removed_link(t2, x2, t3, x3) / U;
            if (draw() < p) {
              field[t2][x2] = 0;
              worm_link_off(t3, x3, dir2);
              worm_link_on(t2, x2, dir);
              field[t4][x4] = SOURCE_MONOMER;
              removed_link(t2, x2, dir);
              t2 = t4;
              x2 = x4;
            }
          }
        } else if (field[t3][x3] == 0) {
          /* the hop to an empty neighbour (:565-577) */
          const double det = fabs(det_moved_monomer(t2, x2, t3, x3));
          if (draw() < det) {
            moved_source_monomer(t2, x2, t3, x3);
            t2 = t3;
            x2 = x3;
          } else {
            not_moved_source_monomer(t2, x2, t3, x3);
          }
        }
        steps += current_sign;
      }
      update_linklists();
      n_links = worm_links(links);
    }
  free(links);
  if (value) *value = (double)steps * scale;
  return steps;
}

__attribute__((weak)) void measure_susceptibility(void)
{
  double v = 0.0;
  thirring_bag_worm(20, &v);
  printf("Susceptibility %g \n", v);
}
