/*
 * th2d_oracle.c -- CPU restatement of the Thirring2D fermion hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library in
 * thirring2d_amd/.  Nothing in the product links or loads this file.
 *
 * Written from the mathematics of the reference's algorithms; each
 * function cites the reference file:line whose behaviour it restates.
 * Plain C99, single thread, FP64 throughout.
 */
#include "th2d_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------ */
/* MT19937 (mersenne_inline.c:59-114, mersenne.h:11)                   */
/* ------------------------------------------------------------------ */
enum { MT_N = 624, MT_M = 397 };

void orc_mt_seed(orc_mt *st, long seed)
{
  /* sgenrand: even seeds are made odd, then Knuth's LCG 69069 fills the
   * state (mersenne_inline.c:67-70). */
  if (seed % 2 == 0) seed++;
  st->mt[0] = (uint32_t)(seed & 0xffffffffL);
  for (int k = 1; k < MT_N; k++) st->mt[k] = (uint32_t)(69069u * st->mt[k - 1]);
  st->i = 0;
}

static void mt_refill(orc_mt *st)
{
  uint32_t *s = st->mt;
  for (int k = 0; k < MT_N; k++) {
    uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % MT_N] & 0x7fffffffu);
    uint32_t v = s[(k + MT_M) % MT_N] ^ (y >> 1);
    if (y & 1u) v ^= 0x9908b0dfu;
    s[k] = v;
  }
  for (int k = 0; k < MT_N; k++) {
    uint32_t y = s[k];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    st->out[k] = (double)y * (1.0 / 4294967296.0);
  }
  st->i = MT_N;
}

double orc_mt_next(orc_mt *st)
{
  /* blocks are consumed from the top down (mersenne.h:11) */
  if (st->i <= 0) mt_refill(st);
  return st->out[--st->i];
}

void orc_mt_skip(orc_mt *st, long n)
{
  for (long k = 0; k < n; k++) (void)orc_mt_next(st);
}

void orc_gauss_pair(orc_mt *st, double *c, double *s)
{
  double x1 = orc_mt_next(st);
  double x2 = orc_mt_next(st);
  double r = sqrt(-2.0 * log(x1));
  *c = r * cos(2 * M_PI * x2);
  *s = r * sin(2 * M_PI * x2);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 and the device noise mapping                          */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;  /* Weyl key schedule */
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static double u53(uint32_t a, uint32_t b)
{
  uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)v + 0.5) / 9007199254740992.0;
}

void orc_philox_uniform2(uint64_t seed, uint32_t stream, uint32_t idx, uint32_t ctr, double *u0,
                         double *u1)
{
  const uint32_t c[4] = {idx, ctr, stream, 0x7468u};
  const uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  orc_philox4x32_10(c, k, r);
  *u0 = u53(r[0], r[1]);
  *u1 = u53(r[2], r[3]);
}

void orc_philox_gauss2(uint64_t seed, uint32_t stream, uint32_t idx, uint32_t ctr, double *re,
                       double *im)
{
  const uint32_t c[4] = {idx, ctr, stream, 0x7468u};
  const uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  orc_philox4x32_10(c, k, r);
  double u1 = u53(r[0], r[1]), u2 = u53(r[2], r[3]);
  double rad = sqrt(-2.0 * log(u1));
  *re = rad * cos(2 * M_PI * u2);
  *im = rad * sin(2 * M_PI * u2);
}

/* ------------------------------------------------------------------ */
/* U(1) staggered operator (hmc.c)                                     */
/* ------------------------------------------------------------------ */
int orc_eta(int t, int x, int dir)
{
  (void)t;
  if (dir == 1) return 1;
  return (x % 2 == 0) ? 1 : -1;
}

static inline cplx ld(const double *v, int s) { return v[2 * s] + I * v[2 * s + 1]; }
static inline void st_(double *v, int s, cplx z) { v[2 * s] = creal(z); v[2 * s + 1] = cimag(z); }

/* Sign of a hop that crosses the boundary in direction dir. */
static inline double wrap_sign(const orc_lattice *L, int dir)
{
  return (dir == 0 ? L->anti_t : L->anti_x) ? -1.0 : 1.0;
}

/* Matrix elements of M (hmc.c:139-180):
 *   M[x, x+mu] = +1/2 eta_mu(x) e^{+mu_mu} s U_mu(x)
 *   M[x, x-mu] = -1/2 eta_mu(x-mu) e^{-mu_mu} s conj U_mu(x-mu)
 * s = -1 for a hop through an antiperiodic boundary; mu_t = mu, mu_x = 0. */
static inline cplx hop_fwd(const orc_lattice *L, const double *A, int t, int x, int dir)
{
  int V_nt = L->nt, V_nx = L->nx;
  int s = t * V_nx + x;
  double a = A[2 * s + dir];
  int wraps = (dir == 0) ? (t == V_nt - 1) : (x == V_nx - 1);
  double sg = wraps ? wrap_sign(L, dir) : 1.0;
  double em = (dir == 0) ? exp(L->mu) : 1.0;
  return 0.5 * sg * orc_eta(t, x, dir) * em * (cos(a) + I * sin(a));
}

/* Coefficient of psi(x-mu) in (M psi)(x); the link lives at x-mu. */
static inline cplx hop_bwd(const orc_lattice *L, const double *A, int t, int x, int dir)
{
  int tb = (dir == 0) ? (t - 1 + L->nt) % L->nt : t;
  int xb = (dir == 1) ? (x - 1 + L->nx) % L->nx : x;
  int s = tb * L->nx + xb;
  double a = A[2 * s + dir];
  int wraps = (dir == 0) ? (t == 0) : (x == 0);
  double sg = wraps ? wrap_sign(L, dir) : 1.0;
  double em = (dir == 0) ? exp(-L->mu) : 1.0;
  return -0.5 * sg * orc_eta(tb, xb, dir) * em * (cos(a) - I * sin(a));
}

void orc_fm_mul(const orc_lattice *L, const double *A, const double *in, double *out)
{
  int nt = L->nt, nx = L->nx;
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      int tu = (t + 1) % nt, td = (t - 1 + nt) % nt;
      int xu = (x + 1) % nx, xd = (x - 1 + nx) % nx;
      cplx v = L->m * ld(in, s);
      v += hop_fwd(L, A, t, x, 0) * ld(in, tu * nx + x);
      v += hop_bwd(L, A, t, x, 0) * ld(in, td * nx + x);
      v += hop_fwd(L, A, t, x, 1) * ld(in, t * nx + xu);
      v += hop_bwd(L, A, t, x, 1) * ld(in, t * nx + xd);
      st_(out, s, v);
    }
}

/* (M^dag psi)(x) = sum_y conj(M[y,x]) psi(y): the element of M that hops
 * from y = x-mu forward to x, and from y = x+mu backward to x. */
void orc_fm_dag_mul(const orc_lattice *L, const double *A, const double *in, double *out)
{
  int nt = L->nt, nx = L->nx;
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      int tu = (t + 1) % nt, td = (t - 1 + nt) % nt;
      int xu = (x + 1) % nx, xd = (x - 1 + nx) % nx;
      cplx v = L->m * ld(in, s);
      /* y = x - t: M[y, x] = hop_fwd at y */
      v += conj(hop_fwd(L, A, td, x, 0)) * ld(in, td * nx + x);
      /* y = x + t: M[y, x] = hop_bwd at y */
      v += conj(hop_bwd(L, A, tu, x, 0)) * ld(in, tu * nx + x);
      v += conj(hop_fwd(L, A, t, xd, 1)) * ld(in, t * nx + xd);
      v += conj(hop_bwd(L, A, t, xu, 1)) * ld(in, t * nx + xu);
      st_(out, s, v);
    }
}

void orc_fermion_matrix(const orc_lattice *L, const double *A, double *M)
{
  int nt = L->nt, nx = L->nx, V = nt * nx;
  memset(M, 0, sizeof(double) * 2 * (size_t)V * V);
  /* Elements accumulate, as the apply does (for an extent of 2 the
   * reference builder overwrites instead; extents here are >= 4). */
#define ADD(r, c, z)                                               \
  do {                                                             \
    size_t k_ = (size_t)(r) + (size_t)(c) * V;                     \
    cplx z_ = (z);                                                 \
    M[2 * k_] += creal(z_);                                        \
    M[2 * k_ + 1] += cimag(z_);                                    \
  } while (0)
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      int tu = (t + 1) % nt, td = (t - 1 + nt) % nt;
      int xu = (x + 1) % nx, xd = (x - 1 + nx) % nx;
      ADD(s, s, L->m);
      ADD(s, tu * nx + x, hop_fwd(L, A, t, x, 0));
      ADD(s, td * nx + x, hop_bwd(L, A, t, x, 0));
      ADD(s, t * nx + xu, hop_fwd(L, A, t, x, 1));
      ADD(s, t * nx + xd, hop_bwd(L, A, t, x, 1));
    }
#undef ADD
}

int orc_zlogdet(int n, double *Mr, double *logabs, double *arg)
{
  cplx *M = (cplx *)Mr;  /* interleaved double pairs == C99 complex layout */
  double la = 0.0;
  cplx ph = 1.0;
  int info = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = cabs(M[k + (size_t)k * n]);
    for (int r = k + 1; r < n; r++) {
      double v = cabs(M[r + (size_t)k * n]);
      if (v > best) { best = v; p = r; }
    }
    if (best == 0.0) { info = k + 1; break; }
    if (p != k) {
      for (int c = 0; c < n; c++) {
        cplx tmp = M[k + (size_t)c * n];
        M[k + (size_t)c * n] = M[p + (size_t)c * n];
        M[p + (size_t)c * n] = tmp;
      }
      ph = -ph;
    }
    cplx piv = M[k + (size_t)k * n];
    la += log(cabs(piv));
    ph *= piv / cabs(piv);
    for (int r = k + 1; r < n; r++) {
      cplx l = M[r + (size_t)k * n] / piv;
      M[r + (size_t)k * n] = l;
      if (l != 0.0)
        for (int c = k + 1; c < n; c++) M[r + (size_t)c * n] -= l * M[k + (size_t)c * n];
    }
  }
  *logabs = info ? -INFINITY : la;
  *arg = carg(ph);
  return info;
}

/* ------------------------------------------------------------------ */
/* CG on M^dag M (hmc.c:341-404)                                       */
/* ------------------------------------------------------------------ */
static double dotr(int V, const double *a, const double *b)
{
  double s = 0.0;
  for (int k = 0; k < 2 * V; k++) s += a[k] * b[k];
  return s;
}

int orc_cg_mdm(const orc_lattice *L, const double *A, const double *b, double *x,
               double tol_abs_rr, double tol_rel, int max_iter, int *status)
{
  int V = L->nt * L->nx;
  size_t bytes = sizeof(double) * 2 * (size_t)V;
  double *r = malloc(bytes), *p = malloc(bytes), *Mp = malloc(bytes), *MMp = malloc(bytes);
  memset(x, 0, bytes);
  memcpy(r, b, bytes);
  memcpy(p, b, bytes);
  double rr_old = dotr(V, r, r), rr0 = rr_old;
  int k = 0;
  *status = 0;
  double stop = tol_abs_rr;
  if (tol_rel > 0) {
    double rel = tol_rel * tol_rel * rr0;
    stop = (tol_abs_rr > 0 && tol_abs_rr > rel) ? tol_abs_rr : rel;
  }
  if (rr_old < stop || rr_old == 0.0) goto done;
  for (k = 1; k < max_iter; k++) {
    orc_fm_mul(L, A, p, Mp);
    orc_fm_dag_mul(L, A, Mp, MMp);
    double pMp = dotr(V, p, MMp);
    double a = rr_old / pMp;
    for (int i = 0; i < 2 * V; i++) x[i] += a * p[i];
    for (int i = 0; i < 2 * V; i++) r[i] -= a * MMp[i];
    double rr = dotr(V, r, r);
    if (rr < stop) break;
    if (rr / rr0 > 1e10) { *status = 2; break; }
    double be = rr / rr_old;
    for (int i = 0; i < 2 * V; i++) p[i] = r[i] + be * p[i];
    rr_old = rr;
  }
  if (k >= max_iter) *status = 1;
done:
  free(r); free(p); free(Mp); free(MMp);
  return k;
}

/* ------------------------------------------------------------------ */
/* Noise, actions, MD (hmc.c:73-101, 418-499, 504-746)                 */
/* ------------------------------------------------------------------ */
static void gauss_vector(const orc_lattice *L, orc_mt *st, double *v, int phys_norm)
{
  int V = L->nt * L->nx;
  double sc = phys_norm ? sqrt(0.5) : 1.0;
  for (int s = 0; s < V; s++) {
    double c, sn;
    orc_gauss_pair(st, &c, &sn);
    v[2 * s] = sc * c;
    v[2 * s + 1] = sc * sn;
  }
}

double orc_random_pseudofermion(const orc_lattice *L, orc_mt *st, const double *A,
                                double *xi, double *psi, int phys_norm)
{
  int V = L->nt * L->nx;
  gauss_vector(L, st, xi, phys_norm);
  orc_fm_dag_mul(L, A, xi, psi);
  return dotr(V, xi, xi);
}

void orc_stochastic_vector(const orc_lattice *L, orc_mt *st, double *v, int phys_norm)
{
  gauss_vector(L, st, v, phys_norm);
}

double orc_random_momentum(const orc_lattice *L, orc_mt *st, double *mom, int phys_norm)
{
  /* cos -> dir 0, sin -> dir 1 (hmc.c:490-491) */
  gauss_vector(L, st, mom, phys_norm);
  int V = L->nt * L->nx;
  return dotr(V, mom, mom);
}

double orc_gauge_action(const orc_lattice *L, const double *A)
{
  int V = L->nt * L->nx;
  double S = 0.0;
  for (int k = 0; k < 2 * V; k++) S += 1.0 - cos(A[k]);
  return L->nf / L->g * S;
}

double orc_pseudofermion_action(const orc_lattice *L, const double *A, const double *psi,
                                double tol_abs_rr, double tol_rel, int max_iter, int *iters)
{
  int V = L->nt * L->nx, status;
  double *chi = malloc(sizeof(double) * 2 * (size_t)V);
  int it = orc_cg_mdm(L, A, psi, chi, tol_abs_rr, tol_rel, max_iter, &status);
  if (iters) *iters = it;
  double S = dotr(V, psi, chi);
  free(chi);
  return S;
}

double orc_stochastic_md_action(const orc_lattice *L, const double *A, const double *st)
{
  int V = L->nt * L->nx;
  double *tmp = malloc(sizeof(double) * 2 * (size_t)V);
  orc_fm_dag_mul(L, A, st, tmp);
  double S = dotr(V, st, tmp);
  free(tmp);
  return S;
}

void orc_gauge_step(const orc_lattice *L, double *A, const double *mom, double eps)
{
  int V = L->nt * L->nx;
  for (int k = 0; k < 2 * V; k++) A[k] += 2.0 * eps * mom[k];
}

int orc_momentum_step(const orc_lattice *L, double *mom, const double *A, const double *psi,
                      const double *st, double eps, double tol_abs_rr, double tol_rel,
                      int max_iter, double *force_out)
{
  int nt = L->nt, nx = L->nx, V = nt * nx;
  size_t bytes = sizeof(double) * 2 * (size_t)V;
  double *F = calloc(2 * (size_t)V, sizeof(double));
  /* gauge part, dS/dA = Nf/g sin A (hmc.c:507-510) */
  for (int k = 0; k < 2 * V; k++) F[k] = L->nf / L->g * sin(A[k]);

  /* pseudofermion part (hmc.c:513-605) */
  double *chi = malloc(bytes), *Mchi = malloc(bytes);
  int status;
  int iters = orc_cg_mdm(L, A, psi, chi, tol_abs_rr, tol_rel, max_iter, &status);
  orc_fm_mul(L, A, chi, Mchi);
  double emu = exp(L->mu), emmu = exp(-L->mu);
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      for (int dir = 0; dir < 2; dir++) {
        int t2 = dir == 0 ? (t + 1) % nt : t;
        int x2 = dir == 1 ? (x + 1) % nx : x;
        int s2 = t2 * nx + x2;
        int wraps = dir == 0 ? (t == nt - 1) : (x == nx - 1);
        double sg = wraps ? wrap_sign(L, dir) : 1.0;
        double a = A[2 * s + dir], sa = sin(a), ca = cos(a);
        double eta = orc_eta(t, x, dir);
        double ef = dir == 0 ? emu : 1.0, eb = dir == 0 ? emmu : 1.0;
        /* forward hop x -> x+mu: conj(M chi)(x) chi(x+mu) */
        cplx c1 = conj(ld(Mchi, s)) * ld(chi, s2);
        double dm1 = eta * ef * (-creal(c1) * sa - cimag(c1) * ca);
        /* backward hop x+mu -> x: conj(chi)(x) (M chi)(x+mu) */
        cplx c2 = conj(ld(chi, s)) * ld(Mchi, s2);
        double dm2 = -eta * eb * (-creal(c2) * sa - cimag(c2) * ca);
        F[2 * s + dir] += -sg * dm1 - sg * dm2;
      }
    }

  /* stochastic part (hmc.c:608-661); identically zero at mu = 0.  The
   * x branch pairs st(x) with the t neighbour, as the reference does (an
   * index slip there); its two terms cancel for every input. */
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      int st2 = ((t + 1) % nt) * nx + x;
      cplx c = conj(ld(st, s)) * ld(st, st2);
      for (int dir = 0; dir < 2; dir++) {
        int wraps = dir == 0 ? (t == nt - 1) : (x == nx - 1);
        double sg = wraps ? wrap_sign(L, dir) : 1.0;
        double a = A[2 * s + dir], sa = sin(a), ca = cos(a);
        double eta = orc_eta(t, x, dir);
        double eb = dir == 0 ? emmu : 1.0, ef = dir == 0 ? emu : 1.0;
        double X = creal(c) * sa + cimag(c) * ca;
        double dm1 = -0.5 * eta * eb * X;
        double dm2 = 0.5 * eta * ef * X;
        F[2 * s + dir] += -sg * dm1 - sg * dm2;
      }
    }

  for (int k = 0; k < 2 * V; k++) mom[k] -= eps * F[k];
  if (force_out) memcpy(force_out, F, bytes);
  free(F); free(chi); free(Mchi);
  return iters;
}

int orc_update_gauge(const orc_lattice *L, orc_mt *st, double *A, int nsteps, double traj_len,
                     double tol_abs_rr, double tol_rel, int max_iter, int phys_norm,
                     double *dS_out, double *acts, int *cg_iters_total)
{
  int V = L->nt * L->nx;
  size_t bytes = sizeof(double) * 2 * (size_t)V;
  double *xi = malloc(bytes), *psi = malloc(bytes), *chi = malloc(bytes);
  double *mom = malloc(bytes), *nA = malloc(bytes);
  int its = 0, it;

  double Spf0 = orc_random_pseudofermion(L, st, A, xi, psi, phys_norm);
  double Smom0 = orc_random_momentum(L, st, mom, phys_norm);
  double Sg0 = orc_gauge_action(L, A);
  orc_stochastic_vector(L, st, chi, phys_norm);
  double Smd0 = orc_stochastic_md_action(L, A, chi);

  memcpy(nA, A, bytes);
  double h = traj_len / nsteps;
  for (int i = 0; i < nsteps; i++) {
    orc_gauge_step(L, nA, mom, 0.5 * h);
    its += orc_momentum_step(L, mom, nA, psi, chi, h, tol_abs_rr, tol_rel, max_iter, NULL);
    orc_gauge_step(L, nA, mom, 0.5 * h);
  }
  double Spf1 = orc_pseudofermion_action(L, nA, psi, tol_abs_rr, tol_rel, max_iter, &it);
  its += it;
  double Smom1 = dotr(V, mom, mom);
  double Sg1 = orc_gauge_action(L, nA);
  double Smd1 = orc_stochastic_md_action(L, nA, chi);
  double dS = (Smom1 - Smom0) + (Spf1 - Spf0) + (Smd1 - Smd0) + (Sg1 - Sg0);

  int acc = 0;
  if (exp(-dS) > orc_mt_next(st)) {
    memcpy(A, nA, bytes);
    acc = 1;
  }
  if (dS_out) *dS_out = dS;
  if (acts) {
    acts[0] = Sg0; acts[1] = Spf0; acts[2] = Smd0; acts[3] = Smom0;
    acts[4] = Sg1; acts[5] = Spf1; acts[6] = Smd1; acts[7] = Smom1;
  }
  if (cg_iters_total) *cg_iters_total = its;
  free(xi); free(psi); free(chi); free(mom); free(nA);
  return acc;
}

void orc_puregauge_sweep(const orc_lattice *L, orc_mt *st, double *A)
{
  int V = L->nt * L->nx;
  for (int k = 0; k < 2 * V; k++) {
    double s1 = -L->nf / L->g * cos(A[k]);
    double nw = 2 * M_PI * orc_mt_next(st) - M_PI;
    double s2 = -L->nf / L->g * cos(nw);
    if (orc_mt_next(st) < exp(-s2 + s1)) A[k] = nw;
  }
}

/* ------------------------------------------------------------------ */
/* Fermion-bag family                                                  */
/* ------------------------------------------------------------------ */
/* Neighbour in direction dir (0 tup, 1 xup, 2 tdn, 3 xdn), or -1 when
 * OPENX leaves the lattice (fermionbag.c:753-765). */
static inline int bag_nb_t(const orc_bag_lattice *B, int t, int dir)
{
  if (dir == 0) return (t + 1) % B->nt;
  if (dir == 2) return (t - 1 + B->nt) % B->nt;
  return t;
}
static inline int bag_nb_x(const orc_bag_lattice *B, int x, int dir)
{
  if (dir == 1) {
    if (B->bc == ORC_BC_OPENX && x == B->nx - 1) return -1;
    return (x + 1) % B->nx;
  }
  if (dir == 3) {
    if (B->bc == ORC_BC_OPENX && x == 0) return -1;
    return (x - 1 + B->nx) % B->nx;
  }
  return x;
}

double orc_fM_index(const orc_bag_lattice *B, int t1, int x1, int t2, int x2, double mu)
{
  int nt = B->nt, nx = B->nx;
  int eta_t = (x1 % 2 == 0) ? 1 : -1;  /* eta[t1][x1][0] */
  int xup = bag_nb_x(B, x1, 1), xdn = bag_nb_x(B, x1, 3);
  int tup = (t1 + 1) % nt, tdn = (t1 - 1 + nt) % nt;
  if (x1 < 0 || x2 < 0 || x1 >= nx || x2 >= nx) return 0.0;
  if (t1 == t2) {
    if (!B->with_mass_monomers && x1 == x2) return B->m;
    if (B->bc == ORC_BC_OPENX) {
      /* fermion_matrix.c:160-162: no wrap, no sign test */
      if (x2 == xup) return 0.5;
      if (x2 == xdn) return -0.5;
      return 0.0;
    }
    if (x2 == xup || x2 == xdn) {
      /* fermion_matrix.c:36-43 */
      if (nx == 2) return x2 > x1 ? 1.0 : -1.0;
      return x2 > x1 ? 0.5 : -0.5;
    }
    return 0.0;
  }
  if (x1 == x2) {
    if (t2 == tup) {
      if (nt == 2) return (t2 > t1 ? 1.0 : -1.0) * (exp(mu) + exp(-mu)) * eta_t;
      return (t2 > t1 ? 0.5 : -0.5) * exp(mu) * eta_t;
    }
    if (t2 == tdn && nt > 2) return (t2 > t1 ? 0.5 : -0.5) * exp(-mu) * eta_t;
    return 0.0;
  }
  return 0.0;
}

/* Masked apply (vec_ops.c:96-172).  A hop from (t,x) to a neighbour
 * (t2,x2) carries 1/2 eta e^{+mu} (t forward), 1/2 eta e^{-mu} (t backward)
 * or 1/2 (x), with sign + when the neighbour coordinate is larger and -
 * otherwise; the transpose flips that sign and swaps e^{+mu} <-> e^{-mu}.
 * Occupied rows are identity rows; hops into occupied sites are dropped. */
static void bag_apply(const orc_bag_lattice *B, const int *field, double *chi, const double *psi,
                      int transpose)
{
  int nt = B->nt, nx = B->nx;
  double ep = exp(B->mu), em = exp(-B->mu);
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      if (field[s] != 0) { chi[s] = psi[s]; continue; }
      double v = B->with_mass_monomers ? 0.0 : B->m * psi[s];
      double eta_t = (x % 2 == 0) ? 1.0 : -1.0;
      for (int dir = 0; dir < 4; dir++) {
        int t2 = bag_nb_t(B, t, dir), x2 = bag_nb_x(B, x, dir);
        if (x2 < 0) continue;
        int s2 = t2 * nx + x2;
        if (field[s2] != 0) continue;
        double mag;
        int up;
        if (dir == 0 || dir == 2) {
          int fwd = (dir == 0);
          mag = 0.5 * eta_t * ((fwd != transpose) ? ep : em);
          up = t2 > t;
        } else {
          mag = 0.5;
          up = x2 > x;
        }
        double sg = up ? 1.0 : -1.0;
        if (transpose) sg = -sg;
        v += sg * mag * psi[s2];
      }
      chi[s] = v;
    }
}

void orc_fM(const orc_bag_lattice *B, const int *field, double *chi, const double *psi)
{
  bag_apply(B, field, chi, psi, 0);
}

void orc_fM_transpose(const orc_bag_lattice *B, const int *field, double *chi, const double *psi)
{
  bag_apply(B, field, chi, psi, 1);
}

int orc_cg_propagator(const orc_bag_lattice *B, const int *field, double *x, const double *src,
                      double tol_abs_rr, int max_iter, int *status)
{
  int V = B->nt * B->nx;
  size_t bytes = sizeof(double) * (size_t)V;
  double *b = malloc(bytes), *r = malloc(bytes), *p = malloc(bytes);
  double *Mp = malloc(bytes), *MMp = malloc(bytes);
  orc_fM_transpose(B, field, b, src);
  memset(x, 0, bytes);
  memcpy(r, b, bytes);
  memcpy(p, b, bytes);
  double rr_old = 0.0;
  for (int i = 0; i < V; i++) rr_old += r[i] * r[i];
  double rr0 = rr_old;
  int k = 0;
  *status = 0;
  if (rr_old < tol_abs_rr) goto done;
  for (k = 1; k < max_iter; k++) {
    orc_fM(B, field, Mp, p);
    orc_fM_transpose(B, field, MMp, Mp);
    double pMp = 0.0;
    for (int i = 0; i < V; i++) pMp += p[i] * MMp[i];
    double a = rr_old / pMp;
    for (int i = 0; i < V; i++) x[i] += a * p[i];
    for (int i = 0; i < V; i++) r[i] -= a * MMp[i];
    double rr = 0.0;
    for (int i = 0; i < V; i++) rr += r[i] * r[i];
    if (rr < tol_abs_rr) break;
    if (rr / rr0 > 1e10) {
      for (int i = 0; i < V; i++) x[i] = 1e50;
      *status = 2;
      break;
    }
    double be = rr / rr_old;
    for (int i = 0; i < V; i++) p[i] = r[i] + be * p[i];
    rr_old = rr;
  }
  if (k >= max_iter) *status = 1;
done:
  free(b); free(r); free(p); free(Mp); free(MMp);
  return k;
}

int orc_dlogdet(int n, double *M, double *logabs, int *sign)
{
  double la = 0.0;
  int sg = 1, info = 0;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(M[(size_t)k * n + k]);
    for (int r = k + 1; r < n; r++) {
      double v = fabs(M[(size_t)r * n + k]);
      if (v > best) { best = v; p = r; }
    }
    if (best == 0.0) { info = k + 1; break; }
    if (p != k) {
      for (int c = 0; c < n; c++) {
        double tmp = M[(size_t)k * n + c];
        M[(size_t)k * n + c] = M[(size_t)p * n + c];
        M[(size_t)p * n + c] = tmp;
      }
      sg = -sg;
    }
    double piv = M[(size_t)k * n + k];
    la += log(fabs(piv));
    if (piv < 0) sg = -sg;
    for (int r = k + 1; r < n; r++) {
      double l = M[(size_t)r * n + k] / piv;
      M[(size_t)r * n + k] = l;
      if (l != 0.0)
        for (int c = k + 1; c < n; c++) M[(size_t)r * n + c] -= l * M[(size_t)k * n + c];
    }
  }
  *logabs = info ? -INFINITY : la;
  *sign = info ? 0 : sg;
  return info;
}

int orc_dinverse(int n, double *M)
{
  /* Gauss-Jordan with partial pivoting on [M | I]. */
  double *W = malloc(sizeof(double) * (size_t)n * 2 * n);
  int w = 2 * n, info = 0;
  for (int r = 0; r < n; r++)
    for (int c = 0; c < w; c++)
      W[(size_t)r * w + c] = c < n ? M[(size_t)r * n + c] : (c - n == r ? 1.0 : 0.0);
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(W[(size_t)k * w + k]);
    for (int r = k + 1; r < n; r++)
      if (fabs(W[(size_t)r * w + k]) > best) { best = fabs(W[(size_t)r * w + k]); p = r; }
    if (best == 0.0) { info = k + 1; break; }
    if (p != k)
      for (int c = 0; c < w; c++) {
        double tmp = W[(size_t)k * w + c];
        W[(size_t)k * w + c] = W[(size_t)p * w + c];
        W[(size_t)p * w + c] = tmp;
      }
    double inv = 1.0 / W[(size_t)k * w + k];
    for (int c = 0; c < w; c++) W[(size_t)k * w + c] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == k) continue;
      double f = W[(size_t)r * w + k];
      if (f != 0.0)
        for (int c = 0; c < w; c++) W[(size_t)r * w + c] -= f * W[(size_t)k * w + c];
    }
  }
  if (!info)
    for (int r = 0; r < n; r++)
      for (int c = 0; c < n; c++) M[(size_t)r * n + c] = W[(size_t)r * w + n + c];
  free(W);
  return info;
}

int orc_reduced_matrix(const orc_bag_lattice *B, const int *field, double mu, double *M)
{
  int nx = B->nx, V = B->nt * nx;
  int *fr = malloc(sizeof(int) * V);
  int n = 0;
  for (int s = 0; s < V; s++)
    if (field[s] == 0) fr[n++] = s;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      M[(size_t)i * n + j] = orc_fM_index(B, fr[i] / nx, fr[i] % nx, fr[j] / nx, fr[j] % nx, mu);
  free(fr);
  return n;
}

int orc_determinant_mu(const orc_bag_lattice *B, const int *field, double mu,
                       double *logabs, int *sign)
{
  int V = B->nt * B->nx, nfree = 0;
  for (int s = 0; s < V; s++) nfree += field[s] == 0;
  if (nfree == 0) { *logabs = 0.0; *sign = 1; return 0; }
  double *M = malloc(sizeof(double) * (size_t)nfree * nfree);
  int n = orc_reduced_matrix(B, field, mu, M);
  orc_dlogdet(n, M, logabs, sign);
  free(M);
  return n;
}

int orc_calc_Dinv(const orc_bag_lattice *B, double *Dinv)
{
  int nt = B->nt, nx = B->nx, V = nt * nx, h = V / 2;
  double *M = malloc(sizeof(double) * (size_t)V * V);
  for (int i = 0; i < V; i++)
    for (int j = 0; j < V; j++)
      M[(size_t)i * V + j] = orc_fM_index(B, i / nx, i % nx, j / nx, j % nx, B->mu);
  int info = orc_dinverse(V, M);
  if (info) { free(M); return info; }
  for (int i = 0; i < V; i++) {
    int pi = ((i / nx) + (i % nx)) & 1, ei = i / 2;
    for (int j = 0; j < V; j++) {
      int pj = ((j / nx) + (j % nx)) & 1, ej = j / 2;
      double v = M[(size_t)i * V + j];  /* (D^-1)_{i,j} */
      if (pi == 0 && pj == 1) Dinv[(size_t)ei * h + ej] = v;                      /* (e,o) */
      if (pi == 1 && pj == 0) Dinv[(size_t)h * h + (size_t)ei * h + ej] = v;      /* (o,e) */
    }
  }
  free(M);
  return 0;
}

int orc_bag_weight(const orc_bag_lattice *B, const int *field, const double *Dinv,
                   double *logabs, int *sign, int *n_out)
{
  int nt = B->nt, nx = B->nx, V = nt * nx, h = V / 2;
  int *ev = malloc(sizeof(int) * h), *od = malloc(sizeof(int) * h);
  int ne = 0, no = 0;
  for (int t = 0; t < nt; t++)
    for (int x = 0; x < nx; x++) {
      int s = t * nx + x;
      if (field[s] > 0) {
        if (((t + x) & 1) == 0) ev[ne++] = s / 2;
        else od[no++] = s / 2;
      }
    }
  *n_out = ne;
  int ret = 0;
  if (ne != no) { *logabs = -INFINITY; *sign = 0; ret = 1; goto out; }
  if (ne == 0) { *logabs = 0.0; *sign = 1; goto out; }
  {
    int n = ne;
    double *Ge = malloc(sizeof(double) * (size_t)n * n), *Go = malloc(sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        Ge[(size_t)i * n + j] = Dinv[(size_t)ev[i] * h + od[j]];
        Go[(size_t)i * n + j] = -Dinv[(size_t)h * h + (size_t)od[i] * h + ev[j]];
      }
    double le, lo;
    int se, so;
    int ie = orc_dlogdet(n, Ge, &le, &se), io = orc_dlogdet(n, Go, &lo, &so);
    if (ie || io) { *logabs = -INFINITY; *sign = 0; ret = 2; }
    else { *logabs = le + lo; *sign = se * so; }
    free(Ge); free(Go);
  }
out:
  free(ev); free(od);
  return ret;
}

/* Exact enumeration (tests/thirring_brute.c:456-494). */
typedef struct {
  const orc_bag_lattice *B;
  int *field;
  double *Dinv;
  double *Z, *Zabs;
  long visited;
} brute_ctx;

static void brute_visit(brute_ctx *c)
{
  const orc_bag_lattice *B = c->B;
  int V = B->nt * B->nx;
  int nl = 0, nm = 0;
  for (int s = 0; s < V; s++) {
    if (c->field[s] == 2 || c->field[s] == 3) nl++;   /* LINK_TUP, LINK_XUP */
    if (c->field[s] == 1) nm++;                        /* MONOMER */
  }
  double det = 1.0;
  if (nl + nm > 0) {
    if (nm % 2 == 0) {
      double la; int sg, n;
      int r = orc_bag_weight(B, c->field, c->Dinv, &la, &sg, &n);
      det = r ? 0.0 : sg * exp(la);
    } else {
      det = 0.0;
    }
  }
  c->Z[(size_t)nl * (V + 1) + nm] += det;
  c->Zabs[(size_t)nl * (V + 1) + nm] += fabs(det);
  c->visited++;
}

static void brute_rec(brute_ctx *c, int i0)
{
  const orc_bag_lattice *B = c->B;
  int nx = B->nx, V = B->nt * nx;
  brute_visit(c);
  for (int i = i0; i < V; i++) {
    int t = i / nx, x = i % nx;
    if (c->field[i] != 0) continue;
    for (int nu = 0; nu < 2; nu++) {
      int t2 = bag_nb_t(B, t, nu), x2 = bag_nb_x(B, x, nu);
      if (x2 < 0) continue;
      int s2 = t2 * nx + x2;
      if (c->field[s2] != 0) continue;
      c->field[i] = 2 + nu;          /* link_on: dir+2 and opp_dir+2 */
      c->field[s2] = 2 + nu + 2;
      brute_rec(c, i);
      c->field[i] = 0;
      c->field[s2] = 0;
    }
    c->field[i] = 1;
    brute_rec(c, i);
    c->field[i] = 0;
  }
}

long orc_brute_enumerate(const orc_bag_lattice *B, double *Z, double *Zabs)
{
  int V = B->nt * B->nx, h = V / 2;
  brute_ctx c;
  c.B = B;
  c.field = calloc(V, sizeof(int));
  c.Dinv = malloc(sizeof(double) * 2 * (size_t)h * h);
  c.Z = Z;
  c.Zabs = Zabs;
  c.visited = 0;
  memset(Z, 0, sizeof(double) * (size_t)(h + 1) * (V + 1));
  memset(Zabs, 0, sizeof(double) * (size_t)(h + 1) * (V + 1));
  if (orc_calc_Dinv(B, c.Dinv) == 0) brute_rec(&c, 0);
  free(c.field);
  free(c.Dinv);
  return c.visited;
}

/* ------------------------------------------------------------------ */
/* Worm estimate of the susceptibility (fermionbag.c:450-590)          */
/* ------------------------------------------------------------------ */
enum { W_LINK_TUP = 2, W_SOURCE = 6, W_NDIRS = 4 };

static int w_nb(const orc_bag_lattice *B, int s, int dir)
{
  int t = s / B->nx, x = s % B->nx;
  if (dir == 0) t = (t + 1) % B->nt;
  else if (dir == 2) t = (t - 1 + B->nt) % B->nt;
  else if (dir == 1) x = (x + 1) % B->nx;
  else x = (x - 1 + B->nx) % B->nx;
  return t * B->nx + x;
}

/* links in the order update_linklists lists them (fermionbag.c:166-178):
 * tails (values 2, 3) in site order, as dir * V + site */
static int w_links(const orc_bag_lattice *B, const int *f, int *links)
{
  int V = B->nt * B->nx, n = 0;
  for (int s = 0; s < V; s++)
    if (f[s] >= W_LINK_TUP && f[s] < W_LINK_TUP + 2) links[n++] = (f[s] - W_LINK_TUP) * V + s;
  return n;
}

static void w_link_on(const orc_bag_lattice *B, int *f, int s, int dir)
{
  int s2 = w_nb(B, s, dir);
  f[s] = dir + 2;
  f[s2] = (dir + 2) % 4 + 2;
}

/* |det D_F'F' / det D_FF| with F' = F minus occ plus fr (natural sites) */
static double w_ratio(const orc_bag_lattice *B, int *f, const int *occ, int no, const int *fr, int nf,
                      int *new_sign)
{
  int V = B->nt * B->nx;
  int *g = malloc(sizeof(int) * V);
  memcpy(g, f, sizeof(int) * V);
  for (int a = 0; a < nf; a++) g[fr[a]] = 0;
  for (int a = 0; a < no; a++) g[occ[a]] = 1;
  double l0, l1;
  int s0, s1;
  orc_determinant_mu(B, f, B->mu, &l0, &s0);
  orc_determinant_mu(B, g, B->mu, &l1, &s1);
  free(g);
  if (new_sign) *new_sign = s1;
  return s1 == 0 ? 0.0 : exp(l1 - l0);
}

long orc_worm(const orc_bag_lattice *B, int *f, orc_mt *st, double U, int n_attempts,
              int *n_links, int *n_monomers, int *sign, double *value)
{
  int V = B->nt * B->nx;
  double ld;
  int sd;
  int *empty = calloc(V, sizeof(int));
  orc_determinant_mu(B, empty, B->mu, &ld, &sd);
  free(empty);
  int *links = malloc(sizeof(int) * 2 * V);
  long steps = 0;
  double scale = (double)*n_links / (W_NDIRS * n_attempts * V);
  *n_links = w_links(B, f, links);
  if (*n_links > 0)
    for (int attempt = 0; attempt < n_attempts; attempt++) {
      int l = links[(int)(orc_mt_next(st) * *n_links)];
      int s1 = l % V, nu = l / V;
      if (orc_mt_next(st) > 0.5) s1 = w_nb(B, s1, nu);
      int dir = f[s1] - W_LINK_TUP;
      int s2 = w_nb(B, s1, dir);
      f[s1] = W_SOURCE;
      f[s2] = W_SOURCE;
      *n_monomers += 2;
      *n_links -= 1;
      steps += *sign;
      for (;;) {
        int d = (int)(W_NDIRS * orc_mt_next(st));
        int s3 = w_nb(B, s2, d);
        if (orc_mt_next(st) > 0.5) {
          if (s3 == s1) {  /* back at the start: the link again */
            f[s1] = 0;
            f[s2] = 0;
            w_link_on(B, f, s2, d);
            *n_monomers -= 2;
            *n_links += 1;
            break;
          } else if (f[s3] >= W_LINK_TUP) {  /* flip the link */
            int ld3 = f[s3] - W_LINK_TUP;
            int s4 = w_nb(B, s3, ld3);
            f[s2] = 0;
            f[s3] = 0;
            f[s4] = 0;
            w_link_on(B, f, s2, d);
            f[s4] = W_SOURCE;
            s2 = s4;
          } else if (f[s3] == 1) {  /* swap with the monomer */
            f[s2] = 1;
            f[s3] = W_SOURCE;
            s2 = s3;
          }
        } else if (B->with_mass_monomers) {
          int d2;
          do d2 = (int)(W_NDIRS * orc_mt_next(st));
          while (d2 == (d + 2) % 4);
          int s4 = w_nb(B, s3, d2);
          (void)orc_mt_next(st); /* the hop of :532-544 is never taken */
          if (f[s3] == 0 && f[s4] == 0) {  /* a new link (s3, s4), the source to s4 */
            int occ[2] = {s3, s4}, ns;
            double p = U * w_ratio(B, f, occ, 2, NULL, 0, &ns);
            if (orc_mt_next(st) < p) {
              *sign = ns * sd;  /* new_link commits the occupation of s3, s4 */
              f[s2] = 0;
              f[s3] = 0;
              f[s4] = 0;
              w_link_on(B, f, s2, d);
              f[s4] = W_SOURCE;
              s2 = s4;
            }
          } else if (f[s3] == W_LINK_TUP + d2) {  /* remove the link (s3, s4) */
            int fr[2] = {s2, s3}, ns;
            double p = w_ratio(B, f, NULL, 0, fr, 2, &ns) / U;
            if (orc_mt_next(st) < p) {
              f[s2] = 0;
              f[s3] = 0;
              f[s4] = 0;
              f[s4] = W_SOURCE;  /* the link (s2, s3) goes on and off again */
              *n_links -= 1;
              *sign = ns * sd;
              s2 = s4;
            }
          }
        } else {
          /* the hop to an empty neighbour (:565-577): the neighbour has the
           * other parity, so the ratio is 0 and the move is rejected */
          if (f[s3] == 0) (void)orc_mt_next(st);
        }
        steps += *sign;
      }
      *n_links = w_links(B, f, links);
    }
  free(links);
  *value = (double)steps * scale;
  return steps;
}
