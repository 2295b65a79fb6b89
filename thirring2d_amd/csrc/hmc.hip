// hmc.hip -- device-resident HMC trajectory (update_gauge, hmc.c:671-746).
//
// Order of one trajectory, as the reference:
//   refresh: xi -> psi = M^dag xi, S_pf0 = |xi|^2; momenta, S_mom0 = sum p^2;
//            S_g0; stochastic vector st, S_md0 = Re st^dag M^dag st
//   nsteps x { A += 2 (h/2) p;  p -= h dS/dA;  A += 2 (h/2) p }
//   S_pf1 = Re psi^dag (M^dag M)^{-1} psi, S_mom1, S_g1, S_md1
//   dH = sum of differences; accept if exp(-dH) > u.
// Fields stay on the device for the whole trajectory; the host sees the
// solver's iteration counts and the action sums.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>

namespace th2d {
int reduce_to_host(th2d_ctx *c, int op, long n, const double *a, const double *b, double *out,
                   double scale);
int gauss_fill(th2d_ctx *c, double *v, long n_complex, unsigned long long seed, uint32_t stream,
               uint32_t ctr, double scale);
int force_launch(th2d_ctx *c, const Geo &G, int anti_t, int anti_x, double *mom, const double *K,
                 const double *chi, const double *st, double eps, double gc, int nf, double *fout,
                 double *mchi_scratch);
}  // namespace th2d

namespace {

struct HmcBufs {
  double *nA, *K, *mom, *xi, *psi, *st, *chi, *tmp, *mchi, *stage;
  float *K32;
};

int hmc_bufs(th2d_ctx *c, long V, HmcBufs &b)
{
  size_t cv = sizeof(double) * 2 * V;  // one complex vector / one link field
  size_t need = cv * 13 + sizeof(float) * 4 * V + 8 * 256;
  TH2D_TRY(th2d_buf_reserve(c->hmc, need));
  char *p = (char *)c->hmc.ptr;
  auto take = [&](size_t n) {
    char *r = p;
    p += (n + 255) & ~(size_t)255;
    return r;
  };
  b.nA = (double *)take(cv);
  b.K = (double *)take(2 * cv);
  b.mom = (double *)take(cv);
  b.xi = (double *)take(cv);
  b.psi = (double *)take(cv);
  b.st = (double *)take(cv);
  b.chi = (double *)take(cv);
  b.tmp = (double *)take(cv);
  b.mchi = (double *)take(cv);
  b.stage = (double *)take(2 * cv);
  b.K32 = (float *)take(sizeof(float) * 4 * V);
  if ((size_t)(p - (char *)c->hmc.ptr) > need) return TH2D_EARG;  // layout exceeds the block
  return TH2D_OK;
}

enum { RED_DOT = 0, RED_COSGAP = 2 };

}  // namespace

extern "C" int th2d_hmc_trajectory(th2d_ctx *c, const th2d_geom *geo, const th2d_hmc_params *hp,
                                   double *A, const double *xi_h, const double *mom_h,
                                   const double *st_h, double u_accept, unsigned long long seed,
                                   unsigned long long counter, th2d_hmc_result *res)
{
  TH2D_TRY(th2d_check_geom(geo));
  if (!c || !hp || !A || !res || hp->nsteps < 1 || hp->g == 0.0) return TH2D_EARG;
  if ((xi_h != nullptr) != (mom_h != nullptr) || (xi_h != nullptr) != (st_h != nullptr))
    return TH2D_EARG;
  Geo G = make_geo(geo);
  const long V = 2L * G.vh;
  const size_t cv = sizeof(double) * 2 * V;
  const bool mixed = hp->cg.precision == TH2D_PREC_MIXED;
  HmcBufs b;
  TH2D_TRY(hmc_bufs(c, V, b));
  float *K32 = mixed ? b.K32 : nullptr;
  th2d_hmc_result R;
  memset(&R, 0, sizeof(R));
  R.status = TH2D_OK;

  // ---- noise ----
  const double sc = hp->phys_norm ? sqrt(0.5) : 1.0;
  if (xi_h) {
    // host arrays in natural order (reference draw order), scaled here
    TH2D_HIP(hipMemcpyAsync(b.stage, xi_h, cv, hipMemcpyHostToDevice, c->stream));
    TH2D_TRY(th2d_vec_to_eo(c, geo, b.xi, b.stage));
    TH2D_HIP(hipMemcpyAsync(b.stage, mom_h, cv, hipMemcpyHostToDevice, c->stream));
    TH2D_TRY(th2d_links_to_eo(c, geo, b.mom, b.stage));
    TH2D_HIP(hipMemcpyAsync(b.stage, st_h, cv, hipMemcpyHostToDevice, c->stream));
    TH2D_TRY(th2d_vec_to_eo(c, geo, b.st, b.stage));
    if (sc != 1.0) {
      TH2D_TRY(th2d::axpy_like(c, 2, 2 * V, sc, b.xi, b.xi, nullptr));
      TH2D_TRY(th2d::axpy_like(c, 2, 2 * V, sc, b.mom, b.mom, nullptr));
      TH2D_TRY(th2d::axpy_like(c, 2, 2 * V, sc, b.st, b.st, nullptr));
    }
  } else {
    uint32_t ctr = (uint32_t)counter;
    TH2D_TRY(th2d::gauss_fill(c, b.xi, V, seed, 1u, ctr, sc));
    TH2D_TRY(th2d::gauss_fill(c, b.mom, V, seed, 2u, ctr, sc));
    TH2D_TRY(th2d::gauss_fill(c, b.st, V, seed, 3u, ctr, sc));
  }

  const double h = hp->traj_len / hp->nsteps;
  double acts[8];
  // ---- start of trajectory ----
  TH2D_TRY(th2d_fold_links(c, geo, b.K, K32, A));
  TH2D_TRY(th2d::dslash_launch(c, G, b.psi, b.xi, b.K, 1, TH2D_PREC_F64, 1));  // psi = M^dag xi
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.xi, b.xi, &acts[1], 1.0));
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.mom, b.mom, &acts[3], 1.0));
  TH2D_TRY(th2d::reduce_to_host(c, RED_COSGAP, 2 * V, A, nullptr, &acts[0], hp->nf / hp->g));
  TH2D_TRY(th2d::dslash_launch(c, G, b.tmp, b.st, b.K, 1, TH2D_PREC_F64, 1));
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.st, b.tmp, &acts[2], 1.0));
  TH2D_HIP(hipMemcpyAsync(b.nA, A, cv, hipMemcpyDeviceToDevice, c->stream));

  auto solve = [&](double *x, const double *rhs) -> int {
    th2d_cg_info inf;
    int rc = th2d_cg_mdagm(c, geo, x, rhs, b.K, K32, &hp->cg, &inf);
    R.cg_iterations += inf.iterations;
    R.cg_solves += 1;
    R.max_true_rel_resid = std::max(R.max_true_rel_resid, inf.true_rel_resid);
    if (rc == TH2D_ENOCONV || rc == TH2D_EBREAKDOWN) {
      if (R.status == TH2D_OK) R.status = rc;
      return TH2D_OK;  // recorded; the trajectory continues as the reference does
    }
    return rc;
  };

  // ---- leapfrog (hmc.c:708-716) ----
  for (int s = 0; s < hp->nsteps; s++) {
    TH2D_TRY(th2d_gauge_step(c, geo, b.nA, b.mom, 0.5 * h, b.K, K32));
    TH2D_TRY(solve(b.chi, b.psi));
    TH2D_TRY(th2d::force_launch(c, G, geo->anti_t, geo->anti_x, b.mom, b.K, b.chi,
                                geo->mu != 0.0 ? b.st : nullptr, h, hp->g, hp->nf, nullptr,
                                b.mchi));
    TH2D_TRY(th2d_gauge_step(c, geo, b.nA, b.mom, 0.5 * h, b.K, K32));
  }

  // ---- end of trajectory ----
  TH2D_TRY(solve(b.chi, b.psi));
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.psi, b.chi, &acts[5], 1.0));
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.mom, b.mom, &acts[7], 1.0));
  TH2D_TRY(th2d::reduce_to_host(c, RED_COSGAP, 2 * V, b.nA, nullptr, &acts[4], hp->nf / hp->g));
  TH2D_TRY(th2d::dslash_launch(c, G, b.tmp, b.st, b.K, 1, TH2D_PREC_F64, 1));
  TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * V, b.st, b.tmp, &acts[6], 1.0));

  double dS = (acts[7] - acts[3]) + (acts[5] - acts[1]) + (acts[6] - acts[2]) + (acts[4] - acts[0]);
  R.dS = dS;
  memcpy(R.acts, acts, sizeof(acts));
  R.accepted = exp(-dS) > u_accept;
  if (R.accepted) TH2D_HIP(hipMemcpyAsync(A, b.nA, cv, hipMemcpyDeviceToDevice, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  *res = R;
  return R.status;
}
