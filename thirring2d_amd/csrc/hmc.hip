// hmc.hip -- device-resident HMC trajectories (update_gauge, hmc.c:671-746)
// for one chain or a batch of independent chains of any lattice sizes.
//
// Order of one trajectory, as the reference:
//   refresh: xi -> psi = M^dag xi, S_pf0 = |xi|^2; momenta, S_mom0 = sum p^2;
//            S_g0; stochastic vector st, S_md0 = Re st^dag M^dag st
//   nsteps x { A += 2 (h/2) p;  p -= h dS/dA;  A += 2 (h/2) p }
//   S_pf1 = Re psi^dag (M^dag M)^{-1} psi, S_mom1, S_g1, S_md1
//   dH = sum of differences; accept if exp(-dH) > u.
// The chains of a batch advance in lockstep: every inversion is one batched
// solve (cg.hip), the field updates and forces are per-chain launches.
// Fields stay on the device for the whole trajectory; the host sees the
// solver's iteration counts and the action sums.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace th2d {
int reduce_op(th2d_ctx *c, int op, long n, const double *a, const double *b, double *dout,
              double scale);
int gauss_fill(th2d_ctx *c, double *v, long n_complex, unsigned long long seed, uint32_t stream,
               uint32_t ctr, double scale);
int force_launch(th2d_ctx *c, const Geo &G, int anti_t, int anti_x, double *mom, const double *K,
                 const double *chi, const double *st, double eps, double gc, int nf, double *fout,
                 double *mchi_scratch);
}  // namespace th2d

namespace {

enum { RED_DOT = 0, RED_COSGAP = 2 };

struct ChainBufs {
  double *nA, *K, *mom, *xi, *psi, *st, *chi, *tmp, *mchi, *stage;
  float *K32;
};

// 13 field-sized blocks per chain (cv = one complex vector or link field)
size_t chain_bytes(long V)
{
  size_t cv = sizeof(double) * 2 * V;
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  return 9 * al(cv) + 2 * al(2 * cv) + al(sizeof(float) * 4 * V);
}

void carve(char *p, long V, ChainBufs &b)
{
  size_t cv = sizeof(double) * 2 * V;
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
}

}  // namespace

extern "C" int th2d_hmc_trajectory_batched(th2d_ctx *c, int nch, const th2d_geom *geos,
                                           const th2d_hmc_params *hps, double **A,
                                           const double *const *xi_h, const double *const *mom_h,
                                           const double *const *st_h, const double *u_accept,
                                           unsigned long long seed,
                                           const unsigned long long *counters,
                                           th2d_hmc_result *res)
{
  if (!c || nch < 1 || !geos || !hps || !A || !u_accept || !res) return TH2D_EARG;
  const th2d_hmc_params &h0 = hps[0];
  for (int k = 0; k < nch; k++) {
    TH2D_TRY(th2d_check_geom(&geos[k]));
    const th2d_hmc_params &hk = hps[k];
    if (!A[k] || hk.nsteps != h0.nsteps || hk.traj_len != h0.traj_len || hk.g == 0.0 ||
        memcmp(&hk.cg, &h0.cg, sizeof(th2d_cg_params)) != 0 || hk.phys_norm != h0.phys_norm)
      return TH2D_EARG;
    bool host_noise = xi_h && xi_h[k];
    if (host_noise && (!mom_h || !mom_h[k] || !st_h || !st_h[k])) return TH2D_EARG;
  }
  if (h0.nsteps < 1) return TH2D_EARG;
  const bool mixed = h0.cg.precision == TH2D_PREC_MIXED;
  std::vector<Geo> G(nch);
  std::vector<long> V(nch);
  std::vector<size_t> off(nch + 1, 0);
  for (int k = 0; k < nch; k++) {
    G[k] = make_geo(&geos[k]);
    V[k] = 2L * G[k].vh;
    off[k + 1] = off[k] + chain_bytes(V[k]);
  }
  TH2D_TRY(th2d_buf_reserve(c->hmc, off[nch] + sizeof(double) * 8 * nch + 256));
  std::vector<ChainBufs> B(nch);
  for (int k = 0; k < nch; k++) carve((char *)c->hmc.ptr + off[k], V[k], B[k]);
  double *dacts = (double *)((char *)c->hmc.ptr + off[nch]);  // [nch][8]
  std::vector<th2d_hmc_result> R(nch);
  for (auto &r : R) {
    memset(&r, 0, sizeof(r));
    r.status = TH2D_OK;
  }

  // ---- noise ----
  const double sc = h0.phys_norm ? sqrt(0.5) : 1.0;
  for (int k = 0; k < nch; k++) {
    const size_t cv = sizeof(double) * 2 * V[k];
    ChainBufs &b = B[k];
    if (xi_h && xi_h[k]) {
      // host arrays in natural order (reference draw order), scaled here
      TH2D_HIP(hipMemcpyAsync(b.stage, xi_h[k], cv, hipMemcpyHostToDevice, c->stream));
      TH2D_TRY(th2d_vec_to_eo(c, &geos[k], b.xi, b.stage));
      TH2D_HIP(hipMemcpyAsync(b.stage, mom_h[k], cv, hipMemcpyHostToDevice, c->stream));
      TH2D_TRY(th2d_links_to_eo(c, &geos[k], b.mom, b.stage));
      TH2D_HIP(hipMemcpyAsync(b.stage, st_h[k], cv, hipMemcpyHostToDevice, c->stream));
      TH2D_TRY(th2d_vec_to_eo(c, &geos[k], b.st, b.stage));
      TH2D_HIP(hipStreamSynchronize(c->stream));  // host buffers may go away
      if (sc != 1.0) {
        TH2D_TRY(th2d::axpy_like(c, 2, 2 * V[k], sc, b.xi, b.xi, nullptr));
        TH2D_TRY(th2d::axpy_like(c, 2, 2 * V[k], sc, b.mom, b.mom, nullptr));
        TH2D_TRY(th2d::axpy_like(c, 2, 2 * V[k], sc, b.st, b.st, nullptr));
      }
    } else {
      uint32_t ctr = (uint32_t)(counters ? counters[k] : 0);
      // streams 1..3 per chain: xi, momenta, stochastic vector
      TH2D_TRY(th2d::gauss_fill(c, b.xi, V[k], seed, 1u + 4u * (uint32_t)k, ctr, sc));
      TH2D_TRY(th2d::gauss_fill(c, b.mom, V[k], seed, 2u + 4u * (uint32_t)k, ctr, sc));
      TH2D_TRY(th2d::gauss_fill(c, b.st, V[k], seed, 3u + 4u * (uint32_t)k, ctr, sc));
    }
  }

  const double h = h0.traj_len / h0.nsteps;
  // ---- start of trajectory ----
  for (int k = 0; k < nch; k++) {
    ChainBufs &b = B[k];
    const double gnf = hps[k].nf / hps[k].g;
    double *a = dacts + 8 * k;
    TH2D_TRY(th2d_fold_links(c, &geos[k], b.K, mixed ? b.K32 : nullptr, A[k]));
    TH2D_TRY(th2d::dslash_launch(c, G[k], b.psi, b.xi, b.K, 1, TH2D_PREC_F64, 1));
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.xi, b.xi, a + 1, 1.0));
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.mom, b.mom, a + 3, 1.0));
    TH2D_TRY(th2d::reduce_op(c, RED_COSGAP, 2 * V[k], A[k], nullptr, a + 0, gnf));
    TH2D_TRY(th2d::dslash_launch(c, G[k], b.tmp, b.st, b.K, 1, TH2D_PREC_F64, 1));
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.st, b.tmp, a + 2, 1.0));
    TH2D_HIP(hipMemcpyAsync(b.nA, A[k], sizeof(double) * 2 * V[k], hipMemcpyDeviceToDevice,
                            c->stream));
  }

  std::vector<double *> xs(nch);
  std::vector<const double *> bs(nch), ks(nch);
  std::vector<const float *> k32(nch);
  for (int k = 0; k < nch; k++) {
    xs[k] = B[k].chi;
    bs[k] = B[k].psi;
    ks[k] = B[k].K;
    k32[k] = B[k].K32;
  }
  std::vector<th2d_cg_info> inf(nch);
  auto solve = [&]() -> int {
    int rc = th2d_cg_mdagm_batched(c, nch, geos, xs.data(), bs.data(), ks.data(),
                                   mixed ? k32.data() : nullptr, &h0.cg, inf.data());
    for (int k = 0; k < nch; k++) {
      R[k].cg_iterations += inf[k].iterations;
      R[k].cg_solves += 1;
      R[k].max_true_rel_resid = std::max(R[k].max_true_rel_resid, inf[k].true_rel_resid);
      if ((inf[k].status == TH2D_ENOCONV || inf[k].status == TH2D_EBREAKDOWN) &&
          R[k].status == TH2D_OK)
        R[k].status = inf[k].status;
    }
    // non-convergence is recorded per chain; the trajectory continues as
    // the reference does
    return (rc == TH2D_ENOCONV || rc == TH2D_EBREAKDOWN) ? TH2D_OK : rc;
  };

  // ---- leapfrog (hmc.c:708-716) ----
  for (int s = 0; s < h0.nsteps; s++) {
    for (int k = 0; k < nch; k++)
      TH2D_TRY(th2d_gauge_step(c, &geos[k], B[k].nA, B[k].mom, 0.5 * h, B[k].K,
                               mixed ? B[k].K32 : nullptr));
    TH2D_TRY(solve());
    for (int k = 0; k < nch; k++)
      TH2D_TRY(th2d::force_launch(c, G[k], geos[k].anti_t, geos[k].anti_x, B[k].mom, B[k].K,
                                  B[k].chi, geos[k].mu != 0.0 ? B[k].st : nullptr, h, hps[k].g,
                                  hps[k].nf, nullptr, B[k].mchi));
    for (int k = 0; k < nch; k++)
      TH2D_TRY(th2d_gauge_step(c, &geos[k], B[k].nA, B[k].mom, 0.5 * h, B[k].K,
                               mixed ? B[k].K32 : nullptr));
  }

  // ---- end of trajectory ----
  TH2D_TRY(solve());
  for (int k = 0; k < nch; k++) {
    ChainBufs &b = B[k];
    double *a = dacts + 8 * k;
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.psi, b.chi, a + 5, 1.0));
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.mom, b.mom, a + 7, 1.0));
    TH2D_TRY(th2d::reduce_op(c, RED_COSGAP, 2 * V[k], b.nA, nullptr, a + 4,
                             hps[k].nf / hps[k].g));
    TH2D_TRY(th2d::dslash_launch(c, G[k], b.tmp, b.st, b.K, 1, TH2D_PREC_F64, 1));
    TH2D_TRY(th2d::reduce_op(c, RED_DOT, 2 * V[k], b.st, b.tmp, a + 6, 1.0));
  }
  std::vector<double> acts(8 * (size_t)nch);
  TH2D_TRY(th2d::fetch_doubles(c, acts.data(), dacts, 8 * nch));
  for (int k = 0; k < nch; k++) {
    const double *a = &acts[8 * k];
    double dS = (a[7] - a[3]) + (a[5] - a[1]) + (a[6] - a[2]) + (a[4] - a[0]);
    R[k].dS = dS;
    memcpy(R[k].acts, a, sizeof(double) * 8);
    R[k].accepted = exp(-dS) > u_accept[k];
    if (R[k].accepted)
      TH2D_HIP(hipMemcpyAsync(A[k], B[k].nA, sizeof(double) * 2 * V[k], hipMemcpyDeviceToDevice,
                              c->stream));
  }
  TH2D_HIP(hipStreamSynchronize(c->stream));
  int worst = TH2D_OK;
  for (int k = 0; k < nch; k++) {
    res[k] = R[k];
    if (R[k].status != TH2D_OK && worst == TH2D_OK) worst = R[k].status;
  }
  return worst;
}

extern "C" int th2d_hmc_trajectory(th2d_ctx *c, const th2d_geom *geo, const th2d_hmc_params *hp,
                                   double *A, const double *xi_h, const double *mom_h,
                                   const double *st_h, double u_accept, unsigned long long seed,
                                   unsigned long long counter, th2d_hmc_result *res)
{
  if ((xi_h != nullptr) != (mom_h != nullptr) || (xi_h != nullptr) != (st_h != nullptr))
    return TH2D_EARG;
  double *As[1] = {A};
  const double *xs[1] = {xi_h}, *ms[1] = {mom_h}, *ss[1] = {st_h};
  unsigned long long ctr[1] = {counter};
  return th2d_hmc_trajectory_batched(c, 1, geo, hp, As, xi_h ? xs : nullptr, xi_h ? ms : nullptr,
                                     xi_h ? ss : nullptr, &u_accept, seed, ctr, res);
}
