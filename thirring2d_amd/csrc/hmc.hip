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
// solve (cg.hip), and every field update, force, operator application and
// action sum is one launch over all chains (batch.hip).
// Fields stay on the device for the whole trajectory; the host sees the
// solver's iteration counts and the action sums.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace th2d {
int gauss_fill(th2d_ctx *c, double *v, long n_complex, unsigned long long seed, uint32_t stream,
               uint32_t ctr, double scale);
}  // namespace th2d

namespace {

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
  // ---- descriptor tables of the batched site operations (batch.hip) ----
  // one table per operation, all chains in each; uploaded once
  enum { T_FOLD, T_PSI, T_DXI, T_DMOM0, T_COS0, T_TMP, T_DST, T_COPY, T_GS, T_MCHI, T_FORCE,
         T_DPC, T_DMOM1, T_COS1, T_DST1, T_N };
  std::vector<std::vector<BOp>> T(T_N, std::vector<BOp>(nch));
  for (int k = 0; k < nch; k++) {
    ChainBufs &b = B[k];
    BOp o;
    memset(&o, 0, sizeof(o));
    o.G = G[k];
    o.bc = Bc{geos[k].anti_t, geos[k].anti_x};
    double *a = dacts + 8 * k;
    const double gnf = hps[k].nf / hps[k].g;
    float *k32 = mixed ? b.K32 : nullptr;
    auto set = [&](int t, double *d0, double *d1, double *d2, double *d3, double *d4, float *f0,
                   double s0, double s1, double *out) {
      BOp q = o;
      q.d0 = d0;
      q.d1 = d1;
      q.d2 = d2;
      q.d3 = d3;
      q.d4 = d4;
      q.f0 = f0;
      q.s0 = s0;
      q.s1 = s1;
      q.out = out;
      T[t][k] = q;
    };
    set(T_FOLD, A[k], nullptr, b.K, nullptr, nullptr, k32, 0, 0, nullptr);
    set(T_PSI, b.psi, b.xi, b.K, nullptr, nullptr, nullptr, 0, 1.0, nullptr);  // M^dag xi
    set(T_DXI, b.xi, b.xi, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 1);
    set(T_DMOM0, b.mom, b.mom, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 3);
    set(T_COS0, A[k], nullptr, nullptr, nullptr, nullptr, nullptr, 0, gnf, a + 0);
    set(T_TMP, b.tmp, b.st, b.K, nullptr, nullptr, nullptr, 0, 1.0, nullptr);  // M^dag st
    set(T_DST, b.st, b.tmp, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 2);
    set(T_COPY, b.nA, A[k], nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr);
    set(T_GS, b.nA, b.mom, b.K, nullptr, nullptr, k32, 0.5 * h, 0, nullptr);
    set(T_MCHI, b.mchi, b.chi, b.K, nullptr, nullptr, nullptr, 0, 0.0, nullptr);  // M chi
    set(T_FORCE, b.mom, b.chi, b.K, b.mchi, geos[k].mu != 0.0 ? b.st : nullptr, nullptr, h, gnf,
        nullptr);
    set(T_DPC, b.psi, b.chi, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 5);
    set(T_DMOM1, b.mom, b.mom, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 7);
    set(T_COS1, b.nA, nullptr, nullptr, nullptr, nullptr, nullptr, 0, gnf, a + 4);
    set(T_DST1, b.st, b.tmp, nullptr, nullptr, nullptr, nullptr, 0, 1.0, a + 6);
  }
  long nblk = 0;
  std::vector<BOp> all;
  for (auto &t : T) {
    nblk = th2d::bop_finish(t);
    all.insert(all.end(), t.begin(), t.end());
  }
  TH2D_TRY(th2d::bop_upload(c, c->btab[2], all));
  const BOp *dt = (const BOp *)c->btab[2].ptr;
  auto run = [&](int t, int op) { return th2d::bop_launch(c, op, dt + (size_t)t * nch, nch, nblk); };

  // ---- start of trajectory ----
  TH2D_TRY(run(T_FOLD, th2d::BOP_FOLD));
  TH2D_TRY(run(T_PSI, th2d::BOP_DSLASH));
  TH2D_TRY(run(T_DXI, th2d::BOP_DOT));
  TH2D_TRY(run(T_DMOM0, th2d::BOP_DOT));
  TH2D_TRY(run(T_COS0, th2d::BOP_COSGAP));
  TH2D_TRY(run(T_TMP, th2d::BOP_DSLASH));
  TH2D_TRY(run(T_DST, th2d::BOP_DOT));
  TH2D_TRY(run(T_COPY, th2d::BOP_COPY));

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
  // Small lattices in FP64: one MD step (gauge step, the whole solve in one
  // launch, M chi, force, gauge step) has no host decision in it, so it is
  // captured once as a graph and replayed nsteps times; the solver's
  // statistics accumulate on the device and are read once.
  bool graphed = false;
  if (c->hmc_graph && th2d::md_graph_fits(c, nch, G.data(), &h0.cg)) {
    void *gs[3];
    TH2D_TRY(th2d::md_graph_prepare(c, nch, G.data(), xs.data(), bs.data(), ks.data(), &h0.cg, gs));
    if (!c->cap_stream) TH2D_HIP(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      hipStream_t own = c->stream;
      c->stream = c->cap_stream;  // the site operations enqueue into the capture
      int rc = run(T_GS, th2d::BOP_GAUGE_STEP);
      th2d::md_graph_solve(c->cap_stream, nch, gs);
      if (rc == TH2D_OK) rc = run(T_MCHI, th2d::BOP_DSLASH);
      if (rc == TH2D_OK) rc = run(T_FORCE, th2d::BOP_FORCE);
      if (rc == TH2D_OK) rc = run(T_GS, th2d::BOP_GAUGE_STEP);
      c->stream = own;
      const hipError_t e = hipStreamEndCapture(c->cap_stream, &graph);
      TH2D_TRY(rc);
      TH2D_HIP(e);
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      TH2D_HIP(ei);
      hipError_t el = hipSuccess;
      for (int s = 0; s < h0.nsteps && el == hipSuccess; s++) el = hipGraphLaunch(ex, c->stream);
      std::vector<int> its(nch), nsol(nch), st(nch);
      std::vector<double> tr(nch);
      const int rs = el == hipSuccess ? th2d::md_graph_stats(c, nch, gs, its.data(), nsol.data(),
                                                              st.data(), tr.data())
                                      : TH2D_EHIP;
      (void)hipStreamSynchronize(c->stream);
      (void)hipGraphExecDestroy(ex);
      TH2D_HIP(el);
      TH2D_TRY(rs);
      for (int k = 0; k < nch; k++) {
        R[k].cg_iterations += its[k];
        R[k].cg_solves += nsol[k];
        R[k].max_true_rel_resid = std::max(R[k].max_true_rel_resid, tr[k]);
        if (st[k] != TH2D_OK && R[k].status == TH2D_OK) R[k].status = st[k];
      }
      graphed = true;
    } else {
      (void)hipGetLastError();
    }
  }
  for (int s = 0; s < h0.nsteps && !graphed; s++) {
    TH2D_TRY(run(T_GS, th2d::BOP_GAUGE_STEP));
    TH2D_TRY(solve());
    TH2D_TRY(run(T_MCHI, th2d::BOP_DSLASH));
    TH2D_TRY(run(T_FORCE, th2d::BOP_FORCE));
    TH2D_TRY(run(T_GS, th2d::BOP_GAUGE_STEP));
  }

  // ---- end of trajectory ----
  TH2D_TRY(solve());
  TH2D_TRY(run(T_DPC, th2d::BOP_DOT));
  TH2D_TRY(run(T_DMOM1, th2d::BOP_DOT));
  TH2D_TRY(run(T_COS1, th2d::BOP_COSGAP));
  TH2D_TRY(run(T_TMP, th2d::BOP_DSLASH));
  TH2D_TRY(run(T_DST1, th2d::BOP_DOT));
  std::vector<double> acts(8 * (size_t)nch);
  TH2D_TRY(th2d::fetch_doubles(c, acts.data(), dacts, 8 * nch));
  for (int k = 0; k < nch; k++) {
    const double *a = &acts[8 * k];
    double dS = (a[7] - a[3]) + (a[5] - a[1]) + (a[6] - a[2]) + (a[4] - a[0]);
    R[k].dS = dS;
    memcpy(R[k].acts, a, sizeof(double) * 8);
    R[k].accepted = exp(-dS) > u_accept[k];
  }
  // accepted chains take the evolved field (one batched copy)
  std::vector<BOp> tacc;
  for (int k = 0; k < nch; k++)
    if (R[k].accepted) {
      BOp q;
      memset(&q, 0, sizeof(q));
      q.G = G[k];
      q.d0 = A[k];
      q.d1 = B[k].nA;
      tacc.push_back(q);
    }
  if (!tacc.empty()) {
    const long nb = th2d::bop_finish(tacc);
    TH2D_TRY(th2d::bop_upload(c, c->btab[3], tacc));
    TH2D_TRY(th2d::bop_launch(c, th2d::BOP_COPY, (const BOp *)c->btab[3].ptr, (int)tacc.size(),
                              nb));
  }
  TH2D_HIP(hipStreamSynchronize(c->stream));
  int worst = TH2D_OK;
  for (int k = 0; k < nch; k++) {
    res[k] = R[k];
    if (R[k].status != TH2D_OK && worst == TH2D_OK) worst = R[k].status;
  }
  return worst;
}

extern "C" int th2d_set_hmc_graph(th2d_ctx *c, int on)
{
  if (!c) return TH2D_EARG;
  c->hmc_graph = on ? 1 : 0;
  return TH2D_OK;
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
