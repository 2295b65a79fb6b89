// force.hip -- HMC force dS/dA on gfx950 (momentum_step, hmc.c:504-665).
//
// With chi = (M^dag M)^{-1} psi, Mchi = M chi and K = 1/2 eta s e^{iA}:
//   dS/dA_mu(x) = Nf/g sin A_mu(x)
//     - 2 Re[ i K_mu(x) ( e^{+mu_mu} conj(Mchi(x)) chi(x+mu)
//                        - e^{-mu_mu} conj(chi(x)) Mchi(x+mu) ) ]
//     + (e^{mu_mu} - e^{-mu_mu}) Re[ i K_mu(x) conj(st(x)) st(x+mu) ]
// (mu_t = mu, mu_x = 0).  The first line is the gauge force (hmc.c:507-510),
// the second the pseudofermion force (hmc.c:517-605) and the third the
// stochastic-term force (hmc.c:608-661), which vanishes at mu = 0 and in x
// for every mu.  sin A is recovered from K exactly: sin A = 2 eta s Im K.
#include "th2d_internal.h"

namespace {

__global__ void __launch_bounds__(256) k_force(Geo G, int anti_t, int anti_x,
                                               const double2 *__restrict__ K,
                                               const double2 *__restrict__ chi,
                                               const double2 *__restrict__ mchi,
                                               const double2 *__restrict__ st, double cg,
                                               double eps, double *__restrict__ mom,
                                               double *__restrict__ fout)
{
  const int V = 2 * G.vh;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  double F[2];
  force_site(G, Bc{anti_t, anti_x}, K, chi, mchi, st, cg, i, F);
  if (fout) {
    fout[i] = F[0];
    fout[V + i] = F[1];
  } else {
    mom[i] -= eps * F[0];
    mom[V + i] -= eps * F[1];
  }
}

}  // namespace

namespace th2d {
int force_launch(th2d_ctx *c, const Geo &G, int anti_t, int anti_x, double *mom, const double *K,
                 const double *chi, const double *st, double eps, double gc, int nf, double *fout,
                 double *mchi_scratch)
{
  int V = 2 * G.vh;
  TH2D_TRY(dslash_launch(c, G, mchi_scratch, chi, K, 0, TH2D_PREC_F64, 1));
  hipLaunchKernelGGL(k_force, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, anti_t, anti_x,
                     (const double2 *)K, (const double2 *)chi, (const double2 *)mchi_scratch,
                     (const double2 *)st, nf / gc, eps, mom, fout);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}
}  // namespace th2d

extern "C" int th2d_force(th2d_ctx *c, const th2d_geom *g, double *mom, const double *A,
                          const double *K, const double *chi, const double *st, double eps,
                          double gc, int nf, double *fout)
{
  TH2D_TRY(th2d_check_geom(g));
  (void)A;  // the gauge force is taken from K (sin A = 2 eta s Im K)
  if (!c || !K || !chi || gc == 0.0 || (!mom && !fout)) return TH2D_EARG;
  Geo G = make_geo(g);
  size_t vb = sizeof(double) * 2 * (size_t)(2 * G.vh);
  TH2D_TRY(th2d_buf_reserve(c->force, vb));
  return th2d::force_launch(c, G, g->anti_t, g->anti_x, mom, K, chi, st, eps, gc, nf, fout,
                            (double *)c->force.ptr);
}
