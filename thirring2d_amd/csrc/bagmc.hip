// bagmc.hip -- many fermion-bag configurations advanced by Metropolis
// proposals in lockstep on the device (the update loop of fermionbag.c:
// 284-379 with the determinant ratios and accept hooks of
// fluctuation_determinant.c:305-871, for a batch of matrices).
//
// Every matrix b keeps B_b = (D_FF)^-1 over the free sites of its
// configuration (V_b x V_b, identity rows and columns at occupied sites),
// its field and log|det D_FF| with its sign, all on the device.  One step
// proposes one change per matrix and applies the accepted ones:
//   k_bmc_propose   one thread per matrix: a Philox draw picks a site and a
//                   move (monomer on/off: rank 1; link on/off, monomer hop
//                   to a neighbour: rank 2), the ratio follows from a few
//                   entries of B (freeing R: det(D_RR - D_RF B D_FR);
//                   occupying O: det of the updated B on O), and a second
//                   draw makes the Metropolis decision;
//   k_bmc_snap      the columns and rows of B that the update needs;
//   k_bmc_update    B += U C W over the whole matrix (the HBM-bound part);
//   k_bmc_fix       the unit diagonal / identity rows and columns;
//   k_bmc_commit    field, log-determinant, sign and the accept counter.
// A change that frees and occupies sites (the hop) applies the freeing
// update first (pass 0), then the occupying one (pass 1).  Matrices whose
// proposal was rejected leave every kernel at once.  The host enqueues the
// steps and never waits in between.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/th2d_bag.h"
#include "bag_geo.h"

namespace {

struct BMove {
  int no, nr;
  int occ[2], fr[2], vals[2];
  int acc;
  int nbn[2];
  int nbf[2][4];
  double din[2][4], dout[2][4];
  double Sinv[4];  // (D_RR - D_RF B D_FR)^-1, nr x nr
  double Cocc[4];  // -(B1_OO)^-1, no x no
  double ratio;
};

struct BView {
  int count;
  const BagGeo *G;
  const int *V;
  const long *boff;  // start of B_b (elements)
  const long *foff;  // start of field_b; U_b, W_b start at 2 foff
  double *B;
  int *field;
  BMove *mv;
  double *U, *W;
  double *logabs;
  int *sign;
  long *nacc;
};

__device__ inline double det_small(const double *a, int n)
{
  return n == 1 ? a[0] : a[0] * a[3] - a[1] * a[2];
}

__device__ inline void inv_small(const double *a, int n, double *o)
{
  if (n == 1) {
    o[0] = 1.0 / a[0];
  } else {
    const double d = a[0] * a[3] - a[1] * a[2];
    o[0] = a[3] / d;
    o[1] = -a[1] / d;
    o[2] = -a[2] / d;
    o[3] = a[0] / d;
  }
}

__global__ void __launch_bounds__(64) k_bmc_propose(BView v, unsigned long long seed, uint32_t ctr,
                                                    th2d_bag_move *trace)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= v.count) return;
  const BagGeo G = v.G[b];
  const int V = v.V[b];
  const int *f = v.field + v.foff[b];
  const double *B = v.B + v.boff[b];
  BMove m;
  m.no = m.nr = 0;
  m.acc = 0;
  m.ratio = 0.0;
  m.nbn[0] = m.nbn[1] = 0;
  const uint2 key = make_uint2((uint32_t)seed, (uint32_t)(seed >> 32));
  const uint4 r0 = philox10(make_uint4((uint32_t)b, ctr, 0x42414731u, 0x7468u), key);
  const uint4 r1 = philox10(make_uint4((uint32_t)b, ctr, 0x42414732u, 0x7468u), key);
  const double u_site = u01(r0.x, r0.y), u_kind = u01(r0.z, r0.w);
  const double u_dir = u01(r1.x, r1.y), u_acc = u01(r1.z, r1.w);
  const int s = min((int)(u_site * V), V - 1);
  const int t = s / G.nx, x = s - t * G.nx;
  if (u_kind < 0.3) {  // a monomer on or off: rank 1
    if (f[s] == 0) {
      m.no = 1;
      m.occ[0] = s;
      m.vals[0] = 1;
    } else if (f[s] == 1) {
      m.nr = 1;
      m.fr[0] = s;
    }
  } else {  // a link on or off, or a monomer hop onto s: rank 2
    const int d = u_dir < 0.5 ? 0 : 1;
    int t2, x2;
    bag_nb(G, t, x, d, t2, x2);
    if (x2 >= 0) {
      const int s2 = t2 * G.nx + x2;
      if (f[s] == 0 && f[s2] == 0) {
        m.no = 2;
        m.occ[0] = s;
        m.occ[1] = s2;
        m.vals[0] = 2 + d;
        m.vals[1] = 4 + d;
      } else if (f[s] == 2 + d && f[s2] == 4 + d) {
        m.nr = 2;
        m.fr[0] = s;
        m.fr[1] = s2;
      } else if (f[s] == 0 && f[s2] == 1) {
        m.no = 1;
        m.occ[0] = s;
        m.vals[0] = 1;
        m.nr = 1;
        m.fr[0] = s2;
      }
    }
  }
  double r_free = 1.0, r_occ = 1.0;
  bool ok = m.no + m.nr > 0;
  if (ok && m.nr) {
    // free neighbours of each freed site, with D[nb][R] and D[R][nb]
    for (int a = 0; a < m.nr; a++) {
      const int ra = m.fr[a], ta = ra / G.nx, xa = ra - ta * G.nx;
      int n = 0;
      for (int dir = 0; dir < 4; dir++) {
        int t2, x2;
        bag_nb(G, ta, xa, dir, t2, x2);
        if (x2 < 0) continue;
        const int s2 = t2 * G.nx + x2;
        if (f[s2] != 0) continue;
        bool dup = false;
        for (int q = 0; q < n; q++) dup |= m.nbf[a][q] == s2;
        if (dup) continue;
        m.nbf[a][n] = s2;
        m.din[a][n] = bag_elem(G, s2, ra);
        m.dout[a][n] = bag_elem(G, ra, s2);
        n++;
      }
      m.nbn[a] = n;
    }
    double S[4];
    for (int a = 0; a < m.nr; a++)
      for (int c = 0; c < m.nr; c++) {
        double val = bag_elem(G, m.fr[a], m.fr[c]);
        for (int p = 0; p < m.nbn[a]; p++)
          for (int q = 0; q < m.nbn[c]; q++)
            val -= m.dout[a][p] * B[(long)m.nbf[a][p] * V + m.nbf[c][q]] * m.din[c][q];
        S[a * m.nr + c] = val;
      }
    r_free = det_small(S, m.nr);
    if (r_free == 0.0) ok = false;
    else inv_small(S, m.nr, m.Sinv);
  }
  if (ok && m.no) {
    // B after freeing R, on the sites O (all free now)
    double Boo[4];
    for (int a = 0; a < m.no; a++)
      for (int c = 0; c < m.no; c++) {
        const int oa = m.occ[a], oc = m.occ[c];
        double val = B[(long)oa * V + oc];
        for (int p = 0; p < m.nr; p++) {
          double u = 0.0;
          for (int q = 0; q < m.nbn[p]; q++) u += B[(long)oa * V + m.nbf[p][q]] * m.din[p][q];
          for (int q = 0; q < m.nr; q++) {
            double w = 0.0;
            for (int r = 0; r < m.nbn[q]; r++) w += m.dout[q][r] * B[(long)m.nbf[q][r] * V + oc];
            val += u * m.Sinv[p * m.nr + q] * w;
          }
        }
        Boo[a * m.no + c] = val;
      }
    r_occ = det_small(Boo, m.no);
    if (r_occ == 0.0) ok = false;
    else {
      inv_small(Boo, m.no, m.Cocc);
      for (int k = 0; k < m.no * m.no; k++) m.Cocc[k] = -m.Cocc[k];
    }
  }
  if (ok) {
    m.ratio = r_free * r_occ;
    m.acc = u_acc < fabs(m.ratio) ? 1 : 0;
  }
  v.mv[b] = m;
  if (trace) {
    th2d_bag_move &o = trace[b];
    o.no = m.no;
    o.nr = m.nr;
    for (int a = 0; a < 2; a++) {
      o.occ[a] = a < m.no ? m.occ[a] : -1;
      o.fr[a] = a < m.nr ? m.fr[a] : -1;
      o.vals[a] = a < m.no ? m.vals[a] : 0;
    }
    o.accepted = m.acc;
    o.pad = 0;
    o.ratio = m.ratio;
  }
}

__device__ inline bool pass_active(const BMove &m, int pass)
{
  return m.acc && (pass == 0 ? m.nr > 0 : m.no > 0);
}

// U[a][i] = column combination a of B, W[a][i] = row combination a.
__global__ void __launch_bounds__(256) k_bmc_snap(BView v, int pass)
{
  const int b = blockIdx.y;
  const BMove &m = v.mv[b];
  if (!pass_active(m, pass)) return;
  const int V = v.V[b];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const double *B = v.B + v.boff[b];
  double *U = v.U + 2 * v.foff[b], *W = v.W + 2 * v.foff[b];
  if (pass == 0) {
    // B D_FR and D_RF B, less the unit vectors of R (the rows and columns
    // of occupied sites in B are unit vectors)
    for (int a = 0; a < m.nr; a++) {
      double u = 0.0, w = 0.0;
      for (int q = 0; q < m.nbn[a]; q++) {
        u += B[(long)i * V + m.nbf[a][q]] * m.din[a][q];
        w += m.dout[a][q] * B[(long)m.nbf[a][q] * V + i];
      }
      const double e = i == m.fr[a] ? 1.0 : 0.0;
      U[(long)a * V + i] = u - e;
      W[(long)a * V + i] = w - e;
    }
  } else {
    for (int a = 0; a < m.no; a++) {
      U[(long)a * V + i] = B[(long)i * V + m.occ[a]];
      W[(long)a * V + i] = B[(long)m.occ[a] * V + i];
    }
  }
}

// B += U C W (C = Sinv in pass 0, -(B1_OO)^-1 in pass 1), rank 1 or 2.
// A thread owns column j of a band of BMC_ROWS rows: it forms (C W)_j once,
// issues the band's loads of B together, then its stores, so each row's
// accesses are contiguous across the wavefront and U is a broadcast.
constexpr int BMC_ROWS = 16;

__global__ void __launch_bounds__(256) k_bmc_update(BView v, int pass)
{
  const int b = blockIdx.y;
  const BMove &m = v.mv[b];
  if (!pass_active(m, pass)) return;
  const int V = v.V[b];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = blockIdx.z * BMC_ROWS;
  if (j >= V || i0 >= V) return;
  const int nr = min(BMC_ROWS, V - i0);
  const int k = pass == 0 ? m.nr : m.no;
  const double *C = pass == 0 ? m.Sinv : m.Cocc;
  double *__restrict__ B = v.B + v.boff[b] + (long)i0 * V + j;
  const double *__restrict__ U = v.U + 2 * v.foff[b] + i0;
  const double *__restrict__ W = v.W + 2 * v.foff[b];
  const double w0 = W[j], w1 = k == 2 ? W[V + j] : 0.0;
  const double cw0 = k == 2 ? C[0] * w0 + C[1] * w1 : C[0] * w0;
  const double cw1 = k == 2 ? C[2] * w0 + C[3] * w1 : 0.0;
  double bv[BMC_ROWS];
#pragma unroll
  for (int r = 0; r < BMC_ROWS; r++)
    if (r < nr) bv[r] = B[(long)r * V];
#pragma unroll
  for (int r = 0; r < BMC_ROWS; r++)
    if (r < nr) {
      const double acc = k == 2 ? U[r] * cw0 + U[V + r] * cw1 : U[r] * cw0;
      B[(long)r * V] = bv[r] + acc;
    }
}

// Pass 0: the R diagonal of B was 1 and is not part of B1 (B1 = B +
// (u - e_R) S^-1 (w - e_R) - I_R).  Pass 1: rows and columns of O become
// unit vectors.
__global__ void __launch_bounds__(256) k_bmc_fix(BView v, int pass)
{
  const int b = blockIdx.y;
  const BMove &m = v.mv[b];
  if (!pass_active(m, pass)) return;
  const long V = v.V[b];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  double *B = v.B + v.boff[b];
  if (pass == 0) {
    for (int a = 0; a < m.nr; a++)
      if (i == m.fr[a]) B[(long)i * V + i] -= 1.0;
  } else {
    for (int a = 0; a < m.no; a++) {
      const long o = m.occ[a];
      const double e = i == o ? 1.0 : 0.0;
      B[o * V + i] = e;
      B[(long)i * V + o] = e;
    }
  }
}

__global__ void __launch_bounds__(64) k_bmc_commit(BView v)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= v.count) return;
  const BMove &m = v.mv[b];
  if (!m.acc) return;
  int *f = v.field + v.foff[b];
  for (int a = 0; a < m.nr; a++) f[m.fr[a]] = 0;
  for (int a = 0; a < m.no; a++) f[m.occ[a]] = m.vals[a];
  v.logabs[b] += log(fabs(m.ratio));
  if (m.ratio < 0) v.sign[b] = -v.sign[b];
  v.nacc[b] += 1;
}

}  // namespace

struct th2d_bag_batch {
  th2d_ctx *ctx = nullptr;
  int count = 0, maxV = 0;
  std::vector<int> hV;
  std::vector<long> hboff, hfoff;
  long nB = 0, nF = 0;
  th2d_bag_geom *dgeom = nullptr;
  BagGeo *dG = nullptr;
  int *dV = nullptr;
  long *dboff = nullptr, *dfoff = nullptr;
  double *B = nullptr, *work = nullptr, *U = nullptr, *W = nullptr;
  int *field = nullptr, *piv = nullptr, *info = nullptr;
  BMove *mv = nullptr;
  double *logabs = nullptr;
  int *sign = nullptr;
  long *nacc = nullptr;

  BView view() const
  {
    return BView{count, dG, dV, dboff, dfoff, B, field, mv, U, W, logabs, sign, nacc};
  }
};

namespace {

template <typename T>
int dalloc(T **p, size_t n)
{
  return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)) == hipSuccess ? TH2D_OK
                                                                                 : TH2D_ENOMEM;
}

// B_b = (D_FF)^-1 of every configuration from scratch, with log|det D_FF|
// and its sign (update_background, fluctuation_determinant.c:126-301).
int bmc_refresh(th2d_bag_batch *s)
{
  th2d_ctx *c = s->ctx;
  th2d_batch bt{s->count, s->maxV, s->dboff, s->dV, s->dfoff};
  TH2D_TRY(th2d_bag_build_masked(c, s->dgeom, s->field, s->dfoff, &bt, s->work, TH2D_PREC_F64));
  TH2D_TRY(th2d_lu_batched(c, &bt, s->work, TH2D_PREC_F64, s->piv, s->info));
  TH2D_TRY(th2d_lu_logdet(c, &bt, s->work, TH2D_PREC_F64, s->piv, s->info, s->logabs, s->sign));
  std::vector<int> info(s->count);
  TH2D_HIP(hipMemcpyAsync(info.data(), s->info, sizeof(int) * s->count, hipMemcpyDeviceToHost,
                          c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  for (int v : info)
    if (v) return TH2D_ESINGULAR;
  return th2d_lu_inverse(c, &bt, s->work, TH2D_PREC_F64, s->piv, s->B);
}

}  // namespace

extern "C" int th2d_bag_batch_destroy(th2d_bag_batch *s)
{
  if (!s) return TH2D_OK;
  if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
  for (void *p : {(void *)s->dgeom, (void *)s->dG, (void *)s->dV, (void *)s->dboff,
                  (void *)s->dfoff, (void *)s->B, (void *)s->work, (void *)s->U, (void *)s->W,
                  (void *)s->field, (void *)s->piv, (void *)s->info, (void *)s->mv,
                  (void *)s->logabs, (void *)s->sign, (void *)s->nacc})
    if (p) (void)hipFree(p);
  delete s;
  return TH2D_OK;
}

extern "C" int th2d_bag_batch_create(th2d_ctx *c, int count, const th2d_bag_geom *geoms,
                                     const int *fields, th2d_bag_batch **out)
{
  // count is a grid dimension of every batch launch
  if (!c || count < 1 || count > 65535 || !geoms || !fields || !out) return TH2D_EARG;
  for (int b = 0; b < count; b++) TH2D_TRY(check_bag_geom(&geoms[b]));
  th2d_bag_batch *s = new th2d_bag_batch();
  s->ctx = c;
  s->count = count;
  std::vector<BagGeo> hG(count);
  for (int b = 0; b < count; b++) {
    const int V = geoms[b].nt * geoms[b].nx;
    s->hV.push_back(V);
    s->hboff.push_back(s->nB);
    s->hfoff.push_back(s->nF);
    s->nB += (long)V * V;
    s->nF += V;
    s->maxV = std::max(s->maxV, V);
    hG[b] = bag_geo(&geoms[b]);
  }
  int rc = TH2D_OK;
  auto A = [&](int r) {
    if (rc == TH2D_OK) rc = r;
  };
  A(dalloc(&s->dgeom, count));
  A(dalloc(&s->dG, count));
  A(dalloc(&s->dV, count));
  A(dalloc(&s->dboff, count));
  A(dalloc(&s->dfoff, count));
  A(dalloc(&s->B, s->nB));
  A(dalloc(&s->work, s->nB));
  A(dalloc(&s->U, 2 * s->nF));
  A(dalloc(&s->W, 2 * s->nF));
  A(dalloc(&s->field, s->nF));
  A(dalloc(&s->piv, s->nF));
  A(dalloc(&s->info, count));
  A(dalloc(&s->mv, count));
  A(dalloc(&s->logabs, count));
  A(dalloc(&s->sign, count));
  A(dalloc(&s->nacc, count));
  if (rc != TH2D_OK) {
    th2d_bag_batch_destroy(s);
    return rc;
  }
  hipStream_t st = c->stream;
  rc = hipMemcpyAsync(s->dgeom, geoms, sizeof(th2d_bag_geom) * count, hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(s->dG, hG.data(), sizeof(BagGeo) * count, hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(s->dV, s->hV.data(), sizeof(int) * count, hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(s->dboff, s->hboff.data(), sizeof(long) * count,
                              hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(s->dfoff, s->hfoff.data(), sizeof(long) * count,
                              hipMemcpyHostToDevice, st) ||
               hipMemcpyAsync(s->field, fields, sizeof(int) * s->nF, hipMemcpyHostToDevice, st) ||
               hipMemsetAsync(s->nacc, 0, sizeof(long) * count, st)
           ? TH2D_EHIP
           : TH2D_OK;
  if (rc == TH2D_OK) rc = bmc_refresh(s);
  if (rc != TH2D_OK) {
    th2d_bag_batch_destroy(s);
    return rc;
  }
  *out = s;
  return TH2D_OK;
}

extern "C" int th2d_bag_batch_refresh(th2d_bag_batch *s)
{
  if (!s) return TH2D_EARG;
  return bmc_refresh(s);
}

extern "C" int th2d_bag_batch_sweep(th2d_bag_batch *s, int nsteps, unsigned long long seed,
                                    unsigned long long counter, th2d_bag_move *trace)
{
  if (!s || nsteps < 0) return TH2D_EARG;
  th2d_ctx *c = s->ctx;
  const BView v = s->view();
  const unsigned g1 = (unsigned)((s->count + 63) / 64);
  const dim3 gv((unsigned)((s->maxV + 255) / 256), (unsigned)s->count);
  const dim3 gu((unsigned)((s->maxV + 255) / 256), (unsigned)s->count,
                (unsigned)((s->maxV + BMC_ROWS - 1) / BMC_ROWS));
  for (int k = 0; k < nsteps; k++) {
    th2d_bag_move *tr = trace ? trace + (size_t)k * s->count : nullptr;
    hipLaunchKernelGGL(k_bmc_propose, dim3(g1), dim3(64), 0, c->stream, v, seed,
                       (uint32_t)(counter + (unsigned long long)k), tr);
    for (int pass = 0; pass < 2; pass++) {
      hipLaunchKernelGGL(k_bmc_snap, gv, dim3(256), 0, c->stream, v, pass);
      hipLaunchKernelGGL(k_bmc_update, gu, dim3(256), 0, c->stream, v, pass);
      hipLaunchKernelGGL(k_bmc_fix, gv, dim3(256), 0, c->stream, v, pass);
    }
    hipLaunchKernelGGL(k_bmc_commit, dim3(g1), dim3(64), 0, c->stream, v);
  }
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_bag_batch_state(th2d_bag_batch *s, double *logabs, int *sign, int *fields,
                                    long *accepted)
{
  if (!s) return TH2D_EARG;
  hipStream_t st = s->ctx->stream;
  if (logabs)
    TH2D_HIP(hipMemcpyAsync(logabs, s->logabs, sizeof(double) * s->count, hipMemcpyDeviceToHost, st));
  if (sign) TH2D_HIP(hipMemcpyAsync(sign, s->sign, sizeof(int) * s->count, hipMemcpyDeviceToHost, st));
  if (fields)
    TH2D_HIP(hipMemcpyAsync(fields, s->field, sizeof(int) * s->nF, hipMemcpyDeviceToHost, st));
  if (accepted)
    TH2D_HIP(hipMemcpyAsync(accepted, s->nacc, sizeof(long) * s->count, hipMemcpyDeviceToHost, st));
  TH2D_HIP(hipStreamSynchronize(st));
  return TH2D_OK;
}
