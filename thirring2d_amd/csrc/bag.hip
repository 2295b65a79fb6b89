// bag.hip -- fermion-bag operator, solver and determinant ratios on gfx950
// (vec_ops.c:96-321, fermion_matrix.c:28-427, full_determinant.c:144-244,
// fluctuation_determinant.c:126-871).
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "../../include/th2d_bag.h"
#include "bag_geo.h"

namespace {

// --------------------------------------------------------- apply -------
__global__ void __launch_bounds__(256) k_bag_apply(BagGeo G, const int *field, double *out,
                                                   const double *in, int transpose, int nrhs)
{
  const int V = G.nt * G.nx;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  int t = s / G.nx, x = s - t * G.nx;
  for (int r = 0; r < nrhs; r++) {
    const double *src = in + (long)r * V;
    double v;
    if (field[s] != 0) v = src[s];
    else {
      v = G.wmm ? 0.0 : G.m * src[s];
      for (int dir = 0; dir < 4; dir++) {
        int t2, x2;
        bag_nb(G, t, x, dir, t2, x2);
        if (x2 < 0) continue;
        int s2 = t2 * G.nx + x2;
        if (field[s2] != 0) continue;
        v += bag_hop(G, t, x, dir, t2, x2, transpose) * src[s2];
      }
    }
    out[(long)r * V + s] = v;
  }
}

// ----------------------------------------------------------- solver ----
// One block solves one right-hand side completely: V is at most a few
// thousand sites here, so the block sweeps the lattice with __syncthreads
// between the two operator applications and the fused vector updates.
__device__ double blk_sum(double v, double *sh)
{
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  int nw = (blockDim.x + 63) >> 6;
  for (int k = 0; k < nw; k++) s += sh[k];  // every thread, same order
  __syncthreads();
  return s;
}

__device__ void blk_apply(const BagGeo &G, const int *field, double *out, const double *in,
                          int transpose)
{
  const int V = G.nt * G.nx;
  for (int s = threadIdx.x; s < V; s += blockDim.x) {
    int t = s / G.nx, x = s - t * G.nx;
    double v;
    if (field[s] != 0) v = in[s];
    else {
      v = G.wmm ? 0.0 : G.m * in[s];
      for (int dir = 0; dir < 4; dir++) {
        int t2, x2;
        bag_nb(G, t, x, dir, t2, x2);
        if (x2 < 0) continue;
        int s2 = t2 * G.nx + x2;
        if (field[s2] != 0) continue;
        v += bag_hop(G, t, x, dir, t2, x2, transpose) * in[s2];
      }
    }
    out[s] = v;
  }
}

__global__ void __launch_bounds__(1024) k_bag_cg(BagGeo G, const int *field, double *X,
                                                 const double *SRC, double *work, double tol,
                                                 int max_iter, int *iters, int *status)
{
  __shared__ double sh[16];
  const int V = G.nt * G.nx;
  const int rhs = blockIdx.x;
  double *x = X + (long)rhs * V;
  const double *src = SRC + (long)rhs * V;
  double *r = work + (long)rhs * 4 * V, *p = r + V, *q = p + V, *z = q + V;
  // b = M^T src; x = 0; r = p = b
  blk_apply(G, field, r, src, 1);
  __syncthreads();
  double acc = 0.0;
  for (int s = threadIdx.x; s < V; s += blockDim.x) {
    x[s] = 0.0;
    p[s] = r[s];
    acc += r[s] * r[s];
  }
  double rr = blk_sum(acc, sh), rr0 = rr;
  int k = 0, st = TH2D_OK;
  if (!(rr < tol)) {
    for (k = 1; k < max_iter; k++) {
      blk_apply(G, field, q, p, 0);
      __syncthreads();
      blk_apply(G, field, z, q, 1);
      __syncthreads();
      acc = 0.0;
      for (int s = threadIdx.x; s < V; s += blockDim.x) acc += p[s] * z[s];
      double pz = blk_sum(acc, sh);
      double a = rr / pz;
      acc = 0.0;
      for (int s = threadIdx.x; s < V; s += blockDim.x) {
        x[s] += a * p[s];
        double rn = r[s] - a * z[s];
        r[s] = rn;
        acc += rn * rn;
      }
      double rn2 = blk_sum(acc, sh);
      if (rn2 < tol) break;
      if (rn2 / rr0 > 1e10 || !isfinite(rn2)) {  // vec_ops.c:292-296
        for (int s = threadIdx.x; s < V; s += blockDim.x) x[s] = 1e50;
        st = TH2D_EBREAKDOWN;
        break;
      }
      double be = rn2 / rr;
      for (int s = threadIdx.x; s < V; s += blockDim.x) p[s] = r[s] + be * p[s];
      rr = rn2;
      __syncthreads();
    }
    if (k >= max_iter) st = TH2D_ENOCONV;
  }
  if (threadIdx.x == 0) {
    iters[rhs] = k;
    status[rhs] = st;
  }
}

// --------------------------------------------------------- builders ----
struct BagBatchView {
  const th2d_bag_geom *geoms;
  const int *fields;
  const long *foff;
  const long *off;
  const int *n;
};

template <typename T>
__global__ void __launch_bounds__(256) k_build_masked(BagBatchView bv, T *A)
{
  const int b = blockIdx.y;
  const th2d_bag_geom gg = bv.geoms[b];
  BagGeo G{gg.nt, gg.nx, gg.bc, gg.with_mass_monomers, gg.m, exp(gg.mu), exp(-gg.mu)};
  const int V = G.nt * G.nx;
  const int *field = bv.fields + bv.foff[b];
  T *M = A + bv.off[b];
  long total = (long)V * V;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    int i = (int)(e / V), j = (int)(e % V);
    double v;
    if (field[i] != 0 || field[j] != 0) v = (i == j) ? 1.0 : 0.0;
    else v = bag_elem(G, i, j);
    M[e] = (T)v;
  }
}

// Dinv in the reference layout from the full inverse X = D^{-1}:
// Dinv[e*h + o] = X[e_site][o_site], Dinv[h*h + o*h + e] = X[o_site][e_site]
// with eo index (nx*t + x)/2 (fermion_matrix.c:230-309).
__global__ void k_extract_dinv(int nt, int nx, const double *X, double *Dinv)
{
  const int V = nt * nx, h = V / 2;
  long total = (long)V * V;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    int i = (int)(e / V), j = (int)(e % V);
    int pi = ((i / nx) + (i % nx)) & 1, pj = ((j / nx) + (j % nx)) & 1;
    if (pi == 0 && pj == 1) Dinv[(long)(i / 2) * h + j / 2] = X[e];
    else if (pi == 1 && pj == 0) Dinv[(long)h * h + (long)(i / 2) * h + j / 2] = X[e];
  }
}

// Bag weight det_e * det_o of many small configurations (tests/
// thirring_brute.c:273-345): one thread per configuration, occupied-site
// blocks of Dinv gathered and factored in registers (n <= 16).
constexpr int SMALL_N = 16;
__device__ double small_logdet(double *a, int n, int &sign)
{
  double la = 0.0;
  int sg = 1;
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(a[k * n + k]);
    for (int r = k + 1; r < n; r++)
      if (fabs(a[r * n + k]) > best) {
        best = fabs(a[r * n + k]);
        p = r;
      }
    if (best == 0.0) {
      sign = 0;
      return -INFINITY;
    }
    if (p != k) {
      for (int c = 0; c < n; c++) {
        double t0 = a[k * n + c];
        a[k * n + c] = a[p * n + c];
        a[p * n + c] = t0;
      }
      sg = -sg;
    }
    double pv = a[k * n + k];
    la += log(fabs(pv));
    if (pv < 0) sg = -sg;
    for (int r = k + 1; r < n; r++) {
      double l = a[r * n + k] / pv;
      for (int c = k + 1; c < n; c++) a[r * n + c] -= l * a[k * n + c];
    }
  }
  sign = sg;
  return la;
}

__global__ void k_small_weights(int nt, int nx, const double *Dinv, const unsigned char *cfg,
                                long count, double *w)
{
  long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= count) return;
  const int V = nt * nx, h = V / 2;
  const unsigned char *f = cfg + c * V;
  int ev[SMALL_N], od[SMALL_N], ne = 0, no = 0, nl = 0, nm = 0;
  for (int s = 0; s < V; s++) {
    int v = f[s];
    if (v == 2 || v == 3) nl++;
    if (v == 1) nm++;
    if (v > 0) {
      int t = s / nx, x = s - t * nx;
      if (((t + x) & 1) == 0) { if (ne < SMALL_N) ev[ne] = s / 2; ne++; }
      else { if (no < SMALL_N) od[no] = s / 2; no++; }
    }
  }
  double wt;
  if (nl + nm == 0) wt = 1.0;
  else if ((nm & 1) || ne != no || ne > SMALL_N) wt = 0.0;
  else {
    int n = ne;
    double a[SMALL_N * SMALL_N];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) a[i * n + j] = Dinv[(long)ev[i] * h + od[j]];
    int s1, s2;
    double l1 = small_logdet(a, n, s1);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) a[i * n + j] = -Dinv[(long)h * h + (long)od[i] * h + ev[j]];
    double l2 = small_logdet(a, n, s2);
    wt = (s1 == 0 || s2 == 0) ? 0.0 : s1 * s2 * exp(l1 + l2);
  }
  w[c] = wt;
}

// -------------------------------------------------- bag ratio state ----
// B_ij for the index list idx (k x k), gathered into out.
__global__ void k_gather(const double *B, int V, const int *idx, int k, double *out)
{
  int e = threadIdx.x;
  if (e < k * k) out[e] = B[(long)idx[e / k] * V + idx[e % k]];
}

// cols[:, a] = sum_f B[:, f] cf[a][f]  and  rows[a, :] = sum_g cg[a][g] B[g, :]
// over up to 8 (index, coefficient) terms per a (neighbour sums of D).
struct Combo {
  int k;
  int nterm[4];
  int idx[4][8];
  double coef[4][8];
};

__global__ void k_colrow_combos(const double *B, int V, Combo cc, Combo rr, double *U, double *W)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  for (int a = 0; a < cc.k; a++) {
    double u = 0.0, w = 0.0;
    for (int q = 0; q < cc.nterm[a]; q++) u += B[(long)i * V + cc.idx[a][q]] * cc.coef[a][q];
    for (int q = 0; q < rr.nterm[a]; q++) w += rr.coef[a][q] * B[(long)rr.idx[a][q] * V + i];
    U[(long)a * V + i] = u;
    W[(long)a * V + i] = w;
  }
}

// B += U C W (U: k columns, W: k rows, C: k x k), one pass over B.
__global__ void __launch_bounds__(256) k_rank_update(double *B, int V, const double *U,
                                                     const double *W, int k, const double *C)
{
  __shared__ double sC[16];
  if (threadIdx.x < k * k) sC[threadIdx.x] = C[threadIdx.x];
  __syncthreads();
  const int i = blockIdx.y;
  double uc[4] = {0, 0, 0, 0};
  for (int b = 0; b < k; b++)
    for (int a = 0; a < k; a++) uc[b] += U[(long)a * V + i] * sC[a * k + b];
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < V; j += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < k; b++) s += uc[b] * W[(long)b * V + j];
    B[(long)i * V + j] += s;
  }
}

// Rows and columns of `idx` set to the identity (occupied sites).
__global__ void k_set_identity(double *B, int V, const int *idx, int k)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= V) return;
  for (int a = 0; a < k; a++) {
    int s = idx[a];
    B[(long)s * V + j] = (j == s) ? 1.0 : 0.0;
    B[(long)j * V + s] = (j == s) ? 1.0 : 0.0;
  }
}

// U[a][R_a] -= 1, W[a][R_a] -= 1 (columns / rows stored with stride V)
__global__ void k_sub_unit(double *U, double *W, int V, const int *R, int k)
{
  int a = threadIdx.x;
  if (a < k) {
    U[(long)a * V + R[a]] -= 1.0;
    W[(long)a * V + R[a]] -= 1.0;
  }
}

// B[s][s] += v for listed diagonal entries
__global__ void k_add_diag(double *B, int V, const int *idx, const double *v, int k)
{
  int a = threadIdx.x;
  if (a < k) B[(long)idx[a] * V + idx[a]] += v[a];
}

}  // namespace

// ================================================================ ABI ===
extern "C" int th2d_bag_apply(th2d_ctx *c, const th2d_bag_geom *g, const int *field, double *out,
                              const double *in, int transpose, int nrhs)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !field || !out || !in || nrhs < 1 || out == in) return TH2D_EARG;
  BagGeo G = bag_geo(g);
  int V = g->nt * g->nx;
  hipLaunchKernelGGL(k_bag_apply, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, field, out,
                     in, transpose ? 1 : 0, nrhs);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_bag_cg(th2d_ctx *c, const th2d_bag_geom *g, const int *field, double *x,
                           const double *src, int nrhs, double tol, int max_iter, int *iters_h,
                           int *status_h)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !field || !x || !src || nrhs < 1 || max_iter < 1 || x == src) return TH2D_EARG;
  BagGeo G = bag_geo(g);
  long V = (long)g->nt * g->nx;
  size_t work = sizeof(double) * 4 * V * nrhs;
  size_t meta = sizeof(int) * 2 * (size_t)nrhs + 256;
  TH2D_TRY(th2d_buf_reserve(c->bag, work + meta));
  double *wk = (double *)c->bag.ptr;
  int *it = (int *)((char *)c->bag.ptr + work);
  int *st = it + nrhs;
  int threads = V >= 1024 ? 1024 : 256;
  hipLaunchKernelGGL(k_bag_cg, dim3(nrhs), dim3(threads), 0, c->stream, G, field, x, src, wk, tol,
                     max_iter, it, st);
  TH2D_HIP(hipGetLastError());
  std::vector<int> h(2 * (size_t)nrhs);
  TH2D_HIP(hipMemcpyAsync(h.data(), it, sizeof(int) * 2 * nrhs, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  if (iters_h) memcpy(iters_h, h.data(), sizeof(int) * nrhs);
  if (status_h) memcpy(status_h, h.data() + nrhs, sizeof(int) * nrhs);
  for (int r = 0; r < nrhs; r++)
    if (h[nrhs + r] != TH2D_OK) return h[nrhs + r];
  return TH2D_OK;
}

extern "C" int th2d_bag_build_masked(th2d_ctx *c, const th2d_bag_geom *geoms, const int *fields,
                                     const long *foff, const th2d_batch *b, void *A, int prec)
{
  if (!c || !geoms || !fields || !foff || !b || !A || b->count < 1) return TH2D_EARG;
  BagBatchView bv{geoms, fields, foff, b->off, b->n};
  long N = b->max_n;
  dim3 g((unsigned)std::min(4096L, (N * N + 255) / 256), b->count);
  if (prec == TH2D_PREC_F64)
    hipLaunchKernelGGL(k_build_masked<double>, g, dim3(256), 0, c->stream, bv, (double *)A);
  else if (prec == TH2D_PREC_F32)
    hipLaunchKernelGGL(k_build_masked<float>, g, dim3(256), 0, c->stream, bv, (float *)A);
  else return TH2D_EARG;
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// ------------------------------------------------------ bag state ------
struct th2d_bag {
  th2d_ctx *ctx;
  th2d_bag_geom geom;
  BagGeo G;
  int V;
  double *B = nullptr;      // V x V
  double *work = nullptr;   // V x V scratch (LU)
  int *field = nullptr;     // device copy
  std::vector<int> hfield;  // host copy
  th2d_bag_geom *dgeom = nullptr;
  long *boff = nullptr;     // batch of one: off, poff
  int *bn = nullptr;
  int *piv = nullptr, *info = nullptr;
  double *dlog = nullptr;
  int *dsign = nullptr;
  double *small = nullptr;  // gather / combo scratch: U, W (4 V each), C, idx
  int *didx = nullptr;
  double logabs = 0.0, logabs_d = 0.0;
  int sign = 1, sign_d = 1;
};

static int bag_lu_inverse(th2d_bag *bg, double *logabs, int *sign)
{
  th2d_ctx *c = bg->ctx;
  const int V = bg->V;
  th2d_batch bt{1, V, bg->boff, bg->bn, bg->boff};
  TH2D_TRY(th2d_bag_build_masked(c, bg->dgeom, bg->field, bg->boff, &bt, bg->work, TH2D_PREC_F64));
  TH2D_TRY(th2d_lu_batched(c, &bt, bg->work, TH2D_PREC_F64, bg->piv, bg->info));
  TH2D_TRY(th2d_lu_logdet(c, &bt, bg->work, TH2D_PREC_F64, bg->piv, bg->info, bg->dlog, bg->dsign));
  double la;
  int sg, inf;
  TH2D_HIP(hipMemcpyAsync(&la, bg->dlog, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipMemcpyAsync(&sg, bg->dsign, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipMemcpyAsync(&inf, bg->info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  *logabs = la;
  *sign = sg;
  if (inf) return TH2D_ESINGULAR;
  return th2d_lu_inverse(c, &bt, bg->work, TH2D_PREC_F64, bg->piv, bg->B);
}

extern "C" int th2d_bag_create(th2d_ctx *c, const th2d_bag_geom *g, th2d_bag **out)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !out) return TH2D_EARG;
  th2d_bag *bg = new th2d_bag();
  bg->ctx = c;
  bg->geom = *g;
  bg->G = bag_geo(g);
  bg->V = g->nt * g->nx;
  const long V = bg->V;
  bg->hfield.assign(V, 0);
  bool ok = hipMalloc(&bg->B, sizeof(double) * V * V) == hipSuccess &&
            hipMalloc(&bg->work, sizeof(double) * V * V) == hipSuccess &&
            hipMalloc(&bg->field, sizeof(int) * V) == hipSuccess &&
            hipMalloc(&bg->dgeom, sizeof(th2d_bag_geom)) == hipSuccess &&
            hipMalloc(&bg->boff, sizeof(long) * 2) == hipSuccess &&
            hipMalloc(&bg->bn, sizeof(int) * 2) == hipSuccess &&
            hipMalloc(&bg->piv, sizeof(int) * V) == hipSuccess &&
            hipMalloc(&bg->info, sizeof(int) * 2) == hipSuccess &&
            hipMalloc(&bg->dlog, sizeof(double) * 2) == hipSuccess &&
            hipMalloc(&bg->dsign, sizeof(int) * 2) == hipSuccess &&
            hipMalloc(&bg->small, sizeof(double) * (8 * V + 64)) == hipSuccess &&
            hipMalloc(&bg->didx, sizeof(int) * 64) == hipSuccess;
  if (!ok) {
    th2d_bag_destroy(bg);
    return TH2D_ENOMEM;
  }
  long zero[2] = {0, 0};
  int nn[2] = {bg->V, bg->V};
  TH2D_HIP(hipMemcpyAsync(bg->boff, zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
  TH2D_HIP(hipMemcpyAsync(bg->bn, nn, sizeof(nn), hipMemcpyHostToDevice, c->stream));
  TH2D_HIP(hipMemcpyAsync(bg->dgeom, g, sizeof(*g), hipMemcpyHostToDevice, c->stream));
  TH2D_HIP(hipMemcpyAsync(bg->field, bg->hfield.data(), sizeof(int) * V, hipMemcpyHostToDevice,
                          c->stream));
  // det D of the empty configuration, then B = D^{-1}
  int rc = bag_lu_inverse(bg, &bg->logabs_d, &bg->sign_d);
  if (rc != TH2D_OK) {
    th2d_bag_destroy(bg);
    return rc;
  }
  bg->logabs = bg->logabs_d;
  bg->sign = bg->sign_d;
  *out = bg;
  return TH2D_OK;
}

extern "C" int th2d_bag_destroy(th2d_bag *bg)
{
  if (!bg) return TH2D_OK;
  (void)hipStreamSynchronize(bg->ctx->stream);
  for (void *p : {(void *)bg->B, (void *)bg->work, (void *)bg->field, (void *)bg->dgeom,
                  (void *)bg->boff, (void *)bg->bn, (void *)bg->piv, (void *)bg->info,
                  (void *)bg->dlog, (void *)bg->dsign, (void *)bg->small, (void *)bg->didx})
    if (p) (void)hipFree(p);
  delete bg;
  return TH2D_OK;
}

extern "C" int th2d_bag_refresh(th2d_bag *bg)
{
  if (!bg) return TH2D_EARG;
  return bag_lu_inverse(bg, &bg->logabs, &bg->sign);
}

extern "C" int th2d_bag_set_field(th2d_bag *bg, const int *field_host)
{
  if (!bg || !field_host) return TH2D_EARG;
  bg->hfield.assign(field_host, field_host + bg->V);
  TH2D_HIP(hipMemcpyAsync(bg->field, bg->hfield.data(), sizeof(int) * bg->V, hipMemcpyHostToDevice,
                          bg->ctx->stream));
  return th2d_bag_refresh(bg);
}

extern "C" int th2d_bag_logdet(th2d_bag *bg, double *la, int *sg, double *lad, int *sgd)
{
  if (!bg) return TH2D_EARG;
  if (la) *la = bg->logabs;
  if (sg) *sg = bg->sign;
  if (lad) *lad = bg->logabs_d;
  if (sgd) *sgd = bg->sign_d;
  return TH2D_OK;
}

extern "C" int th2d_bag_views(th2d_bag *bg, double **B, int **field, int *V)
{
  if (!bg) return TH2D_EARG;
  if (B) *B = bg->B;
  if (field) *field = bg->field;
  if (V) *V = bg->V;
  return TH2D_OK;
}

namespace {

// small dense helpers on the host (k <= 8)
double small_det(std::vector<double> a, int n)
{
  double d = 1.0;
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int r = k + 1; r < n; r++)
      if (fabs(a[r * n + k]) > fabs(a[p * n + k])) p = r;
    if (a[p * n + k] == 0.0) return 0.0;
    if (p != k) {
      for (int c = 0; c < n; c++) std::swap(a[k * n + c], a[p * n + c]);
      d = -d;
    }
    d *= a[k * n + k];
    for (int r = k + 1; r < n; r++) {
      double l = a[r * n + k] / a[k * n + k];
      for (int c = k; c < n; c++) a[r * n + c] -= l * a[k * n + c];
    }
  }
  return d;
}

bool small_inv(std::vector<double> &a, int n)
{
  std::vector<double> w(n * 2 * n);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < 2 * n; c++) w[r * 2 * n + c] = c < n ? a[r * n + c] : (c - n == r);
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int r = k + 1; r < n; r++)
      if (fabs(w[r * 2 * n + k]) > fabs(w[p * 2 * n + k])) p = r;
    if (w[p * 2 * n + k] == 0.0) return false;
    for (int c = 0; c < 2 * n; c++) std::swap(w[k * 2 * n + c], w[p * 2 * n + c]);
    double iv = 1.0 / w[k * 2 * n + k];
    for (int c = 0; c < 2 * n; c++) w[k * 2 * n + c] *= iv;
    for (int r = 0; r < n; r++)
      if (r != k) {
        double f = w[r * 2 * n + k];
        for (int c = 0; c < 2 * n; c++) w[r * 2 * n + c] -= f * w[k * 2 * n + c];
      }
  }
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) a[r * n + c] = w[r * 2 * n + n + c];
  return true;
}

// Neighbours of site s that are free, with D[f][s] (into s) and D[s][f].
struct NbList {
  int n = 0;
  int f[4];
  double d_in[4], d_out[4];
};

NbList free_nbrs(const th2d_bag *bg, int s, const std::vector<int> &free_now)
{
  NbList L;
  const BagGeo &G = bg->G;
  int t = s / G.nx, x = s - t * G.nx;
  for (int dir = 0; dir < 4; dir++) {
    int t2, x2;
    bag_nb(G, t, x, dir, t2, x2);
    if (x2 < 0) continue;
    int s2 = t2 * G.nx + x2;
    if (!free_now[s2]) continue;
    bool dup = false;
    for (int q = 0; q < L.n; q++) dup |= L.f[q] == s2;
    if (dup) continue;
    L.f[L.n] = s2;
    L.d_in[L.n] = bag_elem(G, s2, s);
    L.d_out[L.n] = bag_elem(G, s, s2);
    L.n++;
  }
  return L;
}

// Gather B on an index list (k x k) to the host.
int gather(th2d_bag *bg, const std::vector<int> &idx, std::vector<double> &out)
{
  int k = (int)idx.size();
  out.assign((size_t)k * k, 0.0);
  if (!k) return TH2D_OK;
  if (k * k > 1024) return TH2D_EARG;
  th2d_ctx *c = bg->ctx;
  double *dout = bg->small + 8L * bg->V;
  TH2D_HIP(hipMemcpyAsync(bg->didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, c->stream));
  // k <= 32: the gather kernel takes (k*k) threads of one block
  hipLaunchKernelGGL(k_gather, dim3(1), dim3(1024), 0, c->stream, bg->B, bg->V, bg->didx, k, dout);
  TH2D_HIP(hipMemcpyAsync(out.data(), dout, sizeof(double) * k * k, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  return TH2D_OK;
}

// Ratio and (if apply) the update for: free the sites R (occupied now),
// then occupy the sites O (free after step 1).  |R|, |O| <= 4.
int bag_change(th2d_bag *bg, const int *O, int no, const int *R, int nr, double *ratio, bool apply,
               const int *fvals)
{
  const int V = bg->V;
  if (no < 0 || nr < 0 || no > 4 || nr > 4) return TH2D_EARG;
  std::vector<int> freev(V);
  for (int s = 0; s < V; s++) freev[s] = bg->hfield[s] == 0;
  for (int a = 0; a < nr; a++)
    if (R[a] < 0 || R[a] >= V || freev[R[a]]) return TH2D_EARG;
  for (int a = 0; a < no; a++) {
    bool in_r = false;
    for (int q = 0; q < nr; q++) in_r |= R[q] == O[a];
    if (O[a] < 0 || O[a] >= V || (!freev[O[a]] && !in_r)) return TH2D_EARG;
  }
  double r1 = 1.0, r2 = 1.0;
  // ---- step 1: free R (bordering): S = D_RR - D_RF B D_FR
  std::vector<NbList> nb(nr);
  std::vector<int> idx;  // union of free neighbours of R, then O
  auto pos = [&](int s) {
    for (size_t q = 0; q < idx.size(); q++)
      if (idx[q] == s) return (int)q;
    idx.push_back(s);
    return (int)idx.size() - 1;
  };
  for (int a = 0; a < nr; a++) {
    nb[a] = free_nbrs(bg, R[a], freev);
    for (int q = 0; q < nb[a].n; q++) pos(nb[a].f[q]);
  }
  for (int a = 0; a < no; a++)
    if (freev[O[a]]) pos(O[a]);
  std::vector<double> Bs;
  TH2D_TRY(gather(bg, idx, Bs));
  const int K = (int)idx.size();
  auto Bv = [&](int i, int j) { return Bs[(size_t)pos(i) * K + pos(j)]; };
  std::vector<double> S(nr * nr), Sinv;
  for (int a = 0; a < nr; a++)
    for (int b = 0; b < nr; b++) {
      double v = bag_elem(bg->G, R[a], R[b]);
      for (int p = 0; p < nb[a].n; p++)
        for (int q = 0; q < nb[b].n; q++)
          v -= nb[a].d_out[p] * Bv(nb[a].f[p], nb[b].f[q]) * nb[b].d_in[q];
      S[a * nr + b] = v;
    }
  if (nr) {
    r1 = small_det(S, nr);
    Sinv = S;
    if (r1 == 0.0 || !small_inv(Sinv, nr)) {
      *ratio = 0.0;
      return apply ? TH2D_ESINGULAR : TH2D_OK;
    }
  }
  // B1 restricted to the O sites: for o, o' in F: B + (B D_FR) S^-1 (D_RF B);
  // for o in R: the R-blocks of B1 (-(B D_FR) S^-1, -S^-1 (D_RF B), S^-1)
  auto u_of = [&](int i, int a) {  // (B D_FR)[i][a], i free
    double u = 0.0;
    for (int q = 0; q < nb[a].n; q++) u += Bv(i, nb[a].f[q]) * nb[a].d_in[q];
    return u;
  };
  auto w_of = [&](int a, int j) {  // (D_RF B)[a][j], j free
    double w = 0.0;
    for (int q = 0; q < nb[a].n; q++) w += nb[a].d_out[q] * Bv(nb[a].f[q], j);
    return w;
  };
  auto ridx = [&](int s) {
    for (int q = 0; q < nr; q++)
      if (R[q] == s) return q;
    return -1;
  };
  std::vector<double> Boo(no * no);
  for (int a = 0; a < no; a++)
    for (int b = 0; b < no; b++) {
      int oa = O[a], ob = O[b], ra = ridx(oa), rb = ridx(ob);
      double v;
      if (ra < 0 && rb < 0) {
        v = Bv(oa, ob);
        for (int p = 0; p < nr; p++)
          for (int q = 0; q < nr; q++) v += u_of(oa, p) * Sinv[p * nr + q] * w_of(q, ob);
      } else if (ra < 0) {
        v = 0.0;
        for (int p = 0; p < nr; p++) v -= u_of(oa, p) * Sinv[p * nr + rb];
      } else if (rb < 0) {
        v = 0.0;
        for (int q = 0; q < nr; q++) v -= Sinv[ra * nr + q] * w_of(q, ob);
      } else {
        v = Sinv[ra * nr + rb];
      }
      Boo[a * no + b] = v;
    }
  if (no) r2 = small_det(Boo, no);
  *ratio = r1 * r2;
  if (!apply) return TH2D_OK;
  if (*ratio == 0.0) return TH2D_ESINGULAR;

  th2d_ctx *c = bg->ctx;
  double *U = bg->small, *W = bg->small + 4L * V;
  double *dC = bg->small + 8L * V;
  const unsigned nblk = (unsigned)((V + 255) / 256);
  if (nr) {
    // B1 = B + (u - e_R) S^-1 (w - e_R^T) - I_R (the R diagonal of B was 1)
    Combo cc{}, rw{};
    cc.k = rw.k = nr;
    for (int a = 0; a < nr; a++) {
      cc.nterm[a] = rw.nterm[a] = nb[a].n;
      for (int q = 0; q < nb[a].n; q++) {
        cc.idx[a][q] = nb[a].f[q];
        cc.coef[a][q] = nb[a].d_in[q];
        rw.idx[a][q] = nb[a].f[q];
        rw.coef[a][q] = nb[a].d_out[q];
      }
    }
    hipLaunchKernelGGL(k_colrow_combos, dim3(nblk), dim3(256), 0, c->stream, bg->B, V, cc, rw, U, W);
    // subtract e_R: U[R_a][a] -= 1, W[a][R_a] -= 1 (rows and columns of R
    // in B are unit vectors, so u and w vanish there)
    std::vector<int> rv(R, R + nr);
    TH2D_HIP(hipMemcpyAsync(bg->didx, rv.data(), sizeof(int) * nr, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_sub_unit, dim3(1), dim3(64), 0, c->stream, U, W, V, bg->didx, nr);
    TH2D_HIP(hipMemcpyAsync(dC, Sinv.data(), sizeof(double) * nr * nr, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_rank_update, dim3(std::max(1u, (unsigned)((V + 255) / 256)), V), dim3(256), 0,
                       c->stream, bg->B, V, U, W, nr, dC);
    std::vector<double> mone(nr, -1.0);
    TH2D_HIP(hipMemcpyAsync(dC + 16, mone.data(), sizeof(double) * nr, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_add_diag, dim3(1), dim3(64), 0, c->stream, bg->B, V, bg->didx, dC + 16, nr);
    TH2D_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < nr; a++) bg->hfield[R[a]] = 0;
  }
  if (no) {
    // B2 = B1 - B1[:, O] (B1_OO)^-1 B1[O, :], then identity on O
    std::vector<double> Binv = Boo;
    if (!small_inv(Binv, no)) return TH2D_ESINGULAR;
    for (double &v : Binv) v = -v;
    Combo cc{}, rw{};
    cc.k = rw.k = no;
    for (int a = 0; a < no; a++) {
      cc.nterm[a] = rw.nterm[a] = 1;
      cc.idx[a][0] = rw.idx[a][0] = O[a];
      cc.coef[a][0] = rw.coef[a][0] = 1.0;
    }
    hipLaunchKernelGGL(k_colrow_combos, dim3(nblk), dim3(256), 0, c->stream, bg->B, V, cc, rw, U, W);
    TH2D_HIP(hipMemcpyAsync(dC, Binv.data(), sizeof(double) * no * no, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_rank_update, dim3(std::max(1u, (unsigned)((V + 255) / 256)), V), dim3(256), 0,
                       c->stream, bg->B, V, U, W, no, dC);
    std::vector<int> ov(O, O + no);
    TH2D_HIP(hipMemcpyAsync(bg->didx, ov.data(), sizeof(int) * no, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_set_identity, dim3(nblk), dim3(256), 0, c->stream, bg->B, V, bg->didx, no);
    for (int a = 0; a < no; a++) bg->hfield[O[a]] = fvals ? fvals[a] : 1;
  }
  TH2D_HIP(hipMemcpyAsync(bg->field, bg->hfield.data(), sizeof(int) * V, hipMemcpyHostToDevice,
                          c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  TH2D_HIP(hipGetLastError());
  bg->logabs += log(fabs(*ratio));
  if (*ratio < 0) bg->sign = -bg->sign;
  return TH2D_OK;
}

}  // namespace

extern "C" int th2d_bag_ratio(th2d_bag *bg, const int *occ, int no, const int *fr, int nf,
                              double *ratio)
{
  if (!bg || !ratio || (no && !occ) || (nf && !fr)) return TH2D_EARG;
  return bag_change(bg, occ, no, fr, nf, ratio, false, nullptr);
}

extern "C" int th2d_bag_accept(th2d_bag *bg, const int *occ, int no, const int *fvals,
                               const int *fr, int nf)
{
  if (!bg || (no && !occ) || (nf && !fr)) return TH2D_EARG;
  double r;
  return bag_change(bg, occ, no, fr, nf, &r, true, fvals);
}

// calc_Dinv (fermion_matrix.c:191-351): full inverse of the unmasked
// operator by LU, written in the reference's eo block layout.
extern "C" int th2d_bag_calc_dinv(th2d_ctx *c, const th2d_bag_geom *g, double *Dinv)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !Dinv) return TH2D_EARG;
  th2d_bag *bg;
  TH2D_TRY(th2d_bag_create(c, g, &bg));
  long total = (long)bg->V * bg->V;
  hipLaunchKernelGGL(k_extract_dinv, dim3((unsigned)std::min(4096L, (total + 255) / 256)),
                     dim3(256), 0, c->stream, g->nt, g->nx, bg->B, Dinv);
  TH2D_HIP(hipGetLastError());
  TH2D_HIP(hipStreamSynchronize(c->stream));
  return th2d_bag_destroy(bg);
}

// Weights det_e det_o of `count` configurations (uint8 field per site,
// natural order) from Dinv (tests/thirring_brute.c:273-345).
extern "C" int th2d_bag_small_weights(th2d_ctx *c, const th2d_bag_geom *g, const double *Dinv,
                                      const unsigned char *cfg, long count, double *w)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !Dinv || !cfg || !w || count < 0) return TH2D_EARG;
  if (g->nt * g->nx / 2 > SMALL_N) return TH2D_EARG;
  if (!count) return TH2D_OK;
  hipLaunchKernelGGL(k_small_weights, dim3((unsigned)((count + 127) / 128)), dim3(128), 0,
                     c->stream, g->nt, g->nx, Dinv, cfg, count, w);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// Exact enumeration of every link/monomer configuration (tests/
// thirring_brute.c:402-494): depth-first on the host in the reference's
// order (link in direction t, then x, then a monomer at each site >= the
// last placed one), weights det_e det_o on the device in chunks, sums per
// (links, monomers) in enumeration order.
namespace {
struct BruteEnum {
  BagGeo G;
  int V;
  std::vector<int> f;
  std::vector<unsigned char> chunk;
  std::vector<int> bins;
  long n = 0, cap = 0, visited = 0;
  int rc = TH2D_OK;
  std::function<int()> flush;
  void visit()
  {
    int nl = 0, nm = 0;
    for (int s = 0; s < V; s++) {
      nl += f[s] == 2 || f[s] == 3;
      nm += f[s] == 1;
    }
    unsigned char *dst = chunk.data() + (size_t)n * V;
    for (int s = 0; s < V; s++) dst[s] = (unsigned char)f[s];
    bins[n] = nl * (V + 1) + nm;
    n++;
    visited++;
    if (n == cap && rc == TH2D_OK) rc = flush();
  }
  void rec(int i0)
  {
    visit();
    for (int i = i0; i < V && rc == TH2D_OK; i++) {
      if (f[i]) continue;
      int t = i / G.nx, x = i - t * G.nx;
      for (int nu = 0; nu < 2; nu++) {
        int t2, x2;
        bag_nb(G, t, x, nu, t2, x2);
        if (x2 < 0) continue;
        int s2 = t2 * G.nx + x2;
        if (f[s2]) continue;
        f[i] = 2 + nu;
        f[s2] = 4 + nu;
        rec(i);
        f[i] = 0;
        f[s2] = 0;
      }
      f[i] = 1;
      rec(i);
      f[i] = 0;
    }
  }
};
}  // namespace

extern "C" int th2d_bag_brute_enumerate(th2d_ctx *c, const th2d_bag_geom *g, double *Z,
                                        double *Zabs, long *visited)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !Z || !Zabs) return TH2D_EARG;
  const int V = g->nt * g->nx, h = V / 2;
  if (h > SMALL_N) return TH2D_EARG;
  const long cap = 1L << 20;
  double *Dinv = nullptr, *w = nullptr;
  unsigned char *cfg = nullptr;
  if (hipMalloc(&Dinv, sizeof(double) * 2 * h * h) != hipSuccess ||
      hipMalloc(&w, sizeof(double) * cap) != hipSuccess ||
      hipMalloc(&cfg, (size_t)cap * V) != hipSuccess) {
    if (Dinv) (void)hipFree(Dinv);
    if (w) (void)hipFree(w);
    return TH2D_ENOMEM;
  }
  int rc = th2d_bag_calc_dinv(c, g, Dinv);
  const size_t nbins = (size_t)(h + 1) * (V + 1);
  for (size_t k = 0; k < nbins; k++) Z[k] = Zabs[k] = 0.0;
  if (rc == TH2D_OK) {
    BruteEnum E;
    E.G = bag_geo(g);
    E.V = V;
    E.f.assign(V, 0);
    E.cap = cap;
    E.chunk.resize((size_t)cap * V);
    E.bins.resize(cap);
    std::vector<double> hw(cap);
    E.flush = [&]() -> int {
      if (!E.n) return TH2D_OK;
      TH2D_HIP(hipMemcpyAsync(cfg, E.chunk.data(), (size_t)E.n * V, hipMemcpyHostToDevice,
                              c->stream));
      TH2D_TRY(th2d_bag_small_weights(c, g, Dinv, cfg, E.n, w));
      TH2D_HIP(hipMemcpyAsync(hw.data(), w, sizeof(double) * E.n, hipMemcpyDeviceToHost,
                              c->stream));
      TH2D_HIP(hipStreamSynchronize(c->stream));
      for (long k = 0; k < E.n; k++) {
        Z[E.bins[k]] += hw[k];
        Zabs[E.bins[k]] += fabs(hw[k]);
      }
      E.n = 0;
      return TH2D_OK;
    };
    E.rec(0);
    if (E.rc == TH2D_OK) E.rc = E.flush();
    rc = E.rc;
    if (visited) *visited = E.visited;
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(Dinv);
  (void)hipFree(w);
  (void)hipFree(cfg);
  return rc;
}

// Matrix element of fM_index (fermion_matrix.c:28-64, 152-182) at the
// given mu, evaluated on the host by the function the kernels use.
extern "C" int th2d_bag_matrix_element(const th2d_bag_geom *g, int t1, int x1, int t2, int x2,
                                       double mu, double *out)
{
  TH2D_TRY(check_bag_geom(g));
  if (!out || t1 < 0 || t2 < 0 || x1 < 0 || x2 < 0 || t1 >= g->nt || t2 >= g->nt ||
      x1 >= g->nx || x2 >= g->nx)
    return TH2D_EARG;
  th2d_bag_geom h = *g;
  h.mu = mu;
  const BagGeo G = bag_geo(&h);
  *out = bag_elem(G, t1 * g->nx + x1, t2 * g->nx + x2);
  return TH2D_OK;
}

// det of the matrix over the free sites of a configuration
// (determinant_mu, full_determinant.c:144-186) by LU on the device.
extern "C" int th2d_bag_determinant(th2d_ctx *c, const th2d_bag_geom *g, const int *field_host,
                                    double *logabs, int *sign)
{
  TH2D_TRY(check_bag_geom(g));
  if (!c || !field_host || !logabs || !sign) return TH2D_EARG;
  th2d_bag *bg = new th2d_bag();
  bg->ctx = c;
  bg->geom = *g;
  bg->G = bag_geo(g);
  bg->V = g->nt * g->nx;
  const long V = bg->V;
  bool ok = hipMalloc(&bg->work, sizeof(double) * V * V) == hipSuccess &&
            hipMalloc(&bg->field, sizeof(int) * V) == hipSuccess &&
            hipMalloc(&bg->dgeom, sizeof(th2d_bag_geom)) == hipSuccess &&
            hipMalloc(&bg->boff, sizeof(long) * 2) == hipSuccess &&
            hipMalloc(&bg->bn, sizeof(int) * 2) == hipSuccess &&
            hipMalloc(&bg->piv, sizeof(int) * V) == hipSuccess &&
            hipMalloc(&bg->info, sizeof(int) * 2) == hipSuccess &&
            hipMalloc(&bg->dlog, sizeof(double) * 2) == hipSuccess &&
            hipMalloc(&bg->dsign, sizeof(int) * 2) == hipSuccess;
  int rc = ok ? TH2D_OK : TH2D_ENOMEM;
  if (rc == TH2D_OK) {
    long zero[2] = {0, 0};
    int nn[2] = {bg->V, bg->V};
    TH2D_HIP(hipMemcpyAsync(bg->boff, zero, sizeof(zero), hipMemcpyHostToDevice, c->stream));
    TH2D_HIP(hipMemcpyAsync(bg->bn, nn, sizeof(nn), hipMemcpyHostToDevice, c->stream));
    TH2D_HIP(hipMemcpyAsync(bg->dgeom, g, sizeof(*g), hipMemcpyHostToDevice, c->stream));
    TH2D_HIP(hipMemcpyAsync(bg->field, field_host, sizeof(int) * V, hipMemcpyHostToDevice,
                            c->stream));
    th2d_batch bt{1, bg->V, bg->boff, bg->bn, bg->boff};
    rc = th2d_bag_build_masked(c, bg->dgeom, bg->field, bg->boff, &bt, bg->work, TH2D_PREC_F64);
    if (rc == TH2D_OK) rc = th2d_lu_batched(c, &bt, bg->work, TH2D_PREC_F64, bg->piv, bg->info);
    if (rc == TH2D_OK)
      rc = th2d_lu_logdet(c, &bt, bg->work, TH2D_PREC_F64, bg->piv, bg->info, bg->dlog, bg->dsign);
    if (rc == TH2D_OK) {
      TH2D_HIP(hipMemcpyAsync(logabs, bg->dlog, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      TH2D_HIP(hipMemcpyAsync(sign, bg->dsign, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      TH2D_HIP(hipStreamSynchronize(c->stream));
    }
  }
  th2d_bag_destroy(bg);
  return rc;
}
