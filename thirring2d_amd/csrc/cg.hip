// cg.hip -- conjugate-gradient solver for M^dag M on gfx950
// (fmdm_invert_cg, hmc.c:341-404; same Krylov iterates, different
// schedule).
//
// Schedule: Chronopoulos-Gear CG, whose two inner products (r.r, r.Ar) are
// taken together, so one grid-wide synchronisation per iteration
// suffices; it is the boundary between two launches of one fused step
// kernel.  Each workgroup owns TT x TX tiles of the lattice.  Iteration j
// (one launch) does, per tile:
//   1. reads r_j, w_j = A r_j and s_{j-1} on the tile plus a 2-site halo,
//      forms s_j = w_j + beta s_{j-1} and r_{j+1} = r_j - alpha s_j there
//      (the halo values are recomputed, not exchanged);
//   2. applies M to r_{j+1} on the tile plus a 1-site halo (LDS);
//   3. applies M^dag to that on the tile: w_{j+1} = A r_{j+1};
//   4. updates p_j = r_j + beta p_{j-1}, x_{j+1} = x_j + alpha p_j in place;
//   5. writes its partial (r.r, r.w).
// The next launch starts by adding all partials in a fixed order (every
// block the same way, so all take bitwise the same alpha/beta and stop
// decision).  r, s, w, the partials and the scalar state are
// double-buffered by iteration parity; p, x are only touched by their
// owner.  The host enqueues launches in chunks and reads the stop flag once
// per chunk; launches after the stop return at once.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>

namespace {

struct CgState {
  double gam_old, alpha_old;
  int j, pad;
};

struct CgResult {
  int iters, status;
  double rr;
  int done, pad;
};

template <typename T> struct CgArgs {
  Geo G;
  int TT, TX, tiles_x, ntiles;
  const c2<T> *b;
  c2<T> *x, *p;
  c2<T> *r[2], *s[2], *w[2];
  const c2<T> *K;  // [2][V]
  double *part;    // [2][gridDim][2]
  CgState *state;  // [2]
  CgResult *res;   // result block and sticky stop flag
  double stop_rr;
  int max_iter;
};

__device__ __forceinline__ int wrapc(int v, int n)
{
  v %= n;
  return v < 0 ? v + n : v;
}

template <typename T> struct TileCtx {
  using V2 = c2<T>;
  V2 *R2, *Y1, *LKt, *LKx;
  double *red;
  int W2, W1;  // row widths of E2 and E1
  int E2, E1;
};

// Load the folded links of a tile on its 2-halo region into LDS.
template <typename T>
__device__ void load_links(const CgArgs<T> &a, const TileCtx<T> &c, int t0, int x0)
{
  const Geo &G = a.G;
  const int V = 2 * G.vh;
  for (int e = threadIdx.x; e < c.E2; e += blockDim.x) {
    int et = e / c.W2, ex = e - et * c.W2;
    int i = eo_index(G, wrapc(t0 - 2 + et, G.nt), wrapc(x0 - 2 + ex, G.nx));
    c.LKt[e] = a.K[i];
    c.LKx[e] = a.K[V + i];
  }
}

// M on the 1-halo region from R2, then M^dag on the tile; returns the
// partial sums of (r.r, r.w) over the owned sites of the tile and writes
// w.  R2 must hold r_{j+1} on E2.
template <typename T>
__device__ void apply_and_reduce(const CgArgs<T> &a, const TileCtx<T> &c, int t0, int x0,
                                 c2<T> *wout, double &gacc, double &dacc)
{
  using V2 = c2<T>;
  const Geo &G = a.G;
  const T m = (T)G.m, ep = (T)G.ep, em = (T)G.em;
  // Y1 = M R2 on E1 (E1 local (ft,fx) == E2 local (ft+1,fx+1))
  for (int e = threadIdx.x; e < c.E1; e += blockDim.x) {
    int ft = e / c.W1, fx = e - ft * c.W1;
    int q = (ft + 1) * c.W2 + (fx + 1);
    V2 kt = c.LKt[q], kx = c.LKx[q], ktd = c.LKt[q - c.W2], kxd = c.LKx[q - 1];
    V2 v = cscale(c.R2[q], m);
    v = cadd(v, cscale(cmul(kt, c.R2[q + c.W2]), ep));
    v = csub(v, cscale(cjmul(ktd, c.R2[q - c.W2]), em));
    v = cadd(v, cmul(kx, c.R2[q + 1]));
    v = csub(v, cjmul(kxd, c.R2[q - 1]));
    c.Y1[e] = v;
  }
  __syncthreads();
  // w = M^dag Y1 on the owned tile sites
  for (int e = threadIdx.x; e < a.TT * a.TX; e += blockDim.x) {
    int lt = e / a.TX, lx = e - lt * a.TX;
    int tg = t0 + lt, xg = x0 + lx;
    if (tg >= G.nt || xg >= G.nx) continue;
    int q1 = (lt + 1) * c.W1 + (lx + 1);
    int q2 = (lt + 2) * c.W2 + (lx + 2);
    V2 kt = c.LKt[q2], kx = c.LKx[q2], ktd = c.LKt[q2 - c.W2], kxd = c.LKx[q2 - 1];
    V2 v = cscale(c.Y1[q1], m);
    v = csub(v, cscale(cmul(kt, c.Y1[q1 + c.W1]), em));
    v = cadd(v, cscale(cjmul(ktd, c.Y1[q1 - c.W1]), ep));
    v = csub(v, cmul(kx, c.Y1[q1 + 1]));
    v = cadd(v, cjmul(kxd, c.Y1[q1 - 1]));
    int i = eo_index(G, tg, xg);
    wout[i] = v;
    V2 r = c.R2[q2];
    gacc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
    dacc += (double)r.x * (double)v.x + (double)r.y * (double)v.y;
  }
}

template <typename T>
__device__ __forceinline__ TileCtx<T> tile_ctx(const CgArgs<T> &a)
{
  using V2 = c2<T>;
  extern __shared__ double lds_raw[];
  TileCtx<T> c;
  c.W2 = a.TX + 4;
  c.W1 = a.TX + 2;
  c.E2 = (a.TT + 4) * c.W2;
  c.E1 = (a.TT + 2) * c.W1;
  c.R2 = (V2 *)lds_raw;
  c.Y1 = c.R2 + c.E2;
  c.LKt = c.Y1 + c.E1;
  c.LKx = c.LKt + c.E2;
  c.red = (double *)(c.LKx + c.E2);
  return c;
}

template <typename T>
__device__ __forceinline__ void write_partials(const CgArgs<T> &a, const TileCtx<T> &c, int slot,
                                               double gacc, double dacc)
{
  double gs = block_sum(gacc, c.red);
  __syncthreads();
  double ds = block_sum(dacc, c.red + 16);
  if (threadIdx.x == 0) {
    a.part[(slot * gridDim.x + blockIdx.x) * 2 + 0] = gs;
    a.part[(slot * gridDim.x + blockIdx.x) * 2 + 1] = ds;
  }
}

// r_0 = b, w_0 = A b, s_{-1} = p_{-1} = x_0 = 0; partials of slot 0.
template <typename T>
__global__ void __launch_bounds__(256) k_cg_init(CgArgs<T> a)
{
  using V2 = c2<T>;
  const Geo &G = a.G;
  TileCtx<T> c = tile_ctx(a);
  double gacc = 0.0, dacc = 0.0;
  for (int tau = blockIdx.x; tau < a.ntiles; tau += gridDim.x) {
    int t0 = (tau / a.tiles_x) * a.TT, x0 = (tau % a.tiles_x) * a.TX;
    __syncthreads();
    load_links(a, c, t0, x0);
    for (int e = threadIdx.x; e < c.E2; e += blockDim.x) {
      int et = e / c.W2, ex = e - et * c.W2;
      int tg = t0 - 2 + et, xg = x0 - 2 + ex;
      int i = eo_index(G, wrapc(tg, G.nt), wrapc(xg, G.nx));
      V2 bv = a.b[i];
      c.R2[e] = bv;
      bool own = et >= 2 && et < a.TT + 2 && ex >= 2 && ex < a.TX + 2 && tg < G.nt && xg < G.nx;
      if (own) {
        V2 z = cmk<V2>((T)0, (T)0);
        a.r[0][i] = bv;
        a.s[0][i] = z;
        a.p[i] = z;
        a.x[i] = z;
      }
    }
    __syncthreads();
    apply_and_reduce(a, c, t0, x0, a.w[0], gacc, dacc);
  }
  write_partials(a, c, 0, gacc, dacc);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.state[0].gam_old = 1.0;
    a.state[0].alpha_old = 1.0;
    a.state[0].j = 0;
    a.res->done = 0;
    a.res->iters = 0;
    a.res->status = TH2D_OK;
    a.res->rr = 0.0;
  }
}

// Iteration j = state[slot].j; reads the slot's buffers, writes slot^1.
template <typename T>
__global__ void __launch_bounds__(256) k_cg_step(CgArgs<T> a, int slot)
{
  using V2 = c2<T>;
  const Geo &G = a.G;
  if (a.res->done) return;
  TileCtx<T> c = tile_ctx(a);
  const int nb = gridDim.x, nxt = slot ^ 1;
  // every block adds the partials of the previous launch in one order
  double pg = 0.0, pd = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    pg += a.part[(slot * nb + k) * 2 + 0];
    pd += a.part[(slot * nb + k) * 2 + 1];
  }
  double gs = block_sum(pg, c.red);
  __syncthreads();
  double ds = block_sum(pd, c.red + 16);
  if (threadIdx.x == 0) {
    c.red[32] = gs;
    c.red[33] = ds;
  }
  __syncthreads();
  const double gam = c.red[32], del = c.red[33];
  const CgState st = a.state[slot];
  const int j = st.j;
  bool stop = true;
  int status = TH2D_OK;
  double beta = 0.0, alpha = 0.0;
  if (gam < a.stop_rr) status = TH2D_OK;
  else if (j >= a.max_iter) status = TH2D_ENOCONV;
  else if (!(del > 0.0) || !isfinite(gam)) status = TH2D_EBREAKDOWN;
  else {
    beta = j == 0 ? 0.0 : gam / st.gam_old;
    alpha = j == 0 ? gam / del : gam / (del - beta * gam / st.alpha_old);
    if (!(alpha > 0.0) || !isfinite(alpha)) status = TH2D_EBREAKDOWN;
    else stop = false;
  }
  if (stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.res->iters = j;
      a.res->status = status;
      a.res->rr = gam;
      a.res->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.state[nxt].gam_old = gam;
    a.state[nxt].alpha_old = alpha;
    a.state[nxt].j = j + 1;
  }
  const T ta = (T)alpha, tb = (T)beta;
  double gacc = 0.0, dacc = 0.0;
  for (int tau = blockIdx.x; tau < a.ntiles; tau += gridDim.x) {
    int t0 = (tau / a.tiles_x) * a.TT, x0 = (tau % a.tiles_x) * a.TX;
    __syncthreads();
    load_links(a, c, t0, x0);
    for (int e = threadIdx.x; e < c.E2; e += blockDim.x) {
      int et = e / c.W2, ex = e - et * c.W2;
      int tg = t0 - 2 + et, xg = x0 - 2 + ex;
      int i = eo_index(G, wrapc(tg, G.nt), wrapc(xg, G.nx));
      V2 rv = a.r[slot][i], wv = a.w[slot][i], sv = a.s[slot][i];
      V2 sn = cadd(wv, cscale(sv, tb));
      V2 rn = csub(rv, cscale(sn, ta));
      c.R2[e] = rn;
      bool own = et >= 2 && et < a.TT + 2 && ex >= 2 && ex < a.TX + 2 && tg < G.nt && xg < G.nx;
      if (own) {
        V2 pn = cadd(rv, cscale(a.p[i], tb));
        a.p[i] = pn;
        a.x[i] = cadd(a.x[i], cscale(pn, ta));
        a.s[nxt][i] = sn;
        a.r[nxt][i] = rn;
      }
    }
    __syncthreads();
    apply_and_reduce(a, c, t0, x0, a.w[nxt], gacc, dacc);
  }
  write_partials(a, c, nxt, gacc, dacc);
}

struct TileChoice {
  int tt, tx;
};

// Tile shape: the first candidate whose tile count fits one wave of
// co-resident blocks, preferring counts close to the CU count.
TileChoice choose_tile(th2d_ctx *ctx, int nt, int nx)
{
  if (ctx->tile_t > 0 && ctx->tile_x > 0) return {ctx->tile_t, ctx->tile_x};
  // all candidates keep the FP64 LDS footprint under 64 KB
  static const TileChoice cand[] = {{16, 16}, {8, 16}, {16, 32}, {8, 8}, {4, 8}};
  TileChoice best = {16, 16};
  double best_score = 1e30;
  for (auto c : cand) {
    int n = ((nt + c.tt - 1) / c.tt) * ((nx + c.tx - 1) / c.tx);
    double halo = double((c.tt + 4) * (c.tx + 4)) / (c.tt * c.tx);
    // one block per CU is the target; fewer tiles idle CUs, more tiles
    // than co-resident blocks serialise
    double fill = n >= ctx->num_cu ? double(n) / ctx->num_cu : double(ctx->num_cu) / n;
    if (n > 2 * ctx->num_cu) fill *= 2.0;
    double score = fill * halo * (c.tt * c.tx <= 256 ? 1.0 : 1.25);
    if (nt * nx <= 1024) score = (n == 1) ? 0.0 : score;  // tiny lattice: one block
    if (score < best_score) {
      best_score = score;
      best = c;
    }
  }
  return best;
}

template <typename T> size_t lds_bytes(int tt, int tx)
{
  size_t e2 = (size_t)(tt + 4) * (tx + 4), e1 = (size_t)(tt + 2) * (tx + 2);
  return sizeof(c2<T>) * (3 * e2 + e1) + sizeof(double) * 48;
}

// Scratch layout of one solve of precision T: 7 vectors, partials,
// scalar state and the result block.
template <typename T> struct SolveBufs {
  c2<T> *x, *p, *r[2], *s[2], *w[2];
  double *part;
  CgState *state;
  CgResult *res;
};

template <typename T>
int solve_bufs(th2d_ctx *ctx, th2d_buf &buf, long V, int nblocks, SolveBufs<T> &o)
{
  (void)ctx;
  size_t vec = sizeof(c2<T>) * (size_t)V;
  vec = (vec + 255) & ~(size_t)255;
  size_t part = sizeof(double) * 4 * (size_t)nblocks;
  part = (part + 255) & ~(size_t)255;
  // x, p, r[2], s[2], w[2]: eight vectors, then partials, state, result
  const size_t nvec = 8;
  size_t bytes = nvec * vec + part + 256;
  TH2D_TRY(th2d_buf_reserve(buf, bytes));
  char *p = (char *)buf.ptr;
  o.x = (c2<T> *)p;
  p += vec;
  o.p = (c2<T> *)p;
  p += vec;
  for (int k = 0; k < 2; k++) {
    o.r[k] = (c2<T> *)p;
    p += vec;
    o.s[k] = (c2<T> *)p;
    p += vec;
    o.w[k] = (c2<T> *)p;
    p += vec;
  }
  o.part = (double *)p;
  p += part;
  o.state = (CgState *)p;          // 2 x 24 bytes
  o.res = (CgResult *)(p + 64);    // 24 bytes
  if ((size_t)((char *)(o.res + 1) - (char *)buf.ptr) > bytes) return TH2D_EARG;
  return TH2D_OK;
}

struct InnerResult {
  int iters;
  double rr;
  int status;
};

// One solve A y = rhs from y = 0 with recursive stop rr < stop; y is the
// internal x of the scratch block, returned in *xout.
template <typename T>
int inner_solve(th2d_ctx *ctx, const Geo &G, th2d_buf &buf, const c2<T> *rhs, const c2<T> *K,
                double stop_rr, int max_iter, c2<T> **xout, InnerResult *res)
{
  long V = 2L * G.vh;
  TileChoice tc = choose_tile(ctx, G.nt, G.nx);
  int tiles_x = (G.nx + tc.tx - 1) / tc.tx;
  int ntiles = ((G.nt + tc.tt - 1) / tc.tt) * tiles_x;
  size_t lds = lds_bytes<T>(tc.tt, tc.tx);
  // one tile per block up to a cap of blocks; more tiles loop in a block
  int nblocks = std::min(ntiles, ctx->cg_max_blocks > 0 ? ctx->cg_max_blocks : 4 * ctx->num_cu);
  SolveBufs<T> sb;
  TH2D_TRY(solve_bufs<T>(ctx, buf, V, nblocks, sb));
  CgArgs<T> a;
  a.G = G;
  a.TT = tc.tt;
  a.TX = tc.tx;
  a.tiles_x = tiles_x;
  a.ntiles = ntiles;
  a.b = rhs;
  a.x = sb.x;
  a.p = sb.p;
  for (int k = 0; k < 2; k++) {
    a.r[k] = sb.r[k];
    a.s[k] = sb.s[k];
    a.w[k] = sb.w[k];
  }
  a.K = K;
  a.part = sb.part;
  a.state = sb.state;
  a.res = sb.res;
  a.stop_rr = stop_rr;
  a.max_iter = max_iter;
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  hipLaunchKernelGGL(k_cg_init<T>, dim3(nblocks), dim3(256), lds, ctx->stream, a);
  TH2D_HIP(hipGetLastError());
  TH2D_TRY(th2d_host_reserve(ctx->redh, 64));
  CgResult h;
  memset(&h, 0, sizeof(h));
  long launched = 0;
  // first chunk: the iteration count of the previous solve of this
  // precision (solves along a trajectory are alike), then steps of 32
  int &last = ctx->cg_last_iters[sizeof(T) == 8 ? 1 : 0];
  int chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : std::max(16, last + 4);
  for (;;) {
    for (int k = 0; k < chunk; k++) {
      hipLaunchKernelGGL(k_cg_step<T>, dim3(nblocks), dim3(256), lds, ctx->stream, a,
                         (int)((launched + k) & 1));
    }
    TH2D_HIP(hipGetLastError());
    launched += chunk;
    TH2D_HIP(hipMemcpyAsync(ctx->redh.ptr, sb.res, sizeof(CgResult), hipMemcpyDeviceToHost,
                            ctx->stream));
    TH2D_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(&h, ctx->redh.ptr, sizeof(h));
    if (h.done) break;
    if (launched > (long)max_iter + 1) {  // cannot happen: the kernel stops at max_iter
      h.status = TH2D_ENOCONV;
      break;
    }
    chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : 32;
  }
  last = h.iters;
  if (ctx->ktiming) {
    TH2D_HIP(hipEventRecord(ctx->k1, ctx->stream));
    TH2D_HIP(hipEventSynchronize(ctx->k1));
    float ms = 0.f;
    TH2D_HIP(hipEventElapsedTime(&ms, ctx->k0, ctx->k1));
    ctx->ktime_ms += ms;
    ctx->klaunches += launched + 1;
    // algorithmic bytes of one CG iteration per site: two operator
    // applications (in + out + 2 links each: 8 complex) and the fused
    // vector updates (7 complex): 240 B in FP64, 120 B in FP32.
    double vb = sizeof(c2<T>);
    ctx->kbytes += (double)h.iters * V * 15.0 * vb;
    ctx->kiters += h.iters;
  }
  res->iters = h.iters;
  res->rr = h.rr;
  res->status = h.status;
  *xout = sb.x;
  return TH2D_OK;
}

__global__ void k_scale_f64_to_f32(float2 *d, const double2 *s, long n, double sc)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2((float)(s[i].x * sc), (float)(s[i].y * sc));
}

__global__ void k_axpy_f32_into_f64(double2 *y, const float2 *x, long n, double a)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double2 v = y[i];
    y[i] = make_double2(v.x + a * (double)x[i].x, v.y + a * (double)x[i].y);
  }
}

__global__ void k_axpy_f64(double2 *y, const double2 *x, long n, double a)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double2 v = y[i];
    y[i] = make_double2(v.x + a * x[i].x, v.y + a * x[i].y);
  }
}

}  // namespace

extern "C" int th2d_set_cg_tile(th2d_ctx *ctx, int tt, int tx)
{
  if (!ctx || tt < 0 || tx < 0 || (tx & 1)) return TH2D_EARG;
  if (tt * tx > 1024) return TH2D_EARG;
  ctx->tile_t = tt;
  ctx->tile_x = tx;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_max_blocks(th2d_ctx *ctx, int n)
{
  if (!ctx || n < 0) return TH2D_EARG;
  ctx->cg_max_blocks = n;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_chunk(th2d_ctx *ctx, int n)
{
  if (!ctx || n < 0) return TH2D_EARG;
  ctx->cg_chunk = n;
  return TH2D_OK;
}

// Outer loop: true residual in FP64, inner persistent solves in FP64 or
// FP32 on the normalised residual (defect correction).
extern "C" int th2d_cg_mdagm(th2d_ctx *ctx, const th2d_geom *g, double *x, const double *b,
                             const double *K, const float *K32, const th2d_cg_params *p,
                             th2d_cg_info *info)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!ctx || !x || !b || !K || !p || p->max_iter < 1) return TH2D_EARG;
  const bool mixed = p->precision == TH2D_PREC_MIXED;
  if (mixed && !K32) return TH2D_EARG;
  if (p->precision != TH2D_PREC_F64 && !mixed) return TH2D_EARG;
  Geo G = make_geo(g);
  const long V = 2L * G.vh;
  const size_t vb = sizeof(double) * 2 * V;
  // FP64 work vectors: r, tmp
  TH2D_TRY(th2d_buf_reserve(ctx->misc, 2 * vb + 256));
  double *r = (double *)ctx->misc.ptr;
  double *tmp = (double *)((char *)ctx->misc.ptr + vb);
  double *dsc = (double *)((char *)ctx->misc.ptr + 2 * vb);
  TH2D_TRY(th2d_buf_reserve(ctx->cgmix, sizeof(float) * 2 * V + 256));
  float2 *r32 = (float2 *)ctx->cgmix.ptr;
  const unsigned nblk = (unsigned)((V + 255) / 256);

  th2d_cg_info inf;
  memset(&inf, 0, sizeof(inf));
  double bb;
  TH2D_TRY(th2d::reduce_dot(ctx, 2 * V, b, b, dsc));
  TH2D_TRY(th2d::fetch_doubles(ctx, &bb, dsc, 1));
  TH2D_TRY(th2d_memset(ctx, x, 0, vb));
  if (bb == 0.0) {
    inf.status = TH2D_OK;
    if (info) *info = inf;
    return TH2D_OK;
  }
  // target on the true residual
  double target = 0.0;
  if (p->tol_rel > 0) target = p->tol_rel * p->tol_rel * bb;
  if (p->tol_abs_rr > 0) target = std::max(target, p->tol_abs_rr);
  if (!(target > 0)) return TH2D_EARG;
  const bool abs_mode = p->tol_rel <= 0;  // reference criterion on the recursion
  double delta = p->delta > 0 ? p->delta : 1e-4;

  double rr = bb;
  const double *rhs = b;  // current residual
  int left = p->max_iter;
  int status = TH2D_OK;
  for (int outer = 0; outer < 64; outer++) {
    if (rr <= target) break;
    double nr = sqrt(rr);
    InnerResult ir;
    if (mixed) {
      hipLaunchKernelGGL(k_scale_f64_to_f32, dim3(nblk), dim3(256), 0, ctx->stream, r32,
                         (const double2 *)rhs, V, 1.0 / nr);
      double stop = std::max(delta * delta, target / rr);
      float2 *e32;
      TH2D_TRY(inner_solve<float>(ctx, G, ctx->cg32, r32, (const float2 *)K32, stop, left, &e32,
                                  &ir));
      hipLaunchKernelGGL(k_axpy_f32_into_f64, dim3(nblk), dim3(256), 0, ctx->stream,
                         (double2 *)x, e32, V, nr);
    } else {
      // FP64 inner solve straight to the target (relative to this rhs)
      double stop = target / rr;
      if (!abs_mode) stop = std::max(stop, 1e-30);
      double2 *e64;
      const double2 *src = (const double2 *)rhs;
      if (outer > 0) {
        // normalise the residual so the recursion scale is 1
        TH2D_TRY(th2d::axpy_like(ctx, 2, 2 * V, 1.0 / nr, tmp, rhs, nullptr));
        src = (const double2 *)tmp;
      } else {
        stop = target;  // first pass: unscaled, exact reference criterion
      }
      TH2D_TRY(inner_solve<double>(ctx, G, ctx->cg64, src, (const double2 *)K, stop, left, &e64,
                                   &ir));
      hipLaunchKernelGGL(k_axpy_f64, dim3(nblk), dim3(256), 0, ctx->stream, (double2 *)x, e64, V,
                         outer > 0 ? nr : 1.0);
    }
    TH2D_HIP(hipGetLastError());
    inf.iterations += ir.iters;
    left -= ir.iters;
    inf.rr = ir.rr * (mixed || outer > 0 ? rr : 1.0);
    if (ir.status == TH2D_EBREAKDOWN) {
      status = TH2D_EBREAKDOWN;
    }
    // true residual r = b - A x in FP64
    TH2D_TRY(th2d::dslash_launch(ctx, G, tmp, x, K, 0, TH2D_PREC_F64, 1));
    TH2D_TRY(th2d::dslash_launch(ctx, G, r, tmp, K, 1, TH2D_PREC_F64, 1));
    TH2D_TRY(th2d::axpy_like(ctx, 1, 2 * V, 0.0, r, b, r));
    TH2D_TRY(th2d::reduce_dot(ctx, 2 * V, r, r, dsc));
    double rr_new;
    TH2D_TRY(th2d::fetch_doubles(ctx, &rr_new, dsc, 1));
    rhs = r;
    if (outer > 0) inf.restarts++;
    if (abs_mode && !mixed && ir.status == TH2D_OK) {  // recursion met the reference stop
      rr = rr_new;
      break;
    }
    if (status == TH2D_EBREAKDOWN && !(rr_new < rr)) {
      rr = rr_new;
      break;
    }
    status = TH2D_OK;
    if (!(rr_new < rr) && outer > 2) {  // stagnation at round-off
      rr = rr_new;
      break;
    }
    rr = rr_new;
    if (left <= 0) {
      status = TH2D_ENOCONV;
      break;
    }
  }
  inf.true_rel_resid = sqrt(rr / bb);
  if (status == TH2D_OK && rr > target && !abs_mode) status = TH2D_ENOCONV;
  inf.status = status;
  if (info) *info = inf;
  return status;
}
