// cg.hip -- conjugate-gradient solver for M^dag M on gfx950, for one
// lattice or a batch of independent lattices of any sizes
// (fmdm_invert_cg, hmc.c:341-404; same Krylov iterates, different
// schedule).
//
// Schedule: Chronopoulos-Gear CG, whose two inner products (r.r, r.Ar) are
// taken together, so one grid-wide synchronisation per iteration
// suffices; it is the boundary between two launches of one fused step
// kernel.  Every lattice ("chain") is cut into TT x TX tiles; one launch
// serves the tiles of all chains of the batch.  Iteration j (one launch)
// does, per tile:
//   1. loads the chain's tile partials of iteration j and, at the same
//      time, r_j, w_j = A r_j, s_{j-1} and the folded links on the tile plus
//      a 2-site halo;
//   2. adds the partials in a fixed order (every tile of a chain the same
//      way, so all take bitwise the same alpha/beta and stop decision);
//   3. forms s_j = w_j + beta s_{j-1}, r_{j+1} = r_j - alpha s_j on the
//      halo region (recomputed, not exchanged), p_j and x_{j+1} in place;
//   4. applies M on the tile plus a 1-site halo and M^dag on the tile:
//      w_{j+1} = A r_{j+1} (LDS);
//   5. writes its partial (r.r, r.w) for iteration j+1.
// r, s, w, the partials and the chain state are double-buffered by
// iteration parity; p, x are only touched by the tile that owns them.  The
// host enqueues launches in chunks and reads the per-chain stop flags once
// per chunk; tiles of a stopped chain return at once.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

constexpr int MAXE = 4;  // halo-region elements per thread (E2 <= 1024)

struct CgChain {
  Geo G;
  int tbeg, tend;   // tile range of the chain
  long voff;        // element offset of the chain in the work vectors
  const void *b;    // right-hand side (eo, precision T)
  const void *K;    // folded links ([2][V], precision T)
  double stop_rr;
  int max_iter, pad;
};

struct CgTile {
  int chain, t0, x0, pad;
};

struct CgState {
  double gam_old, alpha_old;
  int j, pad;
};

struct CgResult {
  int iters, status;
  double rr;
  int done, pad;
};

template <typename T> struct CgBatch {
  const CgChain *chains;
  const CgTile *tiles;
  int ntiles, nchains, TT, TX;
  c2<T> *x, *p, *r0, *r1, *s0, *s1, *w0, *w1;
  double *part;    // [2][ntiles][2]
  CgState *state;  // [2][nchains]
  CgResult *res;   // [nchains]
};

__device__ __forceinline__ int wrapn(int v, int n)
{
  while (v < 0) v += n;
  while (v >= n) v -= n;
  return v;
}

template <typename T> struct Lds {
  c2<T> *R2, *Y1, *LKt, *LKx;
  double *red;
  int W2, W1, E2, E1;
};

template <typename T> __device__ __forceinline__ Lds<T> lds_view(int TT, int TX)
{
  using V2 = c2<T>;
  extern __shared__ double lds_raw[];
  Lds<T> c;
  c.W2 = TX + 4;
  c.W1 = TX + 2;
  c.E2 = (TT + 4) * c.W2;
  c.E1 = (TT + 2) * c.W1;
  c.R2 = (V2 *)lds_raw;
  c.Y1 = c.R2 + c.E2;
  c.LKt = c.Y1 + c.E1;
  c.LKx = c.LKt + c.E2;
  c.red = (double *)(c.LKx + c.E2);
  return c;
}

// M on the 1-halo region from R2, then M^dag on the owned tile sites:
// w_out = A r; adds the tile's (r.r, r.w) into gacc/dacc.
template <typename T>
__device__ void apply_and_reduce(const Geo &G, const Lds<T> &c, int TT, int TX, int t0, int x0,
                                 c2<T> *wout, double &gacc, double &dacc)
{
  using V2 = c2<T>;
  const T m = (T)G.m, ep = (T)G.ep, em = (T)G.em;
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
  for (int e = threadIdx.x; e < TT * TX; e += blockDim.x) {
    int lt = e / TX, lx = e - lt * TX;
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
    wout[eo_index(G, tg, xg)] = v;
    V2 r = c.R2[q2];
    gacc += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
    dacc += (double)r.x * (double)v.x + (double)r.y * (double)v.y;
  }
}

template <typename T>
__device__ __forceinline__ void write_tile_partial(const CgBatch<T> &a, const Lds<T> &c, int slot,
                                                   int tau, double gacc, double dacc)
{
  double gs = block_sum(gacc, c.red);
  __syncthreads();
  double ds = block_sum(dacc, c.red + 16);
  if (threadIdx.x == 0) {
    a.part[((long)slot * a.ntiles + tau) * 2 + 0] = gs;
    a.part[((long)slot * a.ntiles + tau) * 2 + 1] = ds;
  }
  __syncthreads();
}

// r_0 = b, w_0 = A b, s_{-1} = p_{-1} = x_0 = 0; partials of slot 0.
template <typename T>
__global__ void __launch_bounds__(256) k_cg_init(CgBatch<T> a)
{
  using V2 = c2<T>;
  Lds<T> c = lds_view<T>(a.TT, a.TX);
  for (int tau = blockIdx.x; tau < a.ntiles; tau += gridDim.x) {
    const CgTile tl = a.tiles[tau];
    const CgChain ch = a.chains[tl.chain];
    const Geo &G = ch.G;
    const int V = 2 * G.vh;
    const V2 *K = (const V2 *)ch.K;
    const V2 *b = (const V2 *)ch.b;
    V2 *r0 = a.r0 + ch.voff, *s0 = a.s0 + ch.voff, *p = a.p + ch.voff, *x = a.x + ch.voff;
    __syncthreads();
    for (int e = threadIdx.x; e < c.E2; e += blockDim.x) {
      int et = e / c.W2, ex = e - et * c.W2;
      int tg = tl.t0 - 2 + et, xg = tl.x0 - 2 + ex;
      int i = eo_index(G, wrapn(tg, G.nt), wrapn(xg, G.nx));
      c.LKt[e] = K[i];
      c.LKx[e] = K[V + i];
      V2 bv = b[i];
      c.R2[e] = bv;
      bool own = et >= 2 && et < a.TT + 2 && ex >= 2 && ex < a.TX + 2 && tg < G.nt && xg < G.nx;
      if (own) {
        V2 z = cmk<V2>((T)0, (T)0);
        r0[i] = bv;
        s0[i] = z;
        p[i] = z;
        x[i] = z;
      }
    }
    __syncthreads();
    double gacc = 0.0, dacc = 0.0;
    apply_and_reduce(G, c, a.TT, a.TX, tl.t0, tl.x0, a.w0 + ch.voff, gacc, dacc);
    write_tile_partial(a, c, 0, tau, gacc, dacc);
    if (tau == ch.tbeg && threadIdx.x == 0) {
      a.state[tl.chain] = CgState{1.0, 1.0, 0, 0};
      a.res[tl.chain] = CgResult{0, TH2D_OK, 0.0, 0, 0};
    }
  }
}

// One iteration of every chain that has not stopped; reads the buffers of
// `slot`, writes those of slot^1.
template <typename T>
__global__ void __launch_bounds__(256) k_cg_step(CgBatch<T> a, int slot)
{
  using V2 = c2<T>;
  Lds<T> c = lds_view<T>(a.TT, a.TX);
  const int nxt = slot ^ 1;
  for (int tau = blockIdx.x; tau < a.ntiles; tau += gridDim.x) {
    const CgTile tl = a.tiles[tau];
    // the stop flag of a chain is set during a launch by its first tile:
    // read it once per block so that every thread takes the same branch
    __syncthreads();
    if (threadIdx.x == 0) c.red[40] = (double)a.res[tl.chain].done;
    __syncthreads();
    if (c.red[40] != 0.0) continue;
    const CgChain ch = a.chains[tl.chain];
    const Geo &G = ch.G;
    const int V = 2 * G.vh;
    const V2 *K = (const V2 *)ch.K;
    const V2 *rc = (slot ? a.r1 : a.r0) + ch.voff, *wc = (slot ? a.w1 : a.w0) + ch.voff;
    const V2 *sc = (slot ? a.s1 : a.s0) + ch.voff;
    V2 *rn_ = (slot ? a.r0 : a.r1) + ch.voff, *sn_ = (slot ? a.s0 : a.s1) + ch.voff;
    V2 *wn_ = (slot ? a.w0 : a.w1) + ch.voff;
    V2 *p = a.p + ch.voff, *x = a.x + ch.voff;
    // 1. issue the loads: chain partials, halo-region vectors and links
    double pg = 0.0, pd = 0.0;
    for (int k = ch.tbeg + threadIdx.x; k < ch.tend; k += blockDim.x) {
      pg += a.part[((long)slot * a.ntiles + k) * 2 + 0];
      pd += a.part[((long)slot * a.ntiles + k) * 2 + 1];
    }
    V2 er[MAXE], ew[MAXE], es[MAXE], ep_[MAXE], ex_[MAXE];
    int ei[MAXE];
    bool eown[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      int e = threadIdx.x + q * blockDim.x;
      ei[q] = -1;
      eown[q] = false;
      if (e < c.E2) {
        int et = e / c.W2, exx = e - et * c.W2;
        int tg = tl.t0 - 2 + et, xg = tl.x0 - 2 + exx;
        int i = eo_index(G, wrapn(tg, G.nt), wrapn(xg, G.nx));
        ei[q] = i;
        er[q] = rc[i];
        ew[q] = wc[i];
        es[q] = sc[i];
        c.LKt[e] = K[i];
        c.LKx[e] = K[V + i];
        eown[q] = et >= 2 && et < a.TT + 2 && exx >= 2 && exx < a.TX + 2 && tg < G.nt &&
                  xg < G.nx;
        if (eown[q]) {
          ep_[q] = p[i];
          ex_[q] = x[i];
        }
      }
    }
    // 2. scalars of iteration j, identical in every tile of the chain
    double gs = block_sum(pg, c.red);
    __syncthreads();
    double ds = block_sum(pd, c.red + 16);
    if (threadIdx.x == 0) {
      c.red[32] = gs;
      c.red[33] = ds;
    }
    __syncthreads();
    const double gam = c.red[32], del = c.red[33];
    const CgState st = a.state[(long)slot * a.nchains + tl.chain];
    const int j = st.j;
    bool stop = true;
    int status = TH2D_OK;
    double beta = 0.0, alpha = 0.0;
    if (gam < ch.stop_rr) status = TH2D_OK;
    else if (j >= ch.max_iter) status = TH2D_ENOCONV;
    else if (!(del > 0.0) || !isfinite(gam)) status = TH2D_EBREAKDOWN;
    else {
      beta = j == 0 ? 0.0 : gam / st.gam_old;
      alpha = j == 0 ? gam / del : gam / (del - beta * gam / st.alpha_old);
      if (!(alpha > 0.0) || !isfinite(alpha)) status = TH2D_EBREAKDOWN;
      else stop = false;
    }
    if (stop) {
      if (tau == ch.tbeg && threadIdx.x == 0) a.res[tl.chain] = CgResult{j, status, gam, 1, 0};
      __syncthreads();
      continue;
    }
    if (tau == ch.tbeg && threadIdx.x == 0)
      a.state[(long)nxt * a.nchains + tl.chain] = CgState{gam, alpha, j + 1, 0};
    // 3. s_j, r_{j+1} on the halo region; p_j, x_{j+1} on owned sites
    const T ta = (T)alpha, tb = (T)beta;
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      int e = threadIdx.x + q * blockDim.x;
      if (ei[q] < 0) continue;
      V2 sn = cadd(ew[q], cscale(es[q], tb));
      V2 rn = csub(er[q], cscale(sn, ta));
      c.R2[e] = rn;
      if (eown[q]) {
        int i = ei[q];
        V2 pn = cadd(er[q], cscale(ep_[q], tb));
        p[i] = pn;
        x[i] = cadd(ex_[q], cscale(pn, ta));
        sn_[i] = sn;
        rn_[i] = rn;
      }
    }
    __syncthreads();
    // 4./5. w_{j+1} = A r_{j+1}, partials of iteration j+1
    double gacc = 0.0, dacc = 0.0;
    apply_and_reduce(G, c, a.TT, a.TX, tl.t0, tl.x0, wn_, gacc, dacc);
    write_tile_partial(a, c, nxt, tau, gacc, dacc);
  }
}

struct TileChoice {
  int tt, tx;
};

// Tile shape for a batch whose largest lattice is nt x nx holding
// `sites` sites in all: near one tile per CU for a single lattice,
// 16 x 16 when the batch fills the device anyway.
TileChoice choose_tile(th2d_ctx *ctx, int nt, int nx, long sites)
{
  if (ctx->tile_t > 0 && ctx->tile_x > 0) return {ctx->tile_t, ctx->tile_x};
  static const TileChoice cand[] = {{16, 16}, {8, 16}, {16, 32}, {8, 8}, {4, 8}};
  if (sites >= 2L * 256 * ctx->num_cu) return {16, 16};
  TileChoice best = {16, 16};
  double best_score = 1e30;
  for (auto c : cand) {
    double n = double(sites) / (c.tt * c.tx);
    double halo = double((c.tt + 4) * (c.tx + 4)) / (c.tt * c.tx);
    double fill = n >= ctx->num_cu ? n / ctx->num_cu : double(ctx->num_cu) / n;
    if (n > 2 * ctx->num_cu) fill *= 2.0;
    double score = fill * halo * (c.tt * c.tx <= 256 ? 1.0 : 1.25);
    if (nt * nx <= 1024 && sites <= 1024) score = (c.tt >= nt && c.tx >= nx) ? 0.0 : score;
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

struct InnerResult {
  int iters;
  double rr;
  int status;
};

// Batched inner solve: A_c y_c = rhs_c from y_c = 0, stop r.r < stop_c.
// The solutions are the internal x vectors, returned in xout[c].
template <typename T>
int inner_solve(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G, const c2<T> *const *rhs,
                const c2<T> *const *K, const double *stop_rr, const int *max_iter, c2<T> **xout,
                InnerResult *res)
{
  int max_nt = 0, max_nx = 0;
  long sites = 0;
  for (int c = 0; c < nch; c++) {
    max_nt = std::max(max_nt, G[c].nt);
    max_nx = std::max(max_nx, G[c].nx);
    sites += 2L * G[c].vh;
  }
  TileChoice tc = choose_tile(ctx, max_nt, max_nx, sites);
  if ((tc.tt + 4) * (tc.tx + 4) > MAXE * 256) return TH2D_EARG;
  // host tables
  std::vector<CgChain> ch(nch);
  std::vector<CgTile> tiles;
  long voff = 0;
  for (int c = 0; c < nch; c++) {
    ch[c].G = G[c];
    ch[c].tbeg = (int)tiles.size();
    for (int t0 = 0; t0 < G[c].nt; t0 += tc.tt)
      for (int x0 = 0; x0 < G[c].nx; x0 += tc.tx) tiles.push_back(CgTile{c, t0, x0, 0});
    ch[c].tend = (int)tiles.size();
    ch[c].voff = voff;
    voff += (2L * G[c].vh + 63) & ~63L;
    ch[c].b = rhs[c];
    ch[c].K = K[c];
    ch[c].stop_rr = stop_rr[c];
    ch[c].max_iter = max_iter[c];
    ch[c].pad = 0;
  }
  const int ntiles = (int)tiles.size();
  // device block: 8 vectors, partials, state, results, chain and tile tables
  size_t vec = sizeof(c2<T>) * (size_t)voff;
  vec = (vec + 255) & ~(size_t)255;
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  size_t part = al(sizeof(double) * 4 * (size_t)ntiles);
  size_t stb = al(sizeof(CgState) * 2 * (size_t)nch);
  size_t rsb = al(sizeof(CgResult) * (size_t)nch);
  size_t chb = al(sizeof(CgChain) * (size_t)nch);
  size_t tlb = al(sizeof(CgTile) * (size_t)ntiles);
  size_t bytes = 8 * vec + part + stb + rsb + chb + tlb;
  TH2D_TRY(th2d_buf_reserve(buf, bytes));
  char *pb = (char *)buf.ptr;
  CgBatch<T> a;
  c2<T> **vp[8] = {&a.x, &a.p, &a.r0, &a.r1, &a.s0, &a.s1, &a.w0, &a.w1};
  for (int k = 0; k < 8; k++) {
    *vp[k] = (c2<T> *)pb;
    pb += vec;
  }
  a.part = (double *)pb;
  pb += part;
  a.state = (CgState *)pb;
  pb += stb;
  a.res = (CgResult *)pb;
  pb += rsb;
  CgChain *dch = (CgChain *)pb;
  pb += chb;
  CgTile *dtl = (CgTile *)pb;
  pb += tlb;
  if ((size_t)(pb - (char *)buf.ptr) > bytes) return TH2D_EARG;
  a.chains = dch;
  a.tiles = dtl;
  a.ntiles = ntiles;
  a.nchains = nch;
  a.TT = tc.tt;
  a.TX = tc.tx;
  // tables through a pinned staging block
  size_t hbytes = std::max(chb + tlb, rsb);
  TH2D_TRY(th2d_host_reserve(ctx->redh, hbytes + 64));
  memcpy(ctx->redh.ptr, ch.data(), sizeof(CgChain) * nch);
  memcpy((char *)ctx->redh.ptr + chb, tiles.data(), sizeof(CgTile) * ntiles);
  TH2D_HIP(hipMemcpyAsync(dch, ctx->redh.ptr, chb + tlb, hipMemcpyHostToDevice, ctx->stream));
  const size_t lds = lds_bytes<T>(tc.tt, tc.tx);
  const int cap = ctx->cg_max_blocks > 0 ? ctx->cg_max_blocks : 8 * ctx->num_cu;
  const int nblocks = std::max(1, std::min(ntiles, cap));
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  hipLaunchKernelGGL(k_cg_init<T>, dim3(nblocks), dim3(256), lds, ctx->stream, a);
  TH2D_HIP(hipGetLastError());
  std::vector<CgResult> h(nch);
  long launched = 0;
  int &last = ctx->cg_last_iters[sizeof(T) == 8 ? 1 : 0];
  int chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : std::max(16, last + 4);
  int most = 0;
  for (int c = 0; c < nch; c++) most = std::max(most, max_iter[c]);
  for (;;) {
    for (int k = 0; k < chunk; k++)
      hipLaunchKernelGGL(k_cg_step<T>, dim3(nblocks), dim3(256), lds, ctx->stream, a,
                         (int)((launched + k) & 1));
    TH2D_HIP(hipGetLastError());
    launched += chunk;
    TH2D_HIP(hipMemcpyAsync(ctx->redh.ptr, a.res, sizeof(CgResult) * nch, hipMemcpyDeviceToHost,
                            ctx->stream));
    TH2D_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(h.data(), ctx->redh.ptr, sizeof(CgResult) * nch);
    bool all = true;
    for (int c = 0; c < nch; c++) all &= h[c].done != 0;
    if (all) break;
    if (launched > (long)most + 1) {  // the kernel stops every chain at its max_iter
      for (int c = 0; c < nch; c++)
        if (!h[c].done) h[c].status = TH2D_ENOCONV;
      break;
    }
    chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : 32;
  }
  int itmax = 0;
  for (int c = 0; c < nch; c++) itmax = std::max(itmax, h[c].iters);
  last = itmax;
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
    for (int c = 0; c < nch; c++) {
      ctx->kbytes += (double)h[c].iters * 2.0 * G[c].vh * 15.0 * vb;
      ctx->kiters += h[c].iters;
    }
  }
  for (int c = 0; c < nch; c++) {
    res[c].iters = h[c].iters;
    res[c].rr = h[c].rr;
    res[c].status = h[c].status;
    xout[c] = a.x + ch[c].voff;
  }
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
  if ((tt || tx) && (tt + 4) * (tx + 4) > MAXE * 256) return TH2D_EARG;
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

// Outer loop per chain: true residual in FP64, batched inner solves in
// FP64 or FP32 on the normalised residual (defect correction); chains that
// have met their target leave the batch.
extern "C" int th2d_cg_mdagm_batched(th2d_ctx *ctx, int nch, const th2d_geom *geoms, double **x,
                                     const double *const *b, const double *const *K,
                                     const float *const *K32, const th2d_cg_params *p,
                                     th2d_cg_info *infos)
{
  if (!ctx || nch < 1 || !geoms || !x || !b || !K || !p || p->max_iter < 1) return TH2D_EARG;
  const bool mixed = p->precision == TH2D_PREC_MIXED;
  if (p->precision != TH2D_PREC_F64 && !mixed) return TH2D_EARG;
  std::vector<Geo> G(nch);
  std::vector<long> V(nch), voff(nch + 1, 0);
  for (int c = 0; c < nch; c++) {
    TH2D_TRY(th2d_check_geom(&geoms[c]));
    if (!x[c] || !b[c] || !K[c] || (mixed && (!K32 || !K32[c]))) return TH2D_EARG;
    G[c] = make_geo(&geoms[c]);
    V[c] = 2L * G[c].vh;
    voff[c + 1] = voff[c] + ((V[c] + 63) & ~63L);
  }
  const long Vt = voff[nch];
  // FP64 work: residuals r (all chains), tmp, per-chain scalars
  TH2D_TRY(th2d_buf_reserve(ctx->misc, sizeof(double) * 2 * 2 * Vt + sizeof(double) * (nch + 64)));
  double *R = (double *)ctx->misc.ptr;
  double *TMP = R + 2 * Vt;
  double *dsc = TMP + 2 * Vt;
  TH2D_TRY(th2d_buf_reserve(ctx->cgmix, sizeof(float) * 2 * Vt + 256));
  float2 *R32 = (float2 *)ctx->cgmix.ptr;

  std::vector<th2d_cg_info> inf(nch);
  std::vector<double> bb(nch), rr(nch), target(nch);
  std::vector<int> left(nch, p->max_iter), active(nch, 1), status(nch, TH2D_OK);
  const bool abs_mode = p->tol_rel <= 0;
  const double delta = p->delta > 0 ? p->delta : 1e-4;
  for (int c = 0; c < nch; c++) {
    memset(&inf[c], 0, sizeof(th2d_cg_info));
    TH2D_TRY(th2d::reduce_dot(ctx, 2 * V[c], b[c], b[c], dsc + c));
    TH2D_TRY(th2d_memset(ctx, x[c], 0, sizeof(double) * 2 * V[c]));
  }
  TH2D_TRY(th2d::fetch_doubles(ctx, bb.data(), dsc, nch));
  for (int c = 0; c < nch; c++) {
    double t = 0.0;
    if (p->tol_rel > 0) t = p->tol_rel * p->tol_rel * bb[c];
    if (p->tol_abs_rr > 0) t = std::max(t, p->tol_abs_rr);
    if (!(t > 0)) return TH2D_EARG;
    target[c] = t;
    rr[c] = bb[c];
    if (bb[c] == 0.0 || rr[c] <= target[c]) active[c] = 0;
  }
  for (int outer = 0; outer < 64; outer++) {
    std::vector<int> idx;
    for (int c = 0; c < nch; c++)
      if (active[c]) idx.push_back(c);
    if (idx.empty()) break;
    const int na = (int)idx.size();
    std::vector<Geo> g2(na);
    std::vector<double> stop(na);
    std::vector<int> mi(na);
    std::vector<InnerResult> ir(na);
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      g2[k] = G[c];
      mi[k] = left[c];
      const double *rhs = outer == 0 ? b[c] : R + 2 * voff[c];
      double nr = sqrt(rr[c]);
      if (mixed) {
        hipLaunchKernelGGL(k_scale_f64_to_f32, dim3((unsigned)((V[c] + 255) / 256)), dim3(256), 0,
                           ctx->stream, R32 + voff[c], (const double2 *)rhs, V[c], 1.0 / nr);
        stop[k] = std::max(delta * delta, target[c] / rr[c]);
      } else if (outer == 0) {
        stop[k] = target[c];  // unscaled first pass: the reference criterion
      } else {
        TH2D_TRY(th2d::axpy_like(ctx, 2, 2 * V[c], 1.0 / nr, TMP + 2 * voff[c], rhs, nullptr));
        stop[k] = abs_mode ? target[c] / rr[c] : std::max(target[c] / rr[c], 1e-30);
      }
    }
    if (mixed) {
      std::vector<const float2 *> rh(na), kk(na);
      std::vector<float2 *> xo(na);
      for (int k = 0; k < na; k++) {
        rh[k] = R32 + voff[idx[k]];
        kk[k] = (const float2 *)K32[idx[k]];
      }
      TH2D_TRY(inner_solve<float>(ctx, ctx->cg32, na, g2.data(), rh.data(), kk.data(), stop.data(),
                                  mi.data(), xo.data(), ir.data()));
      for (int k = 0; k < na; k++) {
        int c = idx[k];
        hipLaunchKernelGGL(k_axpy_f32_into_f64, dim3((unsigned)((V[c] + 255) / 256)), dim3(256), 0,
                           ctx->stream, (double2 *)x[c], xo[k], V[c], sqrt(rr[c]));
      }
    } else {
      std::vector<const double2 *> rh(na), kk(na);
      std::vector<double2 *> xo(na);
      for (int k = 0; k < na; k++) {
        int c = idx[k];
        rh[k] = (const double2 *)(outer == 0 ? b[c] : TMP + 2 * voff[c]);
        kk[k] = (const double2 *)K[c];
      }
      TH2D_TRY(inner_solve<double>(ctx, ctx->cg64, na, g2.data(), rh.data(), kk.data(),
                                   stop.data(), mi.data(), xo.data(), ir.data()));
      for (int k = 0; k < na; k++) {
        int c = idx[k];
        hipLaunchKernelGGL(k_axpy_f64, dim3((unsigned)((V[c] + 255) / 256)), dim3(256), 0,
                           ctx->stream, (double2 *)x[c], xo[k], V[c],
                           outer > 0 ? sqrt(rr[c]) : 1.0);
      }
    }
    TH2D_HIP(hipGetLastError());
    // true residuals r = b - A x in FP64
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      double *r = R + 2 * voff[c], *tmp = TMP + 2 * voff[c];
      TH2D_TRY(th2d::dslash_launch(ctx, G[c], tmp, x[c], K[c], 0, TH2D_PREC_F64, 1));
      TH2D_TRY(th2d::dslash_launch(ctx, G[c], r, tmp, K[c], 1, TH2D_PREC_F64, 1));
      TH2D_TRY(th2d::axpy_like(ctx, 1, 2 * V[c], 0.0, r, b[c], r));
      TH2D_TRY(th2d::reduce_dot(ctx, 2 * V[c], r, r, dsc + k));
    }
    std::vector<double> rn(na);
    TH2D_TRY(th2d::fetch_doubles(ctx, rn.data(), dsc, na));
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      inf[c].iterations += ir[k].iters;
      left[c] -= ir[k].iters;
      inf[c].rr = ir[k].rr * (mixed || outer > 0 ? rr[c] : 1.0);
      if (outer > 0) inf[c].restarts++;
      double rnew = rn[k];
      bool breakdown = ir[k].status == TH2D_EBREAKDOWN;
      if (abs_mode && !mixed && ir[k].status == TH2D_OK) {  // reference stop met
        rr[c] = rnew;
        active[c] = 0;
        continue;
      }
      if (breakdown && !(rnew < rr[c])) {
        status[c] = TH2D_EBREAKDOWN;
        rr[c] = rnew;
        active[c] = 0;
        continue;
      }
      if (!(rnew < rr[c]) && outer > 2) {  // stagnation at round-off
        rr[c] = rnew;
        active[c] = 0;
        continue;
      }
      rr[c] = rnew;
      if (rr[c] <= target[c]) active[c] = 0;
      else if (left[c] <= 0) {
        status[c] = TH2D_ENOCONV;
        active[c] = 0;
      }
    }
  }
  int worst = TH2D_OK;
  for (int c = 0; c < nch; c++) {
    inf[c].true_rel_resid = bb[c] > 0 ? sqrt(rr[c] / bb[c]) : 0.0;
    if (status[c] == TH2D_OK && rr[c] > target[c] && !abs_mode && bb[c] > 0)
      status[c] = TH2D_ENOCONV;
    inf[c].status = status[c];
    if (status[c] != TH2D_OK && worst == TH2D_OK) worst = status[c];
    if (infos) infos[c] = inf[c];
  }
  return worst;
}

extern "C" int th2d_cg_mdagm(th2d_ctx *ctx, const th2d_geom *g, double *x, const double *b,
                             const double *K, const float *K32, const th2d_cg_params *p,
                             th2d_cg_info *info)
{
  double *xs[1] = {x};
  const double *bs[1] = {b};
  const double *ks[1] = {K};
  const float *k32[1] = {K32};
  return th2d_cg_mdagm_batched(ctx, 1, g, xs, bs, ks, K32 ? k32 : nullptr, p, info);
}
