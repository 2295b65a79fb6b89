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
#include "res_layout.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

namespace {

constexpr int CG_MAX_THREADS = 512;

// Tile shapes of the step kernel (TT x TX sites, NT threads), compiled
// one by one so that every index division is by a constant and every
// element loop unrolls.
struct ShapeEntry {
  int tt, tx, nt;
};
constexpr ShapeEntry kShapes[] = {{16, 64, 512}, {8, 64, 256}, {16, 32, 256}, {8, 32, 256},
                                  {16, 16, 256}, {8, 16, 256}, {8, 8, 128},   {4, 8, 64},
                                  {4, 4, 64},    {2, 4, 64},   {2, 2, 64}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

inline int shape_index(int tt, int tx)
{
  for (int k = 0; k < kNumShapes; k++)
    if (kShapes[k].tt == tt && kShapes[k].tx == tx) return k;
  return -1;
}

template <int TT_, int TX_, int NT_> struct Shape {
  static constexpr int TT = TT_, TX = TX_, NT = NT_;
  static constexpr int W2 = TX + 4, E2 = (TT + 4) * W2;  // 2-site halo region
  static constexpr int W1 = TX + 2, E1 = (TT + 2) * W1;  // 1-site halo region
  static constexpr int NE = (E2 + NT - 1) / NT, N1 = (E1 + NT - 1) / NT;
  static constexpr int NO = (TT * TX + NT - 1) / NT;
};

struct CgChain {
  Geo G;
  int tbeg, tend;   // tile range of the chain
  long voff;        // element offset of the chain in the work vectors
  const void *b;    // right-hand side (eo, precision T)
  const void *K;    // folded links ([2][V], precision T)
  double stop_rr;
  int max_iter;
  int pad;          // resident solver: first partial slot of the chain (8-aligned)
  // reliable update (resident solver): p, s of the previous pass in and
  // out (chain vectors, FP32; null: none), and with cscale > 0 the solve
  // continues that pass's Krylov sequence with p, s scaled by cscale and
  // gam_old = gam0 cscale^2, alpha_old = alpha0
  void *cp, *cs;
  double cscale, gam0, alpha0;
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
  double gam_old, alpha_old;  // resident solver: the last step's gam, alpha
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

// Wrap of a coordinate at most one period outside [0, n) (tiles no larger
// than the lattice).
__device__ __forceinline__ int wrap1(int v, int n)
{
  v = v < 0 ? v + n : v;
  return v >= n ? v - n : v;
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

// Stores the tile's (r.r, r.w) as partial `tau` of `slot` (fixed order:
// lanes, then waves).  c.red[0..15] must be free: its last reader is
// thread 0 of the previous tile, before that tile's closing barrier.  The
// closing barrier (LDS reuse by the next tile) is left out on the block's
// last tile.
template <typename T>
__device__ __forceinline__ void write_tile_partial(const CgBatch<T> &a, double *red, int slot,
                                                   int tau, double gacc, double dacc, bool last)
{
  gacc = wave_sum(gacc);
  dacc = wave_sum(dacc);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wid] = gacc;
    red[2 * wid + 1] = dacc;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double g = 0.0, d = 0.0;
    for (int k = 0; k < (int)((blockDim.x + 63) >> 6); k++) {
      g += red[2 * k];
      d += red[2 * k + 1];
    }
    a.part[((long)slot * a.ntiles + tau) * 2 + 0] = g;
    a.part[((long)slot * a.ntiles + tau) * 2 + 1] = d;
  }
  if (!last) __syncthreads();
}

// Tile served by virtual block v.  Workgroups are dispatched round-robin
// over the 8 XCDs (v mod 8 for a grid that is a multiple of 8); each XCD
// gets a contiguous range of tiles, so the halo of a tile is mostly read
// by blocks of the same XCD and shared through its L2.  A bijection of
// [0, ntiles) for any ntiles.
__device__ __forceinline__ int tile_of(int v, int ntiles)
{
  const int q = ntiles >> 3, rem = ntiles & 7, x = v & 7, k = v >> 3;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + k;
}

// r_0 = b, w_0 = A b, s_{-1} = p_{-1} = x_0 = 0; partials of slot 0.
template <typename T>
__global__ void __launch_bounds__(CG_MAX_THREADS) k_cg_init(CgBatch<T> a)
{
  using V2 = c2<T>;
  Lds<T> c = lds_view<T>(a.TT, a.TX);
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x) {
    const int tau = tile_of(v, a.ntiles);
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
    write_tile_partial(a, c.red, 0, tau, gacc, dacc, v + (int)gridDim.x >= a.ntiles);
    if (tau == ch.tbeg && threadIdx.x == 0) {
      a.state[tl.chain] = CgState{1.0, 1.0, 0, 0};
      a.res[tl.chain] = CgResult{0, TH2D_OK, 0.0, 0, 0, 0.0, 0.0};
    }
  }
}

// One iteration of every chain that has not stopped; reads the buffers of
// `slot`, writes those of slot^1.  Every lattice of the batch is at least
// S::TT x S::TX.
template <typename T, typename S>
__global__ void __launch_bounds__(S::NT) k_cg_step(CgBatch<T> a, int slot)
{
  using V2 = c2<T>;
  constexpr int TT = S::TT, TX = S::TX, NT = S::NT, W2 = S::W2, W1 = S::W1;
  constexpr int E2 = S::E2, E1 = S::E1, NE = S::NE;
  constexpr int LDS_D = (3 * E2 + E1) * (int)sizeof(V2) / 8;
  __shared__ double lds[LDS_D + 40];
  V2 *R2 = (V2 *)lds, *Y1 = R2 + E2, *LKt = Y1 + E1, *LKx = LKt + E2;
  double *red = lds + LDS_D;
  const int tid = threadIdx.x;
  const int nxt = slot ^ 1;
  for (int v = blockIdx.x; v < a.ntiles; v += gridDim.x) {
    const bool last = v + (int)gridDim.x >= a.ntiles;
    const int tau = tile_of(v, a.ntiles);
    const CgTile tl = a.tiles[tau];
    const CgChain ch = a.chains[tl.chain];
    const Geo &G = ch.G;
    const int V = 2 * G.vh;
    const V2 *K = (const V2 *)ch.K;
    const V2 *rc = (slot ? a.r1 : a.r0) + ch.voff, *wc = (slot ? a.w1 : a.w0) + ch.voff;
    const V2 *sc = (slot ? a.s1 : a.s0) + ch.voff;
    V2 *rn_ = (slot ? a.r0 : a.r1) + ch.voff, *sn_ = (slot ? a.s0 : a.s1) + ch.voff;
    V2 *wn_ = (slot ? a.w0 : a.w1) + ch.voff;
    V2 *p = a.p + ch.voff, *x = a.x + ch.voff;
    // 1. issue the loads: chain state, halo-region vectors and links, then
    //    (wave 0) the chain's tile partials and its stop flag
    const CgState st = a.state[(long)slot * a.nchains + tl.chain];
    V2 er[NE], ew[NE], es[NE], ep_[NE], ex_[NE];
    int ei[NE];
    bool eown[NE];
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + q * NT;
      ei[q] = -1;
      eown[q] = false;
      if (E2 % NT == 0 || e < E2) {
        const int et = e / W2, exx = e - et * W2;
        const int tg = tl.t0 - 2 + et, xg = tl.x0 - 2 + exx;
        const int i = eo_index(G, wrap1(tg, G.nt), wrap1(xg, G.nx));
        ei[q] = i;
        er[q] = rc[i];
        ew[q] = wc[i];
        es[q] = sc[i];
        LKt[e] = K[i];
        LKx[e] = K[V + i];
        eown[q] = et >= 2 && et < TT + 2 && exx >= 2 && exx < TX + 2 && tg < G.nt && xg < G.nx;
        if (eown[q]) {
          ep_[q] = p[i];
          ex_[q] = x[i];
        }
      }
    }
    // 2. scalars of iteration j, identical in every tile of the chain.  The
    //    stop flag is set during a launch by the chain's first tile, so
    //    lanes may see different values: lane 0's is broadcast.
    if (tid < 64) {
      const double *pp = a.part + (long)slot * a.ntiles * 2;
      double pg = 0.0, pd = 0.0;
      for (int k = ch.tbeg + tid; k < ch.tend; k += 64) {
        pg += pp[2 * k];
        pd += pp[2 * k + 1];
      }
      pg = wave_sum(pg);
      pd = wave_sum(pd);
      if (tid == 0) {
        red[32] = pg;
        red[33] = pd;
        red[34] = (double)a.res[tl.chain].done;
      }
    }
    __syncthreads();
    const double gam = red[32], del = red[33];
    if (red[34] != 0.0) {
      __syncthreads();  // red[32..34] is rewritten by the next tile
      continue;
    }
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
      if (tau == ch.tbeg && tid == 0) a.res[tl.chain] = CgResult{j, status, gam, 1, 0, 0.0, 0.0};
      __syncthreads();
      continue;
    }
    if (tau == ch.tbeg && tid == 0)
      a.state[(long)nxt * a.nchains + tl.chain] = CgState{gam, alpha, j + 1, 0};
    // 3. s_j, r_{j+1} on the halo region; p_j, x_{j+1} on owned sites
    const T ta = (T)alpha, tb = (T)beta;
#pragma unroll
    for (int q = 0; q < NE; q++) {
      const int e = tid + q * NT;
      if (!(E2 % NT == 0 || e < E2)) continue;
      V2 sn = cadd(ew[q], cscale(es[q], tb));
      V2 rn = csub(er[q], cscale(sn, ta));
      R2[e] = rn;
      if (eown[q]) {
        const int i = ei[q];
        V2 pn = cadd(er[q], cscale(ep_[q], tb));
        p[i] = pn;
        x[i] = cadd(ex_[q], cscale(pn, ta));
        sn_[i] = sn;
        rn_[i] = rn;
      }
    }
    __syncthreads();
    // 4. y = M r_{j+1} on the 1-site halo region
    const T m = (T)G.m, epm = (T)G.ep, emm = (T)G.em;
#pragma unroll
    for (int q = 0; q < S::N1; q++) {
      const int e = tid + q * NT;
      if (!(E1 % NT == 0 || e < E1)) continue;
      const int ft = e / W1, fx = e - ft * W1;
      const int c = (ft + 1) * W2 + (fx + 1);
      V2 kt = LKt[c], kx = LKx[c], ktd = LKt[c - W2], kxd = LKx[c - 1];
      V2 y = cscale(R2[c], m);
      y = cadd(y, cscale(cmul(kt, R2[c + W2]), epm));
      y = csub(y, cscale(cjmul(ktd, R2[c - W2]), emm));
      y = cadd(y, cmul(kx, R2[c + 1]));
      y = csub(y, cjmul(kxd, R2[c - 1]));
      Y1[e] = y;
    }
    __syncthreads();
    // 5. w_{j+1} = M^dag y on the owned sites, partials of iteration j+1
    T gacc = 0, dacc = 0;
#pragma unroll
    for (int q = 0; q < S::NO; q++) {
      const int e = tid + q * NT;
      if (!((TT * TX) % NT == 0 || e < TT * TX)) continue;
      const int lt = e / TX, lx = e - lt * TX;
      const int tg = tl.t0 + lt, xg = tl.x0 + lx;
      if (tg >= G.nt || xg >= G.nx) continue;
      const int c1 = (lt + 1) * W1 + (lx + 1);
      const int c2i = (lt + 2) * W2 + (lx + 2);
      V2 kt = LKt[c2i], kx = LKx[c2i], ktd = LKt[c2i - W2], kxd = LKx[c2i - 1];
      V2 w = cscale(Y1[c1], m);
      w = csub(w, cscale(cmul(kt, Y1[c1 + W1]), emm));
      w = cadd(w, cscale(cjmul(ktd, Y1[c1 - W1]), epm));
      w = csub(w, cmul(kx, Y1[c1 + 1]));
      w = cadd(w, cjmul(kxd, Y1[c1 - 1]));
      wn_[eo_index(G, tg, xg)] = w;
      V2 r = R2[c2i];
      gacc += r.x * r.x + r.y * r.y;
      dacc += r.x * w.x + r.y * w.y;
    }
    write_tile_partial(a, red, nxt, tau, (double)gacc, (double)dacc, last);
  }
}

// ------------------------------------------------------ resident solver --
// One workgroup per 64 x 64 region of a lattice for a whole inner solve,
// every workgroup resident at once (cooperative launch).  The region's
// r, w, s, p, x stay in registers and its folded links, with the 2-site
// halo, in LDS, so an iteration reads and writes only the 2-site boundary
// layers of r, s, w and one partial per region (the step kernel moves
// 96 B per site and iteration in FP32).  Iteration j of a region:
//   1. arrives on its chain's counter (iteration j-1 done) and waits until
//      every region of the chain has;
//   2. reads the chain's partials, and the halo layers of r_j, w_j, s_{j-1}
//      that the neighbouring regions published;
//   3. as the step kernel: s_j, r_{j+1} on region and halo, p_j, x_{j+1},
//      y = M r_{j+1} on the 1-halo, w_{j+1} = M^dag y on the region;
//   4. publishes its boundary layers of r_{j+1}, s_j, w_{j+1} and its
//      partial of iteration j+1.
// The regions of a chain are placed on one XCD (place_regions), so the
// exchange stays in that XCD's L2 (see ResMem); a chain larger than an XCD
// is solved by the step kernel.  A site is published when it lies within 2
// sites of its region's edge; every halo value that reaches an owned
// site's result is such a site (a halo site two steps from a region lies
// within two steps of the edge of the region it belongs to).  Sites are
// published into a compact block per region (res_layout.h), not at their
// places in the lattice vectors: neighbours then read whole rows and column
// pairs from consecutive addresses, and the exchange of the 32 regions an
// XCD holds (24 KB each) stays in its L2.  Through the lattice vectors a
// column layer costs a 128-B line per site and the exchange missed L2.  All regions
// of a chain take bitwise the same stop decision and leave together; every
// wait is bounded (a wait that times out sets the error flag and ends its
// workgroup).
// Regions of 64 x 64 sites, and of 32 x 64 for batches small enough that
// halving the regions still leaves one region per CU (one chain of 256^2:
// 32 regions instead of 16).
using RShape = Shape<64, 64, 512>;
using RShapeH = Shape<32, 64, 512>;
static_assert(RShape::TT == RES_T && RShape::TX == RES_T && RShapeH::TX == RES_T,
              "regions of the publication layout");
constexpr unsigned RES_SPIN_LIMIT = 1u << 20;            // polls of one wait (at most about 1 s)
constexpr int RES_PART_LANE = 4;  // partial slots a lane reads: at most 256 a chain (host check)
constexpr unsigned kNoLoad = 0xFFFFFFFFu;                 // halo element no result needs

struct ResCtl {
  const int *wg_tile;  // [grid] region of each workgroup, -1: idle (see place_regions)
  unsigned *bar;     // [nchains][32] arrival counters of the iterations (own cache line each)
  unsigned *bar0;    // [nchains][32] set-up counters (agent scope)
  int *err;          // set by a wait that timed out
  int *misplaced;    // set when a chain's regions do not share an XCD
  int *xcc;          // [ntiles] XCD of the workgroup of each region
  double *part;      // [2][nparts][waves][4] partials; a chain's start on a 2-KB boundary
  void *pub;         // [2][2][ntiles][RES_PUB] published boundary layers per parity
                     // (res_layout.h): (r, s) pairs, then w (16 B per slot each)
  int nparts;        // partial slots per parity
  unsigned blk_bytes;  // bytes of the solver block from CgBatch::x (buffer range)
  unsigned spin;     // polls of one wait before it gives up (RES_SPIN_LIMIT)
};

__device__ __forceinline__ void pub_store(float2 *p, float2 v)
{
  __hip_atomic_store((unsigned long long *)p, __builtin_bit_cast(unsigned long long, v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float2 pub_load(const float2 *p)
{
  return __builtin_bit_cast(float2, __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void pub_store(double *p, double v)
{
  __hip_atomic_store((unsigned long long *)p, __builtin_bit_cast(unsigned long long, v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double pub_load(const double *p)
{
  return __builtin_bit_cast(double, __hip_atomic_load((unsigned long long *)p, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT));
}

// Element at byte offset `off` of a base pointer (a 32-bit offset per lane).
template <typename V> __device__ __forceinline__ V *at(V *base, unsigned off)
{
  return (V *)((char *)base + off);
}
template <typename V> __device__ __forceinline__ const V *at(const V *base, unsigned off)
{
  return (const V *)((const char *)base + off);
}

// Starts the invalidation of this CU's vector L1, so that later loads read
// L2; it has completed when the wave's vector memory counter reaches 0.
__device__ __forceinline__ void l1_invalidate_issue()
{
  __asm__ volatile("buffer_inv sc0" ::: "memory");
}

// Memory of the resident solver: the solver block at `base`, addressed by
// a uniform byte offset plus a per-lane byte offset (buffer accesses).  All
// regions of a chain run on one XCD and exchange through its L2: plain
// stores (the vector L1 writes through to L2), arrival counters updated by
// L2 atomics and polled with agent-scope loads (cache policy sc1: they
// miss in the vector L1 and read L2; a plain or workgroup-scope load of the
// counter may be served by a stale L1 line), then an L1 invalidation before
// the reads (loads with cache policy sc0).
typedef unsigned b64v __attribute__((ext_vector_type(2)));
typedef unsigned b128v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
struct ResMem {
  char *base;
  __amdgpu_buffer_rsrc_t rs;

  __device__ __forceinline__ float2 ld2(unsigned soff, unsigned voff) const
  {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 1));
  }
  __device__ __forceinline__ void st2(unsigned soff, unsigned voff, float2 v) const
  {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64v, v), rs, voff, soff, 0);
  }
  // 16 bytes: two FP32 complex values, or two doubles
  __device__ __forceinline__ f4v ld4(unsigned soff, unsigned voff) const
  {
    return __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 1));
  }
  __device__ __forceinline__ d2v ldd2(unsigned soff, unsigned voff) const
  {
    return __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 1));
  }
  __device__ __forceinline__ void st4(unsigned soff, unsigned voff, f4v v) const
  {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(b128v, v), rs, voff, soff, 0);
  }
  __device__ __forceinline__ void std2(unsigned soff, unsigned voff, d2v v) const
  {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(b128v, v), rs, voff, soff, 0);
  }
  __device__ __forceinline__ double ldd(unsigned soff, unsigned voff) const
  {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 1));
  }
  __device__ __forceinline__ void std_(unsigned soff, unsigned voff, double v) const
  {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64v, v), rs, voff, soff, 0);
  }
  // Arrive on the counter at byte offset `obar` and wait until it reaches
  // `target`; false when the wait timed out here or elsewhere (then the
  // error flag is set).  Every wave first waits for its own stores.
  __device__ bool sync(unsigned obar, unsigned target, int *err, double *flag,
                       unsigned spin) const
  {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned *bar = (unsigned *)(base + obar);
      __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      // The L1 invalidation runs under the wait: one L1 per CU, this
      // workgroup is alone on it, and until the other waves pass the
      // barrier below nothing here loads through it (the polls and the
      // error flag are agent-scope loads, which miss in it).
      l1_invalidate_issue();
      bool ok = true;
      // back-to-back polls, each one L2 round trip (a sleep between them,
      // or a second poll in flight, measured slower)
      for (unsigned n = 0;; n++) {
        if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
        if (n >= spin ||
            ((n & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = false;
          break;
        }
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the invalidation has completed
      *flag = ok ? 1.0 : 0.0;
    }
    __syncthreads();
    return *flag != 0.0;
  }
};

// Agent-scope arrival counter for the set-up, before the regions of a chain
// are known to share an XCD: the caller's stores have completed before it
// arrives, and after the wait an acquire fence (L1 and L2 invalidation)
// orders the reads that follow.
__device__ bool agent_sync(unsigned *bar, unsigned target, int *err, double *flag, unsigned spin)
{
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores have completed
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = true;
    for (unsigned n = 0; __hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target;
         n++) {
      if (n >= spin ||
          ((n & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *flag = ok ? 1.0 : 0.0;
  }
  __syncthreads();
  return *flag != 0.0;
}

// Position (row, column) in the (TT+4) x (TX+4) halo frame of halo element
// q: two rows above, two below, then two columns left and right.
template <typename S>
__device__ __forceinline__ void res_ring_pos(int q, int &et, int &ex)
{
  constexpr int W2 = S::W2, TT = S::TT, TX = S::TX;
  if (q < 2 * W2) {
    et = q / W2;
    ex = q - et * W2;
  } else if (q < 4 * W2) {
    q -= 2 * W2;
    et = TT + 2 + q / W2;
    ex = q - (q / W2) * W2;
  } else {
    q -= 4 * W2;
    const int side = q / (2 * TT), k = q - side * 2 * TT;
    et = 2 + (k >> 1);
    ex = (k & 1) + (side ? TX + 2 : 0);
  }
}

// a b + c by the three-address FMA.  Written as a b + c, the compiler
// picks the two-address form, whose destination must hold c first: where a
// and c are loop-carried values and the result replaces a, that costs a
// register copy of both every iteration.
__device__ __forceinline__ float fma3(float a, float b, float c)
{
  float d;
  __asm__("v_fma_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}

// y += s a b and y += s conj(a) b for FP32 complex values, as chains of
// fused multiply-adds into y: written as y + (a b - c d) the compiler may
// not reassociate, and spends a separate add per term.  s = +-1 folds into
// the operand's sign.
template <typename V>
__device__ __forceinline__ void cfma(V &y, V a, V b, float s)
{
  const float ax = s * a.x, ay = s * a.y;
  y.x = __builtin_fmaf(ax, b.x, y.x);
  y.x = __builtin_fmaf(-ay, b.y, y.x);
  y.y = __builtin_fmaf(ax, b.y, y.y);
  y.y = __builtin_fmaf(ay, b.x, y.y);
}
template <typename V>
__device__ __forceinline__ void cjfma(V &y, V a, V b, float s)
{
  const float ax = s * a.x, ay = s * a.y;
  y.x = __builtin_fmaf(ax, b.x, y.x);
  y.x = __builtin_fmaf(ay, b.y, y.x);
  y.y = __builtin_fmaf(ax, b.y, y.y);
  y.y = __builtin_fmaf(-ay, b.x, y.y);
}

// Wave64 total of a double by DPP moves within the wave (no LDS round
// trips): pairs, quads (quad permutations), half-rows and rows (mirrors),
// then rows 0+1 into row 1 and 2+3 into row 3 (broadcast of lane 15), and
// rows 0..3 into row 3 (broadcast of lane 31); lane 63 then holds the total
// in a fixed order, and readlane gives it to every lane.  Lanes a
// broadcast does not write keep an undefined value (no register has to be
// cleared first): no such lane feeds lane 63.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_total(double v)
{
  v += dpp_d<0xB1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141, 0xf>(v);  // row_half_mirror
  v += dpp_d<0x140, 0xf>(v);  // row_mirror
  v += dpp_d<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_d<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | lo);
}

// The same for a float, each step one add with a DPP operand (the result
// is the float total in a fixed order, widened).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
__device__ __forceinline__ double wave_total_f(float v)
{
  v += dpp_f<0xB1, 0xf>(v);
  v += dpp_f<0x4E, 0xf>(v);
  v += dpp_f<0x141, 0xf>(v);
  v += dpp_f<0x140, 0xf>(v);
  v += dpp_f<0x142, 0xa>(v);
  v += dpp_f<0x143, 0xc>(v);
  return (double)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// a / b from the hardware reciprocal refined by two Newton steps and a
// corrected quotient (a few fused multiply-adds instead of the scaled
// division sequence).  Every region computes it bitwise alike; b = 0 gives
// NaN, which the breakdown checks catch as they catch an infinite quotient.
__device__ __forceinline__ double qdiv(double a, double b)
{
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  return __builtin_fma(r, __builtin_fma(-b, q, a), q);
}

// XCD of the calling workgroup (hardware register XCC_ID, bits 3:0).
__device__ __forceinline__ int xcc_id()
{
  return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u);
}

// MU0: every chain of the batch has mu = 0 (the e^{+-mu} factors are 1);
// S: the region shape (RShape or RShapeH); CLIP: some region of the batch
// is larger than what is left of its lattice (a lattice of one region
// smaller than the shape).  Without clipped regions every element is owned
// and no load of another region's values precedes the halo's: the waits of
// 3a are then for the halo loads alone.
template <typename T, bool MU0, typename S, bool CLIP>
__global__ void __launch_bounds__(S::NT) k_cg_resident(CgBatch<T> a, ResCtl ctl)
{
  using V2 = c2<T>;
  static_assert(sizeof(V2) == 8, "the resident solver is FP32");
  constexpr int TT = S::TT, TX = S::TX, NT = S::NT, W2 = S::W2, W1 = S::W1;
  constexpr int E2 = S::E2, E1 = S::E1, NO = S::NO;
  constexpr int LDS_D = (3 * E2 + E1) * (int)sizeof(V2) / 8;
  constexpr int NR = 4 * W2 + 4 * TT;   // halo ring elements of a region
  constexpr int NQ = (NR + NT - 1) / NT;
  // more than half of a CU's LDS, so that one workgroup runs per CU (the
  // 32-row regions would otherwise share CUs and gain nothing)
  constexpr int LDS_W = LDS_D + 40 > 82 * 1024 / 8 ? LDS_D + 40 : 82 * 1024 / 8;
  __shared__ double rlds[LDS_W];
  V2 *R2 = (V2 *)rlds, *Y1 = R2 + E2, *LKt = Y1 + E1, *LKx = LKt + E2;
  double *red = rlds + LDS_D;
  const int tid = threadIdx.x;
  // a wave holds NO consecutive rows of the region: element q of a thread
  // lies at row wrow + q (uniform in the wave, kept in a scalar register so
  // that the row tests branch on the scalar unit) and column lane
  static_assert(TX == 64, "a wave covers whole rows of a region");
  static_assert(NO * (NT / 64) == TT, "the waves' rows tile the region");
  const int wrow = __builtin_amdgcn_readfirstlane(tid >> 6) * NO, lane = tid & 63;
  const int tau = ctl.wg_tile[blockIdx.x];  // uniform: idle workgroups leave at once
  if (tau < 0) return;
  const CgTile tl = a.tiles[tau];
  const CgChain ch = a.chains[tl.chain];
  const Geo G = ch.G;
  const int V = 2 * G.vh, ntc = ch.tend - ch.tbeg;
  const bool first = tau == ch.tbeg;
  const int tte = min(TT, G.nt - tl.t0), txe = min(TX, G.nx - tl.x0);
  const V2 *K = (const V2 *)ch.K, *b = (const V2 *)ch.b;
  const T m = (T)G.m, epm = MU0 ? (T)1 : (T)G.ep, emm = MU0 ? (T)1 : (T)G.em;
  ResMem M;
  M.base = (char *)a.x;
  M.rs = __builtin_amdgcn_make_buffer_rsrc(a.x, 0, (int)ctl.blk_bytes, 0x00020000);
  auto boff = [&](const void *p) { return (unsigned)((const char *)p - M.base); };
  // publication planes per parity, (r, s) pairs and w, 16 B per slot (one
  // load or store for r and s), and this region's block
  constexpr unsigned PS = 16, PB = RES_PUB * PS;
  const unsigned plane = (unsigned)a.ntiles * PB, o_pub = boff(ctl.pub);
  const unsigned o_rs[2] = {o_pub, o_pub + 2 * plane};
  const unsigned o_w[2] = {o_pub + plane, o_pub + 3 * plane};
  const unsigned my_blk = (unsigned)tau * PB;
  // publication offset of owned element (lt, lx) within 2 sites of the edge
  auto pub_at = [&](int lt, int lx) {
    return my_blk + (unsigned)res_pub_slot(lt, lx, tte, txe) * PS;
  };
  // publication offset of the site at region coordinates (lt, lx) outside
  // the owned block, or kNoLoad when no owned result depends on it
  auto src_at = [&](int lt, int lx) {
    unsigned off = kNoLoad;
    if (lt < tte + 2 && lx < txe + 2) {
      int k;
      const int sl =
          res_site_slot(wrapn(tl.t0 + lt, G.nt), wrapn(tl.x0 + lx, G.nx), G.nt, G.nx, k, TT);
      if (sl >= 0) off = (unsigned)(ch.tbeg + k) * PB + (unsigned)sl * PS;
    }
    return off;
  };
  // partial slots of 4 doubles: (r.r, r.w) each iteration, and in the
  // set-up of a continued pass also (r.s, p.s) of the carried p, s
  constexpr int NW = NT / 64;  // waves per workgroup: partial slots per region
  const unsigned o_part[2] = {boff(ctl.part) + (unsigned)(32 * NW * ch.pad),
                              boff(ctl.part) + (unsigned)(32 * NW * (ctl.nparts + ch.pad))};
  const unsigned o_bar = boff(ctl.bar + 32 * tl.chain);
  const unsigned my_part = (unsigned)(32 * NW * (tau - ch.tbeg));

  // links and r_0 = b on the region and its halo
  for (int e = tid; e < E2; e += NT) {
    const int et = e / W2, ex = e - et * W2;
    const int i = eo_index(G, wrapn(tl.t0 - 2 + et, G.nt), wrapn(tl.x0 - 2 + ex, G.nx));
    LKt[e] = K[i];
    LKx[e] = K[V + i];
    R2[e] = b[i];
  }
  // owned elements e = tid + q NT of the region; those outside the lattice
  // (a region larger than what is left of it) are halo, reloaded each time
  // (byte offset of element q in a chain vector)
  auto oi_of = [&](int q) {
    const int lt = wrow + q, lx = lane;
    return (unsigned)eo_index(G, wrapn(tl.t0 + lt, G.nt), wrapn(tl.x0 + lx, G.nx)) *
           (unsigned)sizeof(V2);
  };
  // A buffer access at an offset at or past the end of the block (nol) is
  // out of the descriptor's range: a load returns 0 and a store is dropped,
  // without a memory access.  The publication stores and the halo and
  // partial loads are unconditional (nol where an element has none), so that
  // a wait for a load counts exactly the accesses issued after it and does
  // not wait for later stores.
  const unsigned nol = ctl.blk_bytes;
  unsigned inl = 0;
  unsigned pso[NO];  // publication offset of owned element q (nol: none)
  V2 r[NO], w[NO], s[NO], p[NO], x[NO];
  const V2 zero = cmk<V2>((T)0, (T)0);
#pragma unroll
  for (int q = 0; q < NO; q++) {
    const int lt = wrow + q, lx = lane;
    const unsigned oi = oi_of(q);
    pso[q] = nol;
    if (lt < tte && lx < txe) {
      inl |= 1u << q;
      if (lt < 2 || lt >= tte - 2 || lx < 2 || lx >= txe - 2) pso[q] = pub_at(lt, lx);
    }
    r[q] = *at(b, oi);
    w[q] = s[q] = p[q] = x[q] = zero;
    if (ch.cscale > 0.0 && lt < tte && lx < txe) {  // continue the previous pass
      const T cf = (T)ch.cscale;
      p[q] = cscale(*at((const V2 *)ch.cp, oi), cf);
      s[q] = cscale(*at((const V2 *)ch.cs, oi), cf);
    }
  }
  unsigned ri[NQ];  // publication offsets of the halo elements (nol: none)
  int re[NQ];       // their places in R2
#pragma unroll
  for (int h = 0; h < NQ; h++) {
    ri[h] = nol;
    re[h] = 0;
    if (tid + h * NT < NR) {
      int et, ex;
      res_ring_pos<S>(tid + h * NT, et, ex);
      const unsigned o = src_at(et - 2, ex - 2);
      ri[h] = o == kNoLoad ? nol : o;
      re[h] = et * W2 + ex;
    }
  }
  // the regions of the chain were placed on one XCD (workgroups are
  // dispatched round-robin over the XCDs); check it with the XCD register
  // (agent scope), and leave without solving if not (the host then solves
  // the batch with the step kernel)
  if (tid == 0)
    __hip_atomic_store(ctl.xcc + tau, xcc_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // A solve that does not run or ends on a wait that timed out stores x = 0
  // (with the error flag set): a zero correction for a caller that has
  // queued its update of the solution before reading the flags.
  bool failed = !agent_sync(ctl.bar0 + 32 * tl.chain, (unsigned)ntc, ctl.err, red + 36, ctl.spin);
  if (!failed) {
    if (tid == 0) {
      const int mine = __hip_atomic_load(ctl.xcc + tau, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bool same = true;
      for (int k = ch.tbeg; k < ch.tend; k++)
        same &= __hip_atomic_load(ctl.xcc + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == mine;
      red[37] = same ? 1.0 : 0.0;
      if (!same) __hip_atomic_store(ctl.misplaced, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    failed = red[37] == 0.0;  // the same decision in every region of the chain
  }

  // The links of the wave's rows (K_t, K_x and the x neighbour's K_x) stay
  // in registers for both stencils, loaded once (the barriers of the set-up
  // wait above have ordered the LDS staging).  The mu != 0 instance with
  // clipped regions of 64 rows keeps reading them from LDS: in registers it
  // would spill.
  constexpr bool LREG = MU0 || !CLIP || NO <= 4;
  constexpr int NL = LREG ? NO : 1;
  V2 lkt[NL], lkx[NL], lkxd[NL], lktd0;
  const int cl0 = (wrow + 2) * W2 + (lane + 2);  // link index of (wrow, lane)
  lktd0 = LKt[cl0 - W2];
  if constexpr (LREG) {
#pragma unroll
    for (int q = 0; q < NO; q++) {
      lkt[q] = LKt[cl0 + q * W2];
      lkx[q] = LKx[cl0 + q * W2];
      lkxd[q] = LKx[cl0 + q * W2 - 1];
    }
  }
  auto link_t = [&](int q) { return LREG ? lkt[LREG ? q : 0] : LKt[cl0 + q * W2]; };
  auto link_x = [&](int q) { return LREG ? lkx[LREG ? q : 0] : LKx[cl0 + q * W2]; };
  auto link_xd = [&](int q) { return LREG ? lkxd[LREG ? q : 0] : LKx[cl0 + q * W2 - 1]; };
  const bool cont = ch.cscale > 0.0;
  double gam_old = cont ? ch.gam0 * ch.cscale * ch.cscale : 1.0;
  double alpha_old = cont ? ch.alpha0 : 1.0;
  int slot = 0;
  // iteration j; j = -1 is the set-up, w_0 = A b
  T ta = 0, tb = 0;
  V2 hr[NQ], hw[NQ], hs[NQ];  // halo values of the next iteration
  V2 yv[NO];                  // y = M r on the wave's rows (3b, used in 3c)
  // One iteration; false when the chain leaves.  The set-up is an instance
  // of its own, so that in the loop every vector flows through one path (no
  // register copies where a set-up path and an iteration path would join).
  auto iteration = [&](auto setup_c, int j) __attribute__((always_inline)) -> bool {
    constexpr bool setup = decltype(setup_c)::value;
    const int ns = setup ? 0 : slot ^ 1;  // parity written by this pass
    if constexpr (!setup) {
      // 3a. s_j, r_{j+1} on the region and the halo; p_j, x_{j+1}.  Elements
      // outside the lattice part of the region (a region larger than what is
      // left of the lattice) first take the owners' values, in a pass of its
      // own: the wait for those loads then comes before this iteration's
      // stores (a wait for the counter of loads also waits for the stores
      // issued before it, one store round trip per element).  The region
      // before the halo ring: it does not wait for the halo loads.
#pragma unroll
      for (int q = 0; q < NO; q++) {
        if (!CLIP || ((inl >> q) & 1u)) continue;
        const int lt = wrow + q, lx = lane;
        const unsigned so = src_at(lt, lx);
        r[q] = w[q] = s[q] = zero;
        if (so != kNoLoad) {
          const f4v v = M.ld4(o_rs[slot], so);
          r[q] = cmk<V2>(v.x, v.y);
          s[q] = cmk<V2>(v.z, v.w);
          w[q] = M.ld2(o_w[slot], so);
        }
      }
#pragma unroll
      for (int q = 0; q < NO; q++) {
        const int lt = wrow + q, lx = lane;
        const bool own = !CLIP || ((inl >> q) & 1u);
        // s, p and r by three-address FMAs (fma3), each result in the
        // register of the value it replaces
        const V2 sn = cmk<V2>(fma3(s[q].x, tb, w[q].x), fma3(s[q].y, tb, w[q].y));
        if (own) {
          p[q] = cmk<V2>(fma3(p[q].x, tb, r[q].x), fma3(p[q].y, tb, r[q].y));
          x[q] = cadd(x[q], cscale(p[q], ta));
        }
        const V2 rn = cmk<V2>(fma3(sn.x, -ta, r[q].x), fma3(sn.y, -ta, r[q].y));
        R2[(lt + 2) * W2 + (lx + 2)] = rn;
        r[q] = rn;
        s[q] = sn;
        M.st4(o_rs[ns], pso[q], f4v{rn.x, rn.y, sn.x, sn.y});  // nol: dropped
      }
#pragma unroll
      for (int h = 0; h < NQ; h++) {
        if (tid + h * NT >= NR) continue;
        const V2 sn = cadd(hw[h], cscale(hs[h], tb));
        R2[re[h]] = csub(hr[h], cscale(sn, ta));
      }
    } else {
#pragma unroll
      for (int q = 0; q < NO; q++) {
        M.st4(o_rs[0], pso[q], f4v{r[q].x, r[q].y, s[q].x, s[q].y});  // nol: dropped
      }
    }
    __syncthreads();
    // 3b. y = M r_{j+1} on the 1-halo.  A wave takes the NO rows wrow + 1 ..
    // wrow + NO of the 1-halo frame at column lane + 1, its own sites: r(t),
    // r(t +- 1) come from r[] and the links from registers, so an element
    // loads its two x neighbours from LDS (the rows above and below the
    // wave's once).  Then the frame's first and
    // last rows and columns without the corners, at most one element per
    // thread.
    {
      const int c = (wrow + 2) * W2 + (lane + 2);  // R2 index of 1-halo (wrow + 1, lane + 1)
      // r on the wave's rows is r[] (3a wrote the same values to R2); the
      // rows above and below come from LDS
      const V2 rtop = R2[c - W2], rbot = R2[c + NO * W2];
      V2 ktd = lktd0;
      V2 *const yo = Y1 + (wrow + 1) * W1 + (lane + 1);
      // unrolled, one element ahead: the loads of element q + 1 are issued
      // before the arithmetic of element q, which waits for its own loads only
      V2 rxp = R2[c + 1], rxm = R2[c - 1];
#pragma unroll
      for (int q = 0; q < NO; q++) {
        const int cn = c + (q + 1) * W2;
        V2 nrxp, nrxm;
        if (q + 1 < NO) {
          nrxp = R2[cn + 1];
          nrxm = R2[cn - 1];
        }
        const V2 kt = link_t(q), kx = link_x(q), kxd = link_xd(q);
        __builtin_amdgcn_sched_barrier(0);
        const V2 r0 = r[q], rtp = q + 1 < NO ? r[q + 1] : rbot, rtm = q > 0 ? r[q - 1] : rtop;
        // y = m r + e^mu K_t r(t+1) - e^-mu K_t(t-1)* r(t-1) + K_x r(x+1) - K_x(x-1)* r(x-1)
        V2 y = cscale(r0, m);
        cfma(y, kt, rtp, epm);
        cjfma(y, ktd, rtm, -emm);
        cfma(y, kx, rxp, 1.0f);
        cjfma(y, kxd, rxm, -1.0f);
        yo[q * W1] = y;
        yv[q] = y;
        ktd = kt;
        if (q + 1 < NO) {
          rxp = nrxp;
          rxm = nrxm;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      // (3c reads no corner of the frame: 2 TX + 2 TT elements, one per
      // thread of the first waves, one wave per SIMD at 64 x 64 with 512
      // threads)
      for (int e = tid; e < 2 * TX + 2 * TT; e += NT) {
        int ft, fx;
        if (e < 2 * TX) {  // rows 0 and TT + 1, columns 1 .. TX
          ft = e < TX ? 0 : TT + 1;
          fx = 1 + (e < TX ? e : e - TX);
        } else {  // columns 0 and W1 - 1 of rows 1 .. TT
          const int k = e - 2 * TX;
          ft = 1 + (k >> 1);
          fx = (k & 1) ? W1 - 1 : 0;
        }
        const int cc = (ft + 1) * W2 + (fx + 1);
        const V2 kt = LKt[cc], kx = LKx[cc], ktd2 = LKt[cc - W2], kxd = LKx[cc - 1];
        const V2 r00 = R2[cc], rtp = R2[cc + W2], rtm2 = R2[cc - W2], rxp = R2[cc + 1],
                 rxm = R2[cc - 1];
        __builtin_amdgcn_sched_barrier(0);
        V2 y = cscale(r00, m);
        cfma(y, kt, rtp, epm);
        cjfma(y, ktd2, rtm2, -emm);
        cfma(y, kx, rxp, 1.0f);
        cjfma(y, kxd, rxm, -1.0f);
        Y1[ft * W1 + fx] = y;
      }
    }
    __syncthreads();
    // 3c. w_{j+1} = M^dag y on the region; partials of iteration j+1 (a
    // thread's sites in FP32, then FP64).  y(t), y(t +- 1) come from yv[]
    // (the rows above and below the wave's from LDS) and the links from
    // registers, as in 3b.  Elements outside the lattice (a region larger
    // than the lattice) are computed too, to keep that chain, and their
    // results dropped.
    T gacc32 = 0, dacc32 = 0, racc32 = 0, sacc32 = 0;
    {
      int c1 = (wrow + 1) * W1 + (lane + 1);   // Y1 index of (wrow, lane)
      int c2i = (wrow + 2) * W2 + (lane + 2);  // R2 / link index of (wrow, lane)
      // y on the wave's rows is yv[] (3b); the rows above and below from LDS
      const V2 ytop = Y1[c1 - W1], ybot = Y1[c1 + NO * W1];
      V2 ktd = lktd0;
#pragma unroll
      for (int q = 0; q < NO; q++, c1 += W1, c2i += W2) {
        const V2 kt = link_t(q), kx = link_x(q), kxd = link_xd(q);
        const V2 yxp = Y1[c1 + 1], yxm = Y1[c1 - 1];
        __builtin_amdgcn_sched_barrier(0);  // one LDS latency per element (as in 3b)
        const V2 y0 = yv[q], ytp = q + 1 < NO ? yv[q + 1] : ybot, ytm = q > 0 ? yv[q - 1] : ytop;
        // w = m y - e^-mu K_t y(t+1) + e^mu K_t(t-1)* y(t-1) - K_x y(x+1) + K_x(x-1)* y(x-1)
        V2 v = cscale(y0, m);
        cfma(v, kt, ytp, -emm);
        cjfma(v, ktd, ytm, epm);
        cfma(v, kx, yxp, -1.0f);
        cjfma(v, kxd, yxm, 1.0f);
        ktd = kt;
        if (CLIP && !((inl >> q) & 1u)) continue;
        w[q] = v;
        M.st2(o_w[ns], pso[q], v);  // nol: dropped
        gacc32 += r[q].x * r[q].x + r[q].y * r[q].y;
        dacc32 += r[q].x * v.x + r[q].y * v.y;
        if (setup && cont) {
          racc32 += r[q].x * s[q].x + r[q].y * s[q].y;
          sacc32 += p[q].x * s[q].x + p[q].y * s[q].y;
        }
      }
    }
    // 4. the partials of each wave (fixed lane order), stored by its lane 0:
    // a chain's partials are summed by the readers, so no cross-wave
    // reduction (LDS round trip and barrier) is needed here
    const int npart = setup && cont ? 4 : 2;
    {
      const double g = wave_total_f(gacc32), d = wave_total_f(dacc32);
      double rs = 0.0, ps = 0.0;
      if (npart == 4) {
        rs = wave_total_f(racc32);
        ps = wave_total_f(sacc32);
      }
      if ((tid & 63) == 0) {
        const unsigned mp = my_part + 32u * (unsigned)(tid >> 6);
        M.std2(o_part[ns], mp, d2v{g, d});
        if (npart == 4) M.std2(o_part[ns], mp + 16, d2v{rs, ps});
      }
    }
    slot = ns;
    // 1. every region of the chain has finished iteration j
    if (!M.sync(o_bar, (unsigned)(j + 2) * (unsigned)ntc, ctl.err, red + 36, ctl.spin)) {
      failed = true;
      return false;
    }
    // 2. the chain's partials of iteration j+1, summed by every wave in the
    // same order (a butterfly leaves the total in every lane), so that all
    // waves of all regions of the chain hold bitwise the same gam, del; then
    // the halo layers of iteration j+1 (used from step 3a on: the sums only
    // wait for the partials)
    const int jn = j + 1;
    const bool first_cont = setup && cont;  // jn == 0
    // Every load is issued before the first wait: the partials (gam, del of
    // RES_PART_LANE slots per lane, 0 past the chain's), the halo values of
    // the first ring element, then in a wave-uniform branch those of the
    // other ring elements.  A wait for a load then never waits for the halo
    // loads issued after it.
    const int npt = ntc * NW;
    const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);
    d2v pgd[RES_PART_LANE], prs[RES_PART_LANE];
#pragma unroll
    for (int u = 0; u < RES_PART_LANE; u++) {
      prs[u] = d2v{0.0, 0.0};
      const int k = (tid & 63) + 64 * u;
      const unsigned po = k < npt ? 32u * (unsigned)k : nol;  // nol: 0
      pgd[u] = M.ldd2(o_part[slot], po);
      if (first_cont) prs[u] = M.ldd2(o_part[slot], po + 16);
    }
#pragma unroll
    for (int h = 0; h < NQ; h++) {
      hr[h] = hw[h] = hs[h] = zero;
      if (h == 0 || wave0 + h * NT < NR) {  // (uniform)
        const f4v v = M.ld4(o_rs[slot], ri[h]);  // nol: 0
        hr[h] = cmk<V2>(v.x, v.y);
        hs[h] = cmk<V2>(v.z, v.w);
        hw[h] = M.ld2(o_w[slot], ri[h]);
      }
    }
    double gam = 0.0, del = 0.0, rho = 0.0, sig = 0.0;  // a lane's partials in order
#pragma unroll
    for (int u = 0; u < RES_PART_LANE; u++) {
      gam += pgd[u].x;
      del += pgd[u].y;
      if (first_cont) {
        rho += prs[u].x;
        sig += prs[u].y;
      }
    }
    gam = wave_total(gam);
    del = wave_total(del);
    if (first_cont) {
      rho = wave_total(rho);
      sig = wave_total(sig);
    }
    int status = 1;  // 1: continue
    double alpha = 0.0, beta = 0.0;
    if (gam < ch.stop_rr) status = TH2D_OK;
    else if (jn >= ch.max_iter) status = TH2D_ENOCONV;
    else if (!(del > 0.0) || !isfinite(gam)) status = TH2D_EBREAKDOWN;
    else {
      const bool fresh = setup && !cont;  // jn == 0
      beta = fresh ? 0.0 : qdiv(gam, gam_old);
      // p.Ap from the CG relations, except after a residual replacement,
      // which breaks them: there p.Ap = r.w + 2 beta r.s + beta^2 p.s
      alpha = fresh ? qdiv(gam, del)
                    : qdiv(gam, first_cont ? del + 2.0 * beta * rho + beta * beta * sig
                                           : del - qdiv(beta * gam, alpha_old));
      if (!(alpha > 0.0) || !isfinite(alpha)) status = TH2D_EBREAKDOWN;
    }
    if (status != 1) {
      if (first && tid == 0)
        a.res[tl.chain] = CgResult{jn, status, gam, 1, 0, gam_old, alpha_old};
      return false;
    }
    gam_old = gam;
    alpha_old = alpha;
    ta = (T)alpha;
    tb = (T)beta;
    return true;
  };
  if (!failed && iteration(std::true_type{}, -1))
    for (int j = 0; iteration(std::false_type{}, j); j++) {
    }
  // x of the owned sites (0 after a failure), and p, s for a reliable update
  V2 *xo = a.x + ch.voff;
#pragma unroll
  for (int q = 0; q < NO; q++)
    if (!CLIP || ((inl >> q) & 1u)) {
      const unsigned oi = oi_of(q);
      *at(xo, oi) = failed ? zero : x[q];
      if (ch.cp) {
        *at((V2 *)ch.cp, oi) = p[q];
        *at((V2 *)ch.cs, oi) = s[q];
      }
    }
}

// Default tile shape for a batch: among the compiled shapes no larger
// than any lattice of the batch, the one with the least modelled load
// cost, loaded elements per owned site (2-site halo and padding of
// lattices that are not multiples of the tile) times a coalescing factor
// that favours wide rows (a row of the eo layout splits into two
// half-rows), with a penalty when the batch has fewer tiles than CUs.
// th2d_set_cg_tile overrides it (the autotuner measures instead).
int choose_shape(th2d_ctx *ctx, int nch, const Geo *G)
{
  int min_nt = 1 << 30, min_nx = 1 << 30;
  for (int k = 0; k < nch; k++) {
    min_nt = std::min(min_nt, G[k].nt);
    min_nx = std::min(min_nx, G[k].nx);
  }
  if (ctx->tile_t > 0 && ctx->tile_x > 0) {
    int s = shape_index(ctx->tile_t, ctx->tile_x);
    if (s >= 0 && ctx->tile_t <= min_nt && ctx->tile_x <= min_nx) return s;
  }
  int best = -1;
  double best_cost = 1e300;
  for (int s = 0; s < kNumShapes; s++) {
    const ShapeEntry c = kShapes[s];
    if (c.tt > min_nt || c.tx > min_nx) continue;
    double tiles = 0.0, loaded = 0.0, owned = 0.0;
    for (int k = 0; k < nch; k++) {
      double n = double((G[k].nt + c.tt - 1) / c.tt) * double((G[k].nx + c.tx - 1) / c.tx);
      tiles += n;
      loaded += n * (c.tt + 4) * (c.tx + 4);
      owned += 2.0 * G[k].vh;
    }
    double cost = loaded / owned * (1.0 + 48.0 / (c.tx + 4));
    if (tiles < ctx->num_cu) cost *= ctx->num_cu / tiles;
    if (cost < best_cost) {
      best_cost = cost;
      best = s;
    }
  }
  return best;
}

template <typename T> size_t lds_bytes(int tt, int tx)
{
  size_t e2 = (size_t)(tt + 4) * (tx + 4), e1 = (size_t)(tt + 2) * (tx + 2);
  return sizeof(c2<T>) * (3 * e2 + e1) + sizeof(double) * 48;
}

template <typename T, int K>
void launch_shape(int nblocks, hipStream_t st, const CgBatch<T> &a, int slot)
{
  using S = Shape<kShapes[K].tt, kShapes[K].tx, kShapes[K].nt>;
  hipLaunchKernelGGL((k_cg_step<T, S>), dim3(nblocks), dim3(S::NT), 0, st, a, slot);
}

template <typename T>
void launch_step(int shape, int nblocks, hipStream_t st, const CgBatch<T> &a, int slot)
{
  switch (shape) {
  case 0: launch_shape<T, 0>(nblocks, st, a, slot); break;
  case 1: launch_shape<T, 1>(nblocks, st, a, slot); break;
  case 2: launch_shape<T, 2>(nblocks, st, a, slot); break;
  case 3: launch_shape<T, 3>(nblocks, st, a, slot); break;
  case 4: launch_shape<T, 4>(nblocks, st, a, slot); break;
  case 5: launch_shape<T, 5>(nblocks, st, a, slot); break;
  case 6: launch_shape<T, 6>(nblocks, st, a, slot); break;
  case 7: launch_shape<T, 7>(nblocks, st, a, slot); break;
  case 8: launch_shape<T, 8>(nblocks, st, a, slot); break;
  case 9: launch_shape<T, 9>(nblocks, st, a, slot); break;
  default: launch_shape<T, 10>(nblocks, st, a, slot); break;
  }
}
static_assert(kNumShapes == 11, "launch_step lists every shape");

// A run of kStepGraph step launches (parities 0, 1, 0, ...) captured as a
// graph, cached per context by the launch arguments (element size, tile
// shape, grid and the batch descriptor's bytes); at most 16 are kept.  The
// capture runs on a private stream of the context (the context's own
// stream may be the null stream, which cannot be captured); the graph is
// then launched into the context's stream.  *out stays null when the
// capture is refused: the caller then launches eagerly.
constexpr int kStepGraph = 8;
template <typename T>
int step_graph(th2d_ctx *ctx, int shape, int nblocks, const CgBatch<T> &a, hipGraphExec_t *out)
{
  *out = nullptr;
  const int hdr[3] = {(int)sizeof(T), shape, nblocks};
  std::vector<unsigned char> key(sizeof(hdr) + sizeof(a));
  memcpy(key.data(), hdr, sizeof(hdr));
  memcpy(key.data() + sizeof(hdr), &a, sizeof(a));
  for (const auto &g : ctx->cg_graphs)
    if (g.key == key) {
      *out = g.exec;
      return TH2D_OK;
    }
  if (!ctx->cap_stream)
    TH2D_HIP(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
  if (hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return TH2D_OK;
  }
  for (int k = 0; k < kStepGraph; k++) launch_step(shape, nblocks, ctx->cap_stream, a, k & 1);
  hipGraph_t graph = nullptr;
  TH2D_HIP(hipStreamEndCapture(ctx->cap_stream, &graph));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return TH2D_EHIP;
  if (ctx->cg_graphs.size() >= 16) {
    TH2D_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipGraphExecDestroy(ctx->cg_graphs.front().exec);
    ctx->cg_graphs.erase(ctx->cg_graphs.begin());
  }
  ctx->cg_graphs.push_back(th2d_ctx::CgGraph{std::move(key), ex});
  *out = ex;
  return TH2D_OK;
}

struct InnerResult {
  int iters;
  double rr;
  int status;
  double gam, alpha;  // resident solver: state for a reliable update
  int resident;       // solved by the resident kernel
};

// Carried Krylov state of a chain between passes of the mixed solver (a
// reliable update): p, s in and out; cscale > 0 continues.
struct Carry {
  void *p, *s;
  double cscale, gam0, alpha0;
};

// ---- small lattices: a whole FP64 inner solve in one workgroup ----------
// One workgroup of NT threads per chain, SPT sites per thread (V <= NT SPT,
// at most 2048 sites):
// x, r, p and the site's four folded links stay in registers; the operator
// reads its neighbours from one LDS copy of the vector it is applied to;
// plain CG (two block sums per iteration) runs until r.r < stop or
// max_iter, in one launch for every chain of the batch.  For the small
// lattices of the reference's own runs (16^2: one site per thread) this
// replaces a launch and a host check per iteration.
struct SmallChain {
  Geo G;
  const double2 *b, *K;
  double2 *x;
  double stop;
  int max_iter, pad;
};

template <int NW>
__device__ __forceinline__ double small_sum(double v, double *red)
{
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NW; w++) s += red[w];  // every thread, same order
  return s;
}

template <int NT, int SPT, bool KREG>
__global__ void __launch_bounds__(NT) k_cg_small(const SmallChain *chains, CgResult *out)
{
  constexpr int NW = NT / 64;
  __shared__ double2 buf[NT * SPT];
  __shared__ double red[2][NW];
  const SmallChain C = chains[blockIdx.x];
  const Geo G = C.G;
  const int V = 2 * G.vh;
  double2 x[SPT], r[SPT], p[SPT], q[SPT], kt[SPT], kx[SPT], ktd[SPT], kxd[SPT];
  int itu[SPT], itd[SPT], ixu[SPT], ixd[SPT];
  double rr = 0.0;
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const int i = threadIdx.x + k * NT;
    x[k] = r[k] = p[k] = q[k] = make_double2(0.0, 0.0);
    kt[k] = kx[k] = ktd[k] = kxd[k] = make_double2(0.0, 0.0);
    itu[k] = itd[k] = ixu[k] = ixd[k] = 0;
    if (i < V) {
      const EoSite st = eo_site(G, i);
      itu[k] = st.itu;
      itd[k] = st.itd;
      ixu[k] = st.ixu;
      ixd[k] = st.ixd;
      if (KREG) {  // else (1024 threads, 4 sites each) read in every apply
        kt[k] = C.K[i];
        kx[k] = C.K[V + i];
        ktd[k] = C.K[st.itd];
        kxd[k] = C.K[V + st.ixd];
      }
      r[k] = p[k] = C.b[i];
      rr += r[k].x * r[k].x + r[k].y * r[k].y;
    }
  }
  rr = small_sum<NW>(rr, red[0]);
  int it = 0, status = TH2D_OK;
  auto apply = [&](int dagger, double2 *dst) {
    const double fu = dagger ? -G.em : G.ep, fd = dagger ? -G.ep : G.em, sx = dagger ? -1.0 : 1.0;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      const int i = threadIdx.x + k * NT;
      if (i >= V) continue;
      const double2 a = KREG ? kt[k] : C.K[i], b = KREG ? ktd[k] : C.K[itd[k]];
      const double2 c = KREG ? kx[k] : C.K[V + i], d = KREG ? kxd[k] : C.K[V + ixd[k]];
      double2 v = cscale(buf[i], G.m);
      v = cadd(v, cscale(cmul(a, buf[itu[k]]), fu));
      v = csub(v, cscale(cjmul(b, buf[itd[k]]), fd));
      v = cadd(v, cscale(cmul(c, buf[ixu[k]]), sx));
      v = csub(v, cscale(cjmul(d, buf[ixd[k]]), sx));
      dst[k] = v;
    }
  };
  while (!(rr < C.stop) && it < C.max_iter) {
    double2 y[SPT];
#pragma unroll
    for (int k = 0; k < SPT; k++)
      if (threadIdx.x + k * NT < V) buf[threadIdx.x + k * NT] = p[k];
    __syncthreads();
    apply(0, y);  // y = M p
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SPT; k++)
      if (threadIdx.x + k * NT < V) buf[threadIdx.x + k * NT] = y[k];
    __syncthreads();
    apply(1, q);  // q = M^dag y
    double pq = 0.0;
#pragma unroll
    for (int k = 0; k < SPT; k++)
      if (threadIdx.x + k * NT < V) pq += p[k].x * q[k].x + p[k].y * q[k].y;
    pq = small_sum<NW>(pq, red[0]);
    if (!(pq > 0.0) || !isfinite(pq)) {
      status = TH2D_EBREAKDOWN;
      break;
    }
    const double alpha = rr / pq;
    double rn = 0.0;
#pragma unroll
    for (int k = 0; k < SPT; k++) {
      x[k] = cadd(x[k], cscale(p[k], alpha));
      r[k] = csub(r[k], cscale(q[k], alpha));
      rn += r[k].x * r[k].x + r[k].y * r[k].y;
    }
    rn = small_sum<NW>(rn, red[1]);
    it++;
    const double beta = rn / rr;
    rr = rn;
    if (rr < C.stop) break;
#pragma unroll
    for (int k = 0; k < SPT; k++) p[k] = cadd(r[k], cscale(p[k], beta));
  }
  if (status == TH2D_OK && !(rr < C.stop)) status = TH2D_ENOCONV;
#pragma unroll
  for (int k = 0; k < SPT; k++)
    if (threadIdx.x + k * NT < V) C.x[threadIdx.x + k * NT] = x[k];
  if (threadIdx.x == 0) {
    CgResult o;
    memset(&o, 0, sizeof(o));
    o.iters = it;
    o.status = status;
    o.rr = rr;
    o.done = 1;
    out[blockIdx.x] = o;
  }
}

// The whole FP64 solve of th2d_cg_mdagm_batched on a small lattice in one
// workgroup: |b|^2, the stop target, CG on (M^dag M) x = b from x = 0 in
// the Chronopoulos-Gear form (both inner products of an iteration in one
// block sum), then the true residual b - M^dag M x in FP64 and, while it
// misses the target, restarts from it, as the host loop does for larger
// lattices.  x is written to the caller's vector.
struct SmallFull {
  Geo G;
  const double2 *b, *K;
  double2 *x;
  double tol_rel, tol_abs_rr;
  int max_iter, pad;
};

struct SmallInfo {
  int iters, restarts, status, pad;
  double true_rel, rr;
};

template <int NT, int SPT>
__global__ void __launch_bounds__(NT) k_cg_small_full(const SmallFull *chains, SmallInfo *out,
                                                      int accumulate)
{
  constexpr int NW = NT / 64;
  __shared__ double2 buf[NT * SPT];
  __shared__ double red[2][2 * NW];
  const SmallFull C = chains[blockIdx.x];
  const Geo G = C.G;
  const int V = 2 * G.vh;
  double2 x[SPT], r[SPT], p[SPT], s[SPT], kt[SPT], kx[SPT], ktd[SPT], kxd[SPT];
  int itu[SPT], itd[SPT], ixu[SPT], ixd[SPT];
  double bb = 0.0;
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const int i = threadIdx.x + k * NT;
    x[k] = r[k] = p[k] = s[k] = make_double2(0.0, 0.0);
    kt[k] = kx[k] = ktd[k] = kxd[k] = make_double2(0.0, 0.0);
    itu[k] = itd[k] = ixu[k] = ixd[k] = 0;
    if (i < V) {
      const EoSite st = eo_site(G, i);
      itu[k] = st.itu;
      itd[k] = st.itd;
      ixu[k] = st.ixu;
      ixd[k] = st.ixd;
      kt[k] = C.K[i];
      kx[k] = C.K[V + i];
      ktd[k] = C.K[st.itd];
      kxd[k] = C.K[V + st.ixd];
      r[k] = C.b[i];
      bb += r[k].x * r[k].x + r[k].y * r[k].y;
    }
  }
  int rb = 0;  // which reduction buffer; alternates, a barrier between reuses
  auto sum2 = [&](double a, double b2) {
    a = wave_sum(a);
    b2 = wave_sum(b2);
    double *rd = red[rb];
    rb ^= 1;
    if ((threadIdx.x & 63) == 0) {
      rd[2 * (threadIdx.x >> 6)] = a;
      rd[2 * (threadIdx.x >> 6) + 1] = b2;
    }
    __syncthreads();
    double2 o = make_double2(0.0, 0.0);
#pragma unroll
    for (int w = 0; w < NW; w++) {
      o.x += rd[2 * w];
      o.y += rd[2 * w + 1];
    }
    return o;
  };
  // w = M^dag M v for the thread's sites, v given per site
  auto mdagm = [&](const double2 *v, double2 *w) {
    double2 y[SPT];
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
      const double2 *src = pass == 0 ? v : y;
#pragma unroll
      for (int k = 0; k < SPT; k++)
        if (threadIdx.x + k * NT < V) buf[threadIdx.x + k * NT] = src[k];
      __syncthreads();
      const double fu = pass ? -G.em : G.ep, fd = pass ? -G.ep : G.em, sx = pass ? -1.0 : 1.0;
      double2 o[SPT];
#pragma unroll
      for (int k = 0; k < SPT; k++) {
        const int i = threadIdx.x + k * NT;
        o[k] = make_double2(0.0, 0.0);
        if (i >= V) continue;
        double2 t = cscale(buf[i], G.m);
        t = cadd(t, cscale(cmul(kt[k], buf[itu[k]]), fu));
        t = csub(t, cscale(cjmul(ktd[k], buf[itd[k]]), fd));
        t = cadd(t, cscale(cmul(kx[k], buf[ixu[k]]), sx));
        t = csub(t, cscale(cjmul(kxd[k], buf[ixd[k]]), sx));
        o[k] = t;
      }
      __syncthreads();  // every thread has read buf before it is rewritten
#pragma unroll
      for (int k = 0; k < SPT; k++) (pass == 0 ? y : w)[k] = o[k];
    }
  };
  bb = sum2(bb, 0.0).x;
  const bool abs_mode = C.tol_rel <= 0;
  double target = 0.0;
  if (C.tol_rel > 0) target = C.tol_rel * C.tol_rel * bb;
  if (C.tol_abs_rr > 0) target = fmax(target, C.tol_abs_rr);
  int it = 0, restarts = 0, status = TH2D_OK;
  double rtrue = bb, rr = bb;
  if (bb > 0.0 && !(bb <= target)) {
    for (int pass = 0; pass < 64; pass++) {
      // CG from x with r the current residual: p = s = 0, then per
      // iteration w = A r, gamma = r.r, delta = r.w (one block sum)
      double gam_old = 1.0, alpha_old = 1.0;
      bool broke = false;
#pragma unroll
      for (int k = 0; k < SPT; k++) p[k] = s[k] = make_double2(0.0, 0.0);
      for (int j = 0; it < C.max_iter; j++) {
        double2 w[SPT];
        mdagm(r, w);
        double g = 0.0, d = 0.0;
#pragma unroll
        for (int k = 0; k < SPT; k++) {
          g += r[k].x * r[k].x + r[k].y * r[k].y;
          d += r[k].x * w[k].x + r[k].y * w[k].y;
        }
        const double2 gd = sum2(g, d);
        rr = gd.x;
        if (rr < target) break;  // the recursive residual meets the target
        const double beta = j == 0 ? 0.0 : gd.x / gam_old;
        const double den = gd.y - beta * gd.x / alpha_old;
        if (!(den > 0.0) || !isfinite(den)) {
          broke = true;
          break;
        }
        const double alpha = gd.x / den;
#pragma unroll
        for (int k = 0; k < SPT; k++) {
          p[k] = cadd(r[k], cscale(p[k], beta));
          s[k] = cadd(w[k], cscale(s[k], beta));
          x[k] = cadd(x[k], cscale(p[k], alpha));
          r[k] = csub(r[k], cscale(s[k], alpha));
        }
        gam_old = gd.x;
        alpha_old = alpha;
        it++;
      }
      // the true residual in FP64
      double2 ax[SPT];
      mdagm(x, ax);
      double rn = 0.0;
#pragma unroll
      for (int k = 0; k < SPT; k++) {
        const int i = threadIdx.x + k * NT;
        r[k] = i < V ? csub(C.b[i], ax[k]) : make_double2(0.0, 0.0);
        rn += r[k].x * r[k].x + r[k].y * r[k].y;
      }
      rn = sum2(rn, 0.0).x;
      if (pass > 0) restarts++;
      if (abs_mode && !broke) {  // the reference's stop on the recursive r.r
        rtrue = rn;
        break;
      }
      if (broke && !(rn < rtrue)) {
        status = TH2D_EBREAKDOWN;
        rtrue = rn;
        break;
      }
      if (!(rn < rtrue) && pass > 2) {  // stagnation at round-off
        rtrue = rn;
        break;
      }
      rtrue = rn;
      if (rtrue <= target) break;
      if (it >= C.max_iter) {
        status = TH2D_ENOCONV;
        break;
      }
    }
  }
  if (status == TH2D_OK && rtrue > target && !abs_mode && bb > 0) status = TH2D_ENOCONV;
#pragma unroll
  for (int k = 0; k < SPT; k++)
    if (threadIdx.x + k * NT < V) C.x[threadIdx.x + k * NT] = x[k];
  if (threadIdx.x == 0) {
    const double tr = bb > 0 ? sqrt(rtrue / bb) : 0.0;
    SmallInfo &o = out[blockIdx.x];
    if (accumulate) {  // successive solves of one chain: sums, worst, last
      o.iters += it;
      o.restarts += restarts;
      if (o.status == TH2D_OK) o.status = status;
      o.pad += 1;  // solves
      o.true_rel = fmax(o.true_rel, tr);
      o.rr = rr;
    } else {
      o.iters = it;
      o.restarts = restarts;
      o.status = status;
      o.pad = 1;
      o.true_rel = tr;
      o.rr = rr;
    }
  }
}

__global__ void k_small_info_zero(SmallInfo *out, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = SmallInfo{0, 0, TH2D_OK, 0, 0.0, 0.0};
}

constexpr long kSmallMaxSites = 2048;  // 1024 threads x 2 sites without spills

bool small_fits(int nch, const Geo *G)
{
  for (int c = 0; c < nch; c++)
    if (2L * G[c].vh > kSmallMaxSites) return false;
  return nch > 0;
}

// The batched inner solve (A_c y_c = rhs_c from y_c = 0, stop r.r < stop_c)
// on small lattices, one workgroup per chain; solutions in xout[c].
int small_solve(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G,
                const double2 *const *rhs, const double2 *const *K, const double *stop_rr,
                const int *max_iter, double2 **xout, InnerResult *res)
{
  long maxV = 0, voff = 0;
  std::vector<long> off(nch);
  for (int c = 0; c < nch; c++) {
    const long V = 2L * G[c].vh;
    maxV = std::max(maxV, V);
    off[c] = voff;
    voff += (V + 63) & ~63L;
  }
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  const size_t vb = al(sizeof(double2) * (size_t)voff), cb = al(sizeof(SmallChain) * nch);
  TH2D_TRY(th2d_buf_reserve(buf, vb + cb + sizeof(CgResult) * nch));
  double2 *X = (double2 *)buf.ptr;
  SmallChain *dch = (SmallChain *)((char *)buf.ptr + vb);
  CgResult *dres = (CgResult *)((char *)buf.ptr + vb + cb);
  std::vector<SmallChain> ch(nch);
  for (int c = 0; c < nch; c++) {
    memset(&ch[c], 0, sizeof(SmallChain));
    ch[c].G = G[c];
    ch[c].b = rhs[c];
    ch[c].K = K[c];
    ch[c].x = X + off[c];
    ch[c].stop = stop_rr[c];
    ch[c].max_iter = max_iter[c];
    xout[c] = X + off[c];
  }
  TH2D_TRY(th2d::stage_h2d(ctx, dch, ch.data(), sizeof(SmallChain) * nch));
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  if (maxV <= 256)
    hipLaunchKernelGGL((k_cg_small<256, 1, true>), dim3(nch), dim3(256), 0, ctx->stream, dch, dres);
  else if (maxV <= 512)
    hipLaunchKernelGGL((k_cg_small<256, 2, true>), dim3(nch), dim3(256), 0, ctx->stream, dch, dres);
  else if (maxV <= 1024)
    hipLaunchKernelGGL((k_cg_small<256, 4, true>), dim3(nch), dim3(256), 0, ctx->stream, dch, dres);
  else
    hipLaunchKernelGGL((k_cg_small<1024, 2, true>), dim3(nch), dim3(1024), 0, ctx->stream, dch, dres);
  TH2D_HIP(hipGetLastError());
  const size_t rb = sizeof(CgResult) * nch;
  TH2D_TRY(th2d_host_reserve(ctx->redh, rb + 64));
  TH2D_HIP(hipMemcpyAsync(ctx->redh.ptr, dres, rb, hipMemcpyDeviceToHost, ctx->stream));
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k1, ctx->stream));
  TH2D_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<CgResult> h(nch);
  memcpy(h.data(), ctx->redh.ptr, rb);
  if (ctx->ktiming) {
    float ms = 0.f;
    TH2D_HIP(hipEventElapsedTime(&ms, ctx->k0, ctx->k1));
    ctx->ktime_ms += ms;
    ctx->klaunches += 1;
    for (int c = 0; c < nch; c++) {
      ctx->kbytes += (double)h[c].iters * 2.0 * G[c].vh * 240.0;
      ctx->kiters += h[c].iters;
    }
  }
  for (int c = 0; c < nch; c++) {
    memset(&res[c], 0, sizeof(InnerResult));
    res[c].iters = h[c].iters;
    res[c].rr = h[c].rr;
    res[c].status = h[c].status;
  }
  ctx->cg_last_resident = 3;
  return TH2D_OK;
}

// th2d_cg_mdagm_batched in FP64 for small lattices: one launch.  Set-up
// (descriptors into `buf`) and launch are separate so that the HMC
// trajectory can put the launch into a captured graph.
int small_full_prepare(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G, double *const *x,
                       const double *const *b, const double *const *K, const th2d_cg_params *p,
                       SmallFull **dch, SmallInfo **dinf, long *maxV)
{
  *maxV = 0;
  for (int c = 0; c < nch; c++) *maxV = std::max(*maxV, 2L * G[c].vh);
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  const size_t cb = al(sizeof(SmallFull) * nch);
  TH2D_TRY(th2d_buf_reserve(buf, cb + sizeof(SmallInfo) * nch));
  *dch = (SmallFull *)buf.ptr;
  *dinf = (SmallInfo *)((char *)buf.ptr + cb);
  std::vector<SmallFull> ch(nch);
  for (int c = 0; c < nch; c++) {
    memset(&ch[c], 0, sizeof(SmallFull));
    ch[c].G = G[c];
    ch[c].b = (const double2 *)b[c];
    ch[c].K = (const double2 *)K[c];
    ch[c].x = (double2 *)x[c];
    ch[c].tol_rel = p->tol_rel;
    ch[c].tol_abs_rr = p->tol_abs_rr;
    ch[c].max_iter = p->max_iter;
  }
  TH2D_TRY(th2d::stage_h2d(ctx, *dch, ch.data(), sizeof(SmallFull) * nch));
  return TH2D_OK;
}

void small_full_launch(hipStream_t st, int nch, long maxV, const SmallFull *dch, SmallInfo *dinf,
                       int accumulate)
{
  if (maxV <= 256)
    hipLaunchKernelGGL((k_cg_small_full<256, 1>), dim3(nch), dim3(256), 0, st, dch, dinf,
                       accumulate);
  else if (maxV <= 512)
    hipLaunchKernelGGL((k_cg_small_full<256, 2>), dim3(nch), dim3(256), 0, st, dch, dinf,
                       accumulate);
  else if (maxV <= 1024)
    hipLaunchKernelGGL((k_cg_small_full<256, 4>), dim3(nch), dim3(256), 0, st, dch, dinf,
                       accumulate);
  else
    hipLaunchKernelGGL((k_cg_small_full<1024, 2>), dim3(nch), dim3(1024), 0, st, dch, dinf,
                       accumulate);
}

int small_full_solve(th2d_ctx *ctx, int nch, const Geo *G, double *const *x,
                     const double *const *b, const double *const *K, const th2d_cg_params *p,
                     th2d_cg_info *infos)
{
  SmallFull *dch;
  SmallInfo *dinf;
  long maxV;
  TH2D_TRY(small_full_prepare(ctx, ctx->cg64, nch, G, x, b, K, p, &dch, &dinf, &maxV));
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  small_full_launch(ctx->stream, nch, maxV, dch, dinf, 0);
  TH2D_HIP(hipGetLastError());
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k1, ctx->stream));
  const size_t rb = sizeof(SmallInfo) * nch;
  TH2D_TRY(th2d_host_reserve(ctx->redh, rb + 64));
  TH2D_HIP(hipMemcpyAsync(ctx->redh.ptr, dinf, rb, hipMemcpyDeviceToHost, ctx->stream));
  TH2D_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<SmallInfo> h(nch);
  memcpy(h.data(), ctx->redh.ptr, rb);
  if (ctx->ktiming) {
    float ms = 0.f;
    TH2D_HIP(hipEventElapsedTime(&ms, ctx->k0, ctx->k1));
    ctx->ktime_ms += ms;
    ctx->klaunches += 1;
    for (int c = 0; c < nch; c++) {
      ctx->kbytes += (double)h[c].iters * 2.0 * G[c].vh * 240.0;
      ctx->kiters += h[c].iters;
    }
  }
  ctx->cg_last_resident = 3;
  int worst = TH2D_OK;
  for (int c = 0; c < nch; c++) {
    if (infos) {
      infos[c].iterations = h[c].iters;
      infos[c].restarts = h[c].restarts;
      infos[c].status = h[c].status;
      infos[c].true_rel_resid = h[c].true_rel;
      infos[c].rr = h[c].rr;
    }
    if (h[c].status != TH2D_OK && worst == TH2D_OK) worst = h[c].status;
  }
  return worst;
}

// Batched inner solve: A_c y_c = rhs_c from y_c = 0, stop r.r < stop_c.
// The solutions are the internal x vectors, returned in xout[c].
// force_shape: -1 automatic (resident kernel where it applies), -2 step
// kernel with an automatic tile shape, >= 0 that step-kernel shape.
template <typename T>
int inner_solve(th2d_ctx *ctx, th2d_buf &buf, int pass, int nch, const Geo *G,
                const c2<T> *const *rhs,
                const c2<T> *const *K, const double *stop_rr, const int *max_iter, c2<T> **xout,
                InnerResult *res, int force_shape = -1, const Carry *carry = nullptr);

// Key of a batch for the autotuner: the distinct lattice shapes and the
// order of magnitude of the site count (a restart pass over a subset of
// the chains reuses the choice made for the whole batch).
unsigned long long tune_key(int nch, const Geo *G)
{
  std::vector<std::pair<int, int>> d;
  long sites = 0;
  for (int c = 0; c < nch; c++) {
    d.emplace_back(G[c].nt, G[c].nx);
    sites += 2L * G[c].vh;
  }
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  unsigned long long h = 1469598103934665603ull;
  auto mix = [&](unsigned long long v) { h = (h ^ v) * 1099511628211ull; };
  for (auto &q : d) mix(((unsigned long long)q.first << 32) | (unsigned)q.second);
  int lg = 0;
  while ((1L << lg) < sites) lg++;
  mix(lg);
  return h;
}

// Measuring autotuner: runs `iters` solver iterations of the actual batch
// with every compiled tile shape that fits, times them with HIP events and
// returns the fastest (the solver state is rebuilt by the real solve).
template <typename T>
int autotune_shape(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G, const c2<T> *const *rhs,
                   const c2<T> *const *K, int iters)
{
  int min_nt = 1 << 30, min_nx = 1 << 30;
  for (int k = 0; k < nch; k++) {
    min_nt = std::min(min_nt, G[k].nt);
    min_nx = std::min(min_nx, G[k].nx);
  }
  std::vector<double> stop(nch, 0.0);
  std::vector<int> mi(nch, iters);
  std::vector<c2<T> *> xo(nch);
  std::vector<InnerResult> ir(nch);
  const int saved_timing = ctx->ktiming, saved_chunk = ctx->cg_chunk;
  ctx->ktiming = 0;
  ctx->cg_chunk = iters + 1;
  int best = -1;
  float best_ms = 1e30f;
  int rc = TH2D_OK;
  for (int s = 0; s < kNumShapes && rc == TH2D_OK; s++) {
    if (kShapes[s].tt > min_nt || kShapes[s].tx > min_nx) continue;
    float ms = 0.f;
    for (int rep = 0; rep < 2 && rc == TH2D_OK; rep++) {  // the first run warms up
      if (hipEventRecord(ctx->a0, ctx->stream) != hipSuccess) {
        rc = TH2D_EHIP;
        break;
      }
      rc = inner_solve<T>(ctx, buf, 7, nch, G, rhs, K, stop.data(), mi.data(), xo.data(),
                          ir.data(), s);
      if (rc != TH2D_OK) break;
      if (hipEventRecord(ctx->a1, ctx->stream) != hipSuccess ||
          hipEventSynchronize(ctx->a1) != hipSuccess ||
          hipEventElapsedTime(&ms, ctx->a0, ctx->a1) != hipSuccess)
        rc = TH2D_EHIP;
    }
    if (rc == TH2D_OK && ms < best_ms) {
      best_ms = ms;
      best = s;
    }
  }
  ctx->ktiming = saved_timing;
  ctx->cg_chunk = saved_chunk;
  if (rc != TH2D_OK) return rc;
  return best >= 0 ? best : TH2D_EARG;
}

// Chain and tile tables of a batch cut into tt x tx tiles, and the device
// block of the solver (8 vectors, partials, state, results, tables and
// `extra` bytes), in `buf`; the tables are uploaded on the context stream.
template <typename T>
int build_batch(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G, const c2<T> *const *rhs,
                const c2<T> *const *K, const double *stop_rr, const int *max_iter, int tt, int tx,
                size_t extra, CgBatch<T> &a, std::vector<CgChain> &ch, void **extra_ptr,
                const Carry *carry = nullptr, const void *extra_init = nullptr,
                size_t extra_init_bytes = 0)
{
  memset(&a, 0, sizeof(a));  // padding included: the bytes key captured graphs
  ch.assign(nch, CgChain{});
  std::vector<CgTile> tiles;
  long voff = 0;
  int pbeg = 0;
  for (int c = 0; c < nch; c++) {
    ch[c].G = G[c];
    ch[c].tbeg = (int)tiles.size();
    for (int t0 = 0; t0 < G[c].nt; t0 += tt)
      for (int x0 = 0; x0 < G[c].nx; x0 += tx) tiles.push_back(CgTile{c, t0, x0, 0});
    ch[c].tend = (int)tiles.size();
    ch[c].voff = voff;
    voff += (2L * G[c].vh + 63) & ~63L;
    ch[c].b = rhs[c];
    ch[c].K = K[c];
    ch[c].stop_rr = stop_rr[c];
    ch[c].max_iter = max_iter[c];
    ch[c].pad = pbeg;  // partial slot of the chain's first tile (8-aligned: a cache-line boundary)
    pbeg += (ch[c].tend - ch[c].tbeg + 7) & ~7;
    ch[c].cp = carry ? carry[c].p : nullptr;
    ch[c].cs = carry ? carry[c].s : nullptr;
    ch[c].cscale = carry ? carry[c].cscale : 0.0;
    ch[c].gam0 = carry ? carry[c].gam0 : 0.0;
    ch[c].alpha0 = carry ? carry[c].alpha0 : 0.0;
  }
  const int ntiles = (int)tiles.size();
  size_t vec = sizeof(c2<T>) * (size_t)voff;
  vec = (vec + 255) & ~(size_t)255;
  auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
  size_t part = al(sizeof(double) * 4 * (size_t)ntiles);
  size_t stb = al(sizeof(CgState) * 2 * (size_t)nch);
  // + the resident solver's flags, and |r|^2 of each chain for a caller
  // that reads them with the results (resident_collect)
  size_t rsb = al(sizeof(CgResult) * (size_t)nch + 64 + sizeof(double) * (size_t)nch);
  size_t chb = al(sizeof(CgChain) * (size_t)nch);
  size_t tlb = al(sizeof(CgTile) * (size_t)ntiles);
  size_t exb = al(extra);
  size_t bytes = 8 * vec + part + stb + rsb + chb + tlb + exb;
  TH2D_TRY(th2d_buf_reserve(buf, bytes));
  char *pb = (char *)buf.ptr;
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
  *extra_ptr = pb;
  pb += exb;
  if ((size_t)(pb - (char *)buf.ptr) > bytes) return TH2D_EARG;
  a.chains = dch;
  a.tiles = dtl;
  a.ntiles = ntiles;
  a.nchains = nch;
  a.TT = tt;
  a.TX = tx;
  // tables (and the first bytes of `extra`, which follow them) through a
  // pinned staging block, in one copy
  if (extra_init_bytes > exb) return TH2D_EARG;
  size_t hbytes = std::max(chb + tlb + extra_init_bytes, rsb);
  TH2D_TRY(th2d_host_reserve(ctx->redh, hbytes + 64));
  memcpy(ctx->redh.ptr, ch.data(), sizeof(CgChain) * nch);
  memcpy((char *)ctx->redh.ptr + chb, tiles.data(), sizeof(CgTile) * ntiles);
  if (extra_init_bytes) memcpy((char *)ctx->redh.ptr + chb + tlb, extra_init, extra_init_bytes);
  TH2D_HIP(hipMemcpyAsync(dch, ctx->redh.ptr, chb + tlb + extra_init_bytes, hipMemcpyHostToDevice,
                          ctx->stream));
  return TH2D_OK;
}

// Per-chain results into the caller's arrays, the solver's kernel time and
// algorithmic bytes into the context counters.
template <typename T>
int finish_solve(th2d_ctx *ctx, int nch, const Geo *G, const CgBatch<T> &a,
                 const std::vector<CgChain> &ch, const std::vector<CgResult> &h, long launches,
                 c2<T> **xout, InnerResult *res)
{
  if (ctx->ktiming) {
    TH2D_HIP(hipEventRecord(ctx->k1, ctx->stream));
    TH2D_HIP(hipEventSynchronize(ctx->k1));
    float ms = 0.f;
    TH2D_HIP(hipEventElapsedTime(&ms, ctx->k0, ctx->k1));
    ctx->ktime_ms += ms;
    ctx->klaunches += launches;
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
    res[c].gam = h[c].gam_old;
    res[c].alpha = h[c].alpha_old;
    res[c].resident = 0;
    xout[c] = a.x + ch[c].voff;
  }
  return TH2D_OK;
}

__global__ void k_zero_u32(unsigned *p, int n, unsigned *q, int nq)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) q[i] = 0u;
}

// Regions of a lattice, and of a batch, for regions of tt rows.
inline long res_chain_regions(const Geo &g, int tt)
{
  return (long)((g.nt + tt - 1) / tt) * ((g.nx + RES_T - 1) / RES_T);
}
int res_regions(int nch, const Geo *G, int tt = RES_T)
{
  long n = 0;
  for (int c = 0; c < nch; c++) n += res_chain_regions(G[c], tt);
  return n > (1L << 30) ? (1 << 30) : (int)n;
}

// Lattices the resident solver takes: one region (both sides at most 64),
// or sides that are multiples of 64.  A lattice cut into several regions
// with a clipped edge region broke down now and then on the GPU when the
// XCDs were full (wrong halo values in those chains only), while a CPU
// emulation of the kernel with random schedules solved the same batches
// exactly; until the cause is known the step kernel solves such batches.
bool res_eligible(const Geo &g, int tt = RES_T)
{
  const bool one = g.nt <= tt && g.nx <= RES_T;
  return one || (g.nt % tt == 0 && g.nx % RES_T == 0);
}

// Resident workgroups the device holds (cooperative launch bound); the
// same for both region shapes (one workgroup per CU).
int res_capacity(th2d_ctx *ctx)
{
  if (ctx->res_capacity < 0) {
    int occ = 0, occh = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, (const void *)k_cg_resident<float, true, RShape, true>, RShape::NT, 0) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occh, (const void *)k_cg_resident<float, true, RShapeH, true>, RShapeH::NT, 0) !=
            hipSuccess)
      occ = occh = 0;
    ctx->res_capacity = std::min(occ, occh) * ctx->num_cu;
  }
  return ctx->res_capacity;
}

// Places the regions of every chain on one XCD.  Workgroups are dispatched
// round-robin over the XCDs, so workgroups x, x + 8, x + 16, ... of a
// launch run on XCD x.  Chains are packed largest first: into the first
// group (one cooperative launch) with an XCD that has room for the whole
// chain (at most per_xcd regions per XCD), onto its least-loaded such XCD;
// a new group when none has.  Fills wg_tile with every group's table (the
// tile of each workgroup, -1: idle) and `groups` with each group's grid and
// table offset; false when a chain does not fit on one XCD.  The kernel
// checks the placement with the XCD register.
constexpr int kNumXcd = 8;
struct ResGroup {
  int grid;        // workgroups of the launch (a multiple of 8)
  size_t tab;      // offset of its table in wg_tile
  long sites;      // lattice sites it solves
};
bool place_regions(int nch, const Geo *G, int per_xcd, int tt, std::vector<int> &wg_tile,
                   std::vector<ResGroup> &groups)
{
  std::vector<int> ntc(nch), tbeg(nch + 1, 0), order(nch);
  for (int c = 0; c < nch; c++) {
    ntc[c] = (int)res_chain_regions(G[c], tt);
    if (ntc[c] > per_xcd) return false;
    tbeg[c + 1] = tbeg[c] + ntc[c];
    order[c] = c;
  }
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return ntc[p] > ntc[q]; });
  std::vector<std::vector<int>> load;                // [group][xcd]
  std::vector<std::vector<std::vector<int>>> on;     // [group][xcd] chains
  std::vector<long> sites;
  for (int c : order) {
    int gb = -1, best = -1;
    for (int g = 0; g < (int)load.size() && gb < 0; g++)
      for (int x = 0; x < kNumXcd; x++)
        if (load[g][x] + ntc[c] <= per_xcd && (best < 0 || load[g][x] < load[g][best])) {
          gb = g;
          best = x;
        }
    if (gb < 0) {
      gb = (int)load.size();
      best = 0;
      load.emplace_back(kNumXcd, 0);
      on.emplace_back(kNumXcd);
      sites.push_back(0);
    }
    on[gb][best].push_back(c);
    load[gb][best] += ntc[c];
    sites[gb] += 2L * G[c].vh;
  }
  wg_tile.clear();
  groups.clear();
  for (size_t g = 0; g < load.size(); g++) {
    const int slots = *std::max_element(load[g].begin(), load[g].end());
    const size_t tab = wg_tile.size();
    wg_tile.resize(tab + (size_t)kNumXcd * slots, -1);
    for (int x = 0; x < kNumXcd; x++) {
      int slot = 0;
      for (int c : on[g][x])
        for (int t = tbeg[c]; t < tbeg[c + 1]; t++) wg_tile[tab + (size_t)slot++ * kNumXcd + x] = t;
    }
    groups.push_back(ResGroup{kNumXcd * slots, tab, sites[g]});
  }
  return true;
}

// One inner solve as cooperative launches of the resident kernel (FP32),
// one per group of chains that fits the device's resident workgroups;
// TH2D_ELAUNCH when a chain does not fit on one XCD, or when the groups are
// modelled slower than the step kernel: an iteration of a group costs
// about 8.5 us (measured at 16 lattices of 256^2, set by the per-iteration
// synchronisation), one of the step kernel about 10 us plus 120 B per site
// at 4.5 TB/s (measured).  Many small lattices favour the step kernel.
template <typename S>
int resident_solve_s(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G,
                     const c2<float> *const *rhs, const c2<float> *const *K, const double *stop_rr,
                     const int *max_iter, c2<float> **xout, InnerResult *res, const Carry *carry)
{
  const int nreg = res_regions(nch, G, S::TT);
  const int cap = res_capacity(ctx);
  if (nreg < 1 || cap < kNumXcd) return TH2D_ELAUNCH;
  for (int c = 0; c < nch; c++)
    if (!res_eligible(G[c], S::TT) ||
        res_chain_regions(G[c], S::TT) * (S::NT / 64) > 64 * RES_PART_LANE)
      return TH2D_ELAUNCH;
  std::vector<int> wg_tile;
  std::vector<ResGroup> groups;
  if (!place_regions(nch, G, cap / kNumXcd, S::TT, wg_tile, groups)) return TH2D_ELAUNCH;
  long sites = 0;
  for (const ResGroup &g : groups) {
    if (g.grid < 1 || g.grid > cap) return TH2D_ELAUNCH;
    sites += g.sites;
  }
  if (groups.size() > 1 && 8.5 * (double)groups.size() > 10.0 + (double)sites * 120.0 / 4.5e6)
    return TH2D_ELAUNCH;  // (us per iteration) the step kernel is faster
  // control block: the workgroup table, per chain an iteration counter and a
  // set-up counter (a cache line each), a spare line, the XCD table, the
  // partials and the publication planes (the error flags follow the results)
  int nparts = 0;
  for (int c = 0; c < nch; c++)
    nparts += ((int)res_chain_regions(G[c], S::TT) + 7) & ~7;
  const size_t ntab = (wg_tile.size() + 31) & ~(size_t)31;
  const size_t nzero = 64 * (size_t)nch + 32;  // counters and error flags, zeroed
  const size_t nxcc = ((size_t)nreg + 31) & ~(size_t)31;
  const size_t pub_bytes = 4 * (size_t)nreg * RES_PUB * 16;  // 2 x ((r, s), w), 16 B a slot
  constexpr size_t NW = S::NT / 64;  // partial slots of a region: one per wave
  const size_t extra_bytes = 4 * (ntab + nzero + nxcc) + 64 * NW * (size_t)nparts + pub_bytes;
  CgBatch<float> a;
  std::vector<CgChain> ch;
  void *extra = nullptr;
  TH2D_TRY(build_batch<float>(ctx, buf, nch, G, rhs, K, stop_rr, max_iter, S::TT, S::TX,
                              extra_bytes, a, ch, &extra, carry, wg_tile.data(),
                              sizeof(int) * wg_tile.size()));
  if (a.ntiles != nreg) return TH2D_EARG;
  ResCtl ctl;
  ctl.wg_tile = (const int *)extra;
  ctl.bar = (unsigned *)extra + ntab;
  ctl.bar0 = ctl.bar + 32 * nch;
  // the two flags follow the results, so that one copy brings both back
  ctl.err = (int *)((char *)a.res + sizeof(CgResult) * nch);
  ctl.misplaced = ctl.err + 1;
  ctl.xcc = (int *)(ctl.bar0 + 32 * nch) + 32;
  ctl.part = (double *)(ctl.xcc + nxcc);  // 128-B aligned: every count above is a multiple of 32
  ctl.pub = (char *)ctl.part + 64 * NW * (size_t)nparts;  // 512-B aligned: nparts is a multiple of 8
  ctl.nparts = nparts;
  const size_t blk = (size_t)((char *)extra + extra_bytes - (char *)a.x);
  if (blk >= (size_t)1 << 31) return TH2D_ELAUNCH;  // buffer offsets are 32-bit
  ctl.blk_bytes = (unsigned)blk;
  ctl.spin = ctx->res_spin_limit < 0 ? RES_SPIN_LIMIT : (unsigned)ctx->res_spin_limit;
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  hipLaunchKernelGGL(k_zero_u32, dim3(std::min<int>(64, (int)(nzero + 255) / 256)), dim3(256), 0,
                     ctx->stream, ctl.bar, (int)nzero, (unsigned *)ctl.err, 2);
  TH2D_HIP(hipGetLastError());
  bool mu0 = true, clip = false;
  for (int c = 0; c < nch; c++) {
    mu0 &= G[c].ep == 1.0 && G[c].em == 1.0;
    clip |= G[c].nt % S::TT != 0 || G[c].nx % S::TX != 0;
  }
  const void *kern = mu0 ? (clip ? (const void *)k_cg_resident<float, true, S, true>
                                 : (const void *)k_cg_resident<float, true, S, false>)
                         : (clip ? (const void *)k_cg_resident<float, false, S, true>
                                 : (const void *)k_cg_resident<float, false, S, false>);
  const int *tab0 = ctl.wg_tile;
  for (const ResGroup &g : groups) {  // one launch per group (each chain in one group)
    ctl.wg_tile = tab0 + g.tab;
    void *args[] = {(void *)&a, (void *)&ctl};
    TH2D_HIP(hipLaunchCooperativeKernel(kern, dim3(g.grid), dim3(S::NT), args, 0, ctx->stream));
  }
  const size_t rb = sizeof(CgResult) * nch;
  if (ctx->cg_defer == 1) {
    // left in flight: resident_collect copies the results, the flags and the
    // caller's |r|^2 (written behind them) back after the caller's work
    if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k1, ctx->stream));
    ctx->res_pend_dev = a.res;
    for (int c = 0; c < nch; c++) {
      xout[c] = a.x + ch[c].voff;
      res[c] = InnerResult{0, 0.0, TH2D_OK, 0.0, 0.0, 1};
    }
    ctx->cg_defer = 2;
    ctx->res_pend_nch = nch;
    ctx->res_pend_tt = S::TT;
    ctx->res_pend_tx = S::TX;
    ctx->res_pend_launches = (long)groups.size();
    return TH2D_OK;
  }
  // the results and the two flags after them into one pinned block, then one wait
  TH2D_TRY(th2d_host_reserve(ctx->redh, rb + 64));
  TH2D_HIP(hipMemcpyAsync(ctx->redh.ptr, a.res, rb + 2 * sizeof(int), hipMemcpyDeviceToHost,
                          ctx->stream));
  TH2D_HIP(hipStreamSynchronize(ctx->stream));
  ctx->hstage_off = 0;  // the stream is idle: every staged copy has been read
  std::vector<CgResult> h(nch);
  memcpy(h.data(), ctx->redh.ptr, rb);
  int flags[2] = {0, 0};
  memcpy(flags, (char *)ctx->redh.ptr + rb, sizeof(flags));
  if (flags[1]) return TH2D_ELAUNCH;  // placement not as dispatched: the step kernel solves
  if (flags[0]) return TH2D_ESYNC;
  ctx->cg_last_tile[0] = S::TT;
  ctx->cg_last_tile[1] = S::TX;
  ctx->cg_last_resident = 1;
  TH2D_TRY(finish_solve<float>(ctx, nch, G, a, ch, h, (long)groups.size(), xout, res));
  for (int c = 0; c < nch; c++) res[c].resident = 1;
  return TH2D_OK;
}

// The resident solve with regions of 32 rows when every chain of the batch
// can be cut so (sides multiples of 32 x 64, or one region) and twice the
// regions still fit the device in one launch; of 64 rows otherwise.
// th2d_set_cg_resident_rows forces either.
int resident_solve(th2d_ctx *ctx, th2d_buf &buf, int nch, const Geo *G,
                   const c2<float> *const *rhs, const c2<float> *const *K, const double *stop_rr,
                   const int *max_iter, c2<float> **xout, InnerResult *res, const Carry *carry)
{
  bool half = ctx->res_rows == RShapeH::TT;
  if (ctx->res_rows == 0) {
    const int cap = res_capacity(ctx);
    half = res_regions(nch, G, RShapeH::TT) <= cap;
    for (int c = 0; c < nch && half; c++)
      half = res_eligible(G[c], RShapeH::TT) && res_eligible(G[c], RShape::TT) &&
             res_chain_regions(G[c], RShapeH::TT) <= cap / kNumXcd;
  }
  if (half)
    return resident_solve_s<RShapeH>(ctx, buf, nch, G, rhs, K, stop_rr, max_iter, xout, res,
                                     carry);
  return resident_solve_s<RShape>(ctx, buf, nch, G, rhs, K, stop_rr, max_iter, xout, res, carry);
}

// Where the caller of a resident solve left in flight writes |r|^2 of its
// chain k (slot k), so that resident_collect brings it back with the results.
double *resident_rr_slots(th2d_ctx *ctx)
{
  if (ctx->cg_defer != 2) return nullptr;
  return (double *)((char *)ctx->res_pend_dev + sizeof(CgResult) * ctx->res_pend_nch + 64);
}

// Results of the resident solve left in flight (cg_defer 2), after the
// caller's work on the stream: one copy back and one wait.  TH2D_ESYNC
// after a wait that timed out, TH2D_ELAUNCH after a misplacement (x is 0
// then), or the per-chain results in `res`; `rr` (if set) receives the
// caller's |r|^2 slots either way.
int resident_collect(th2d_ctx *ctx, int nch, const Geo *G, InnerResult *res, double *rr)
{
  if (ctx->cg_defer != 2 || ctx->res_pend_nch != nch) return TH2D_EARG;
  ctx->cg_defer = 0;
  const size_t rb = sizeof(CgResult) * nch, bytes = rb + 64 + sizeof(double) * nch;
  TH2D_TRY(th2d_host_reserve(ctx->resh, bytes));
  TH2D_HIP(hipMemcpyAsync(ctx->resh.ptr, ctx->res_pend_dev, bytes, hipMemcpyDeviceToHost,
                          ctx->stream));
  // The host spins on an event instead of a blocking stream wait: the next
  // pass is launched sooner (a failed device reports its error here too).
  TH2D_HIP(hipEventRecord(ctx->sync_ev, ctx->stream));
  {
    hipError_t q;
    while ((q = hipEventQuery(ctx->sync_ev)) == hipErrorNotReady) {
    }
    TH2D_HIP(q);
  }
  ctx->hstage_off = 0;  // the stream is idle: every staged copy has been read
  if (rr) memcpy(rr, (char *)ctx->resh.ptr + rb + 64, sizeof(double) * nch);
  std::vector<CgResult> h(nch);
  memcpy(h.data(), ctx->resh.ptr, rb);
  int flags[2] = {0, 0};
  memcpy(flags, (char *)ctx->resh.ptr + rb, sizeof(flags));
  if (flags[1]) return TH2D_ELAUNCH;
  if (flags[0]) return TH2D_ESYNC;
  ctx->cg_last_tile[0] = ctx->res_pend_tt;
  ctx->cg_last_tile[1] = ctx->res_pend_tx;
  ctx->cg_last_resident = 1;
  if (ctx->ktiming) {
    float ms = 0.f;
    TH2D_HIP(hipEventElapsedTime(&ms, ctx->k0, ctx->k1));
    ctx->ktime_ms += ms;
    ctx->klaunches += ctx->res_pend_launches;
    for (int c = 0; c < nch; c++) {  // algorithmic bytes as finish_solve counts them
      ctx->kbytes += (double)h[c].iters * 2.0 * G[c].vh * 15.0 * sizeof(c2<float>);
      ctx->kiters += h[c].iters;
    }
  }
  for (int c = 0; c < nch; c++)
    res[c] = InnerResult{h[c].iters, h[c].rr, h[c].status, h[c].gam_old, h[c].alpha_old, 1};
  return TH2D_OK;
}

// Whether the resident solver takes a lattice: eligible, and no more
// regions than one XCD holds.
bool res_takes(const Geo &g, int cap)
{
  return res_eligible(g) && res_chain_regions(g, RShape::TT) <= cap / kNumXcd;
}

template <typename T>
int inner_solve(th2d_ctx *ctx, th2d_buf &buf, int pass, int nch, const Geo *G,
                const c2<T> *const *rhs, const c2<T> *const *K, const double *stop_rr,
                const int *max_iter, c2<T> **xout, InnerResult *res, int force_shape,
                const Carry *carry);

// A batch split between the kernels: the chains `ra` in the resident
// kernel (block in `buf`), the chains `rb` in the step kernel (block in the
// context's second FP32 buffer), one after the other on the stream.
// TH2D_ELAUNCH when the resident kernel declines its part.
int split_solve(th2d_ctx *ctx, th2d_buf &buf, int pass, const Geo *G,
                const c2<float> *const *rhs, const c2<float> *const *K, const double *stop_rr,
                const int *max_iter, c2<float> **xout, InnerResult *res, const Carry *carry,
                const std::vector<int> &ra, const std::vector<int> &rb)
{
  for (int part = 0; part < 2; part++) {
    const std::vector<int> &ix = part == 0 ? ra : rb;
    const int n = (int)ix.size();
    std::vector<Geo> g(n);
    std::vector<const c2<float> *> r(n), k(n);
    std::vector<double> st(n);
    std::vector<int> mi(n);
    std::vector<c2<float> *> xo(n);
    std::vector<InnerResult> ir(n);
    std::vector<Carry> cy(n);
    for (int q = 0; q < n; q++) {
      const int c = ix[q];
      g[q] = G[c];
      r[q] = rhs[c];
      k[q] = K[c];
      st[q] = stop_rr[c];
      mi[q] = max_iter[c];
      if (carry) cy[q] = carry[c];
    }
    if (part == 0) {
      const int rc = resident_solve(ctx, buf, n, g.data(), r.data(), k.data(), st.data(), mi.data(),
                                    xo.data(), ir.data(), carry ? cy.data() : nullptr);
      if (rc != TH2D_OK) return rc;
    } else {
      TH2D_TRY(inner_solve<float>(ctx, ctx->cg32b, pass, n, g.data(), r.data(), k.data(), st.data(),
                                  mi.data(), xo.data(), ir.data(), -2, nullptr));
    }
    for (int q = 0; q < n; q++) {
      xout[ix[q]] = xo[q];
      res[ix[q]] = ir[q];
    }
  }
  ctx->cg_last_resident = 2;
  return TH2D_OK;
}

template <typename T>
int inner_solve(th2d_ctx *ctx, th2d_buf &buf, int pass, int nch, const Geo *G,
                const c2<T> *const *rhs,
                const c2<T> *const *K, const double *stop_rr, const int *max_iter, c2<T> **xout,
                InnerResult *res, int force_shape, const Carry *carry)
{
  if constexpr (sizeof(T) == 4) {
    if (force_shape == -1 && ctx->cg_resident && !(ctx->tile_t > 0 && ctx->tile_x > 0)) {
      // the chains the resident kernel takes; a batch with others is split
      const int cap = res_capacity(ctx);
      std::vector<int> ra, rb;
      for (int c = 0; c < nch; c++) (res_takes(G[c], cap) ? ra : rb).push_back(c);
      int rc = TH2D_ELAUNCH;
      if (rb.empty()) {
        rc = resident_solve(ctx, buf, nch, G, rhs, K, stop_rr, max_iter, xout, res, carry);
      } else if (!ra.empty()) {
        const int defer = ctx->cg_defer;  // a split batch is solved in place
        ctx->cg_defer = 0;
        rc = split_solve(ctx, buf, pass, G, rhs, K, stop_rr, max_iter, xout, res, carry, ra, rb);
        ctx->cg_defer = defer;
      }
      // A wait that timed out (TH2D_ESYNC) left every workgroup exiting
      // normally and the right-hand sides untouched: the step kernel
      // re-solves the batch, as it does for a batch the resident kernel
      // declines (TH2D_ELAUNCH).
      if (rc == TH2D_ESYNC) ctx->cg_sync_fallbacks++;
      if (rc != TH2D_ELAUNCH && rc != TH2D_ESYNC) return rc;
    }
  }
  ctx->cg_last_resident = 0;
  int shape = force_shape;
  if (shape < 0 && ctx->cg_autotune && !(ctx->tile_t > 0 && ctx->tile_x > 0)) {
    auto &cache = ctx->cg_tuned[sizeof(T) == 8 ? 1 : 0];
    const unsigned long long key = tune_key(nch, G);
    auto it = cache.find(key);
    if (it != cache.end()) {
      shape = it->second;
    } else {
      int s = autotune_shape<T>(ctx, buf, nch, G, rhs, K, 12);
      if (s < 0) return s;
      cache[key] = s;
      shape = s;
    }
    for (int c = 0; c < nch; c++)  // a shape larger than a lattice of this subset
      if (kShapes[shape].tt > G[c].nt || kShapes[shape].tx > G[c].nx) shape = -1;
  }
  if (shape < 0) shape = choose_shape(ctx, nch, G);
  if (shape < 0) return TH2D_EARG;
  const ShapeEntry tc = kShapes[shape];
  const int nthreads = tc.nt;
  ctx->cg_last_tile[0] = tc.tt;
  ctx->cg_last_tile[1] = tc.tx;
  CgBatch<T> a;
  std::vector<CgChain> ch;
  void *extra = nullptr;
  TH2D_TRY(build_batch<T>(ctx, buf, nch, G, rhs, K, stop_rr, max_iter, tc.tt, tc.tx, 0, a, ch,
                          &extra));
  const int ntiles = a.ntiles;
  const size_t lds = lds_bytes<T>(tc.tt, tc.tx);
  // one block per tile unless capped; a capped grid is a multiple of 8 so
  // that a block keeps its XCD over the grid-stride loop
  int nblocks = ntiles;
  if (ctx->cg_max_blocks > 0 && ctx->cg_max_blocks < ntiles)
    nblocks = std::max(8, ctx->cg_max_blocks & ~7);
  if (ctx->ktiming) TH2D_HIP(hipEventRecord(ctx->k0, ctx->stream));
  hipLaunchKernelGGL(k_cg_init<T>, dim3(nblocks), dim3(nthreads), lds, ctx->stream, a);
  TH2D_HIP(hipGetLastError());
  std::vector<CgResult> h(nch);
  long launched = 0;
  // launches per stop-flag read: the iteration count this restart pass took
  // in the previous solve (solves of one trajectory are alike), then 8
  int &last = ctx->cg_pass_iters[sizeof(T) == 8 ? 1 : 0][std::min(pass, 7)];
  int chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : (last > 0 ? last + 1 : 32);
  int most = 0;
  for (int c = 0; c < nch; c++) most = std::max(most, max_iter[c]);
  hipGraphExec_t gx = nullptr;  // the captured run of step launches
  if (ctx->cg_graph) TH2D_TRY(step_graph<T>(ctx, shape, nblocks, a, &gx));
  for (;;) {
    if (gx) {  // whole runs (launched stays even; surplus launches find every chain done)
      const int reps = (chunk + kStepGraph - 1) / kStepGraph;
      for (int r = 0; r < reps; r++) TH2D_HIP(hipGraphLaunch(gx, ctx->stream));
      chunk = reps * kStepGraph;
    } else {
      for (int k = 0; k < chunk; k++)
        launch_step(shape, nblocks, ctx->stream, a, (int)((launched + k) & 1));
    }
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
    chunk = ctx->cg_chunk > 0 ? ctx->cg_chunk : 8;
  }
  int itmax = 0;
  for (int c = 0; c < nch; c++) itmax = std::max(itmax, h[c].iters);
  last = itmax;
  return finish_solve<T>(ctx, nch, G, a, ch, h, launched + 1, xout, res);
}




}  // namespace

extern "C" int th2d_set_cg_graph(th2d_ctx *ctx, int on)
{
  if (!ctx) return TH2D_EARG;
  ctx->cg_graph = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_tile(th2d_ctx *ctx, int tt, int tx)
{
  if (!ctx || tt < 0 || tx < 0) return TH2D_EARG;
  if ((tt || tx) && shape_index(tt, tx) < 0) return TH2D_EARG;
  ctx->tile_t = tt;
  ctx->tile_x = tx;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_autotune(th2d_ctx *ctx, int on)
{
  if (!ctx) return TH2D_EARG;
  ctx->cg_autotune = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_reliable(th2d_ctx *ctx, int on)
{
  if (!ctx) return TH2D_EARG;
  ctx->cg_reliable = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_resident(th2d_ctx *ctx, int on)
{
  if (!ctx) return TH2D_EARG;
  ctx->cg_resident = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_cg_last_resident(th2d_ctx *ctx, int *on)
{
  if (!ctx || !on) return TH2D_EARG;
  *on = ctx->cg_last_resident;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_small(th2d_ctx *ctx, int on)
{
  if (!ctx) return TH2D_EARG;
  ctx->cg_small = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_resident_rows(th2d_ctx *ctx, int rows)
{
  if (!ctx || (rows != 0 && rows != RShape::TT && rows != RShapeH::TT)) return TH2D_EARG;
  ctx->res_rows = rows;
  return TH2D_OK;
}

extern "C" int th2d_set_cg_spin_limit(th2d_ctx *ctx, long polls)
{
  if (!ctx || polls > (long)0xFFFFFFFFu) return TH2D_EARG;
  ctx->res_spin_limit = polls < 0 ? -1 : polls;
  return TH2D_OK;
}

extern "C" int th2d_cg_sync_fallbacks(th2d_ctx *ctx, long *n)
{
  if (!ctx || !n) return TH2D_EARG;
  *n = ctx->cg_sync_fallbacks;
  return TH2D_OK;
}

extern "C" int th2d_cg_last_tile(th2d_ctx *ctx, int *tt, int *tx)
{
  if (!ctx || !tt || !tx) return TH2D_EARG;
  *tt = ctx->cg_last_tile[0];
  *tx = ctx->cg_last_tile[1];
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
  if (!mixed && ctx->cg_small && small_fits(nch, G.data()) && !(ctx->tile_t > 0 && ctx->tile_x > 0))
    return small_full_solve(ctx, nch, G.data(), x, b, K, p, infos);
  // FP64 work: residuals r (all chains), tmp, per-chain scalars
  TH2D_TRY(th2d_buf_reserve(ctx->misc, sizeof(double) * 2 * 2 * Vt + sizeof(double) * (nch + 64)));
  double *R = (double *)ctx->misc.ptr;
  double *TMP = R + 2 * Vt;
  double *dsc = TMP + 2 * Vt;
  // FP32 right-hand sides, and p, s carried between passes (reliable updates)
  TH2D_TRY(th2d_buf_reserve(ctx->cgmix, sizeof(float) * 2 * 3 * Vt + 256));
  float2 *R32 = (float2 *)ctx->cgmix.ptr;
  float2 *P32 = R32 + Vt, *S32 = P32 + Vt;
  std::vector<char> carry_ok(nch, 0);
  std::vector<double> cgam(nch, 0.0), calpha(nch, 0.0), nprev(nch, 0.0);

  std::vector<th2d_cg_info> inf(nch);
  std::vector<double> bb(nch), rr(nch), target(nch);
  std::vector<int> left(nch, p->max_iter), active(nch, 1), status(nch, TH2D_OK);
  const bool abs_mode = p->tol_rel <= 0;
  const double delta = p->delta > 0 ? p->delta : 1e-4;
  // per-chain descriptor of a batched site operation (batch.hip)
  auto op_of = [&](int c) {
    BOp o;
    memset(&o, 0, sizeof(o));
    o.G = G[c];
    o.bc = Bc{geoms[c].anti_t, geoms[c].anti_x};
    return o;
  };
  {
    // |b|^2 and x = 0 for every chain: two batched launches
    std::vector<BOp> t0(nch), t1(nch);
    for (int c = 0; c < nch; c++) {
      memset(&inf[c], 0, sizeof(th2d_cg_info));
      t0[c] = op_of(c);
      t0[c].d0 = (double *)b[c];
      t0[c].d1 = (double *)b[c];
      t0[c].s1 = 1.0;
      t0[c].out = dsc + c;
      t1[c] = op_of(c);  // x = b - b
      t1[c].d0 = x[c];
      t1[c].d1 = (double *)b[c];
      t1[c].d2 = (double *)b[c];
    }
    const long nblk = th2d::bop_finish(t0);
    th2d::bop_finish(t1);
    std::vector<BOp> all(t0);
    all.insert(all.end(), t1.begin(), t1.end());
    TH2D_TRY(th2d::bop_upload(ctx, ctx->btab[0], all));
    const BOp *dt = (const BOp *)ctx->btab[0].ptr;
    TH2D_TRY(th2d::bop_launch(ctx, th2d::BOP_DOT, dt, nch, nblk));
    TH2D_TRY(th2d::bop_launch(ctx, th2d::BOP_SUB, dt + nch, nch, nblk));
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
  bool res_defer = true;  // resident inner solves left in flight (mixed)
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
    // right-hand sides of the inner solves: r / |r| (FP32 or FP64)
    std::vector<BOp> tin(na);
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      g2[k] = G[c];
      mi[k] = left[c];
      const double *rhs = outer == 0 ? b[c] : R + 2 * voff[c];
      const double nr = sqrt(rr[c]);
      tin[k] = op_of(c);
      tin[k].d1 = (double *)rhs;
      tin[k].s0 = 1.0 / nr;
      if (mixed) {
        // the scale from the chain's |r|^2 on the device: the table stays
        // the same from pass to pass (uploaded again only when it changes)
        tin[k].s0 = 0.0;
        tin[k].sp = dsc + c;
        tin[k].f0 = (float *)(R32 + voff[c]);
        stop[k] = std::max(delta * delta, target[c] / rr[c]);
      } else if (outer == 0) {
        stop[k] = target[c];  // unscaled first pass: the reference criterion
      } else {
        tin[k].d0 = TMP + 2 * voff[c];
        stop[k] = abs_mode ? target[c] / rr[c] : std::max(target[c] / rr[c], 1e-30);
      }
    }
    const long nblk_in = th2d::bop_finish(tin);
    if (mixed || outer > 0) {
      if (mixed) TH2D_TRY(th2d::bop_upload_key(ctx, 0, tin));
      else TH2D_TRY(th2d::bop_upload(ctx, ctx->btab[0], tin));
      TH2D_TRY(th2d::bop_launch(ctx, mixed ? th2d::BOP_TO_F32 : th2d::BOP_SCALE,
                                (const BOp *)ctx->btab[0].ptr, na, nblk_in));
    }
    // inner solve, then x += |r| y for every chain and the true residual
    // r = b - M^dag M x with |r|^2 in FP64: three batched launches
    std::vector<BOp> tx(na), tm(na), tres(na);
    std::vector<char> was_cont(na, 0);
    bool deferred = false;
    auto note_carry = [&](int n, const InnerResult *q) {  // (kept only if the pass stays consistent)
      for (int k = 0; k < n; k++) {
        const int c = idx[k];
        carry_ok[c] = q[k].resident && q[k].status != TH2D_EBREAKDOWN;
        cgam[c] = q[k].gam;
        calpha[c] = q[k].alpha;
      }
    };
    if (mixed) {
      std::vector<const float2 *> rh(na), kk(na);
      std::vector<float2 *> xo(na);
      // reliable update: a chain whose previous pass left its p, s continues
      // that Krylov sequence with r replaced by the true residual (rescaled
      // from the previous pass's normalisation to this one's)
      std::vector<Carry> cy(na);
      for (int k = 0; k < na; k++) {
        const int c = idx[k];
        rh[k] = R32 + voff[c];
        kk[k] = (const float2 *)K32[c];
        cy[k] = Carry{P32 + voff[c], S32 + voff[c], 0.0, 0.0, 0.0};
        if (outer > 0 && carry_ok[c] && ctx->cg_reliable && cgam[c] > 0.0 && calpha[c] > 0.0) {
          cy[k].cscale = nprev[c] / sqrt(rr[c]);
          cy[k].gam0 = cgam[c];
          cy[k].alpha0 = calpha[c];
          was_cont[k] = 1;
        }
        nprev[c] = sqrt(rr[c]);
      }
      // A resident solve may be left in flight (its x is then 0 if it
      // fails): the FP64 update and residual below are queued behind it,
      // and one wait reads the solver's results with |r|^2.
      ctx->cg_defer = res_defer ? 1 : 0;
      const int irc = inner_solve<float>(ctx, ctx->cg32, outer, na, g2.data(), rh.data(),
                                         kk.data(), stop.data(), mi.data(), xo.data(), ir.data(),
                                         -1, cy.data());
      deferred = ctx->cg_defer == 2;
      if (!deferred) ctx->cg_defer = 0;
      TH2D_TRY(irc);
      if (!deferred) note_carry(na, ir.data());
      for (int k = 0; k < na; k++) {
        tx[k] = op_of(idx[k]);
        tx[k].d0 = x[idx[k]];
        tx[k].f0 = (float *)xo[k];
        tx[k].sp = dsc + idx[k];  // |r| of the pass's right-hand side, before RESID replaces it
      }
    } else {
      std::vector<const double2 *> rh(na), kk(na);
      std::vector<double2 *> xo(na);
      for (int k = 0; k < na; k++) {
        int c = idx[k];
        rh[k] = (const double2 *)(outer == 0 ? b[c] : TMP + 2 * voff[c]);
        kk[k] = (const double2 *)K[c];
      }
      if (ctx->cg_small && small_fits(na, g2.data()) && !(ctx->tile_t > 0 && ctx->tile_x > 0))
        TH2D_TRY(small_solve(ctx, ctx->cg64, na, g2.data(), rh.data(), kk.data(), stop.data(),
                             mi.data(), xo.data(), ir.data()));
      else
        TH2D_TRY(inner_solve<double>(ctx, ctx->cg64, outer, na, g2.data(), rh.data(), kk.data(),
                                     stop.data(), mi.data(), xo.data(), ir.data()));
      for (int k = 0; k < na; k++) {
        tx[k] = op_of(idx[k]);
        tx[k].d0 = x[idx[k]];
        tx[k].d1 = (double *)xo[k];
        tx[k].s0 = outer > 0 ? sqrt(rr[idx[k]]) : 1.0;
      }
    }
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      double *r = R + 2 * voff[c], *tmp = TMP + 2 * voff[c];
      tm[k] = op_of(c);  // tmp = M x
      tm[k].d0 = tmp;
      tm[k].d1 = x[c];
      tm[k].d2 = (double *)K[c];
      tres[k] = op_of(c);  // r = b - M^dag tmp, |r|^2
      tres[k].d0 = r;
      tres[k].d1 = (double *)b[c];
      tres[k].d2 = (double *)K[c];
      tres[k].d3 = tmp;
      tres[k].s1 = 1.0;
      tres[k].out = dsc + c;
      if (deferred) tres[k].d4 = resident_rr_slots(ctx) + k;  // also behind the results
    }
    const long nblk = th2d::bop_finish(tx);
    for (auto *t : {&tm, &tres}) th2d::bop_finish(*t);
    std::vector<BOp> all(tx);
    for (auto *t : {&tm, &tres}) all.insert(all.end(), t->begin(), t->end());
    if (mixed) TH2D_TRY(th2d::bop_upload_key(ctx, 1, all));
    else TH2D_TRY(th2d::bop_upload(ctx, ctx->btab[1], all));
    const BOp *dt = (const BOp *)ctx->btab[1].ptr;
    TH2D_TRY(th2d::bop_launch(ctx, mixed ? th2d::BOP_ADD_F32 : th2d::BOP_AXPY, dt, na, nblk));
    TH2D_TRY(th2d::bop_launch(ctx, th2d::BOP_DSLASH, dt + na, na, nblk));
    TH2D_TRY(th2d::bop_launch(ctx, th2d::BOP_RESID, dt + 2 * na, na, nblk));
    std::vector<double> rn(na);
    int crc = TH2D_OK;
    if (deferred) {  // one copy back and one wait for the results and |r|^2
      crc = resident_collect(ctx, na, g2.data(), ir.data(), rn.data());
    } else {
      std::vector<double> rnc(nch);  // |r|^2 per chain (slot c)
      TH2D_TRY(th2d::fetch_doubles(ctx, rnc.data(), dsc, nch));
      for (int k = 0; k < na; k++) rn[k] = rnc[idx[k]];
    }
    if (deferred) {
      if (crc == TH2D_ESYNC || crc == TH2D_ELAUNCH) {
        // x took a zero correction: the pass is repeated, in place (with the
        // step kernel as the fallback), and counts nothing
        if (crc == TH2D_ESYNC) ctx->cg_sync_fallbacks++;
        res_defer = false;
        for (int k = 0; k < na; k++) carry_ok[idx[k]] = 0;
        continue;
      }
      TH2D_TRY(crc);
      note_carry(na, ir.data());
    }
    for (int k = 0; k < na; k++) {
      int c = idx[k];
      inf[c].iterations += ir[k].iters;
      left[c] -= ir[k].iters;
      inf[c].rr = ir[k].rr * (mixed || outer > 0 ? rr[c] : 1.0);
      if (outer > 0) inf[c].restarts++;
      double rnew = rn[k];
      bool breakdown = ir[k].status == TH2D_EBREAKDOWN;
      // the carried Krylov state is consistent with the true residual only
      // when the FP32 recursive residual has not drifted from it:
      // |r_true|^2 <= 1.1 |r_rec|^2 (drift below a third of the residual);
      // otherwise the next pass restarts
      if (mixed && !(rnew <= 1.1 * ir[k].rr * rr[c])) carry_ok[c] = 0;
      if (abs_mode && !mixed && ir[k].status == TH2D_OK) {  // reference stop met
        rr[c] = rnew;
        active[c] = 0;
        continue;
      }
      if (breakdown && was_cont[k]) {  // the continuation broke down: restart afresh
        rr[c] = std::min(rr[c], rnew);
        if (rr[c] <= target[c]) active[c] = 0;
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

// ---- one MD step of small lattices as a captured graph (hmc.hip) ----------
namespace th2d {
bool md_graph_fits(th2d_ctx *ctx, int nch, const Geo *G, const th2d_cg_params *p)
{
  return ctx->cg_small && p->precision == TH2D_PREC_F64 && small_fits(nch, G) &&
         !(ctx->tile_t > 0 && ctx->tile_x > 0);
}

// Descriptors of the step's solve (chi_c = (M^dag M)^-1 psi_c) and a zeroed
// per-chain accumulator, into ctx->cg64; the launch itself is enqueued by
// md_graph_solve (inside the capture).
int md_graph_prepare(th2d_ctx *ctx, int nch, const Geo *G, double *const *x,
                     const double *const *b, const double *const *K, const th2d_cg_params *p,
                     void **state)
{
  SmallFull *dch;
  SmallInfo *dinf;
  long maxV;
  TH2D_TRY(small_full_prepare(ctx, ctx->cg64, nch, G, x, b, K, p, &dch, &dinf, &maxV));
  hipLaunchKernelGGL(k_small_info_zero, dim3((nch + 63) / 64), dim3(64), 0, ctx->stream, dinf, nch);
  TH2D_HIP(hipGetLastError());
  state[0] = dch;
  state[1] = dinf;
  state[2] = (void *)maxV;
  return TH2D_OK;
}

void md_graph_solve(hipStream_t st, int nch, void *const *state)
{
  small_full_launch(st, nch, (long)state[2], (const SmallFull *)state[0], (SmallInfo *)state[1], 1);
}

// The accumulated solver statistics: iterations, solves, worst status and
// the largest true relative residual per chain.
int md_graph_stats(th2d_ctx *ctx, int nch, void *const *state, int *iters, int *solves,
                   int *status, double *max_true_rel)
{
  std::vector<SmallInfo> h(nch);
  TH2D_HIP(hipMemcpyAsync(h.data(), state[1], sizeof(SmallInfo) * nch, hipMemcpyDeviceToHost,
                          ctx->stream));
  TH2D_HIP(hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < nch; c++) {
    iters[c] = h[c].iters;
    solves[c] = h[c].pad;
    status[c] = h[c].status;
    max_true_rel[c] = h[c].true_rel;
  }
  ctx->cg_last_resident = 3;
  return TH2D_OK;
}
}  // namespace th2d
