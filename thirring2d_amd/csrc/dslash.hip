// dslash.hip -- the U(1) staggered fermion matrix M and its adjoint on
// gfx950 (fm_mul, hmc.c:123-184; the true adjoint that fm_conjugate_mul,
// hmc.c:188-249, stands for).
//
// With folded links K_mu(x) = 1/2 eta_mu(x) s_mu(x) e^{iA_mu(x)}:
//   (M psi)(x)   = m psi(x) + e^{mu} K_t(x) psi(x+t) - e^{-mu} K_t(x-t)* psi(x-t)
//                           +       K_x(x) psi(x+x) -         K_x(x-x)* psi(x-x)
//   (M^dag psi)(x) = m psi(x) - e^{-mu} K_t(x) psi(x+t) + e^{mu} K_t(x-t)* psi(x-t)
//                             -         K_x(x) psi(x+x) +        K_x(x-x)* psi(x-x)
// Two variants: a direct one (one thread per eo site; the t and x
// neighbours of a wavefront are contiguous runs of the other parity) and an
// LDS-tiled one that stages a (rows+2) x (cols+2) block of the input once
// and serves all four neighbour reads from LDS.  V = 2 G.vh < 2^31.
#include "th2d_internal.h"

template <typename T> struct HopCoef {
  c2<T> tu, td, xu, xd;
};

template <typename T, int DAG>
__device__ __forceinline__ HopCoef<T> hop_coefs(const Geo &G, c2<T> kt, c2<T> kx, c2<T> ktd,
                                                c2<T> kxd)
{
  using V2 = c2<T>;
  T ep = (T)G.ep, em = (T)G.em;
  HopCoef<T> h;
  if (!DAG) {
    h.tu = cscale(kt, ep);
    h.td = cmk<V2>(-em * ktd.x, em * ktd.y);  // -e^{-mu} conj(ktd)
    h.xu = kx;
    h.xd = cmk<V2>(-kxd.x, kxd.y);            // -conj(kxd)
  } else {
    h.tu = cscale(kt, -em);
    h.td = cmk<V2>(ep * ktd.x, -ep * ktd.y);  // e^{mu} conj(ktd)
    h.xu = cscale(kx, (T)-1);
    h.xd = cmk<V2>(kxd.x, -kxd.y);            // conj(kxd)
  }
  return h;
}

// ------------------------------------------------------------- direct ---
// One thread per eo site; a workgroup covers NT consecutive sites of one
// parity row.  Rows are ordered (t, parity) so that both parities of a t
// run back to back: an output row of parity p reads the rows t-1, t, t+1
// and the neighbour links of parity 1-p, which the other parity's rows of
// the same t read (or wrote) a moment earlier, so they are still in L2.  A
// parity-major order fetches the input and the links from HBM twice on
// lattices larger than the MALL.  Workgroups are dispatched round-robin
// over the 8 XCDs: with a multiple of 8 workgroups per row each XCD walks
// a column strip of all rows; otherwise each XCD gets a contiguous range
// of rows (xcd_block).
__device__ __forceinline__ int xcd_block(int v, int n)
{
  const int q = n >> 3, rem = n & 7, x = v & 7, k = v >> 3;
  return (x < rem ? x * (q + 1) : rem * (q + 1) + (x - rem) * q) + k;
}

template <typename T, int DAG, int NT>
__global__ void __launch_bounds__(NT) k_dslash_direct(Geo G, c2<T> *__restrict__ out,
                                                      const c2<T> *__restrict__ in,
                                                      const c2<T> *__restrict__ K, int nrhs,
                                                      int bpr)
{
  using V2 = c2<T>;
  const int V = 2 * G.vh;
  // with a multiple of 8 workgroups per row the dispatch order already
  // keeps a column strip of rows on one XCD
  const int lb = (bpr & 7) ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int row = lb / bpr, bx = lb - row * bpr;
  const int p = row & 1, t = row >> 1;
  const int j = bx * NT + threadIdx.x;
  if (j >= G.nxh) return;
  const int i = p * G.vh + t * G.nxh + j, q = (1 - p) * G.vh;
  const int x = 2 * j + ((t + p) & 1);
  const int tu = t + 1 == G.nt ? 0 : t + 1, td = t == 0 ? G.nt - 1 : t - 1;
  const int xu = x + 1 == G.nx ? 0 : x + 1, xd = x == 0 ? G.nx - 1 : x - 1;
  const int itu = q + tu * G.nxh + j, itd = q + td * G.nxh + j;
  const int ixu = q + t * G.nxh + (xu >> 1), ixd = q + t * G.nxh + (xd >> 1);
  HopCoef<T> c = hop_coefs<T, DAG>(G, K[i], K[V + i], K[itd], K[V + ixd]);
  const T m = (T)G.m;
  for (int r = 0; r < nrhs; r++) {
    const V2 *src = in + (long)r * V;
    V2 v = cscale(src[i], m);
    v = cadd(v, cmul(c.tu, src[itu]));
    v = cadd(v, cmul(c.td, src[itd]));
    v = cadd(v, cmul(c.xu, src[ixu]));
    v = cadd(v, cmul(c.xd, src[ixd]));
    out[(long)r * V + i] = v;
  }
}

// ---------------------------------------------------------- LDS tiled ---
// Block = ROWS x COLS output sites in natural coordinates (COLS even);
// LDS holds psi on (ROWS+2) x (COLS+2) and the links on ROWS+1 x COLS+1.
template <typename T, int DAG, int ROWS, int COLS>
__global__ void __launch_bounds__(ROWS *COLS) k_dslash_lds(Geo G, c2<T> *__restrict__ out,
                                                          const c2<T> *__restrict__ in,
                                                          const c2<T> *__restrict__ K, int nrhs)
{
  using V2 = c2<T>;
  constexpr int LW = COLS + 2, LH = ROWS + 2;
  __shared__ V2 sp[LH * LW];
  __shared__ V2 skt[(ROWS + 1) * (COLS + 1)];  // K_t on rows t0-1.., cols x0..
  __shared__ V2 skx[(ROWS + 1) * (COLS + 1)];  // K_x on rows t0.., cols x0-1..
  const int V = 2 * G.vh;
  const int tiles_x = (G.nx + COLS - 1) / COLS;
  const int t0 = (blockIdx.x / tiles_x) * ROWS, x0 = (blockIdx.x % tiles_x) * COLS;
  const int nthr = ROWS * COLS;
  // links: K_t at (t0-1+a, x0+b) for a in [0,ROWS], b in [0,COLS); K_x at
  // (t0+a, x0-1+b) for a in [0,ROWS), b in [0,COLS]
  for (int e = threadIdx.x; e < (ROWS + 1) * COLS; e += nthr) {
    int a = e / COLS, b = e - a * COLS;
    int t = t0 - 1 + a;
    t = t < 0 ? t + G.nt : (t >= G.nt ? t - G.nt : t);
    int x = x0 + b;
    x = x >= G.nx ? x - G.nx : x;
    skt[a * (COLS + 1) + b] = K[eo_index(G, t, x)];
  }
  for (int e = threadIdx.x; e < ROWS * (COLS + 1); e += nthr) {
    int a = e / (COLS + 1), b = e - a * (COLS + 1);
    int t = t0 + a;
    t = t >= G.nt ? t - G.nt : t;
    int x = x0 - 1 + b;
    x = x < 0 ? x + G.nx : (x >= G.nx ? x - G.nx : x);
    skx[a * (COLS + 1) + b] = K[V + eo_index(G, t, x)];
  }
  const int lt = threadIdx.x / COLS, lx = threadIdx.x - lt * COLS;
  const int tg = t0 + lt, xg = x0 + lx;
  const bool own = tg < G.nt && xg < G.nx;
  T m = (T)G.m;
  for (int r = 0; r < nrhs; r++) {
    const V2 *src = in + (long)r * V;
    __syncthreads();
    for (int e = threadIdx.x; e < LH * LW; e += nthr) {
      int a = e / LW, b = e - a * LW;
      int t = t0 - 1 + a, x = x0 - 1 + b;
      t = t < 0 ? t + G.nt : (t >= G.nt ? t - G.nt : t);
      x = x < 0 ? x + G.nx : (x >= G.nx ? x - G.nx : x);
      sp[e] = src[eo_index(G, t, x)];
    }
    __syncthreads();
    if (own) {
      int c = (lt + 1) * LW + (lx + 1);
      HopCoef<T> h = hop_coefs<T, DAG>(G, skt[(lt + 1) * (COLS + 1) + lx],
                                       skx[lt * (COLS + 1) + lx + 1], skt[lt * (COLS + 1) + lx],
                                       skx[lt * (COLS + 1) + lx]);
      V2 v = cscale(sp[c], m);
      v = cadd(v, cmul(h.tu, sp[c + LW]));
      v = cadd(v, cmul(h.td, sp[c - LW]));
      v = cadd(v, cmul(h.xu, sp[c + 1]));
      v = cadd(v, cmul(h.xd, sp[c - 1]));
      out[(long)r * V + eo_index(G, tg, xg)] = v;
    }
  }
}

// Variant selector: 0 = direct, 1 = LDS tiled (4 x 64).  The default is
// picked by the tuner (tune_dslash) and can be forced with
// TH2D_DSLASH_VARIANT.
static int g_dslash_variant = -1;

static int dslash_variant()
{
  if (g_dslash_variant < 0) {
    const char *e = getenv("TH2D_DSLASH_VARIANT");
    g_dslash_variant = e ? atoi(e) : 0;
  }
  return g_dslash_variant;
}

extern "C" int th2d_set_dslash_variant(int v)
{
  if (v < 0 || v > 1) return TH2D_EARG;
  g_dslash_variant = v;
  return TH2D_OK;
}

namespace th2d {

template <typename T, int DAG>
static void launch_t(th2d_ctx *c, const Geo &G, void *out, const void *in, const void *K, int nrhs)
{
  using V2 = c2<T>;
  if (dslash_variant() == 1) {
    constexpr int R = 4, CC = 64;
    int tiles = ((G.nt + R - 1) / R) * ((G.nx + CC - 1) / CC);
    hipLaunchKernelGGL((k_dslash_lds<T, DAG, R, CC>), dim3(tiles), dim3(R * CC), 0, c->stream, G,
                       (V2 *)out, (const V2 *)in, (const V2 *)K, nrhs);
  } else {
    // FP64: 256 threads per workgroup, FP32: 128 (measured at 2048^2 and
    // 4096^2); narrower rows take the smallest multiple of 64 that covers them
    const int nt_max = sizeof(T) == 8 ? 256 : 128;
    const int nt = G.nxh >= nt_max ? nt_max : (G.nxh > 64 ? 128 : 64);
    const int bpr = (G.nxh + nt - 1) / nt;
    const dim3 grid((unsigned)((long)bpr * 2 * G.nt));
    if (nt == 256)
      hipLaunchKernelGGL((k_dslash_direct<T, DAG, 256>), grid, dim3(256), 0, c->stream, G,
                         (V2 *)out, (const V2 *)in, (const V2 *)K, nrhs, bpr);
    else if (nt == 128)
      hipLaunchKernelGGL((k_dslash_direct<T, DAG, 128>), grid, dim3(128), 0, c->stream, G,
                         (V2 *)out, (const V2 *)in, (const V2 *)K, nrhs, bpr);
    else
      hipLaunchKernelGGL((k_dslash_direct<T, DAG, 64>), grid, dim3(64), 0, c->stream, G,
                         (V2 *)out, (const V2 *)in, (const V2 *)K, nrhs, bpr);
  }
}

int dslash_launch(th2d_ctx *c, const Geo &G, void *out, const void *in, const void *K, int dagger,
                  int prec, int nrhs)
{
  if (prec == TH2D_PREC_F32) {
    if (dagger) launch_t<float, 1>(c, G, out, in, K, nrhs);
    else launch_t<float, 0>(c, G, out, in, K, nrhs);
  } else {
    if (dagger) launch_t<double, 1>(c, G, out, in, K, nrhs);
    else launch_t<double, 0>(c, G, out, in, K, nrhs);
  }
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

}  // namespace th2d

extern "C" int th2d_dslash(th2d_ctx *c, const th2d_geom *g, void *out, const void *in,
                           const void *K, int dagger, int precision, int nrhs)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !out || !in || !K || nrhs < 1 || out == in) return TH2D_EARG;
  if (precision != TH2D_PREC_F64 && precision != TH2D_PREC_F32) return TH2D_EARG;
  Geo G = make_geo(g);
  return th2d::dslash_launch(c, G, out, in, K, dagger ? 1 : 0, precision, nrhs);
}

extern "C" int th2d_mdagm(th2d_ctx *c, const th2d_geom *g, double *out, const double *in,
                          const double *K, double *tmp)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !out || !in || !K || !tmp) return TH2D_EARG;
  Geo G = make_geo(g);
  TH2D_TRY(th2d::dslash_launch(c, G, tmp, in, K, 0, TH2D_PREC_F64, 1));
  return th2d::dslash_launch(c, G, out, tmp, K, 1, TH2D_PREC_F64, 1);
}
