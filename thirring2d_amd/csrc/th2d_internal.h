// th2d_internal.h -- shared definitions of the gfx950 kernels and the host
// runtime of libth2d.  Not installed; the public ABI is include/th2d.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/th2d.h"

// ---------------------------------------------------------------- errors --
#define TH2D_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "th2d: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
              __FILE__, __LINE__);                                              \
      return TH2D_EHIP;                                                         \
    }                                                                           \
  } while (0)

#define TH2D_TRY(call)          \
  do {                          \
    int rc_ = (call);           \
    if (rc_ != TH2D_OK) return rc_; \
  } while (0)

// -------------------------------------------------------------- context --
// Scratch buffers grow on demand and are reused; the hot path allocates
// nothing once a lattice size has been seen.
struct th2d_buf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct th2d_ctx {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cu = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // per-kernel timing of the solver kernel
  int ktiming = 0;
  hipEvent_t k0 = nullptr, k1 = nullptr;
  hipEvent_t a0 = nullptr, a1 = nullptr;  // autotuner
  hipEvent_t sync_ev = nullptr;           // polled by a host wait that spins (no timing)
  double ktime_ms = 0.0;
  long klaunches = 0;
  double kbytes = 0.0;
  // scratch
  th2d_buf cg64, cg32, cg32b, cgmix, red, redh, hmc, force, bag, misc;  // cg32b: step-kernel
                                                                       // part of a split batch
  th2d_buf bpart, btab[4];  // batched site operations: partials, descriptor tables
  std::vector<unsigned char> btab_key[4];  // bytes last uploaded to btab[i] by bop_upload_key
  th2d_buf hstage;          // pinned staging ring of small host-to-device copies
  size_t hstage_off = 0;    // next free byte of the ring
  th2d_buf dsl_tab;                    // th2d_dslash_batched: its descriptor table,
  std::vector<unsigned char> dsl_key;  // uploaded again only when it changes
  // tile configuration of the persistent solver (0 = autotuned default)
  int tile_t = 0, tile_x = 0;
  int cg_max_blocks = 0;  // 0 = one workgroup per tile
  int cg_chunk = 0;       // step launches per stop-flag check (0 = from history)
  long kiters = 0;
  int cg_pass_iters[2][8] = {};   // iterations per restart pass (FP32, FP64)
  int cg_last_tile[2] = {0, 0};   // tile shape of the last inner solve
  int cg_autotune = 0;            // measure tile shapes per batch signature
  int cg_resident = 1;            // FP32 inner solves in the resident kernel when they fit
  int cg_reliable = 0;            // mixed solves continue the Krylov sequence across passes
  int res_capacity = -1;          // resident workgroups the device holds (-1: not queried)
  int cg_last_resident = 0;       // the last inner solve: 0 step kernel, 1 resident, 2 both
  // look-ahead LU (dense.hip): panel and update streams on disjoint CUs
  int lu_stream_state = 0;  // 0 not tried, 1 made, -1 unavailable
  hipStream_t lu_ps = nullptr, lu_us = nullptr;
  hipEvent_t lu_ev[3] = {nullptr, nullptr, nullptr};
  int hmc_graph = 1;              // small-lattice FP64 MD steps as graph replays
  int cg_small = 1;               // FP64 solves of lattices up to 2048 sites in one launch
  long res_spin_limit = -1;       // polls of one resident-kernel wait (-1: the default)
  int res_rows = 0;               // resident regions of 64 or 32 rows (0: chosen per batch)
  long cg_sync_fallbacks = 0;     // resident solves that timed out and were re-solved
  // A resident inner solve left in flight for the mixed solver's outer loop
  // (cg.hip): 1 asks for it, 2 says one was launched; its results and flags
  // are copied into `resh` behind it.
  int cg_defer = 0;
  int res_pend_nch = 0, res_pend_tt = 0, res_pend_tx = 0;
  long res_pend_launches = 0;
  void *res_pend_dev = nullptr;  // its result block (device)
  th2d_buf resh;
  std::map<unsigned long long, int> cg_tuned[2];  // signature -> shape (FP32, FP64)
  // step-kernel launches captured as graphs (cg.hip), keyed by the launch
  // arguments; cg_graph 0 (default) launches eagerly
  int cg_graph = 0;
  hipStream_t cap_stream = nullptr;  // private stream the graphs are captured on
  struct CgGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec;
  };
  std::vector<CgGraph> cg_graphs;
};

int th2d_buf_reserve(th2d_buf &b, size_t bytes);
int th2d_host_reserve(th2d_buf &b, size_t bytes);

// ------------------------------------------------------------- geometry --
static inline int th2d_check_geom(const th2d_geom *g)
{
  if (!g || g->nt < 2 || g->nx < 2 || (g->nt & 1) || (g->nx & 1)) return TH2D_EARG;
  return TH2D_OK;
}

// Device view of the lattice.
struct Geo {
  int nt, nx, nxh, vh;  // nxh = nx/2, vh = V/2
  double m;
  double ep, em;        // e^{+mu}, e^{-mu}
};

static inline Geo make_geo(const th2d_geom *g)
{
  Geo o;
  o.nt = g->nt;
  o.nx = g->nx;
  o.nxh = g->nx / 2;
  o.vh = g->nt * g->nx / 2;
  o.m = g->m;
  o.ep = exp(g->mu);
  o.em = exp(-g->mu);
  return o;
}

// eo index of site (t, x)
__device__ __forceinline__ int eo_index(const Geo &G, int t, int x)
{
  return ((t + x) & 1) * G.vh + t * G.nxh + (x >> 1);
}

// ------------------------------------------------------- complex helpers --
template <typename T> struct cpx;
template <> struct cpx<double> { using t = double2; };
template <> struct cpx<float> { using t = float2; };
template <typename T> using c2 = typename cpx<T>::t;

template <typename V> __device__ __forceinline__ V cmk(decltype(V::x) a, decltype(V::x) b)
{
  V r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmk<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmk<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cscale(V a, decltype(V::x) s) { return cmk<V>(a.x * s, a.y * s); }
// a*b
template <typename V> __device__ __forceinline__ V cmul(V a, V b)
{
  return cmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a)*b
template <typename V> __device__ __forceinline__ V cjmul(V a, V b)
{
  return cmk<V>(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// ------------------------------------------------------------ reductions --
// Wave64 sum with a fixed shuffle order (deterministic).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of up to 16 waves; result valid in thread 0.  `sh` holds
// blockDim/64 doubles.  Fixed order: lanes, then waves by index.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
  v = wave_sum(v);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) s += sh[k];
  }
  return s;
}

// Block sums of two values at once (fixed order as block_sum); results
// valid in thread 0.  `sh` holds 2 * blockDim/64 doubles.
__device__ __forceinline__ double2 block_sum2(double a, double b, double *sh)
{
  a = wave_sum(a);
  b = wave_sum(b);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    sh[2 * wid] = a;
    sh[2 * wid + 1] = b;
  }
  __syncthreads();
  double2 s = make_double2(0.0, 0.0);
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) {
      s.x += sh[2 * k];
      s.y += sh[2 * k + 1];
    }
  }
  return s;
}

// ------------------------------------------------------- per-site code --
// Site of eo index i: (t, x) and the eo indices of its t- and x-neighbours.
struct EoSite {
  int p, t, x, j;
  int itu, itd, ixu, ixd;  // neighbours (other parity)
};

__device__ __forceinline__ EoSite eo_site(const Geo &G, int i)
{
  EoSite s;
  s.p = i >= G.vh;
  const int h = i - s.p * G.vh;
  s.t = h / G.nxh;
  s.j = h - s.t * G.nxh;
  s.x = 2 * s.j + ((s.t + s.p) & 1);
  const int q = (1 - s.p) * G.vh;
  const int tu = s.t + 1 == G.nt ? 0 : s.t + 1, td = s.t == 0 ? G.nt - 1 : s.t - 1;
  const int xu = s.x + 1 == G.nx ? 0 : s.x + 1, xd = s.x == 0 ? G.nx - 1 : s.x - 1;
  s.itu = q + tu * G.nxh + s.j;
  s.itd = q + td * G.nxh + s.j;
  s.ixu = q + s.t * G.nxh + (xu >> 1);
  s.ixd = q + s.t * G.nxh + (xd >> 1);
  return s;
}

// (M psi)(i) (dagger 0) or (M^dag psi)(i) (dagger 1) in FP64 with folded K.
__device__ __forceinline__ double2 dslash_site(const Geo &G, const double2 *__restrict__ in,
                                               const double2 *__restrict__ K, int i, int dagger)
{
  const int V = 2 * G.vh;
  const EoSite s = eo_site(G, i);
  const double2 kt = K[i], kx = K[V + i], ktd = K[s.itd], kxd = K[V + s.ixd];
  // M:     + e^{mu} K_t psi(x+t) - e^{-mu} K_t(x-t)* psi(x-t) + K_x psi(x+x) - K_x(x-x)* psi(x-x)
  // M^dag: - e^{-mu} K_t psi(x+t) + e^{mu} K_t(x-t)* psi(x-t) - K_x psi(x+x) + K_x(x-x)* psi(x-x)
  const double fu = dagger ? -G.em : G.ep, fd = dagger ? -G.ep : G.em, sx = dagger ? -1.0 : 1.0;
  double2 v = cscale(in[i], G.m);
  v = cadd(v, cscale(cmul(kt, in[s.itu]), fu));
  v = csub(v, cscale(cjmul(ktd, in[s.itd]), fd));
  v = cadd(v, cscale(cmul(kx, in[s.ixu]), sx));
  v = csub(v, cscale(cjmul(kxd, in[s.ixd]), sx));
  return v;
}

// K_mu(x) = 1/2 eta_mu(x) s_mu(x) e^{i A_mu(x)}: eta_t = (-1)^x, eta_x = 1
// (hmc.c:914-921); s = -1 for the hop leaving the last row/column when
// that direction is antiperiodic (hmc.c:143-180).
struct Bc {
  int anti_t, anti_x;
};

__device__ __forceinline__ void fold_one(const Geo &G, Bc bc, int i, const double *A, double *K,
                                         float *K32)
{
  const int V = 2 * G.vh;
  const EoSite s = eo_site(G, i);
  double ht = 0.5 * ((s.x & 1) ? -1.0 : 1.0);
  if (bc.anti_t && s.t == G.nt - 1) ht = -ht;
  double hx = 0.5;
  if (bc.anti_x && s.x == G.nx - 1) hx = -hx;
  double st, ct, sx, cx;
  sincos(A[i], &st, &ct);
  sincos(A[V + i], &sx, &cx);
  double2 kt = make_double2(ht * ct, ht * st), kx = make_double2(hx * cx, hx * sx);
  ((double2 *)K)[i] = kt;
  ((double2 *)K)[V + i] = kx;
  if (K32) {
    ((float2 *)K32)[i] = make_float2((float)kt.x, (float)kt.y);
    ((float2 *)K32)[V + i] = make_float2((float)kx.x, (float)kx.y);
  }
}

// dS/dA of both links of site i (momentum_step, hmc.c:504-665):
//   Nf/g sin A - 2 Re[i K (e^{+} conj(Mchi(x)) chi(x+mu) - e^{-} conj(chi(x)) Mchi(x+mu))]
//   + (e^{mu} - e^{-mu}) Re[i K conj(st(x)) st(x+mu)]   (t only)
// with sin A = 2 eta s Im K.  cg = Nf/g; st may be NULL.
__device__ __forceinline__ void force_site(const Geo &G, Bc bc, const double2 *__restrict__ K,
                                           const double2 *__restrict__ chi,
                                           const double2 *__restrict__ mchi,
                                           const double2 *__restrict__ st, double cg, int i,
                                           double F[2])
{
  const int V = 2 * G.vh;
  const EoSite s = eo_site(G, i);
  const double2 c0 = chi[i], m0 = mchi[i];
  for (int dir = 0; dir < 2; dir++) {
    const int in = dir == 0 ? s.itu : s.ixu;
    const double2 k = K[dir * V + i];
    const double2 c1 = chi[in], m1 = mchi[in];
    const double ef = dir == 0 ? G.ep : 1.0, eb = dir == 0 ? G.em : 1.0;
    const double2 a = cjmul(m0, c1), b = cjmul(c0, m1);
    const double2 z = make_double2(ef * a.x - eb * b.x, ef * a.y - eb * b.y);
    double f = 2.0 * (k.x * z.y + k.y * z.x);  // -2 Re[i k z] = 2 Im(k z)
    const double eta = dir == 0 ? ((s.x & 1) ? -1.0 : 1.0) : 1.0;
    const int wraps = dir == 0 ? (bc.anti_t && s.t == G.nt - 1) : (bc.anti_x && s.x == G.nx - 1);
    f += cg * 2.0 * (wraps ? -eta : eta) * k.y;
    if (st && dir == 0 && ef != eb) {
      const double2 cst = cjmul(st[i], st[in]);
      f -= (ef - eb) * (k.x * cst.y + k.y * cst.x);
    }
    F[dir] = f;
  }
}

// ------------------------------------------------ batched site operations --
// One launch applies a per-site operation to every chain of a batch: the
// host builds a table of per-chain descriptors (lattice, pointers,
// scalars) with the first 256-site block of each chain; a block finds its
// chain by binary search.
struct BOp {
  Geo G;
  Bc bc;
  long blk0;                   // first block of the chain
  double s0, s1;               // scalars of the operation
  double *d0, *d1, *d2, *d3, *d4;
  float *f0;
  double *out;                 // reduction result (device)
  const double *sp;            // TO_F32, ADD_F32: s0 = 1 / sqrt(*sp), sqrt(*sp) (device)
};

__device__ __forceinline__ int bop_chain(const BOp *ops, int nch, long blk)
{
  int lo = 0, hi = nch - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ops[mid].blk0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------- RNG ------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), keyed by a 64-bit seed
// and a 128-bit counter; uniform doubles in (0,1), never 0.
__device__ __forceinline__ uint4 philox10(uint4 c, uint2 k)
{
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += W0;
    k.y += W1;
  }
  return c;
}

__device__ __forceinline__ double u01(uint32_t a, uint32_t b)
{
  uint64_t v = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);  // 53 bits
  return ((double)v + 0.5) * (1.0 / 9007199254740992.0);
}


// --------------------------------------------------------- kernel entry --
// Host-side launchers shared between translation units.
namespace th2d {
int fold_links(th2d_ctx *ctx, const Geo &G, double *K, float *K32, const double *A);
int dslash_launch(th2d_ctx *ctx, const Geo &G, void *out, const void *in, const void *K,
                  int dagger, int prec, int nrhs);
int axpy_like(th2d_ctx *ctx, int op, long n, double a, double *y, const double *x,
              const double *z);
int reduce_dot(th2d_ctx *ctx, long n_real, const double *a, const double *b, double *dev_out);
int fetch_doubles(th2d_ctx *ctx, double *host, const double *dev, int n);
// Host-to-device copy of a small block through the context's pinned
// staging ring: the call returns once the copy is queued (a pageable
// source would make it wait for the staging).
int stage_h2d(th2d_ctx *ctx, void *dst, const void *src, size_t bytes);
// Device address of the static scratch array of the allocation-free LU
// inverse on the current device, looked up once per device (dense.hip).
int noalloc_scratch_init(int device);
// One MD step of a batch of small lattices as a captured graph (cg.hip):
// whether the batch qualifies, the solve's set-up (state: 3 pointers), its
// launch into a stream (accumulating per-chain statistics on the device)
// and the statistics after the replays.
bool md_graph_fits(th2d_ctx *ctx, int nch, const Geo *G, const th2d_cg_params *p);
int md_graph_prepare(th2d_ctx *ctx, int nch, const Geo *G, double *const *x,
                     const double *const *b, const double *const *K, const th2d_cg_params *p,
                     void **state);
void md_graph_solve(hipStream_t st, int nch, void *const *state);
int md_graph_stats(th2d_ctx *ctx, int nch, void *const *state, int *iters, int *solves,
                   int *status, double *max_true_rel);

// Batched site operations over a device table of nch descriptors covering
// nblk blocks of 256 sites (table built by bop_finish).
enum {
  BOP_GAUGE_STEP,  // A(d0) += 2 s0 mom(d1); K(d2), K32(f0) refolded (f0 may be NULL)
  BOP_FOLD,        // K(d2), K32(f0) from A(d0)
  BOP_DSLASH,      // d0 = M d1 (s1 == 0) or M^dag d1 (s1 != 0), K = d2
  BOP_FORCE,       // mom(d0) -= s0 F; K = d2, chi = d1, Mchi = d3, st = d4 (may be NULL); s1 = Nf/g
  BOP_AXPY,        // d0 += s0 d1
  BOP_SUB,         // d0 = d1 - d2
  BOP_TO_F32,      // f0 = (float)(s0 d1); s0 = 1 / sqrt(*sp) when sp is set
  BOP_ADD_F32,     // d0 += s0 f0; s0 = sqrt(*sp) when sp is set
  BOP_COPY,        // d0 = d1 (both link fields: 2 reals per direction)
  BOP_SCALE,       // d0 = s0 d1
  BOP_DOT,         // out = s1 * sum_sites Re(d0 . d1)  (two-pass, fixed order)
  BOP_COSGAP,      // out = s1 * sum_links (1 - cos d0)
  BOP_RESID,       // d0 = d1 - M^dag d3 (K = d2); out (and d4 if set) = s1 * sum_sites |d0|^2
};
// Fills blk0 and returns the block count.
long bop_finish(std::vector<BOp> &ops);
int bop_upload(th2d_ctx *ctx, th2d_buf &dev, const std::vector<BOp> &ops);
// The same into btab[slot], skipped when the bytes equal the last upload's.
int bop_upload_key(th2d_ctx *ctx, int slot, const std::vector<BOp> &ops);
int bop_launch(th2d_ctx *ctx, int op, const BOp *dev, int nch, long nblk);
}  // namespace th2d
