// th2d_internal.h -- shared definitions of the gfx950 kernels and the host
// runtime of libth2d.  Not installed; the public ABI is include/th2d.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>

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
  double ktime_ms = 0.0;
  long klaunches = 0;
  double kbytes = 0.0;
  // scratch
  th2d_buf cg64, cg32, cgmix, red, redh, hmc, force, bag, misc;
  // tile configuration of the persistent solver (0 = autotuned default)
  int tile_t = 0, tile_x = 0;
  int cg_max_blocks = 0;  // 0 = one workgroup per tile
  int cg_chunk = 0;       // step launches per stop-flag check (0 = from history)
  long kiters = 0;
  int cg_pass_iters[2][8] = {};   // iterations per restart pass (FP32, FP64)
  int cg_last_tile[2] = {0, 0};   // tile shape of the last inner solve
  int cg_autotune = 0;            // measure tile shapes per batch signature
  std::map<unsigned long long, int> cg_tuned[2];  // signature -> shape (FP32, FP64)
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
}  // namespace th2d
