// field.hip -- context, memory, layout, link folding, gauge updates,
// reductions and the device random number generator of libth2d (gfx950).
#include "th2d_internal.h"

#include <algorithm>

#include <cmath>
#include <cstdlib>

// =========================================================== context ====
int th2d_buf_reserve(th2d_buf &b, size_t bytes)
{
  if (b.bytes >= bytes && b.ptr) return TH2D_OK;
  if (b.ptr) {
    if (hipFree(b.ptr) != hipSuccess) return TH2D_EHIP;
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipMalloc(&b.ptr, want) != hipSuccess) return TH2D_ENOMEM;
  b.bytes = want;
  return TH2D_OK;
}

int th2d_host_reserve(th2d_buf &b, size_t bytes)
{
  if (b.bytes >= bytes && b.ptr) return TH2D_OK;
  if (b.ptr) (void)hipHostFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  size_t want = bytes < 4096 ? 4096 : bytes;
  if (hipHostMalloc(&b.ptr, want, hipHostMallocDefault) != hipSuccess) return TH2D_ENOMEM;
  b.bytes = want;
  return TH2D_OK;
}

static void buf_release(th2d_buf &b, bool host = false)
{
  if (b.ptr) {
    if (host) (void)hipHostFree(b.ptr);
    else (void)hipFree(b.ptr);
  }
  b.ptr = nullptr;
  b.bytes = 0;
}

extern "C" int th2d_ctx_create(th2d_ctx **out, void *stream)
{
  if (!out) return TH2D_EARG;
  th2d_ctx *c = new th2d_ctx();
  c->stream = (hipStream_t)stream;
  TH2D_HIP(hipGetDevice(&c->device));
  hipDeviceProp_t prop;
  TH2D_HIP(hipGetDeviceProperties(&prop, c->device));
  c->num_cu = prop.multiProcessorCount;
  TH2D_HIP(hipEventCreate(&c->t0));
  TH2D_HIP(hipEventCreate(&c->t1));
  TH2D_HIP(hipEventCreate(&c->k0));
  TH2D_HIP(hipEventCreate(&c->k1));
  TH2D_HIP(hipEventCreate(&c->a0));
  TH2D_HIP(hipEventCreate(&c->a1));
  TH2D_HIP(hipEventCreateWithFlags(&c->sync_ev, hipEventDisableTiming));
  TH2D_TRY(th2d::noalloc_scratch_init(c->device));
  *out = c;
  return TH2D_OK;
}

extern "C" int th2d_ctx_destroy(th2d_ctx *c)
{
  if (!c) return TH2D_OK;
  (void)hipStreamSynchronize(c->stream);
  for (auto &g : c->cg_graphs) (void)hipGraphExecDestroy(g.exec);
  c->cg_graphs.clear();
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->lu_ps) (void)hipStreamDestroy(c->lu_ps);
  if (c->lu_us) (void)hipStreamDestroy(c->lu_us);
  for (hipEvent_t e : c->lu_ev)
    if (e) (void)hipEventDestroy(e);
  buf_release(c->cg64);
  buf_release(c->cg32);
  buf_release(c->cg32b);
  buf_release(c->cgmix);
  buf_release(c->red);
  buf_release(c->redh, true);
  buf_release(c->hstage, true);
  buf_release(c->resh, true);
  buf_release(c->hmc);
  buf_release(c->force);
  buf_release(c->bag);
  buf_release(c->misc);
  buf_release(c->bpart);
  for (auto &b : c->btab) buf_release(b);
  buf_release(c->dsl_tab);
  (void)hipEventDestroy(c->t0);
  (void)hipEventDestroy(c->t1);
  (void)hipEventDestroy(c->k0);
  (void)hipEventDestroy(c->k1);
  (void)hipEventDestroy(c->a0);
  (void)hipEventDestroy(c->a1);
  if (c->sync_ev) (void)hipEventDestroy(c->sync_ev);
  delete c;
  return TH2D_OK;
}

extern "C" const char *th2d_strerror(int code)
{
  switch (code) {
    case TH2D_OK: return "ok";
    case TH2D_EARG: return "bad argument";
    case TH2D_EHIP: return "HIP runtime error";
    case TH2D_ENOCONV: return "solver did not converge within max_iter";
    case TH2D_EBREAKDOWN: return "solver breakdown (non-positive curvature, NaN or divergence)";
    case TH2D_ESINGULAR: return "singular matrix (zero pivot)";
    case TH2D_ENOMEM: return "device allocation failed";
    case TH2D_ELAUNCH: return "cooperative launch does not fit the device";
    case TH2D_ESYNC: return "a bounded wait between workgroups timed out";
    default: return "unknown th2d status";
  }
}

extern "C" int th2d_device_count(int *n)
{
  if (!n) return TH2D_EARG;
  if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
  return TH2D_OK;
}

extern "C" int th2d_device_props(int *num_cu, int *num_xcd, size_t *lds, char *name, int name_len)
{
  int dev;
  TH2D_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  TH2D_HIP(hipGetDeviceProperties(&p, dev));
  if (num_cu) *num_cu = p.multiProcessorCount;
  if (num_xcd) *num_xcd = 8;  // MI355X: 8 XCDs of 32 CUs
  if (lds) *lds = p.maxSharedMemoryPerMultiProcessor;
  if (name && name_len > 0) {
    strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  return TH2D_OK;
}

extern "C" int th2d_malloc(void **d, size_t bytes)
{
  if (!d) return TH2D_EARG;
  if (hipMalloc(d, bytes ? bytes : 1) != hipSuccess) return TH2D_ENOMEM;
  return TH2D_OK;
}

extern "C" int th2d_free(void *d)
{
  if (d) TH2D_HIP(hipFree(d));
  return TH2D_OK;
}

__global__ void k_fill_bytes(unsigned char *p, unsigned char v, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Memory clear as a kernel (safe inside captured graphs).
extern "C" int th2d_memset(th2d_ctx *c, void *d, int value, size_t bytes)
{
  if (!c || (!d && bytes)) return TH2D_EARG;
  if (!bytes) return TH2D_OK;
  unsigned char v = (unsigned char)value;
  if ((bytes & 3) == 0 && ((uintptr_t)d & 3) == 0) {
    uint32_t w = v * 0x01010101u;
    size_t n = bytes / 4;
    int blocks = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, c->stream, (uint32_t *)d, w, n);
  } else {
    int blocks = (int)((bytes + 255) / 256 < 2048 ? (bytes + 255) / 256 : 2048);
    hipLaunchKernelGGL(k_fill_bytes, dim3(blocks), dim3(256), 0, c->stream, (unsigned char *)d, v,
                       bytes);
  }
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_h2d(th2d_ctx *c, void *dst, const void *src, size_t bytes)
{
  if (!c) return TH2D_EARG;
  TH2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  return TH2D_OK;
}

extern "C" int th2d_d2h(th2d_ctx *c, void *dst, const void *src, size_t bytes)
{
  if (!c) return TH2D_EARG;
  TH2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  return TH2D_OK;
}

extern "C" int th2d_d2d(th2d_ctx *c, void *dst, const void *src, size_t bytes)
{
  if (!c) return TH2D_EARG;
  TH2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return TH2D_OK;
}

extern "C" int th2d_sync(th2d_ctx *c)
{
  if (!c) return TH2D_EARG;
  TH2D_HIP(hipStreamSynchronize(c->stream));
  return TH2D_OK;
}

extern "C" int th2d_timer_start(th2d_ctx *c)
{
  if (!c) return TH2D_EARG;
  TH2D_HIP(hipEventRecord(c->t0, c->stream));
  return TH2D_OK;
}

extern "C" int th2d_timer_stop(th2d_ctx *c, float *ms)
{
  if (!c || !ms) return TH2D_EARG;
  TH2D_HIP(hipEventRecord(c->t1, c->stream));
  TH2D_HIP(hipEventSynchronize(c->t1));
  TH2D_HIP(hipEventElapsedTime(ms, c->t0, c->t1));
  return TH2D_OK;
}

extern "C" int th2d_kernel_timing(th2d_ctx *c, int enable)
{
  if (!c) return TH2D_EARG;
  c->ktiming = enable;
  c->ktime_ms = 0.0;
  c->klaunches = 0;
  c->kbytes = 0.0;
  return TH2D_OK;
}

extern "C" int th2d_kernel_time(th2d_ctx *c, double *ms, long *n, double *bytes)
{
  if (!c) return TH2D_EARG;
  if (ms) *ms = c->ktime_ms;
  if (n) *n = c->klaunches;
  if (bytes) *bytes = c->kbytes;
  return TH2D_OK;
}

// ============================================================ layout ====
__global__ void k_vec_to_eo(Geo G, double2 *eo, const double2 *nat, int inverse)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int V = 2 * G.vh;
  if (s >= V) return;
  int t = s / G.nx, x = s - t * G.nx;
  int i = eo_index(G, t, x);
  if (inverse) eo[s] = nat[i];  // here eo is the natural destination
  else eo[i] = nat[s];
}

__global__ void k_links_to_eo(Geo G, double *eo, const double *nat, int inverse)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int V = 2 * G.vh;
  if (s >= V) return;
  int t = s / G.nx, x = s - t * G.nx;
  int i = eo_index(G, t, x);
  if (inverse) {
    eo[2 * s] = nat[i];
    eo[2 * s + 1] = nat[V + i];
  } else {
    eo[i] = nat[2 * s];
    eo[V + i] = nat[2 * s + 1];
  }
}

extern "C" int th2d_vec_to_eo(th2d_ctx *c, const th2d_geom *g, double *eo, const double *nat)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !eo || !nat) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_vec_to_eo, dim3((V + 255) / 256), dim3(256), 0, c->stream, G,
                     (double2 *)eo, (const double2 *)nat, 0);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_vec_from_eo(th2d_ctx *c, const th2d_geom *g, double *nat, const double *eo)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !eo || !nat) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_vec_to_eo, dim3((V + 255) / 256), dim3(256), 0, c->stream, G,
                     (double2 *)nat, (const double2 *)eo, 1);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_links_to_eo(th2d_ctx *c, const th2d_geom *g, double *eo, const double *nat)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !eo || !nat) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_links_to_eo, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, eo, nat, 0);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_links_from_eo(th2d_ctx *c, const th2d_geom *g, double *nat, const double *eo)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !eo || !nat) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_links_to_eo, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, nat, eo, 1);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

__global__ void k_f64_to_f32(float2 *d, const double2 *s, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_float2((float)s[i].x, (float)s[i].y);
}

__global__ void k_f32_to_f64(double2 *d, const float2 *s, long n)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = make_double2((double)s[i].x, (double)s[i].y);
}

extern "C" int th2d_vec_f64_to_f32(th2d_ctx *c, float *d, const double *s, long n)
{
  if (!c || n < 0) return TH2D_EARG;
  if (!n) return TH2D_OK;
  hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (float2 *)d, (const double2 *)s, n);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_vec_f32_to_f64(th2d_ctx *c, double *d, const float *s, long n)
{
  if (!c || n < 0) return TH2D_EARG;
  if (!n) return TH2D_OK;
  hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (double2 *)d, (const float2 *)s, n);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// ======================================================= gauge field ====
// Folding of the links: fold_one (th2d_internal.h).
__global__ void k_fold(Geo G, Bc bc, const double *A, double *K, float *K32)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * G.vh) fold_one(G, bc, i, A, K, K32);
}

// A += 2 eps mom and refold, one pass over the links.
__global__ void k_gauge_step(Geo G, Bc bc, double *A, const double *mom, double eps, double *K,
                             float *K32)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int V = 2 * G.vh;
  if (i >= V) return;
  A[i] += 2.0 * eps * mom[i];
  A[V + i] += 2.0 * eps * mom[V + i];
  if (K) fold_one(G, bc, i, A, K, K32);
}

static Bc bc_of(const th2d_geom *g) { return Bc{g->anti_t, g->anti_x}; }

extern "C" int th2d_fold_links(th2d_ctx *c, const th2d_geom *g, double *K, float *K32,
                               const double *A)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !K || !A) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_fold, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, bc_of(g), A, K,
                     K32);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_gauge_step(th2d_ctx *c, const th2d_geom *g, double *A, const double *mom,
                               double eps, double *K, float *K32)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !A || !mom) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_gauge_step, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, bc_of(g), A,
                     mom, eps, K, K32);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// ========================================================= reductions ===
// Two-pass deterministic sums: pass 1 gives one partial per block over a
// fixed index partition, pass 2 adds the partials in block order.
enum { RED_DOT = 0, RED_CROSS = 1, RED_COSGAP = 2 };
static const int RED_BLOCKS = 512;

__global__ void __launch_bounds__(256) k_reduce1(int op, long n, const double *a, const double *b,
                                                 double *part)
{
  __shared__ double sh[4];
  double acc = 0.0;
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (op == RED_DOT) acc += a[i] * b[i];
    else if (op == RED_CROSS) {
      // Im(conj(a) b) over complex pairs, i counts complex numbers
      double2 x = ((const double2 *)a)[i], y = ((const double2 *)b)[i];
      acc += x.x * y.y - x.y * y.x;
    } else acc += 1.0 - cos(a[i]);
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_reduce2(int nparts, const double *part, double *out,
                                                 double scale)
{
  __shared__ double sh[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += part[i];
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) *out = s * scale;
}

namespace th2d {

int reduce_op(th2d_ctx *c, int op, long n, const double *a, const double *b, double *dout,
                     double scale)
{
  TH2D_TRY(th2d_buf_reserve(c->red, sizeof(double) * (RED_BLOCKS + 64)));
  double *part = (double *)c->red.ptr;
  long blocks = (n + 255) / 256;
  if (blocks > RED_BLOCKS) blocks = RED_BLOCKS;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_reduce1, dim3((unsigned)blocks), dim3(256), 0, c->stream, op, n, a, b,
                     part);
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, c->stream, (int)blocks, part, dout, scale);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

int reduce_dot(th2d_ctx *c, long n_real, const double *a, const double *b, double *dev_out)
{
  return reduce_op(c, RED_DOT, n_real, a, b, dev_out, 1.0);
}

int fetch_doubles(th2d_ctx *c, double *host, const double *dev, int n)
{
  TH2D_TRY(th2d_host_reserve(c->redh, sizeof(double) * n));
  TH2D_HIP(hipMemcpyAsync(c->redh.ptr, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  c->hstage_off = 0;  // the stream is idle: every staged copy has been read
  memcpy(host, c->redh.ptr, sizeof(double) * n);
  return TH2D_OK;
}

int stage_h2d(th2d_ctx *c, void *dst, const void *src, size_t bytes)
{
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (c->hstage_off + need > c->hstage.bytes) {
    // the ring is reused from its start once every copy out of it is done
    TH2D_HIP(hipStreamSynchronize(c->stream));
    c->hstage_off = 0;
    if (need > c->hstage.bytes)
      TH2D_TRY(th2d_host_reserve(c->hstage, std::max(need, (size_t)1 << 18)));
  }
  char *p = (char *)c->hstage.ptr + c->hstage_off;
  memcpy(p, src, bytes);
  TH2D_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, c->stream));
  c->hstage_off += need;
  return TH2D_OK;
}

int reduce_to_host(th2d_ctx *c, int op, long n, const double *a, const double *b, double *out,
                   double scale)
{
  TH2D_TRY(th2d_buf_reserve(c->red, sizeof(double) * (RED_BLOCKS + 64)));
  double *slot = (double *)c->red.ptr + RED_BLOCKS;
  TH2D_TRY(reduce_op(c, op, n, a, b, slot, scale));
  return fetch_doubles(c, out, slot, 1);
}

}  // namespace th2d

extern "C" int th2d_dot(th2d_ctx *c, long n, const double *a, const double *b, double *re,
                        double *im)
{
  if (!c || n < 0 || !a || !b) return TH2D_EARG;
  if (re) TH2D_TRY(th2d::reduce_to_host(c, RED_DOT, 2 * n, a, b, re, 1.0));
  if (im) TH2D_TRY(th2d::reduce_to_host(c, RED_CROSS, n, a, b, im, 1.0));
  return TH2D_OK;
}

extern "C" int th2d_norm2(th2d_ctx *c, long n, const double *a, double *out)
{
  if (!c || n < 0 || !a || !out) return TH2D_EARG;
  return th2d::reduce_to_host(c, RED_DOT, n, a, a, out, 1.0);
}

extern "C" int th2d_gauge_action(th2d_ctx *c, const th2d_geom *g, const double *A, double gc,
                                 int nf, double *out)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !A || !out || gc == 0.0) return TH2D_EARG;
  long V = (long)g->nt * g->nx;
  return th2d::reduce_to_host(c, RED_COSGAP, 2 * V, A, nullptr, out, nf / gc);
}

// Fused BLAS-1 helpers used by the outer (FP64) solver loop.
//  op 0: y += a*x      op 1: y = x - z (residual)      op 2: y = a*x
__global__ void k_axpy_like(int op, long n, double a, double *y, const double *x, const double *z)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (op == 0) y[i] += a * x[i];
  else if (op == 1) y[i] = x[i] - z[i];
  else y[i] = a * x[i];
}

namespace th2d {
int axpy_like(th2d_ctx *c, int op, long n, double a, double *y, const double *x, const double *z)
{
  hipLaunchKernelGGL(k_axpy_like, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, op,
                     n, a, y, x, z);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}
}  // namespace th2d

// ============================================================ RNG =======
// Two independent N(0,1) for (seed, stream, index).
__device__ __forceinline__ double2 gauss2(unsigned long long seed, uint32_t stream, uint32_t idx,
                                          uint32_t ctr)
{
  uint4 r = philox10(make_uint4(idx, ctr, stream, 0x7468u),
                     make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
  double u1 = u01(r.x, r.y), u2 = u01(r.z, r.w);
  double rad = sqrt(-2.0 * log(u1)), s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_double2(rad * c, rad * s);
}

// Raw generator output for known-answer checks: item j has counter
// ck[6j..6j+3] and key ck[6j+4..6j+5]; out[4j..4j+3] receives the block.
__global__ void k_philox_blocks(long n, const uint32_t *ck, uint32_t *out)
{
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t *q = ck + 6 * j;
  uint4 r = philox10(make_uint4(q[0], q[1], q[2], q[3]), make_uint2(q[4], q[5]));
  out[4 * j] = r.x;
  out[4 * j + 1] = r.y;
  out[4 * j + 2] = r.z;
  out[4 * j + 3] = r.w;
}

extern "C" int th2d_philox4x32_10(th2d_ctx *c, long n, const unsigned *ctr_key, unsigned *out)
{
  if (!c || n < 0 || (n > 0 && (!ctr_key || !out))) return TH2D_EARG;
  if (n == 0) return TH2D_OK;
  hipLaunchKernelGGL(k_philox_blocks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     n, (const uint32_t *)ctr_key, (uint32_t *)out);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// Gaussian fill of n complex numbers (or 2n reals) scaled by `scale`.
__global__ void k_gauss_fill(double2 *v, long n, unsigned long long seed, uint32_t stream,
                             uint32_t ctr, double scale)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 z = gauss2(seed, stream, (uint32_t)i, ctr);
  v[i] = make_double2(scale * z.x, scale * z.y);
}

namespace th2d {
int gauss_fill(th2d_ctx *c, double *v, long n_complex, unsigned long long seed, uint32_t stream,
               uint32_t ctr, double scale)
{
  hipLaunchKernelGGL(k_gauss_fill, dim3((unsigned)((n_complex + 255) / 256)), dim3(256), 0,
                     c->stream, (double2 *)v, n_complex, seed, stream, ctr, scale);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}
}  // namespace th2d

extern "C" int th2d_random_links(th2d_ctx *c, const th2d_geom *g, double *A,
                                 unsigned long long seed)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !A) return TH2D_EARG;
  long V = (long)g->nt * g->nx;
  return th2d::gauss_fill(c, A, V, seed, 0x4c4e4b53u, 0, 1.0);  // 2V reals
}

// Pure-gauge Metropolis, each link independent given its two draws
// (hmc.c:82-93: proposal 2 pi u - pi, accept if u' < exp(-s2 + s1)).
__global__ void k_puregauge(Geo G, double *A, const double *draws, double cg)
{
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int V = 2 * G.vh;
  if (s >= V) return;
  int t = s / G.nx, x = s - t * G.nx;
  int i = eo_index(G, t, x);
  for (int dir = 0; dir < 2; dir++) {
    double *a = A + dir * V + i;
    double u0 = draws[4 * s + 2 * dir], u1 = draws[4 * s + 2 * dir + 1];
    double s1 = -cg * cos(*a);
    double nw = 2 * M_PI * u0 - M_PI;
    double s2 = -cg * cos(nw);
    if (u1 < exp(-s2 + s1)) *a = nw;
  }
}

// n pure-gauge Metropolis sweeps with device draws.  The action
// Nf/g sum(1 - cos A) is a sum of single-link terms, so every link runs its
// n sweeps in registers; sweep k of link l draws one Philox block at
// (index l, counter ctr0 + k): the first 53-bit uniform is the proposal,
// the second the accept draw (hmc.c:82-93 order).
__device__ __forceinline__ void uniform2(unsigned long long seed, uint32_t stream, uint32_t idx,
                                         uint32_t ctr, double *u0, double *u1)
{
  uint4 r = philox10(make_uint4(idx, ctr, stream, 0x7468u),
                     make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
  *u0 = u01(r.x, r.y);
  *u1 = u01(r.z, r.w);
}

__global__ void k_puregauge_n(long nlinks, double *A, double cg, int nsweeps,
                              unsigned long long seed, uint32_t ctr0)
{
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  double a = A[l];
  for (int k = 0; k < nsweeps; k++) {
    double u0, u1;
    uniform2(seed, 0x50475357u, (uint32_t)l, ctr0 + (uint32_t)k, &u0, &u1);
    const double nw = 2 * M_PI * u0 - M_PI;
    if (u1 < exp(-cg * cos(a) + cg * cos(nw))) a = nw;
  }
  A[l] = a;
}

extern "C" int th2d_puregauge_sweeps(th2d_ctx *c, const th2d_geom *g, double *A, double gc,
                                     int nf, int nsweeps, unsigned long long seed,
                                     unsigned long long counter)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !A || gc == 0.0 || nsweeps < 0) return TH2D_EARG;
  if (nsweeps == 0) return TH2D_OK;
  const long nl = 2L * g->nt * g->nx;
  hipLaunchKernelGGL(k_puregauge_n, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, c->stream,
                     nl, A, nf / gc, nsweeps, seed, (uint32_t)counter);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_puregauge_sweep(th2d_ctx *c, const th2d_geom *g, double *A,
                                    const double *draws, double gc, int nf)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!c || !A || !draws || gc == 0.0) return TH2D_EARG;
  Geo G = make_geo(g);
  int V = 2 * G.vh;
  hipLaunchKernelGGL(k_puregauge, dim3((V + 255) / 256), dim3(256), 0, c->stream, G, A, draws,
                     nf / gc);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}
