// zdense.hip -- dense complex fermion matrix and its determinant on gfx950
// (fermion_matrix, hmc.c:269-310; determinant, hmc.c:314-336).
//
// The matrix is V x V complex, column-major M[row + col*V] with row/col the
// natural site index t*nx + x, exactly as the reference builds it.  The
// determinant comes from partial-pivot LU in one workgroup per matrix
// (pivot = largest |Re|+|Im| as LAPACK's izamax), accumulated as
// log|det| and phase so that it neither overflows nor underflows.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>

namespace {

__global__ void k_zero_f64(double *p, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Row s of M: m on the diagonal, the four hops of fm_mul (K folded, eo).
__global__ void k_fermion_matrix(Geo G, const double2 *K, double2 *M)
{
  const int V = 2 * G.vh;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const int t = s / G.nx, x = s - t * G.nx;
  const int tp = t + 1 == G.nt ? 0 : t + 1, tm = t == 0 ? G.nt - 1 : t - 1;
  const int xp = x + 1 == G.nx ? 0 : x + 1, xm = x == 0 ? G.nx - 1 : x - 1;
  const double2 kt = K[eo_index(G, t, x)], kx = K[V + eo_index(G, t, x)];
  const double2 ktd = K[eo_index(G, tm, x)], kxd = K[V + eo_index(G, t, xm)];
  auto at = [&](int col) -> double2 & { return M[(long)s + (long)col * V]; };
  at(s) = make_double2(G.m, 0.0);
  // += so that lattices of extent 2 (both neighbours the same site) add up
  double2 v;
  v = at(tp * G.nx + x);
  at(tp * G.nx + x) = make_double2(v.x + G.ep * kt.x, v.y + G.ep * kt.y);
  v = at(tm * G.nx + x);
  at(tm * G.nx + x) = make_double2(v.x - G.em * ktd.x, v.y + G.em * ktd.y);
  v = at(t * G.nx + xp);
  at(t * G.nx + xp) = make_double2(v.x + kx.x, v.y + kx.y);
  v = at(t * G.nx + xm);
  at(t * G.nx + xm) = make_double2(v.x - kxd.x, v.y + kxd.y);
}

__device__ __forceinline__ double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }

// In-place LU of one column-major n x n complex matrix by one workgroup;
// out[0] = log|det|, out[1] = arg(det) in (-pi, pi], out[2] = first zero
// pivot + 1 (0 if none).
__global__ void __launch_bounds__(1024) k_zlu_det(double2 *A, int n, double *out)
{
  __shared__ double sv[1024];
  __shared__ int si[1024];
  __shared__ double acc[2];
  const int tid = threadIdx.x, nt = blockDim.x;
  double la = 0.0, ph = 0.0;
  int info = 0;
  for (int j = 0; j < n; j++) {
    double best = -1.0;
    int bi = n;
    for (int i = j + tid; i < n; i += nt) {
      double v = cabs1(A[i + (long)j * n]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int s = nt >> 1; s > 0; s >>= 1) {
      if (tid < s && (sv[tid + s] > sv[tid] || (sv[tid + s] == sv[tid] && si[tid + s] < si[tid]))) {
        sv[tid] = sv[tid + s];
        si[tid] = si[tid + s];
      }
      __syncthreads();
    }
    const int p = si[0];
    const double pabs = sv[0];
    __syncthreads();
    if (pabs == 0.0) {
      if (!info) info = j + 1;
      continue;
    }
    if (p != j)
      for (int c = tid; c < n; c += nt) {
        double2 a = A[j + (long)c * n];
        A[j + (long)c * n] = A[p + (long)c * n];
        A[p + (long)c * n] = a;
      }
    __syncthreads();
    const double2 pv = A[j + (long)j * n];
    if (tid == 0) {
      la += 0.5 * log(pv.x * pv.x + pv.y * pv.y);
      ph += atan2(pv.y, pv.x) + (p != j ? M_PI : 0.0);
    }
    const double den = pv.x * pv.x + pv.y * pv.y;
    const double2 inv = make_double2(pv.x / den, -pv.y / den);
    for (int i = j + 1 + tid; i < n; i += nt) {
      double2 a = A[i + (long)j * n];
      A[i + (long)j * n] = make_double2(a.x * inv.x - a.y * inv.y, a.x * inv.y + a.y * inv.x);
    }
    __syncthreads();
    const int w = n - j - 1;
    for (long e = tid; e < (long)w * w; e += nt) {
      const int i = j + 1 + (int)(e % w), c = j + 1 + (int)(e / w);
      const double2 l = A[i + (long)j * n], u = A[j + (long)c * n];
      double2 a = A[i + (long)c * n];
      a.x -= l.x * u.x - l.y * u.y;
      a.y -= l.x * u.y + l.y * u.x;
      A[i + (long)c * n] = a;
    }
    __syncthreads();
  }
  if (tid == 0) {
    acc[0] = la;
    acc[1] = remainder(ph, 2.0 * M_PI);
    out[0] = info ? -INFINITY : acc[0];
    out[1] = info ? 0.0 : acc[1];
    out[2] = info;
  }
}

}  // namespace

extern "C" int th2d_fermion_matrix(th2d_ctx *ctx, const th2d_geom *g, const double *K, double *M)
{
  TH2D_TRY(th2d_check_geom(g));
  if (!ctx || !K || !M) return TH2D_EARG;
  const Geo G = make_geo(g);
  const long V = 2L * G.vh;
  hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)std::min(4096L, (2 * V * V + 255) / 256)), dim3(256),
                     0, ctx->stream, M, 2 * V * V);
  hipLaunchKernelGGL(k_fermion_matrix, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, ctx->stream, G,
                     (const double2 *)K, (double2 *)M);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_zdet(th2d_ctx *ctx, int n, double *M, double *logabs, double *phase)
{
  if (!ctx || n < 1 || !M || !logabs || !phase) return TH2D_EARG;
  TH2D_TRY(th2d_buf_reserve(ctx->red, 64 * sizeof(double)));
  double *d = (double *)ctx->red.ptr;
  hipLaunchKernelGGL(k_zlu_det, dim3(1), dim3(1024), 0, ctx->stream, (double2 *)M, n, d);
  TH2D_HIP(hipGetLastError());
  double h[3];
  TH2D_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  TH2D_HIP(hipStreamSynchronize(ctx->stream));
  *logabs = h[0];
  *phase = h[1];
  return h[2] != 0.0 ? TH2D_ESINGULAR : TH2D_OK;
}
