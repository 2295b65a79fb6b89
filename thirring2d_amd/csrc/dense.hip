// dense.hip -- batched dense linear algebra on gfx950 for the fermion-bag
// determinants (full_determinant.c:144-186, fluctuation_determinant.c:
// 126-301, fermion_matrix.c:355-427): partial-pivot LU, log|det| and sign,
// inverse from the LU factors, GEMM updates on the matrix cores.
//
// A batch is a set of row-major square matrices in one device buffer:
// matrix b has order n[b], starts at element off[b], its pivots at poff[b].
// Every kernel takes the whole batch; matrices of different orders share a
// launch (blocks of a matrix that is already finished return at once).
//
// LU: right-looking with partial pivoting (the factors equal those of
// LAPACK's whole-row interchanges).  Up to order PR*PT, blocks of LOB = 128
// columns, each factored as register panels of 16 columns; then
//   trsm   -- U12 = L11^{-1} A12 (unit lower, one thread per column);
//   update -- A22 -= L21 U12 of depth 128, a GEMM on MFMA 16x16x4 tiles
//             (FP64 or FP32).
// Larger orders use panels of NB = 32 columns reduced through LDS, with a
// trailing update per panel.
#include "th2d_internal.h"

#include <mutex>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/th2d_bag.h"

namespace {

constexpr int NB = 32;    // LU panel width
constexpr int BM = 64;    // GEMM tile rows
constexpr int BN = 64;    // GEMM tile cols
constexpr int BK = 16;    // GEMM k chunk

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Operand layout of both 16x16x4 instructions (lane l): A[l%16][l/16],
// B[l/16][l%16].  Result register r of lane l holds row row(l, r) and
// column l%16 -- rows are blocked by lane group in FP32 (4*(l/16) + r)
// but interleaved in FP64 (4*r + l/16), as measured on gfx950.
template <> struct Mfma<double> {
  using acc = v4d;
  static __device__ __forceinline__ acc op(double a, double b, acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <> struct Mfma<float> {
  using acc = v4f;
  static __device__ __forceinline__ acc op(float a, float b, acc c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// C[i][j] += alpha * sum_k A[i][k] B[k][j] for one BM x BN tile at
// (i0, j0); rows >= M, cols >= N, k >= K are skipped.  256 threads: wave w
// owns the 32x32 quadrant (w/2, w%2) as 2x2 MFMA blocks of 16x16; a
// thread's results sit at rows Mfma<T>::row(lane, r) of each block.
template <typename T, bool USE_MFMA>
__device__ void gemm_tile(const T *A, long lda, const T *B, long ldb, T *C, long ldc, int M, int N,
                          int K, T alpha, int i0, int j0, T *sA, T *sB)
{
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;
  using Acc = typename Mfma<T>::acc;
  Acc acc[2][2];
  T vacc[4][4];
  if (USE_MFMA) {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = Acc{0, 0, 0, 0};
  } else {
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) vacc[a][b] = 0;
  }
  // software pipeline: the next chunk's A and B entries are loaded into
  // registers while the current chunk is multiplied out of LDS
  constexpr int EA = BM * BK / 256, EB = BK * BN / 256;
  T ra[EA], rb[EB];
  // an FP64 tile inside the matrix loads its chunks without bounds checks
  // (a uniform branch per chunk), so the loads issue back to back; in FP32
  // the second path costs registers, and so occupancy
  const bool inner = sizeof(T) == 8 && i0 + BM <= M && j0 + BN <= N;
  auto fetch = [&](int k0) {
    if (inner && k0 + BK <= K) {
#pragma unroll
      for (int q = 0; q < EA; q++) {
        const int e = tid + q * 256, r = e / BK, c = e - r * BK;
        ra[q] = A[(i0 + r) * lda + k0 + c];
      }
#pragma unroll
      for (int q = 0; q < EB; q++) {
        const int e = tid + q * 256, r = e / BN, c = e - r * BN;
        rb[q] = B[(k0 + r) * ldb + j0 + c];
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < EA; q++) {
      const int e = tid + q * 256, r = e / BK, c = e - r * BK;
      const int gi = i0 + r, gk = k0 + c;
      ra[q] = (gi < M && gk < K) ? A[gi * lda + gk] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < EB; q++) {
      const int e = tid + q * 256, r = e / BN, c = e - r * BN;
      const int gk = k0 + r, gj = j0 + c;
      rb[q] = (gk < K && gj < N) ? B[gk * ldb + gj] : (T)0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EA; q++) {
      const int e = tid + q * 256, r = e / BK, c = e - r * BK;
      sA[r * (BK + 1) + c] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < EB; q++) {
      const int e = tid + q * 256, r = e / BN, c = e - r * BN;
      sB[r * (BN + 1) + c] = rb[q];
    }
    __syncthreads();
    if (k0 + BK < K) fetch(k0 + BK);
    if (USE_MFMA) {
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T av[2], bv[2];
        for (int a = 0; a < 2; a++) av[a] = sA[(wr + a * 16 + (lane & 15)) * (BK + 1) + kk + (lane >> 4)];
        for (int b = 0; b < 2; b++) bv[b] = sB[(kk + (lane >> 4)) * (BN + 1) + wc + b * 16 + (lane & 15)];
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) acc[a][b] = Mfma<T>::op(av[a], bv[b], acc[a][b]);
      }
    } else {
      // thread owns the entries of the MFMA result layout:
      // rows wr + a*16 + row(lane, r), cols wc + b*16 + lane%16
      for (int kk = 0; kk < BK; kk++)
        for (int a = 0; a < 2; a++)
          for (int r = 0; r < 4; r++) {
            T av = sA[(wr + a * 16 + Mfma<T>::row(lane, r)) * (BK + 1) + kk];
            for (int b = 0; b < 2; b++)
              vacc[a * 2 + b][r] += av * sB[kk * (BN + 1) + wc + b * 16 + (lane & 15)];
          }
    }
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        int gi = i0 + wr + a * 16 + Mfma<T>::row(lane, r);
        int gj = j0 + wc + b * 16 + (lane & 15);
        if (inner || (gi < M && gj < N)) {
          T v = USE_MFMA ? acc[a][b][r] : vacc[a * 2 + b][r];
          C[gi * ldc + gj] += alpha * v;
        }
      }
}

struct BatchView {
  const long *off;
  const int *n;
  const long *poff;
  int count;
};

// ----------------------------------------------------------------- LU ---
template <typename T>
__global__ void __launch_bounds__(256) k_lu_panel(T *base, BatchView bv, int *pivb, int *info,
                                                  int k)
{
  __shared__ double sval[256];
  __shared__ int sidx[256];
  const int b = blockIdx.x;
  const int n = bv.n[b];
  if (k >= n) return;
  T *A = base + bv.off[b];
  int *piv = pivb + bv.poff[b];
  const int jb = min(NB, n - k), tid = threadIdx.x;
  for (int j = k; j < k + jb; j++) {
    // pivot: largest |a_ij|, i >= j; ties to the smaller row
    double best = -1.0;
    int bi = n;
    for (int i = j + tid; i < n; i += 256) {
      double v = fabs((double)A[(long)i * n + j]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    sval[tid] = best;
    sidx[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) {
        double v2 = sval[tid + s];
        int i2 = sidx[tid + s];
        if (v2 > sval[tid] || (v2 == sval[tid] && i2 < sidx[tid])) {
          sval[tid] = v2;
          sidx[tid] = i2;
        }
      }
      __syncthreads();
    }
    const int p = sidx[0];
    const double pabs = sval[0];
    __syncthreads();
    if (tid == 0) {
      piv[j] = p;
      if (pabs == 0.0 && info[b] == 0) info[b] = j + 1;
    }
    if (p != j)
      for (int c = tid; c < n; c += 256) {
        T t0 = A[(long)j * n + c];
        A[(long)j * n + c] = A[(long)p * n + c];
        A[(long)p * n + c] = t0;
      }
    __syncthreads();
    const T pv = A[(long)j * n + j];
    if (pv != (T)0)
      for (int i = j + 1 + tid; i < n; i += 256) A[(long)i * n + j] /= pv;
    __syncthreads();
    const int w = k + jb - 1 - j;  // remaining panel columns
    if (w > 0)
      for (long e = tid; e < (long)(n - j - 1) * w; e += 256) {
        int i = j + 1 + (int)(e / w), c = j + 1 + (int)(e % w);
        A[(long)i * n + c] -= A[(long)i * n + j] * A[(long)j * n + c];
      }
    __syncthreads();
  }
}

// Panel of width PNB held in registers: one block of NT threads per matrix
// (NT = 256, 512 or 1024, the fewest that hold the rows left), row
// k + t + q*NT of the panel in thread t, slot q < PR (orders up to PR*PT).
// Per column, with one barrier: each wave finds its largest |a_ij| (a max
// by DPP moves, then the smallest row attaining it by ballots, so ties go
// to the smaller row as in k_lu_panel) and posts it with that row's panel
// part to LDS, and the thread holding the diagonal row posts that row;
// after the barrier every wave reduces the NT/64 posts to the pivot (DPP
// moves again) and every thread does the exchange, the scaling of the
// column and the rank-1 update of the panel's later columns on registers.  Only the panel's columns are swapped
// here; k_lu_swap_trsm applies the panel's interchanges to the other columns
// afterwards, in order, so the factors equal those of whole-row swaps.
// register panel width: 16 columns in FP64, 32 in FP32 (the same bytes of
// registers per thread; half the panels, interchange and update launches)
template <typename T> constexpr int pnb = sizeof(T) == 4 ? 32 : 16;
constexpr int PT = 1024;  // threads of the register panel
constexpr int PR = 3;     // panel rows per thread
constexpr int LOB = 128;  // outer block of the register-panel LU

// DPP move of a double across lanes; lanes the move does not write keep v.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_keep(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf,
                                             false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Wave64 maximum: quads, half-rows, rows, then rows 0..3 into lane 63,
// which readlane gives to every lane.
__device__ __forceinline__ double wave_max(double v)
{
  v = fmax(v, dpp_keep<0xB1, 0xf>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_keep<0x4E, 0xf>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_keep<0x141, 0xf>(v));  // row_half_mirror
  v = fmax(v, dpp_keep<0x140, 0xf>(v));  // row_mirror
  v = fmax(v, dpp_keep<0x142, 0xa>(v));  // row_bcast:15 into rows 1, 3
  v = fmax(v, dpp_keep<0x143, 0xc>(v));  // row_bcast:31 into rows 2, 3
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | lo);
}

// Maximum (minimum) over lanes 0..15 by DPP moves within the first row of
// the wave, read from lane 0.
__device__ __forceinline__ double row16_max(double v)
{
  v = fmax(v, dpp_keep<0xB1, 0xf>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_keep<0x4E, 0xf>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_keep<0x141, 0xf>(v));  // row_half_mirror
  v = fmax(v, dpp_keep<0x140, 0xf>(v));  // row_mirror
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, 0);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), 0);
  return __longlong_as_double(((long long)hi << 32) | lo);
}

__device__ __forceinline__ int row16_min(int v)
{
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xf, 0xf, false));
  return __builtin_amdgcn_readlane(v, 0);
}

// One element of a matrix by a buffer access at a byte offset from its
// first element.  An offset at or beyond the resource's size (the matrix's
// n*n elements) reads 0 and drops the store.
constexpr unsigned BUF_OOR = 0x80000000u;
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t rs, unsigned voff)
{
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0));
}
template <typename T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t rs, unsigned voff, T v)
{
  typedef unsigned u2v __attribute__((ext_vector_type(2)));
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), rs, voff, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rs, voff, 0, 0);
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_lu_panel_reg(T *base, BatchView bv, int *pivb, int *info,
                                                     int k)
{
  constexpr int PNB = pnb<T>;
  // double-buffered by column parity: a buffer is rewritten two columns
  // later, after every thread has passed the barrier between
  __shared__ double wbest[2][NT / 64];
  __shared__ int widx[2][NT / 64];
  __shared__ T wrow[2][NT / 64][PNB];  // each wave's candidate row
  __shared__ T drow[2][PNB];      // the diagonal row before the exchange
  __shared__ T stage[NT * (PNB + 1)];  // one slab of NT rows (at most 136 KiB)
  const int b = blockIdx.x;
  const int n = bv.n[b];
  if (k >= n) return;
  T *A = base + bv.off[b];
  int *piv = pivb + bv.poff[b];
  const int jb = min(PNB, n - k), tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  T a[PR][PNB];
#pragma unroll
  for (int q = 0; q < PR; q++)
#pragma unroll
    for (int c = 0; c < PNB; c++) a[q][c] = (T)0;
  // rows arrive through LDS in slabs of NT rows, one per q, read
  // row-contiguously (PNB consecutive lanes per row) rather than one row per
  // lane; a thread's loads of a slab are all issued before the first wait.
  // Thread t moves column t % PNB of rows t / PNB + m * NT / PNB; elements
  // outside the matrix get an out-of-range offset.
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(A, 0, (int)((long)n * n * sizeof(T)), 0x00020000);
  const int sc = tid % PNB, sr = tid / PNB;
  constexpr int SRS = NT / PNB;  // rows between a thread's elements
  auto soff = [&](int sr, int sc, int r0, int m) {
    const int i = k + r0 + sr + m * SRS;
    return (sc < jb && i < n) ? (unsigned)(((long)i * n + k + sc) * sizeof(T)) : BUF_OOR;
  };
#pragma unroll
  for (int q = 0; q < PR; q++) {
    const int r0 = q * NT;  // first panel row of the slab, from k
    if (k + r0 >= n) break;
    if (q) __syncthreads();  // the previous slab has been read
    T v[PNB];
#pragma unroll
    for (int m = 0; m < PNB; m++) v[m] = buf_ld<T>(rs, soff(sr, sc, r0, m));
#pragma unroll
    for (int m = 0; m < PNB; m++) stage[(sr + m * SRS) * (PNB + 1) + sc] = v[m];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < PNB; c++) a[q][c] = stage[tid * (PNB + 1) + c];
  }
#pragma clang loop unroll(full)
  for (int j = 0; j < PNB; j++) {
    if (j >= jb) continue;  // (a full unroll keeps a[][] in registers)
    const int gj = k + j, par = j & 1;
    // the thread's largest |a_ij| over rows i >= gj, first row on ties
    double best = -1.0;
    int bq = 0;
#pragma unroll
    for (int q = 0; q < PR; q++) {
      const int i = k + tid + q * NT;
      const double v = fabs((double)a[q][j]);
      if (i >= gj && i < n && v > best) {
        best = v;
        bq = q;
      }
    }
    // the wave's: rows ascend with (q, lane), so the first q whose ballot
    // of lanes at the maximum is not empty, and its lowest lane
    const double wm = wave_max(best);
    int qs = -1, ls = 0;
#pragma unroll
    for (int q = PR - 1; q >= 0; q--) {
      const unsigned long long m = __ballot(best == wm && bq == q);
      if (m) {
        qs = q;
        ls = __builtin_ctzll(m);
      }
    }
    if (wm < 0.0) qs = -1;  // no rows of this wave are candidates
#pragma unroll
    for (int q = 0; q < PR; q++)
      if (q == qs && lane == ls) {
#pragma unroll
        for (int c = 0; c < PNB; c++) wrow[par][wv][c] = a[q][c];
      }
    if (lane == 0) {
      wbest[par][wv] = qs < 0 ? -1.0 : wm;
      widx[par][wv] = qs < 0 ? n : k + wv * 64 + ls + qs * NT;
    }
    if (tid == j) {  // row gj = k + j is slot 0 of thread j
#pragma unroll
      for (int c = 0; c < PNB; c++) drow[par][c] = a[0][c];
    }
    __syncthreads();
    // every wave reduces the NT/64 posts itself: lane w reads post w, then
    // the largest value and the smallest row at it, by DPP moves over the
    // first row of 16 lanes (two LDS reads per wave instead of a scan of
    // every post by every thread)
    double pv2 = -2.0;
    int pi2 = 0x7fffffff;
    if (lane < NT / 64) {
      pv2 = wbest[par][lane];
      pi2 = widx[par][lane];
    }
    const double bv2 = row16_max(pv2);
    const int bi = row16_min(pv2 == bv2 ? pi2 : 0x7fffffff);
    const int bw = __builtin_ctzll(__ballot(lane < NT / 64 && pv2 == bv2 && pi2 == bi));
    const int p = bi;  // >= gj: row gj itself is a candidate
    if (tid == 0) {
      piv[gj] = p;
      if (bv2 == 0.0 && info[b] == 0) info[b] = gj + 1;
    }
    const T *prow = wrow[par][bw], *xrow = drow[par];
    // exchange rows gj and p of the panel
#pragma unroll
    for (int q = 0; q < PR; q++) {
      const int i = k + tid + q * NT;
      if (i == gj) {
#pragma unroll
        for (int c = 0; c < PNB; c++) a[q][c] = prow[c];
      } else if (i == p) {
#pragma unroll
        for (int c = 0; c < PNB; c++) a[q][c] = xrow[c];
      }
    }
    // L column and rank-1 update of the later panel columns
    const T pv = prow[j];
#pragma unroll
    for (int q = 0; q < PR; q++) {
      const int i = k + tid + q * NT;
      if (i > gj && i < n) {
        T l = a[q][j];
        if (pv != (T)0) l /= pv;
        a[q][j] = l;
#pragma unroll
        for (int c = 0; c < PNB; c++)  // constant trip count: indices fold after unrolling
          if (c > j) a[q][c] -= l * prow[c];
      }
    }
  }
  // the thread's row and column in a slab, recomputed here rather than
  // kept in registers through the column loop
  int t2 = tid;
  __asm__ volatile("" : "+v"(t2));
  const int sc2 = t2 % PNB, sr2 = t2 / PNB;
#pragma unroll
  for (int q = 0; q < PR; q++) {
    const int r0 = q * NT;
    if (k + r0 >= n) break;
    if (q) __syncthreads();  // the slab before has been stored (the column loop's
                             // barriers separate the first from the loads' reads)
#pragma unroll
    for (int c = 0; c < PNB; c++) stage[tid * (PNB + 1) + c] = a[q][c];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PNB; m++)
      buf_st<T>(rs, soff(sr2, sc2, r0, m), stage[(sr2 + m * SRS) * (PNB + 1) + sc2]);
  }
}

// The row interchanges of the panel at k (pivots k..k+jb-1, in order) on
// every column outside it, one thread per column; then, for the block's
// later columns k + PNB <= c < c1, U of the panel's rows (as
// k_lu_trsm<T, PNB>): the thread of a column applies both, so neither
// waits for another thread.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_swap_trsm(T *base, BatchView bv, const int *pivb,
                                                      int k, int c1, int sw0, int sw1, int cbase)
{
  constexpr int PNB = pnb<T>;
  __shared__ T sL[PNB * PNB];
  const int b = blockIdx.y;
  const int n = bv.n[b];
  if (k >= n) return;  // the whole workgroup
  const int jb = min(PNB, n - k);
  T *A = base + bv.off[b];
  const int *piv = pivb + bv.poff[b];
  const int c = cbase + blockIdx.x * 256 + threadIdx.x;
  if (c < n && c >= sw0 && c < sw1 && !(c >= k && c < k + jb)) {
    for (int j = k; j < k + jb; j++) {
      const int p = piv[j];
      if (p != j) {
        const T t0 = A[(long)j * n + c];
        A[(long)j * n + c] = A[(long)p * n + c];
        A[(long)p * n + c] = t0;
      }
    }
  }
  const int cb = k + PNB, ce = min(c1, n);
  const int w0 = cbase + blockIdx.x * 256;
  if (cb >= ce || w0 + 256 <= cb || w0 >= ce) return;  // the whole workgroup
  for (int e = threadIdx.x; e < jb * jb; e += 256) {
    const int r = e / jb, q = e - r * jb;
    sL[r * PNB + q] = A[(long)(k + r) * n + k + q];
  }
  __syncthreads();
  if (c < cb || c >= ce) return;
  T x[PNB];
#pragma unroll
  for (int r = 0; r < PNB; r++) x[r] = r < jb ? A[(long)(k + r) * n + c] : (T)0;
#pragma unroll
  for (int r = 1; r < PNB; r++) {
    if (r < jb) {
      T sum = x[r];
#pragma unroll
      for (int q = 0; q < r; q++) sum -= sL[r * PNB + q] * x[q];
      x[r] = sum;
      A[(long)(k + r) * n + c] = sum;
    }
  }
}

// The row interchanges of the pivots k0..k1-1, in order, on every column
// outside [k0, k1): the part of a block's interchanges that the look-ahead
// defers until the previous block's trailing update has left those columns.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_swap_out(T *base, BatchView bv, const int *pivb,
                                                     int k0, int k1)
{
  const int b = blockIdx.y;
  const int n = bv.n[b];
  if (k0 >= n) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n || (c >= k0 && c < k1)) return;
  T *A = base + bv.off[b];
  const int *piv = pivb + bv.poff[b];
  const int je = min(k1, n);
  for (int j = k0; j < je; j++) {
    const int p = piv[j];
    if (p != j) {
      const T t0 = A[(long)j * n + c];
      A[(long)j * n + c] = A[(long)p * n + c];
      A[(long)p * n + c] = t0;
    }
  }
}

// U12 = L11^{-1} A12 on rows k..k+jb-1 of the columns c0 <= c < c1 (all
// from k + jb when c1 <= 0), unit lower L11 of order jb <= JB.  One thread
// per column; its column of U12 lives in registers (the loops over the JB
// rows unroll fully, so every index is a constant), L11 in LDS.
template <typename T, int JB>
__global__ void __launch_bounds__(256) k_lu_trsm(T *base, BatchView bv, int k, int c0, int c1)
{
  __shared__ T sL[JB * JB];
  const int b = blockIdx.y;
  const int n = bv.n[b];
  if (k >= n) return;
  const int jb = min(JB, n - k);
  const int cb = c1 > 0 ? c0 : k + jb, ce = c1 > 0 ? min(c1, n) : n;
  if (cb >= ce) return;
  T *A = base + bv.off[b];
  for (int e = threadIdx.x; e < jb * jb; e += 256) {
    const int r = e / jb, q = e - r * jb;
    sL[r * JB + q] = A[(long)(k + r) * n + k + q];
  }
  __syncthreads();
  const int c = cb + blockIdx.x * 256 + threadIdx.x;
  if (c >= ce) return;
  T x[JB];
#pragma unroll
  for (int r = 0; r < JB; r++) x[r] = r < jb ? A[(long)(k + r) * n + c] : (T)0;
#pragma unroll
  for (int r = 1; r < JB; r++) {
    if (r < jb) {
      T s = x[r];
#pragma unroll
      for (int q = 0; q < r; q++) s -= sL[r * JB + q] * x[q];
      x[r] = s;
      A[(long)(k + r) * n + c] = s;
    }
  }
}

// A[r0:r1, c0:c1] -= A[r0:r1, kk:kk+K] A[kk:kk+K, c0:c1], every range
// clipped to the matrix's order: the trailing and in-block updates of the
// LU (L21 U12).
// (FP64 on the matrix cores: at most 128 VGPRs, four waves per SIMD)
template <typename T, bool MF>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && MF ? 4 : 1)))
    k_lu_update(T *base, BatchView bv, int kk, int K, int r0, int r1, int c0, int c1)
{
  __shared__ T sA[BM * (BK + 1)];
  __shared__ T sB[BK * (BN + 1)];
  const int b = blockIdx.z;
  const int n = bv.n[b];
  const int kn = min(K, n - kk), m = min(r1, n) - r0, mc = min(c1, n) - c0;
  if (kn <= 0 || m <= 0 || mc <= 0) return;
  const int ti = blockIdx.y * BM, tj = blockIdx.x * BN;
  if (ti >= m || tj >= mc) return;
  T *A = base + bv.off[b];
  gemm_tile<T, MF>(A + (long)r0 * n + kk, n, A + (long)kk * n + c0, n, A + (long)r0 * n + c0, n,
                   m, mc, kn, (T)-1, ti, tj, sA, sB);
}

// log|det| and sign from the factors: sign = prod sign(u_ii) * (-1)^{#swaps}
template <typename T>
__global__ void __launch_bounds__(256) k_lu_logdet(const T *base, BatchView bv, const int *pivb,
                                                   const int *info, double *logabs, int *sign)
{
  __shared__ double sh[4];
  __shared__ int shn[256];
  const int b = blockIdx.x;
  const int n = bv.n[b];
  const T *A = base + bv.off[b];
  const int *piv = pivb + bv.poff[b];
  double acc = 0.0;
  int neg = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double u = (double)A[(long)i * n + i];
    acc += log(fabs(u));
    neg += (u < 0) + (piv[i] != i);
  }
  double s = block_sum(acc, sh);
  shn[threadIdx.x] = neg;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int q = 0; q < 256; q++) t += shn[q];
    if (info[b]) {
      logabs[b] = -INFINITY;
      sign[b] = 0;
    } else {
      logabs[b] = n ? s : 0.0;
      sign[b] = (t & 1) ? -1 : 1;
    }
  }
}

// ------------------------------------------------------------ inverse ---
// X = P (row permutation of the identity that the pivots define).
template <typename T>
__global__ void __launch_bounds__(256) k_perm_identity(T *X, BatchView bv, const int *pivb,
                                                       int *permb, long pstride)
{
  const int b = blockIdx.y;
  const int n = bv.n[b];
  const int *piv = pivb + bv.poff[b];
  int *perm = permb + (pstride > 0 ? pstride * b : bv.poff[b]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int i = 0; i < n; i++) {
      int p = piv[i];
      int t0 = perm[i];
      perm[i] = perm[p];
      perm[p] = t0;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill_perm(T *X, BatchView bv, const int *permb,
                                                   long pstride)
{
  const int b = blockIdx.y;
  const int n = bv.n[b];
  const int *perm = permb + (pstride > 0 ? pstride * b : bv.poff[b]);
  T *Xb = X + bv.off[b];
  long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    int i = (int)(e / n), c = (int)(e % n);
    Xb[e] = (c == perm[i]) ? (T)1 : (T)0;
  }
}

// Row block [k, k+jb) of X: X_blk = L11^{-1} X_blk (lower unit) or
// U11^{-1} X_blk (upper), all n columns; one thread per column.
template <typename T, bool UPPER>
__global__ void __launch_bounds__(256) k_trsm_rows(const T *LU, T *X, BatchView bv, int k)
{
  __shared__ T sT[NB * NB];
  const int b = blockIdx.y;
  const int n = bv.n[b];
  if (k >= n) return;
  const int jb = min(NB, n - k);
  const T *F = LU + bv.off[b];
  T *Xb = X + bv.off[b];
  for (int e = threadIdx.x; e < jb * jb; e += 256) {
    int r = e / jb, q = e - r * jb;
    sT[r * NB + q] = F[(long)(k + r) * n + k + q];
  }
  __syncthreads();
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  T x[NB];
  for (int r = 0; r < jb; r++) x[r] = Xb[(long)(k + r) * n + c];
  if (!UPPER) {
    for (int r = 1; r < jb; r++) {
      T s = x[r];
      for (int q = 0; q < r; q++) s -= sT[r * NB + q] * x[q];
      x[r] = s;
    }
  } else {
    for (int r = jb - 1; r >= 0; r--) {
      T s = x[r];
      for (int q = r + 1; q < jb; q++) s -= sT[r * NB + q] * x[q];
      x[r] = s / sT[r * NB + r];
    }
  }
  for (int r = 0; r < jb; r++) Xb[(long)(k + r) * n + c] = x[r];
}

// Forward: X[k+jb:, :] -= L[k+jb:, k:k+jb] X[k:k+jb, :]
// Backward: X[0:k, :] -= U[0:k, k:k+jb] X[k:k+jb, :]
template <typename T, bool MF, bool BACK>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && MF ? 4 : 1)))
    k_inv_update(const T *LU, T *X, BatchView bv, int k)
{
  __shared__ T sA[BM * (BK + 1)];
  __shared__ T sB[BK * (BN + 1)];
  const int b = blockIdx.z;
  const int n = bv.n[b];
  if (k >= n) return;
  const int jb = min(NB, n - k);
  const int m = BACK ? k : n - k - jb;
  if (m <= 0) return;
  const int ti = blockIdx.y * BM, tj = blockIdx.x * BN;
  if (ti >= m || tj >= n) return;
  const T *F = LU + bv.off[b];
  T *Xb = X + bv.off[b];
  const int r0 = BACK ? 0 : k + jb;
  gemm_tile<T, MF>(F + (long)r0 * n + k, n, Xb + (long)k * n, n, Xb + (long)r0 * n, n, m, n, jb,
                   (T)-1, ti, tj, sA, sB);
}

// generic batched GEMM: C += alpha A B (jobs in device memory)
template <typename T, bool MF>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 && MF ? 4 : 1)))
    k_gemm_jobs(T *base, const th2d_gemm_job *jobs, double alpha)
{
  __shared__ T sA[BM * (BK + 1)];
  __shared__ T sB[BK * (BN + 1)];
  const th2d_gemm_job J = jobs[blockIdx.z];
  const int ti = blockIdx.y * BM, tj = blockIdx.x * BN;
  if (ti >= J.M || tj >= J.N) return;
  gemm_tile<T, MF>(base + J.a, J.lda, base + J.b, J.ldb, base + J.c, J.ldc, J.M, J.N, J.K,
                   (T)alpha, ti, tj, sA, sB);
}

__global__ void k_f64_to_f32_n(float *d, const double *s, long n)
{
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    d[i] = (float)s[i];
}

// ---- FP64 correction of an FP32 LU's log-determinant ---------------------
// The promoted factors: L (unit lower) and U (upper) as dense FP64
// matrices, and G = 0, all in the batch layout.
__global__ void __launch_bounds__(256) k_expand_factors(const float *LU32, BatchView bv, double *L,
                                                        double *U, double *G)
{
  const int b = blockIdx.y;
  const long n = bv.n[b], off = bv.off[b];
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n * n; e += (long)gridDim.x * 256) {
    const long i = e / n, j = e - i * n;
    const double v = (double)LU32[off + e];
    L[off + e] = i > j ? v : (i == j ? 1.0 : 0.0);
    U[off + e] = i <= j ? v : 0.0;
    G[off + e] = 0.0;
  }
}

// Partial sums of tr(X E) = sum_{r,i} X[i][perm[r]] (A[perm[r]][i] - G[r][i])
// with E = A - P L U the FP64 residual of the FP32 factorisation (rows of
// A permuted as the factorisation permuted them: row r of P^T A is
// A[perm[r]]) and X the FP32 inverse of P L U.
__global__ void __launch_bounds__(256) k_trace_correction(const double *A, const float *X,
                                                          const double *G, const int *permb,
                                                          BatchView bv, double *part, int nparts)
{
  __shared__ double sh[4];
  const int b = blockIdx.y;
  const long n = bv.n[b], off = bv.off[b];
  const int *perm = permb + bv.poff[b];
  double acc = 0.0;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n * n; e += (long)gridDim.x * 256) {
    const long r = e / n, i = e - r * n, pr = perm[r];
    acc += (double)X[off + i * n + pr] * (A[off + pr * n + i] - G[off + e]);
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(long)b * nparts + blockIdx.x] = s;
}

__global__ void k_add_correction(const double *part, int nparts, int count, const int *info,
                                 double *logabs)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= count || info[b]) return;
  double s = 0.0;
  for (int q = 0; q < nparts; q++) s += part[(long)b * nparts + q];
  logabs[b] += s;
}

static int g_use_mfma = -1;
static int g_lu_reg_panel = 1;  // register panel when it fits (th2d_set_lu_panel)
bool use_mfma()
{
  if (g_use_mfma < 0) {
    const char *e = getenv("TH2D_DENSE_MFMA");
    g_use_mfma = e ? atoi(e) != 0 : 1;
  }
  return g_use_mfma != 0;
}

BatchView view(const th2d_batch *b) { return BatchView{b->off, b->n, b->poff, b->count}; }

template <typename T>
void launch_update(th2d_ctx *c, bool mf, T *A, BatchView bv, int count, int kk, int K, int r0,
                   int r1, int c0, int c1, int N, hipStream_t st = nullptr)
{
  const int rows = std::min(r1, N) - r0, cols = std::min(c1, N) - c0;
  if (rows <= 0 || cols <= 0) return;
  if (!st) st = c->stream;
  dim3 g((cols + BN - 1) / BN, (rows + BM - 1) / BM, count);
  if (mf)
    hipLaunchKernelGGL((k_lu_update<T, true>), g, dim3(256), 0, st, A, bv, kk, K, r0, r1, c0, c1);
  else
    hipLaunchKernelGGL((k_lu_update<T, false>), g, dim3(256), 0, st, A, bv, kk, K, r0, r1, c0, c1);
}

static int g_lu_lookahead = 0;  // th2d_set_lu_lookahead (measured slower: off)

// The two streams of the look-ahead: panels on a few CUs of every XCD,
// updates on the rest, so that neither waits for the other's workgroups.
// CU-mask bit i is taken for the panels when i mod 32 < 8: 8 of every 32
// CUs, 64 of 256, whether the runtime numbers the CUs XCD by XCD or
// round-robin over the XCDs.  0 when the streams cannot be made.
int lu_streams(th2d_ctx *c, hipStream_t *ps, hipStream_t *us)
{
  if (c->lu_stream_state == 0) {
    c->lu_stream_state = -1;
    const int ncu = c->num_cu > 0 ? c->num_cu : 256;
    const int words = (ncu + 31) / 32;
    std::vector<uint32_t> mp(words, 0), mu(words, 0);
    for (int i = 0; i < ncu; i++) ((i % 32) < 8 ? mp : mu)[i / 32] |= 1u << (i % 32);
    if (hipExtStreamCreateWithCUMask(&c->lu_ps, words, mp.data()) == hipSuccess &&
        hipExtStreamCreateWithCUMask(&c->lu_us, words, mu.data()) == hipSuccess &&
        hipEventCreateWithFlags(&c->lu_ev[0], hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&c->lu_ev[1], hipEventDisableTiming) == hipSuccess &&
        hipEventCreateWithFlags(&c->lu_ev[2], hipEventDisableTiming) == hipSuccess)
      c->lu_stream_state = 1;
  }
  *ps = c->lu_ps;
  *us = c->lu_us;
  return c->lu_stream_state == 1;
}

template <typename T>
int lu_run(th2d_ctx *c, const th2d_batch *bt, T *A, int *piv, int *info)
{
  BatchView bv = view(bt);
  TH2D_TRY(th2d_memset(c, info, 0, sizeof(int) * bt->count));
  const bool mf = use_mfma();
  const int N = bt->max_n, cnt = bt->count;
  constexpr int END = 1 << 30;  // a range bound past every order
  constexpr int PNB = pnb<T>;
  hipStream_t ps = nullptr, us = nullptr;
  if (bt->max_n <= PR * PT && g_lu_reg_panel && g_lu_lookahead && N > LOB &&
      lu_streams(c, &ps, &us)) {
    // The same blocks with a look-ahead.  Panel stream: a block's panels,
    // their interchanges inside the block, its 16-row solves and updates.
    // Update stream: the block's interchanges on the other columns, U12,
    // the trailing update of the next block's columns -- after which the
    // panel stream factors that block -- and then the rest of the trailing
    // update, which runs beside those panels.
    hipEvent_t e_start = c->lu_ev[0], e_panel = c->lu_ev[1], e_next = c->lu_ev[2];
    TH2D_HIP(hipEventRecord(e_start, c->stream));
    TH2D_HIP(hipStreamWaitEvent(ps, e_start, 0));
    TH2D_HIP(hipStreamWaitEvent(us, e_start, 0));
    for (int k = 0; k < N; k += LOB) {
      const int ce = std::min(k + LOB, N);
      for (int kk = k; kk < ce; kk += PNB) {
        const int rows = N - kk;
        if (rows <= PR * 256)
          hipLaunchKernelGGL((k_lu_panel_reg<T, 256>), dim3(cnt), dim3(256), 0, ps, A, bv, piv,
                             info, kk);
        else if (rows <= PR * 512)
          hipLaunchKernelGGL((k_lu_panel_reg<T, 512>), dim3(cnt), dim3(512), 0, ps, A, bv, piv,
                             info, kk);
        else
          hipLaunchKernelGGL((k_lu_panel_reg<T, PT>), dim3(cnt), dim3(PT), 0, ps, A, bv, piv,
                             info, kk);
        hipLaunchKernelGGL(k_lu_swap_trsm<T>, dim3((ce - k + 255) / 256, cnt), dim3(256), 0, ps,
                           A, bv, piv, kk, ce, k, ce, k);
        const int cb = kk + PNB;
        if (cb < ce) launch_update<T>(c, mf, A, bv, cnt, kk, PNB, cb, END, cb, ce, N, ps);
      }
      TH2D_HIP(hipEventRecord(e_panel, ps));
      TH2D_HIP(hipStreamWaitEvent(us, e_panel, 0));
      hipLaunchKernelGGL(k_lu_swap_out<T>, dim3((N + 255) / 256, cnt), dim3(256), 0, us, A, bv,
                         piv, k, ce);
      if (ce >= N) break;  // (the last block's interchanges on the earlier columns: above)
      for (int s = k; s < ce; s += NB) {
        hipLaunchKernelGGL((k_lu_trsm<T, NB>), dim3((N - ce + 255) / 256, cnt), dim3(256), 0, us,
                           A, bv, s, ce, END);
        if (s + NB < ce) launch_update<T>(c, mf, A, bv, cnt, s, NB, s + NB, ce, ce, END, N, us);
      }
      const int ne = std::min(ce + LOB, N);
      launch_update<T>(c, mf, A, bv, cnt, k, LOB, ce, END, ce, ne, N, us);
      TH2D_HIP(hipEventRecord(e_next, us));
      TH2D_HIP(hipStreamWaitEvent(ps, e_next, 0));
      if (ne < N) launch_update<T>(c, mf, A, bv, cnt, k, LOB, ce, END, ne, END, N, us);
    }
    TH2D_HIP(hipEventRecord(e_panel, ps));
    TH2D_HIP(hipStreamWaitEvent(c->stream, e_panel, 0));
    TH2D_HIP(hipEventRecord(e_next, us));
    TH2D_HIP(hipStreamWaitEvent(c->stream, e_next, 0));
  } else if (bt->max_n <= PR * PT && g_lu_reg_panel) {
    // blocks of LOB columns.  Inside a block: register panels of PNB
    // columns, each followed by its row interchanges on the other columns,
    // U of its rows for the block's later columns and their update.  After
    // the block: U12 of the block's rows for every later column, by
    // triangular solves on NB-row slices with the updates between them,
    // then one trailing update A22 -= L21 U12 of depth LOB.
    for (int k = 0; k < N; k += LOB) {
      const int ce = std::min(k + LOB, N);
      for (int kk = k; kk < ce; kk += PNB) {
        const int rows = N - kk;  // the most rows a matrix has left
        if (rows <= PR * 256)
          hipLaunchKernelGGL((k_lu_panel_reg<T, 256>), dim3(cnt), dim3(256), 0, c->stream, A, bv,
                             piv, info, kk);
        else if (rows <= PR * 512)
          hipLaunchKernelGGL((k_lu_panel_reg<T, 512>), dim3(cnt), dim3(512), 0, c->stream, A, bv,
                             piv, info, kk);
        else
          hipLaunchKernelGGL((k_lu_panel_reg<T, PT>), dim3(cnt), dim3(PT), 0, c->stream, A, bv,
                             piv, info, kk);
        hipLaunchKernelGGL(k_lu_swap_trsm<T>, dim3((N + 255) / 256, cnt), dim3(256), 0,
                           c->stream, A, bv, piv, kk, ce, 0, END, 0);
        const int cb = kk + PNB;
        if (cb < ce) launch_update<T>(c, mf, A, bv, cnt, kk, PNB, cb, END, cb, ce, N);
      }
      if (ce >= N) break;
      for (int s = k; s < ce; s += NB) {
        hipLaunchKernelGGL((k_lu_trsm<T, NB>), dim3((N - ce + 255) / 256, cnt), dim3(256), 0,
                           c->stream, A, bv, s, ce, END);
        if (s + NB < ce) launch_update<T>(c, mf, A, bv, cnt, s, NB, s + NB, ce, ce, END, N);
      }
      launch_update<T>(c, mf, A, bv, cnt, k, LOB, ce, END, ce, END, N);
    }
  } else {
    // the LDS-reduction panel of width NB, which swaps whole rows itself
    for (int k = 0; k < N; k += NB) {
      hipLaunchKernelGGL(k_lu_panel<T>, dim3(cnt), dim3(256), 0, c->stream, A, bv, piv, info, k);
      const int rest = N - k - NB;
      if (rest > 0) {
        hipLaunchKernelGGL((k_lu_trsm<T, NB>), dim3((rest + 255) / 256, cnt), dim3(256), 0,
                           c->stream, A, bv, k, 0, 0);
        launch_update<T>(c, mf, A, bv, cnt, k, NB, k + NB, END, k + NB, END, N);
      }
    }
  }
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// The substitutions of the inverse after the permutation scratch is set:
// perm holds each matrix's permutation at pivot offset poff[b]
// (pstride 0) or at b * pstride.
template <typename T>
int inverse_launch(th2d_ctx *c, const th2d_batch *bt, const T *LU, const int *piv, T *X,
                   int *perm, long pstride)
{
  BatchView bv = view(bt);
  const int N = bt->max_n;
  const bool mf = use_mfma();
  hipLaunchKernelGGL(k_perm_identity<T>, dim3(1, bt->count), dim3(256), 0, c->stream, X, bv, piv,
                     perm, pstride);
  hipLaunchKernelGGL(k_fill_perm<T>, dim3(std::min(1024, (N * N + 255) / 256), bt->count),
                     dim3(256), 0, c->stream, X, bv, perm, pstride);
  // forward substitution with L (unit lower)
  for (int k = 0; k < N; k += NB) {
    hipLaunchKernelGGL((k_trsm_rows<T, false>), dim3((N + 255) / 256, bt->count), dim3(256), 0,
                       c->stream, LU, X, bv, k);
    int rest = N - k - NB;
    if (rest > 0) {
      dim3 g((N + BN - 1) / BN, (rest + BM - 1) / BM, bt->count);
      if (mf) hipLaunchKernelGGL((k_inv_update<T, true, false>), g, dim3(256), 0, c->stream, LU,
                                 X, bv, k);
      else hipLaunchKernelGGL((k_inv_update<T, false, false>), g, dim3(256), 0, c->stream, LU, X,
                              bv, k);
    }
  }
  // back substitution with U: the last block first
  int klast = ((N - 1) / NB) * NB;
  for (int k = klast; k >= 0; k -= NB) {
    hipLaunchKernelGGL((k_trsm_rows<T, true>), dim3((N + 255) / 256, bt->count), dim3(256), 0,
                       c->stream, LU, X, bv, k);
    if (k > 0) {
      dim3 g((N + BN - 1) / BN, (k + BM - 1) / BM, bt->count);
      if (mf) hipLaunchKernelGGL((k_inv_update<T, true, true>), g, dim3(256), 0, c->stream, LU, X,
                                 bv, k);
      else hipLaunchKernelGGL((k_inv_update<T, false, true>), g, dim3(256), 0, c->stream, LU, X,
                              bv, k);
    }
  }
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

template <typename T>
int inverse_run(th2d_ctx *c, const th2d_batch *bt, const T *LU, const int *piv, T *X)
{
  // scratch: permutation per matrix (same offsets as the pivots)
  long total = 0;
  std::vector<long> h(bt->count);
  TH2D_HIP(hipMemcpyAsync(h.data(), bt->poff, sizeof(long) * bt->count, hipMemcpyDeviceToHost,
                          c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < bt->count; i++) total = std::max(total, h[i]);
  TH2D_TRY(th2d_buf_reserve(c->bag, sizeof(int) * (size_t)(total + bt->max_n + 64)));
  return inverse_launch<T>(c, bt, LU, piv, X, (int *)c->bag.ptr, 0);
}

}  // namespace

// ---- the allocation-free inverse ------------------------------------------
// Its permutation scratch is a file-scope device array: each GPU holds its
// own copy until the process ends.  Host code uses the per-device address
// from hipGetSymbolAddress (looked up when a context is created on that
// device), never the array's name.  Calls that overlap on one GPU share
// the array and race.
constexpr long kNoAllocPermInts = 1L << 22;  // 16 MiB: count * max_n <= 4,194,304
static __device__ __align__(256) int scratch[kNoAllocPermInts];
static std::mutex g_scratch_mu;
static int *g_scratch_ptr[64];

namespace th2d {
int noalloc_scratch_init(int device)
{
  if (device < 0 || device >= 64) return TH2D_EARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return TH2D_OK;
  void *p = nullptr;
  TH2D_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = (int *)p;
  return TH2D_OK;
}
}  // namespace th2d

static int *noalloc_scratch(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < 64) ? g_scratch_ptr[device] : nullptr;
}


// ================================================================ ABI ===
extern "C" int th2d_lu_batched(th2d_ctx *c, const th2d_batch *b, void *A, int precision, int *piv,
                               int *info)
{
  if (!c || !b || !A || !piv || !info || b->count < 1 || b->max_n < 0) return TH2D_EARG;
  if (precision == TH2D_PREC_F64) return lu_run<double>(c, b, (double *)A, piv, info);
  if (precision == TH2D_PREC_F32) return lu_run<float>(c, b, (float *)A, piv, info);
  return TH2D_EARG;
}

extern "C" int th2d_lu_logdet(th2d_ctx *c, const th2d_batch *b, const void *LU, int precision,
                              const int *piv, const int *info, double *logabs, int *sign)
{
  if (!c || !b || !LU || !piv || !info || !logabs || !sign || b->count < 1) return TH2D_EARG;
  BatchView bv = view(b);
  if (precision == TH2D_PREC_F64)
    hipLaunchKernelGGL(k_lu_logdet<double>, dim3(b->count), dim3(256), 0, c->stream,
                       (const double *)LU, bv, piv, info, logabs, sign);
  else if (precision == TH2D_PREC_F32)
    hipLaunchKernelGGL(k_lu_logdet<float>, dim3(b->count), dim3(256), 0, c->stream,
                       (const float *)LU, bv, piv, info, logabs, sign);
  else return TH2D_EARG;
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

// log|det A| of FP32 factors, corrected in FP64: log det A = log det B +
// tr(B^-1 E) + O(|B^-1 E|^2) with B = P L U (the FP32 factors) and
// E = A - B formed in FP64 from the promoted factors (a triangle-aware
// MFMA product: the tile (I, J) of L U needs depth min(I, J) + 64 only)
// and B^-1 the FP32 inverse from the same factors.
extern "C" int th2d_lu_logdet_refined(th2d_ctx *c, const th2d_batch *b, const double *A,
                                      const float *LU32, const int *piv, const int *info,
                                      double *logabs, int *sign)
{
  if (!c || !b || !A || !LU32 || !piv || !info || !logabs || !sign || b->count < 1)
    return TH2D_EARG;
  TH2D_TRY(th2d_lu_logdet(c, b, LU32, TH2D_PREC_F32, piv, info, logabs, sign));
  const int count = b->count;
  std::vector<long> off(count), poff(count);
  std::vector<int> n(count);
  TH2D_HIP(hipMemcpyAsync(off.data(), b->off, sizeof(long) * count, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipMemcpyAsync(poff.data(), b->poff, sizeof(long) * count, hipMemcpyDeviceToHost,
                          c->stream));
  TH2D_HIP(hipMemcpyAsync(n.data(), b->n, sizeof(int) * count, hipMemcpyDeviceToHost, c->stream));
  TH2D_HIP(hipStreamSynchronize(c->stream));
  long total = 0, ptotal = 0;
  for (int q = 0; q < count; q++) {
    total = std::max(total, off[q] + (long)n[q] * n[q]);
    ptotal = std::max(ptotal, poff[q] + (long)n[q]);
  }
  // per-tile GEMM jobs of G = L U
  std::vector<th2d_gemm_job> jobs;
  for (int q = 0; q < count; q++)
    for (int I = 0; I < n[q]; I += BM)
      for (int J = 0; J < n[q]; J += BN) {
        th2d_gemm_job j;
        j.M = std::min(BM, n[q] - I);
        j.N = std::min(BN, n[q] - J);
        j.K = std::min(n[q], std::min(I, J) + BM);
        j.a = total + off[q] + (long)I * n[q];             // L rows I..
        j.b = 2 * total + off[q] + J;                        // U columns J..
        j.c = off[q] + (long)I * n[q] + J;                   // G
        j.lda = j.ldb = j.ldc = n[q];
        jobs.push_back(j);
      }
  const int nparts = 64;
  // scratch: G, L, U (FP64, `total` each), X (FP32), jobs, perm, partials
  const size_t bytes = sizeof(double) * 3 * (size_t)total + sizeof(float) * (size_t)total +
                       sizeof(th2d_gemm_job) * jobs.size() + sizeof(int) * (size_t)(ptotal + 64) +
                       sizeof(double) * (size_t)count * nparts + 1024;
  TH2D_TRY(th2d_buf_reserve(c->misc, bytes));
  char *p = (char *)c->misc.ptr;
  double *G = (double *)p;
  double *L = G + total, *U = L + total;
  float *X = (float *)(U + total);
  th2d_gemm_job *djobs = (th2d_gemm_job *)(((uintptr_t)(X + total) + 255) & ~(uintptr_t)255);
  int *perm = (int *)(djobs + jobs.size());
  double *part = (double *)(((uintptr_t)(perm + ptotal + 64) + 255) & ~(uintptr_t)255);
  BatchView bv = view(b);
  const int N = b->max_n;
  TH2D_HIP(hipMemcpyAsync(djobs, jobs.data(), sizeof(th2d_gemm_job) * jobs.size(),
                          hipMemcpyHostToDevice, c->stream));
  TH2D_TRY(inverse_launch<float>(c, b, LU32, piv, X, perm, 0));
  hipLaunchKernelGGL(k_expand_factors, dim3(std::min(2048, (N * N + 255) / 256), count), dim3(256),
                     0, c->stream, LU32, bv, L, U, G);
  const bool mf = use_mfma();
  for (size_t j0 = 0; j0 < jobs.size(); j0 += 65535) {
    const unsigned nj = (unsigned)std::min<size_t>(65535, jobs.size() - j0);
    if (mf)
      hipLaunchKernelGGL((k_gemm_jobs<double, true>), dim3(1, 1, nj), dim3(256), 0, c->stream, G,
                         djobs + j0, 1.0);
    else
      hipLaunchKernelGGL((k_gemm_jobs<double, false>), dim3(1, 1, nj), dim3(256), 0, c->stream, G,
                         djobs + j0, 1.0);
  }
  // perm (inverse_launch left it there) pairs each row of L U with its row of A
  hipLaunchKernelGGL(k_trace_correction, dim3(nparts, count), dim3(256), 0, c->stream, A, X, G,
                     perm, bv, part, nparts);
  hipLaunchKernelGGL(k_add_correction, dim3((count + 63) / 64), dim3(64), 0, c->stream, part,
                     nparts, count, info, logabs);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_lu_inverse(th2d_ctx *c, const th2d_batch *b, const void *LU, int precision,
                               const int *piv, void *X)
{
  if (!c || !b || !LU || !piv || !X || LU == X || b->count < 1) return TH2D_EARG;
  if (precision == TH2D_PREC_F64) return inverse_run<double>(c, b, (const double *)LU, piv, (double *)X);
  if (precision == TH2D_PREC_F32) return inverse_run<float>(c, b, (const float *)LU, piv, (float *)X);
  return TH2D_EARG;
}

extern "C" int th2d_lu_inverse_noalloc(th2d_ctx *c, const th2d_batch *b, const void *LU,
                                       int precision, const int *piv, void *X)
{
  if (!c || !b || !LU || !piv || !X || LU == X || b->count < 1 || b->max_n < 1) return TH2D_EARG;
  if ((long)b->count * b->max_n > kNoAllocPermInts) return TH2D_EARG;  // beyond the scratch
  int *perm = noalloc_scratch(c->device);
  if (!perm) return TH2D_EARG;  // no context was created on this device
  if (precision == TH2D_PREC_F64)
    return inverse_launch<double>(c, b, (const double *)LU, piv, (double *)X, perm, b->max_n);
  if (precision == TH2D_PREC_F32)
    return inverse_launch<float>(c, b, (const float *)LU, piv, (float *)X, perm, b->max_n);
  return TH2D_EARG;
}

extern "C" int th2d_gemm_batched(th2d_ctx *c, void *base, int precision,
                                 const th2d_gemm_job *jobs_dev, int count, int max_m, int max_n,
                                 double alpha)
{
  if (!c || !base || !jobs_dev || count < 1 || max_m < 0 || max_n < 0) return TH2D_EARG;
  if (!max_m || !max_n) return TH2D_OK;
  dim3 g((max_n + BN - 1) / BN, (max_m + BM - 1) / BM, count);
  const bool mf = use_mfma();
  if (precision == TH2D_PREC_F64) {
    if (mf) hipLaunchKernelGGL((k_gemm_jobs<double, true>), g, dim3(256), 0, c->stream, (double *)base, jobs_dev, alpha);
    else hipLaunchKernelGGL((k_gemm_jobs<double, false>), g, dim3(256), 0, c->stream, (double *)base, jobs_dev, alpha);
  } else if (precision == TH2D_PREC_F32) {
    if (mf) hipLaunchKernelGGL((k_gemm_jobs<float, true>), g, dim3(256), 0, c->stream, (float *)base, jobs_dev, alpha);
    else hipLaunchKernelGGL((k_gemm_jobs<float, false>), g, dim3(256), 0, c->stream, (float *)base, jobs_dev, alpha);
  } else return TH2D_EARG;
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

extern "C" int th2d_set_dense_mfma(int on)
{
  g_use_mfma = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_lu_lookahead(int on)
{
  g_lu_lookahead = on ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_set_lu_panel(int reg)
{
  g_lu_reg_panel = reg ? 1 : 0;
  return TH2D_OK;
}

extern "C" int th2d_copy_f64_to_f32(th2d_ctx *c, float *d, const double *s, long n)
{
  if (!c || (!d && n) || (!s && n) || n < 0) return TH2D_EARG;
  if (!n) return TH2D_OK;
  long blocks = std::min(4096L, (n + 255) / 256);
  hipLaunchKernelGGL(k_f64_to_f32_n, dim3((unsigned)blocks), dim3(256), 0, c->stream, d, s, n);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}
