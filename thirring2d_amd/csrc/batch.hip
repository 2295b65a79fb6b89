// batch.hip -- per-site operations on every chain of a batch in one launch
// (the field updates, operator applications, forces, vector updates and
// reductions around the solver in update_gauge, hmc.c:671-746, and in the
// FP64 outer loop of the solver).
//
// A launch covers sum_k ceil(V_k / 256) blocks; block b belongs to the
// chain whose first block blk0 is the largest <= b, and handles 256 of its
// eo sites.  Reductions are two-pass and deterministic: one partial per
// block, then one workgroup per chain adds its partials in block order.
#include "th2d_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

using namespace th2d;
constexpr int BT = 256;

template <int OP>
__global__ void __launch_bounds__(BT) k_bop(const BOp *__restrict__ ops, int nch,
                                            double *__restrict__ part)
{
  const int k = bop_chain(ops, nch, blockIdx.x);
  const BOp o = ops[k];
  const int V = 2 * o.G.vh;
  const int i = (int)(blockIdx.x - o.blk0) * BT + threadIdx.x;
  const bool in = i < V;
  if (OP == BOP_DOT || OP == BOP_COSGAP || OP == BOP_RESID) {
    __shared__ double sh[8];
    double acc = 0.0;
    if (in) {
      if (OP == BOP_DOT) {
        const double2 a = ((const double2 *)o.d0)[i], b = ((const double2 *)o.d1)[i];
        acc = a.x * b.x + a.y * b.y;
      } else if (OP == BOP_RESID) {
        const double2 b = ((const double2 *)o.d1)[i];
        const double2 m = dslash_site(o.G, (const double2 *)o.d3, (const double2 *)o.d2, i, 1);
        const double2 r = make_double2(b.x - m.x, b.y - m.y);
        ((double2 *)o.d0)[i] = r;
        acc = r.x * r.x + r.y * r.y;
      } else {
        acc = (1.0 - cos(o.d0[i])) + (1.0 - cos(o.d0[V + i]));
      }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    return;
  }
  if (!in) return;
  if (OP == BOP_GAUGE_STEP) {
    o.d0[i] += 2.0 * o.s0 * o.d1[i];
    o.d0[V + i] += 2.0 * o.s0 * o.d1[V + i];
    fold_one(o.G, o.bc, i, o.d0, o.d2, o.f0);
  } else if (OP == BOP_FOLD) {
    fold_one(o.G, o.bc, i, o.d0, o.d2, o.f0);
  } else if (OP == BOP_DSLASH) {
    ((double2 *)o.d0)[i] =
        dslash_site(o.G, (const double2 *)o.d1, (const double2 *)o.d2, i, o.s1 != 0.0);
  } else if (OP == BOP_FORCE) {
    double F[2];
    force_site(o.G, o.bc, (const double2 *)o.d2, (const double2 *)o.d1, (const double2 *)o.d3,
               (const double2 *)o.d4, o.s1, i, F);
    o.d0[i] -= o.s0 * F[0];
    o.d0[V + i] -= o.s0 * F[1];
  } else if (OP == BOP_AXPY) {
    double2 y = ((double2 *)o.d0)[i];
    const double2 x = ((const double2 *)o.d1)[i];
    ((double2 *)o.d0)[i] = make_double2(y.x + o.s0 * x.x, y.y + o.s0 * x.y);
  } else if (OP == BOP_SUB) {
    const double2 a = ((const double2 *)o.d1)[i], b = ((const double2 *)o.d2)[i];
    ((double2 *)o.d0)[i] = make_double2(a.x - b.x, a.y - b.y);
  } else if (OP == BOP_TO_F32) {
    const double s0 = o.sp ? 1.0 / sqrt(*o.sp) : o.s0;
    const double2 a = ((const double2 *)o.d1)[i];
    ((float2 *)o.f0)[i] = make_float2((float)(a.x * s0), (float)(a.y * s0));
  } else if (OP == BOP_ADD_F32) {
    const double s0 = o.sp ? sqrt(*o.sp) : o.s0;
    double2 y = ((double2 *)o.d0)[i];
    const float2 x = ((const float2 *)o.f0)[i];
    ((double2 *)o.d0)[i] = make_double2(y.x + s0 * (double)x.x, y.y + s0 * (double)x.y);
  } else if (OP == BOP_COPY) {
    o.d0[i] = o.d1[i];
    o.d0[V + i] = o.d1[V + i];
  } else if (OP == BOP_SCALE) {
    const double2 a = ((const double2 *)o.d1)[i];
    ((double2 *)o.d0)[i] = make_double2(o.s0 * a.x, o.s0 * a.y);
  }
}

// One workgroup per chain: out = s1 * (sum of the chain's block partials).
__global__ void __launch_bounds__(BT) k_bop_finish(const BOp *__restrict__ ops, int nch,
                                                   long nblk, const double *__restrict__ part)
{
  __shared__ double sh[8];
  const int k = blockIdx.x;
  const BOp o = ops[k];
  const long end = k + 1 < nch ? ops[k + 1].blk0 : nblk;
  double acc = 0.0;
  for (long b = o.blk0 + threadIdx.x; b < end; b += BT) acc += part[b];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    *o.out = s * o.s1;
    if (o.d4) *o.d4 = s * o.s1;  // a second copy (RESID; d4 is unused by the other reductions)
  }
}

template <int OP>
void launch(th2d_ctx *c, const BOp *dev, int nch, long nblk, double *part)
{
  hipLaunchKernelGGL(k_bop<OP>, dim3((unsigned)nblk), dim3(BT), 0, c->stream, dev, nch, part);
}

}  // namespace

namespace th2d {

long bop_finish(std::vector<BOp> &ops)
{
  long b = 0;
  for (auto &o : ops) {
    o.blk0 = b;
    b += (2L * o.G.vh + BT - 1) / BT;
  }
  return b;
}

int bop_upload(th2d_ctx *c, th2d_buf &dev, const std::vector<BOp> &ops)
{
  for (int k = 0; k < 4; k++)
    if (&dev == &c->btab[k]) c->btab_key[k].clear();  // its contents change
  const size_t bytes = sizeof(BOp) * ops.size();
  TH2D_TRY(th2d_buf_reserve(dev, bytes));
  return stage_h2d(c, dev.ptr, ops.data(), bytes);
}

int bop_upload_key(th2d_ctx *c, int slot, const std::vector<BOp> &ops)
{
  const size_t bytes = sizeof(BOp) * ops.size();
  const unsigned char *raw = (const unsigned char *)ops.data();
  std::vector<unsigned char> &key = c->btab_key[slot];
  if (key.size() == bytes && memcmp(key.data(), raw, bytes) == 0) return TH2D_OK;
  TH2D_TRY(bop_upload(c, c->btab[slot], ops));
  key.assign(raw, raw + bytes);
  return TH2D_OK;
}

int bop_launch(th2d_ctx *c, int op, const BOp *dev, int nch, long nblk)
{
  if (nch < 1 || nblk < 1) return TH2D_OK;
  double *part = nullptr;
  const bool red = op == BOP_DOT || op == BOP_COSGAP || op == BOP_RESID;
  if (red) {
    TH2D_TRY(th2d_buf_reserve(c->bpart, sizeof(double) * (size_t)nblk));
    part = (double *)c->bpart.ptr;
  }
  switch (op) {
  case BOP_GAUGE_STEP: launch<BOP_GAUGE_STEP>(c, dev, nch, nblk, part); break;
  case BOP_FOLD: launch<BOP_FOLD>(c, dev, nch, nblk, part); break;
  case BOP_DSLASH: launch<BOP_DSLASH>(c, dev, nch, nblk, part); break;
  case BOP_FORCE: launch<BOP_FORCE>(c, dev, nch, nblk, part); break;
  case BOP_AXPY: launch<BOP_AXPY>(c, dev, nch, nblk, part); break;
  case BOP_SUB: launch<BOP_SUB>(c, dev, nch, nblk, part); break;
  case BOP_TO_F32: launch<BOP_TO_F32>(c, dev, nch, nblk, part); break;
  case BOP_ADD_F32: launch<BOP_ADD_F32>(c, dev, nch, nblk, part); break;
  case BOP_COPY: launch<BOP_COPY>(c, dev, nch, nblk, part); break;
  case BOP_SCALE: launch<BOP_SCALE>(c, dev, nch, nblk, part); break;
  case BOP_DOT: launch<BOP_DOT>(c, dev, nch, nblk, part); break;
  case BOP_COSGAP: launch<BOP_COSGAP>(c, dev, nch, nblk, part); break;
  case BOP_RESID: launch<BOP_RESID>(c, dev, nch, nblk, part); break;
  default: return TH2D_EARG;
  }
  if (red)
    hipLaunchKernelGGL(k_bop_finish, dim3(nch), dim3(BT), 0, c->stream, dev, nch, nblk, part);
  TH2D_HIP(hipGetLastError());
  return TH2D_OK;
}

}  // namespace th2d

// out_c = M in_c (dagger 0) or M^dag in_c (dagger 1) for every chain c of a
// batch of lattices of any sizes, in one launch (FP64, eo order, as
// th2d_dslash).  The descriptor table is uploaded only when it differs from
// the previous call's on this context.
extern "C" int th2d_dslash_batched(th2d_ctx *c, int nch, const th2d_geom *geoms,
                                   double *const *out, const double *const *in,
                                   const double *const *K, int dagger)
{
  if (!c || nch < 1 || !geoms || !out || !in || !K) return TH2D_EARG;
  std::vector<BOp> ops(nch);
  for (int k = 0; k < nch; k++) {
    TH2D_TRY(th2d_check_geom(&geoms[k]));
    if (!out[k] || !in[k] || !K[k] || out[k] == in[k]) return TH2D_EARG;
    memset(&ops[k], 0, sizeof(BOp));  // padding too: the bytes are the cache key
    ops[k].G = make_geo(&geoms[k]);
    ops[k].bc = Bc{geoms[k].anti_t, geoms[k].anti_x};
    ops[k].d0 = out[k];
    ops[k].d1 = (double *)in[k];
    ops[k].d2 = (double *)K[k];
    ops[k].s1 = dagger ? 1.0 : 0.0;
  }
  const long nblk = th2d::bop_finish(ops);
  const size_t bytes = sizeof(BOp) * ops.size();
  const unsigned char *raw = (const unsigned char *)ops.data();
  if (c->dsl_key.size() != bytes || memcmp(c->dsl_key.data(), raw, bytes) != 0) {
    TH2D_TRY(th2d::bop_upload(c, c->dsl_tab, ops));
    c->dsl_key.assign(raw, raw + bytes);
  }
  return th2d::bop_launch(c, th2d::BOP_DSLASH, (const BOp *)c->dsl_tab.ptr, nch, nblk);
}
