// bag_geo.h -- lattice geometry and matrix elements of the fermion-bag
// operator (fermion_matrix.c:28-64, 152-182; vec_ops.c:96-172), shared by
// the bag kernels (bag.hip) and the batched bag driver (bagmc.hip).
#pragma once

#include <cmath>

#include "../../include/th2d_bag.h"

namespace {

struct BagGeo {
  int nt, nx, bc, wmm;
  double m, ep, em;
};

BagGeo bag_geo(const th2d_bag_geom *g)
{
  return BagGeo{g->nt, g->nx, g->bc, g->with_mass_monomers, g->m, exp(g->mu), exp(-g->mu)};
}

int check_bag_geom(const th2d_bag_geom *g)
{
  if (!g || g->nt < 4 || g->nx < 4 || (g->nt & 1) || (g->nx & 1)) return TH2D_EARG;
  if (g->bc != TH2D_BC_ANTISYMMETRIC && g->bc != TH2D_BC_OPENX) return TH2D_EARG;
  return TH2D_OK;
}

// Neighbour of (t, x) in direction dir (0 tup, 1 xup, 2 tdn, 3 xdn); x = -1
// when an open x boundary is crossed (fermionbag.c:753-765).
__host__ __device__ inline void bag_nb(const BagGeo &G, int t, int x, int dir, int &t2, int &x2)
{
  t2 = t;
  x2 = x;
  if (dir == 0) t2 = t + 1 == G.nt ? 0 : t + 1;
  else if (dir == 2) t2 = t == 0 ? G.nt - 1 : t - 1;
  else if (dir == 1) x2 = x + 1 == G.nx ? (G.bc == TH2D_BC_OPENX ? -1 : 0) : x + 1;
  else x2 = x == 0 ? (G.bc == TH2D_BC_OPENX ? -1 : G.nx - 1) : x - 1;
}

// Element of fM from (t,x) to its neighbour in direction dir (vec_ops.c:
// 105-128): 1/2 eta_t e^{+mu} (t up), 1/2 eta_t e^{-mu} (t down), 1/2 (x),
// + when the neighbour's coordinate is larger, - otherwise.  The
// transpose flips the sign and swaps e^{+mu} and e^{-mu}.
__host__ __device__ inline double bag_hop(const BagGeo &G, int t, int x, int dir, int t2, int x2,
                                          int transpose)
{
  double mag, sg;
  if (dir == 0 || dir == 2) {
    double eta = (x & 1) ? -1.0 : 1.0;
    bool fwd = dir == 0;
    mag = 0.5 * eta * ((fwd != (bool)transpose) ? G.ep : G.em);
    sg = t2 > t ? 1.0 : -1.0;
  } else {
    mag = 0.5;
    sg = x2 > x ? 1.0 : -1.0;
  }
  return transpose ? -sg * mag : sg * mag;
}

// Matrix element D[s1][s2] of the unmasked operator (fM_index semantics for
// extents >= 4; fermion_matrix.c:28-64, 152-182).
__host__ __device__ inline double bag_elem(const BagGeo &G, int s1, int s2)
{
  int t1 = s1 / G.nx, x1 = s1 - t1 * G.nx;
  if (s1 == s2) return G.wmm ? 0.0 : G.m;
  for (int dir = 0; dir < 4; dir++) {
    int t2, x2;
    bag_nb(G, t1, x1, dir, t2, x2);
    if (x2 >= 0 && t2 * G.nx + x2 == s2) return bag_hop(G, t1, x1, dir, t2, x2, 0);
  }
  return 0.0;
}

}  // namespace
