// Boundary exchange layout of the resident solver (k_cg_resident).
//
// A lattice is cut into regions of RES_T (or RES_T / 2) rows and RES_T
// columns, numbered t-major within a chain (region k covers rows
// tt (k / ntx) .., columns 64 (k % ntx) ..; the last row and column of
// regions are clipped to the lattice).  Each
// region publishes the sites within 2 sites of its clipped edge into a
// block of RES_PUB elements of its own: the top two rows, the bottom two
// rows, then for every other row its two left and its two right columns,
// so that a neighbour reads whole rows and whole column pairs from
// consecutive addresses.  Every halo value that reaches a region's result
// lies within two sites of the edge of the region that owns it, so it has
// a slot here.  Used by the device code and by host-side checks.
#pragma once

#ifdef __HIPCC__
#define RES_HD __host__ __device__
#else
#define RES_HD
#endif

constexpr int RES_T = 64;      // region rows and columns
constexpr int RES_PUB = 512;   // elements of a publication block (496 used at most)

// Slot of element (lt, lx) of a region with clipped extents te x xe, or -1
// when the element is not within 2 sites of the region's edge.
RES_HD inline int res_pub_slot(int lt, int lx, int te, int xe)
{
  if (lt < 2) return lt * RES_T + lx;
  if (lt >= te - 2) return 2 * RES_T + (lt - (te - 2)) * RES_T + lx;
  const int r = lt - 2;  // 0 <= r < te - 4
  if (lx < 2) return 4 * RES_T + 2 * r + lx;
  if (lx >= xe - 2) return 4 * RES_T + 2 * (te - 4) + 2 * r + (lx - (xe - 2));
  return -1;
}

// Region (index within the chain, into k) and publication slot of lattice
// site (t, x), 0 <= t < nt, 0 <= x < nx, for regions of tt rows (64 or 32)
// and RES_T columns; -1 when the site is not published.
RES_HD inline int res_site_slot(int t, int x, int nt, int nx, int &k, int tt = RES_T)
{
  const int rt = t / tt, rx = x / RES_T;
  k = rt * ((nx + RES_T - 1) / RES_T) + rx;
  const int te = nt - rt * tt < tt ? nt - rt * tt : tt;
  const int xe = nx - rx * RES_T < RES_T ? nx - rx * RES_T : RES_T;
  return res_pub_slot(t - rt * tt, x - rx * RES_T, te, xe);
}
