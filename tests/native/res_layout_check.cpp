// Host check of the resident solver's publication layout (res_layout.h),
// for regions of 64 x 64 and 32 x 64 sites and every even lattice up to
// NMAX x NMAX: a region publishes exactly the
// sites within 2 of its clipped edge, into unique slots below 496; every
// site that an owned site's M^dag M result depends on (L1 distance <= 2)
// and that the region does not own is published by its owner; every
// element of the 2-site halo rectangle the kernel may load has a slot.
#include <cstdio>
#include <vector>
#include <set>
#include <cstdlib>
#include <algorithm>
#include "res_layout.h"
int main(int argc, char **argv)
{
  const int nmax = argc > 1 ? atoi(argv[1]) : 200;
  long bad = 0, checked = 0, dummy = 0;
  for (int tt : {64, 32})  // regions of 64 x 64 and 32 x 64 sites
  for (int nt = 2; nt <= nmax; nt += 2)
    for (int nx = 2; nx <= nmax; nx += 2) {
      const int ntt = (nt + tt - 1) / tt, ntx = (nx + 63) / 64;
      // uniqueness and range
      for (int rt = 0; rt < ntt; rt++)
        for (int rx = 0; rx < ntx; rx++) {
          int te = std::min(tt, nt - rt * tt), xe = std::min(64, nx - rx * 64);
          std::set<int> used;
          for (int lt = 0; lt < te; lt++)
            for (int lx = 0; lx < xe; lx++) {
              bool edge = lt < 2 || lt >= te - 2 || lx < 2 || lx >= xe - 2;
              int s = res_pub_slot(lt, lx, te, xe);
              if (edge != (s >= 0)) { if (bad++ < 10) printf("edge/slot mismatch %d %d (%d,%d) te %d xe %d s %d\n", nt, nx, lt, lx, te, xe, s); }
              if (s >= 496 || (s >= 0 && !used.insert(s).second)) { if (bad++ < 10) printf("slot range/dup %d %d (%d,%d) s %d\n", nt, nx, lt, lx, s); }
            }
          // every site within L1 distance 2 of an owned site and not owned is published
          for (int lt = -2; lt < te + 2; lt++)
            for (int lx = -2; lx < xe + 2; lx++) {
              bool own = lt >= 0 && lt < te && lx >= 0 && lx < xe;
              if (own) continue;
              bool need = false;
              for (int dt = -2; dt <= 2 && !need; dt++)
                for (int dx = -2; dx <= 2; dx++) {
                  if (abs(dt) + abs(dx) > 2) continue;
                  int a = lt + dt, b = lx + dx;
                  if (a >= 0 && a < te && b >= 0 && b < xe) { need = true; break; }
                }
              int t = ((rt * tt + lt) % nt + nt) % nt, x = ((rx * 64 + lx) % nx + nx) % nx, k;
              int s = res_site_slot(t, x, nt, nx, k, tt);
              checked++;
              if (need && s < 0) { if (bad++ < 10) printf("needed but unpublished %d %d region (%d,%d) (%d,%d)\n", nt, nx, rt, rx, lt, lx); }
              if (!need && s < 0) dummy++;
              if (k < 0 || k >= ntt * ntx) { if (bad++ < 10) printf("owner out of range\n"); }
            }
        }
    }
  printf("checked %ld halo elements, %ld bad, %ld not needed and unpublished\n", checked, bad, dummy);
  return bad != 0 || dummy != 0;
}
