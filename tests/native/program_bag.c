/*
 * A program written against fermionbag.c's interface, linked as
 * INTEGRATION.md shows (-lthirring_bag -lth2d): it owns field, Dinv, m,
 * mu, U, the counters, max_changes and update_linklists, which must
 * replace the library's weak ones.  6 x 6 lattice, antiperiodic, no mass
 * monomers (the full-inverse calc_Dinv), m = 0.4.
 *
 * Prints key=value lines for tests/test_gpu_boundary.py:
 *   fM        fM applied to psi(s) = sin(1 + s) with one link occupied
 *   dinv      the first 8 entries of the program's Dinv after calc_Dinv
 *   added     det_added_link of a free pair, then new_link commits it
 *   removed   det_removed_link of that pair
 *   linklists calls of the program's update_linklists
 *   sign      the program's current_sign after the commits
 *   field     the program's field after the commits
 */
#include <stdio.h>
#include <stdlib.h>

#include "thirring_bag.h"

#define NT 6
#define NX 6

int **field;
double m = 0.4, mu = 0.0, U = 1.0;
double *Dinv;
int n_monomers = 0, n_links = 0, current_sign = 7, max_changes = 2;
int linklists_calls = 0;
void update_linklists(void) { linklists_calls++; }

int main(void)
{
  field = malloc(NT * sizeof(*field));
  for (int t = 0; t < NT; t++) field[t] = calloc(NX + 1, sizeof(int));
  Dinv = malloc(sizeof(double) * NT * NX * NT * NX);
  if (thirring_bag_init(NT, NX, 0, 0) != 0) {
    printf("init failed\n");
    return 1;
  }
  calc_Dinv();
  printf("dinv=");
  for (int k = 0; k < 8; k++) printf("%.17g ", Dinv[k]);
  printf("\n");

  double **psi = malloc(NT * sizeof(*psi)), **chi = malloc(NT * sizeof(*chi));
  for (int t = 0; t < NT; t++) {
    psi[t] = calloc(NX + 1, sizeof(double));
    chi[t] = calloc(NX + 1, sizeof(double));
    for (int x = 0; x < NX; x++) psi[t][x] = __builtin_sin(1.0 + t * NX + x);
  }
  field[2][3] = 2; /* a link in direction t from (2,3) to (3,3) */
  field[3][3] = 4;
  fM(chi, psi);
  printf("fM=");
  for (int t = 0; t < NT; t++)
    for (int x = 0; x < NX; x++) printf("%.17g ", chi[t][x]);
  printf("\n");

  update_background();
  double a = det_added_link(0, 0, 0, 1);
  new_link(0, 0, 1);
  n_links++;
  double r = det_removed_link(0, 0, 0, 1);
  double b = det_added_link(4, 1, 5, 1);
  new_link(4, 1, 0);
  n_links++;
  printf("added=%.17g %.17g\nremoved=%.17g\n", a, b, r);
  printf("linklists=%d\nsign=%d\nstatus=%d\n", linklists_calls, current_sign,
         thirring_bag_last_status());
  printf("field=");
  for (int t = 0; t < NT; t++)
    for (int x = 0; x < NX; x++) printf("%d ", field[t][x]);
  printf("\n");
  thirring_bag_finalize();
  return 0;
}
