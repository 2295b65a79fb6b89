/*
 * A program written against hmc.c's interface, linked as INTEGRATION.md
 * shows (-lthirring_hmc -lth2d): it owns the globals m, g, mu, A and its
 * own MT19937 stream (mersenne_i, mersenne_array, seed_mersenne,
 * mersenne_generate), which must replace the library's weak ones.
 *
 * Prints key=value lines for tests/test_gpu_boundary.py:
 *   mass_response   (M v)(0) at m = 0.3, then at m = 0.7, for v = delta_0
 *   generate_calls  blocks drawn from this program's generator
 *   traj k: accepted dH  for trajectories from seed 4001 (nsteps 20)
 *   ckpt_read       status of reading back a checkpoint (the program's
 *                   generator: only its current block can be restored)
 *   A               the final field, natural order
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "thirring_hmc.h"

#define NT 8
#define NX 8

double m = 0.3, g = 0.5, mu = 0.0;
double ***A;

/* ---- the program's own MT19937 (Matsumoto & Nishimura 1998, sgenrand
 * seeding, blocks consumed from the top as mersenne.h's macro does) ---- */
int mersenne_i = -1;
double mersenne_array[624];
static unsigned int words[624];
int generate_calls = 0;

void seed_mersenne(long seed)
{
  if (seed % 2 == 0) seed++;
  words[0] = (unsigned int)seed;
  for (int k = 1; k < 624; k++) words[k] = 69069u * words[k - 1];
  mersenne_i = 0;
}

double mersenne_generate(void)
{
  generate_calls++;
  if (mersenne_i < 0) seed_mersenne(4357);
  for (int k = 0; k < 624; k++) {
    unsigned int y = (words[k] & 0x80000000u) | (words[(k + 1) % 624] & 0x7fffffffu);
    words[k] = words[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  for (int k = 0; k < 624; k++) {
    unsigned int y = words[k];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    mersenne_array[k] = y * (1.0 / 4294967296.0);
  }
  mersenne_i = 624;
  return mersenne_array[--mersenne_i];
}

static _Complex double **alloc_vector(void)
{
  _Complex double **v = malloc(NT * sizeof(*v));
  for (int t = 0; t < NT; t++) v[t] = calloc(NX + 1, sizeof(**v));
  return v;
}

int main(int argc, char **argv)
{
  const char *ckpt = argc > 1 ? argv[1] : "hmc.ckpt";
  if (thirring_hmc_init(NT, NX) != 0) {
    printf("init failed\n");
    return 1;
  }
  A = malloc(NT * sizeof(*A));
  for (int t = 0; t < NT; t++) {
    A[t] = malloc((NX + 1) * sizeof(**A));
    for (int x = 0; x < NX + 1; x++) {
      A[t][x] = malloc(2 * sizeof(double));
      A[t][x][0] = 0.3 * sin(1.0 + t + 2.0 * x);
      A[t][x][1] = 0.3 * cos(2.0 * t - x);
    }
  }
  _Complex double **v = alloc_vector(), **w = alloc_vector();
  v[0][0] = 1.0;
  fm_mul(v, w, A);
  double r0 = creal(w[0][0]);
  m = 0.7; /* the library reads the program's m */
  fm_mul(v, w, A);
  printf("mass_response=%.17g %.17g\n", r0, creal(w[0][0]));
  m = 0.3;

  thirring_hmc_set_cg(1e-12, 0.0, 100000, 0);
  thirring_hmc_set_md(20, 1.0);
  seed_mersenne(4001);
  for (int k = 0; k < 3; k++) {
    update_gauge(A);
    int acc, its;
    double dH;
    thirring_hmc_last_trajectory(&acc, &dH, &its);
    printf("traj=%d %d %.17g\n", k, acc, dH);
  }
  printf("generate_calls=%d\n", generate_calls);
  printf("status=%d\n", thirring_hmc_last_status());
  int rc = thirring_hmc_write_checkpoint(ckpt, A, 3);
  long n = 0;
  int rd = thirring_hmc_read_checkpoint(ckpt, A, &n);
  printf("ckpt_write=%d\nckpt_read=%d %ld\n", rc, rd, n);
  printf("A=");
  for (int t = 0; t < NT; t++)
    for (int x = 0; x < NX; x++) printf("%.17g %.17g ", A[t][x][0], A[t][x][1]);
  printf("\n");
  thirring_hmc_finalize();
  return 0;
}
