/* The C ABI from plain C (gcc, C99): a host-only plan on the unit circle, its
 * size, stats, sorting and correction rows, and the error codes of bad
 * arguments. Compiled and run by tests/test_host.py; needs no GPU. Prints
 * "ok" and exits 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ifgf2d.h"

static const double kPi = 3.14159265358979323846;

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      printf("check failed at line %d: %s\n", __LINE__, #c);          \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main(void) {
  const int npatches = 40;
  double lims[41];
  for (int p = 0; p <= npatches; p++) lims[p] = 2.0 * kPi * p / npatches;
  const double circle[3] = {0.0, 0.0, 1.0};
  const double delta = 0.75 * 2.0 * sin(kPi / npatches);
  ifgf_surface s = {IFGF_CURVE_CIRCLE, circle, 3, lims, npatches, 10, delta, 6.0};
  ifgf_params prm = {IFGF_COMBINED, IFGF_FP64, 10.0, 10.0, 5, 5, 3, 1, IFGF_PLAN_HOST_ONLY};
  ifgf_plan* plan = NULL;
  CHECK(ifgf_plan_create(&s, &prm, &plan) == IFGF_OK && plan != NULL);
  CHECK(ifgf_plan_size(plan) == 400);
  long long st[20];
  CHECK(ifgf_plan_stats(plan, st) == IFGF_OK);
  CHECK(st[6] == 3 && st[7] == 400 && st[4] > 0);
  long long* sorting = (long long*)malloc(400 * sizeof(long long));
  CHECK(ifgf_plan_sorting(plan, sorting) == IFGF_OK);
  int seen[400] = {0};
  for (int i = 0; i < 400; i++) {
    CHECK(sorting[i] >= 0 && sorting[i] < 400 && !seen[sorting[i]]);
    seen[sorting[i]] = 1;
  }
  const long long ne = ifgf_plan_correction(plan, NULL, NULL, NULL);
  CHECK(ne >= 400);
  /* no GPU work on a host-only plan: a status code, never a crash */
  double x[800] = {0}, y[800];
  CHECK(ifgf_apply(plan, x, y, 1, 400, 0, NULL) == IFGF_ERR_NO_DEVICE);
  CHECK(ifgf_plan_error(plan)[0] != '\0');
  /* field of a host-only plan: tables and counts, no evaluation */
  double tx[3] = {1.5, 0.0, -3.0}, ty[3] = {0.0, 1.5, 2.0};
  ifgf_field* field = NULL;
  CHECK(ifgf_field_create(plan, tx, ty, 3, 5, 5, 5, IFGF_FP64, &field) == IFGF_OK && field != NULL);
  long long fs[8];
  CHECK(ifgf_field_stats(field, fs) == IFGF_OK);
  CHECK(fs[0] == 3 && fs[2] > 0 && fs[5] > 0 && fs[6] == 5);
  CHECK(ifgf_field_apply(field, x, y, 0, NULL) == IFGF_ERR_NO_DEVICE);
  ifgf_field_destroy(field);
  CHECK(ifgf_field_create(plan, tx, ty, 0, 5, 5, 5, IFGF_FP64, &field) == IFGF_ERR_ARG);
  CHECK(ifgf_field_create(plan, tx, ty, 3, 2, 5, 5, IFGF_FP64, &field) == IFGF_ERR_LEVELS);
  ifgf_plan_destroy(plan);
  /* argument errors */
  prm.nlevels = 2;
  CHECK(ifgf_plan_create(&s, &prm, &plan) == IFGF_ERR_LEVELS);
  prm.nlevels = 3;
  prm.Ps = 1;
  CHECK(ifgf_plan_create(&s, &prm, &plan) == IFGF_ERR_ARG);
  CHECK(ifgf_last_error()[0] != '\0');
  CHECK(ifgf_morton_sort(x, x, 4, 2, sorting, NULL) == IFGF_ERR_LEVELS);
  free(sorting);
  printf("ok\n");
  return 0;
}
