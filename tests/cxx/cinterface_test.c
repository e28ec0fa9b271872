/* A C99 caller of the reference's C interface, shaped like
 * src_tests/test9-Cinterface.c: the 7x7 BABD system (x = 1..7) and its
 * bordered 8x8 extension (x = (1..7, 0)), through BABD_factorize,
 * BABD_factorize_bordered, BABD_solve and BABD_get_last_error as declared in
 * include/alglin_amd.h.  Exit status 0 when both solutions match. */
#include <math.h>
#include <stdio.h>

#include "alglin_amd.h"

static int check(char const* what, BABD_realType const* x, BABD_realType const* ref, int m) {
  int i, bad = 0;
  for (i = 0; i < m; ++i)
    if (fabs(x[i] - ref[i]) > 1e-9) {
      printf("%s: x[%d] = %.17g expected %.17g\n", what, i, x[i], ref[i]);
      bad = 1;
    }
  return bad;
}

int main(void) {
  /* rows of A: [D_k | E_k] for k = 0, 1, column-major with ldDE = 2 */
  static BABD_realType const DE[] = {10, 2, 2, 40, 2, 1, 3, -1, 10, 2, 2, 40, 2, 1, 3, -1};
  static BABD_realType const H0[] = {0, 1, 1, 1, 0, 2};
  static BABD_realType const HN[] = {20, 0, 2, 2, -10, 2};
  static BABD_realType const Hq[] = {1, 1, 30};
  static BABD_realType const B[] = {1, 1, 1, 1, 1, 1, 1};
  static BABD_realType const C[] = {1, -1, 1, -1, 1, -1, 1};
  static BABD_realType const D[] = {-1};
  BABD_realType rhs[] = {32, 81, 66, 165, 121, -52, 237};
  BABD_realType rhs1[] = {32, 81, 66, 165, 121, -52, 237, 4};
  BABD_realType const x7[] = {1, 2, 3, 4, 5, 6, 7};
  BABD_realType const x8[] = {1, 2, 3, 4, 5, 6, 7, 0};
  int bad = 0;

  if (BABD_factorize(1, 0, 0, 2, 2, 1, 1, DE, 2, H0, 3, HN, 3, Hq, 3) != 0) {
    printf("BABD_factorize: %s\n", BABD_get_last_error());
    return 2;
  }
  if (BABD_solve(1, rhs) != 0) {
    printf("BABD_solve: %s\n", BABD_get_last_error());
    return 2;
  }
  bad |= check("7x7", rhs, x7, 7);

  if (BABD_factorize_bordered(1, 0, 0, 2, 2, 1, 1, 1, 1, DE, 2, H0, 3, HN, 3, Hq, 3, B, 7, C, 1, D, 1) != 0) {
    printf("BABD_factorize_bordered: %s\n", BABD_get_last_error());
    return 2;
  }
  if (BABD_solve(1, rhs1) != 0) {
    printf("BABD_solve: %s\n", BABD_get_last_error());
    return 2;
  }
  bad |= check("8x8", rhs1, x8, 8);

  /* an unknown id and a rejected mat_fact leave a message */
  if (BABD_solve(77, rhs1) == 0) { printf("BABD_solve on an unknown id succeeded\n"); bad = 1; }
  if (BABD_factorize(2, 1, 0, 2, 2, 1, 1, DE, 2, H0, 3, HN, 3, Hq, 3) == 0) {
    printf("mat_fact = 1 (Diaz) was accepted\n");
    bad = 1;
  }
  if (BABD_get_last_error()[0] == 0) { printf("no last error\n"); bad = 1; }
  BABD_free(1);
  BABD_free(2);
  printf(bad ? "FAILED\n" : "ok\n");
  return bad;
}
