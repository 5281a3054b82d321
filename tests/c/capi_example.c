/* A plain C99 caller of the qsx C-ABI (host-only entry points): checks that
   include/qsx.h is a valid C header and that a C program links against
   libqsx.so and gets the documented answers.  Exit status 0 on success. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "qsx.h"

static int failures = 0;

static void check(int ok, const char* what) {
  if (!ok) {
    fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

int main(void) {
  const char* v = qsx_version();
  check(v != NULL && v[0] != '\0', "qsx_version");

  /* contractQS index strings: '13*' -> legs 0 and 2 of the conjugated operand */
  int32_t idx[8], conj = -1;
  int32_t n = qsx_parse_ctridx("13*", idx, 8, &conj);
  check(n == 2 && idx[0] == 0 && idx[1] == 2 && conj == 1, "qsx_parse_ctridx(\"13*\")");
  n = qsx_parse_ctridx("[2 1]", idx, 8, &conj);
  check(n == 2 && idx[0] == 1 && idx[1] == 0 && conj == 0, "qsx_parse_ctridx(\"[2 1]\")");
  check(qsx_parse_ctridx("0", idx, 8, &conj) < 0, "qsx_parse_ctridx rejects a 0 index");

  /* <1/2 1/2; 1/2 -1/2 | 0 0> = 1/sqrt(2) */
  check(fabs(qsx_su2_cg(1, 1, 1, -1, 0, 0) - sqrt(0.5)) < 1e-15, "qsx_su2_cg singlet");
  /* four spin-1/2 legs couple to a singlet in two ways: outer multiplicity 2 */
  {
    const int32_t tj[4] = {1, 1, 1, 1}, dir[4] = {1, 1, -1, -1};
    char err[256];
    check(qsx_su2_om_basis(4, tj, dir, NULL, 0, err, sizeof err) == 2, "qsx_su2_om_basis count");
  }

  /* packed offsets of a two-block rank-2 tensor: 2x3 and 4x1 */
  {
    const int32_t Q[4] = {1, 1, -1, -1};
    const int64_t dims[4] = {2, 3, 4, 1};
    qsx_tensor T;
    int64_t offs[3];
    memset(&T, 0, sizeof T);
    T.nblk = 2;
    T.rank = 2;
    T.qdim = 1;
    T.Q = Q;
    T.dims = dims;
    T.qtype = "A";
    check(qsx_packed_offsets(&T, offs) == 0 && offs[0] == 0 && offs[1] == 6 && offs[2] == 10,
          "qsx_packed_offsets");
  }

  if (failures == 0) printf("capi_example: ok (%s)\n", v);
  return failures == 0 ? 0 : 1;
}
