// Host check of the detonation program's exp (exp_glibc in
// sweeping_fd_weno_amd/csrc/swf_ex7.h) against the host C library's exp, which the
// reference build of ex7_detonation calls: bitwise, on seeded arguments over the
// ranges the program's source term and its large-|x| paths see.  Built (g++ against
// tests/emu's HIP runtime header) and run by tests/test_ex7_exp.py.
//   ex7_exp_check table       prints the 256 table words
//   ex7_exp_check N           prints "lo hi mismatches samples" per range
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "swf_ex7.h"

static uint64_t s_state = 88172645463325252ull;
static uint64_t rnd() {
  s_state ^= s_state << 13;
  s_state ^= s_state >> 7;
  s_state ^= s_state << 17;
  return s_state;
}

int main(int argc, char **argv) {
  if (argc > 1 && strcmp(argv[1], "table") == 0) {
    for (int i = 0; i < 256; ++i) printf("%016llx\n", (unsigned long long)swf::ex7::kExpTab[i]);
    return 0;
  }
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  const double ranges[][2] = {{-60, 0}, {-1, 1}, {-745.2, -700}, {-1100, 10}, {-1e-10, 1e-10}, {700, 709.8}};
  for (const auto &r : ranges) {
    long bad = 0;
    for (long i = 0; i < n; ++i) {
      const double x = r[0] + (r[1] - r[0]) * ((rnd() >> 11) * 0x1.0p-53);
      const double a = exp(x), b = swf::ex7::exp_glibc(x);
      uint64_t ua, ub;
      memcpy(&ua, &a, 8);
      memcpy(&ub, &b, 8);
      if (ua != ub) {
        if (bad < 3) fprintf(stderr, "x=%a libm=%a restated=%a\n", x, a, b);
        ++bad;
      }
    }
    printf("%g %g %ld %ld\n", r[0], r[1], bad, n);
  }
  // special arguments
  const double sp[] = {0.0, -0.0, 1e-300, -1e-300, -745.1332191019411, -745.1332191019412, -708.3964185322641,
                       709.782712893384, 709.7827128933841, -1024.0, -INFINITY, INFINITY, NAN};
  long bad = 0;
  for (double x : sp) {
    const double a = exp(x), b = swf::ex7::exp_glibc(x);
    uint64_t ua, ub;
    memcpy(&ua, &a, 8);
    memcpy(&ub, &b, 8);
    if (ua != ub && !(a != a && b != b)) ++bad;
  }
  printf("special special %ld %d\n", bad, (int)(sizeof(sp) / sizeof(sp[0])));
  return 0;
}
