// refrandom.cc -- golden-vector driver around the REFERENCE's own random.hh.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from the
// reference header where it lies.  Prints the streams the oracle and the device
// generator must reproduce: Splitmix64 (random.hh:160-185), Xoshiro128++ seeding,
// randomInt32, jump and float-rounded random() (random.hh:333-465), Poisson
// (random.hh:42-83).  JSON on stdout.
#include <cassert>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>
#include "random.hh"

using namespace morefit;

int main() {
  const uint64_t seed = 229387429ULL;  // benchmarking/morefit_massfit.cc:59-60
  printf("{\n  \"generator\": \"oracle/ref/refrandom.cc over the reference random.hh\",\n  \"seed\": %llu,\n", (unsigned long long)seed);
  {
    Splitmix64 s(seed);
    printf("  \"splitmix64\": [");
    for (int i = 0; i < 16; i++) printf("%s\"%llu\"", i ? ", " : "", (unsigned long long)s.randomInt64());
    printf("],\n");
  }
  {
    Xoshiro128pp x;
    x.setSeed(seed);
    printf("  \"xoshiro128pp_state\": [%u, %u, %u, %u],\n", x.getSeed(0), x.getSeed(1), x.getSeed(2), x.getSeed(3));
    printf("  \"xoshiro128pp_int32\": [");
    for (int i = 0; i < 32; i++) printf("%s%u", i ? ", " : "", x.randomInt32());
    printf("],\n");
    x.jump();
    printf("  \"xoshiro128pp_after_jump_state\": [%u, %u, %u, %u],\n", x.getSeed(0), x.getSeed(1), x.getSeed(2), x.getSeed(3));
    printf("  \"xoshiro128pp_after_jump_int32\": [");
    for (int i = 0; i < 8; i++) printf("%s%u", i ? ", " : "", x.randomInt32());
    printf("],\n");
    printf("  \"xoshiro128pp_random\": [");
    for (int i = 0; i < 16; i++) printf("%s%.17g", i ? ", " : "", x.random());
    printf("],\n");
  }
  {
    const double means[4] = {3.5, 24.0, 30.0, 5000.0};
    printf("  \"poisson\": {");
    for (int m = 0; m < 4; m++) {
      Xoshiro128pp x;
      x.setSeed(seed + m);
      printf("%s\"%g\": [", m ? ", " : "", means[m]);
      for (int i = 0; i < 64; i++) printf("%s%u", i ? ", " : "", x.poisson(means[m]));
      printf("]");
    }
    printf("}\n");
  }
  printf("}\n");
  return 0;
}
