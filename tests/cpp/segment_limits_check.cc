// segment_limits_check.cc -- the 32-bit event indices of a segment are guarded on the host: a
// segment of more than MF_MAX_SEGMENT_EVENTS events (or a negative count) is refused before any
// table is built, and the largest allowed segment gets a chunk table whose bounds and the
// kernels' furthest index (end + unroll x block) stay below 2^31.  Prints "ok" or "FAIL ...".
#include <climits>
#include <cstdio>
#include <stdexcept>

#include "morefit/compute_hip.hh"

using namespace morefit;

int main() {
  int fails = 0;
  auto refused = [&](long long n) {
    std::vector<mf_chunk> ch;
    mf_segment sg{};
    try {
      append_chunks(ch, sg, 0, n, 256 * 64);
    } catch (const std::invalid_argument&) {
      return ch.empty();
    }
    return false;
  };
  for (long long n : {MF_MAX_SEGMENT_EVENTS + 1, 3000000000LL, (long long)INT_MAX, -1LL})
    if (!refused(n)) {
      std::printf("FAIL segment of %lld events accepted\n", n);
      fails++;
    }
  {
    std::vector<mf_chunk> ch;
    mf_segment sg{};
    const int block = 1024, unroll = 8;
    append_chunks(ch, sg, 0, MF_MAX_SEGMENT_EVENTS, block * 64);
    long long prev = 0;
    for (auto& c : ch) {
      if (c.begin != prev || c.end <= c.begin) fails++;
      if ((long long)c.end + (long long)unroll * block > INT_MAX) fails++;
      prev = c.end;
    }
    if (prev != MF_MAX_SEGMENT_EVENTS) {
      std::printf("FAIL chunks cover %lld events\n", prev);
      fails++;
    }
  }
  if (fails == 0) std::printf("ok\n");
  return fails ? 1 : 0;
}
