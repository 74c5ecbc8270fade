// CPU unit test of the serving loop's arena and size classes
// (fishfry_amd/csrc/ff_arena.hpp): randomized allocate / release sequences
// checked against a model of 256-byte units -- no overlap, 256-byte
// alignment, first fit, full coalescing, used / peak accounting.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "ff_arena.hpp"

#define REQUIRE(c)                                                    \
  do {                                                                \
    if (!(c)) {                                                       \
      std::fprintf(stderr, "FAILED %s (line %d)\n", #c, __LINE__);    \
      std::exit(1);                                                   \
    }                                                                 \
  } while (0)

int main() {
  const size_t unit = 256;
  std::mt19937_64 rng(20261202);
  for (int trial = 0; trial < 200; ++trial) {
    const size_t units = 64 + rng() % 4096;
    ff::Arena a;
    a.reset(units * unit);
    std::vector<int> owner(units, -1);  // model: which allocation holds each unit
    std::map<size_t, size_t> live;      // offset -> rounded bytes
    size_t used = 0, peak = 0;
    int id = 0;
    for (int step = 0; step < 2000; ++step) {
      if (live.empty() || rng() % 3) {
        const size_t bytes = 1 + rng() % (unit * 64);
        const size_t need = ff::align_up(std::max<size_t>(bytes, unit), unit);
        // first fit in the model: lowest offset with need/unit free units
        size_t want = units, run = 0;
        for (size_t u = 0; u < units; ++u) {
          run = owner[u] < 0 ? run + 1 : 0;
          if (run == need / unit) {
            want = u + 1 - run;
            break;
          }
        }
        size_t off = 0;
        const bool ok = a.alloc(bytes, off);
        REQUIRE(ok == (want < units));
        if (!ok) continue;
        REQUIRE(off == want * unit);
        REQUIRE(off % unit == 0 && off + need <= units * unit);
        for (size_t u = off / unit; u < (off + need) / unit; ++u) {
          REQUIRE(owner[u] < 0);
          owner[u] = id;
        }
        ++id;
        live[off] = need;
        used += need;
        peak = std::max(peak, used);
      } else {
        auto it = live.begin();
        std::advance(it, rng() % live.size());
        a.release(it->first);
        for (size_t u = it->first / unit; u < (it->first + it->second) / unit; ++u) owner[u] = -1;
        used -= it->second;
        live.erase(it);
      }
      REQUIRE(a.used() == used && a.peak() == peak && a.size() == units * unit);
    }
    // releasing everything coalesces back to one block: the whole arena fits again
    for (auto& kv : live) a.release(kv.first);
    REQUIRE(a.used() == 0);
    size_t off = 1;
    REQUIRE(a.alloc(units * unit, off) && off == 0);
    a.release(12345);  // unknown offsets are ignored
    REQUIRE(a.used() == units * unit);
  }
  // size classes: grouped by precision and refinement steps, and by the
  // octave of the mean edge
  const int small[3] = {32, 32, 32}, near[3] = {40, 40, 40}, big[3] = {256, 256, 256}, flat[3] = {64, 64, 32},
            next[3] = {64, 64, 64};
  REQUIRE(ff::size_class(0, 0, small) == ff::size_class(0, 0, near));
  REQUIRE(ff::size_class(0, 0, small) < ff::size_class(0, 0, big));
  REQUIRE(ff::size_class(0, 0, flat) == ff::size_class(0, 0, small));     // mean edge 50.8: same octave
  REQUIRE(ff::size_class(0, 0, next) == ff::size_class(0, 0, small) + 1);  // mean edge 64: next octave
  REQUIRE(ff::size_class(0, 0, small) != ff::size_class(1, 0, small));
  REQUIRE(ff::size_class(1, 1, small) != ff::size_class(1, 2, small));
  std::printf("arena ok\n");
  return 0;
}
