// Host-only check of the ensemble launch-table cache (mom5_amd/csrc/
// table_cache.h): tables seen while capturing are never evicted; eviction
// takes the oldest unpinned table; a cache full of pinned tables refuses to
// grow.  Built with g++ by tests/test_table_cache.py; prints OK or the
// first failed check and exits non-zero.
#include <cstdio>
#include <cstdlib>

#include "table_cache.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAILED line %d: %s\n", __LINE__, #c);           \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  TableCache<int> tc;
  tc.cap = 4;
  int buf[8];
  for (int i = 0; i < 8; ++i) buf[i] = 100 + i;
  // fill: no victim until full
  for (int i = 0; i < 4; ++i) {
    CHECK(tc.lookup(&buf[i], sizeof(int), false) == -1);
    CHECK(tc.victim() == -1);
    tc.add(&buf[i], sizeof(int), i);
  }
  // a capture uses tables 0 and 2: they are pinned
  CHECK(tc.lookup(&buf[0], sizeof(int), true) == 0);
  CHECK(tc.lookup(&buf[2], sizeof(int), true) == 2);
  CHECK(tc.entries[0].pinned && tc.entries[2].pinned);
  CHECK(!tc.entries[1].pinned && !tc.entries[3].pinned);
  // an eager lookup does not pin
  CHECK(tc.lookup(&buf[3], sizeof(int), false) == 3 && !tc.entries[3].pinned);
  // full: the oldest unpinned is table 1 (device handle 1)
  int v = tc.victim();
  CHECK(v == 1 && tc.entries[v].dev == 1);
  tc.erase(v);
  tc.add(&buf[4], sizeof(int), 4);
  v = tc.victim();
  CHECK(v >= 0 && tc.entries[v].dev == 3);
  tc.erase(v);
  tc.add(&buf[5], sizeof(int), 5);
  // pin the rest: nothing may be evicted any more
  CHECK(tc.lookup(&buf[4], sizeof(int), true) >= 0);
  CHECK(tc.lookup(&buf[5], sizeof(int), true) >= 0);
  CHECK(tc.victim() == -2);
  // pinned tables are still found by content, with their device handles
  CHECK(tc.entries[tc.lookup(&buf[0], sizeof(int), false)].dev == 0);
  CHECK(tc.entries[tc.lookup(&buf[2], sizeof(int), false)].dev == 2);
  // different sizes never match
  long long wide = 100;
  CHECK(tc.lookup(&wide, sizeof wide, false) == -1);
  std::printf("OK\n");
  return 0;
}
