// Device launch tables of an ensemble, shared by content (the two parities of
// two-level time stepping alternate two tables per phase).  Plain C++ with
// no HIP dependency: the device handle is a template parameter, and the
// caller allocates, copies and frees.
//
// Lifetime rule: a table handed out while the lead stream is being captured
// may be read by every later replay of that graph, so it is pinned and only
// the ensemble's destruction frees it.  Eviction (to bound the cache) takes
// the oldest unpinned table; when every table is pinned the cache refuses to
// grow instead of freeing memory a graph may still read.
#pragma once

#include <cstddef>
#include <cstring>
#include <vector>

template <typename Dev>
struct TableCache {
  struct Entry {
    std::vector<unsigned char> host;
    Dev dev;
    bool pinned;
  };
  std::vector<Entry> entries;
  size_t cap = 64;

  // Index of the table with these contents (pinning it when capturing), or -1.
  int lookup(const void *host, size_t bytes, bool capturing) {
    for (size_t i = 0; i < entries.size(); ++i) {
      Entry &t = entries[i];
      if (t.host.size() == bytes && std::memcmp(t.host.data(), host, bytes) == 0) {
        if (capturing) t.pinned = true;
        return (int)i;
      }
    }
    return -1;
  }

  // Index of the table to evict before adding one (the oldest unpinned), -1
  // when there is room already, -2 when the cache is full of pinned tables.
  int victim() const {
    if (entries.size() < cap) return -1;
    for (size_t i = 0; i < entries.size(); ++i)
      if (!entries[i].pinned) return (int)i;
    return -2;
  }

  void erase(int i) { entries.erase(entries.begin() + i); }

  void add(const void *host, size_t bytes, Dev dev) {
    const unsigned char *h = (const unsigned char *)host;
    entries.push_back({std::vector<unsigned char>(h, h + bytes), dev, false});
  }
};
