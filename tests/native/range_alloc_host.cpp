// range_alloc_host.cpp - C entry for the CPU tests of the batched arena's
// range allocator (ccfd_amd/csrc/range_alloc.hpp). Test infrastructure;
// built by ccfd_amd/build.py into tests/native/.
#include "range_alloc.hpp"

using ccfd::RangeAlloc;

extern "C" void* ra_new(int cap) {
  auto* a = new RangeAlloc();
  a->init(cap);
  return a;
}
extern "C" void ra_delete(void* h) { delete (RangeAlloc*)h; }
extern "C" int ra_alloc(void* h, int n) { return ((RangeAlloc*)h)->alloc(n); }
extern "C" int ra_can(void* h, int n) { return ((RangeAlloc*)h)->can(n); }
extern "C" void ra_release(void* h, int s, int n) {
  ((RangeAlloc*)h)->release(s, n);
}
extern "C" int ra_high(void* h) { return ((RangeAlloc*)h)->high(); }
extern "C" int ra_used(void* h) { return ((RangeAlloc*)h)->used; }
// number of free ranges (coalesced: no two touch)
extern "C" int ra_ranges(void* h) { return (int)((RangeAlloc*)h)->freeR.size(); }
