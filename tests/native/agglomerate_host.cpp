// agglomerate_host.cpp - C entry for the CPU tests of the agglomeration
// (ccfd_amd/csrc/agglomerate.cpp). Test infrastructure; built by
// ccfd_amd/build.py into tests/native/.
#include <cstring>

#include "agglomerate.hpp"

using namespace ccfd;

extern "C" void* agg_build(int N, int Np, int S, const int* link,
                           const double* w, int minCells, int maxLevels) {
  auto* v = new std::vector<AggLevel>();
  build_agglomeration(N, Np, S, link, w, minCells, maxLevels, *v);
  return v;
}

extern "C" void agg_free(void* h) { delete (std::vector<AggLevel>*)h; }

extern "C" int agg_levels(void* h) {
  return (int)((std::vector<AggLevel>*)h)->size();
}

// sizes of: aggOf, rowStart, col, dSrcStart, dSrc, oSrcStart, oSrc,
// colorStart, colorCells (9 entries)
extern "C" void agg_sizes(void* h, int l, long* sz) {
  const AggLevel& L = (*(std::vector<AggLevel>*)h)[l];
  const size_t v[9] = {L.aggOf.size(), L.rowStart.size(), L.col.size(),
                       L.dSrcStart.size(), L.dSrc.size(), L.oSrcStart.size(),
                       L.oSrc.size(), L.colorStart.size(), L.colorCells.size()};
  for (int k = 0; k < 9; ++k) sz[k] = (long)v[k];
}

// copy array k (order as in agg_sizes; ints, or int64 for 3-6) to out
extern "C" void agg_get(void* h, int l, int k, void* out) {
  const AggLevel& L = (*(std::vector<AggLevel>*)h)[l];
  switch (k) {
    case 0: std::memcpy(out, L.aggOf.data(), L.aggOf.size() * 4); break;
    case 1: std::memcpy(out, L.rowStart.data(), L.rowStart.size() * 4); break;
    case 2: std::memcpy(out, L.col.data(), L.col.size() * 4); break;
    case 3: std::memcpy(out, L.dSrcStart.data(), L.dSrcStart.size() * 8); break;
    case 4: std::memcpy(out, L.dSrc.data(), L.dSrc.size() * 8); break;
    case 5: std::memcpy(out, L.oSrcStart.data(), L.oSrcStart.size() * 8); break;
    case 6: std::memcpy(out, L.oSrc.data(), L.oSrc.size() * 8); break;
    case 7: std::memcpy(out, L.colorStart.data(), L.colorStart.size() * 4); break;
    case 8: std::memcpy(out, L.colorCells.data(), L.colorCells.size() * 4); break;
  }
}
