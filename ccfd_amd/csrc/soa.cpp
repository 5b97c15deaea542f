// soa.cpp - host-side structure-of-arrays image of a mesh (see soa.hpp).
// Plain C++; compiled with g++ into libccfd_gpu.so and into the CPU test
// library that checks the halo lists.
#include "soa.hpp"

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <utility>

namespace ccfd {

// Hilbert curve index of (x, y) on an n x n grid (n a power of two)
static uint32_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
  uint32_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    uint32_t rx = (x & s) > 0;
    uint32_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

void sfc_order(const ccfd_mesh* m, std::vector<int>& perm) {
  long N = m->nElems;
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (long e = 0; e < N; ++e) {
    x0 = std::min(x0, m->bary[2 * e]);
    x1 = std::max(x1, m->bary[2 * e]);
    y0 = std::min(y0, m->bary[2 * e + 1]);
    y1 = std::max(y1, m->bary[2 * e + 1]);
  }
  double span = std::max(x1 - x0, y1 - y0);
  if (!(span > 0)) span = 1.0;
  std::vector<std::pair<uint32_t, int>> key(N);
  for (long e = 0; e < N; ++e) {
    uint32_t ix = (uint32_t)std::min(65535.0, (m->bary[2 * e] - x0) / span * 65535.0);
    uint32_t iy = (uint32_t)std::min(65535.0, (m->bary[2 * e + 1] - y0) / span * 65535.0);
    key[e] = {hilbert_index(65536u, ix, iy), (int)e};
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint32_t, int>& a,
                      const std::pair<uint32_t, int>& b) { return a.first < b.first; });
  perm.resize(N);
  for (long k = 0; k < N; ++k) perm[k] = key[k].second;
}

void build_soa(const ccfd_mesh* m, int S, int elemBase, int ghostBase,
               int bcBase, MeshSoA& o) {
  const int N = (int)m->nElems;
  const int Np = ((N + kBlock - 1) / kBlock) * kBlock;
  const int G = (int)m->nGhosts;
  o.N = N;
  o.Np = Np;
  o.S = S;
  o.G = G;
  o.nb = Np / kBlock;
  sfc_order(m, o.perm);
  o.inv.assign(N, 0);
  for (int k = 0; k < N; ++k) o.inv[o.perm[k]] = k;

  const size_t sl = (size_t)S * Np;
  o.link.assign(sl, kNoSide);
  o.w.assign(2 * sl, 0.0);
  o.gp.assign(2 * sl, 0.0);
  o.fn.assign(2 * sl, 0.0);
  o.flen.assign(sl, 1.0);
  o.bbv.assign(2 * sl, 0.0);
  o.bbd.assign(sl, 1.0);
  o.areaq.assign(Np, 0.0);
  o.area.assign(Np, 0.0);
  o.sx.assign(Np, 1.0);
  o.sy.assign(Np, 1.0);
  o.bary.assign(2 * (size_t)Np, 0.0);
  for (int i = 0; i < N; ++i) {
    const int e = o.perm[i];
    o.areaq[i] = m->areaq[e];
    o.area[i] = m->area[e];
    o.sx[i] = m->sx[e];
    o.sy[i] = m->sy[e];
    o.bary[i] = m->bary[2 * e];
    o.bary[Np + i] = m->bary[2 * e + 1];
    const long s0 = m->slotStart[e];
    const int ns = (int)(m->slotStart[e + 1] - s0);
    for (int k = 0; k < ns; ++k) {
      const long s = s0 + k;
      const size_t c = (size_t)k * Np + i;
      const long nb = m->nbr[s];
      if (nb >= 0) {
        const long ps = m->partner[s] - m->slotStart[nb];
        o.link[c] = ((o.inv[nb] + elemBase) << 2) | (int)ps;
      } else {
        o.link[c] = (int)nb - ghostBase;   // -(g + ghostBase + 1)
      }
      o.w[c] = m->w[2 * s];
      o.w[sl + c] = m->w[2 * s + 1];
      o.gp[c] = m->GP[2 * s];
      o.gp[sl + c] = m->GP[2 * s + 1];
      o.fn[c] = m->fn[2 * s];
      o.fn[sl + c] = m->fn[2 * s + 1];
      const bool own = m->owner[s] || nb < 0;
      o.flen[c] = own ? m->flen[s] : -m->flen[s];
      o.bbv[c] = m->bbv[2 * s];
      o.bbv[sl + c] = m->bbv[2 * s + 1];
      o.bbd[c] = m->bbd[s];
    }
  }

  // ghosts, in the mesh's ghost order
  o.ghostBC.assign(G, 0);
  o.ghostCell.assign(G, 0);
  o.ghostK.assign(G, 0);
  o.ghostBary.assign(2 * (size_t)G, 0.0);
  for (int g = 0; g < G; ++g) {
    const long s = m->ghostSlot[g];
    const long e = std::upper_bound(m->slotStart, m->slotStart + N + 1, s) -
                   m->slotStart - 1;
    o.ghostCell[g] = o.inv[e] + elemBase;
    o.ghostK[g] = (int)(s - m->slotStart[e]);
    o.ghostBC[g] = m->ghostBC[g] + bcBase;
    o.ghostBary[g] = m->ghostBary[2 * g];
    o.ghostBary[G + g] = m->ghostBary[2 * g + 1];
  }

  // halo of each block: elements of other blocks sharing a face with it
  build_halo(o, kBlock, elemBase, o.haloStart, o.haloCells, o.maxHalo);
}

void build_halo(const MeshSoA& o, int tileElems, int elemBase,
                std::vector<int>& start, std::vector<int>& cells,
                int& maxHalo) {
  const int N = o.N, Np = o.Np, S = o.S;
  const int nt = (N + tileElems - 1) / tileElems;
  start.assign(nt + 1, 0);
  cells.clear();
  maxHalo = 0;
  std::vector<int> tmp;
  for (int b = 0; b < nt; ++b) {
    tmp.clear();
    const int lo = b * tileElems + elemBase, hi = lo + tileElems;
    for (int i = b * tileElems; i < std::min(N, (b + 1) * tileElems); ++i)
      for (int k = 0; k < S; ++k) {
        const int l = o.link[(size_t)k * Np + i];
        if (l < 0) continue;
        const int j = l >> 2;
        if (j < lo || j >= hi) tmp.push_back(j);
      }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cells.insert(cells.end(), tmp.begin(), tmp.end());
    start[b + 1] = (int)cells.size();
    maxHalo = std::max(maxHalo, (int)tmp.size());
  }
}

}  // namespace ccfd
