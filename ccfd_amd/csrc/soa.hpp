// soa.hpp - host-side structure-of-arrays image of one mesh in the device
// layout: elements renumbered along a Hilbert curve, padded to whole blocks,
// slot arrays [component][slot][Np], ghost arrays [component][G], and the
// per-block halo lists of the fused operator.
#pragma once

#include <vector>

#include "ccfd_mesh.h"
#include "layout.hpp"

namespace ccfd {

struct MeshSoA {
  int N = 0, Np = 0, S = 0, G = 0, nb = 0;
  std::vector<int> perm, inv;  // new -> old, old -> new element
  // slot arrays [c][S][Np]; link holds (neighbour + elemBase) << 2 | partner
  // slot, -(ghost + ghostBase + 1) or kNoSide
  std::vector<int> link;
  std::vector<double> w, gp, fn, flen, bbv, bbd;
  // element arrays [Np] / [2][Np]
  std::vector<double> areaq, area, sx, sy, bary;
  // ghosts [G] / [2][G]; ghostCell is offset by elemBase, ghostBC by bcBase
  std::vector<int> ghostBC, ghostCell, ghostK;
  std::vector<double> ghostBary;
  // per block: halo element ids (offset by elemBase), ascending
  std::vector<int> haloStart, haloCells;
  int maxHalo = 0;
};

// Hilbert order of the element barycentres (stable for equal keys)
void sfc_order(const ccfd_mesh* m, std::vector<int>& perm);

// Build the image of mesh m with S slots per element (>= the mesh's largest
// element). Element, ghost and boundary-condition ids written into the
// arrays are offset by elemBase, ghostBase and bcBase (all 0 for a mesh of
// its own; the batched arena places several meshes side by side).
void build_soa(const ccfd_mesh* m, int S, int elemBase, int ghostBase,
               int bcBase, MeshSoA& out);

// halo lists of tiles of tileElems consecutive elements (a multiple of the
// block size): per tile the ascending ids (offset by elemBase) of the
// elements outside it that share a face with it; start has one entry per
// tile plus one
void build_halo(const MeshSoA& o, int tileElems, int elemBase,
                std::vector<int>& start, std::vector<int>& cells,
                int& maxHalo);

}  // namespace ccfd
