// agglomerate.hpp - element agglomeration for the multigrid preconditioner
// of the implicit solver (host side).
//
// Gauss-Seidel sweeps reduce the error between neighbouring elements fast
// and the smooth, domain-wide part of it hardly at all; near the steady
// state that smooth part is what is left. Agglomeration multigrid solves
// for it on coarser levels: each coarse element is a group of fine ones,
// its 4x4 blocks are sums of the fine blocks (the Galerkin operator R A P
// with piecewise-constant prolongation), and a V-cycle smooths on every
// level with the same multicolour Gauss-Seidel sweeps.
//
// Groups are formed by two passes of heavy-edge matching on the element
// graph, each pass pairing every element with its most strongly coupled
// unpaired neighbour (weight: face length / barycentre distance, summed
// over the faces between groups on coarser levels). On stretched meshes
// that pairs elements across the stretched direction first, which is the
// direction the smooth error is hardest to remove in.
#pragma once

#include <cstdint>
#include <vector>

namespace ccfd {

// one coarse level, built from the level above it (level 0: the elements)
struct AggLevel {
  int n = 0;                        // elements of this level
  int nnz = 0;                      // off-diagonal blocks
  std::vector<int> aggOf;           // [elements above] -> element here
  std::vector<int> memStart, mem;   // members (elements above) of each
  std::vector<int> rowStart, col;   // off-diagonal pattern, CSR by row
  // the blocks above that sum into each diagonal / off-diagonal block:
  // (index << 1) | 1 for a diagonal block of the level above, (index << 1)
  // for an off-diagonal one. Level 1 indexes the level-0 storage: element
  // i for a diagonal block, side s of element i as s * Np + i otherwise.
  std::vector<int64_t> dSrcStart, dSrc;
  std::vector<int64_t> oSrcStart, oSrc;
  std::vector<int> colorStart, colorCells;   // elements by colour
};

// link: [S][Np] (neighbour << 2 | partner side) or negative; w: [S][Np]
// coupling weights (read where link >= 0). Levels are added until one has
// at most minCells elements or stops shrinking, at most maxLevels of them.
void build_agglomeration(int N, int Np, int S, const int* link,
                         const double* w, int minCells, int maxLevels,
                         std::vector<AggLevel>& levels);

}  // namespace ccfd
