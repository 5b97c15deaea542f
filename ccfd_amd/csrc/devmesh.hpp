// devmesh.hpp - request-time mesh preprocessing on the device: from raw
// connectivity (ccfd_raw_mesh) to the batched arena's structure-of-arrays
// image, without the host's ccfd_mesh_build / build_soa / build_halo.
//
// The steps are those of the host path, with the same floating-point
// expressions in the same order, so every array is bit-identical:
//  * side list and pairing (src/mesh.c:1532-1566, 1713-1755): one entry per
//    element side (creation order: elements in order, sides in node order)
//    and per boundary edge, keyed by (smaller node, larger node); a stable
//    radix sort orders them as the host's comparison sort does (boundary
//    edges are created last, so creation order also orders (node pair, bc));
//    each consecutive pair is one face, the earlier-created side owns it;
//  * createElemInfo / createSideInfo (src/mesh.c:250-341) per element and
//    per owning side, the ghost barycentre and GP of boundary sides;
//  * connectPeriodicBC (src/mesh.c:172-244): each side of an x1 periodic
//    boundary takes the first x2 side of its group, in boundary-side
//    order, within 1e-7 of its position plus the connection vector; the
//    remaining boundary sides are renumbered as ghosts in that order;
//  * createReconstructionInfo (src/mesh.c:80-114) per element;
//  * Hilbert order of the barycentres (stable radix sort of the curve
//    index) and the arena scatter with renumbered links;
//  * per 256-element block the ascending distinct ids of outside
//    neighbours (one workgroup per block: LDS gather and bitonic sort).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ccfd_mesh.h"

namespace ccfd {

// where the image goes: the arena's arrays, components strided by E (by
// capG for ghosts), this instance at element offset e0 and ghost offset g0;
// ids written into them are offset by e0 (elements), g0 (ghosts) and bcBase
struct DevArenaDst {
  int S = 3, E = 0, e0 = 0, capG = 0, g0 = 0, bcBase = 0;
  int ts = 8;             // 8: double operator arrays, 4: float
  double venkK = 0.0;     // Venkatakrishnan K for the eps^2 array
  int* link = nullptr;
  void *w = nullptr, *gp = nullptr, *fn = nullptr, *flen = nullptr,
       *areaq = nullptr, *venk = nullptr;
  double *area = nullptr, *sx = nullptr, *sy = nullptr, *bary = nullptr;
  int *ghostBC = nullptr, *ghostCell = nullptr, *ghostK = nullptr;
  double* ghostBary = nullptr;
};

class DevMesh {
 public:
  DevMesh() = default;
  DevMesh(const DevMesh&) = delete;
  DevMesh& operator=(const DevMesh&) = delete;
  ~DevMesh();

  // check the input on the host, then pair sides, build the geometry and
  // the element order on the device (stream st). Returns CCFD_OK or a
  // status with the message in err (at least 256 bytes).
  int build(const ccfd_raw_mesh& m, hipStream_t st, char* err);

  // after build
  int N = 0;              // elements
  int G = 0;              // ghosts (boundary sides)
  int sides = 3;          // 3, or 4 with quads
  std::vector<int> perm;  // image position -> element (reference order)
  double totalArea_q = 0.0;

  // write the image of the elements, slots and ghosts into d
  int write(const DevArenaDst& d, char* err);
  // halo list sizes per block of the written image, then the lists
  // themselves at dst + start[block] (start: one entry per block)
  int halo_counts(const DevArenaDst& d, std::vector<int>& counts, char* err);
  int halo_write(const DevArenaDst& d, const std::vector<int>& start,
                 int* dst, char* err);

 private:
  void* alloc(size_t bytes);
  std::vector<void*> allocs_;
  bool oom_ = false;
  hipStream_t st_ = nullptr;
  int nTri_ = 0, nSlots_ = 0;
  double* xy_ = nullptr;
  int* en_ = nullptr;
  int *nbr_ = nullptr, *partner_ = nullptr;
  unsigned char* owner_ = nullptr;
  double *len_ = nullptr, *flen_ = nullptr;
  double *n_ = nullptr, *gp_ = nullptr, *gpc_ = nullptr, *fn_ = nullptr,
         *w_ = nullptr;
  double *bary_ = nullptr, *area_ = nullptr, *areaq_ = nullptr, *sx_ = nullptr,
         *sy_ = nullptr;
  int *gSlot_ = nullptr, *gBC_ = nullptr;
  double *gBary_ = nullptr, *gGP_ = nullptr;
  int *perm_ = nullptr, *inv_ = nullptr;
};

}  // namespace ccfd
