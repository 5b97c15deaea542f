// host_sanitize_cxx.cpp - drives the host C++ of the device library
// (soa.cpp: device layout and halo lists; agglomerate.cpp: multigrid
// levels; range_alloc.hpp: the batched arena's allocator) on the gmsh
// meshes given as arguments; tests/test_host_sanitizers.py builds it with
// AddressSanitizer, UndefinedBehaviorSanitizer and leak detection. Test
// infrastructure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "agglomerate.hpp"
#include "ccfd_mesh.h"
#include "range_alloc.hpp"
#include "soa.hpp"

using namespace ccfd;

static void mesh_file(const char* path) {
  char err[256];
  long nv, nt, nq, nb;
  double* xy = nullptr;
  long *tri = nullptr, *quad = nullptr, *bce = nullptr;
  const size_t L = std::strlen(path);
  const bool emc2 = L > 5 && !std::strcmp(path + L - 5, ".mesh");
  if ((emc2 ? ccfd_emc2_read : ccfd_gmsh_read)(path, &nv, &xy, &nt, &tri, &nq,
                                               &quad, &nb, &bce, err)) {
    std::printf("read %s: %s\n", path, err);
    return;
  }
  ccfd_bc bc[32];
  int nbc = 0;
  std::memset(bc, 0, sizeof(bc));
  for (long e = 0; e < nb; ++e) {
    const int code = (int)bce[3 * e + 2];
    bool seen = false;
    for (int k = 0; k < nbc; ++k) seen |= bc[k].type * 100 + bc[k].id == code;
    if (!seen && nbc < 32) {
      bc[nbc].type = code / 100;
      bc[nbc].id = code % 100;
      bc[nbc].pVar[0] = bc[nbc].pVar[3] = 1.0;
      nbc++;
    }
  }
  ccfd_mesh* m = ccfd_mesh_build(nv, xy, nt, tri, nq, quad, nb, bce, nbc, bc,
                                 err);
  std::free(xy);
  std::free(tri);
  std::free(quad);
  std::free(bce);
  if (!m) {
    std::printf("mesh %s: %s\n", path, err);
    return;
  }
  const int S = nq > 0 ? 4 : 3;
  MeshSoA o;
  build_soa(m, S, 0, 0, 0, o);
  // face weights as the implicit solver forms them: length / distance
  std::vector<double> w((size_t)S * o.Np, 0.0);
  for (size_t c = 0; c < w.size(); ++c)
    if (o.link[c] >= 0) w[c] = std::fabs(o.flen[c]) / o.bbd[c];
  std::vector<AggLevel> levels;
  build_agglomeration(o.N, o.Np, S, o.link.data(), w.data(), 256, 12, levels);
  std::printf("mesh %s: %d elements, %d blocks, max halo %d, %zu levels\n",
              path, o.N, o.nb, o.maxHalo, levels.size());
  ccfd_mesh_free(m);
}

static void allocator(unsigned seed) {
  std::mt19937 rng(seed);
  RangeAlloc a;
  a.init(1000);
  std::vector<std::pair<int, int>> live;
  for (int step = 0; step < 20000; ++step) {
    if (!live.empty() && rng() % 2) {
      const size_t k = rng() % live.size();
      a.release(live[k].first, live[k].second);
      live.erase(live.begin() + k);
    } else {
      const int n = (int)(rng() % 120);
      const int s = a.alloc(n);
      if (s >= 0 && n > 0) live.push_back({s, n});
    }
  }
  for (auto& r : live) a.release(r.first, r.second);
  std::printf("allocator %u: used %d high %d ranges %zu\n", seed, a.used,
              a.high(), a.freeR.size());
}

int main(int argc, char* argv[]) {
  for (int i = 1; i < argc; ++i) mesh_file(argv[i]);
  for (unsigned s = 1; s <= 3; ++s) allocator(s);
  return 0;
}
