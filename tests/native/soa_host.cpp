// soa_host.cpp - C entry for the CPU tests of the device layout builder
// (ccfd_amd/csrc/soa.cpp). Test infrastructure; built by ccfd_amd/build.py
// into tests/native/.
#include <cstring>

#include "soa.hpp"

using namespace ccfd;

extern "C" void* soa_build(const ccfd_mesh* m, int S, int elemBase,
                           int ghostBase, int bcBase) {
  auto* o = new MeshSoA();
  build_soa(m, S, elemBase, ghostBase, bcBase, *o);
  return o;
}

extern "C" void soa_free(void* h) { delete (MeshSoA*)h; }

// N, Np, S, G, nb, maxHalo, number of halo entries
extern "C" void soa_sizes(void* h, long* out) {
  const MeshSoA& o = *(MeshSoA*)h;
  const long v[7] = {o.N, o.Np, o.S, o.G, o.nb, o.maxHalo,
                     (long)o.haloCells.size()};
  std::memcpy(out, v, sizeof(v));
}

static const std::vector<int>* ints(const MeshSoA& o, int k) {
  const std::vector<int>* v[8] = {&o.perm, &o.inv, &o.link, &o.ghostBC,
                                  &o.ghostCell, &o.ghostK, &o.haloStart,
                                  &o.haloCells};
  return k >= 0 && k < 8 ? v[k] : nullptr;
}

static const std::vector<double>* dbls(const MeshSoA& o, int k) {
  const std::vector<double>* v[12] = {&o.w, &o.gp, &o.fn, &o.flen,
                                      &o.bbv, &o.bbd, &o.areaq, &o.area,
                                      &o.sx, &o.sy, &o.bary, &o.ghostBary};
  return k >= 0 && k < 12 ? v[k] : nullptr;
}

// which: perm, inv, link, ghostBC, ghostCell, ghostK, haloStart, haloCells
extern "C" long soa_int_len(void* h, int k) {
  const auto* v = ints(*(MeshSoA*)h, k);
  return v ? (long)v->size() : -1;
}
extern "C" void soa_get_int(void* h, int k, int* out) {
  const auto* v = ints(*(MeshSoA*)h, k);
  if (v && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(int));
}

// which: w, gp, fn, flen, bbv, bbd, areaq, area, sx, sy, bary, ghostBary
extern "C" long soa_dbl_len(void* h, int k) {
  const auto* v = dbls(*(MeshSoA*)h, k);
  return v ? (long)v->size() : -1;
}
extern "C" void soa_get_dbl(void* h, int k, double* out) {
  const auto* v = dbls(*(MeshSoA*)h, k);
  if (v && !v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(double));
}
