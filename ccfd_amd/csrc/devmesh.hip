// devmesh.hip - request-time mesh preprocessing on the device (devmesh.hpp).
// Each kernel restates one step of the host builder (ccfd_amd/csrc/mesh.c,
// soa.cpp) with its expressions in the same order; -ffp-contract=off and
// IEEE division and square root keep the results bit-identical.
#include "devmesh.hpp"

#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "ccfd_gpu.h"
#include "layout.hpp"

namespace ccfd {
namespace {

constexpr int kT = 256;
inline unsigned nblk(long n) { return (unsigned)((n + kT - 1) / kT); }

#define DM_ERR(code, ...)                                  \
  do {                                                     \
    if (err) snprintf(err, 256, __VA_ARGS__);              \
    return (code);                                         \
  } while (0)

// slots: the 3 of each triangle, then the 4 of each quad, in element order
__device__ __forceinline__ int elem_type(int e, int nTri) {
  return e < nTri ? 3 : 4;
}
__device__ __forceinline__ int slot_start(int e, int nTri) {
  return e < nTri ? 3 * e : 3 * nTri + 4 * (e - nTri);
}
__device__ __forceinline__ int slot_elem(int s, int nTri) {
  return s < 3 * nTri ? s / 3 : nTri + (s - 3 * nTri) / 4;
}
// side is of element e is created as entry slot_start(e) + is and sits at
// list position type - 1 - is (sides are prepended, src/mesh.c:1532-1566)
__device__ __forceinline__ int create_slot(int c, int nTri) {
  const int e = slot_elem(c, nTri), s0 = slot_start(e, nTri);
  return s0 + (elem_type(e, nTri) - 1 - (c - s0));
}

// one side-list entry per element side and per boundary edge
__global__ void k_entries(int nTri, int nSlots, int nB, unsigned long long nV,
                          const int* __restrict__ en, const int* __restrict__ bce,
                          unsigned long long* __restrict__ key,
                          int* __restrict__ val) {
  const int c = blockIdx.x * kT + threadIdx.x;
  if (c >= nSlots + nB) return;
  int a, b;
  if (c < nSlots) {
    const int e = slot_elem(c, nTri), t = elem_type(e, nTri);
    const int is = c - slot_start(e, nTri);
    a = en[4 * e + is];
    b = en[4 * e + (is + 1) % t];
  } else {
    const int k = c - nSlots;
    a = bce[2 * k];
    b = bce[2 * k + 1];
  }
  const unsigned long long lo = (unsigned long long)(a < b ? a : b);
  const unsigned long long hi = (unsigned long long)(a < b ? b : a);
  key[c] = lo * nV + hi;
  val[c] = c;
}

// consecutive sorted entries pair up; flag boundary pairs. err[0]: 1 an
// unmatched side, 2 two boundary edges on one face; err[1]: first pair
__global__ void k_pair_check(int nPairs, int nSlots,
                             const unsigned long long* __restrict__ key,
                             const int* __restrict__ val, int* __restrict__ isBC,
                             int* __restrict__ err) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p >= nPairs) return;
  const bool a = val[2 * p] >= nSlots, b = val[2 * p + 1] >= nSlots;
  isBC[p] = (a || b) ? 1 : 0;
  int code = 0;
  if (key[2 * p] != key[2 * p + 1]) code = 1;
  else if (a && b) code = 2;
  if (code) {
    atomicMax(err, code);
    atomicMin(err + 1, p);
  }
}

// pairing (src/mesh.c:1713-1755): the earlier-created side owns the face;
// a boundary pair makes ghost kb[p] (pair order) of its side
__global__ void k_pair_link(int nPairs, int nSlots, int nTri,
                            const int* __restrict__ val,
                            const int* __restrict__ isBC,
                            const int* __restrict__ kb,
                            const int* __restrict__ edgeBC,
                            int* __restrict__ nbr, int* __restrict__ partner,
                            unsigned char* __restrict__ owner,
                            int* __restrict__ gSlot, int* __restrict__ gBC) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p >= nPairs) return;
  int cA = val[2 * p], cB = val[2 * p + 1];
  if (isBC[p]) {
    if (cA >= nSlots) {
      const int t = cA;
      cA = cB;
      cB = t;
    }
    const int s = create_slot(cA, nTri), g = kb[p];
    nbr[s] = -(g + 1);
    partner[s] = -1;
    owner[s] = 1;
    gSlot[g] = s;
    gBC[g] = edgeBC[cB - nSlots];
  } else {
    const int sa = create_slot(cA, nTri), sb = create_slot(cB, nTri);
    nbr[sa] = slot_elem(sb, nTri);
    nbr[sb] = slot_elem(sa, nTri);
    partner[sa] = sb;
    partner[sb] = sa;
    owner[sa] = 1;
    owner[sb] = 0;
  }
}

// createElemInfo (src/mesh.c:296-341) without the running total
__global__ void k_elem_info(int N, int nTri, const double* __restrict__ xy,
                            const int* __restrict__ en, double* __restrict__ bary,
                            double* __restrict__ area, double* __restrict__ areaq,
                            double* __restrict__ sx, double* __restrict__ sy) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= N) return;
  const int type = elem_type(e, nTri);
  const int* nd = en + 4 * e;
  double bx = 0.0, by = 0.0;
  for (int i = 0; i < type; ++i) {
    bx += xy[2 * nd[i]] / type;
    by += xy[2 * nd[i] + 1] / type;
  }
  bary[2 * e] = bx;
  bary[2 * e + 1] = by;
  double len[4] = {0.0, 0.0, 0.0, 0.0}, nx[4] = {0.0, 0.0, 0.0, 0.0},
         ny[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = 0; i < type; ++i) {
    const int j = (i + 1) % type;
    const double xi = xy[2 * nd[i]], yi = xy[2 * nd[i] + 1];
    const double xj = xy[2 * nd[j]], yj = xy[2 * nd[j] + 1];
    const double vx = xj - xi, vy = yj - yi;
    len[i] = sqrt(vx * vx + vy * vy);
    nx[i] = yj - yi;
    ny[i] = xi - xj;
    const double t = sqrt(nx[i] * nx[i] + ny[i] * ny[i]);
    nx[i] /= t;
    ny[i] /= t;
  }
  double a = 0.5 * len[0] * len[1] * fabs(nx[0] * ny[1] - nx[1] * ny[0]);
  if (type == 4) a += 0.5 * len[2] * len[3] * fabs(nx[2] * ny[3] - nx[3] * ny[2]);
  area[e] = a;
  areaq[e] = 1.0 / a;
  double s_x = 0.0, s_y = 0.0;
  for (int i = 0; i < type; ++i) {
    s_x += fabs(nx[i]) * len[i] / 2.0;
    s_y += fabs(ny[i]) * len[i] / 2.0;
  }
  sx[e] = s_x;
  sy[e] = s_y;
}

// createSideInfo (src/mesh.c:250-289) for every owning side: its own
// normal, length and GP, the connected side's (or the ghost's) GP, normal
// and length, and the ghost barycentre (src/mesh.c:275-282)
__global__ void k_side_info(int nSlots, int nTri, const double* __restrict__ xy,
                            const int* __restrict__ en,
                            const double* __restrict__ bary,
                            const int* __restrict__ nbr,
                            const int* __restrict__ partner,
                            const unsigned char* __restrict__ owner,
                            double* __restrict__ n, double* __restrict__ len,
                            double* __restrict__ GP, double* __restrict__ gBary,
                            double* __restrict__ gGP) {
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s >= nSlots || !owner[s]) return;
  const int e = slot_elem(s, nTri), type = elem_type(e, nTri);
  const int is = type - 1 - (s - slot_start(e, nTri));
  const int a = en[4 * e + is], b = en[4 * e + (is + 1) % type];
  const double x0 = xy[2 * a], y0 = xy[2 * a + 1];
  const double x1 = xy[2 * b], y1 = xy[2 * b + 1];
  double nx = y1 - y0, ny = x0 - x1;
  const double l = sqrt(nx * nx + ny * ny);
  nx /= l;
  ny /= l;
  const double gx = 0.5 * (x0 + x1), gy = 0.5 * (y0 + y1);
  const double GPx = gx - bary[2 * e], GPy = gy - bary[2 * e + 1];
  if ((nx * GPx + ny * GPy) < 0.0) {
    nx *= -1.0;
    ny *= -1.0;
  }
  n[2 * s] = nx;
  n[2 * s + 1] = ny;
  len[s] = l;
  GP[2 * s] = GPx;
  GP[2 * s + 1] = GPy;
  const int c = nbr[s];
  if (c < 0) {
    const int g = -c - 1;
    const double tmp = 2.0 * fabs(GPx * nx + GPy * ny);
    const double cbx = bary[2 * e] + tmp * nx;
    const double cby = bary[2 * e + 1] + tmp * ny;
    gBary[2 * g] = cbx;
    gBary[2 * g + 1] = cby;
    gGP[2 * g] = gx - cbx;
    gGP[2 * g + 1] = gy - cby;
  } else {
    const int t = partner[s];
    const double cbx = bary[2 * c], cby = bary[2 * c + 1];
    n[2 * t] = -nx;
    n[2 * t + 1] = -ny;
    len[t] = l;
    GP[2 * t] = gx - cbx;
    GP[2 * t + 1] = gy - cby;
  }
}

// connectPeriodicBC (src/mesh.c:172-244): per boundary condition its
// periodic class (1: the side to connect, 2: a candidate, 0: neither), its
// group (id / 10) and connection vector
struct PerBC {
  int cls[kMaxBC], group[kMaxBC];
  double conn[kMaxBC][2];
};

__global__ void k_per_flag(int G, const int* __restrict__ gBC, PerBC pb,
                           int* __restrict__ isCand) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  const int b = gBC[g];
  isCand[g] = (b >= 0 && pb.cls[b] == 2) ? 1 : 0;
}

__global__ void k_per_gather(int G, const int* __restrict__ isCand,
                             const int* __restrict__ pos, int* __restrict__ cand) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g < G && isCand[g]) cand[pos[g]] = g;
}

// each side to connect takes the first candidate of its group (boundary-
// side order) within 1e-7 (L1) of its GP position plus the connection;
// hit[g] = the candidate's ghost index, -1 if none (err[0] = 3, err[1] = g)
__global__ void k_per_match(int G, int nCand, const int* __restrict__ gBC,
                            PerBC pb, const int* __restrict__ cand,
                            const double* __restrict__ gGP,
                            const double* __restrict__ gBary,
                            int* __restrict__ hit, int* __restrict__ err) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  const int b = gBC[g];
  hit[g] = -2;   // not a side to connect
  if (b < 0 || pb.cls[b] != 1) return;
  const double ax = gGP[2 * g] + gBary[2 * g];
  const double ay = gGP[2 * g + 1] + gBary[2 * g + 1];
  int h = -1;
  for (int j = 0; j < nCand && h < 0; ++j) {
    const int kk = cand[j];
    if (pb.group[gBC[kk]] != pb.group[b]) continue;
    const double sxp = gGP[2 * kk] + gBary[2 * kk];
    const double syp = gGP[2 * kk + 1] + gBary[2 * kk + 1];
    if ((fabs(ax + pb.conn[b][0] - sxp) + fabs(ay + pb.conn[b][1] - syp)) <= 1e-7)
      h = kk;
  }
  hit[g] = h;
  if (h < 0) {
    atomicMax(err, 3);
    atomicMin(err + 1, g);
  }
}

// connect each matched pair (the side to connect owns the face) and mark
// both ghosts as gone
__global__ void k_per_link(int G, int nTri, const int* __restrict__ hit,
                           const int* __restrict__ gSlot, int* __restrict__ nbr,
                           int* __restrict__ partner,
                           unsigned char* __restrict__ owner,
                           int* __restrict__ keep) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= G || hit[g] < 0) return;
  const int h = hit[g];
  const int ur = gSlot[g], tg = gSlot[h];
  nbr[ur] = slot_elem(tg, nTri);
  nbr[tg] = slot_elem(ur, nTri);
  partner[ur] = tg;
  partner[tg] = ur;
  owner[ur] = 1;
  owner[tg] = 0;
  keep[g] = 0;
  keep[h] = 0;
}

// the remaining ghosts, renumbered in boundary-side order
__global__ void k_ghost_compact(int G, const int* __restrict__ keep,
                                const int* __restrict__ pos,
                                const int* __restrict__ gSlot,
                                const int* __restrict__ gBC,
                                const double* __restrict__ gBary,
                                const double* __restrict__ gGP,
                                int* __restrict__ nbr, int* __restrict__ oSlot,
                                int* __restrict__ oBC, double* __restrict__ oBary,
                                double* __restrict__ oGP) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= G || !keep[g]) return;
  const int n = pos[g], s = gSlot[g];
  oSlot[n] = s;
  oBC[n] = gBC[g];
  oBary[2 * n] = gBary[2 * g];
  oBary[2 * n + 1] = gBary[2 * g + 1];
  oGP[2 * n] = gGP[2 * g];
  oGP[2 * n + 1] = gGP[2 * g + 1];
  nbr[s] = -(n + 1);
}

// connection GP and the face's normal and length on every slot
__global__ void k_slot_conn(int nSlots, const int* __restrict__ nbr,
                            const int* __restrict__ partner,
                            const unsigned char* __restrict__ owner,
                            const double* __restrict__ n,
                            const double* __restrict__ len,
                            const double* __restrict__ GP,
                            const double* __restrict__ gGP,
                            double* __restrict__ GPc, double* __restrict__ fn,
                            double* __restrict__ flen) {
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s >= nSlots) return;
  const int c = nbr[s];
  if (c < 0) {
    const int g = -c - 1;
    GPc[2 * s] = gGP[2 * g];
    GPc[2 * s + 1] = gGP[2 * g + 1];
  } else {
    const int t = partner[s];
    GPc[2 * s] = GP[2 * t];
    GPc[2 * s + 1] = GP[2 * t + 1];
  }
  const int o = (owner[s] || c < 0) ? s : partner[s];
  fn[2 * s] = n[2 * o];
  fn[2 * s + 1] = n[2 * o + 1];
  flen[s] = len[o];
}

// createReconstructionInfo (src/mesh.c:80-114)
__global__ void k_recon_w(int N, int nTri, const double* __restrict__ GP,
                          const double* __restrict__ GPc, double* __restrict__ w) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= N) return;
  const int s0 = slot_start(e, nTri), s1 = s0 + elem_type(e, nTri);
  double r11 = 0.0, r12 = 0.0, r22 = 0.0;
  for (int s = s0; s < s1; ++s) {
    const double bx = GP[2 * s] - GPc[2 * s];
    const double by = GP[2 * s + 1] - GPc[2 * s + 1];
    r11 += bx * bx;
    r12 += bx * by;
    r22 += by * by;
  }
  r11 = sqrt(r11);
  r12 = r12 / r11;
  r22 = sqrt(r22 - r12 * r12);
  for (int s = s0; s < s1; ++s) {
    const double bx = GP[2 * s] - GPc[2 * s];
    const double by = GP[2 * s + 1] - GPc[2 * s + 1];
    const double alpha1 = bx / (r11 * r11);
    const double alpha2 = (by - r12 / r11 * bx) / (r22 * r22);
    w[2 * s] = alpha1 - r12 / r11 * alpha2;
    w[2 * s + 1] = alpha2;
  }
}

// bounding box of the barycentres: per-block partials, then one block
__device__ inline void box_merge(double* a, const double* b) {
  a[0] = fmin(a[0], b[0]);
  a[1] = fmax(a[1], b[1]);
  a[2] = fmin(a[2], b[2]);
  a[3] = fmax(a[3], b[3]);
}

__global__ void k_box(int n, int stride, const double* __restrict__ in,
                      double* __restrict__ out) {
  // in: n boxes ([4] each at in + 4k) when stride == 4, else n points
  // (x, y) at in + 2k
  __shared__ double sh[kT][4];
  double b[4] = {1e300, -1e300, 1e300, -1e300};
  for (int k = blockIdx.x * kT + threadIdx.x; k < n; k += gridDim.x * kT) {
    if (stride == 4) {
      box_merge(b, in + 4 * k);
    } else {
      const double p[4] = {in[2 * k], in[2 * k], in[2 * k + 1], in[2 * k + 1]};
      box_merge(b, p);
    }
  }
  for (int v = 0; v < 4; ++v) sh[threadIdx.x][v] = b[v];
  __syncthreads();
  for (int h = kT / 2; h > 0; h /= 2) {
    if ((int)threadIdx.x < h) box_merge(sh[threadIdx.x], sh[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < 4; ++v) out[4 * blockIdx.x + v] = sh[0][v];
}

// Hilbert curve index of (x, y) on an n x n grid (soa.cpp)
__device__ uint32_t hilbert_index(uint32_t n, uint32_t x, uint32_t y) {
  uint32_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) > 0;
    const uint32_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return d;
}

__global__ void k_sfc_key(int N, const double* __restrict__ bary,
                          const double* __restrict__ box,
                          uint32_t* __restrict__ key, int* __restrict__ idx) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= N) return;
  const double x0 = box[0], x1 = box[1], y0 = box[2], y1 = box[3];
  const double ax = x1 - x0, ay = y1 - y0;
  double span = (ax < ay) ? ay : ax;
  if (!(span > 0)) span = 1.0;
  const double fx = (bary[2 * e] - x0) / span * 65535.0;
  const double fy = (bary[2 * e + 1] - y0) / span * 65535.0;
  const uint32_t ix = (uint32_t)((fx < 65535.0) ? fx : 65535.0);
  const uint32_t iy = (uint32_t)((fy < 65535.0) ? fy : 65535.0);
  key[e] = hilbert_index(65536u, ix, iy);
  idx[e] = e;
}

__global__ void k_fill_one(int n, int* __restrict__ p) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < n) p[i] = 1;
}

__global__ void k_inverse(int N, const int* __restrict__ perm,
                          int* __restrict__ inv) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i < N) inv[perm[i]] = i;
}

// the arena image of elements and slots (build_soa, soa.cpp); padding
// elements and missing slots get the host builder's fill values
template <typename T>
__global__ void k_write_elems(int N, int Np, int nTri, DevArenaDst d,
                              const int* __restrict__ perm,
                              const int* __restrict__ inv,
                              const int* __restrict__ nbr,
                              const int* __restrict__ partner,
                              const unsigned char* __restrict__ owner,
                              const double* __restrict__ w,
                              const double* __restrict__ GP,
                              const double* __restrict__ fn,
                              const double* __restrict__ flen,
                              const double* __restrict__ areaq,
                              const double* __restrict__ area,
                              const double* __restrict__ sx,
                              const double* __restrict__ sy,
                              const double* __restrict__ bary) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= Np) return;
  const size_t E = (size_t)d.E, c0 = (size_t)d.e0 + i;
  const int S = d.S;
  T* W = (T*)d.w;
  T* Gp = (T*)d.gp;
  T* Fn = (T*)d.fn;
  T* Fl = (T*)d.flen;
  int e = -1, type = 0, s0 = 0;
  if (i < N) {
    e = perm[i];
    type = elem_type(e, nTri);
    s0 = slot_start(e, nTri);
    const double a = area[e], kk = d.venkK;
    ((T*)d.areaq)[c0] = T(areaq[e]);
    ((T*)d.venk)[c0] = T((kk * sqrt(a)) * (kk * sqrt(a)) * (kk * sqrt(a)));
    d.area[c0] = a;
    d.sx[c0] = sx[e];
    d.sy[c0] = sy[e];
    d.bary[c0] = bary[2 * e];
    d.bary[E + c0] = bary[2 * e + 1];
  } else {
    ((T*)d.areaq)[c0] = T(0.0);
    ((T*)d.venk)[c0] = T(0.0);
    d.area[c0] = 0.0;
    d.sx[c0] = 1.0;
    d.sy[c0] = 1.0;
    d.bary[c0] = 0.0;
    d.bary[E + c0] = 0.0;
  }
  for (int k = 0; k < S; ++k) {
    const size_t c = (size_t)k * E + c0, cy = (size_t)(S + k) * E + c0;
    if (k < type) {
      const int s = s0 + k, nb = nbr[s];
      d.link[c] = nb >= 0
                      ? ((inv[nb] + d.e0) << 2) | (partner[s] - slot_start(nb, nTri))
                      : nb - d.g0;
      W[c] = T(w[2 * s]);
      W[cy] = T(w[2 * s + 1]);
      Gp[c] = T(GP[2 * s]);
      Gp[cy] = T(GP[2 * s + 1]);
      Fn[c] = T(fn[2 * s]);
      Fn[cy] = T(fn[2 * s + 1]);
      Fl[c] = T((owner[s] || nb < 0) ? flen[s] : -flen[s]);
    } else {
      d.link[c] = kNoSide;
      W[c] = W[cy] = T(0.0);
      Gp[c] = Gp[cy] = T(0.0);
      Fn[c] = Fn[cy] = T(0.0);
      Fl[c] = T(1.0);
    }
  }
}

__global__ void k_write_ghosts(int G, int nTri, DevArenaDst d,
                               const int* __restrict__ inv,
                               const int* __restrict__ gSlot,
                               const int* __restrict__ gBC,
                               const double* __restrict__ gBary) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g >= G) return;
  const int s = gSlot[g], e = slot_elem(s, nTri);
  const size_t o = (size_t)d.g0 + g;
  d.ghostCell[o] = inv[e] + d.e0;
  d.ghostK[o] = s - slot_start(e, nTri);
  d.ghostBC[o] = gBC[g] + d.bcBase;
  d.ghostBary[o] = gBary[2 * g];
  d.ghostBary[(size_t)d.capG + o] = gBary[2 * g + 1];
}

// halo of block jb (build_halo, soa.cpp): the ascending distinct ids of
// elements outside [lo, lo + kBlock) linked from the block's elements.
// Counts only (out == nullptr) or writes them at out + start[jb].
constexpr int kHaloCand = kBlock * kMaxSlots;

__global__ __launch_bounds__(kBlock) void k_halo(int N, DevArenaDst d,
                                                 int* __restrict__ count,
                                                 const int* __restrict__ start,
                                                 int* __restrict__ out) {
  __shared__ int cand[kHaloCand];
  __shared__ int nc;
  const int jb = blockIdx.x, t = threadIdx.x;
  if (t == 0) nc = 0;
  __syncthreads();
  const int i = jb * kBlock + t;
  const int lo = jb * kBlock + d.e0, hi = lo + kBlock;
  if (i < N)
    for (int k = 0; k < d.S; ++k) {
      const int l = d.link[(size_t)k * d.E + d.e0 + i];
      if (l < 0) continue;
      const int j = l >> 2;
      if (j < lo || j >= hi) cand[atomicAdd(&nc, 1)] = j;
    }
  __syncthreads();
  const int n = nc;
  for (int q = t; q < kHaloCand; q += kBlock)
    if (q >= n) cand[q] = INT_MAX;
  __syncthreads();
  // bitonic sort of the kHaloCand entries, ascending
  for (int size = 2; size <= kHaloCand; size <<= 1)
    for (int stride = size / 2; stride > 0; stride >>= 1) {
      for (int p = t; p < kHaloCand / 2; p += kBlock) {
        const int a = 2 * p - (p & (stride - 1)), b = a + stride;
        const bool up = (a & size) == 0;
        const int x = cand[a], y = cand[b];
        if ((x > y) == up) {
          cand[a] = y;
          cand[b] = x;
        }
      }
      __syncthreads();
    }
  if (t == 0) {
    int u = 0;
    for (int q = 0; q < n; ++q)
      if (q == 0 || cand[q] != cand[q - 1]) {
        if (out) out[start[jb] + u] = cand[q];
        ++u;
      }
    if (count) count[jb] = u;
  }
}

}  // namespace

DevMesh::~DevMesh() {
  if (st_) hipStreamSynchronize(st_);
  for (void* p : allocs_) hipFree(p);
}

void* DevMesh::alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) {
    oom_ = true;
    return nullptr;
  }
  allocs_.push_back(p);
  return p;
}

int DevMesh::build(const ccfd_raw_mesh& m, hipStream_t st, char* err) {
  st_ = st;
  if (m.nVert <= 0 || m.nTri < 0 || m.nQuad < 0 || m.nBCedge < 0 || !m.xy ||
      (m.nTri && !m.tri) || (m.nQuad && !m.quad) || (m.nBCedge && !m.bcEdge) ||
      (m.nBC && !m.bc) || m.nBC < 0)
    DM_ERR(CCFD_EINVAL, "bad raw mesh arguments");
  const long nE = m.nTri + m.nQuad;
  if (nE <= 0) DM_ERR(CCFD_EINVAL, "empty mesh");
  if (nE > kMaxIndexedElems / kMaxSlots || m.nVert >= INT_MAX ||
      m.nBCedge > INT_MAX / 4)
    DM_ERR(CCFD_EINVAL, "mesh too large for 32-bit indices");
  const int N = (int)nE, nTri = (int)m.nTri, nB = (int)m.nBCedge;
  const int nSlots = 3 * nTri + 4 * (int)m.nQuad;
  if ((nSlots + (long)nB) % 2 != 0)
    DM_ERR(CCFD_EINVAL, "side count %d + %d BC edges is odd: open mesh", nSlots, nB);
  // host: node ids in range, boundary codes resolved (as ccfd_mesh_build)
  std::vector<int> en(4 * (size_t)N, -1), bce(2 * (size_t)nB), edgeBC(nB);
  for (int e = 0; e < N; ++e) {
    const int type = e < nTri ? 3 : 4;
    const long* src = e < nTri ? m.tri + 4 * (size_t)e : m.quad + 5 * (size_t)(e - nTri);
    for (int i = 0; i < type; ++i) {
      if (src[i] < 0 || src[i] >= m.nVert)
        DM_ERR(CCFD_EINVAL, "element %d references node %ld (of %ld)", e, src[i],
               m.nVert);
      en[4 * (size_t)e + i] = (int)src[i];
    }
  }
  for (int k = 0; k < nB; ++k) {
    const long a = m.bcEdge[3 * (size_t)k], b = m.bcEdge[3 * (size_t)k + 1];
    const long code = m.bcEdge[3 * (size_t)k + 2];
    const int type = (int)(code / 100), id = (int)(code % 100);
    edgeBC[k] = -1;
    for (int q = 0; q < m.nBC; ++q)
      if (m.bc[q].type == type && m.bc[q].id == id) {
        edgeBC[k] = q;
        break;
      }
    if (type != 0 && edgeBC[k] < 0)
      DM_ERR(CCFD_EINVAL, "BC %ld from mesh file not defined in parameter file", code);
    if (a < 0 || a >= m.nVert || b < 0 || b >= m.nVert)
      DM_ERR(CCFD_EINVAL, "BC edge %d references a missing node", k);
    bce[2 * (size_t)k] = (int)a;
    bce[2 * (size_t)k + 1] = (int)b;
  }
  this->N = N;
  nTri_ = nTri;
  nSlots_ = nSlots;
  sides = m.nQuad > 0 ? 4 : 3;

  const int nEnt = nSlots + nB, nPairs = nEnt / 2;
  xy_ = (double*)alloc(16 * (size_t)m.nVert);
  en_ = (int*)alloc(en.size() * 4);
  int* bceD = (int*)alloc(bce.size() * 4 + 4);
  int* ebcD = (int*)alloc(edgeBC.size() * 4 + 4);
  auto* keyA = (unsigned long long*)alloc(8 * (size_t)nEnt);
  auto* keyB = (unsigned long long*)alloc(8 * (size_t)nEnt);
  int* valA = (int*)alloc(4 * (size_t)nEnt);
  int* valB = (int*)alloc(4 * (size_t)nEnt);
  int* isBC = (int*)alloc(4 * (size_t)nPairs + 4);
  int* kb = (int*)alloc(4 * (size_t)nPairs + 4);
  int* errD = (int*)alloc(16);
  nbr_ = (int*)alloc(4 * (size_t)nSlots);
  partner_ = (int*)alloc(4 * (size_t)nSlots);
  owner_ = (unsigned char*)alloc(nSlots);
  n_ = (double*)alloc(16 * (size_t)nSlots);
  len_ = (double*)alloc(8 * (size_t)nSlots);
  gp_ = (double*)alloc(16 * (size_t)nSlots);
  gpc_ = (double*)alloc(16 * (size_t)nSlots);
  fn_ = (double*)alloc(16 * (size_t)nSlots);
  flen_ = (double*)alloc(8 * (size_t)nSlots);
  w_ = (double*)alloc(16 * (size_t)nSlots);
  bary_ = (double*)alloc(16 * (size_t)N);
  area_ = (double*)alloc(8 * (size_t)N);
  areaq_ = (double*)alloc(8 * (size_t)N);
  sx_ = (double*)alloc(8 * (size_t)N);
  sy_ = (double*)alloc(8 * (size_t)N);
  gSlot_ = (int*)alloc(4 * (size_t)nB + 4);
  gBC_ = (int*)alloc(4 * (size_t)nB + 4);
  gBary_ = (double*)alloc(16 * (size_t)nB + 16);
  gGP_ = (double*)alloc(16 * (size_t)nB + 16);
  uint32_t* sk = (uint32_t*)alloc(4 * (size_t)N);
  uint32_t* sk2 = (uint32_t*)alloc(4 * (size_t)N);
  int* idx = (int*)alloc(4 * (size_t)N);
  perm_ = (int*)alloc(4 * (size_t)N);
  inv_ = (int*)alloc(4 * (size_t)N);
  const unsigned nbox = std::min(nblk(N), 1024u);
  double* box = (double*)alloc(32 * (size_t)nbox + 32);
  // radix sort / scan scratch
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  int endBits = 1;
  while (endBits < 64 && ((unsigned long long)m.nVert * (unsigned long long)m.nVert >> endBits))
    ++endBits;
  hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, keyA, keyB, valA, valB, nEnt, 0,
                                     endBits, st);
  hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, isBC, kb, nPairs, st);
  hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, sk, sk2, idx, perm_, N, 0, 32, st);
  void* tmp = alloc(std::max(tb1, std::max(tb2, tb3)));
  if (oom_) DM_ERR(CCFD_ENOMEM, "device allocation for the mesh build failed");

  hipMemcpyAsync(xy_, m.xy, 16 * (size_t)m.nVert, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(en_, en.data(), en.size() * 4, hipMemcpyHostToDevice, st);
  if (nB) {
    hipMemcpyAsync(bceD, bce.data(), bce.size() * 4, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(ebcD, edgeBC.data(), edgeBC.size() * 4, hipMemcpyHostToDevice, st);
  }
  const int err0[2] = {0, INT_MAX};
  hipMemcpyAsync(errD, err0, 8, hipMemcpyHostToDevice, st);

  // side list, sorted, paired
  hipLaunchKernelGGL(k_entries, dim3(nblk(nEnt)), dim3(kT), 0, st, nTri, nSlots, nB,
                     (unsigned long long)m.nVert, en_, bceD, keyA, valA);
  size_t tbs = tb1;
  hipcub::DeviceRadixSort::SortPairs(tmp, tbs, keyA, keyB, valA, valB, nEnt, 0, endBits,
                                     st);
  hipLaunchKernelGGL(k_pair_check, dim3(nblk(nPairs)), dim3(kT), 0, st, nPairs, nSlots,
                     keyB, valB, isBC, errD);
  int herr[2] = {0, 0};
  hipMemcpyAsync(herr, errD, 8, hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess)
    DM_ERR(CCFD_EHIP, "side pairing on the device failed");
  if (herr[0]) {
    unsigned long long k2 = 0;
    hipMemcpy(&k2, keyB + 2 * (size_t)herr[1], 8, hipMemcpyDeviceToHost);
    const long n0 = (long)(k2 / (unsigned long long)m.nVert);
    const long n1 = (long)(k2 % (unsigned long long)m.nVert);
    if (herr[0] == 1) DM_ERR(CCFD_EINVAL, "problem in connect: unmatched side (%ld,%ld)", n0, n1);
    DM_ERR(CCFD_EINVAL, "two boundary edges share nodes (%ld,%ld)", n0, n1);
  }
  tbs = tb2;
  hipcub::DeviceScan::ExclusiveSum(tmp, tbs, isBC, kb, nPairs, st);
  hipLaunchKernelGGL(k_pair_link, dim3(nblk(nPairs)), dim3(kT), 0, st, nPairs, nSlots,
                     nTri, valB, isBC, kb, ebcD, nbr_, partner_, owner_, gSlot_, gBC_);

  // geometry
  hipLaunchKernelGGL(k_elem_info, dim3(nblk(N)), dim3(kT), 0, st, N, nTri, xy_, en_,
                     bary_, area_, areaq_, sx_, sy_);
  hipLaunchKernelGGL(k_side_info, dim3(nblk(nSlots)), dim3(kT), 0, st, nSlots, nTri,
                     xy_, en_, bary_, nbr_, partner_, owner_, n_, len_, gp_, gBary_,
                     gGP_);
  // periodic hookup and ghost compaction
  int G0 = nPairs > 0 ? -1 : 0;
  {
    PerBC pb;
    memset(&pb, 0, sizeof(pb));
    bool periodic = false;
    for (int q = 0; q < m.nBC && q < kMaxBC; ++q) {
      if (m.bc[q].type != CCFD_BC_PERIODIC) continue;
      pb.cls[q] = m.bc[q].id % 10 == 1 ? 1 : m.bc[q].id % 10 == 2 ? 2 : 0;
      pb.group[q] = m.bc[q].id / 10;
      pb.conn[q][0] = m.bc[q].connection[0];
      pb.conn[q][1] = m.bc[q].connection[1];
      periodic = periodic || pb.cls[q] != 0;
    }
    // the boundary pairs are known once the pairing's scan has run
    if (nPairs > 0) {
      int lk = 0, lb = 0;
      hipMemcpyAsync(&lk, kb + nPairs - 1, 4, hipMemcpyDeviceToHost, st);
      hipMemcpyAsync(&lb, isBC + nPairs - 1, 4, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess)
        DM_ERR(CCFD_EHIP, "side pairing on the device failed");
      G0 = lk + lb;
    }
    if (G0 != nB) DM_ERR(CCFD_EINVAL, "%d boundary edges but %d boundary sides", nB, G0);
    G = G0;
    if (periodic && G0 > 0) {
      int* isCand = (int*)alloc(4 * (size_t)G0);
      int* pos = (int*)alloc(4 * (size_t)G0);
      int* cand = (int*)alloc(4 * (size_t)G0);
      int* hit = (int*)alloc(4 * (size_t)G0);
      int* keep = (int*)alloc(4 * (size_t)G0);
      int* oSlot = (int*)alloc(4 * (size_t)G0);
      int* oBC = (int*)alloc(4 * (size_t)G0);
      double* oBary = (double*)alloc(16 * (size_t)G0);
      double* oGP = (double*)alloc(16 * (size_t)G0);
      size_t tb4 = 0;
      hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, isCand, pos, G0, st);
      void* tmp4 = alloc(tb4);
      if (oom_) DM_ERR(CCFD_ENOMEM, "device allocation for the periodic hookup failed");
      hipLaunchKernelGGL(k_per_flag, dim3(nblk(G0)), dim3(kT), 0, st, G0, gBC_, pb, isCand);
      hipcub::DeviceScan::ExclusiveSum(tmp4, tb4, isCand, pos, G0, st);
      int nCand = 0, lastC = 0;
      hipMemcpyAsync(&nCand, pos + G0 - 1, 4, hipMemcpyDeviceToHost, st);
      hipMemcpyAsync(&lastC, isCand + G0 - 1, 4, hipMemcpyDeviceToHost, st);
      hipStreamSynchronize(st);
      nCand += lastC;
      hipLaunchKernelGGL(k_per_gather, dim3(nblk(G0)), dim3(kT), 0, st, G0, isCand, pos, cand);
      hipLaunchKernelGGL(k_per_match, dim3(nblk(G0)), dim3(kT), 0, st, G0, nCand, gBC_, pb,
                         cand, gGP_, gBary_, hit, errD);
      int perr[2] = {0, 0};
      hipMemcpyAsync(perr, errD, 8, hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess)
        DM_ERR(CCFD_EHIP, "periodic hookup on the device failed");
      if (perr[0] == 3) {
        double gp[2] = {0, 0}, gb[2] = {0, 0};
        hipMemcpy(gp, gGP_ + 2 * (size_t)perr[1], 16, hipMemcpyDeviceToHost);
        hipMemcpy(gb, gBary_ + 2 * (size_t)perr[1], 16, hipMemcpyDeviceToHost);
        DM_ERR(CCFD_EINVAL, "connectPeriodicBC: no connection for side GP %g %g",
               gp[0] + gb[0], gp[1] + gb[1]);
      }
      // keep = 1 everywhere, then the matched pairs leave
      hipLaunchKernelGGL(k_fill_one, dim3(nblk(G0)), dim3(kT), 0, st, G0, keep);
      hipLaunchKernelGGL(k_per_link, dim3(nblk(G0)), dim3(kT), 0, st, G0, nTri, hit, gSlot_,
                         nbr_, partner_, owner_, keep);
      hipcub::DeviceScan::ExclusiveSum(tmp4, tb4, keep, pos, G0, st);
      int nk = 0, lastK = 0;
      hipMemcpyAsync(&nk, pos + G0 - 1, 4, hipMemcpyDeviceToHost, st);
      hipMemcpyAsync(&lastK, keep + G0 - 1, 4, hipMemcpyDeviceToHost, st);
      hipLaunchKernelGGL(k_ghost_compact, dim3(nblk(G0)), dim3(kT), 0, st, G0, keep, pos,
                         gSlot_, gBC_, gBary_, gGP_, nbr_, oSlot, oBC, oBary, oGP);
      if (hipStreamSynchronize(st) != hipSuccess)
        DM_ERR(CCFD_EHIP, "ghost compaction on the device failed");
      G = nk + lastK;
      gSlot_ = oSlot;
      gBC_ = oBC;
      gBary_ = oBary;
      gGP_ = oGP;
    }
  }
  hipLaunchKernelGGL(k_slot_conn, dim3(nblk(nSlots)), dim3(kT), 0, st, nSlots, nbr_,
                     partner_, owner_, n_, len_, gp_, gGP_, gpc_, fn_, flen_);
  hipLaunchKernelGGL(k_recon_w, dim3(nblk(N)), dim3(kT), 0, st, N, nTri, gp_, gpc_, w_);

  // element order
  hipLaunchKernelGGL(k_box, dim3(nbox), dim3(kT), 0, st, N, 2, bary_, box + 4);
  hipLaunchKernelGGL(k_box, dim3(1), dim3(kT), 0, st, (int)nbox, 4, box + 4, box);
  hipLaunchKernelGGL(k_sfc_key, dim3(nblk(N)), dim3(kT), 0, st, N, bary_, box, sk, idx);
  tbs = tb3;
  hipcub::DeviceRadixSort::SortPairs(tmp, tbs, sk, sk2, idx, perm_, N, 0, 32, st);
  hipLaunchKernelGGL(k_inverse, dim3(nblk(N)), dim3(kT), 0, st, N, perm_, inv_);

  // the host keeps the order and the serial area total (src/mesh.c:328-332)
  perm.resize(N);
  std::vector<double> area(N);
  hipMemcpyAsync(perm.data(), perm_, 4 * (size_t)N, hipMemcpyDeviceToHost, st);
  hipMemcpyAsync(area.data(), area_, 8 * (size_t)N, hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess)
    DM_ERR(CCFD_EHIP, "mesh geometry on the device failed");
  double tq = 0.0;
  for (int e = 0; e < N; ++e) {
    if (tq != 0.0) tq = 1.0 / (1.0 / tq + area[e]);
    else tq = 1.0 / area[e];
  }
  totalArea_q = tq;
  return CCFD_OK;
}

int DevMesh::write(const DevArenaDst& d, char* err) {
  const int Np = ((N + kBlock - 1) / kBlock) * kBlock;
  if (d.S < sides) DM_ERR(CCFD_ENOTSUP, "mesh has %d-sided elements, image holds %d", sides, d.S);
  if (d.ts == 8)
    hipLaunchKernelGGL(k_write_elems<double>, dim3(nblk(Np)), dim3(kT), 0, st_, N, Np, nTri_,
                       d, perm_, inv_, nbr_, partner_, owner_, w_, gp_, fn_, flen_, areaq_,
                       area_, sx_, sy_, bary_);
  else
    hipLaunchKernelGGL(k_write_elems<float>, dim3(nblk(Np)), dim3(kT), 0, st_, N, Np, nTri_,
                       d, perm_, inv_, nbr_, partner_, owner_, w_, gp_, fn_, flen_, areaq_,
                       area_, sx_, sy_, bary_);
  if (G > 0)
    hipLaunchKernelGGL(k_write_ghosts, dim3(nblk(G)), dim3(kT), 0, st_, G, nTri_, d, inv_,
                       gSlot_, gBC_, gBary_);
  if (hipGetLastError() != hipSuccess) DM_ERR(CCFD_EHIP, "image write launch failed");
  return CCFD_OK;
}

int DevMesh::halo_counts(const DevArenaDst& d, std::vector<int>& counts, char* err) {
  const int nbk = (N + kBlock - 1) / kBlock;
  int* c = (int*)alloc(4 * (size_t)nbk);
  if (!c) DM_ERR(CCFD_ENOMEM, "device allocation for the halo counts failed");
  hipLaunchKernelGGL(k_halo, dim3(nbk), dim3(kBlock), 0, st_, N, d, c, nullptr, nullptr);
  counts.resize(nbk);
  hipMemcpyAsync(counts.data(), c, 4 * (size_t)nbk, hipMemcpyDeviceToHost, st_);
  if (hipStreamSynchronize(st_) != hipSuccess)
    DM_ERR(CCFD_EHIP, "halo count on the device failed");
  return CCFD_OK;
}

int DevMesh::halo_write(const DevArenaDst& d, const std::vector<int>& start, int* dst,
                        char* err) {
  const int nbk = (N + kBlock - 1) / kBlock;
  if ((int)start.size() < nbk) DM_ERR(CCFD_EINVAL, "halo starts missing");
  int* s = (int*)alloc(4 * (size_t)nbk);
  if (!s) DM_ERR(CCFD_ENOMEM, "device allocation for the halo starts failed");
  hipMemcpyAsync(s, start.data(), 4 * (size_t)nbk, hipMemcpyHostToDevice, st_);
  hipLaunchKernelGGL(k_halo, dim3(nbk), dim3(kBlock), 0, st_, N, d, nullptr, s, dst);
  if (hipStreamSynchronize(st_) != hipSuccess)
    DM_ERR(CCFD_EHIP, "halo lists on the device failed");
  return CCFD_OK;
}

}  // namespace ccfd
