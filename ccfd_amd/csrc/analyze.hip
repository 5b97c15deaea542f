// analyze.hip - the reference's analysis consumers and source term on the
// device (src/analyze.c, src/source.c):
//   * calcSource: the source term's quadrature per element, evaluated by
//     every operator call at its stage time when a source function is set;
//   * calcCoef: lift, drag and pressure coefficients from the interior side
//     states at the wall that the last operator call reconstructed (the
//     reference reads side->pVar, src/analyze.c:314,344);
//   * calcErrors: L1, L2 and Linf norms of the reconstructed state against
//     an exact function at the quadrature points;
//   * record points: the primitive state of a few elements.
// Quadrature points (src/mesh.c:117-165): a triangle's three side
// midpoints (barycentre + GP) with weight area / 3; a quadrilateral's four
// side midpoints with area / 6 and the midpoint of its nodes 0 and 2 with
// area / 3. Elements keep the reference's side order, so the sums run in
// the reference's order within an element.
#include <math.h>

#include <algorithm>
#include <vector>

#include "ctx.hpp"
#include "operator.hpp"

using namespace ccfd;

namespace {

constexpr int kMaxGP = 5;

// quadrature points and weights of device element i; returns their count
__device__ inline int quad_points(int i, int Np, int S, const int* link,
                                  const double* bary, const double* area,
                                  const double* gp, const double* q5,
                                  double xg[kMaxGP][2], double wg[kMaxGP]) {
  int ns = 0;
  for (int s = 0; s < S; ++s)
    if (link[s * Np + i] != kNoSide) ++ns;
  const double bx = bary[i], by = bary[Np + i], ar = area[i];
  const int nside = ns == 3 ? 3 : 4;
  for (int k = 0; k < nside; ++k) {
    xg[k][0] = bx + gp[k * Np + i];
    xg[k][1] = by + gp[(S + k) * Np + i];
    wg[k] = ns == 3 ? ar / 3.0 : ar / 6.0;
  }
  if (ns == 3) return 3;
  xg[4][0] = q5[i];
  xg[4][1] = q5[Np + i];
  wg[4] = ar / 3.0;
  return 5;
}

// calcSource (src/source.c:66-85) at the time of the coming operator call;
// the point values carry over between points, as the reference's local
// array does
template <typename T>
__global__ __launch_bounds__(256) void k_source(
    int N, int Np, int S, const int* __restrict__ link,
    const double* __restrict__ bary, const double* __restrict__ area,
    const double* __restrict__ gp, const double* __restrict__ q5,
    Eq<double> q, int f, int ns, const double* tdev, double tcoef,
    double time0, T* __restrict__ src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double t = tdev ? tdev[0] + tcoef * tdev[1] : time0;
  double xg[kMaxGP][2], wg[kMaxGP];
  const int n = quad_points(i, Np, S, link, bary, area, gp, q5, xg, wg);
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, pt[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    eval_source<double>(q, f, ns, xg[k], t, pt);
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += pt[v] * wg[k];
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) src[v * Np + i] = T(acc[v]);
}

// calcCoef's sums (src/analyze.c:292-387) in the reference's list order:
// the terms n * p * len and cp in parallel, the sums by one thread in list
// order, 256 entries at a time. prm: pInf, qInfQ, qInfLq, cos, sin of alpha.
__global__ __launch_bounds__(256) void k_wing(int n, int G,
                                              const int* __restrict__ ghost,
                                              const double* __restrict__ nl,
                                              const double* __restrict__ bcIn,
                                              double pInf, double qInfQ,
                                              double qInfLq, double ca,
                                              double sa, double* out) {
  __shared__ double ty[256], tx[256];
  __shared__ double acc[2];
  if (threadIdx.x == 0) acc[0] = acc[1] = 0.0;
  for (int base = 0; base < n; base += 256) {
    const int k = base + threadIdx.x;
    if (k < n) {
      const double p0 = bcIn[3 * G + ghost[k]];
      const double nx = nl[k], ny = nl[n + k], len = nl[2 * n + k];
      ty[threadIdx.x] = ny * p0 * len;
      tx[threadIdx.x] = nx * p0 * len;
      out[2 + k] = (p0 - pInf) * qInfQ;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double cl = acc[0], cd = acc[1];
      const int m = min(256, n - base);
      for (int j = 0; j < m; ++j) {
        cl += ty[j];
        cd += tx[j];
      }
      acc[0] = cl;
      acc[1] = cd;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cl = acc[0] * qInfLq, cd = acc[1] * qInfLq;
    out[0] = cl * ca - cd * sa;
    out[1] = cd * ca + cl * sa;
  }
}

// calcErrors (src/analyze.c:445-514): per element, the error at each
// quadrature point in point order; per block, fixed-order sums (L1, L2)
// and maxima (Linf) into part[12][nb]
template <typename T>
__global__ __launch_bounds__(256) void k_err_partial(
    int N, int Np, int S, int order, const int* __restrict__ link,
    const double* __restrict__ bary, const double* __restrict__ area,
    const double* __restrict__ gp, const double* __restrict__ q5,
    const T* __restrict__ prim, const T* __restrict__ ux,
    const T* __restrict__ uy, Eq<double> q, ExactCfg ex, int f, double time,
    double* __restrict__ part) {
  __shared__ double red[12][4];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double l1[4] = {0.0, 0.0, 0.0, 0.0}, l2[4] = {0.0, 0.0, 0.0, 0.0},
         li[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < N) {
    double xg[kMaxGP][2], wg[kMaxGP];
    const int n = quad_points(i, Np, S, link, bary, area, gp, q5, xg, wg);
    double p[4], gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      p[v] = double(prim[v * Np + i]);
      if (order == 2) {
        gx[v] = double(ux[v * Np + i]);
        gy[v] = double(uy[v * Np + i]);
      }
    }
    for (int k = 0; k < n; ++k) {
      double pe[4];
      exact_func<double>(q, ex, f, xg[k], time, pe);
      const double dx = xg[k][0] - bary[i], dy = xg[k][1] - bary[Np + i];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double err = fabs(pe[v] - (p[v] + dx * gx[v] + dy * gy[v]));
        li[v] = fmax(li[v], err);
        l1[v] += err * wg[k];
        l2[v] += err * err * wg[k];
      }
    }
  }
  // wave reductions (fixed order), then the block's four waves in order
  double vals[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    vals[v] = l1[v];
    vals[4 + v] = l2[v];
    vals[8 + v] = li[v];
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double x = vals[k];
    for (int off = 32; off > 0; off >>= 1) {
      const double y = __shfl_down(x, off, 64);
      x = k < 8 ? x + y : fmax(x, y);
    }
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    double x = red[k][0];
    for (int w = 1; w < 4; ++w) x = k < 8 ? x + red[k][w] : fmax(x, red[k][w]);
    part[(size_t)k * gridDim.x + blockIdx.x] = x;
  }
}

// the block partials in a fixed order: one thread per quantity
__global__ void k_err_final(int nb, const double* __restrict__ part,
                            double* out) {
  const int k = threadIdx.x;
  if (k >= 12) return;
  double x = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double y = part[(size_t)k * nb + b];
    x = k < 8 ? x + y : fmax(x, y);
  }
  out[k] = x;
}

// record points: primitive state of listed device elements, as FP64
template <typename T>
__global__ void k_gather_prim(int n, int Np, const int* __restrict__ idx,
                              const T* __restrict__ prim, double* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) out[4 * k + v] = double(prim[v * Np + idx[k]]);
}

int ensure_q5(ccfd_gpu_ctx* c) {
  if (c->hq5.empty() || c->q5) return CCFD_OK;
  c->q5 = (double*)dmalloc(c, c->hq5.size() * 8);
  if (!c->q5) return fail(CCFD_ENOMEM, "quadrature point allocation failed");
  if (hipMemcpyAsync(c->q5, c->hq5.data(), c->hq5.size() * 8,
                     hipMemcpyHostToDevice, c->st) != hipSuccess)
    return fail(CCFD_EHIP, "quadrature point upload failed");
  return CCFD_OK;
}

int sync(ccfd_gpu_ctx* c, const char* what) {
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "%s failed: %s", what,
                hipGetErrorString(hipGetLastError()));
  return CCFD_OK;
}

}  // namespace

template <typename T>
void ccfd::launch_source(ccfd_gpu_ctx* c, OpArgs<T>& a) {
  const double* q5 = c->q5 ? c->q5 : c->g64.gp;   // q5 unread on triangles
  hipLaunchKernelGGL(k_source<T>, dim3((c->N + 255) / 256), dim3(256), 0,
                     c->st, c->N, c->Np, c->S, c->link, c->bary, c->area,
                     c->g64.gp, q5, eqd(c), c->srcFunc,
                     c->prm.navierstokes ? 1 : 0, a.tdev, a.tcoef, a.time0,
                     (T*)c->src);
  a.src = (const T*)c->src;
}
template void ccfd::launch_source<double>(ccfd_gpu_ctx*, OpArgs<double>&);
template void ccfd::launch_source<float>(ccfd_gpu_ctx*, OpArgs<float>&);

extern "C" int ccfd_gpu_set_source(ccfd_gpu_ctx* c, int sourceFunc) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  if (sourceFunc < 0) return fail(CCFD_EINVAL, "source function %d", sourceFunc);
  hipSetDevice(c->device);
  if (sourceFunc && !c->src) {
    c->src = dmalloc(c, 4 * (size_t)c->Np * 8);
    if (!c->src) return fail(CCFD_ENOMEM, "source term allocation failed");
  }
  int rc = sourceFunc ? ensure_q5(c) : CCFD_OK;
  if (rc) return rc;
  c->srcFunc = sourceFunc;
  return sync(c, "set_source");
}

extern "C" int ccfd_gpu_get_bc_sides(ccfd_gpu_ctx* c, double* pVar) {
  if (!c || !c->hasMesh || !pVar) return fail(CCFD_EINVAL, "no mesh / null output");
  hipSetDevice(c->device);
  const int G = c->G;
  std::vector<double> h(4 * (size_t)std::max(G, 1));
  if (G > 0 && hipMemcpyAsync(h.data(), c->bcIn, 4 * (size_t)G * 8,
                              hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return fail(CCFD_EHIP, "boundary side copy failed");
  int rc = sync(c, "get_bc_sides");
  if (rc) return rc;
  for (int g = 0; g < G; ++g)
    for (int v = 0; v < 4; ++v) pVar[4 * (size_t)g + v] = h[(size_t)v * G + g];
  return CCFD_OK;
}

// initWing (src/analyze.c:97-181): the wall sides of the BC with the given
// code, pressure side (ghost normal n_y > 0) then suction side, each in
// the order the reference's insertion builds: by x of the side midpoint,
// later sides after earlier ones of equal x, sides visited in the order of
// its boundary-side list (built by prepending: last ghost first)
extern "C" int ccfd_gpu_wing_init(ccfd_gpu_ctx* c, const ccfd_mesh* m,
                                  int wallCode, double refLength,
                                  long* nPressure, long* nSuction,
                                  double* xyphi) {
  if (!c || !c->hasMesh || !m) return fail(CCFD_EINVAL, "no mesh");
  if (m->nElems != c->N || m->nGhosts != c->G)
    return fail(CCFD_EINVAL, "mesh differs from the context's");
  if (!(refLength > 0.0)) return fail(CCFD_EINVAL, "reference length %g", refLength);
  hipSetDevice(c->device);
  // the reference keeps the last match of its (prepended) BC list: the
  // first in parameter-file order
  int wing = -1;
  for (int b = m->nBC - 1; b >= 0; --b)
    if (m->bc[b].type * 100 + m->bc[b].id == wallCode) wing = b;
  if (wing < 0) return fail(CCFD_EINVAL, "no boundary condition %d", wallCode);
  auto xmid = [&](long s, int d) {
    long lo = 0, hi = m->nElems - 1;   // element of slot s
    while (lo < hi) {
      const long mid = (lo + hi + 1) / 2;
      if (m->slotStart[mid] <= s) lo = mid; else hi = mid - 1;
    }
    return m->GP[2 * s + d] + m->bary[2 * lo + d];
  };
  std::vector<long> pr, su;
  std::vector<int> pg, sg;
  auto insert = [&](std::vector<long>& L, std::vector<int>& Lg, long s, int g) {
    const double xi = xmid(s, 0);
    size_t k = 0;
    while (k < L.size() && !(xi < xmid(L[k], 0))) ++k;
    L.insert(L.begin() + k, s);
    Lg.insert(Lg.begin() + k, g);
  };
  for (long g = m->nGhosts - 1; g >= 0; --g) {
    if (m->ghostBC[g] != wing) continue;
    const long s = m->ghostSlot[g];
    if (-m->n[2 * s + 1] <= 0.0) insert(su, sg, s, (int)g);
    else insert(pr, pg, s, (int)g);
  }
  const int n = (int)(pr.size() + su.size());
  if (n == 0) return fail(CCFD_EINVAL, "boundary %d has no sides", wallCode);
  std::vector<long> sl(pr);
  sl.insert(sl.end(), su.begin(), su.end());
  std::vector<int> gl(pg);
  gl.insert(gl.end(), sg.begin(), sg.end());
  std::vector<double> nl(3 * (size_t)n);
  for (int k = 0; k < n; ++k) {
    nl[k] = m->n[2 * sl[k]];
    nl[n + k] = m->n[2 * sl[k] + 1];
    nl[2 * n + k] = m->len[sl[k]];
    if (xyphi) {
      const double x = xmid(sl[k], 0), y = xmid(sl[k], 1);
      xyphi[3 * k] = x;
      xyphi[3 * k + 1] = y;
      xyphi[3 * k + 2] = atan2(y, x);
    }
  }
  dfree(c, c->wingG);
  dfree(c, c->wingN);
  dfree(c, c->wingOut);
  c->wingG = (int*)dmalloc(c, (size_t)n * 4);
  c->wingN = (double*)dmalloc(c, 3 * (size_t)n * 8);
  c->wingOut = (double*)dmalloc(c, (2 + (size_t)n) * 8);
  c->nWing = 0;
  if (!c->wingG || !c->wingN || !c->wingOut)
    return fail(CCFD_ENOMEM, "wing allocation failed");
  if (hipMemcpyAsync(c->wingG, gl.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(c->wingN, nl.data(), nl.size() * 8, hipMemcpyHostToDevice,
                     c->st) != hipSuccess)
    return fail(CCFD_EHIP, "wing upload failed");
  int rc = sync(c, "wing_init");
  if (rc) return rc;
  c->nWing = n;
  c->nWingP = (int)pr.size();
  c->refLength = refLength;
  if (nPressure) *nPressure = (long)pr.size();
  if (nSuction) *nSuction = (long)su.size();
  return CCFD_OK;
}

// calcCoef (src/analyze.c:292-387): the freestream terms and the rotation
// by the angle of attack are formed on the host with the C library, as the
// reference forms them
extern "C" int ccfd_gpu_wing_coef(ccfd_gpu_ctx* c, const double ref[4],
                                  double alpha, double coef[2], double* cp) {
  if (!c || !c->hasMesh || !ref || !coef) return fail(CCFD_EINVAL, "null argument");
  if (c->nWing <= 0) return fail(CCFD_EINVAL, "ccfd_gpu_wing_init first");
  hipSetDevice(c->device);
  const double pi = acos(-1.0);
  const double v = ref[1] / cos(alpha * pi / 180.0);
  const double qInfQ = 1.0 / (ref[0] * 0.5 * v * v);
  const double qInfLq = qInfQ / c->refLength;
  const double aL = alpha * pi / 180.0;
  hipLaunchKernelGGL(k_wing, dim3(1), dim3(256), 0, c->st, c->nWing, c->G,
                     c->wingG, c->wingN, c->bcIn, ref[3], qInfQ, qInfLq,
                     cos(aL), sin(aL), c->wingOut);
  std::vector<double> h(2 + (size_t)c->nWing);
  if (hipMemcpyAsync(h.data(), c->wingOut, h.size() * 8, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess)
    return fail(CCFD_EHIP, "wing copy failed");
  int rc = sync(c, "wing_coef");
  if (rc) return rc;
  coef[0] = h[0];
  coef[1] = h[1];
  if (cp) std::copy(h.begin() + 2, h.end(), cp);
  return CCFD_OK;
}

template <typename T>
static void errors_t(ccfd_gpu_ctx* c, int f, double time, const double* q5) {
  const int order = c->prm.spatialOrder;
  if (order == 2) recompute_recon<T>(c, time);
  hipLaunchKernelGGL(k_err_partial<T>, dim3(c->nb), dim3(256), 0, c->st, c->N,
                     c->Np, c->S, order, c->link, c->bary, c->area, c->g64.gp,
                     q5, (const T*)c->prim[c->cur], (const T*)c->ux,
                     (const T*)c->uy, eqd(c), c->ex, f, time, c->errPart);
}

extern "C" int ccfd_gpu_calc_errors(ccfd_gpu_ctx* c, int exactFunc,
                                    double time, double L1[4], double L2[4],
                                    double Linf[4]) {
  if (!c || !c->hasMesh || !L1 || !L2 || !Linf)
    return fail(CCFD_EINVAL, "no mesh / null output");
  if (exactFunc < 1 || exactFunc > 6)
    return fail(CCFD_EINVAL, "exact function %d", exactFunc);
  hipSetDevice(c->device);
  int rc = ensure_q5(c);
  if (rc) return rc;
  if (!c->errPart) {
    c->errPart = (double*)dmalloc(c, (12 * (size_t)c->nb + 16) * 8);
    if (!c->errPart) return fail(CCFD_ENOMEM, "error partial allocation failed");
  }
  const double* q5 = c->q5 ? c->q5 : c->g64.gp;
  if (c->prm.precision == CCFD_FP64) errors_t<double>(c, exactFunc, time, q5);
  else errors_t<float>(c, exactFunc, time, q5);
  double* out = c->errPart + 12 * (size_t)c->nb;
  hipLaunchKernelGGL(k_err_final, dim3(1), dim3(64), 0, c->st, c->nb,
                     c->errPart, out);
  double h[12];
  if (hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return fail(CCFD_EHIP, "error copy failed");
  if ((rc = sync(c, "calc_errors"))) return rc;
  for (int v = 0; v < 4; ++v) {
    L1[v] = h[v] * c->totalArea_q;
    L2[v] = sqrt(h[4 + v] * c->totalArea_q);
    Linf[v] = h[8 + v];
  }
  return CCFD_OK;
}

extern "C" int ccfd_gpu_get_prim_at(ccfd_gpu_ctx* c, int n, const long* elems,
                                    double* pVar) {
  if (!c || !c->hasMesh || n < 0 || (n > 0 && (!elems || !pVar)))
    return fail(CCFD_EINVAL, "no mesh / bad arguments");
  if (n == 0) return CCFD_OK;
  hipSetDevice(c->device);
  std::vector<int> idx(n);
  for (int k = 0; k < n; ++k) {
    if (elems[k] < 0 || elems[k] >= c->N)
      return fail(CCFD_EINVAL, "element %ld out of range", elems[k]);
    idx[k] = c->inv[elems[k]];
  }
  if (n > c->recCap) {
    dfree(c, c->recIdx);
    dfree(c, c->recOut);
    c->recCap = 0;
    c->recIdx = (int*)dmalloc(c, (size_t)n * 4);
    c->recOut = (double*)dmalloc(c, 4 * (size_t)n * 8);
    if (!c->recIdx || !c->recOut) return fail(CCFD_ENOMEM, "record allocation failed");
    c->recCap = n;
  }
  if (hipMemcpyAsync(c->recIdx, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess)
    return fail(CCFD_EHIP, "record upload failed");
  const dim3 g((n + 63) / 64), b(64);
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_gather_prim<double>, g, b, 0, c->st, n, c->Np,
                       c->recIdx, (const double*)c->prim[c->cur], c->recOut);
  else
    hipLaunchKernelGGL(k_gather_prim<float>, g, b, 0, c->st, n, c->Np,
                       c->recIdx, (const float*)c->prim[c->cur], c->recOut);
  if (hipMemcpyAsync(pVar, c->recOut, 4 * (size_t)n * 8, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess)
    return fail(CCFD_EHIP, "record copy failed");
  return sync(c, "get_prim_at");
}
