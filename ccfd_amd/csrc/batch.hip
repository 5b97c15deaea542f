// batch.hip - many independent explicit Euler solves in one arena (see
// ccfd_batch.h). The operator kernels are those of the single-mesh path;
// here they run over a block table that lists the arena blocks of all
// active instances, with per-instance time, time step and residual.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ccfd_batch.h"
#include "ctx.hpp"
#include "range_alloc.hpp"
#include "soa.hpp"
#include "devmesh.hpp"

using namespace ccfd;

namespace {

struct Inst {
  bool used = false;
  bool wide = false;   // block halo larger than the fused tile
  int N = 0, nbk = 0, b0 = 0, g0 = 0, ng = 0, h0 = 0, nh = 0, bc0 = 0, nbc = 0;
  int maxHalo = 0;     // largest halo list of the instance's blocks
  std::vector<int> perm;
  double totalArea_q = 0.0;
};

}  // namespace

struct ccfd_batch {
  int device = 0;
  hipStream_t st = nullptr;
  ccfd_batch_cfg cfg{};
  double rk[6] = {0};
  int nst = 1;
  int S = 3;
  size_t ts = 8;
  int capBlocks = 0, capElems = 0, capGhosts = 0, capHalo = 0, capBC = 0;
  int maxInst = 0, tile = 0;
  // LDS columns the fused kernels use now: kBlock + the largest halo of the
  // admitted instances (rounded up to 16), at most tile. Only the LDS
  // layout depends on it, not any result.
  int tileAct = 0;
  // arena arrays
  int* link = nullptr;
  void *w = nullptr, *gp = nullptr, *fn = nullptr, *flen = nullptr,
       *areaq = nullptr, *venk = nullptr;
  double *area = nullptr, *sx = nullptr, *sy = nullptr, *bary = nullptr;
  void* prim[2] = {nullptr, nullptr};
  double* Ubuf[2] = {nullptr, nullptr};
  int cur = 0, ucur = 0;   // prim[cur] and Ubuf[ucur] hold the current state
  void* ut = nullptr;
  void* side = nullptr;
  double* dtLoc = nullptr;
  int *ghostBC = nullptr, *ghostCell = nullptr, *ghostK = nullptr;
  double* ghostBary = nullptr;
  void *ghostP = nullptr, *ghostS = nullptr;
  int *haloStart = nullptr, *haloEnd = nullptr, *haloCells = nullptr;
  int* blockInst = nullptr;
  BCDev* bcp = nullptr;
  // per-instance device scalars
  double *tInst = nullptr, *dtInst = nullptr, *cflInst = nullptr,
         *pTimeInst = nullptr, *stopInst = nullptr, *areaQInst = nullptr,
         *resInst = nullptr, *res0Inst = nullptr, *stepsInst = nullptr,
         *nfInst = nullptr;
  // launch tables
  int* blockMap = nullptr;
  int* actInst = nullptr;
  int* actLB = nullptr;
  int nLaunch = 0, nAct = 0, ghostHigh = 0;
  double* part = nullptr;   // [4][capBlocks]
  double* stage = nullptr;  // host I/O staging
  size_t stageBytes = 0;
  RangeAlloc blocks, ghosts, halos, bcs;
  std::vector<Inst> inst;
  // graphs, one per starting buffer parity
  bool useGraph = true;
  bool dirty = true;
  hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};
  int variant = CCFD_KV_AUTO;
  int wideHalo = 0;   // instances whose block halo exceeds the LDS tile
  std::vector<void*> allocs;
};

// ---------------------------------------------------------------- kernels

// per launched block: CFL step minimum (calcTimeStep, convective part,
// src/timeDiscretization.c:249-289); local steps per element when lts
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bdt_partial(
    const int* __restrict__ blockMap, const int* __restrict__ blockInst,
    int Np, const T* __restrict__ prim, const double* __restrict__ area,
    const double* __restrict__ sx, const double* __restrict__ sy, double gam,
    const double* __restrict__ cflInst, int ts1d, int lts,
    double* __restrict__ dtLoc, double* __restrict__ part) {
  __shared__ double sc[kBlock / 64], sf[kBlock / 64];
  const int m = blockMap[blockIdx.x];
  const int base = (m >> 9) * kBlock, n = m & 511;
  const int i = base + threadIdx.x;
  double dc = 1e150, nf = 0.0;
  if ((int)threadIdx.x < n) {
    const double cfl = cflInst[blockInst[base / kBlock]];
    double r = double(prim[i]), u = double(prim[Np + i]);
    double v = double(prim[2 * Np + i]), p = double(prim[3 * Np + i]);
    double a = sqrt(gam * p / r);
    if (ts1d) {
      dc = cfl * sy[i] / (fabs(u) + a);
    } else {
      double ssr = (fabs(u) + a) * sx[i] + (fabs(v) + a) * sy[i];
      dc = cfl * area[i] / ssr;
    }
    if (!isfinite(dc)) nf = 1.0;
    if (lts) dtLoc[i] = dc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    dc = fmin(dc, __shfl_down(dc, off, 64));
    nf = fmax(nf, __shfl_down(nf, off, 64));
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sc[w] = dc; sf[w] = nf; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; ++k) { dc = fmin(dc, sc[k]); nf = fmax(nf, sf[k]); }
    dc = fmin(dc, sc[0]);
    nf = fmax(nf, sf[0]);
    part[blockIdx.x] = dc;
    part[gridDim.x + blockIdx.x] = nf;
  }
}

// per active instance: dt = min over its blocks, output / stop clamp
// (src/timeDiscretization.c:316-321)
__global__ void k_bdt_final(int nLaunch, const int* __restrict__ actInst,
                            const int* __restrict__ actLB,
                            const double* __restrict__ part, double* tInst,
                            double* dtInst, const double* pTimeInst,
                            const double* stopInst, double* nfInst) {
  __shared__ double sc[256], sf[256];
  const int k = actInst[blockIdx.x];
  const int l0 = actLB[blockIdx.x], l1 = actLB[blockIdx.x + 1];
  double dc = 1e150, nf = 0.0;
  for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x) {
    dc = fmin(dc, part[l]);
    nf = fmax(nf, part[nLaunch + l]);
  }
  sc[threadIdx.x] = dc;
  sf[threadIdx.x] = nf;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sc[threadIdx.x] = fmin(sc[threadIdx.x], sc[threadIdx.x + s]);
      sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double t = tInst[k], pT = pTimeInst[k], sT = stopInst[k];
    double dt = sc[0];
    if ((t + dt > pT) || (t + dt > sT))
      dt = fmin(pT, sT) - t;
    else if ((t + 1.5 * dt > pT) || (t + 1.5 * dt > sT))
      dt = 0.5 * (fmin(pT, sT) - t);
    dtInst[k] = dt;
    nfInst[k] = fmax(nfInst[k], sf[0]);
  }
}

// per active instance: globalResidual (src/analyze.c:521-539) from the
// last stage's block partials, then t += dt and the step count
__global__ void k_bres_final(int nLaunch, const int* __restrict__ actInst,
                             const int* __restrict__ actLB,
                             const double* __restrict__ part,
                             const double* __restrict__ areaQInst,
                             double* resInst, double* res0Inst,
                             double* stepsInst, double* tInst,
                             const double* dtInst) {
  __shared__ double s4[4][256];
  const int k = actInst[blockIdx.x];
  const int l0 = actLB[blockIdx.x], l1 = actLB[blockIdx.x + 1];
  for (int v = 0; v < 4; ++v) {
    double acc = 0.0;
    for (int l = l0 + threadIdx.x; l < l1; l += blockDim.x)
      acc += part[v * nLaunch + l];
    s4[v][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int v = 0; v < 4; ++v) s4[v][threadIdx.x] += s4[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool first = stepsInst[k] == 0.0;
    for (int v = 0; v < 4; ++v) {
      const double r = sqrt(s4[v][0] * areaQInst[k]);
      resInst[4 * k + v] = r;
      if (first) res0Inst[4 * k + v] = r;
    }
    stepsInst[k] += 1.0;
    tInst[k] += dtInst[k];
  }
}

// primitive state (staging, double [4][n]) -> prim_T and U of arena
// elements [e0, e0 + n)
template <typename T>
__global__ void k_bset_prim(int n, int e0, int Np, const double* __restrict__ pin,
                            T* __restrict__ prim, double* __restrict__ U,
                            Eq<double> q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[4], c[4];
  for (int v = 0; v < 4; ++v) p[v] = pin[v * n + i];
  prim_cons<double>(q, p, c);
  for (int v = 0; v < 4; ++v) {
    prim[v * Np + e0 + i] = T(p[v]);
    U[v * Np + e0 + i] = c[v];
  }
}

__global__ void k_bfill_int(int* __restrict__ p, int n, int value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

// ------------------------------------------------------------ host helpers

static Eq<double> beq(const ccfd_batch* b) {
  const double g = b->cfg.scheme.gam;
  Eq<double> q{};
  q.gam = g;
  q.gam1 = g - 1.0;
  q.gam1q = 1.0 / (g - 1.0);
  q.mu = 0.0;
  q.Pr = b->cfg.scheme.Pr;
  return q;
}

static void* bmalloc(ccfd_batch* b, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
  b->allocs.push_back(p);
  hipMemsetAsync(p, 0, bytes ? bytes : 16, b->st);
  return p;
}

// device bytes per element / ghost / halo entry of the arena
static size_t elem_bytes(int S, size_t ts) {
  return (size_t)S * 4 +                 // link
         (size_t)(6 * S + S) * ts +      // w, gp, fn, flen
         2 * ts +                        // areaq, venk
         40 +                            // area, sx, sy, bary
         8 * ts + 64 + 4 * ts +          // prim x2, U x2, ut
         8 +                             // dtLoc
         (size_t)4 * S * ts +            // side (split operator)
         2;                              // halo lists (~0.5 entries)
}
static size_t ghost_bytes(size_t ts) { return 12 + 16 + 8 * ts; }

static void free_graphs(ccfd_batch* b) {
  for (int k = 0; k < 4; ++k)
    if (b->gexec[k]) {
      hipGraphExecDestroy(b->gexec[k]);
      b->gexec[k] = nullptr;
    }
}

template <typename T>
static void upload_range(ccfd_batch* b, void* dst, const std::vector<double>& src,
                         int comps, int n, int off, int stride) {
  std::vector<T> tmp(src.size());
  for (size_t k = 0; k < src.size(); ++k) tmp[k] = T(src[k]);
  for (int c = 0; c < comps; ++c)
    hipMemcpyAsync((T*)dst + (size_t)c * stride + off, tmp.data() + (size_t)c * n,
                   (size_t)n * sizeof(T), hipMemcpyHostToDevice, b->st);
  hipStreamSynchronize(b->st);
}

static void upload_ints(ccfd_batch* b, int* dst, const std::vector<int>& src,
                        int comps, int n, int off, int stride) {
  for (int c = 0; c < comps; ++c)
    hipMemcpyAsync(dst + (size_t)c * stride + off, src.data() + (size_t)c * n,
                   (size_t)n * 4, hipMemcpyHostToDevice, b->st);
  hipStreamSynchronize(b->st);
}

static void set_dbl(ccfd_batch* b, double* dst, double v) {
  hipMemcpyAsync(dst, &v, 8, hipMemcpyHostToDevice, b->st);
  hipStreamSynchronize(b->st);
}

static int ensure_stage(ccfd_batch* b, size_t bytes) {
  if (b->stageBytes >= bytes) return 0;
  if (b->stage) hipFree(b->stage);
  b->stage = nullptr;
  b->stageBytes = 0;
  if (hipMalloc((void**)&b->stage, bytes) != hipSuccess)
    return fail(CCFD_ENOMEM, "staging allocation of %zu bytes failed", bytes);
  b->stageBytes = bytes;
  return 0;
}

// rebuild the launch tables from the active instances (arena order)
static int rebuild_tables(ccfd_batch* b) {
  std::vector<int> order;
  for (int k = 0; k < (int)b->inst.size(); ++k)
    if (b->inst[k].used) order.push_back(k);
  std::sort(order.begin(), order.end(),
            [&](int x, int y) { return b->inst[x].b0 < b->inst[y].b0; });
  std::vector<int> map, act, lb;
  for (int k : order) {
    const Inst& I = b->inst[k];
    act.push_back(k);
    lb.push_back((int)map.size());
    for (int j = 0; j < I.nbk; ++j)
      map.push_back(((I.b0 + j) << 9) | std::min(kBlock, I.N - j * kBlock));
  }
  lb.push_back((int)map.size());
  int halo = 0;
  for (int k : order)
    if (kBlock + b->inst[k].maxHalo <= b->tile) halo = std::max(halo, b->inst[k].maxHalo);
  b->tileAct = std::min(b->tile, kBlock + (halo + 15) / 16 * 16);
  b->nLaunch = (int)map.size();
  b->nAct = (int)act.size();
  b->ghostHigh = b->ghosts.high();
  if (b->nLaunch) {
    hipMemcpyAsync(b->blockMap, map.data(), map.size() * 4, hipMemcpyHostToDevice, b->st);
    hipMemcpyAsync(b->actInst, act.data(), act.size() * 4, hipMemcpyHostToDevice, b->st);
    hipMemcpyAsync(b->actLB, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, b->st);
  }
  free_graphs(b);
  b->dirty = false;
  return hipStreamSynchronize(b->st) == hipSuccess ? 0
                                                   : fail(CCFD_EHIP, "table upload failed");
}

static bool fused_ok(const ccfd_batch* b) {
  return b->cfg.scheme.spatialOrder == 2;
}

// the fused kernels while every instance's halo fits the LDS tile, the
// split pair otherwise (same results)
static int active_variant(const ccfd_batch* b) {
  if (b->variant == CCFD_KV_SPLIT || !fused_ok(b) || b->wideHalo > 0)
    return CCFD_KV_SPLIT;
  return b->variant == CCFD_KV_AUTO ? CCFD_KV_FUSED_FACE : b->variant;
}

template <typename T>
static OpArgs<T> arena_args(ccfd_batch* b) {
  OpArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.N = b->capElems;
  a.Np = b->capElems;
  a.S = b->S;
  a.G = b->capGhosts;   // ghost array stride; kernels cover ghostHigh
  a.order = b->cfg.scheme.spatialOrder;
  a.ns = 0;
  a.link = b->link;
  a.w = (const T*)b->w;
  a.gp = (const T*)b->gp;
  a.fn = (const T*)b->fn;
  a.flen = (const T*)b->flen;
  a.areaq = (const T*)b->areaq;
  a.venk = (const T*)b->venk;
  a.bary = b->bary;
  a.area = b->area;
  a.ghostBC = b->ghostBC;
  a.ghostBary = b->ghostBary;
  a.ghostCell = b->ghostCell;
  a.ghostK = b->ghostK;
  a.ghostP = (T*)b->ghostP;
  a.ghostS = (T*)b->ghostS;
  a.side = (T*)b->side;
  a.ut = (T*)b->ut;
  Eq<double> qd = beq(b);
  a.eq.gam = T(qd.gam);
  a.eq.gam1 = T(qd.gam1);
  a.eq.gam1q = T(qd.gam1q);
  a.eq.mu = T(0.0);
  a.eq.Pr = T(qd.Pr);
  a.eqD = qd;
  a.eqD.mu = 0.0;
  a.blockMap = b->blockMap;
  a.blockInst = b->blockInst;
  a.tInst = b->tInst;
  a.dtInst = b->dtInst;
  a.dtLoc = b->cfg.localTimeStep ? b->dtLoc : nullptr;
  a.bcp = b->bcp;
  return a;
}

// one step of every active instance, launched on b->st (eager or captured)
template <typename T>
static void launch_step(ccfd_batch* b) {
  const int uOld = b->ucur, uNew = 1 - b->ucur;
  const dim3 blk(kBlock), grid(b->nLaunch);
  hipLaunchKernelGGL(k_bdt_partial<T>, grid, blk, 0, b->st, b->blockMap,
                     b->blockInst, b->capElems, (const T*)b->prim[b->cur],
                     b->area, b->sx, b->sy, b->cfg.scheme.gam, b->cflInst,
                     b->cfg.scheme.timeStep1D, b->cfg.localTimeStep, b->dtLoc,
                     b->part);
  hipLaunchKernelGGL(k_bdt_final, dim3(b->nAct), dim3(256), 0, b->st,
                     b->nLaunch, b->actInst, b->actLB, b->part, b->tInst,
                     b->dtInst, b->pTimeInst, b->stopInst, b->nfInst);
  const int kv = active_variant(b);
  const int gBlocks = (b->ghostHigh + 127) / 128;
  for (int k = 1; k <= b->nst; ++k) {
    OpArgs<T> a = arena_args<T>(b);
    a.prim = (const T*)b->prim[b->cur];
    a.tcoef = b->rk[k - 1];
    a.rk = 1;
    a.rkc = b->rk[k];
    a.U0 = b->Ubuf[uOld];
    a.U = b->Ubuf[uNew];
    a.prim_out = (T*)b->prim[1 - b->cur];
    a.write_ut = (k == b->nst) ? 1 : 0;
    a.write_U = (k == b->nst) ? 1 : 0;
    a.res_part = (k == b->nst) ? b->part : nullptr;
    if (a.order == 2) {
      if (b->ghostHigh > 0)
        hipLaunchKernelGGL(bc_bary_kernel<T>, dim3(gBlocks), dim3(128), 0, b->st, a);
      const int lim = b->cfg.scheme.limiter;
      if (kv != CCFD_KV_SPLIT) {
        if (b->ghostHigh > 0) {
          if (b->S == 3) {
            if (lim == 1) hipLaunchKernelGGL((bc_side_recon_kernel<T, 1, 3>), dim3(gBlocks), dim3(128), 0, b->st, a);
            else hipLaunchKernelGGL((bc_side_recon_kernel<T, 2, 3>), dim3(gBlocks), dim3(128), 0, b->st, a);
          } else {
            if (lim == 1) hipLaunchKernelGGL((bc_side_recon_kernel<T, 1, 4>), dim3(gBlocks), dim3(128), 0, b->st, a);
            else hipLaunchKernelGGL((bc_side_recon_kernel<T, 2, 4>), dim3(gBlocks), dim3(128), 0, b->st, a);
          }
        }
        FusedArgs f{b->haloStart, b->haloEnd, b->haloCells, b->tileAct};
        launch_fused_code<T>(b->cfg.scheme.iFlux, a, f, lim,
                             kv == CCFD_KV_FUSED_FACE ? 1 : 0, b->nLaunch,
                             fused_lds_bytes(sizeof(T), b->S, b->tileAct), b->st);
      } else {
        if (b->S == 3) {
          if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 3>), grid, blk, 0, b->st, a);
          else hipLaunchKernelGGL((recon_kernel<T, 2, 3>), grid, blk, 0, b->st, a);
        } else {
          if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 4>), grid, blk, 0, b->st, a);
          else hipLaunchKernelGGL((recon_kernel<T, 2, 4>), grid, blk, 0, b->st, a);
        }
        if (b->ghostHigh > 0)
          hipLaunchKernelGGL(bc_side_kernel<T>, dim3(gBlocks), dim3(128), 0, b->st, a);
        launch_flux_code<T>(b->cfg.scheme.iFlux, a, b->nLaunch, b->st);
      }
    } else {
      if (b->ghostHigh > 0)
        hipLaunchKernelGGL(bc_side_kernel<T>, dim3(gBlocks), dim3(128), 0, b->st, a);
      launch_flux_code<T>(b->cfg.scheme.iFlux, a, b->nLaunch, b->st);
    }
    b->cur = 1 - b->cur;
  }
  hipLaunchKernelGGL(k_bres_final, dim3(b->nAct), dim3(256), 0, b->st, b->nLaunch,
                     b->actInst, b->actLB, b->part, b->areaQInst, b->resInst,
                     b->res0Inst, b->stepsInst, b->tInst, b->dtInst);
  b->ucur = uNew;
}

static void launch_step_any(ccfd_batch* b) {
  if (b->ts == 8) launch_step<double>(b);
  else launch_step<float>(b);
}

// two steps as one graph: after two steps the state is back in the
// buffers it started in, so the graph can be replayed; one graph per
// starting buffer pair
static int graph_two_steps(ccfd_batch* b) {
  const int slot = b->ucur * 2 + b->cur;
  if (!b->gexec[slot]) {
    const int cur0 = b->cur, u0 = b->ucur;
    hipGraph_t g;
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail(CCFD_EHIP, "graph capture failed to start");
    launch_step_any(b);
    launch_step_any(b);
    if (hipStreamEndCapture(b->st, &g) != hipSuccess)
      return fail(CCFD_EHIP, "graph capture failed");
    b->cur = cur0;
    b->ucur = u0;
    hipGraphExec_t ex;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
      hipGraphDestroy(g);
      return fail(CCFD_EHIP, "graph instantiation failed");
    }
    hipGraphDestroy(g);
    b->gexec[slot] = ex;
  }
  if (hipGraphLaunch(b->gexec[slot], b->st) != hipSuccess)
    return fail(CCFD_EHIP, "graph launch failed");
  return 0;
}

// ------------------------------------------------------------------- ABI

extern "C" int ccfd_batch_create(ccfd_batch** out, int device,
                                 const ccfd_batch_cfg* cfg) {
  if (!out || !cfg) return fail(CCFD_EINVAL, "null argument");
  const ccfd_gpu_params& p = cfg->scheme;
  if (p.navierstokes) return fail(CCFD_ENOTSUP, "batches are Euler only");
  if (p.iFlux < 1 || p.iFlux > 11) return fail(CCFD_EINVAL, "unknown flux %d", p.iFlux);
  if (p.spatialOrder != 1 && p.spatialOrder != 2)
    return fail(CCFD_EINVAL, "spatial order must be 1 or 2");
  if (p.spatialOrder == 2 && p.limiter != 1 && p.limiter != 2)
    return fail(CCFD_EINVAL, "limiter must be 1 or 2");
  if (cfg->maxSides != 3 && cfg->maxSides != 4)
    return fail(CCFD_EINVAL, "maxSides must be 3 or 4");
  double rk[6];
  int rc0 = check_params(&p, rk);
  if (rc0) return rc0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return fail(CCFD_EINVAL, "no HIP device %d", device);
  hipSetDevice(device);
  ccfd_batch* b = new ccfd_batch();
  b->device = device;
  b->cfg = *cfg;
  memcpy(b->rk, rk, sizeof(rk));
  b->nst = (p.nRKstages == 1 && p.timeOrder == 1) ? 1 : p.nRKstages;
  b->S = cfg->maxSides;
  b->ts = p.precision == CCFD_FP64 ? 8 : 4;
  b->maxInst = cfg->maxInstances > 0 ? cfg->maxInstances : 4096;
  int ldsMax = 65536;
  hipDeviceGetAttribute(&ldsMax, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  // largest halo whose tile fits the device's LDS (at most 512 columns)
  int fitHalo = 0;
  while (fitHalo + 16 <= 512 &&
         fused_lds_bytes(b->ts, b->S, kBlock + fitHalo + 16) + 256 <= (size_t)ldsMax)
    fitHalo += 16;
  const int maxHalo = cfg->maxHalo > 0 ? cfg->maxHalo : fitHalo;
  b->tile = kBlock + maxHalo;
  if (p.spatialOrder == 2 &&
      fused_lds_bytes(b->ts, b->S, b->tile) + 256 > (size_t)ldsMax) {
    delete b;
    return fail(CCFD_EINVAL, "maxHalo %d needs %zu bytes of LDS per block (limit %d)",
                maxHalo, fused_lds_bytes(b->ts, b->S, b->tile), ldsMax);
  }
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(CCFD_EHIP, "stream creation failed");
  }
  size_t budget = cfg->hbmBytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    budget = fr / 2;
  }
  // arena sizing: elements, plus ghosts at 1/8 of the elements
  const size_t perElem = elem_bytes(b->S, b->ts) + ghost_bytes(b->ts) / 8;
  long capE = (long)(budget / perElem);
  // the kernels index [4][S][E] arrays (side states) with 32-bit ints
  capE = std::min(capE, (long)kMaxIndexedElems / b->S);
  b->capBlocks = (int)(capE / kBlock);
  if (b->capBlocks < 1) {
    hipStreamDestroy(b->st);
    delete b;
    return fail(CCFD_ENOMEM, "HBM budget too small for one block");
  }
  b->capElems = b->capBlocks * kBlock;
  b->capGhosts = std::max(1024, b->capElems / 8);
  b->capHalo = b->capElems / 2 + 1024;
  b->capBC = b->maxInst * 8;
  b->blocks.init(b->capBlocks);
  b->ghosts.init(b->capGhosts);
  b->halos.init(b->capHalo);
  b->bcs.init(b->capBC);
  b->inst.assign(b->maxInst, Inst());

  const size_t E = b->capElems, Sl = (size_t)b->S * E, ts = b->ts, Gc = b->capGhosts;
  b->link = (int*)bmalloc(b, Sl * 4);
  b->w = bmalloc(b, 2 * Sl * ts);
  b->gp = bmalloc(b, 2 * Sl * ts);
  b->fn = bmalloc(b, 2 * Sl * ts);
  b->flen = bmalloc(b, Sl * ts);
  b->areaq = bmalloc(b, E * ts);
  b->venk = bmalloc(b, E * ts);
  b->area = (double*)bmalloc(b, E * 8);
  b->sx = (double*)bmalloc(b, E * 8);
  b->sy = (double*)bmalloc(b, E * 8);
  b->bary = (double*)bmalloc(b, 2 * E * 8);
  b->prim[0] = bmalloc(b, 4 * E * ts);
  b->prim[1] = bmalloc(b, 4 * E * ts);
  b->Ubuf[0] = (double*)bmalloc(b, 4 * E * 8);
  b->Ubuf[1] = (double*)bmalloc(b, 4 * E * 8);
  b->ut = bmalloc(b, 4 * E * ts);
  b->side = bmalloc(b, 4 * Sl * ts);
  b->dtLoc = (double*)bmalloc(b, E * 8);
  b->ghostBC = (int*)bmalloc(b, Gc * 4);
  b->ghostCell = (int*)bmalloc(b, Gc * 4);
  b->ghostK = (int*)bmalloc(b, Gc * 4);
  b->ghostBary = (double*)bmalloc(b, 2 * Gc * 8);
  b->ghostP = bmalloc(b, 4 * Gc * ts);
  b->ghostS = bmalloc(b, 4 * Gc * ts);
  b->haloStart = (int*)bmalloc(b, (size_t)b->capBlocks * 4);
  b->haloEnd = (int*)bmalloc(b, (size_t)b->capBlocks * 4);
  b->haloCells = (int*)bmalloc(b, (size_t)b->capHalo * 4);
  b->blockInst = (int*)bmalloc(b, (size_t)b->capBlocks * 4);
  b->bcp = (BCDev*)bmalloc(b, (size_t)b->capBC * sizeof(BCDev));
  const size_t mi = b->maxInst;
  b->tInst = (double*)bmalloc(b, mi * 8);
  b->dtInst = (double*)bmalloc(b, mi * 8);
  b->cflInst = (double*)bmalloc(b, mi * 8);
  b->pTimeInst = (double*)bmalloc(b, mi * 8);
  b->stopInst = (double*)bmalloc(b, mi * 8);
  b->areaQInst = (double*)bmalloc(b, mi * 8);
  b->resInst = (double*)bmalloc(b, 4 * mi * 8);
  b->res0Inst = (double*)bmalloc(b, 4 * mi * 8);
  b->stepsInst = (double*)bmalloc(b, mi * 8);
  b->nfInst = (double*)bmalloc(b, mi * 8);
  b->blockMap = (int*)bmalloc(b, (size_t)b->capBlocks * 4);
  b->actInst = (int*)bmalloc(b, mi * 4);
  b->actLB = (int*)bmalloc(b, (mi + 1) * 4);
  b->part = (double*)bmalloc(b, 4 * (size_t)b->capBlocks * 8);
  for (void* q : b->allocs)
    if (!q) {
      ccfd_batch_destroy(b);
      return fail(CCFD_ENOMEM, "arena allocation failed (%zu bytes budget)", budget);
    }
  // free ghost slots are skipped by the ghost kernels (ghostBC < 0)
  hipLaunchKernelGGL(k_bfill_int, dim3((b->capGhosts + 255) / 256), dim3(256), 0,
                     b->st, b->ghostBC, b->capGhosts, -1);
  if (hipStreamSynchronize(b->st) != hipSuccess) {
    ccfd_batch_destroy(b);
    return fail(CCFD_EHIP, "arena initialisation failed");
  }
  *out = b;
  return CCFD_OK;
}

extern "C" int ccfd_batch_destroy(ccfd_batch* b) {
  if (!b) return CCFD_OK;
  hipSetDevice(b->device);
  if (b->st) hipStreamSynchronize(b->st);
  free_graphs(b);
  for (void* p : b->allocs)
    if (p) hipFree(p);
  if (b->stage) hipFree(b->stage);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
  return CCFD_OK;
}

static int mesh_sides(const ccfd_mesh* m) {
  int S = 0;
  for (long e = 0; e < m->nElems; ++e) S = std::max(S, m->elemType[e]);
  return S;
}

extern "C" size_t ccfd_batch_bytes(const ccfd_batch* b, const ccfd_mesh* m) {
  if (!b || !m) return 0;
  const size_t nb = (m->nElems + kBlock - 1) / kBlock;
  return nb * kBlock * elem_bytes(b->S, b->ts) + m->nGhosts * ghost_bytes(b->ts);
}

extern "C" int ccfd_batch_fits(ccfd_batch* b, const ccfd_mesh* m, int* fits) {
  if (!b || !m || !fits) return fail(CCFD_EINVAL, "null argument");
  const int nbk = (int)((m->nElems + kBlock - 1) / kBlock);
  bool slot = false;
  for (auto& I : b->inst)
    if (!I.used) { slot = true; break; }
  *fits = slot && b->blocks.can(nbk) && b->ghosts.can((int)m->nGhosts) &&
          b->halos.can((int)(m->nElems / 2 + 1)) && b->bcs.can(m->nBC);
  return CCFD_OK;
}

// the rest of an admission, once the instance's image is in the arena
// (instance k, perm: image position -> reference element): boundary
// conditions, per-instance scalars, the initial state, the launch tables.
// A failure leaves the slot as ccfd_batch_remove would.
static int admit_tail(ccfd_batch* b, int k, int nBC, const ccfd_bc* bc,
                      const ccfd_batch_case* cs, const double* prim,
                      const std::vector<int>& perm, int* id) {
  const Inst& I = b->inst[k];
  const int N = I.N, np = I.nbk * kBlock, E = b->capElems, e0 = I.b0 * kBlock;
  // boundary conditions
  if (nBC > 0) {
    std::vector<BCDev> bcs(nBC);
    for (int q = 0; q < nBC; ++q) {
      memset(&bcs[q], 0, sizeof(BCDev));
      bcs[q].type = bc[q].type;
      bcs[q].exactFunc = bc[q].exactFunc;
      for (int v = 0; v < 4; ++v) bcs[q].pVar[v] = bc[q].pVar[v];
    }
    hipMemcpyAsync(b->bcp + I.bc0, bcs.data(), bcs.size() * sizeof(BCDev),
                   hipMemcpyHostToDevice, b->st);
  }
  // per-instance scalars
  set_dbl(b, b->tInst + k, 0.0);
  set_dbl(b, b->dtInst + k, 0.0);
  set_dbl(b, b->cflInst + k, cs->cfl);
  set_dbl(b, b->pTimeInst + k, cs->pTime);
  set_dbl(b, b->stopInst + k, cs->stopTime);
  set_dbl(b, b->areaQInst + k, I.totalArea_q);
  set_dbl(b, b->stepsInst + k, 0.0);
  set_dbl(b, b->nfInst + k, 0.0);
  // a failure from here on leaves the slot as ccfd_batch_remove would:
  // ranges released, tables rebuilt without it, the first error reported
  auto undo = [&](int code) {
    char msg[512];
    snprintf(msg, sizeof(msg), "%s", ccfd_gpu_last_error());
    ccfd_batch_remove(b, k);
    return fail(code, "%s", msg);
  };
  // initial state (reference order -> renumbered [4][np] staging)
  int rc = ensure_stage(b, 4 * (size_t)np * 8);
  if (rc) return undo(rc);
  std::vector<double> st(4 * (size_t)np, 1.0);
  for (int i = 0; i < N; ++i)
    for (int v = 0; v < 4; ++v) st[(size_t)v * np + i] = prim[4 * (size_t)perm[i] + v];
  hipMemcpyAsync(b->stage, st.data(), st.size() * 8, hipMemcpyHostToDevice, b->st);
  const Eq<double> q = beq(b);
  if (b->ts == 8)
    hipLaunchKernelGGL(k_bset_prim<double>, dim3((np + 255) / 256), dim3(256), 0, b->st,
                       np, e0, E, b->stage, (double*)b->prim[b->cur], b->Ubuf[b->ucur], q);
  else
    hipLaunchKernelGGL(k_bset_prim<float>, dim3((np + 255) / 256), dim3(256), 0, b->st,
                       np, e0, E, b->stage, (float*)b->prim[b->cur], b->Ubuf[b->ucur], q);
  if (hipStreamSynchronize(b->st) != hipSuccess) {
    fail(CCFD_EHIP, "instance upload failed");
    return undo(CCFD_EHIP);
  }
  *id = k;
  return rebuild_tables(b);
}

extern "C" int ccfd_batch_add(ccfd_batch* b, const ccfd_mesh* m,
                              const ccfd_batch_case* cs, const double* prim,
                              int* id) {
  if (!b || !m || !cs || !prim || !id) return fail(CCFD_EINVAL, "null argument");
  if (m->nElems <= 0) return fail(CCFD_EINVAL, "empty mesh");
  if (mesh_sides(m) > b->S)
    return fail(CCFD_ENOTSUP, "mesh has %d-sided elements, batch holds %d",
                mesh_sides(m), b->S);
  if (m->nBC > kMaxBC) return fail(CCFD_ENOTSUP, "at most %d boundary conditions", kMaxBC);
  for (int k = 0; k < m->nBC; ++k)
    if (m->bc[k].exactFunc > 0 &&
        (m->bc[k].type == 5 || m->bc[k].type == 6))
      return fail(CCFD_ENOTSUP, "exact-function boundaries are not batched");
  hipSetDevice(b->device);
  int k = -1;
  for (int q = 0; q < b->maxInst; ++q)
    if (!b->inst[q].used) { k = q; break; }
  if (k < 0) return fail(CCFD_ENOMEM, "all %d instance slots in use", b->maxInst);
  const int N = (int)m->nElems, nbk = (N + kBlock - 1) / kBlock;
  const int G = (int)m->nGhosts;
  // place: blocks, ghosts, BC entries; the halo needs the image first
  const int b0 = b->blocks.alloc(nbk);
  if (b0 < 0) return fail(CCFD_ENOMEM, "no room for %d blocks", nbk);
  const int g0 = b->ghosts.alloc(G);
  if (g0 < 0) {
    b->blocks.release(b0, nbk);
    return fail(CCFD_ENOMEM, "no room for %d ghosts", G);
  }
  const int bc0 = b->bcs.alloc(m->nBC);
  if (bc0 < 0) {
    b->blocks.release(b0, nbk);
    b->ghosts.release(g0, G);
    return fail(CCFD_ENOMEM, "no room for boundary conditions");
  }
  MeshSoA h;
  build_soa(m, b->S, b0 * kBlock, g0, bc0, h);
  const int h0 = b->halos.alloc((int)h.haloCells.size());
  if (h0 < 0) {
    b->blocks.release(b0, nbk);
    b->ghosts.release(g0, G);
    b->bcs.release(bc0, m->nBC);
    return fail(CCFD_ENOMEM, "no room for the halo lists");
  }
  Inst& I = b->inst[k];
  I.used = true;
  I.wide = kBlock + h.maxHalo > b->tile;
  I.maxHalo = h.maxHalo;
  if (I.wide) {
    ++b->wideHalo;
    free_graphs(b);
  }
  I.N = N;
  I.nbk = nbk;
  I.b0 = b0;
  I.g0 = g0;
  I.ng = G;
  I.h0 = h0;
  I.nh = (int)h.haloCells.size();
  I.bc0 = bc0;
  I.nbc = m->nBC;
  I.perm = h.perm;
  I.totalArea_q = m->totalArea_q;

  const int E = b->capElems, e0 = b0 * kBlock, np = h.Np;
  // element and slot arrays
  upload_ints(b, b->link, h.link, b->S, np, e0, E);
  std::vector<double> venk(np, 0.0);
  for (int i = 0; i < N; ++i) {
    const double kk = b->cfg.scheme.venk_k, a = h.area[i];
    venk[i] = (kk * sqrt(a)) * (kk * sqrt(a)) * (kk * sqrt(a));
  }
  if (b->ts == 8) {
    upload_range<double>(b, b->w, h.w, 2 * b->S, np, e0, E);
    upload_range<double>(b, b->gp, h.gp, 2 * b->S, np, e0, E);
    upload_range<double>(b, b->fn, h.fn, 2 * b->S, np, e0, E);
    upload_range<double>(b, b->flen, h.flen, b->S, np, e0, E);
    upload_range<double>(b, b->areaq, h.areaq, 1, np, e0, E);
    upload_range<double>(b, b->venk, venk, 1, np, e0, E);
  } else {
    upload_range<float>(b, b->w, h.w, 2 * b->S, np, e0, E);
    upload_range<float>(b, b->gp, h.gp, 2 * b->S, np, e0, E);
    upload_range<float>(b, b->fn, h.fn, 2 * b->S, np, e0, E);
    upload_range<float>(b, b->flen, h.flen, b->S, np, e0, E);
    upload_range<float>(b, b->areaq, h.areaq, 1, np, e0, E);
    upload_range<float>(b, b->venk, venk, 1, np, e0, E);
  }
  upload_range<double>(b, b->area, h.area, 1, np, e0, E);
  upload_range<double>(b, b->sx, h.sx, 1, np, e0, E);
  upload_range<double>(b, b->sy, h.sy, 1, np, e0, E);
  upload_range<double>(b, b->bary, h.bary, 2, np, e0, E);
  // ghosts
  if (G > 0) {
    upload_ints(b, b->ghostBC, h.ghostBC, 1, G, g0, b->capGhosts);
    upload_ints(b, b->ghostCell, h.ghostCell, 1, G, g0, b->capGhosts);
    upload_ints(b, b->ghostK, h.ghostK, 1, G, g0, b->capGhosts);
    upload_range<double>(b, b->ghostBary, h.ghostBary, 2, G, g0, b->capGhosts);
  }
  // halo lists: block-relative starts shifted into the pool
  std::vector<int> hs(nbk), he(nbk), bi(nbk, k);
  for (int j = 0; j < nbk; ++j) {
    hs[j] = h.haloStart[j] + h0;
    he[j] = h.haloStart[j + 1] + h0;
  }
  upload_ints(b, b->haloStart, hs, 1, nbk, b0, 0);
  upload_ints(b, b->haloEnd, he, 1, nbk, b0, 0);
  upload_ints(b, b->blockInst, bi, 1, nbk, b0, 0);
  if (!h.haloCells.empty())
    upload_ints(b, b->haloCells, h.haloCells, 1, (int)h.haloCells.size(), h0, 0);
  return admit_tail(b, k, m->nBC, m->bc, cs, prim, h.perm, id);
}

static int raw_checks(ccfd_batch* b, const ccfd_raw_mesh* m) {
  if (m->nTri < 0 || m->nQuad < 0 || m->nTri + m->nQuad <= 0)
    return fail(CCFD_EINVAL, "empty mesh");
  if (m->nBC < 0 || m->nBC > kMaxBC || (m->nBC && !m->bc))
    return fail(CCFD_ENOTSUP, "at most %d boundary conditions", kMaxBC);
  if ((m->nQuad > 0 ? 4 : 3) > b->S)
    return fail(CCFD_ENOTSUP, "mesh has 4-sided elements, batch holds %d", b->S);
  for (int k = 0; k < m->nBC; ++k) {
    if (m->bc[k].exactFunc > 0 && (m->bc[k].type == 5 || m->bc[k].type == 6))
      return fail(CCFD_ENOTSUP, "exact-function boundaries are not batched");
  }
  return CCFD_OK;
}

extern "C" int ccfd_batch_fits_raw(ccfd_batch* b, const ccfd_raw_mesh* m, int* fits) {
  if (!b || !m || !fits) return fail(CCFD_EINVAL, "null argument");
  const long N = m->nTri + m->nQuad;
  const int nbk = (int)((N + kBlock - 1) / kBlock);
  bool slot = false;
  for (auto& I : b->inst)
    if (!I.used) { slot = true; break; }
  *fits = slot && N > 0 && b->blocks.can(nbk) && b->ghosts.can((int)m->nBCedge) &&
          b->halos.can((int)(N / 2 + 1)) && b->bcs.can(m->nBC);
  return CCFD_OK;
}

extern "C" int ccfd_batch_add_raw(ccfd_batch* b, const ccfd_raw_mesh* m,
                                  const ccfd_batch_case* cs, const double* prim,
                                  int* id) {
  if (!b || !m || !cs || !prim || !id) return fail(CCFD_EINVAL, "null argument");
  int rc = raw_checks(b, m);
  if (rc) return rc;
  hipSetDevice(b->device);
  int k = -1;
  for (int q = 0; q < b->maxInst; ++q)
    if (!b->inst[q].used) { k = q; break; }
  if (k < 0) return fail(CCFD_ENOMEM, "all %d instance slots in use", b->maxInst);
  // the geometry first: it checks the mesh and counts the ghosts
  DevMesh dm;
  char err[256] = {0};
  hipStreamSynchronize(b->st);
  rc = dm.build(*m, b->st, err);
  if (rc) return fail(rc, "%s", err);
  const int N = dm.N, nbk = (N + kBlock - 1) / kBlock, G = dm.G;
  const int b0 = b->blocks.alloc(nbk);
  if (b0 < 0) return fail(CCFD_ENOMEM, "no room for %d blocks", nbk);
  const int g0 = b->ghosts.alloc(G);
  if (g0 < 0) {
    b->blocks.release(b0, nbk);
    return fail(CCFD_ENOMEM, "no room for %d ghosts", G);
  }
  const int bc0 = b->bcs.alloc(m->nBC);
  if (bc0 < 0) {
    b->blocks.release(b0, nbk);
    b->ghosts.release(g0, G);
    return fail(CCFD_ENOMEM, "no room for boundary conditions");
  }
  DevArenaDst d;
  d.S = b->S;
  d.E = b->capElems;
  d.e0 = b0 * kBlock;
  d.capG = b->capGhosts;
  d.g0 = g0;
  d.bcBase = bc0;
  d.ts = (int)b->ts;
  d.venkK = b->cfg.scheme.venk_k;
  d.link = b->link;
  d.w = b->w;
  d.gp = b->gp;
  d.fn = b->fn;
  d.flen = b->flen;
  d.areaq = b->areaq;
  d.venk = b->venk;
  d.area = b->area;
  d.sx = b->sx;
  d.sy = b->sy;
  d.bary = b->bary;
  d.ghostBC = b->ghostBC;
  d.ghostCell = b->ghostCell;
  d.ghostK = b->ghostK;
  d.ghostBary = b->ghostBary;
  std::vector<int> counts;
  auto release = [&](int code) {
    b->blocks.release(b0, nbk);
    b->ghosts.release(g0, G);
    b->bcs.release(bc0, m->nBC);
    // the image written so far is unreferenced: no launch table lists it
    // and the ghost entries are marked unused as ccfd_batch_remove does
    if (G > 0)
      hipLaunchKernelGGL(k_bfill_int, dim3((G + 255) / 256), dim3(256), 0, b->st,
                         b->ghostBC + g0, G, -1);
    hipStreamSynchronize(b->st);
    return fail(code, "%s", err);
  };
  rc = dm.write(d, err);
  if (!rc) rc = dm.halo_counts(d, counts, err);
  if (rc) return release(rc);
  std::vector<int> start(nbk + 1, 0);
  int maxHalo = 0;
  for (int j = 0; j < nbk; ++j) {
    start[j + 1] = start[j] + counts[j];
    maxHalo = std::max(maxHalo, counts[j]);
  }
  const int nh = start[nbk];
  const int h0 = b->halos.alloc(nh);
  if (h0 < 0) {
    snprintf(err, sizeof(err), "no room for the halo lists");
    return release(CCFD_ENOMEM);
  }
  rc = dm.halo_write(d, start, b->haloCells + h0, err);
  if (rc) {
    b->halos.release(h0, nh);
    return release(rc);
  }
  Inst& I = b->inst[k];
  I.used = true;
  I.wide = kBlock + maxHalo > b->tile;
  I.maxHalo = maxHalo;
  if (I.wide) {
    ++b->wideHalo;
    free_graphs(b);
  }
  I.N = N;
  I.nbk = nbk;
  I.b0 = b0;
  I.g0 = g0;
  I.ng = G;
  I.h0 = h0;
  I.nh = nh;
  I.bc0 = bc0;
  I.nbc = m->nBC;
  I.perm = dm.perm;
  I.totalArea_q = dm.totalArea_q;
  std::vector<int> hs(nbk), he(nbk), bi(nbk, k);
  for (int j = 0; j < nbk; ++j) {
    hs[j] = start[j] + h0;
    he[j] = start[j + 1] + h0;
  }
  upload_ints(b, b->haloStart, hs, 1, nbk, b0, 0);
  upload_ints(b, b->haloEnd, he, 1, nbk, b0, 0);
  upload_ints(b, b->blockInst, bi, 1, nbk, b0, 0);
  return admit_tail(b, k, m->nBC, m->bc, cs, prim, dm.perm, id);
}

extern "C" int ccfd_batch_remove(ccfd_batch* b, int id) {
  if (!b || id < 0 || id >= b->maxInst || !b->inst[id].used)
    return fail(CCFD_EINVAL, "no instance %d", id);
  hipSetDevice(b->device);
  hipStreamSynchronize(b->st);
  Inst& I = b->inst[id];
  if (I.ng > 0)
    hipLaunchKernelGGL(k_bfill_int, dim3((I.ng + 255) / 256), dim3(256), 0, b->st,
                       b->ghostBC + I.g0, I.ng, -1);
  b->blocks.release(I.b0, I.nbk);
  b->ghosts.release(I.g0, I.ng);
  b->halos.release(I.h0, I.nh);
  b->bcs.release(I.bc0, I.nbc);
  if (I.wide) {
    --b->wideHalo;
    free_graphs(b);
  }
  I = Inst();
  return rebuild_tables(b);
}

extern "C" int ccfd_batch_step(ccfd_batch* b, int nSteps) {
  if (!b || nSteps < 0) return fail(CCFD_EINVAL, "bad arguments");
  hipSetDevice(b->device);
  if (b->dirty) {
    int rc = rebuild_tables(b);
    if (rc) return rc;
  }
  if (b->nAct == 0 || nSteps == 0) return CCFD_OK;
  int done = 0;
  if (b->useGraph) {
    while (nSteps - done >= 2) {
      int rc = graph_two_steps(b);
      if (rc) return rc;
      done += 2;
    }
  }
  for (; done < nSteps; ++done) launch_step_any(b);
  if (hipGetLastError() != hipSuccess)
    return fail(CCFD_EHIP, "step launch failed");
  return CCFD_OK;
}

extern "C" int ccfd_batch_sync(ccfd_batch* b) {
  if (!b) return fail(CCFD_EINVAL, "null batch");
  hipSetDevice(b->device);
  if (hipStreamSynchronize(b->st) != hipSuccess)
    return fail(CCFD_EHIP, "batch failed: %s", hipGetErrorString(hipGetLastError()));
  return CCFD_OK;
}

static int read_info(ccfd_batch* b, int k, ccfd_batch_info* o) {
  double v[14];
  hipMemcpyAsync(v + 0, b->tInst + k, 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(v + 1, b->dtInst + k, 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(v + 2, b->stepsInst + k, 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(v + 3, b->nfInst + k, 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(v + 4, b->resInst + 4 * k, 32, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(v + 8, b->res0Inst + 4 * k, 32, hipMemcpyDeviceToHost, b->st);
  if (hipStreamSynchronize(b->st) != hipSuccess)
    return fail(CCFD_EHIP, "info read failed");
  o->nElems = b->inst[k].N;
  o->t = v[0];
  o->dt = v[1];
  o->steps = (long)v[2];
  o->nonfinite = v[3] != 0.0;
  for (int q = 0; q < 4; ++q) {
    o->res[q] = v[4 + q];
    o->res0[q] = v[8 + q];
  }
  return CCFD_OK;
}

extern "C" int ccfd_batch_info_get(ccfd_batch* b, int id, ccfd_batch_info* info) {
  if (!b || !info || id < 0 || id >= b->maxInst || !b->inst[id].used)
    return fail(CCFD_EINVAL, "no instance %d", id);
  hipSetDevice(b->device);
  return read_info(b, id, info);
}

extern "C" int ccfd_batch_poll(ccfd_batch* b, int maxIds, int* ids,
                               ccfd_batch_info* info, int* n) {
  if (!b || !n || (maxIds > 0 && (!ids || !info))) return fail(CCFD_EINVAL, "null argument");
  hipSetDevice(b->device);
  const size_t mi = b->maxInst;
  std::vector<double> t(mi), dt(mi), stp(mi), nf(mi), r(4 * mi), r0(4 * mi);
  hipMemcpyAsync(t.data(), b->tInst, mi * 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(dt.data(), b->dtInst, mi * 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(stp.data(), b->stepsInst, mi * 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(nf.data(), b->nfInst, mi * 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(r.data(), b->resInst, 4 * mi * 8, hipMemcpyDeviceToHost, b->st);
  hipMemcpyAsync(r0.data(), b->res0Inst, 4 * mi * 8, hipMemcpyDeviceToHost, b->st);
  if (hipStreamSynchronize(b->st) != hipSuccess) return fail(CCFD_EHIP, "poll failed");
  int c = 0;
  for (int k = 0; k < b->maxInst && c < maxIds; ++k) {
    if (!b->inst[k].used) continue;
    ids[c] = k;
    ccfd_batch_info& o = info[c];
    o.nElems = b->inst[k].N;
    o.t = t[k];
    o.dt = dt[k];
    o.steps = (long)stp[k];
    o.nonfinite = nf[k] != 0.0;
    for (int q = 0; q < 4; ++q) {
      o.res[q] = r[4 * k + q];
      o.res0[q] = r0[4 * k + q];
    }
    ++c;
  }
  *n = c;
  return CCFD_OK;
}

template <typename T>
__global__ void k_bget(int n, int e0, int Np, const T* __restrict__ prim,
                       const double* __restrict__ U, double* __restrict__ out,
                       int cons) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int v = 0; v < 4; ++v)
    out[v * n + i] = cons ? U[v * Np + e0 + i] : double(prim[v * Np + e0 + i]);
}

static int get_state(ccfd_batch* b, int id, double* outp, int cons) {
  if (!b || !outp || id < 0 || id >= b->maxInst || !b->inst[id].used)
    return fail(CCFD_EINVAL, "no instance %d", id);
  hipSetDevice(b->device);
  const Inst& I = b->inst[id];
  const int np = I.nbk * kBlock, e0 = I.b0 * kBlock;
  int rc = ensure_stage(b, 4 * (size_t)np * 8);
  if (rc) return rc;
  if (b->ts == 8)
    hipLaunchKernelGGL(k_bget<double>, dim3((np + 255) / 256), dim3(256), 0, b->st, np,
                       e0, b->capElems, (const double*)b->prim[b->cur],
                       b->Ubuf[b->ucur], b->stage, cons);
  else
    hipLaunchKernelGGL(k_bget<float>, dim3((np + 255) / 256), dim3(256), 0, b->st, np,
                       e0, b->capElems, (const float*)b->prim[b->cur],
                       b->Ubuf[b->ucur], b->stage, cons);
  std::vector<double> h(4 * (size_t)np);
  hipMemcpyAsync(h.data(), b->stage, h.size() * 8, hipMemcpyDeviceToHost, b->st);
  if (hipStreamSynchronize(b->st) != hipSuccess) return fail(CCFD_EHIP, "state read failed");
  for (int i = 0; i < I.N; ++i)
    for (int v = 0; v < 4; ++v) outp[4 * (size_t)I.perm[i] + v] = h[(size_t)v * np + i];
  return CCFD_OK;
}

extern "C" int ccfd_batch_get_prim(ccfd_batch* b, int id, double* prim) {
  return get_state(b, id, prim, 0);
}

extern "C" int ccfd_batch_get_cons(ccfd_batch* b, int id, double* cons) {
  return get_state(b, id, cons, 1);
}

extern "C" int ccfd_batch_set_variant(ccfd_batch* b, int variant) {
  if (!b) return fail(CCFD_EINVAL, "null batch");
  if (variant < CCFD_KV_AUTO || variant > CCFD_KV_FUSED_FACE)
    return fail(CCFD_EINVAL, "unknown kernel variant %d", variant);
  if ((variant == CCFD_KV_FUSED || variant == CCFD_KV_FUSED_FACE) && !fused_ok(b))
    return fail(CCFD_ENOTSUP, "the fused operator is 2nd order only");
  if (variant != b->variant) free_graphs(b);
  b->variant = variant;
  return CCFD_OK;
}

extern "C" int ccfd_batch_get_variant(ccfd_batch* b, int* variant) {
  if (!b || !variant) return fail(CCFD_EINVAL, "null argument");
  *variant = active_variant(b);
  return CCFD_OK;
}

extern "C" int ccfd_batch_set_graph(ccfd_batch* b, int useGraph) {
  if (!b) return fail(CCFD_EINVAL, "null batch");
  b->useGraph = useGraph != 0;
  return CCFD_OK;
}

extern "C" int ccfd_batch_usage(ccfd_batch* b, long* used, long* cap, int* instances) {
  if (!b) return fail(CCFD_EINVAL, "null batch");
  if (used) *used = (long)b->blocks.used * kBlock;
  if (cap) *cap = b->capElems;
  if (instances) *instances = b->nAct;
  return CCFD_OK;
}

extern "C" int ccfd_batch_autotune(ccfd_batch* b, int nSteps, int* variant,
                                   double* msPerStep) {
  if (!b || nSteps <= 0) return fail(CCFD_EINVAL, "bad arguments");
  hipSetDevice(b->device);
  if (b->dirty) {
    int rc = rebuild_tables(b);
    if (rc) return rc;
  }
  if (msPerStep)
    for (int k = 0; k < 4; ++k) msPerStep[k] = 0.0;
  if (b->nAct == 0) {
    if (variant) *variant = active_variant(b);
    return CCFD_OK;
  }
  // snapshot of everything a step changes
  const size_t E = b->capElems, mi = b->maxInst;
  const size_t pb = 4 * E * b->ts;
  std::vector<void*> snap;
  struct Buf { void* p; size_t n; };
  std::vector<Buf> bufs = {{b->prim[b->cur], pb}, {b->Ubuf[b->ucur], 4 * E * 8},
                           {b->tInst, mi * 8},     {b->dtInst, mi * 8},
                           {b->resInst, 4 * mi * 8}, {b->res0Inst, 4 * mi * 8},
                           {b->stepsInst, mi * 8}, {b->nfInst, mi * 8}};
  for (auto& q : bufs) {
    void* s = nullptr;
    if (hipMalloc(&s, q.n) != hipSuccess) {
      for (void* x : snap) hipFree(x);
      return fail(CCFD_ENOMEM, "autotune snapshot allocation failed");
    }
    hipMemcpyAsync(s, q.p, q.n, hipMemcpyDeviceToDevice, b->st);
    snap.push_back(s);
  }
  const int cur0 = b->cur, u0 = b->ucur, v0 = b->variant;
  const int cand[3] = {CCFD_KV_SPLIT, CCFD_KV_FUSED, CCFD_KV_FUSED_FACE};
  int best = active_variant(b);
  double bestMs = 1e300;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  for (int c : cand) {
    if (c != CCFD_KV_SPLIT && (!fused_ok(b) || b->wideHalo > 0)) continue;
    b->variant = c;
    free_graphs(b);
    ccfd_batch_step(b, 2);   // warm-up and graph capture
    hipEventRecord(e0, b->st);
    ccfd_batch_step(b, nSteps);
    hipEventRecord(e1, b->st);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / nSteps;
    if (msPerStep) msPerStep[c] = per;
    if (per < bestMs) { bestMs = per; best = c; }
    // restore the state for the next candidate
    b->cur = cur0;
    b->ucur = u0;
    for (size_t q = 0; q < bufs.size(); ++q) {
      void* dst = q == 0 ? b->prim[b->cur] : q == 1 ? (void*)b->Ubuf[b->ucur] : bufs[q].p;
      hipMemcpyAsync(dst, snap[q], bufs[q].n, hipMemcpyDeviceToDevice, b->st);
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipStreamSynchronize(b->st);
  for (void* x : snap) hipFree(x);
  (void)v0;
  b->variant = best;
  free_graphs(b);
  if (variant) *variant = best;
  return hipGetLastError() == hipSuccess ? CCFD_OK : fail(CCFD_EHIP, "autotune failed");
}
