// ccfd_gpu.hip - context management, mesh renumbering / upload, reductions,
// explicit time stepping and the C ABI of include/ccfd_gpu.h.
//
// The operator kernels live in operator.hpp (instantiated per flux code in
// op_flux.hip); the implicit solver in implicit.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "ccfd_gpu.h"
#include "ctx.hpp"
#include "soa.hpp"

using namespace ccfd;

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

int ccfd::fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* ccfd_gpu_last_error(void) { return g_err; }

// ---------------------------------------------------- flux launch table
#define DECL_FLUX(k)                                                         \
  void ccfd_launch_flux_##k##_d(const OpArgs<double>&, int, hipStream_t);   \
  void ccfd_launch_flux_##k##_f(const OpArgs<float>&, int, hipStream_t);
DECL_FLUX(1) DECL_FLUX(2) DECL_FLUX(3) DECL_FLUX(4) DECL_FLUX(5) DECL_FLUX(6)
DECL_FLUX(7) DECL_FLUX(8) DECL_FLUX(9) DECL_FLUX(10) DECL_FLUX(11)

typedef void (*flux_fn_d)(const OpArgs<double>&, int, hipStream_t);
typedef void (*flux_fn_f)(const OpArgs<float>&, int, hipStream_t);
static const flux_fn_d kFluxD[12] = {
    nullptr, ccfd_launch_flux_1_d, ccfd_launch_flux_2_d, ccfd_launch_flux_3_d,
    ccfd_launch_flux_4_d, ccfd_launch_flux_5_d, ccfd_launch_flux_6_d,
    ccfd_launch_flux_7_d, ccfd_launch_flux_8_d, ccfd_launch_flux_9_d,
    ccfd_launch_flux_10_d, ccfd_launch_flux_11_d};
static const flux_fn_f kFluxF[12] = {
    nullptr, ccfd_launch_flux_1_f, ccfd_launch_flux_2_f, ccfd_launch_flux_3_f,
    ccfd_launch_flux_4_f, ccfd_launch_flux_5_f, ccfd_launch_flux_6_f,
    ccfd_launch_flux_7_f, ccfd_launch_flux_8_f, ccfd_launch_flux_9_f,
    ccfd_launch_flux_10_f, ccfd_launch_flux_11_f};

#define DECL_FUSED(k)                                                        \
  void ccfd_launch_fused_##k##_d(const OpArgs<double>&, const FusedArgs&, int, \
                                 int, int, size_t, hipStream_t, int);         \
  void ccfd_launch_fused_##k##_f(const OpArgs<float>&, const FusedArgs&, int,  \
                                 int, int, size_t, hipStream_t, int);
DECL_FUSED(1) DECL_FUSED(2) DECL_FUSED(3) DECL_FUSED(4) DECL_FUSED(5)
DECL_FUSED(6) DECL_FUSED(7) DECL_FUSED(8) DECL_FUSED(9) DECL_FUSED(10)
DECL_FUSED(11)

typedef void (*fused_fn_d)(const OpArgs<double>&, const FusedArgs&, int, int,
                           int, size_t, hipStream_t, int);
typedef void (*fused_fn_f)(const OpArgs<float>&, const FusedArgs&, int, int,
                           int, size_t, hipStream_t, int);
static const fused_fn_d kFusedD[12] = {
    nullptr, ccfd_launch_fused_1_d, ccfd_launch_fused_2_d, ccfd_launch_fused_3_d,
    ccfd_launch_fused_4_d, ccfd_launch_fused_5_d, ccfd_launch_fused_6_d,
    ccfd_launch_fused_7_d, ccfd_launch_fused_8_d, ccfd_launch_fused_9_d,
    ccfd_launch_fused_10_d, ccfd_launch_fused_11_d};
static const fused_fn_f kFusedF[12] = {
    nullptr, ccfd_launch_fused_1_f, ccfd_launch_fused_2_f, ccfd_launch_fused_3_f,
    ccfd_launch_fused_4_f, ccfd_launch_fused_5_f, ccfd_launch_fused_6_f,
    ccfd_launch_fused_7_f, ccfd_launch_fused_8_f, ccfd_launch_fused_9_f,
    ccfd_launch_fused_10_f, ccfd_launch_fused_11_f};

template <>
void ccfd::launch_fused_code<double>(int code, const OpArgs<double>& a,
                                     const FusedArgs& f, int lim, int face,
                                     int nb, size_t lds, hipStream_t st,
                                     int tb) {
  kFusedD[code](a, f, lim, face, nb, lds, st, tb);
}
template <>
void ccfd::launch_fused_code<float>(int code, const OpArgs<float>& a,
                                    const FusedArgs& f, int lim, int face,
                                    int nb, size_t lds, hipStream_t st,
                                    int tb) {
  kFusedF[code](a, f, lim, face, nb, lds, st, tb);
}

template <>
void ccfd::launch_flux_code<double>(int code, const OpArgs<double>& a, int nb,
                                    hipStream_t st) {
  kFluxD[code](a, nb, st);
}
template <>
void ccfd::launch_flux_code<float>(int code, const OpArgs<float>& a, int nb,
                                   hipStream_t st) {
  kFluxF[code](a, nb, st);
}

#define DECL_JAC1(k)                                                         \
  void ccfd_launch_jac1_##k##_d(const OpArgs<double>&, const double*, double,  \
                                double, const Eq<double>&, double*, double*,   \
                                hipStream_t);                                  \
  void ccfd_launch_jac1_##k##_f(const OpArgs<float>&, const double*, double,   \
                                double, const Eq<double>&, float*, float*,     \
                                hipStream_t);
DECL_JAC1(1) DECL_JAC1(2) DECL_JAC1(3) DECL_JAC1(4) DECL_JAC1(5) DECL_JAC1(6)
DECL_JAC1(7) DECL_JAC1(8) DECL_JAC1(9) DECL_JAC1(10) DECL_JAC1(11)

typedef void (*jac1_fn_d)(const OpArgs<double>&, const double*, double, double,
                          const Eq<double>&, double*, double*, hipStream_t);
typedef void (*jac1_fn_f)(const OpArgs<float>&, const double*, double, double,
                          const Eq<double>&, float*, float*, hipStream_t);
static const jac1_fn_d kJac1D[12] = {
    nullptr, ccfd_launch_jac1_1_d, ccfd_launch_jac1_2_d, ccfd_launch_jac1_3_d,
    ccfd_launch_jac1_4_d, ccfd_launch_jac1_5_d, ccfd_launch_jac1_6_d,
    ccfd_launch_jac1_7_d, ccfd_launch_jac1_8_d, ccfd_launch_jac1_9_d,
    ccfd_launch_jac1_10_d, ccfd_launch_jac1_11_d};
static const jac1_fn_f kJac1F[12] = {
    nullptr, ccfd_launch_jac1_1_f, ccfd_launch_jac1_2_f, ccfd_launch_jac1_3_f,
    ccfd_launch_jac1_4_f, ccfd_launch_jac1_5_f, ccfd_launch_jac1_6_f,
    ccfd_launch_jac1_7_f, ccfd_launch_jac1_8_f, ccfd_launch_jac1_9_f,
    ccfd_launch_jac1_10_f, ccfd_launch_jac1_11_f};

template <>
void ccfd::launch_jac1_code<double, double>(int code, const OpArgs<double>& a,
                                            const double* X, double reps,
                                            double srEps, const Eq<double>& qd,
                                            double* Dg, double* Blk,
                                            hipStream_t st) {
  kJac1D[code](a, X, reps, srEps, qd, Dg, Blk, st);
}
template <>
void ccfd::launch_jac1_code<float, float>(int code, const OpArgs<float>& a,
                                          const double* X, double reps,
                                          double srEps, const Eq<double>& qd,
                                          float* Dg, float* Blk,
                                          hipStream_t st) {
  kJac1F[code](a, X, reps, srEps, qd, Dg, Blk, st);
}

// --------------------------------------------------------------- kernels

// prim (double, staging) -> prim_T and U = primCons(prim)
template <typename T>
__global__ void k_set_prim(int N, int Np, const double* __restrict__ pin,
                           T* __restrict__ prim, double* __restrict__ U,
                           Eq<double> q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double p[4], c[4];
  for (int v = 0; v < 4; ++v) p[v] = pin[v * Np + i];
  prim_cons<double>(q, p, c);
  for (int v = 0; v < 4; ++v) {
    prim[v * Np + i] = T(p[v]);
    U[v * Np + i] = c[v];
  }
}

// U (double) -> prim_T = consPrim(U)
template <typename T>
__global__ void k_cons_to_prim(int N, int Np, const double* __restrict__ U,
                               T* __restrict__ prim, Eq<double> q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double c[4], p[4];
  for (int v = 0; v < 4; ++v) c[v] = U[v * Np + i];
  cons_prim<double>(q, c, p);
  for (int v = 0; v < 4; ++v) prim[v * Np + i] = T(p[v]);
}

template <typename T>
__global__ void k_to_double(long n, const T* __restrict__ src,
                            double* __restrict__ dst) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = double(src[i]);
}

// calcTimeStep per-block minimum (src/timeDiscretization.c:249-307)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dt_partial(
    int N, int Np, const T* __restrict__ prim, const double* __restrict__ area,
    const double* __restrict__ sx, const double* __restrict__ sy, double gam,
    double mu, double gamPrMax, double cfl, double dfl, int ts1d,
    double* __restrict__ part /* [3][nb]: conv, visc, nonfinite */) {
  __shared__ double sc[kBlock / 64], sv[kBlock / 64], sf[kBlock / 64];
  int i = blockIdx.x * kBlock + threadIdx.x;
  double dc = 1e150, dv = 1e150, nf = 0.0;
  if (i < N) {
    double r = double(prim[i]), u = double(prim[Np + i]);
    double v = double(prim[2 * Np + i]), p = double(prim[3 * Np + i]);
    double a = sqrt(gam * p / r);
    if (ts1d) {
      dc = cfl * sy[i] / (fabs(u) + a);
    } else {
      double ssr = (fabs(u) + a) * sx[i] + (fabs(v) + a) * sy[i];
      dc = cfl * area[i] / ssr;
      if (mu > 1e-10) {
        double ssv = gamPrMax * mu * sx[i] * sx[i] + gamPrMax * mu * sy[i] * sy[i];
        dv = dfl * r * r * area[i] * area[i] / (4.0 * ssv);
        if (!isfinite(dv)) nf = 1.0;
      }
    }
    if (!isfinite(dc)) nf = 1.0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    dc = fmin(dc, __shfl_down(dc, off, 64));
    dv = fmin(dv, __shfl_down(dv, off, 64));
    nf = fmax(nf, __shfl_down(nf, off, 64));
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sc[w] = dc; sv[w] = dv; sf[w] = nf; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; ++k) {
      dc = fmin(dc, sc[k]); dv = fmin(dv, sv[k]); nf = fmax(nf, sf[k]);
    }
    dc = fmin(dc, sc[0]); dv = fmin(dv, sv[0]); nf = fmax(nf, sf[0]);
    part[blockIdx.x] = dc;
    part[gridDim.x + blockIdx.x] = dv;
    part[2 * gridDim.x + blockIdx.x] = nf;
  }
}

// final dt: global minimum, output / stop-time clamp; scal = {t, dt, viscDom,
// nonfinite}. t is read from scal[0] when useDevT, else from tHost.
__global__ void k_dt_final(int nb, const double* __restrict__ part,
                           double* scal, int useDevT, double tHost,
                           double pTime, double stopTime) {
  __shared__ double sc[1024], sv[1024], sf[1024];
  double dc = 1e150, dv = 1e150, nf = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    dc = fmin(dc, part[k]);
    dv = fmin(dv, part[nb + k]);
    nf = fmax(nf, part[2 * nb + k]);
  }
  sc[threadIdx.x] = dc; sv[threadIdx.x] = dv; sf[threadIdx.x] = nf;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sc[threadIdx.x] = fmin(sc[threadIdx.x], sc[threadIdx.x + s]);
      sv[threadIdx.x] = fmin(sv[threadIdx.x], sv[threadIdx.x + s]);
      sf[threadIdx.x] = fmax(sf[threadIdx.x], sf[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double t = useDevT ? scal[0] : tHost;
    double dt = fmin(sc[0], sv[0]);
    double vd = sv[0] < sc[0] ? 1.0 : 0.0;
    if ((t + dt > pTime) || (t + dt > stopTime))
      dt = fmin(pTime, stopTime) - t;
    else if ((t + 1.5 * dt > pTime) || (t + 1.5 * dt > stopTime))
      dt = 0.5 * (fmin(pTime, stopTime) - t);
    scal[0] = t;
    scal[1] = dt;
    scal[2] = vd;
    scal[3] = fmax(scal[3], sf[0]);
  }
}

// first stage of the residual reduction for many blocks: block b of
// kResMid sums a contiguous share of each variable's nb partials into
// mid[v * kResMid + b] (fixed order)
constexpr int kResMid = 64;
__global__ __launch_bounds__(256) void k_res_mid(int nb,
                                                 const double* __restrict__ part,
                                                 double* __restrict__ mid) {
  __shared__ double s4[4][256];
  const int per = (nb + kResMid - 1) / kResMid;
  const int k0 = blockIdx.x * per, k1 = min(nb, k0 + per);
  for (int v = 0; v < 4; ++v) {
    double acc = 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x)
      acc += part[v * nb + k];
    s4[v][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int v = 0; v < 4; ++v) s4[v][threadIdx.x] += s4[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) mid[threadIdx.x * kResMid + blockIdx.x] = s4[threadIdx.x][0];
}

// globalResidual final reduction (src/analyze.c:534-538); res -> scal[4..7]
__global__ void k_res_final(int nb, const double* __restrict__ part,
                            double totalArea_q, double* scal, int advanceT) {
  __shared__ double s4[4][256];
  for (int v = 0; v < 4; ++v) {
    double acc = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) acc += part[v * nb + k];
    s4[v][threadIdx.x] = acc;
  }
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int v = 0; v < 4; ++v) s4[v][threadIdx.x] += s4[v][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int v = 0; v < 4; ++v) scal[4 + v] = sqrt(s4[v][0] * totalArea_q);
    if (advanceT) scal[0] += scal[1];
  }
}

__global__ void k_advance_t(double* scal) {
  if (threadIdx.x == 0 && blockIdx.x == 0) scal[0] += scal[1];
}

template <typename T>
__global__ void k_riemann(int code, long n, const double* __restrict__ L,
                          const double* __restrict__ R, double* __restrict__ F,
                          Eq<T> q) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T l[4], r[4], f[4];
  for (int v = 0; v < 4; ++v) {
    l[v] = T(L[4 * i + v]);
    r[v] = T(R[4 * i + v]);
  }
  conv_flux<T>(q, code, l, r, f);
  for (int v = 0; v < 4; ++v) F[4 * i + v] = double(f[v]);
}

// ------------------------------------------------------------ host helpers

// -------------------------------------------------------- context helpers

void* ccfd::dmalloc(ccfd_gpu_ctx* c, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
  c->allocs.push_back(p);
  hipMemsetAsync(p, 0, bytes ? bytes : 16, c->st);
  return p;
}

void ccfd::dfree(ccfd_gpu_ctx* c, void* p) {
  if (!p) return;
  for (size_t k = 0; k < c->allocs.size(); ++k)
    if (c->allocs[k] == p) {
      hipFree(p);
      c->allocs.erase(c->allocs.begin() + k);
      return;
    }
}

static void free_all(ccfd_gpu_ctx* c) {
  for (void* p : c->allocs) hipFree(p);
  c->allocs.clear();
}

template <typename T>
Geom<T>& geom(ccfd_gpu_ctx* c);
template <>
Geom<double>& geom<double>(ccfd_gpu_ctx* c) { return c->g64; }
template <>
Geom<float>& geom<float>(ccfd_gpu_ctx* c) { return c->g32; }

template <typename T>
static int upload_geom(ccfd_gpu_ctx* c, const std::vector<double>& w,
                       const std::vector<double>& gp, const std::vector<double>& fn,
                       const std::vector<double>& flen,
                       const std::vector<double>& bbv,
                       const std::vector<double>& bbd,
                       const std::vector<double>& areaq,
                       const std::vector<double>& venk) {
  Geom<T>& g = geom<T>(c);
  size_t sl = (size_t)c->S * c->Np;
  auto up = [&](const std::vector<double>& src, T** dst) -> int {
    std::vector<T> tmp(src.size());
    for (size_t k = 0; k < src.size(); ++k) tmp[k] = T(src[k]);
    *dst = (T*)dmalloc(c, tmp.size() * sizeof(T));
    if (!*dst) return fail(CCFD_ENOMEM, "device allocation failed");
    if (hipMemcpyAsync(*dst, tmp.data(), tmp.size() * sizeof(T),
                       hipMemcpyHostToDevice, c->st) != hipSuccess)
      return fail(CCFD_EHIP, "upload failed");
    return hipStreamSynchronize(c->st) == hipSuccess ? 0 : CCFD_EHIP;
  };
  int rc;
  if ((rc = up(w, &g.w)) || (rc = up(gp, &g.gp)) || (rc = up(fn, &g.fn)) ||
      (rc = up(flen, &g.flen)) || (rc = up(bbv, &g.bbv)) ||
      (rc = up(bbd, &g.bbd)) || (rc = up(areaq, &g.areaq)) ||
      (rc = up(venk, &g.venk)))
    return rc;
  (void)sl;
  return 0;
}

// ------------------------------------------------------------------ ABI

extern "C" int ccfd_gpu_create(ccfd_gpu_ctx** out, int device) {
  if (!out) return fail(CCFD_EINVAL, "null context pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(CCFD_EHIP, "no HIP device available");
  if (device < 0 || device >= n)
    return fail(CCFD_EINVAL, "device %d out of range (%d devices)", device, n);
  if (hipSetDevice(device) != hipSuccess)
    return fail(CCFD_EHIP, "hipSetDevice(%d) failed", device);
  ccfd_gpu_ctx* c = new ccfd_gpu_ctx();
  c->device = device;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            device) == hipSuccess && lds > 0)
    c->ldsLimit = (size_t)lds;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(CCFD_EHIP, "stream creation failed");
  }
  if (hipHostMalloc((void**)&c->hscal, 64 * sizeof(double)) != hipSuccess) {
    hipStreamDestroy(c->st);
    delete c;
    return fail(CCFD_EHIP, "pinned allocation failed");
  }
  *out = c;
  return CCFD_OK;
}

extern "C" int ccfd_gpu_destroy(ccfd_gpu_ctx* c) {
  if (!c) return CCFD_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->st);
  ccfd::implicit_free(c);
  free_all(c);
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  if (c->hscal) hipHostFree(c->hscal);
  hipStreamDestroy(c->st);
  delete c;
  return CCFD_OK;
}

int ccfd::check_params(const ccfd_gpu_params* p, double rk[6]) {
  if (p->iFlux < 1 || p->iFlux > 11)
    return fail(CCFD_EINVAL, "unknown flux function %d", p->iFlux);
  if (p->spatialOrder != 1 && p->spatialOrder != 2)
    return fail(CCFD_EINVAL, "spatial order must be 1 or 2");
  if (p->spatialOrder == 2 && p->limiter != 1 && p->limiter != 2)
    return fail(CCFD_EINVAL, "limiter must be 1 or 2");
  if (p->precision != CCFD_FP64 && p->precision != CCFD_FP32_FLUX)
    return fail(CCFD_EINVAL, "unknown precision %d", p->precision);
  for (int k = 0; k < 6; ++k) rk[k] = 0.0;
  if (p->nRKstages == 1 && p->timeOrder == 1) {
    rk[1] = 1.0;
  } else {
    static const double t3[3][3] = {{0.1418, 0.4, 1.0},
                                    {0.1918, 0.4929, 1.0},
                                    {0.2884, 0.5010, 1.0}};
    static const double t4[3][4] = {{0.0833, 0.2069, 0.4265, 1.0},
                                    {0.1084, 0.2602, 0.5052, 1.0},
                                    {0.1666, 0.3027, 0.5275, 1.0}};
    static const double t5[3][5] = {{0.0533, 0.1263, 0.2375, 0.4414, 1.0},
                                    {0.0695, 0.1602, 0.2898, 0.5060, 1.0},
                                    {0.1067, 0.1979, 0.3232, 0.5201, 1.0}};
    int o = p->timeOrder;
    if (o < 1 || o > 3)
      return fail(CCFD_EINVAL, "time order must be 1, 2 or 3");
    if (p->nRKstages == 3) for (int k = 0; k < 3; ++k) rk[k + 1] = t3[o - 1][k];
    else if (p->nRKstages == 4) for (int k = 0; k < 4; ++k) rk[k + 1] = t4[o - 1][k];
    else if (p->nRKstages == 5) for (int k = 0; k < 5; ++k) rk[k + 1] = t5[o - 1][k];
    else return fail(CCFD_EINVAL, "wrong number of RK stages %d", p->nRKstages);
  }
  return CCFD_OK;
}

int ccfd::apply_params(ccfd_gpu_ctx* c, const ccfd_gpu_params* p) {
  double rk[6];
  int rc = check_params(p, rk);
  if (rc) return rc;
  c->prm = *p;
  memcpy(c->rk, rk, sizeof(rk));
  c->ex.refState[0][0] = p->refState[0][0];
  for (int k = 0; k < 2; ++k)
    for (int v = 0; v < 4; ++v) c->ex.refState[k][v] = p->refState[k][v];
  c->ex.rp1Dinterface = p->rp1Dinterface;
  return CCFD_OK;
}

extern "C" int ccfd_gpu_set_params(ccfd_gpu_ctx* c, const ccfd_gpu_params* p) {
  if (!c || !p) return fail(CCFD_EINVAL, "null argument");
  if (!c->hasMesh) return fail(CCFD_EINVAL, "no mesh uploaded");
  int prevPrec = c->prm.precision;
  int rc = apply_params(c, p);
  if (rc) return rc;
  if (prevPrec != p->precision) {
    // re-derive the operator state in the new arithmetic from U
    hipSetDevice(c->device);
    return ccfd::refresh_prim(c);
  }
  // venk K may change: rebuild eps^2
  return ccfd::upload_venk(c);
}

int ccfd::upload_venk(ccfd_gpu_ctx* c) {
  std::vector<double> v(c->Np, 0.0);
  for (int i = 0; i < c->N; ++i) {
    double a = c->hArea[i];
    double k = c->prm.venk_k;
    v[i] = (k * sqrt(a)) * (k * sqrt(a)) * (k * sqrt(a));
  }
  std::vector<float> vf(v.begin(), v.end());
  if (hipMemcpyAsync(c->g64.venk, v.data(), v.size() * 8, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(c->g32.venk, vf.data(), vf.size() * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "venk upload failed");
  return CCFD_OK;
}

extern "C" int ccfd_gpu_set_mesh(ccfd_gpu_ctx* c, const ccfd_mesh* m,
                                 const ccfd_gpu_params* p) {
  if (!c || !m || !p) return fail(CCFD_EINVAL, "null argument");
  if (m->nElems <= 0) return fail(CCFD_EINVAL, "empty mesh");
  if (m->nElems > kMaxIndexedElems / kMaxSlots)
    return fail(CCFD_ENOTSUP, "mesh too large (%ld elements; at most %ld)",
                m->nElems, kMaxIndexedElems / kMaxSlots);
  if (m->nBC > kMaxBC)
    return fail(CCFD_ENOTSUP, "at most %d boundary conditions", kMaxBC);
  hipSetDevice(c->device);
  hipStreamSynchronize(c->st);
  ccfd::implicit_free(c);
  free_all(c);
  c->hasMesh = false;
  c->bcIn = nullptr; c->src = nullptr; c->q5 = nullptr;
  c->wingG = nullptr; c->wingN = nullptr; c->wingOut = nullptr;
  c->nWing = c->nWingP = 0;
  c->recIdx = nullptr; c->recOut = nullptr; c->recCap = 0;
  c->errPart = nullptr;
  c->hq5.clear();
  int rc = apply_params(c, p);
  if (rc) return rc;

  const int N = (int)m->nElems;
  int S = 0;
  for (int e = 0; e < N; ++e) S = std::max(S, m->elemType[e]);
  const int Np = ((N + kBlock - 1) / kBlock) * kBlock;
  const int G = (int)m->nGhosts;
  c->N = N; c->Np = Np; c->S = S; c->G = G;
  c->nb = Np / kBlock;
  c->totalArea_q = m->totalArea_q;
  c->nSlotsHost = m->nSlots;

  MeshSoA h;
  build_soa(m, S, 0, 0, 0, h);
  c->perm = h.perm;
  c->inv = h.inv;
  c->slotStart.assign(m->slotStart, m->slotStart + N + 1);
  c->hArea.assign(h.area.begin(), h.area.begin() + N);
  c->tile = kBlock + h.maxHalo;
  if (h.haloCells.empty()) h.haloCells.push_back(0);
  // halo lists of two-block tiles for the wide face-once kernel
  std::vector<int> hs2, hc2;
  int mh2 = 0;
  build_halo(h, 2 * kBlock, 0, hs2, hc2, mh2);
  c->tile2 = 2 * kBlock + mh2;
  c->nt2 = (int)hs2.size() - 1;
  if (hc2.empty()) hc2.push_back(0);
  std::vector<int> ghostBC(std::max(G, 1), 0), ghostCell(std::max(G, 1), 0),
      ghostK(std::max(G, 1), 0);
  std::vector<double> ghostBary(2 * (size_t)std::max(G, 1), 0.0);
  std::copy(h.ghostBC.begin(), h.ghostBC.end(), ghostBC.begin());
  std::copy(h.ghostCell.begin(), h.ghostCell.end(), ghostCell.begin());
  std::copy(h.ghostK.begin(), h.ghostK.end(), ghostK.begin());
  std::copy(h.ghostBary.begin(), h.ghostBary.end(), ghostBary.begin());
  std::vector<double> venk(Np, 0.0);
  const size_t sl = (size_t)S * Np;
  c->hasExact = false;
  for (int b = 0; b < m->nBC; ++b) {
    if (m->bc[b].type == 6) c->hasExact = true;
    c->bct.bc[b].type = m->bc[b].type;
    c->bct.bc[b].exactFunc = m->bc[b].exactFunc;
    for (int v = 0; v < 4; ++v) c->bct.bc[b].pVar[v] = m->bc[b].pVar[v];
  }
  c->ex.xMin = m->xMin; c->ex.xMax = m->xMax;
  c->ex.yMin = m->yMin; c->ex.yMax = m->yMax;
  c->ex.dxRef = sqrt(1.0 / (m->totalArea_q * m->nElems));

  // device allocations
  auto upd = [&](const void* src, size_t bytes, void** dst) -> int {
    *dst = dmalloc(c, bytes);
    if (!*dst) return fail(CCFD_ENOMEM, "device allocation of %zu bytes failed", bytes);
    if (hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->st) != hipSuccess)
      return fail(CCFD_EHIP, "upload failed");
    return hipStreamSynchronize(c->st) == hipSuccess ? 0
                                                     : fail(CCFD_EHIP, "upload sync failed");
  };
  if ((rc = upd(h.link.data(), h.link.size() * 4, (void**)&c->link)) ||
      (rc = upd(ghostBC.data(), ghostBC.size() * 4, (void**)&c->ghostBC)) ||
      (rc = upd(ghostCell.data(), ghostCell.size() * 4, (void**)&c->ghostCell)) ||
      (rc = upd(ghostK.data(), ghostK.size() * 4, (void**)&c->ghostK)) ||
      (rc = upd(ghostBary.data(), ghostBary.size() * 8, (void**)&c->ghostBary)) ||
      (rc = upd(h.bary.data(), h.bary.size() * 8, (void**)&c->bary)) ||
      (rc = upd(h.area.data(), h.area.size() * 8, (void**)&c->area)) ||
      (rc = upd(h.sx.data(), h.sx.size() * 8, (void**)&c->sx)) ||
      (rc = upd(h.sy.data(), h.sy.size() * 8, (void**)&c->sy)) ||
      (rc = upd(h.haloStart.data(), h.haloStart.size() * 4, (void**)&c->haloStart)) ||
      (rc = upd(h.haloCells.data(), h.haloCells.size() * 4, (void**)&c->haloCells)) ||
      (rc = upd(hs2.data(), hs2.size() * 4, (void**)&c->haloStart2)) ||
      (rc = upd(hc2.data(), hc2.size() * 4, (void**)&c->haloCells2)))
    return rc;
  if ((rc = upload_geom<double>(c, h.w, h.gp, h.fn, h.flen, h.bbv, h.bbd, h.areaq, venk)) ||
      (rc = upload_geom<float>(c, h.w, h.gp, h.fn, h.flen, h.bbv, h.bbd, h.areaq, venk)))
    return rc;

  // state buffers (sized for double, used as T)
  size_t e4 = 4 * (size_t)Np;
  c->prim[0] = dmalloc(c, e4 * 8);
  c->prim[1] = dmalloc(c, e4 * 8);
  c->U = (double*)dmalloc(c, e4 * 8);
  c->U0 = (double*)dmalloc(c, e4 * 8);
  c->ut = dmalloc(c, e4 * 8);
  c->ux = dmalloc(c, e4 * 8);
  c->uy = dmalloc(c, e4 * 8);
  c->side = dmalloc(c, 4 * sl * 8);
  c->ghostP = dmalloc(c, 4 * (size_t)std::max(G, 1) * 8);
  c->ghostS = dmalloc(c, 4 * (size_t)std::max(G, 1) * 8);
  c->stage = (double*)dmalloc(c, std::max(e4, 4 * (size_t)S * Np) * 8);
  // per-block partials (4 per block) and the residual's middle stage
  c->part = (double*)dmalloc(c, (4 * (size_t)c->nb + 4 * kResMid + 8) * 8);
  c->scal = (double*)dmalloc(c, 64 * 8);
  c->bcIn = (double*)dmalloc(c, 4 * (size_t)std::max(G, 1) * 8);
  if (c->srcFunc) c->src = dmalloc(c, e4 * 8);
  if (!c->prim[0] || !c->prim[1] || !c->U || !c->U0 || !c->ut || !c->ux ||
      !c->uy || !c->side || !c->ghostP || !c->ghostS || !c->stage || !c->part ||
      !c->scal || !c->bcIn || (c->srcFunc && !c->src))
    return fail(CCFD_ENOMEM, "device state allocation failed");
  // centre points of the quadrilaterals (the fifth quadrature point,
  // src/mesh.c:157-158), for error norms and source terms
  if (m->nQuads > 0) {
    c->hq5.assign(2 * (size_t)Np, 0.0);
    for (int e = 0; e < N; ++e) {
      if (m->elemType[e] != 4) continue;
      const double* a0 = m->xy + 2 * m->elemNode[4 * (size_t)e];
      const double* a2 = m->xy + 2 * m->elemNode[4 * (size_t)e + 2];
      const int i = c->inv[e];
      c->hq5[i] = 0.5 * (a0[0] + a2[0]);
      c->hq5[Np + i] = 0.5 * (a0[1] + a2[1]);
    }
    // the source kernel reads them at every operator call
    if (c->srcFunc) {
      c->q5 = (double*)dmalloc(c, c->hq5.size() * 8);
      if (!c->q5 || hipMemcpyAsync(c->q5, c->hq5.data(), c->hq5.size() * 8,
                                   hipMemcpyHostToDevice, c->st) != hipSuccess)
        return fail(CCFD_ENOMEM, "quadrature point upload failed");
    }
  }
  c->cur = 0;
  c->hasMesh = true;
  if ((rc = upload_venk(c))) return rc;
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "set_mesh sync failed");
  return CCFD_OK;
}

// host AoS (reference order) <-> device SoA (renumbered) through c->stage
static int upload_aos(ccfd_gpu_ctx* c, const double* h) {
  std::vector<double> tmp(4 * (size_t)c->Np, 0.0);
  for (int i = 0; i < c->N; ++i) {
    int e = c->perm[i];
    for (int v = 0; v < 4; ++v) tmp[(size_t)v * c->Np + i] = h[4 * (size_t)e + v];
  }
  if (hipMemcpyAsync(c->stage, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "state upload failed");
  return CCFD_OK;
}

static int download_aos(ccfd_gpu_ctx* c, const double* dsrc, double* h) {
  std::vector<double> tmp(4 * (size_t)c->Np);
  if (hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * 8, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "state download failed: %s",
                hipGetErrorString(hipGetLastError()));
  for (int i = 0; i < c->N; ++i) {
    int e = c->perm[i];
    for (int v = 0; v < 4; ++v) h[4 * (size_t)e + v] = tmp[(size_t)v * c->Np + i];
  }
  return CCFD_OK;
}

// T-typed device array [4][Np] -> double staging
static const double* as_double(ccfd_gpu_ctx* c, const void* src, size_t n) {
  if (c->prm.precision == CCFD_FP64) return (const double*)src;
  hipLaunchKernelGGL(k_to_double<float>, dim3((n + 255) / 256), dim3(256), 0,
                     c->st, (long)n, (const float*)src, c->stage);
  return c->stage;
}

Eq<double> ccfd::eqd(const ccfd_gpu_ctx* c) {
  Eq<double> q;
  q.gam = c->prm.gam;
  q.gam1 = c->prm.gam - 1.0;
  q.gam1q = 1.0 / q.gam1;
  q.mu = c->prm.navierstokes ? c->prm.mu : 0.0;
  q.Pr = c->prm.Pr;
  return q;
}

int ccfd::refresh_prim(ccfd_gpu_ctx* c) {
  dim3 g((c->N + 255) / 256), b(256);
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_cons_to_prim<double>, g, b, 0, c->st, c->N, c->Np,
                       c->U, (double*)c->prim[c->cur], eqd(c));
  else
    hipLaunchKernelGGL(k_cons_to_prim<float>, g, b, 0, c->st, c->N, c->Np,
                       c->U, (float*)c->prim[c->cur], eqd(c));
  return hipStreamSynchronize(c->st) == hipSuccess
             ? CCFD_OK
             : fail(CCFD_EHIP, "consPrim kernel failed");
}

extern "C" int ccfd_gpu_set_prim(ccfd_gpu_ctx* c, const double* prim) {
  if (!c || !prim || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh / null state");
  hipSetDevice(c->device);
  int rc = upload_aos(c, prim);
  if (rc) return rc;
  dim3 g((c->N + 255) / 256), b(256);
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_set_prim<double>, g, b, 0, c->st, c->N, c->Np,
                       c->stage, (double*)c->prim[c->cur], c->U, eqd(c));
  else
    hipLaunchKernelGGL(k_set_prim<float>, g, b, 0, c->st, c->N, c->Np,
                       c->stage, (float*)c->prim[c->cur], c->U, eqd(c));
  return hipStreamSynchronize(c->st) == hipSuccess
             ? CCFD_OK
             : fail(CCFD_EHIP, "set_prim failed");
}

extern "C" int ccfd_gpu_set_cons(ccfd_gpu_ctx* c, const double* cons) {
  if (!c || !cons || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh / null state");
  hipSetDevice(c->device);
  int rc = upload_aos(c, cons);
  if (rc) return rc;
  if (hipMemcpyAsync(c->U, c->stage, 4 * (size_t)c->Np * 8,
                     hipMemcpyDeviceToDevice, c->st) != hipSuccess)
    return fail(CCFD_EHIP, "copy failed");
  return refresh_prim(c);
}

extern "C" int ccfd_gpu_get_prim(ccfd_gpu_ctx* c, double* prim) {
  if (!c || !prim || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh / null output");
  hipSetDevice(c->device);
  return download_aos(c, as_double(c, c->prim[c->cur], 4 * (size_t)c->Np), prim);
}

extern "C" int ccfd_gpu_get_cons(ccfd_gpu_ctx* c, double* cons) {
  if (!c || !cons || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh / null output");
  hipSetDevice(c->device);
  return download_aos(c, c->U, cons);
}

extern "C" int ccfd_gpu_get_ut(ccfd_gpu_ctx* c, double* ut) {
  if (!c || !ut || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh / null output");
  hipSetDevice(c->device);
  return download_aos(c, as_double(c, c->ut, 4 * (size_t)c->Np), ut);
}

// exact-function ghost states (side 0: barycentre, 1: side midpoint) after
// the boundary kernel that left them out, where the mesh has any
template <typename T>
static void exact_ghosts(ccfd_gpu_ctx* c, const OpArgs<T>& a, int side) {
  if (c->hasExact && c->G > 0)
    hipLaunchKernelGGL(bc_exact_kernel<T>, dim3((c->G + 127) / 128), dim3(128),
                       0, c->st, a, side);
}

// limited gradients and side states of the current state at the given
// time, into c->ux, c->uy, c->side (the fused operator does not materialise
// them)
template <typename T>
void ccfd::recompute_recon(ccfd_gpu_ctx* c, double time) {
  OpArgs<T> a = op_args<T>(c);
  a.time0 = time;
  if (c->G > 0)
    hipLaunchKernelGGL(bc_bary_kernel<T>, dim3((c->G + 127) / 128), dim3(128),
                       0, c->st, a);
  exact_ghosts<T>(c, a, 0);
  dim3 g(c->nb), b(kBlock);
  const int lim = c->prm.limiter;
  if (c->S == 3) {
    if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 3>), g, b, 0, c->st, a);
    else hipLaunchKernelGGL((recon_kernel<T, 2, 3>), g, b, 0, c->st, a);
  } else {
    if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 4>), g, b, 0, c->st, a);
    else hipLaunchKernelGGL((recon_kernel<T, 2, 4>), g, b, 0, c->st, a);
  }
}
template void ccfd::recompute_recon<double>(ccfd_gpu_ctx*, double);
template void ccfd::recompute_recon<float>(ccfd_gpu_ctx*, double);

extern "C" int ccfd_gpu_get_recon(ccfd_gpu_ctx* c, double* u_x, double* u_y,
                                  double* sideP) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  hipSetDevice(c->device);
  int rc;
  if (c->prm.spatialOrder == 2) {
    if (c->prm.precision == CCFD_FP64) recompute_recon<double>(c, c->lastTime);
    else recompute_recon<float>(c, c->lastTime);
  }
  if (u_x && (rc = download_aos(c, as_double(c, c->ux, 4 * (size_t)c->Np), u_x))) return rc;
  if (u_y && (rc = download_aos(c, as_double(c, c->uy, 4 * (size_t)c->Np), u_y))) return rc;
  if (sideP) {
    size_t sl = (size_t)c->S * c->Np;
    std::vector<double> tmp(4 * sl);
    const double* src = c->prm.spatialOrder == 2
                            ? as_double(c, c->side, 4 * sl)
                            : nullptr;
    if (src) {
      if (hipMemcpyAsync(tmp.data(), src, tmp.size() * 8, hipMemcpyDeviceToHost,
                         c->st) != hipSuccess ||
          hipStreamSynchronize(c->st) != hipSuccess)
        return fail(CCFD_EHIP, "side download failed");
      for (int i = 0; i < c->N; ++i) {
        int e = c->perm[i];
        long s0 = c->slotStart[e], ns = c->slotStart[e + 1] - s0;
        for (int k = 0; k < ns; ++k)
          for (int v = 0; v < 4; ++v)
            sideP[4 * (s0 + k) + v] = tmp[((size_t)v * c->S + k) * c->Np + i];
      }
    } else {
      std::vector<double> pr(4 * (size_t)c->N);
      if ((rc = ccfd_gpu_get_prim(c, pr.data()))) return rc;
      for (int e = 0; e < c->N; ++e)
        for (long s = c->slotStart[e]; s < c->slotStart[e + 1]; ++s)
          for (int v = 0; v < 4; ++v) sideP[4 * s + v] = pr[4 * (size_t)e + v];
    }
  }
  return CCFD_OK;
}

// ----------------------------------------------------------- operator

template <typename T>
OpArgs<T> ccfd::op_args(ccfd_gpu_ctx* c) {
  OpArgs<T> a;
  memset(&a, 0, sizeof(a));
  Geom<T>& g = geom<T>(c);
  a.N = c->N; a.Np = c->Np; a.S = c->S; a.G = c->G;
  a.order = c->prm.spatialOrder;
  a.ns = c->prm.navierstokes ? 1 : 0;
  a.link = c->link;
  a.w = g.w; a.gp = g.gp; a.fn = g.fn; a.flen = g.flen;
  a.bbv = g.bbv; a.bbd = g.bbd; a.areaq = g.areaq; a.venk = g.venk;
  a.bary = c->bary; a.area = c->area;
  a.ghostBC = c->ghostBC; a.ghostBary = c->ghostBary;
  a.ghostCell = c->ghostCell; a.ghostK = c->ghostK;
  a.ghostS = (T*)c->ghostS;
  a.bcIn = c->bcIn;
  a.prim = (const T*)c->prim[c->cur];
  a.side = (T*)c->side;
  a.ux = (T*)c->ux; a.uy = (T*)c->uy;
  a.ghostP = (T*)c->ghostP;
  a.ut = (T*)c->ut;
  a.write_ut = 1;
  a.write_U = 1;
  Eq<double> qd = eqd(c);
  a.eq.gam = T(qd.gam); a.eq.gam1 = T(qd.gam1); a.eq.gam1q = T(qd.gam1q);
  a.eq.mu = T(qd.mu); a.eq.Pr = T(qd.Pr);
  a.eqD = qd;
  a.ex = c->ex;
  a.bct = c->bct;
  return a;
}
template OpArgs<double> ccfd::op_args<double>(ccfd_gpu_ctx*);
template OpArgs<float> ccfd::op_args<float>(ccfd_gpu_ctx*);

static hipEvent_t take_event(ccfd_gpu_ctx* c) {
  if (c->evUsed == c->evpool.size()) {
    hipEvent_t e;
    hipEventCreate(&e);
    c->evpool.push_back(e);
  }
  return c->evpool[c->evUsed++];
}

void ccfd::tspan_begin(ccfd_gpu_ctx* c, hipEvent_t* e) {
  if (!c->timing) return;
  *e = take_event(c);
  hipEventRecord(*e, c->st);
}

void ccfd::tspan_end(ccfd_gpu_ctx* c, hipEvent_t a, int kind) {
  if (!c->timing) return;
  hipEvent_t b = take_event(c);
  hipEventRecord(b, c->st);
  c->spans.push_back({a, b, kind});
}

bool ccfd::fused_applies(const ccfd_gpu_ctx* c, size_t tsize) {
  if (c->prm.spatialOrder != 2 || c->prm.navierstokes) return false;
  return fused_lds_bytes(tsize, c->S, c->tile) + 256 <= c->ldsLimit;
}

// elements per block of the face-once kernel: two layout blocks when asked
// for and their LDS fits, else one (the automatic choice: on the 4M-cell
// periodic Delaunay mesh one block of 256 runs 4% faster than 512, whose
// 8-wave barriers cost more than its smaller halo saves)
int ccfd::face_tile(const ccfd_gpu_ctx* c, size_t tsize) {
  const bool wide = fused_lds_bytes(tsize, c->S, c->tile2, 2 * kBlock) + 512 <=
                    c->ldsLimit;
  return c->tileElems == 2 * kBlock && wide ? 2 * kBlock : kBlock;
}

int ccfd::active_variant(const ccfd_gpu_ctx* c, size_t tsize) {
  if (c->variant == CCFD_KV_SPLIT || !fused_applies(c, tsize))
    return CCFD_KV_SPLIT;
  return c->variant == CCFD_KV_AUTO ? CCFD_KV_FUSED_FACE : c->variant;
}

template <typename T>
void ccfd::launch_operator(ccfd_gpu_ctx* c, OpArgs<T>& a) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int kv = active_variant(c, sizeof(T));
  if (c->srcFunc) launch_source<T>(c, a);
  if (kv != CCFD_KV_SPLIT) {
    if (c->G > 0) {
      hipEvent_t eb = nullptr;
      tspan_begin(c, &eb);
      hipLaunchKernelGGL(bc_bary_kernel<T>, dim3((c->G + 127) / 128), dim3(128),
                         0, c->st, a);
      exact_ghosts<T>(c, a, 0);
      const dim3 gg((c->G + 127) / 128), gb(128);
      const int lim = c->prm.limiter;
      if (c->S == 3) {
        if (lim == 1) hipLaunchKernelGGL((bc_side_recon_kernel<T, 1, 3>), gg, gb, 0, c->st, a);
        else hipLaunchKernelGGL((bc_side_recon_kernel<T, 2, 3>), gg, gb, 0, c->st, a);
      } else {
        if (lim == 1) hipLaunchKernelGGL((bc_side_recon_kernel<T, 1, 4>), gg, gb, 0, c->st, a);
        else hipLaunchKernelGGL((bc_side_recon_kernel<T, 2, 4>), gg, gb, 0, c->st, a);
      }
      exact_ghosts<T>(c, a, 1);
      tspan_end(c, eb, KT_OTHER);
    }
    const int face = kv == CCFD_KV_FUSED_FACE ? 1 : 0;
    const int tb = face ? face_tile(c, sizeof(T)) : kBlock;
    tspan_begin(c, &e1);
    if (tb == kBlock) {
      FusedArgs f{c->haloStart, c->haloStart + 1, c->haloCells, c->tile};
      launch_fused_code<T>(c->prm.iFlux, a, f, c->prm.limiter, face, c->nb,
                           fused_lds_bytes(sizeof(T), c->S, c->tile), c->st);
    } else {
      FusedArgs f{c->haloStart2, c->haloStart2 + 1, c->haloCells2, c->tile2};
      launch_fused_code<T>(c->prm.iFlux, a, f, c->prm.limiter, face, c->nt2,
                           fused_lds_bytes(sizeof(T), c->S, c->tile2, tb),
                           c->st, tb);
    }
    tspan_end(c, e1, KT_FLUX);
    return;
  }
  if (a.order == 2) {
    const int lim = c->prm.limiter;
    dim3 g(c->nb), b(kBlock);
    if (c->G > 0) {
      hipEvent_t eb = nullptr;
      tspan_begin(c, &eb);
      hipLaunchKernelGGL(bc_bary_kernel<T>, dim3((c->G + 127) / 128), dim3(128),
                         0, c->st, a);
      exact_ghosts<T>(c, a, 0);
      tspan_end(c, eb, KT_OTHER);
    }
    tspan_begin(c, &e0);
    if (c->S == 3) {
      if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 3>), g, b, 0, c->st, a);
      else hipLaunchKernelGGL((recon_kernel<T, 2, 3>), g, b, 0, c->st, a);
    } else {
      if (lim == 1) hipLaunchKernelGGL((recon_kernel<T, 1, 4>), g, b, 0, c->st, a);
      else hipLaunchKernelGGL((recon_kernel<T, 2, 4>), g, b, 0, c->st, a);
    }
    tspan_end(c, e0, KT_RECON);
  }
  if (c->G > 0) {
    hipEvent_t es = nullptr;
    tspan_begin(c, &es);
    hipLaunchKernelGGL(bc_side_kernel<T>, dim3((c->G + 127) / 128), dim3(128), 0,
                       c->st, a);
    exact_ghosts<T>(c, a, 1);
    tspan_end(c, es, KT_OTHER);
  }
  tspan_begin(c, &e1);
  launch_flux_code<T>(c->prm.iFlux, a, c->nb, c->st);
  tspan_end(c, e1, KT_FLUX);
}
template void ccfd::launch_operator<double>(ccfd_gpu_ctx*, OpArgs<double>&);
template void ccfd::launch_operator<float>(ccfd_gpu_ctx*, OpArgs<float>&);

template <typename T>
static int fvtd_t(ccfd_gpu_ctx* c, double time) {
  OpArgs<T> a = op_args<T>(c);
  a.time0 = time;
  launch_operator<T>(c, a);
  return CCFD_OK;
}

extern "C" int ccfd_gpu_fvtd(ccfd_gpu_ctx* c, double time) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  hipSetDevice(c->device);
  c->lastTime = time;
  int rc = c->prm.precision == CCFD_FP64 ? fvtd_t<double>(c, time)
                                         : fvtd_t<float>(c, time);
  if (rc) return rc;
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "operator failed: %s",
                hipGetErrorString(hipGetLastError()));
  return CCFD_OK;
}

// ------------------------------------------------------------ time step

void ccfd::launch_dt(ccfd_gpu_ctx* c, int useDevT, double tHost, double pTime,
                     double stopTime) {
  hipEvent_t e0 = nullptr;
  tspan_begin(c, &e0);
  double gam = c->prm.gam;
  double mu = c->prm.navierstokes ? c->prm.mu : 0.0;
  double gamPrMax = fmax(4.0 / 3.0, gam / c->prm.Pr);
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_dt_partial<double>, dim3(c->nb), dim3(kBlock), 0, c->st,
                       c->N, c->Np, (const double*)c->prim[c->cur], c->area,
                       c->sx, c->sy, gam, mu, gamPrMax, c->prm.cfl, c->prm.dfl,
                       c->prm.timeStep1D, c->part);
  else
    hipLaunchKernelGGL(k_dt_partial<float>, dim3(c->nb), dim3(kBlock), 0, c->st,
                       c->N, c->Np, (const float*)c->prim[c->cur], c->area,
                       c->sx, c->sy, gam, mu, gamPrMax, c->prm.cfl, c->prm.dfl,
                       c->prm.timeStep1D, c->part);
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(1024), 0, c->st, c->nb, c->part,
                     c->scal, useDevT, tHost, pTime, stopTime);
  tspan_end(c, e0, KT_DT);
}

extern "C" int ccfd_gpu_calc_dt(ccfd_gpu_ctx* c, double t, double pTime,
                                double stopTime, double* dt, int* vd) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  hipSetDevice(c->device);
  hipMemsetAsync(c->scal + 3, 0, 8, c->st);
  launch_dt(c, 0, t, pTime, stopTime);
  if (hipMemcpyAsync(c->hscal, c->scal, 8 * 8, hipMemcpyDeviceToHost, c->st) !=
          hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "dt reduction failed");
  if (dt) *dt = c->hscal[1];
  if (vd) *vd = c->hscal[2] != 0.0;
  if (c->hscal[3] != 0.0)
    return fail(CCFD_ENONFINITE, "convective or viscous time step NaN");
  return CCFD_OK;
}

// one explicit step; dt from the host (dtdev == nullptr) or the device
// globalResidual from the per-block partials in c->part: one workgroup
// for few blocks, else a middle stage of kResMid blocks first
static void launch_res_final(ccfd_gpu_ctx* c, int advanceT) {
  if (c->nb <= 4 * kResMid) {
    hipLaunchKernelGGL(k_res_final, dim3(1), dim3(256), 0, c->st, c->nb,
                       c->part, c->totalArea_q, c->scal, advanceT);
    return;
  }
  double* mid = c->part + 4 * (size_t)c->nb;
  hipLaunchKernelGGL(k_res_mid, dim3(kResMid), dim3(256), 0, c->st, c->nb,
                     c->part, mid);
  hipLaunchKernelGGL(k_res_final, dim3(1), dim3(256), 0, c->st, kResMid, mid,
                     c->totalArea_q, c->scal, advanceT);
}

template <typename T>
static void step_t(ccfd_gpu_ctx* c, double t, double dt, bool devDt, bool res) {
  std::swap(c->U, c->U0);   // U0 <- current state
  int nRK = c->prm.nRKstages;
  bool euler = (nRK == 1 && c->prm.timeOrder == 1);
  int nst = euler ? 1 : nRK;
  for (int k = 1; k <= nst; ++k) {
    OpArgs<T> a = op_args<T>(c);
    if (devDt) {
      a.tdev = c->scal;
      a.tcoef = c->rk[k - 1];
      a.dtdev = c->scal + 1;
    } else {
      a.time0 = t + c->rk[k - 1] * dt;
      a.dt0 = dt;
    }
    a.rk = 1;
    a.rkc = c->rk[k];
    a.U0 = c->U0;
    a.U = c->U;
    a.prim_out = (T*)c->prim[1 - c->cur];
    // dU/dt of the last stage is kept for host readback; earlier stages
    // feed the update only
    a.write_ut = (k == nst) ? 1 : 0;
    a.write_U = (k == nst) ? 1 : 0;
    a.res_part = (res && k == nst) ? c->part : nullptr;
    if (!c->keepGradients && !a.ns) { a.ux = nullptr; a.uy = nullptr; }
    launch_operator<T>(c, a);
    c->cur = 1 - c->cur;
  }
  if (res)
    launch_res_final(c, devDt ? 1 : 0);
  else if (devDt)
    hipLaunchKernelGGL(k_advance_t, dim3(1), dim3(64), 0, c->st, c->scal);
}

extern "C" int ccfd_gpu_explicit_step(ccfd_gpu_ctx* c, double t, double dt,
                                      double res[4]) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  hipSetDevice(c->device);
  if (c->prm.precision == CCFD_FP64) step_t<double>(c, t, dt, false, res != nullptr);
  else step_t<float>(c, t, dt, false, res != nullptr);
  if (res) {
    if (hipMemcpyAsync(c->hscal, c->scal, 8 * 8, hipMemcpyDeviceToHost, c->st) !=
        hipSuccess)
      return fail(CCFD_EHIP, "residual copy failed");
  }
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "explicit step failed: %s",
                hipGetErrorString(hipGetLastError()));
  if (res)
    for (int v = 0; v < 4; ++v) res[v] = c->hscal[4 + v];
  return CCFD_OK;
}

extern "C" int ccfd_gpu_advance(ccfd_gpu_ctx* c, int nSteps, double* t,
                                double pTime, double stopTime, double res[4]) {
  if (!c || !c->hasMesh || !t) return fail(CCFD_EINVAL, "no mesh / null t");
  if (nSteps < 0) return fail(CCFD_EINVAL, "negative step count");
  hipSetDevice(c->device);
  c->hscal[0] = *t;
  hipMemcpyAsync(c->scal, c->hscal, 8, hipMemcpyHostToDevice, c->st);
  hipMemsetAsync(c->scal + 3, 0, 8, c->st);
  // each step: calcTimeStep, the RK stages and, when res is requested,
  // globalResidual (src/timeDiscretization.c:362-401), all on the device
  for (int s = 0; s < nSteps; ++s) {
    launch_dt(c, 1, 0.0, pTime, stopTime);
    const bool r = res != nullptr;
    if (c->prm.precision == CCFD_FP64) step_t<double>(c, 0.0, 0.0, true, r);
    else step_t<float>(c, 0.0, 0.0, true, r);
  }
  if (hipMemcpyAsync(c->hscal, c->scal, 8 * 8, hipMemcpyDeviceToHost, c->st) !=
          hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "advance failed: %s",
                hipGetErrorString(hipGetLastError()));
  *t = c->hscal[0];
  if (res && nSteps > 0)
    for (int v = 0; v < 4; ++v) res[v] = c->hscal[4 + v];
  if (c->hscal[3] != 0.0)
    return fail(CCFD_ENONFINITE, "convective or viscous time step NaN");
  return CCFD_OK;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_res_partial(int N, int Np,
                                                        const T* __restrict__ ut,
                                                        const double* __restrict__ area,
                                                        double* part) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < N) {
    for (int v = 0; v < 4; ++v) {
      double u = double(ut[v * Np + i]);
      r[v] = area[i] * u * u;
    }
  }
  block_sum4<T>(r, part, blockIdx.x, gridDim.x);
}

extern "C" int ccfd_gpu_residual(ccfd_gpu_ctx* c, double res[4]) {
  if (!c || !c->hasMesh || !res) return fail(CCFD_EINVAL, "no mesh / null res");
  hipSetDevice(c->device);
  return residual_of(c, c->prm.precision == CCFD_FP64, res);
}

int ccfd::residual_of(ccfd_gpu_ctx* c, bool fp64, double res[4]) {
  if (fp64)
    hipLaunchKernelGGL(k_res_partial<double>, dim3(c->nb), dim3(kBlock), 0, c->st,
                       c->N, c->Np, (const double*)c->ut, c->area, c->part);
  else
    hipLaunchKernelGGL(k_res_partial<float>, dim3(c->nb), dim3(kBlock), 0, c->st,
                       c->N, c->Np, (const float*)c->ut, c->area, c->part);
  launch_res_final(c, 0);
  if (hipMemcpyAsync(c->hscal, c->scal, 8 * 8, hipMemcpyDeviceToHost, c->st) !=
          hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "residual failed");
  for (int v = 0; v < 4; ++v) res[v] = c->hscal[4 + v];
  return CCFD_OK;
}

// dot product of two [4][Np] vectors: fixed-order block partials + final sum
__global__ __launch_bounds__(kBlock) void k_dot_partial(int N, int Np,
                                                      const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      double* part) {
  __shared__ double red[kBlock / 64];
  int i = blockIdx.x * kBlock + threadIdx.x;
  double x = 0.0;
  if (i < N) {
    x += a[i] * b[i];
    x += a[Np + i] * b[Np + i];
    x += a[2 * Np + i] * b[2 * Np + i];
    x += a[3 * Np + i] * b[3 * Np + i];
  }
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w];
    part[blockIdx.x] = s;
  }
}

__global__ void k_sum_final(int nb, const double* __restrict__ part,
                            double* out) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) acc += part[k];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

void ccfd::launch_dot(ccfd_gpu_ctx* c, const double* a, const double* b,
                      double* out, double* part) {
  hipLaunchKernelGGL(k_dot_partial, dim3(c->nb), dim3(kBlock), 0, c->st, c->N,
                     c->Np, a, b, part);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, c->st, c->nb, part, out);
}

int ccfd::upload_vec(ccfd_gpu_ctx* c, const double* h, double* d) {
  std::vector<double> tmp(4 * (size_t)c->Np, 0.0);
  for (int i = 0; i < c->N; ++i) {
    int e = c->perm[i];
    for (int v = 0; v < 4; ++v) tmp[(size_t)v * c->Np + i] = h[4 * (size_t)e + v];
  }
  if (hipMemcpyAsync(d, tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "vector upload failed");
  return CCFD_OK;
}

int ccfd::download_vec(ccfd_gpu_ctx* c, const double* d, double* h) {
  return download_aos(c, d, h);
}

extern "C" int ccfd_gpu_dot(ccfd_gpu_ctx* c, const double* a, const double* b,
                            double* out) {
  if (!c || !c->hasMesh || !a || !b || !out)
    return fail(CCFD_EINVAL, "bad dot arguments");
  hipSetDevice(c->device);
  size_t bytes = 4 * (size_t)c->Np * 8;
  double *da = nullptr, *db = nullptr;
  if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&db, bytes) != hipSuccess) {
    hipFree(da);
    return fail(CCFD_ENOMEM, "dot allocation failed");
  }
  int rc = upload_vec(c, a, da);
  if (!rc) rc = upload_vec(c, b, db);
  if (!rc) {
    launch_dot(c, da, db, c->scal + 8, c->part);
    if (hipMemcpyAsync(c->hscal, c->scal, 16 * 8, hipMemcpyDeviceToHost,
                       c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      rc = fail(CCFD_EHIP, "dot failed");
    else
      *out = c->hscal[8];
  }
  hipFree(da);
  hipFree(db);
  return rc;
}

extern "C" int ccfd_gpu_riemann_batch(int code, int precision, double gam,
                                      long n, const double* L, const double* R,
                                      double* F) {
  if (code < 1 || code > 11) return fail(CCFD_EINVAL, "unknown flux %d", code);
  if (n < 0 || (n > 0 && (!L || !R || !F))) return fail(CCFD_EINVAL, "bad arrays");
  if (n == 0) return CCFD_OK;
  double *dL = nullptr, *dR = nullptr, *dF = nullptr;
  size_t b = 4 * (size_t)n * 8;
  if (hipMalloc(&dL, b) != hipSuccess || hipMalloc(&dR, b) != hipSuccess ||
      hipMalloc(&dF, b) != hipSuccess) {
    hipFree(dL); hipFree(dR); hipFree(dF);
    return fail(CCFD_ENOMEM, "riemann batch allocation failed");
  }
  hipMemcpy(dL, L, b, hipMemcpyHostToDevice);
  hipMemcpy(dR, R, b, hipMemcpyHostToDevice);
  dim3 g((n + 255) / 256), bl(256);
  if (precision == CCFD_FP64) {
    Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), 0.0, 0.72};
    hipLaunchKernelGGL(k_riemann<double>, g, bl, 0, 0, code, n, dL, dR, dF, q);
  } else {
    Eq<float> q = {float(gam), float(gam - 1.0), float(1.0 / (gam - 1.0)), 0.f, 0.72f};
    hipLaunchKernelGGL(k_riemann<float>, g, bl, 0, 0, code, n, dL, dR, dF, q);
  }
  hipError_t e = hipMemcpy(F, dF, b, hipMemcpyDeviceToHost);
  hipFree(dL); hipFree(dR); hipFree(dF);
  return e == hipSuccess ? CCFD_OK : fail(CCFD_EHIP, "riemann batch failed");
}

// ---- convectiveFlux on n states without allocating (staging in `scratch`)
//
// The device staging of L, R and F lives in a file-scope device array, so
// the call allocates nothing: for callers that cannot call hipMalloc (a
// per-face convectiveFlux loop, code under graph capture, serving loops).
// Host code must not pass the array's name to a kernel or a copy: its
// device address comes from hipGetSymbolAddress, once per device, and is
// kept in g_scratch[device]. Every call on one device uses that device's
// array; calls that overlap on one device (other host threads) race.
constexpr long kNoAllocMaxStates = CCFD_NOALLOC_MAX_STATES;
static __device__ __align__(256) double scratch[kNoAllocMaxStates * 12];

namespace {
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;
}  // namespace

// device address of this device's staging array (set up on first use)
static int scratch_ptr(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(CCFD_EHIP, "no usable current device");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CCFD_EHIP, "staging array address: %s", hipGetErrorString(e));
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return CCFD_OK;
}

extern "C" int ccfd_gpu_riemann_batch_noalloc(int code, int precision,
                                              double gam, long n,
                                              const double* L, const double* R,
                                              double* F) {
  if (code < 1 || code > 11) return fail(CCFD_EINVAL, "unknown flux %d", code);
  if (n < 0 || (n > 0 && (!L || !R || !F))) return fail(CCFD_EINVAL, "bad arrays");
  if (n > kNoAllocMaxStates)
    return fail(CCFD_EINVAL,
                "%ld states exceed the %ld of the preallocated staging array",
                n, kNoAllocMaxStates);
  if (n == 0) return CCFD_OK;
  double* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  double* dL = base;
  double* dR = base + 4 * (size_t)kNoAllocMaxStates;
  double* dF = base + 8 * (size_t)kNoAllocMaxStates;
  const size_t b = 4 * (size_t)n * 8;
  if (hipMemcpy(dL, L, b, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dR, R, b, hipMemcpyHostToDevice) != hipSuccess)
    return fail(CCFD_EHIP, "riemann staging upload failed");
  dim3 g((n + 255) / 256), bl(256);
  if (precision == CCFD_FP64) {
    Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), 0.0, 0.72};
    hipLaunchKernelGGL(k_riemann<double>, g, bl, 0, 0, code, n, dL, dR, dF, q);
  } else {
    Eq<float> q = {float(gam), float(gam - 1.0), float(1.0 / (gam - 1.0)), 0.f, 0.72f};
    hipLaunchKernelGGL(k_riemann<float>, g, bl, 0, 0, code, n, dL, dR, dF, q);
  }
  hipError_t e = hipMemcpy(F, dF, b, hipMemcpyDeviceToHost);
  return e == hipSuccess ? CCFD_OK : fail(CCFD_EHIP, "riemann batch failed");
}

extern "C" int ccfd_gpu_profile_steps(ccfd_gpu_ctx* c, int nSteps, double pTime,
                                     double stopTime, int residual, double* ms,
                                     int* launches, double res[4]) {
  if (!c || !c->hasMesh || !ms || nSteps <= 0)
    return fail(CCFD_EINVAL, "bad profile arguments");
  hipSetDevice(c->device);
  c->timing = true;
  c->evUsed = 0;
  c->spans.clear();
  double t = 0.0;
  hipEvent_t all0 = take_event(c), all1 = nullptr;
  hipEventRecord(all0, c->st);
  double r4[4];
  int rc = ccfd_gpu_advance(c, nSteps, &t, pTime, stopTime,
                            residual ? r4 : nullptr);
  all1 = take_event(c);
  hipEventRecord(all1, c->st);
  c->timing = false;
  if (rc) return rc;
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "profile sync failed");
  double sum[KT_N] = {0}, n[KT_N] = {0};
  for (auto& sp : c->spans) {
    float x = 0.f;
    hipEventElapsedTime(&x, sp.a, sp.b);
    sum[sp.kind] += x;
    n[sp.kind] += 1;
  }
  float tot = 0.f;
  hipEventElapsedTime(&tot, all0, all1);
  for (int k = 0; k < KT_N; ++k) {
    ms[k] = n[k] > 0 ? sum[k] / n[k] : 0.0;
    if (launches) launches[k] = (int)n[k];
  }
  ms[KT_N] = tot / nSteps;
  if (residual && res)
    for (int v = 0; v < 4; ++v) res[v] = r4[v];
  c->spans.clear();
  c->evUsed = 0;
  return CCFD_OK;
}

// plain streaming copy, 16 B per thread per iteration: the calibration
// kernel for HBM byte counters and the reachable-bandwidth reference
__global__ __launch_bounds__(256) void copy_probe_kernel(
    const double2* __restrict__ src, double2* __restrict__ dst, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long st = (long)gridDim.x * blockDim.x;
  for (; i < n; i += st) dst[i] = src[i];
}

// the same copy with one 4 x 16 B tile per thread and one block per 16 KiB
// (no grid-stride loop): all loads of a thread are issued before its stores
__global__ __launch_bounds__(256) void copy_probe_tile_kernel(
    const double2* __restrict__ src, double2* __restrict__ dst, long n) {
  const long i0 = (long)blockIdx.x * 1024 + threadIdx.x;
  double2 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const long i = i0 + 256 * k;
    if (i < n) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const long i = i0 + 256 * k;
    if (i < n) dst[i] = v[k];
  }
}

extern "C" int ccfd_gpu_copy_probe(int device, size_t bytes, int reps,
                                   double* gbs) {
  if (bytes < 4096 || reps <= 0 || !gbs)
    return fail(CCFD_EINVAL, "bad copy probe arguments");
  if (hipSetDevice(device) != hipSuccess)
    return fail(CCFD_EHIP, "hipSetDevice failed");
  long n = (long)(bytes / 16);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    hipFree(a);
    return fail(CCFD_ENOMEM, "copy probe allocation failed");
  }
  hipMemset(a, 0, n * 16);
  // two launch shapes, the faster one counts: a grid-stride loop over
  // 8 waves per CU on 256 CUs, and one 16 KiB tile per block
  const int strideBlocks = 256 * 8 * 4;
  const long tileBlocks = (n + 1023) / 1024;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  double best = 0.0;
  hipError_t err = hipSuccess;
  for (int shape = 0; shape < 2 && err == hipSuccess; ++shape) {
    for (int r = 0; r <= reps; ++r) {   // first launch: warm-up
      if (r == 1) hipEventRecord(e0, 0);
      if (shape == 0)
        hipLaunchKernelGGL(copy_probe_kernel, dim3(strideBlocks), dim3(256), 0,
                           0, a, b, n);
      else
        hipLaunchKernelGGL(copy_probe_tile_kernel, dim3(tileBlocks), dim3(256),
                           0, 0, a, b, n);
    }
    hipEventRecord(e1, 0);
    err = hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess && ms > 0.f)
      best = std::max(best, 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  if (err != hipSuccess) return fail(CCFD_EHIP, "copy probe failed");
  *gbs = best;
  return CCFD_OK;
}

extern "C" int ccfd_gpu_set_tile(ccfd_gpu_ctx* c, int elems) {
  if (!c || !c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  if (elems != 0 && elems != kBlock && elems != 2 * kBlock)
    return fail(CCFD_EINVAL, "tile must be 0 (automatic), %d or %d elements",
                kBlock, 2 * kBlock);
  c->tileElems = elems;
  return CCFD_OK;
}

extern "C" int ccfd_gpu_get_tile(ccfd_gpu_ctx* c, int* elems) {
  if (!c || !c->hasMesh || !elems) return fail(CCFD_EINVAL, "bad arguments");
  const size_t ts = c->prm.precision == CCFD_FP64 ? 8 : 4;
  *elems = active_variant(c, ts) == CCFD_KV_FUSED_FACE ? face_tile(c, ts)
                                                       : kBlock;
  return CCFD_OK;
}

// time each applicable operator implementation and tile on the resident
// state (reps fvTimeDerivative calls each, after one untimed call; the
// state is not changed) and keep the fastest (a later candidate must beat
// the one kept by 2%); ms[k] gets the time per call
// of candidate k: split, fused, face-once 256, face-once 512 (-1 where it
// does not apply)
extern "C" int ccfd_gpu_autotune(ccfd_gpu_ctx* c, int reps, int* variant,
                                 int* tile, double ms[4]) {
  if (!c || !c->hasMesh || reps < 1 || !variant || !tile || !ms)
    return fail(CCFD_EINVAL, "bad autotune arguments");
  hipSetDevice(c->device);
  const size_t ts = c->prm.precision == CCFD_FP64 ? 8 : 4;
  const int cand[4][2] = {{CCFD_KV_SPLIT, kBlock}, {CCFD_KV_FUSED, kBlock},
                          {CCFD_KV_FUSED_FACE, kBlock},
                          {CCFD_KV_FUSED_FACE, 2 * kBlock}};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(CCFD_EHIP, "event creation failed");
  const int sv = c->variant, st = c->tileElems;
  int best = -1;
  for (int k = 0; k < 4; ++k) {
    ms[k] = -1.0;
    c->variant = cand[k][0];
    c->tileElems = cand[k][1];
    if (active_variant(c, ts) != cand[k][0] ||
        (cand[k][0] == CCFD_KV_FUSED_FACE && face_tile(c, ts) != cand[k][1]))
      continue;
    int rc = ccfd_gpu_fvtd(c, c->lastTime);
    hipEventRecord(e0, c->st);
    for (int r = 0; rc == CCFD_OK && r < reps; ++r) rc = ccfd_gpu_fvtd(c, c->lastTime);
    hipEventRecord(e1, c->st);
    float x = 0.f;
    if (rc || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&x, e0, e1) != hipSuccess) {
      c->variant = sv;
      c->tileElems = st;
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return rc ? rc : fail(CCFD_EHIP, "autotune timing failed");
    }
    ms[k] = x / reps;
    // a later candidate has to be 2% faster to replace an earlier one
    // (timing noise must not pick the wider tile over an equal default)
    if (best < 0 || ms[k] < 0.98 * ms[best]) best = k;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->variant = cand[best][0];
  c->tileElems = cand[best][1];
  *variant = c->variant;
  *tile = cand[best][1];
  return CCFD_OK;
}

extern "C" int ccfd_gpu_set_variant(ccfd_gpu_ctx* c, int variant) {
  if (!c) return fail(CCFD_EINVAL, "null context");
  if (variant < CCFD_KV_AUTO || variant > CCFD_KV_FUSED_FACE)
    return fail(CCFD_EINVAL, "unknown kernel variant %d", variant);
  c->variant = variant;
  return CCFD_OK;
}

extern "C" int ccfd_gpu_get_variant(ccfd_gpu_ctx* c, int* variant) {
  if (!c || !variant) return fail(CCFD_EINVAL, "null argument");
  if (!c->hasMesh) return fail(CCFD_EINVAL, "no mesh");
  const size_t ts = c->prm.precision == CCFD_FP64 ? 8 : 4;
  if ((c->variant == CCFD_KV_FUSED || c->variant == CCFD_KV_FUSED_FACE) &&
      !fused_applies(c, ts))
    return fail(CCFD_ENOTSUP,
                "fused operator needs 2nd order Euler and %zu bytes of LDS "
                "(device limit %zu)",
                fused_lds_bytes(ts, c->S, c->tile), c->ldsLimit);
  *variant = active_variant(c, ts);
  return CCFD_OK;
}

extern "C" int ccfd_gpu_synchronize(ccfd_gpu_ctx* c) {
  if (!c) return fail(CCFD_EINVAL, "null context");
  hipSetDevice(c->device);
  return hipStreamSynchronize(c->st) == hipSuccess ? CCFD_OK
                                                   : fail(CCFD_EHIP, "sync failed");
}
