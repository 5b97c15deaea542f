// ctx.hpp - the device context behind ccfd_gpu_ctx and the internal helpers
// shared by ccfd_gpu.hip (explicit path) and implicit.hip (JFNK path).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ccfd_gpu.h"
#include "fused.hpp"

namespace ccfd {

// geometry in the operator's arithmetic
template <typename T>
struct Geom {
  T* w = nullptr;      // [2][S][Np] least-squares weights
  T* gp = nullptr;     // [2][S][Np] side midpoint - barycentre
  T* fn = nullptr;     // [2][S][Np] owner-side normal of the face
  T* flen = nullptr;   // [S][Np] face length, negative if not the owner
  T* bbv = nullptr;    // [2][S][Np] barycentre-to-barycentre vector
  T* bbd = nullptr;    // [S][Np] its length
  T* areaq = nullptr;  // [Np] 1 / area
  T* venk = nullptr;   // [Np] Venkatakrishnan eps^2
};

struct ImplicitState;  // implicit.hip

}  // namespace ccfd

struct ccfd_gpu_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  bool hasMesh = false;
  ccfd_gpu_params prm{};
  double rk[6] = {0};
  int N = 0, Np = 0, S = 0, G = 0, nb = 0;
  long nSlotsHost = 0;
  double totalArea_q = 0.0;
  std::vector<int> perm, inv;       // new -> old, old -> new element
  std::vector<long> slotStart;      // host slot CSR (reference order)
  std::vector<double> hArea;        // renumbered element areas
  // mesh
  int* link = nullptr;
  int* ghostBC = nullptr;
  int* ghostCell = nullptr;
  int* ghostK = nullptr;
  double* ghostBary = nullptr;
  double* bary = nullptr;
  double* area = nullptr;
  double* sx = nullptr;
  double* sy = nullptr;
  ccfd::Geom<double> g64;
  ccfd::Geom<float> g32;
  ccfd::BCTable bct{};
  bool hasExact = false;   // an exact-function boundary (type 6) is present
  ccfd::ExactCfg ex{};
  // state (buffers sized for double, used as the operator type T)
  void* prim[2] = {nullptr, nullptr};
  int cur = 0;
  double* U = nullptr;
  double* U0 = nullptr;
  void* ut = nullptr;
  void* ux = nullptr;
  void* uy = nullptr;
  void* side = nullptr;
  void* ghostP = nullptr;   // ghost cell states [4][G]
  void* ghostS = nullptr;   // ghost side states [4][G]
  double* stage = nullptr;   // staging for host I/O
  double* part = nullptr;    // per-block reduction partials
  double* scal = nullptr;    // device scalars: t, dt, viscDom, nonfinite, res[4]
  double* hscal = nullptr;   // pinned host mirror of scal
  bool keepGradients = false;
  double lastTime = 0.0;   // time of the last ccfd_gpu_fvtd call
  // fused operator: per-block halo element lists and LDS tile columns
  int* haloStart = nullptr;
  int* haloCells = nullptr;
  int tile = 0;
  // the same for tiles of two layout blocks (wide face-once kernel)
  int* haloStart2 = nullptr;
  int* haloCells2 = nullptr;
  int tile2 = 0, nt2 = 0;
  int tileElems = 0;        // face-once tile: 0 automatic, 256 or 512
  int variant = 0;          // ccfd_kernel_variant
  size_t ldsLimit = 65536;  // per-block LDS of the device
  // optional per-kernel timing with HIP events on the context stream
  struct Span { hipEvent_t a, b; int kind; };
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  size_t evUsed = 0;
  std::vector<Span> spans;
  std::vector<void*> allocs;
  ccfd::ImplicitState* imp = nullptr;
  // analysis consumers and source term (analyze.hip)
  double* bcIn = nullptr;     // [4][G] interior side states at the boundary
                              // sides from the last operator evaluation
  int srcFunc = 0;            // source function (0: none)
  void* src = nullptr;        // [4][Np] source term of the last evaluation
  std::vector<double> hq5;    // [2][Np] quadrilateral centre points (host;
                              // empty for triangle meshes)
  double* q5 = nullptr;       // the same on the device, uploaded on first use
  int nWing = 0, nWingP = 0;  // calcCoef sides (pressure side first)
  int* wingG = nullptr;       // ghost of each listed side
  double* wingN = nullptr;    // [3][nWing] nx, ny, len of the interior side
  double* wingOut = nullptr;  // [2 + nWing] cl, cd, cp
  double refLength = 1.0;
  int recCap = 0;             // record points (gather buffers)
  int* recIdx = nullptr;
  double* recOut = nullptr;
  double* errPart = nullptr;  // [12][nb] error-norm partials
};

namespace ccfd {

int fail(int code, const char* fmt, ...);
void* dmalloc(ccfd_gpu_ctx* c, size_t bytes);
// validate parameters; rk receives the low-storage RK coefficients
// (src/timeDiscretization.c:94-178), rk[k] for stage k = 1..nRKstages
int check_params(const ccfd_gpu_params* p, double rk[6]);
int apply_params(ccfd_gpu_ctx* c, const ccfd_gpu_params* p);
int upload_venk(ccfd_gpu_ctx* c);
int refresh_prim(ccfd_gpu_ctx* c);
// globalResidual of c->ut read as FP64 or as FP32 values
int residual_of(ccfd_gpu_ctx* c, bool fp64, double res[4]);
Eq<double> eqd(const ccfd_gpu_ctx* c);
template <typename T>
OpArgs<T> op_args(ccfd_gpu_ctx* c);
template <typename T>
void launch_operator(ccfd_gpu_ctx* c, OpArgs<T>& a);
void launch_dt(ccfd_gpu_ctx* c, int useDevT, double tHost, double pTime,
               double stopTime);
void implicit_free(ccfd_gpu_ctx* c);
// out = sum a.b over [4][Np] vectors (device scalar), part: >= nb scratch
void launch_dot(ccfd_gpu_ctx* c, const double* a, const double* b, double* out,
                double* part);
// host AoS (reference order) <-> device [4][Np] (renumbered)
int upload_vec(ccfd_gpu_ctx* c, const double* h, double* d);
int download_vec(ccfd_gpu_ctx* c, const double* d, double* h);
// kernel kinds for timing spans
enum { KT_DT = 0, KT_RECON = 1, KT_FLUX = 2, KT_OTHER = 3, KT_N = 4 };
void tspan_begin(ccfd_gpu_ctx* c, hipEvent_t* e);
void tspan_end(ccfd_gpu_ctx* c, hipEvent_t a, int kind);
// can the fused kernels run (2nd order Euler, LDS tile fits), and which
// ccfd_kernel_variant the next operator call of type T runs
bool fused_applies(const ccfd_gpu_ctx* c, size_t tsize);
int face_tile(const ccfd_gpu_ctx* c, size_t tsize);
int active_variant(const ccfd_gpu_ctx* c, size_t tsize);
// limited gradients (c->ux, c->uy) and side states (c->side) of the current
// state at the given time, by the split kernels
template <typename T>
void recompute_recon(ccfd_gpu_ctx* c, double time);
// source term of the operator call a is about to make (analyze.hip); a.src
// is set when the context has a source function
template <typename T>
void launch_source(ccfd_gpu_ctx* c, OpArgs<T>& a);
// free one allocation made by dmalloc
void dfree(ccfd_gpu_ctx* c, void* p);

}  // namespace ccfd
