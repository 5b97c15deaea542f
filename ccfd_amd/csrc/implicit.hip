// implicit.hip - backward-Euler Newton with JFNK-GMRES and LU-SGS on the
// device (implicitTimeStep src/timeDiscretization.c:412-552; GMRES_M,
// matrixVector, buildMatrix, calcDinv, LUSGS src/linearSolver.c:115-510).
//
//  * Newton loop on the host, all vectors device-resident ([4][Np], FP64).
//  * GMRES_M: one cycle of at most nKdim iterations (optionally restarted
//    cycles); modified Gram-Schmidt, Givens rotations and back-solve in
//    FP64 on the device, with no host round trip per iteration. Krylov
//    vectors V, Z are FP64, or FP32 in FP32-flux mode.
//  * matrixVector: Frechet difference with epsFD = sqrt(eps)/|v| in FP64
//    (eps of the operator's arithmetic).
//  * buildMatrix: the reference records only the diagonal and face-neighbour
//    4x4 blocks of dR/dU (perturbing the primitive state, src/linearSolver.c:
//    195-224); here those blocks are stored block-sparse and assembled by
//    coloured finite differences: all elements of one colour of a distance-3
//    (2nd order) / distance-2 (1st order) colouring are perturbed together,
//    so each recorded block sees exactly the one perturbation the reference
//    applies for it.
//  * LUSGS (precond = 1): forward / backward Gauss-Seidel sweeps in the
//    reference's element order, executed level by level (cells of one level
//    have no dependence on each other), so each row's arithmetic is the
//    reference's. precond = 2: symmetric Gauss-Seidel sweeps by colour on
//    the Jacobian with respect to the conservative state; precond = 3: a
//    multigrid V-cycle of agglomerated elements around those sweeps.
//  * Steady solves (ccfd_gpu_implicit_steady): local pseudo-time steps, rows
//    scaled by sqrt(area) / dt, CFL number driven by the GMRES convergence,
//    Newton updates bounded to 20% density / pressure change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "agglomerate.hpp"
#include "ctx.hpp"

using namespace ccfd;

namespace ccfd {

struct ImplicitState {
  int nK = 0;
  int prec = -1;          // precision the buffers were sized for
  double *XK = nullptr, *RXK = nullptr, *Q = nullptr, *FXK = nullptr;
  double *dX = nullptr, *W = nullptr, *B = nullptr;
  void* V = nullptr;      // (nK + 1) x [4][Np] Krylov basis (double / float)
  void* Z = nullptr;      // nK x [4][Np] preconditioned basis
  double* H = nullptr;    // device Hessenberg column scratch (nK + 2)
  double* alp = nullptr;  // back-solve coefficients
  double* part = nullptr; // dot partials
  // LU-SGS
  bool haveColors = false;
  int orderColored = 0;
  int nColors = 0;
  int* color = nullptr;         // [Np] colour of each element
  void* Dinv = nullptr;         // [16][Np]
  void* Dg = nullptr;           // [16][Np] accumulation of the diagonal block
  void* Blk = nullptr;          // [16][S][Np] neighbour blocks
  double* primBase = nullptr;   // [4][Np] primitive state at XK
  int* refIdx = nullptr;        // [Np] reference element index
  int* lvlF = nullptr;          // cells of forward levels (CSR items)
  int* lvlB = nullptr;          // backward
  std::vector<int> lvlFStart, lvlBStart;
  // multicolour Gauss-Seidel sweeps (precond = 2, 3): greedy distance-1
  // colouring; elements of one colour are updated together
  bool haveMC = false;
  int* mcCells = nullptr;       // cells by colour (CSR items)
  std::vector<int> mcStart;
  // the sweeps' operands in colour order (row p is element mcCells[p]), so
  // that a colour's sweep reads contiguous rows: neighbour ids [S][N],
  // side blocks [16][S][N] and inverted diagonal blocks [16][N]
  int* mcLink = nullptr;
  void* mcBlk = nullptr;
  void* mcDinv = nullptr;
  // aggregation multigrid (precond = 3): coarse levels below the elements
  struct MGLevel {
    int n = 0, nnz = 0, nPrev = 0;
    int *aggOf = nullptr, *memStart = nullptr, *mem = nullptr;
    int *rowStart = nullptr, *col = nullptr, *colorCells = nullptr;
    int64_t *dSrcStart = nullptr, *dSrc = nullptr;
    int64_t *oSrcStart = nullptr, *oSrc = nullptr;
    std::vector<int> colorStart;
    double *D = nullptr, *Dinv = nullptr, *O = nullptr;   // [16][n], [16][nnz]
    double *b = nullptr, *x = nullptr, *r = nullptr;      // [4][n]
  };
  bool haveMG = false;
  std::vector<MGLevel> mg;
  double* xs = nullptr;         // [4][Np] forward-sweep result
  int* flag = nullptr;          // singular-block flag
  double* dtLoc = nullptr;      // [Np] local pseudo-time steps (steady solves)
  double gmresRel = 0.0;        // |residual| / |rhs| the last GMRES call reached
  // device-resident Arnoldi (GmresDev): H, rotations, gm, alp, scalars
  double* gdev = nullptr;
  double* partA = nullptr;      // [kRedBlocks] partial sums, ping-pong
  double* partB = nullptr;
  double* cgsPart = nullptr;    // [nK + 1][kRedBlocks] CGS2 dot partials
  double* cgsH = nullptr;       // [nK + 8] one CGS2 pass's coefficients
  int* hflag = nullptr;         // pinned: per iteration, 1 = cycle done
  int* hflagDev = nullptr;      // its device address
  std::vector<hipEvent_t> ev;   // end of each queued iteration
  double steadyCfl = 0.0;       // CFL number a steady solve stopped at
  // first-order conservative Jacobian: 1 = element by element (jac1_kernel),
  // 0 = coloured finite differences (the general path)
  int jacMethod = 1;
  std::vector<void*> allocs;
};

}  // namespace ccfd

static void* imalloc(ccfd_gpu_ctx* c, ImplicitState* s, size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
  hipMemsetAsync(p, 0, bytes ? bytes : 16, c->st);
  s->allocs.push_back(p);
  return p;
}

void ccfd::implicit_free(ccfd_gpu_ctx* c) {
  if (!c->imp) return;
  for (void* p : c->imp->allocs) hipFree(p);
  for (hipEvent_t e : c->imp->ev) hipEventDestroy(e);
  if (c->imp->hflag) hipHostFree(c->imp->hflag);
  delete c->imp;
  c->imp = nullptr;
}

// ------------------------------------------------------------------ kernels

#define FOR_CELLS int i = blockIdx.x * blockDim.x + threadIdx.x; if (i >= N) return;

// U <- X, prim <- consPrim(X)
template <typename T>
__global__ void k_load_state(int N, int Np, const double* __restrict__ X,
                             double* __restrict__ U, T* __restrict__ prim,
                             Eq<double> q) {
  FOR_CELLS
  double c[4], p[4];
  for (int v = 0; v < 4; ++v) {
    c[v] = X[v * Np + i];
    U[v * Np + i] = c[v];
  }
  cons_prim<double>(q, c, p);
  for (int v = 0; v < 4; ++v) prim[v * Np + i] = T(p[v]);
}

// F = U - Q - alpha*dt*ut (src/timeDiscretization.c:441-465, 524-537);
// optionally also XK = U, RXK = ut
// alpha*dt of element i: the global step, or alpha times the element's own
// pseudo-time step in steady solves
__device__ inline double elem_adt(double adt, const double* dtv, double alpha,
                                  int i) {
  return dtv ? alpha * dtv[i] : adt;
}

template <typename T>
__global__ void k_newton_residual(int N, int Np, const double* __restrict__ U,
                                  const double* __restrict__ Q,
                                  const T* __restrict__ ut, double adt,
                                  const double* __restrict__ dtv, double alpha,
                                  const double* __restrict__ area,
                                  double* __restrict__ F, double* __restrict__ XK,
                                  double* __restrict__ RXK) {
  FOR_CELLS
  const double a = elem_adt(adt, dtv, alpha, i);
  const double w = dtv ? sqrt(area[i]) : 1.0;
  for (int v = 0; v < 4; ++v) {
    int k = v * Np + i;
    double u = double(ut[k]);
    // local steps: rows scaled as in k_jv_diff
    F[k] = dtv ? w * ((U[k] - Q[k]) / a - u) : U[k] - Q[k] - a * u;
    if (XK) XK[k] = U[k];
    if (RXK) RXK[k] = u;
  }
}

template <typename A, typename B>
__global__ void k_copy_cvt(long n, const A* __restrict__ a, B* __restrict__ b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = B(a[i]);
}

// out = scale * a (scale read from the device: sc[0] / sc[1] or host value)
template <typename K>
__global__ void k_scale(int N, int Np, const double* __restrict__ a,
                        K* __restrict__ out, double s) {
  FOR_CELLS
  for (int v = 0; v < 4; ++v) out[v * Np + i] = K(a[v * Np + i] / s);
}

template <typename K>
__global__ void k_neg_div(int N, int Np, const double* __restrict__ a,
                          K* __restrict__ out, double s) {
  FOR_CELLS
  for (int v = 0; v < 4; ++v) out[v * Np + i] = K(-a[v * Np + i] / s);
}

// ---------------------------------------------- device-resident Arnoldi
//
// One GMRES_M iteration (src/linearSolver.c:422-468) runs without a host
// round trip: the Hessenberg matrix, the Givens rotations, the residual
// estimate and the back-solve live in device memory (GmresDev), every
// reduction is finished on the device, and the host only learns at the
// end (or, lagging behind, while it queues further iterations) whether
// the abort criterion was met. Kernels of an iteration queued after that
// do nothing (they read the done flag first).
//
// Dot products use a fixed grid of kRedBlocks blocks striding over the
// elements, so the partial sums of every product come in the same fixed
// order; the kernels that need a product reduce its partials themselves,
// in the same order in every block.

constexpr int kRedBlocks = 512;   // blocks of the fixed reduction grid

struct GmresDev {
  // H(i, j) = H[i * ld + j], ld = nK + 1
  double* H;
  double* cs;     // [nK] Givens cosines
  double* sn;     // [nK] sines
  double* gm;     // [nK + 1] rotated right-hand side
  double* alp;    // [nK + 1] back-solve coefficients
  double* sc;     // scalars, see below
  int ld;
};
// scalars of GmresDev::sc
enum { GS_DONE = 0, GS_MSTAR, GS_ABORT, GS_NORMB, GS_HNEXT, GS_REL, GS_VV,
       GS_N };

__device__ inline bool gm_done(const double* sc) { return sc[GS_DONE] != 0.0; }

// matrixVector perturbation: epsFD = sqrt(eps) / sqrt(vv); U = XK + epsFD v;
// prim = consPrim(U) (src/linearSolver.c:346-360)
// (nothing once the GMRES cycle is done: sc[GS_DONE], sc may be null)
template <typename T, typename K>
__global__ void k_jv_perturb(int N, int Np, const double* __restrict__ vv,
                             double sqeps, const double* __restrict__ XK,
                             const K* __restrict__ v, double* __restrict__ U,
                             T* __restrict__ prim, double* epsOut, Eq<double> q,
                             const double* sc) {
  FOR_CELLS
  if (sc && gm_done(sc)) return;
  const double eps = sqeps / sqrt(*vv);
  if (i == 0) *epsOut = eps;
  double c[4], p[4];
  for (int k = 0; k < 4; ++k) {
    c[k] = XK[k * Np + i] + eps * double(v[k * Np + i]);
    U[k * Np + i] = c[k];
  }
  cons_prim<double>(q, c, p);
  for (int k = 0; k < 4; ++k) prim[k * Np + i] = T(p[k]);
}

// res = v - alpha*dt*(ut - RXK)/epsFD (src/linearSolver.c:364-372). With
// local steps (dtv) every row is divided by its alpha*dt_i and multiplied
// by sqrt(area_i): sqrt(A_i) (v/(alpha dt_i) - (ut - RXK)/epsFD). The
// 2-norm GMRES minimises is then the area-weighted norm of dU/dt that
// globalResidual measures, not one dominated by the elements with the
// largest steps or the smallest areas.
template <typename T, typename K>
__global__ void k_jv_diff(int N, int Np, const K* __restrict__ v,
                          const T* __restrict__ ut, const double* __restrict__ RXK,
                          const double* __restrict__ epsp, double adt,
                          const double* __restrict__ dtv, double alpha,
                          const double* __restrict__ area,
                          double* __restrict__ res, const double* sc) {
  FOR_CELLS
  if (sc && gm_done(sc)) return;
  const double eps = *epsp;
  const double a = elem_adt(adt, dtv, alpha, i);
  const double w = dtv ? sqrt(area[i]) : 1.0;
  for (int k = 0; k < 4; ++k) {
    int o = k * Np + i;
    res[o] = dtv ? w * (double(v[o]) / a - (double(ut[o]) - RXK[o]) / eps)
                 : double(v[o]) - a * (double(ut[o]) - RXK[o]) / eps;
  }
}


// sum of nr partials, the same fixed-order tree in every block; the result
// is returned to every thread (red: >= 256 doubles of shared memory)
__device__ inline double block_sum_partials(const double* __restrict__ part,
                                            int nr, double* red) {
  double x = 0.0;
  for (int k = threadIdx.x; k < nr; k += blockDim.x) x += part[k];
  red[threadIdx.x] = x;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// block partial of sum x (256 threads) into part[blockIdx.x]
__device__ inline void block_partial(double x, double* red, double* part) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// start of a cycle: H, gm and the flags cleared, gm[0] = beta, the abort
// criterion epsGMRES |B| (src/linearSolver.c:388)
__global__ void k_gm_init(GmresDev g, int nK, double beta, double abortCrit,
                          double normB) {
  for (int k = threadIdx.x; k < (nK + 1) * g.ld; k += blockDim.x) g.H[k] = 0.0;
  for (int k = threadIdx.x; k <= nK; k += blockDim.x) {
    g.gm[k] = k == 0 ? beta : 0.0;
    g.alp[k] = 0.0;
  }
  if (threadIdx.x == 0) {
    g.sc[GS_DONE] = 0.0;
    g.sc[GS_MSTAR] = -1.0;
    g.sc[GS_ABORT] = abortCrit;
    g.sc[GS_NORMB] = normB;
    g.sc[GS_REL] = 0.0;
  }
}

// partials of |v|^2 (FP64 products of the basis values) for the Frechet
// step of matrixVector (src/linearSolver.c:346-347)
template <typename K>
__global__ __launch_bounds__(256) void k_norm_partial(int N, int Np,
                                                      const K* __restrict__ v,
                                                      const double* sc,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  if (gm_done(sc)) return;
  double x = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
    for (int k = 0; k < 4; ++k) {
      const double a = double(v[k * Np + i]);
      x += a * a;
    }
  block_partial(x, red, part);
}

// sum of the partials into sc[GS_VV]
__global__ __launch_bounds__(256) void k_norm_final(const double* __restrict__ part,
                                                    int nr, double* sc) {
  __shared__ double red[256];
  if (gm_done(sc)) return;
  const double s = block_sum_partials(part, nr, red);
  if (threadIdx.x == 0) sc[GS_VV] = s;
}

// first Gram-Schmidt dot: partials of V_0 . W
template <typename K>
__global__ __launch_bounds__(256) void k_mgs_first(int N, int Np,
                                                   const K* __restrict__ V0,
                                                   const double* __restrict__ W,
                                                   const double* sc,
                                                   double* __restrict__ part) {
  __shared__ double red[4];
  if (gm_done(sc)) return;
  double x = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
    for (int k = 0; k < 4; ++k)
      x += double(V0[k * Np + i]) * W[k * Np + i];
  block_partial(x, red, part);
}

// modified Gram-Schmidt step n of iteration m (src/linearSolver.c:440-450):
// h = <V_n, W> from the partials of the previous launch, H(n, m) = h,
// W -= h V_n, then the partials of the next product: <V_{n+1}, W>, or
// <W, W> after the last step (n = m)
template <typename K>
__global__ __launch_bounds__(256) void k_mgs_step(
    int N, int Np, int n, int m, const K* __restrict__ V, size_t vs,
    double* __restrict__ W, const double* __restrict__ partIn,
    double* __restrict__ partOut, GmresDev g) {
  __shared__ double red[256];
  if (gm_done(g.sc)) return;
  const double h = block_sum_partials(partIn, gridDim.x, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) g.H[n * g.ld + m] = h;
  const K* Vn = V + n * vs;
  const K* Vx = n < m ? V + (n + 1) * vs : nullptr;
  double x = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
    for (int k = 0; k < 4; ++k) {
      const int o = k * Np + i;
      const double w = W[o] - h * double(Vn[o]);
      W[o] = w;
      x += Vx ? double(Vx[o]) * w : w * w;
    }
  block_partial(x, red, partOut);
}

// H(m+1, m) = |W|, the previous rotations applied to column m, a new
// rotation, the rotated right-hand side (src/linearSolver.c:451-468); at the
// abort criterion or the last iteration the back-solve for alp
// (src/linearSolver.c:470-482) and the done flag. One block of 256.
__global__ __launch_bounds__(256) void k_givens(GmresDev g, int m, int nK,
                                                const double* __restrict__ part,
                                                int nr, int* hflag) {
  __shared__ double red[256];
  if (gm_done(g.sc)) return;
  const double ww = block_sum_partials(part, nr, red);
  if (threadIdx.x != 0) return;
  double* H = g.H;
  const int ld = g.ld;
#define HM(i, j) H[(i) * ld + (j)]
  const double hnext = sqrt(ww);
  HM(m + 1, m) = hnext;
  g.sc[GS_HNEXT] = hnext;
  for (int n = 0; n <= m - 1; ++n) {
    const double tmp = g.cs[n] * HM(n, m) + g.sn[n] * HM(n + 1, m);
    HM(n + 1, m) = -g.sn[n] * HM(n, m) + g.cs[n] * HM(n + 1, m);
    HM(n, m) = tmp;
  }
  const double bet = sqrt(HM(m, m) * HM(m, m) + HM(m + 1, m) * HM(m + 1, m));
  g.sn[m] = HM(m + 1, m) / bet;
  g.cs[m] = HM(m, m) / bet;
  HM(m, m) = bet;
  g.gm[m + 1] = -g.sn[m] * g.gm[m];
  g.gm[m] = g.cs[m] * g.gm[m];
  const bool conv = fabs(g.gm[m + 1]) <= g.sc[GS_ABORT];
  if (conv || m == nK - 1) {
    for (int n = m; n >= 0; --n) {
      double a = g.gm[n];
      for (int o = n + 1; o <= m; ++o) a -= HM(n, o) * g.alp[o];
      g.alp[n] = a / HM(n, n);
    }
    const double nb = g.sc[GS_NORMB];
    g.sc[GS_REL] = nb > 0.0 ? fabs(g.gm[m + 1]) / nb : 0.0;
    g.sc[GS_MSTAR] = m;
    g.sc[GS_DONE] = conv ? 2.0 : 1.0;
  }
  // the host polls this (pinned) word
  if (hflag)
    __hip_atomic_store(hflag + m, conv || m == nK - 1 ? 1 : 0,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
#undef HM
}

// classical Gram-Schmidt pass over V_0..V_m (steady solves, run twice:
// CGS2): partials of <V_n, W> for every n, part[n * gridDim.x + block]
template <typename K>
__global__ __launch_bounds__(256) void k_cgs_dots(int N, int Np, int m,
                                                  const K* __restrict__ V,
                                                  size_t vs,
                                                  const double* __restrict__ W,
                                                  const double* sc,
                                                  double* __restrict__ part) {
  __shared__ double red[4];
  if (gm_done(sc)) return;
  for (int n = 0; n <= m; ++n) {
    const K* Vn = V + n * vs;
    double x = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
      for (int k = 0; k < 4; ++k) x += double(Vn[k * Np + i]) * W[k * Np + i];
    block_partial(x, red, part + (size_t)n * gridDim.x);
    __syncthreads();
  }
}

// h_n = sum of the partials of <V_n, W> (block n of m + 1); the first pass
// stores H(n, m) = h_n, the second adds its correction
__global__ __launch_bounds__(256) void k_cgs_sum(int m, int nr,
                                                 const double* __restrict__ part,
                                                 double* __restrict__ h,
                                                 GmresDev g, int second) {
  __shared__ double red[256];
  if (gm_done(g.sc)) return;
  const int n = blockIdx.x;
  const double x = block_sum_partials(part + (size_t)n * nr, nr, red);
  if (threadIdx.x == 0) {
    h[n] = x;
    if (second) g.H[n * g.ld + m] += x; else g.H[n * g.ld + m] = x;
  }
}

// W -= sum_n h_n V_n; with wwPart, the partials of |W|^2 afterwards
template <typename K>
__global__ __launch_bounds__(256) void k_cgs_update(
    int N, int Np, int m, const K* __restrict__ V, size_t vs,
    const double* __restrict__ h, double* __restrict__ W, const double* sc,
    double* __restrict__ wwPart) {
  __shared__ double red[4];
  if (gm_done(sc)) return;
  double x = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256)
    for (int k = 0; k < 4; ++k) {
      const int o = k * Np + i;
      double w = W[o];
      for (int n = 0; n <= m; ++n) w -= h[n] * double(V[n * vs + o]);
      W[o] = w;
      x += w * w;
    }
  if (wwPart) block_partial(x, red, wwPart);
}

// V_{m+1} = W / H(m+1, m) (src/linearSolver.c:492-494)
template <typename K>
__global__ void k_scale_dev(int N, int Np, const double* __restrict__ W,
                            K* __restrict__ out, const double* sc) {
  FOR_CELLS
  if (gm_done(sc)) return;
  const double h = sc[GS_HNEXT];
  for (int v = 0; v < 4; ++v) out[v * Np + i] = K(W[v * Np + i] / h);
}

// dX += sum_{n <= m*} alp[n] Z_n (src/linearSolver.c:483-489), m* and alp
// from the device
template <typename K>
__global__ void k_update_dx(int N, int Np, const double* __restrict__ sc,
                            const double* __restrict__ alp,
                            const K* __restrict__ Z, size_t zstride,
                            double* __restrict__ dX) {
  FOR_CELLS
  const int m = (int)sc[GS_MSTAR];
  for (int v = 0; v < 4; ++v) {
    double acc = dX[v * Np + i];
    for (int n = 0; n <= m; ++n)
      acc += alp[n] * double(Z[n * zstride + v * Np + i]);
    dX[v * Np + i] = acc;
  }
}

// x += w d, w from the device when wdev is set
__global__ void k_axpy(int N, int Np, double* __restrict__ x,
                       const double* __restrict__ d, double w,
                       const double* wdev = nullptr) {
  FOR_CELLS
  if (wdev) w = *wdev;
  if (w == 1.0)
    for (int v = 0; v < 4; ++v) x[v * Np + i] += d[v * Np + i];
  else
    for (int v = 0; v < 4; ++v) x[v * Np + i] += w * d[v * Np + i];
}

// largest relative change of density and pressure that the update d makes
// to the conservative state x, per block (non-finite changes count as
// 1e300)
__global__ __launch_bounds__(256) void k_relchange_partial(
    int N, int Np, const double* __restrict__ x, const double* __restrict__ d,
    double gam1, double* __restrict__ part) {
  __shared__ double sm[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  double r = 0.0;
  if (i < N) {
    const double r0 = x[i], mx = x[Np + i], my = x[2 * Np + i];
    const double p0 = gam1 * (x[3 * Np + i] - 0.5 * (mx * mx + my * my) / r0);
    const double r1 = r0 + d[i], nx = mx + d[Np + i], ny = my + d[2 * Np + i];
    const double p1 =
        gam1 * (x[3 * Np + i] + d[3 * Np + i] - 0.5 * (nx * nx + ny * ny) / r1);
    r = fmax(fabs(r1 - r0) / fabs(r0), fabs(p1 - p0) / fabs(p0));
    if (!(r <= 1e300)) r = 1e300;
  }
  sm[threadIdx.x] = r;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// out = max of the partials; with wout, the update weight of a bounded
// Newton update, w = maxChange / max where the max exceeds maxChange (else
// 1), and wmin = min(wmin, w)
__global__ void k_max_final(int nb, const double* __restrict__ part,
                            double* out, double maxChange = 0.0,
                            double* wout = nullptr, double* wmin = nullptr) {
  __shared__ double sm[256];
  double r = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) r = fmax(r, part[k]);
  sm[threadIdx.x] = r;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out = sm[0];
    if (wout) {
      const double w = sm[0] > maxChange ? maxChange / sm[0] : 1.0;
      *wout = w;
      *wmin = fmin(*wmin, w);
    }
  }
}

__global__ void k_set1(double* x, double v) { *x = v; }

// ---- LU-SGS assembly (coloured finite differences) ----

// prim = base, with element i of colour c perturbed in variable iv
template <typename T>
__global__ void k_fd_perturb(int N, int Np, const double* __restrict__ base,
                             const int* __restrict__ color, int c, int iv,
                             double reps, T* __restrict__ prim) {
  FOR_CELLS
  for (int v = 0; v < 4; ++v) {
    double p = base[v * Np + i];
    if (v == iv && color[i] == c) p += reps;
    prim[v * Np + i] = T(p);
  }
}

// prim = consPrim(XK) with the conservative variable iv of the elements of
// colour c perturbed (Jacobian with respect to the conservative state, which
// the LU-SGS sweeps apply it to; the multicolour preconditioner uses it)
template <typename T>
__global__ void k_fd_perturb_cons(int N, int Np, const double* __restrict__ XK,
                                  const int* __restrict__ color, int c, int iv,
                                  double reps, T* __restrict__ prim,
                                  Eq<double> q) {
  FOR_CELLS
  double u[4], p[4];
  for (int v = 0; v < 4; ++v) {
    u[v] = XK[v * Np + i];
    if (v == iv && color[i] == c) u[v] += reps;
  }
  cons_prim<double>(q, u, p);
  for (int v = 0; v < 4; ++v) prim[v * Np + i] = T(p[v]);
}

// record (ut_j - RXK_j) / sqrt(eps) into the diagonal block (j perturbed) or
// the block of the side of j facing the perturbed neighbour
// (src/linearSolver.c:203-222)
template <typename T, typename M>
__global__ void k_fd_record(int N, int Np, int S, const int* __restrict__ link,
                            const int* __restrict__ color, int c, int iv,
                            const T* __restrict__ ut,
                            const double* __restrict__ RXK, double srEps,
                            M* __restrict__ Dg, M* __restrict__ Blk) {
  FOR_CELLS
  double d[4];
  for (int v = 0; v < 4; ++v)
    d[v] = (double(ut[v * Np + i]) - RXK[v * Np + i]) * srEps;
  if (color[i] == c) {
    for (int jv = 0; jv < 4; ++jv)
      Dg[(4 * jv + iv) * Np + i] += M(d[jv]);
  }
  for (int s = 0; s < S; ++s) {
    int lk = link[s * Np + i];
    if (lk < 0) continue;   // ghost or no side
    int j = lk >> 2;
    if (color[j] != c || j == i) continue;
    for (int jv = 0; jv < 4; ++jv)
      Blk[((size_t)(4 * jv + iv) * S + s) * Np + i] += M(d[jv]);
  }
}

// determinant and inverse of a 4x4 block by cofactors (calcDinv,
// src/linearSolver.c:136-179); returns the determinant, Ainv only when it
// is not below DBL_EPSILON in magnitude
__device__ inline double inv4(const double A[4][4], double Ainv[4][4]) {
  double det = A[0][0]*(A[1][1]*(A[2][2]*A[3][3]-A[2][3]*A[3][2])
    +A[1][2]*(A[2][3]*A[3][1]-A[2][1]*A[3][3])+A[1][3]*(A[2][1]
    *A[3][2]-A[2][2]*A[3][1]))-A[0][1]*(A[1][0]*(A[2][2]*A[3][3]
    -A[2][3]*A[3][2])+A[1][2]*(A[2][3]*A[3][0]-A[2][0]*A[3][3])
    +A[1][3]*(A[2][0]*A[3][2]-A[2][2]*A[3][0]))+A[0][2]*(A[1][0]
    *(A[2][1]*A[3][3]-A[2][3]*A[3][1])+A[1][1]*(A[2][3]*A[3][0]
    -A[2][0]*A[3][3])+A[1][3]*(A[2][0]*A[3][1]-A[2][1]*A[3][0]))
    -A[0][3]*(A[1][0]*(A[2][1]*A[3][2]-A[2][2]*A[3][1])+A[1][1]
    *(A[2][2]*A[3][0]-A[2][0]*A[3][2])+A[1][2]*(A[2][0]*A[3][1]
    -A[2][1]*A[3][0]));
  if (fabs(det) <= DBL_EPSILON) return det;
  double C[4][4];
  C[0][0] = A[1][1]*(A[2][2]*A[3][3]-A[2][3]*A[3][2])+A[1][2]*(A[2][3]*A[3][1]-A[2][1]*A[3][3])+A[1][3]*(A[2][1]*A[3][2]-A[2][2]*A[3][1]);
  C[0][1] = A[1][0]*(A[2][3]*A[3][2]-A[2][2]*A[3][3])+A[1][2]*(A[2][0]*A[3][3]-A[2][3]*A[3][0])+A[1][3]*(A[2][2]*A[3][0]-A[2][0]*A[3][2]);
  C[0][2] = A[1][0]*(A[2][1]*A[3][3]-A[2][3]*A[3][1])+A[1][1]*(A[2][3]*A[3][0]-A[2][0]*A[3][3])+A[1][3]*(A[2][0]*A[3][1]-A[2][1]*A[3][0]);
  C[0][3] = A[1][0]*(A[2][2]*A[3][1]-A[2][1]*A[3][2])+A[1][1]*(A[2][0]*A[3][2]-A[2][2]*A[3][0])+A[1][2]*(A[2][1]*A[3][0]-A[2][0]*A[3][1]);
  C[1][0] = A[0][1]*(A[2][3]*A[3][2]-A[2][2]*A[3][3])+A[0][2]*(A[2][1]*A[3][3]-A[2][3]*A[3][1])+A[0][3]*(A[2][2]*A[3][1]-A[2][1]*A[3][2]);
  C[1][1] = A[0][0]*(A[2][2]*A[3][3]-A[2][3]*A[3][2])+A[0][2]*(A[2][3]*A[3][0]-A[2][0]*A[3][3])+A[0][3]*(A[2][0]*A[3][2]-A[2][2]*A[3][0]);
  C[1][2] = A[0][0]*(A[2][3]*A[3][1]-A[2][1]*A[3][3])+A[0][1]*(A[2][0]*A[3][3]-A[2][3]*A[3][0])+A[0][3]*(A[2][1]*A[3][0]-A[2][0]*A[3][1]);
  C[1][3] = A[0][0]*(A[2][1]*A[3][2]-A[2][2]*A[3][1])+A[0][1]*(A[2][2]*A[3][0]-A[2][0]*A[3][2])+A[0][2]*(A[2][0]*A[3][1]-A[2][1]*A[3][0]);
  C[2][0] = A[0][1]*(A[1][2]*A[3][3]-A[1][3]*A[3][2])+A[0][2]*(A[1][3]*A[3][1]-A[1][1]*A[3][3])+A[0][3]*(A[1][1]*A[3][2]-A[1][2]*A[3][1]);
  C[2][1] = A[0][0]*(A[1][3]*A[3][2]-A[1][2]*A[3][3])+A[0][2]*(A[1][0]*A[3][3]-A[1][3]*A[3][0])+A[0][3]*(A[1][2]*A[3][0]-A[1][0]*A[3][2]);
  C[2][2] = A[0][0]*(A[1][1]*A[3][3]-A[1][3]*A[3][1])+A[0][1]*(A[1][3]*A[3][0]-A[1][0]*A[3][3])+A[0][3]*(A[1][0]*A[3][1]-A[1][1]*A[3][0]);
  C[2][3] = A[0][0]*(A[1][2]*A[3][1]-A[1][1]*A[3][2])+A[0][1]*(A[1][0]*A[3][2]-A[1][2]*A[3][0])+A[0][2]*(A[1][1]*A[3][0]-A[1][0]*A[3][1]);
  C[3][0] = A[0][1]*(A[1][3]*A[2][2]-A[1][2]*A[2][3])+A[0][2]*(A[1][1]*A[2][3]-A[1][3]*A[2][1])+A[0][3]*(A[1][2]*A[2][1]-A[1][1]*A[2][2]);
  C[3][1] = A[0][0]*(A[1][2]*A[2][3]-A[1][3]*A[2][2])+A[0][2]*(A[1][3]*A[2][0]-A[1][0]*A[2][3])+A[0][3]*(A[1][0]*A[2][2]-A[1][2]*A[2][0]);
  C[3][2] = A[0][0]*(A[1][3]*A[2][1]-A[1][1]*A[2][3])+A[0][1]*(A[1][0]*A[2][3]-A[1][3]*A[2][0])+A[0][3]*(A[1][1]*A[2][0]-A[1][0]*A[2][1]);
  C[3][3] = A[0][0]*(A[1][1]*A[2][2]-A[1][2]*A[2][1])+A[0][1]*(A[1][2]*A[2][0]-A[1][0]*A[2][2])+A[0][2]*(A[1][0]*A[2][1]-A[1][1]*A[2][0]);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) Ainv[a][b] = C[b][a] / det;
  return det;
}

// D = I - dt J_ii, neighbour blocks -dt J_ij, Dinv = calcDinv(D)
// (src/linearSolver.c:226-251, 136-179); D itself is kept in Dg (the
// multigrid residuals and coarse blocks use it)
template <typename M>
__global__ void k_fd_finish(int N, int Np, int S, const int* __restrict__ link,
                            double dt0, const double* __restrict__ dtv,
                            const double* __restrict__ area,
                            M* __restrict__ Dg, M* __restrict__ Blk,
                            M* __restrict__ Dinv, int* flag) {
  FOR_CELLS
  // local steps (dtv): rows scaled as in k_jv_diff,
  // D = sqrt(A_i) (I/dt_i - J_ii), blocks -sqrt(A_i) J_ij
  const double w = dtv ? sqrt(area[i]) : 1.0;
  const double dt = dtv ? w : dt0;
  const double diag = dtv ? w / dtv[i] : 1.0;
  double A[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double x = double(Dg[(4 * a + b) * Np + i]) * -dt;
      if (a == b) x += diag;
      A[a][b] = x;
    }
  for (int s = 0; s < S; ++s) {
    if (link[s * Np + i] == kNoSide) continue;
    for (int k = 0; k < 16; ++k) {
      size_t o = ((size_t)k * S + s) * Np + i;
      Blk[o] = M(double(Blk[o]) * -dt);
    }
  }
  double E[4][4];
  const double det = inv4(A, E);
  if (fabs(det) <= DBL_EPSILON) {
    *flag = 1;
    return;
  }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      Dinv[(4 * a + b) * Np + i] = M(E[a][b]);
      Dg[(4 * a + b) * Np + i] = M(A[a][b]);
    }
}

// LUSGS forward sweep over the cells of one level (src/linearSolver.c:271-301)
template <typename M, typename K>
__global__ void k_lusgs_fwd(int n, const int* __restrict__ cells, int Np, int S,
                            const int* __restrict__ link,
                            const int* __restrict__ refIdx,
                            const M* __restrict__ Blk, const M* __restrict__ Dinv,
                            const K* __restrict__ b, M* __restrict__ xs) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = cells[t];
  const int ri = refIdx[i];
  M t1[4] = {M(0), M(0), M(0), M(0)};
  for (int s = 0; s < S; ++s) {
    const int lk = link[s * Np + i];
    if (lk < 0) continue;
    const int j = lk >> 2;
    if (refIdx[j] >= ri) continue;
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        t1[a] += Blk[((size_t)(4 * a + c) * S + s) * Np + i] * xs[c * Np + j];
  }
  for (int a = 0; a < 4; ++a) {
    M t2 = M(0);
    for (int c = 0; c < 4; ++c)
      t2 += Dinv[(4 * a + c) * Np + i] * (M(b[c * Np + i]) - t1[c]);
    xs[a * Np + i] = t2;
  }
}

// LUSGS backward sweep (src/linearSolver.c:303-332)
template <typename M, typename K>
__global__ void k_lusgs_bwd(int n, const int* __restrict__ cells, int Np, int S,
                            const int* __restrict__ link,
                            const int* __restrict__ refIdx,
                            const M* __restrict__ Blk, const M* __restrict__ Dinv,
                            const M* __restrict__ xs, K* __restrict__ x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = cells[t];
  const int ri = refIdx[i];
  M t1[4] = {M(0), M(0), M(0), M(0)};
  for (int s = 0; s < S; ++s) {
    const int lk = link[s * Np + i];
    if (lk < 0) continue;
    const int j = lk >> 2;
    if (refIdx[j] <= ri) continue;
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        t1[a] += Blk[((size_t)(4 * a + c) * S + s) * Np + i] * M(x[c * Np + j]);
  }
  for (int a = 0; a < 4; ++a) {
    M t2 = M(0);
    for (int c = 0; c < 4; ++c) t2 += Dinv[(4 * a + c) * Np + i] * t1[c];
    x[a * Np + i] = K(xs[a * Np + i] - t2);
  }
}

// one Gauss-Seidel update of the elements of one colour, rows p = start ..
// start + n - 1 of the colour-ordered operands (element i = cells[p]):
// x_i = D_i^-1 (b_i - sum over the face neighbours j of B_ij x_j), with the
// current values of x (elements of one colour are not neighbours)
template <typename M, typename K>
__global__ __launch_bounds__(128) void k_gs_cells(
    int n, int start, const int* __restrict__ cells, int N, int Np, int S,
    const int* __restrict__ nbr, const M* __restrict__ Blk,
    const M* __restrict__ Dinv, const K* __restrict__ b, K* __restrict__ x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int p = start + t, i = cells[p];
  double r[4];
  for (int v = 0; v < 4; ++v) r[v] = double(b[v * Np + i]);
  for (int s = 0; s < S; ++s) {
    const int j = nbr[(size_t)s * N + p];
    if (j < 0) continue;
    double xj[4];
    for (int v = 0; v < 4; ++v) xj[v] = double(x[v * Np + j]);
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        r[a] -= double(Blk[((size_t)(4 * a + c) * S + s) * N + p]) * xj[c];
  }
  for (int a = 0; a < 4; ++a) {
    double u = 0.0;
    for (int c = 0; c < 4; ++c) u += double(Dinv[(size_t)(4 * a + c) * N + p]) * r[c];
    x[a * Np + i] = K(u);
  }
}

// The same update with four threads per element, one per row a of the
// 4 x 4 system, for FP32 blocks: row a of b_i - sum B_ij x_j accumulates in
// the order of k_gs_cells (neighbours in slot order, columns in order) and
// the rows meet in LDS for the D_i^-1 product, so both kernels give the same
// bits. With FP32 blocks the one-thread kernel waits on its gathers (13.0
// against 8.9 us per colour on the 512 x 256 O-grid); with FP64 blocks the
// four lanes' repeated index and neighbour loads cost more than they hide
// (16.4 against 14.5 us; two lanes of two rows each: 14.7 us, the blocks'
// 512 B per element bound it), so FP64 keeps k_gs_cells.
//
// L lanes per element (2 or 4), each with 4 / L rows; 256 threads
template <typename M, typename K, int L>
__global__ __launch_bounds__(256) void k_gs_rows(
    int n, int start, const int* __restrict__ cells, int N, int Np, int S,
    const int* __restrict__ nbr, const M* __restrict__ Blk,
    const M* __restrict__ Dinv, const K* __restrict__ b, K* __restrict__ x) {
  constexpr int R = 4 / L, C = 256 / L;   // rows per lane, elements per block
  __shared__ double rs[4 * C];
  const int tid = threadIdx.x, lane = tid % L, e = tid / L;
  const int t = blockIdx.x * C + e;
  const bool active = t < n;
  int p = 0, i = 0;
  double r[R];
  for (int q = 0; q < R; ++q) r[q] = 0.0;
  if (active) {
    p = start + t;
    i = cells[p];
    for (int q = 0; q < R; ++q) r[q] = double(b[(lane * R + q) * Np + i]);
    for (int s = 0; s < S; ++s) {
      const int j = nbr[(size_t)s * N + p];
      if (j < 0) continue;
      double xj[4];
      for (int c = 0; c < 4; ++c) xj[c] = double(x[c * Np + j]);
      for (int q = 0; q < R; ++q) {
        const int a = lane * R + q;
        for (int c = 0; c < 4; ++c)
          r[q] -= double(Blk[((size_t)(4 * a + c) * S + s) * N + p]) * xj[c];
      }
    }
  }
  for (int q = 0; q < R; ++q) rs[4 * e + lane * R + q] = r[q];
  __syncthreads();
  if (active) {
    const double* rr = rs + 4 * e;
    for (int q = 0; q < R; ++q) {
      const int a = lane * R + q;
      double u = 0.0;
      for (int c = 0; c < 4; ++c) u += double(Dinv[(size_t)(4 * a + c) * N + p]) * rr[c];
      x[a * Np + i] = K(u);
    }
  }
}

// colour-ordered copies of the side blocks and inverted diagonal blocks
template <typename M>
__global__ void k_gs_order(int N, int Np, int S, const int* __restrict__ cells,
                           const M* __restrict__ Blk, const M* __restrict__ Dinv,
                           M* __restrict__ Bc, M* __restrict__ Dc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int i = cells[p];
  for (int k = 0; k < 16; ++k) {
    for (int s = 0; s < S; ++s)
      Bc[((size_t)k * S + s) * N + p] = Blk[((size_t)k * S + s) * Np + i];
    Dc[(size_t)k * N + p] = Dinv[(size_t)k * Np + i];
  }
}

// ---- aggregation multigrid (precond = 3) ----

// coarse diagonal blocks: sums of the listed blocks of the level above
// (code & 1: a diagonal block, else an off-diagonal one; index code >> 1),
// then inverted
template <typename M>
__global__ void k_mg_diag(int n, const int64_t* __restrict__ start,
                          const int64_t* __restrict__ src,
                          const M* __restrict__ Dp, long dStride,
                          const M* __restrict__ Op, long oStride,
                          double* __restrict__ D, double* __restrict__ Dinv,
                          int* flag) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double A[4][4] = {};
  for (int64_t t = start[c]; t < start[c + 1]; ++t) {
    const int64_t code = src[t], idx = code >> 1;
    for (int k = 0; k < 16; ++k)
      A[k >> 2][k & 3] += (code & 1) ? double(Dp[k * dStride + idx])
                                     : double(Op[k * oStride + idx]);
  }
  double E[4][4];
  if (fabs(inv4(A, E)) <= DBL_EPSILON) {
    *flag = 1;
    return;
  }
  for (int k = 0; k < 16; ++k) {
    D[(size_t)k * n + c] = A[k >> 2][k & 3];
    Dinv[(size_t)k * n + c] = E[k >> 2][k & 3];
  }
}

// coarse off-diagonal blocks
template <typename M>
__global__ void k_mg_off(int nnz, const int64_t* __restrict__ start,
                         const int64_t* __restrict__ src,
                         const M* __restrict__ Dp, long dStride,
                         const M* __restrict__ Op, long oStride,
                         double* __restrict__ O) {
  const int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= nnz) return;
  double a[16] = {};
  for (int64_t t = start[z]; t < start[z + 1]; ++t) {
    const int64_t code = src[t], idx = code >> 1;
    for (int k = 0; k < 16; ++k)
      a[k] += (code & 1) ? double(Dp[k * dStride + idx])
                         : double(Op[k * oStride + idx]);
  }
  for (int k = 0; k < 16; ++k) O[(size_t)k * nnz + z] = a[k];
}

// Gauss-Seidel update of the coarse elements of one colour
__global__ void k_mg_gs(int m, const int* __restrict__ cells, int n, int nnz,
                        const int* __restrict__ rowStart,
                        const int* __restrict__ col,
                        const double* __restrict__ O,
                        const double* __restrict__ Dinv,
                        const double* __restrict__ b, double* __restrict__ x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int i = cells[t];
  double r[4];
  for (int v = 0; v < 4; ++v) r[v] = b[(size_t)v * n + i];
  for (int z = rowStart[i]; z < rowStart[i + 1]; ++z) {
    const int j = col[z];
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        r[a] -= O[(size_t)(4 * a + c) * nnz + z] * x[(size_t)c * n + j];
  }
  for (int a = 0; a < 4; ++a) {
    double u = 0.0;
    for (int c = 0; c < 4; ++c) u += Dinv[(size_t)(4 * a + c) * n + i] * r[c];
    x[(size_t)a * n + i] = u;
  }
}

// r = b - A x on a coarse level
__global__ void k_mg_resid(int n, int nnz, const int* __restrict__ rowStart,
                           const int* __restrict__ col,
                           const double* __restrict__ D,
                           const double* __restrict__ O,
                           const double* __restrict__ b,
                           const double* __restrict__ x,
                           double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double t[4];
  for (int a = 0; a < 4; ++a) {
    t[a] = b[(size_t)a * n + i];
    for (int c = 0; c < 4; ++c)
      t[a] -= D[(size_t)(4 * a + c) * n + i] * x[(size_t)c * n + i];
  }
  for (int z = rowStart[i]; z < rowStart[i + 1]; ++z) {
    const int j = col[z];
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        t[a] -= O[(size_t)(4 * a + c) * nnz + z] * x[(size_t)c * n + j];
  }
  for (int a = 0; a < 4; ++a) r[(size_t)a * n + i] = t[a];
}

// r = b - A x on the elements (diagonal blocks D, side blocks Blk)
template <typename M, typename K>
__global__ void k_mg_resid0(int N, int Np, int S, const int* __restrict__ link,
                            const M* __restrict__ D, const M* __restrict__ Blk,
                            const K* __restrict__ b, const K* __restrict__ x,
                            double* __restrict__ r) {
  FOR_CELLS
  double t[4];
  for (int a = 0; a < 4; ++a) {
    t[a] = double(b[a * Np + i]);
    for (int c = 0; c < 4; ++c)
      t[a] -= double(D[(4 * a + c) * Np + i]) * double(x[c * Np + i]);
  }
  for (int s = 0; s < S; ++s) {
    const int lk = link[s * Np + i];
    if (lk < 0 || (lk >> 2) == i) continue;
    const int j = lk >> 2;
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c)
        t[a] -= double(Blk[((size_t)(4 * a + c) * S + s) * Np + i]) *
                double(x[c * Np + j]);
  }
  for (int a = 0; a < 4; ++a) r[a * Np + i] = t[a];
}

// b_c = sum of r over the members of coarse element c
__global__ void k_mg_restrict(int n, const int* __restrict__ memStart,
                              const int* __restrict__ mem,
                              const double* __restrict__ r, long rStride,
                              double* __restrict__ b) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = memStart[c]; k < memStart[c + 1]; ++k)
    for (int v = 0; v < 4; ++v) t[v] += r[v * rStride + mem[k]];
  for (int v = 0; v < 4; ++v) b[(size_t)v * n + c] = t[v];
}

// x_p += x_c of p's coarse element
template <typename K>
__global__ void k_mg_prolong(int nPrev, const int* __restrict__ aggOf,
                             const double* __restrict__ xc, int n,
                             K* __restrict__ x, long xStride) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nPrev) return;
  const int c = aggOf[p];
  for (int v = 0; v < 4; ++v)
    x[v * xStride + p] = K(double(x[v * xStride + p]) + xc[(size_t)v * n + c]);
}

// local pseudo-time step of every element: the CFL (and, with viscosity,
// diffusive) limit of calcTimeStep (src/timeDiscretization.c:253-308) per
// element instead of its minimum
template <typename T>
__global__ void k_local_dt(int N, int Np, const T* __restrict__ prim,
                           const double* __restrict__ area,
                           const double* __restrict__ sx,
                           const double* __restrict__ sy, double gam, double mu,
                           double gamPrMax, double cfl, double dfl,
                           double* __restrict__ dtv) {
  FOR_CELLS
  const double r = double(prim[i]), u = double(prim[Np + i]);
  const double v = double(prim[2 * Np + i]), p = double(prim[3 * Np + i]);
  const double a = sqrt(gam * p / r);
  double dt = cfl * area[i] / ((fabs(u) + a) * sx[i] + (fabs(v) + a) * sy[i]);
  if (mu > 1e-10) {
    const double ssv = gamPrMax * mu * sx[i] * sx[i] + gamPrMax * mu * sy[i] * sy[i];
    dt = fmin(dt, dfl * r * r * area[i] * area[i] / (4.0 * ssv));
  }
  dtv[i] = dt;
}

// --------------------------------------------------------------- host side

static dim3 cells_grid(const ccfd_gpu_ctx* c) { return dim3((c->N + 255) / 256); }

// doubles of the GmresDev buffer for Krylov dimension nK
static size_t gmres_dev_doubles(int nK) {
  return (size_t)(nK + 1) * (nK + 1) + 2 * (size_t)nK + 2 * (size_t)(nK + 1) +
         GS_N;
}

static GmresDev gmres_view(double* b, int nK) {
  GmresDev g;
  g.ld = nK + 1;
  g.H = b;
  g.cs = g.H + (size_t)(nK + 1) * (nK + 1);
  g.sn = g.cs + nK;
  g.gm = g.sn + nK;
  g.alp = g.gm + nK + 1;
  g.sc = g.alp + nK + 1;
  return g;
}

// blocks of the fixed reduction grid for N elements
static int red_blocks(const ccfd_gpu_ctx* c) {
  return std::max(1, std::min(kRedBlocks, (c->N + 255) / 256));
}

static int ensure(ccfd_gpu_ctx* c, int nK) {
  ImplicitState* s = c->imp;
  int prec = c->prm.precision;
  if (s && s->nK >= nK && s->prec == prec) return CCFD_OK;
  implicit_free(c);
  s = c->imp = new ImplicitState();
  s->nK = nK;
  s->prec = prec;
  size_t e4 = 4 * (size_t)c->Np;
  size_t kb = prec == CCFD_FP64 ? 8 : 4;
  s->XK = (double*)imalloc(c, s, e4 * 8);
  s->RXK = (double*)imalloc(c, s, e4 * 8);
  s->Q = (double*)imalloc(c, s, e4 * 8);
  s->FXK = (double*)imalloc(c, s, e4 * 8);
  s->dX = (double*)imalloc(c, s, e4 * 8);
  s->W = (double*)imalloc(c, s, e4 * 8);
  s->B = (double*)imalloc(c, s, e4 * 8);
  s->V = imalloc(c, s, (nK + 1) * e4 * kb);
  s->Z = imalloc(c, s, (size_t)nK * e4 * kb);
  s->H = (double*)imalloc(c, s, (nK + 8) * 8);
  s->alp = (double*)imalloc(c, s, (nK + 8) * 8);
  s->part = (double*)imalloc(c, s, (c->nb + 64) * 8);
  s->gdev = (double*)imalloc(c, s, gmres_dev_doubles(nK) * 8);
  s->partA = (double*)imalloc(c, s, kRedBlocks * 8);
  s->partB = (double*)imalloc(c, s, kRedBlocks * 8);
  s->cgsPart = (double*)imalloc(c, s, (size_t)(nK + 1) * kRedBlocks * 8);
  s->cgsH = (double*)imalloc(c, s, (nK + 8) * 8);
  if (!s->XK || !s->RXK || !s->Q || !s->FXK || !s->dX || !s->W || !s->B ||
      !s->V || !s->Z || !s->H || !s->alp || !s->part || !s->gdev ||
      !s->partA || !s->partB || !s->cgsPart || !s->cgsH)
    return fail(CCFD_ENOMEM, "implicit solver allocation failed");
  if (hipHostMalloc((void**)&s->hflag, (nK + 8) * sizeof(int),
                    hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
      hipHostGetDevicePointer((void**)&s->hflagDev, s->hflag, 0) != hipSuccess)
    return fail(CCFD_ENOMEM, "pinned GMRES flags allocation failed");
  s->ev.resize(nK);
  for (auto& e : s->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(CCFD_EHIP, "GMRES event creation failed");
  return CCFD_OK;
}

// greedy distance-d colouring of the element graph (host), and the LU-SGS
// level schedules in the reference's element order
static int prepare_lusgs(ccfd_gpu_ctx* c, int order) {
  ImplicitState* s = c->imp;
  if (s->haveColors && s->orderColored == order) return CCFD_OK;
  const int N = c->N, Np = c->Np, S = c->S;
  std::vector<int> link((size_t)S * Np);
  if (hipMemcpyAsync(link.data(), c->link, link.size() * 4,
                     hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "link download failed");
  auto nbrs = [&](int i, std::vector<int>& out) {
    out.clear();
    for (int k = 0; k < S; ++k) {
      int lk = link[(size_t)k * Np + i];
      if (lk >= 0) out.push_back(lk >> 2);
    }
  };
  const int dist = order == 2 ? 3 : 2;
  std::vector<int> color(N, -1), mark(N, -1), frontier, next, nb;
  std::vector<char> used;
  int ncol = 0;
  for (int i = 0; i < N; ++i) {
    // colours within distance `dist` of i
    frontier.assign(1, i);
    mark[i] = i;
    std::vector<int> seen(1, i);
    for (int d = 0; d < dist; ++d) {
      next.clear();
      for (int u : frontier) {
        nbrs(u, nb);
        for (int w : nb)
          if (mark[w] != i) {
            mark[w] = i;
            next.push_back(w);
            seen.push_back(w);
          }
      }
      frontier.swap(next);
    }
    used.assign(ncol + 1, 0);
    for (int w : seen)
      if (w != i && color[w] >= 0) used[color[w]] = 1;
    int cc = 0;
    while (cc < ncol && used[cc]) ++cc;
    color[i] = cc;
    if (cc == ncol) ++ncol;
  }
  // level schedules in reference order
  std::vector<int> ref(N), byRef(N);
  for (int i = 0; i < N; ++i) {
    ref[i] = c->perm[i];
    byRef[c->perm[i]] = i;
  }
  std::vector<int> lf(N, 0), lb(N, 0);
  int nlf = 0, nlb = 0;
  for (int r = 0; r < N; ++r) {
    int i = byRef[r], l = 0;
    nbrs(i, nb);
    for (int j : nb)
      if (ref[j] < r) l = std::max(l, lf[j] + 1);
    lf[i] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int r = N - 1; r >= 0; --r) {
    int i = byRef[r], l = 0;
    nbrs(i, nb);
    for (int j : nb)
      if (ref[j] > r) l = std::max(l, lb[j] + 1);
    lb[i] = l;
    nlb = std::max(nlb, l + 1);
  }
  auto bucket = [&](const std::vector<int>& lv, int nl, std::vector<int>& start,
                    std::vector<int>& items) {
    start.assign(nl + 1, 0);
    for (int i = 0; i < N; ++i) start[lv[i] + 1]++;
    for (int l = 0; l < nl; ++l) start[l + 1] += start[l];
    items.assign(N, 0);
    std::vector<int> pos(start.begin(), start.end() - 1);
    for (int i = 0; i < N; ++i) items[pos[lv[i]]++] = i;
  };
  std::vector<int> itemsF, itemsB;
  bucket(lf, nlf, s->lvlFStart, itemsF);
  bucket(lb, nlb, s->lvlBStart, itemsB);
  std::vector<int> colorP(Np, -1), refP(Np, 0);
  for (int i = 0; i < N; ++i) {
    colorP[i] = color[i];
    refP[i] = ref[i];
  }
  size_t kb = c->prm.precision == CCFD_FP64 ? 8 : 4;
  s->color = (int*)imalloc(c, s, Np * 4);
  s->refIdx = (int*)imalloc(c, s, Np * 4);
  s->lvlF = (int*)imalloc(c, s, N * 4);
  s->lvlB = (int*)imalloc(c, s, N * 4);
  s->Dinv = imalloc(c, s, 16 * (size_t)Np * kb);
  s->Dg = imalloc(c, s, 16 * (size_t)Np * kb);
  s->Blk = imalloc(c, s, 16 * (size_t)S * Np * kb);
  s->primBase = (double*)imalloc(c, s, 4 * (size_t)Np * 8);
  s->xs = (double*)imalloc(c, s, 4 * (size_t)Np * 8);
  s->flag = (int*)imalloc(c, s, 64);
  if (!s->color || !s->refIdx || !s->lvlF || !s->lvlB || !s->Dinv || !s->Dg ||
      !s->Blk || !s->primBase || !s->xs || !s->flag)
    return fail(CCFD_ENOMEM, "LU-SGS allocation failed");
  // on the context stream, after the zero fills imalloc queued there
  if (hipMemcpyAsync(s->color, colorP.data(), Np * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(s->refIdx, refP.data(), Np * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(s->lvlF, itemsF.data(), N * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(s->lvlB, itemsB.data(), N * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "LU-SGS schedule upload failed");
  s->nColors = ncol;
  s->haveColors = true;
  s->orderColored = order;
  return CCFD_OK;
}

// multicolour sweep schedule (precond = 2, 3): greedy distance-1
// colouring; the sweeps run colour by colour
static int prepare_multicolor(ccfd_gpu_ctx* c) {
  ImplicitState* s = c->imp;
  if (s->haveMC) return CCFD_OK;
  const int N = c->N, Np = c->Np, S = c->S;
  std::vector<int> link((size_t)S * Np);
  if (hipMemcpyAsync(link.data(), c->link, link.size() * 4,
                     hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "link download failed");
  // greedy colouring in breadth-first order over the face graph (from the
  // lowest unvisited element, neighbours in side order): on a bipartite
  // mesh such as a structured quadrilateral grid this gives two colours,
  // where element order can give five (the extra small colours cost whole
  // sweep launches)
  std::vector<int> order;
  order.reserve(N);
  {
    std::vector<char> seen(N, 0);
    for (int r = 0; r < N; ++r) {
      if (seen[r]) continue;
      seen[r] = 1;
      size_t head = order.size();
      order.push_back(r);
      while (head < order.size()) {
        const int u = order[head++];
        for (int k = 0; k < S; ++k) {
          const int lk = link[(size_t)k * Np + u];
          if (lk < 0 || seen[lk >> 2]) continue;
          seen[lk >> 2] = 1;
          order.push_back(lk >> 2);
        }
      }
    }
  }
  std::vector<int> col(N, -1);
  std::vector<char> used;
  int ncol = 0;
  for (int i : order) {
    used.assign(ncol + 1, 0);
    for (int k = 0; k < S; ++k) {
      const int lk = link[(size_t)k * Np + i];
      if (lk >= 0 && col[lk >> 2] >= 0) used[col[lk >> 2]] = 1;
    }
    int cc = 0;
    while (cc < ncol && used[cc]) ++cc;
    col[i] = cc;
    if (cc == ncol) ++ncol;
  }
  s->mcStart.assign(ncol + 1, 0);
  for (int i = 0; i < N; ++i) s->mcStart[col[i] + 1]++;
  for (int k = 0; k < ncol; ++k) s->mcStart[k + 1] += s->mcStart[k];
  std::vector<int> items(N), pos(s->mcStart.begin(), s->mcStart.end() - 1);
  for (int i = 0; i < N; ++i) items[pos[col[i]]++] = i;
  // neighbour element of every side in colour order (-1: boundary, absent
  // or a side onto the element itself)
  std::vector<int> nb((size_t)S * std::max(N, 1), -1);
  for (int p = 0; p < N; ++p)
    for (int k = 0; k < S; ++k) {
      const int lk = link[(size_t)k * Np + items[p]];
      if (lk >= 0 && (lk >> 2) != items[p]) nb[(size_t)k * N + p] = lk >> 2;
    }
  const size_t kb = c->prm.precision == CCFD_FP64 ? 8 : 4;
  s->mcCells = (int*)imalloc(c, s, (size_t)std::max(N, 1) * 4);
  s->mcLink = (int*)imalloc(c, s, nb.size() * 4);
  s->mcBlk = imalloc(c, s, 16 * (size_t)S * std::max(N, 1) * kb);
  s->mcDinv = imalloc(c, s, 16 * (size_t)std::max(N, 1) * kb);
  if (!s->mcCells || !s->mcLink || !s->mcBlk || !s->mcDinv)
    return fail(CCFD_ENOMEM, "multicolour schedule allocation failed");
  if (hipMemcpyAsync(s->mcCells, items.data(), N * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(s->mcLink, nb.data(), nb.size() * 4, hipMemcpyHostToDevice,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "multicolour schedule upload failed");
  s->haveMC = true;
  return CCFD_OK;
}

// coarse levels of the multigrid preconditioner (precond = 3): groups of
// elements by heavy-edge matching on face length / barycentre distance,
// down to at most kMGMinCells elements
constexpr int kMGMinCells = 256;
constexpr int kMGMaxLevels = 12;

template <typename X>
static X* upload_new(ccfd_gpu_ctx* c, ImplicitState* s, const std::vector<X>& v,
                     int& rc) {
  X* d = (X*)imalloc(c, s, std::max<size_t>(v.size(), 1) * sizeof(X));
  if (!d) {
    rc = CCFD_ENOMEM;
    return nullptr;
  }
  if (!v.empty() && hipMemcpyAsync(d, v.data(), v.size() * sizeof(X),
                                   hipMemcpyHostToDevice, c->st) != hipSuccess)
    rc = CCFD_EHIP;
  return d;
}

static int prepare_mg(ccfd_gpu_ctx* c) {
  ImplicitState* s = c->imp;
  if (s->haveMG) return CCFD_OK;
  const int N = c->N, Np = c->Np, S = c->S;
  const size_t n = (size_t)S * Np;
  std::vector<int> link(n);
  std::vector<double> flen(n), bbd(n), w(n, 0.0);
  if (hipMemcpyAsync(link.data(), c->link, n * 4, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(flen.data(), c->g64.flen, n * 8, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(bbd.data(), c->g64.bbd, n * 8, hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "multigrid geometry download failed");
  for (size_t k = 0; k < n; ++k)
    if (link[k] >= 0 && bbd[k] > 0.0) w[k] = fabs(flen[k]) / bbd[k];
  std::vector<AggLevel> lv;
  build_agglomeration(N, Np, S, link.data(), w.data(), kMGMinCells,
                      kMGMaxLevels, lv);
  int rc = CCFD_OK, nPrev = N;
  s->mg.resize(lv.size());
  for (size_t l = 0; l < lv.size(); ++l) {
    const AggLevel& A = lv[l];
    ImplicitState::MGLevel& G = s->mg[l];
    G.n = A.n;
    G.nnz = A.nnz;
    G.nPrev = nPrev;
    G.aggOf = upload_new(c, s, A.aggOf, rc);
    G.memStart = upload_new(c, s, A.memStart, rc);
    G.mem = upload_new(c, s, A.mem, rc);
    G.rowStart = upload_new(c, s, A.rowStart, rc);
    G.col = upload_new(c, s, A.col, rc);
    G.colorCells = upload_new(c, s, A.colorCells, rc);
    G.dSrcStart = upload_new(c, s, A.dSrcStart, rc);
    G.dSrc = upload_new(c, s, A.dSrc, rc);
    G.oSrcStart = upload_new(c, s, A.oSrcStart, rc);
    G.oSrc = upload_new(c, s, A.oSrc, rc);
    G.colorStart = A.colorStart;
    G.D = (double*)imalloc(c, s, 16 * (size_t)A.n * 8);
    G.Dinv = (double*)imalloc(c, s, 16 * (size_t)A.n * 8);
    G.O = (double*)imalloc(c, s, 16 * std::max<size_t>(A.nnz, 1) * 8);
    G.b = (double*)imalloc(c, s, 4 * (size_t)A.n * 8);
    G.x = (double*)imalloc(c, s, 4 * (size_t)A.n * 8);
    G.r = (double*)imalloc(c, s, 4 * (size_t)A.n * 8);
    if (!G.D || !G.Dinv || !G.O || !G.b || !G.x || !G.r) rc = CCFD_ENOMEM;
    if (rc) break;
    nPrev = A.n;
  }
  if (rc == CCFD_OK && hipStreamSynchronize(c->st) != hipSuccess) rc = CCFD_EHIP;
  if (rc)
    return fail(rc, "multigrid level setup failed");
  s->haveMG = true;
  return CCFD_OK;
}

// coarse blocks of every level from the element blocks (Dg holds D)
template <typename M>
static void mg_assemble(ccfd_gpu_ctx* c) {
  ImplicitState* s = c->imp;
  for (size_t l = 0; l < s->mg.size(); ++l) {
    ImplicitState::MGLevel& G = s->mg[l];
    const dim3 gd((G.n + 127) / 128), go((G.nnz + 127) / 128), b(128);
    if (l == 0) {
      hipLaunchKernelGGL(k_mg_diag<M>, gd, b, 0, c->st, G.n, G.dSrcStart,
                         G.dSrc, (const M*)s->Dg, (long)c->Np,
                         (const M*)s->Blk, (long)c->S * c->Np, G.D, G.Dinv,
                         s->flag);
      if (G.nnz)
        hipLaunchKernelGGL(k_mg_off<M>, go, b, 0, c->st, G.nnz, G.oSrcStart,
                           G.oSrc, (const M*)s->Dg, (long)c->Np,
                           (const M*)s->Blk, (long)c->S * c->Np, G.O);
    } else {
      const ImplicitState::MGLevel& P = s->mg[l - 1];
      hipLaunchKernelGGL(k_mg_diag<double>, gd, b, 0, c->st, G.n, G.dSrcStart,
                         G.dSrc, (const double*)P.D, (long)P.n,
                         (const double*)P.O, (long)P.nnz, G.D, G.Dinv, s->flag);
      if (G.nnz)
        hipLaunchKernelGGL(k_mg_off<double>, go, b, 0, c->st, G.nnz,
                           G.oSrcStart, G.oSrc, (const double*)P.D, (long)P.n,
                           (const double*)P.O, (long)P.nnz, G.O);
    }
  }
}

// order of the operator whose Jacobian the LU-SGS blocks hold
static int jacobian_order(const ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg) {
  const int o = c->prm.spatialOrder;
  return cfg->jacobianOrder > 0 && cfg->jacobianOrder < o ? cfg->jacobianOrder : o;
}

template <typename T>
static void operator_at(ccfd_gpu_ctx* c, double time) {
  OpArgs<T> a = op_args<T>(c);
  a.time0 = time;
  launch_operator<T>(c, a);
}

// device dot of two FP64 vectors into a device scalar
static void dot_dev(ccfd_gpu_ctx* c, const double* a, const double* b,
                    double* out) {
  launch_dot(c, a, b, out, c->imp->part);
}

static double dot_host(ccfd_gpu_ctx* c, const double* a, const double* b) {
  dot_dev(c, a, b, c->scal + 8);
  hipMemcpyAsync(c->hscal + 8, c->scal + 8, 8, hipMemcpyDeviceToHost, c->st);
  hipStreamSynchronize(c->st);
  return c->hscal[8];
}

// buildMatrix by coloured finite differences at the primitive state of XK
// consJac: perturb the conservative state (precond = 2) instead of the
// primitive state as buildMatrix does (src/linearSolver.c:199-201)
// jacOrder 1 on a 2nd-order context: the blocks are the Jacobian of the
// first-order operator (differences against its own base residual)
template <typename T, typename M>
static int build_matrix(ccfd_gpu_ctx* c, double tGlobal, double dt,
                        const double* dtv, bool consJac, int jacOrder,
                        bool coarse) {
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np, S = c->S;
  const double reps = c->prm.precision == CCFD_FP64 ? sqrt(DBL_EPSILON)
                                                    : sqrt((double)FLT_EPSILON);
  const double srEps = 1.0 / reps;
  size_t kb = sizeof(M);
  hipMemsetAsync(s->Dg, 0, 16 * (size_t)Np * kb, c->st);
  hipMemsetAsync(s->Blk, 0, 16 * (size_t)S * Np * kb, c->st);
  hipMemsetAsync(s->flag, 0, 64, c->st);
  // base primitive state: consPrim(XK), kept in FP64
  hipLaunchKernelGGL(k_load_state<double>, cells_grid(c), dim3(256), 0, c->st,
                     N, Np, s->XK, c->U, s->primBase, eqd(c));
  auto perturb = [&](int col, int iv) {
    if (consJac)
      hipLaunchKernelGGL(k_fd_perturb_cons<T>, cells_grid(c), dim3(256), 0,
                         c->st, N, Np, s->XK, s->color, col, iv, reps,
                         (T*)c->prim[c->cur], eqd(c));
    else
      hipLaunchKernelGGL(k_fd_perturb<T>, cells_grid(c), dim3(256), 0, c->st,
                         N, Np, s->primBase, s->color, col, iv, reps,
                         (T*)c->prim[c->cur]);
  };
  const int opOrder = c->prm.spatialOrder;
  // (the operator and matrix types always agree when this runs; the other
  // instantiations never take the branch)
  if (std::is_same<T, M>::value && consJac && jacOrder == 1 &&
      s->jacMethod == 1) {
    // first-order Jacobian of the conservative state, element by element in
    // one launch (jac1_kernel): the same blocks as the coloured differences
    c->prm.spatialOrder = 1;
    OpArgs<T> a = op_args<T>(c);
    a.time0 = tGlobal;
    c->prm.spatialOrder = opOrder;
    if (c->hasExact && c->G > 0)
      hipLaunchKernelGGL(bc_exact_kernel<T>, dim3((c->G + 127) / 128),
                         dim3(128), 0, c->st, a, 1);
    if constexpr (std::is_same<T, M>::value)
      launch_jac1_code<T, T>(c->prm.iFlux, a, s->XK, reps, srEps, eqd(c),
                             (T*)s->Dg, (T*)s->Blk, c->st);
  } else {
    const double* base = s->RXK;
    if (jacOrder != opOrder) {
      // the first-order operator (split kernels, zero gradients) at XK
      c->prm.spatialOrder = jacOrder;
      perturb(-1, 0);
      operator_at<T>(c, tGlobal);
      hipLaunchKernelGGL((k_copy_cvt<T, double>), dim3((4 * Np + 255) / 256),
                         dim3(256), 0, c->st, 4 * (long)Np, (const T*)c->ut, s->W);
      base = s->W;
    }
    for (int col = 0; col < s->nColors; ++col) {
      for (int iv = 0; iv < 4; ++iv) {
        perturb(col, iv);
        operator_at<T>(c, tGlobal);
        hipLaunchKernelGGL((k_fd_record<T, M>), cells_grid(c), dim3(256), 0,
                           c->st, N, Np, S, c->link, s->color, col, iv,
                           (const T*)c->ut, base, srEps, (M*)s->Dg, (M*)s->Blk);
      }
    }
    c->prm.spatialOrder = opOrder;
  }
  hipLaunchKernelGGL(k_fd_finish<M>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                     S, c->link, dt, dtv, c->area, (M*)s->Dg, (M*)s->Blk,
                     (M*)s->Dinv, s->flag);
  if (coarse) mg_assemble<M>(c);
  if (s->haveMC)
    hipLaunchKernelGGL(k_gs_order<M>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                       S, s->mcCells, (const M*)s->Blk, (const M*)s->Dinv,
                       (M*)s->mcBlk, (M*)s->mcDinv);
  int flag = 0;
  hipMemcpyAsync(&flag, s->flag, 4, hipMemcpyDeviceToHost, c->st);
  if (hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "buildMatrix failed: %s",
                hipGetErrorString(hipGetLastError()));
  if (flag) return fail(CCFD_ESINGULAR, "LUSGS D-Matrix is singular");
  return CCFD_OK;
}

// symmetric Gauss-Seidel sweeps by colour on the elements (colours
// 0..nc-1, then back down), from x = 0 when `zero` (the first sweep is then
// the LU-SGS forward and backward sweep) or from the current x
template <typename M, typename K>
static void gs0(ccfd_gpu_ctx* c, int sweeps, bool zero, const K* b, K* x) {
  ImplicitState* s = c->imp;
  const int Np = c->Np, S = c->S;
  const int nc = (int)s->mcStart.size() - 1;
  if (zero) hipMemsetAsync(x, 0, 4 * (size_t)Np * sizeof(K), c->st);
  auto update = [&](int k) {
    const int n = s->mcStart[k + 1] - s->mcStart[k];
    if (sizeof(M) == 4)
      hipLaunchKernelGGL((k_gs_rows<M, K, 4>), dim3((n + 63) / 64), dim3(256), 0,
                         c->st, n, s->mcStart[k], s->mcCells, c->N, Np, S,
                         s->mcLink, (const M*)s->mcBlk, (const M*)s->mcDinv, b, x);
    else
      hipLaunchKernelGGL((k_gs_cells<M, K>), dim3((n + 127) / 128), dim3(128), 0,
                         c->st, n, s->mcStart[k], s->mcCells, c->N, Np, S,
                         s->mcLink, (const M*)s->mcBlk, (const M*)s->mcDinv, b,
                         x);
  };
  // a colour updated twice in a row gives the same values: skip repeats
  int last = -1;
  for (int sw = 0; sw < std::max(sweeps, 1); ++sw) {
    for (int k = 0; k < nc; ++k)
      if (k != last) update(last = k);
    for (int k = nc - 1; k >= 0; --k)
      if (k != last) update(last = k);
  }
}

// the same on a coarse multigrid level, from x = 0 when `zero`
static void gs_level(ccfd_gpu_ctx* c, ImplicitState::MGLevel& G, int sweeps,
                     bool zero) {
  const int nc = (int)G.colorStart.size() - 1;
  if (zero) hipMemsetAsync(G.x, 0, 4 * (size_t)G.n * 8, c->st);
  auto update = [&](int k) {
    const int m = G.colorStart[k + 1] - G.colorStart[k];
    hipLaunchKernelGGL(k_mg_gs, dim3((m + 127) / 128), dim3(128), 0, c->st, m,
                       G.colorCells + G.colorStart[k], G.n, G.nnz, G.rowStart,
                       G.col, G.O, G.Dinv, G.b, G.x);
  };
  int last = -1;
  for (int sw = 0; sw < std::max(sweeps, 1); ++sw) {
    for (int k = 0; k < nc; ++k)
      if (k != last) update(last = k);
    for (int k = nc - 1; k >= 0; --k)
      if (k != last) update(last = k);
  }
}

// one multigrid V-cycle from x = 0: `sweeps` symmetric Gauss-Seidel sweeps
// before and after the coarse correction on every level, kMGCoarseSweeps on
// the coarsest
constexpr int kMGCoarseSweeps = 8;

template <typename M, typename K>
static void mg_vcycle(ccfd_gpu_ctx* c, int sweeps, const K* b, K* x) {
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np, S = c->S, L = (int)s->mg.size();
  gs0<M, K>(c, sweeps, true, b, x);
  if (L == 0) return;
  // restrict the element residual (s->W is free while GMRES preconditions)
  hipLaunchKernelGGL((k_mg_resid0<M, K>), cells_grid(c), dim3(256), 0, c->st,
                     N, Np, S, c->link, (const M*)s->Dg, (const M*)s->Blk, b, x,
                     s->W);
  hipLaunchKernelGGL(k_mg_restrict, dim3((s->mg[0].n + 127) / 128), dim3(128), 0,
                     c->st, s->mg[0].n, s->mg[0].memStart, s->mg[0].mem, s->W,
                     (long)Np, s->mg[0].b);
  for (int l = 0; l < L; ++l) {
    ImplicitState::MGLevel& G = s->mg[l];
    if (l == L - 1) {
      gs_level(c, G, kMGCoarseSweeps, true);
      break;
    }
    gs_level(c, G, sweeps, true);
    ImplicitState::MGLevel& H = s->mg[l + 1];
    hipLaunchKernelGGL(k_mg_resid, dim3((G.n + 127) / 128), dim3(128), 0, c->st,
                       G.n, G.nnz, G.rowStart, G.col, G.D, G.O, G.b, G.x, G.r);
    hipLaunchKernelGGL(k_mg_restrict, dim3((H.n + 127) / 128), dim3(128), 0,
                       c->st, H.n, H.memStart, H.mem, G.r, (long)G.n, H.b);
  }
  for (int l = L - 2; l >= 0; --l) {
    ImplicitState::MGLevel& G = s->mg[l];
    const ImplicitState::MGLevel& H = s->mg[l + 1];
    hipLaunchKernelGGL(k_mg_prolong<double>, dim3((G.n + 127) / 128), dim3(128),
                       0, c->st, G.n, H.aggOf, H.x, H.n, G.x, (long)G.n);
    gs_level(c, G, sweeps, false);
  }
  hipLaunchKernelGGL(k_mg_prolong<K>, cells_grid(c), dim3(256), 0, c->st, N,
                     s->mg[0].aggOf, s->mg[0].x, s->mg[0].n, x, (long)Np);
  gs0<M, K>(c, sweeps, false, b, x);
}

// precond 2: `sweeps` symmetric Gauss-Seidel sweeps by colour from x = 0;
// precond 3: a multigrid V-cycle with them as smoother; precond 1: the
// reference's LU-SGS in its element order, by dependency levels
template <typename M, typename K>
static void lusgs(ccfd_gpu_ctx* c, int mode, int sweeps, const K* b, K* x) {
  ImplicitState* s = c->imp;
  const int Np = c->Np, S = c->S;
  if (mode == 2) {
    gs0<M, K>(c, sweeps, true, b, x);
    return;
  }
  if (mode == 3) {
    mg_vcycle<M, K>(c, sweeps, b, x);
    return;
  }
  for (size_t l = 0; l + 1 < s->lvlFStart.size(); ++l) {
    int n = s->lvlFStart[l + 1] - s->lvlFStart[l];
    hipLaunchKernelGGL((k_lusgs_fwd<M, K>), dim3((n + 127) / 128), dim3(128), 0,
                       c->st, n, s->lvlF + s->lvlFStart[l], Np, S, c->link,
                       s->refIdx, (const M*)s->Blk, (const M*)s->Dinv, b,
                       (M*)s->xs);
  }
  for (size_t l = 0; l + 1 < s->lvlBStart.size(); ++l) {
    int n = s->lvlBStart[l + 1] - s->lvlBStart[l];
    hipLaunchKernelGGL((k_lusgs_bwd<M, K>), dim3((n + 127) / 128), dim3(128), 0,
                       c->st, n, s->lvlB + s->lvlBStart[l], Np, S, c->link,
                       s->refIdx, (const M*)s->Blk, (const M*)s->Dinv,
                       (const M*)s->xs, x);
  }
}

// GMRES_M (src/linearSolver.c:385-510) with B = s->B, result in s->dX.
// cfg->cycles > 1 restarts it (each cycle from the true residual
// -B - J dX) until the tolerance or the cycle count; one cycle is GMRES_M.
//
// The iterations are queued without waiting for each other: the
// orthogonalisation, the Givens rotations and the back-solve run on the
// device (k_mgs_*, k_givens), and the host waits only for the iteration
// kLookahead places behind the one it queues next, reading the done flag
// k_givens leaves in pinned memory. Iterations queued after the abort
// criterion was met find the flag set and do no Arnoldi work, so the
// result is that of stopping at the first iteration that met it.
constexpr int kLookahead = 2;

template <typename T, typename K>
static int gmres(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
                 double tGlobal, double time, double dt, const double* dtv,
                 double alpha, double normB, int* nIter, bool cgs2) {
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np, nK = cfg->nKdim;
  const size_t vs = 4 * (size_t)Np;
  K* V = (K*)s->V;
  K* Z = (K*)s->Z;
  const double abortCrit = cfg->epsGMRES * normB;
  const double sqeps = c->prm.precision == CCFD_FP64 ? sqrt(DBL_EPSILON)
                                                     : sqrt((double)FLT_EPSILON);
  const int cycles = std::max(cfg->cycles, 1);
  const int R = red_blocks(c);
  const GmresDev g = gmres_view(s->gdev, nK);
  hipMemsetAsync(s->dX, 0, vs * 8, c->st);
  if (cfg->precond) {
    const bool consJac = cfg->precond >= 2;
    const int jo = jacobian_order(c, cfg);
    const bool coarse = cfg->precond == 3;
    int rc = c->prm.precision == CCFD_FP64
                 ? build_matrix<T, double>(c, tGlobal, dt, dtv, consJac, jo, coarse)
                 : build_matrix<T, float>(c, tGlobal, dt, dtv, consJac, jo, coarse);
    if (rc) return rc;
  }
  // matrixVector (src/linearSolver.c:343-373) of v into s->W; |v|^2 in FP64
  // (sc null: outside a cycle, always evaluated)
  double* epsFD = s->H + nK + 5;
  auto jv = [&](const K* v, const double* sc) {
    hipLaunchKernelGGL(k_norm_partial<K>, dim3(R), dim3(256), 0, c->st, N, Np,
                       v, g.sc, s->partA);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(256), 0, c->st, s->partA, R,
                       g.sc);
    hipLaunchKernelGGL((k_jv_perturb<T, K>), cells_grid(c), dim3(256), 0, c->st,
                       N, Np, g.sc + GS_VV, sqeps, s->XK, v, c->U,
                       (T*)c->prim[c->cur], epsFD, eqd(c), sc);
    operator_at<T>(c, time);
    hipLaunchKernelGGL((k_jv_diff<T, K>), cells_grid(c), dim3(256), 0, c->st, N,
                       Np, v, (const T*)c->ut, s->RXK, epsFD, alpha * dt, dtv,
                       alpha, c->area, s->W, sc);
  };
  for (int cyc = 0; cyc < cycles; ++cyc) {
    double beta = normB;
    if (cyc == 0) {
      hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(256), 0, c->st, g, nK, beta,
                         abortCrit, normB);
      hipLaunchKernelGGL(k_neg_div<K>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                         s->B, V, normB);
    } else {
      // r = -B - J dX; V_0 = r / |r| (dX staged in V_0 for the product)
      hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(256), 0, c->st, g, nK, 0.0,
                         abortCrit, normB);
      hipLaunchKernelGGL((k_copy_cvt<double, K>), dim3((vs + 255) / 256),
                         dim3(256), 0, c->st, (long)vs, s->dX, V);
      jv(V, nullptr);
      hipLaunchKernelGGL(k_axpy, cells_grid(c), dim3(256), 0, c->st, N, Np, s->W,
                         s->B, 1.0);
      beta = sqrt(dot_host(c, s->W, s->W));
      if (!(beta > abortCrit)) break;
      hipLaunchKernelGGL(k_gm_init, dim3(1), dim3(256), 0, c->st, g, nK, beta,
                         abortCrit, normB);
      hipLaunchKernelGGL(k_neg_div<K>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                         s->W, V, beta);
    }
    for (int m = 0; m < nK + 8; ++m) s->hflag[m] = 0;
    for (int m = 0; m < nK; ++m) {
      if (m >= kLookahead) {
        const int w = m - kLookahead;
        if (hipEventSynchronize(s->ev[w]) != hipSuccess)
          return fail(CCFD_EHIP, "GMRES iteration failed: %s",
                      hipGetErrorString(hipGetLastError()));
        if (__atomic_load_n(s->hflag + w, __ATOMIC_ACQUIRE)) break;
      }
      K* Vm = V + m * vs;
      K* Zm = Z + m * vs;
      if (cfg->precond) {
        if (c->prm.precision == CCFD_FP64)
          lusgs<double, K>(c, cfg->precond, cfg->sweeps, Vm, Zm);
        else
          lusgs<float, K>(c, cfg->precond, cfg->sweeps, Vm, Zm);
      } else {
        hipMemcpyAsync(Zm, Vm, vs * sizeof(K), hipMemcpyDeviceToDevice, c->st);
      }
      jv(Zm, g.sc);
      double* pin = s->partA;
      if (!cgs2) {
        // modified Gram-Schmidt (src/linearSolver.c:440-450), one launch per
        // basis vector, the last one also giving |W|^2
        hipLaunchKernelGGL(k_mgs_first<K>, dim3(R), dim3(256), 0, c->st, N, Np,
                           V, s->W, g.sc, s->partA);
        double* pout = s->partB;
        for (int n = 0; n <= m; ++n) {
          hipLaunchKernelGGL(k_mgs_step<K>, dim3(R), dim3(256), 0, c->st, N, Np,
                             n, m, V, vs, s->W, pin, pout, g);
          std::swap(pin, pout);
        }
      } else {
        // classical Gram-Schmidt twice (CGS2): three launches per pass
        // whatever m, orthogonal to working precision as MGS is
        for (int pass = 0; pass < 2; ++pass) {
          hipLaunchKernelGGL(k_cgs_dots<K>, dim3(R), dim3(256), 0, c->st, N, Np,
                             m, V, vs, s->W, g.sc, s->cgsPart);
          hipLaunchKernelGGL(k_cgs_sum, dim3(m + 1), dim3(256), 0, c->st, m, R,
                             s->cgsPart, s->cgsH, g, pass);
          hipLaunchKernelGGL(k_cgs_update<K>, dim3(R), dim3(256), 0, c->st, N,
                             Np, m, V, vs, s->cgsH, s->W, g.sc,
                             pass ? s->partA : nullptr);
        }
      }
      hipLaunchKernelGGL(k_givens, dim3(1), dim3(256), 0, c->st, g, m, nK, pin,
                         R, s->hflagDev);
      if (m + 1 < nK)
        hipLaunchKernelGGL(k_scale_dev<K>, cells_grid(c), dim3(256), 0, c->st, N,
                           Np, s->W, V + (m + 1) * vs, g.sc);
      if (hipEventRecord(s->ev[m], c->st) != hipSuccess)
        return fail(CCFD_EHIP, "GMRES event record failed");
    }
    // dX += Z alp over the iterations up to the one that stopped the cycle
    hipLaunchKernelGGL(k_update_dx<K>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                       g.sc, g.alp, Z, vs, s->dX);
    double gsc[GS_N];
    if (hipMemcpyAsync(gsc, g.sc, sizeof(gsc), hipMemcpyDeviceToHost, c->st) !=
            hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return fail(CCFD_EHIP, "GMRES update failed: %s",
                  hipGetErrorString(hipGetLastError()));
    *nIter += (int)gsc[GS_MSTAR] + 1;
    s->gmresRel = gsc[GS_REL];
    if (gsc[GS_DONE] == 2.0) return CCFD_OK;   // abort criterion met
  }
  return CCFD_OK;
}

template <typename T>
static int gmres_any(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
                     double tGlobal, double time, double dt, const double* dtv,
                     double alpha, double normB, int* nIter) {
  // steady solves (local steps) orthogonalise by CGS2; the time-accurate
  // step and GMRES_M itself keep the reference's modified Gram-Schmidt
  const bool cgs2 = dtv != nullptr;
  if (c->prm.precision == CCFD_FP64)
    return gmres<T, double>(c, cfg, tGlobal, time, dt, dtv, alpha, normB, nIter,
                            cgs2);
  return gmres<T, float>(c, cfg, tGlobal, time, dt, dtv, alpha, normB, nIter,
                         cgs2);
}

// Newton residual F(X) = X - Q - alpha*dt*R(X) in FP64 (the outer residual
// of the defect correction: the converged state is the FP64 operator's
// whatever the precision of the Krylov / preconditioner work). U <- X,
// XK <- X when XKout, RXK <- R(X); ut keeps R(X) in FP64.
static void outer_residual(ccfd_gpu_ctx* c, const double* X, double time,
                           double adt, const double* dtv, double alpha,
                           double* XKout) {
  ImplicitState* s = c->imp;
  hipLaunchKernelGGL(k_load_state<double>, cells_grid(c), dim3(256), 0, c->st,
                     c->N, c->Np, X, c->U, (double*)c->prim[c->cur], eqd(c));
  operator_at<double>(c, time);
  hipLaunchKernelGGL(k_newton_residual<double>, cells_grid(c), dim3(256), 0,
                     c->st, c->N, c->Np, c->U, s->Q, (const double*)c->ut, adt,
                     dtv, alpha, c->area, s->FXK, XKout, s->RXK);
}

// FP32-flux mode: R(XK) with the FP32 operator as the base of the Frechet
// differences and the finite-difference Jacobian (differences of one
// arithmetic), from U = XK
static int base_residual_fp32(ccfd_gpu_ctx* c, double time) {
  ImplicitState* s = c->imp;
  int rc = refresh_prim(c);
  if (rc) return rc;
  operator_at<float>(c, time);
  const long n = 4 * (long)c->Np;
  hipLaunchKernelGGL((k_copy_cvt<float, double>), dim3((n + 255) / 256),
                     dim3(256), 0, c->st, n, (const float*)c->ut, s->RXK);
  return CCFD_OK;
}

// dtv (may be null): per-element pseudo-time steps instead of dt; with
// steady set, a Newton loop that stops at nNewtonIter is accepted (one
// Newton iteration per pseudo-time step is the usual steady iteration).
// The Newton residual is FP64 in both precisions; in FP32-flux mode the
// GMRES operator evaluations, the Jacobian and the LU-SGS sweeps are FP32.
// maxChange > 0 (steady solves only): every Newton update is scaled so that
// no element's density or pressure changes by more than that fraction;
// *relax (may be null) receives the smallest scale used.
template <typename T>
static int implicit_step_t(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
                           double t, double dt, const double* dtv, bool steady,
                           double res[4], int* nNewton, int* nGMRES,
                           double maxChange = 0.0, double* relax = nullptr) {
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np;
  const size_t vs = 4 * (size_t)Np;
  const double alpha = 1.0, beta = 1.0;
  const bool mixed = sizeof(T) == 4;
  // Q = U; prim = consPrim(U) (src/timeDiscretization.c:418-428)
  hipMemcpyAsync(s->Q, c->U, vs * 8, hipMemcpyDeviceToDevice, c->st);
  const double time = t + beta * dt;
  outer_residual(c, s->Q, time, alpha * dt, dtv, alpha, s->XK);
  const double n2F0 = dot_host(c, s->FXK, s->FXK);
  const double eps2 = cfg->epsNewton * cfg->epsNewton;
  const double eps2loc = fmax(1e-8 * 1e-8 * N * eps2 / n2F0, eps2);
  double n2FK = n2F0;
  int nIn = 0;
  if (cfg->precond) {
    int rc = prepare_lusgs(c, jacobian_order(c, cfg));
    if (!rc && cfg->precond >= 2) rc = prepare_multicolor(c);
    if (!rc && cfg->precond == 3) rc = prepare_mg(c);
    if (rc) return rc;
  }
  double* wmin = s->H + cfg->nKdim + 3;
  if (maxChange > 0.0)
    hipLaunchKernelGGL(k_set1, dim3(1), dim3(1), 0, c->st, wmin, 1.0);
  while ((n2FK > eps2loc * n2F0) && (nIn < cfg->nNewtonIter)) {
    nIn++;
    if (mixed) {
      int rc = base_residual_fp32(c, time);
      if (rc) return rc;
    }
    hipMemcpyAsync(s->B, s->FXK, vs * 8, hipMemcpyDeviceToDevice, c->st);
    int rc = gmres_any<T>(c, cfg, t, time, dt, dtv, alpha, sqrt(n2FK), nGMRES);
    if (rc) return rc;
    // bounded update (steady solves): the weight stays on the device, the
    // smallest one of the step is read with the residual below
    const double* wdev = nullptr;
    if (maxChange > 0.0) {
      hipLaunchKernelGGL(k_relchange_partial, dim3(c->nb), dim3(256), 0, c->st,
                         N, Np, s->XK, s->dX, c->prm.gam - 1.0, s->part);
      hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, c->st, c->nb,
                         s->part, s->H + cfg->nKdim + 6, maxChange,
                         s->H + cfg->nKdim + 7, wmin);
      wdev = s->H + cfg->nKdim + 7;
    }
    // XK += w dX; U = XK; prim = consPrim(U) (src/timeDiscretization.c:505-520)
    hipLaunchKernelGGL(k_axpy, cells_grid(c), dim3(256), 0, c->st, N, Np, s->XK,
                       s->dX, 1.0, wdev);
    outer_residual(c, s->XK, time, alpha * dt, dtv, alpha, nullptr);
    // a steady step stops after nNewtonIter iterations whatever |F|
    if (!(steady && nIn >= cfg->nNewtonIter))
      n2FK = dot_host(c, s->FXK, s->FXK);
  }
  *nNewton += nIn;
  if (nIn == cfg->nNewtonIter && !steady)
    return fail(CCFD_ENOCONV,
                "Newton NOT converged with %d Newton iterations (|F|^2/|F0|^2 = %g)",
                nIn, n2FK / n2F0);
  if (maxChange > 0.0)
    hipMemcpyAsync(c->hscal + 48, wmin, 8, hipMemcpyDeviceToHost, c->st);
  // globalResidual of the FP64 dU/dt of the new state (ut holds it)
  if (res) {
    int rc = residual_of(c, true, res);
    if (rc) return rc;
  }
  if (maxChange > 0.0) {
    if (hipStreamSynchronize(c->st) != hipSuccess)
      return fail(CCFD_EHIP, "update bound failed");
    if (relax) *relax = fmin(*relax, c->hscal[48]);
  }
  // leave the primitive state in the context's arithmetic
  return mixed ? refresh_prim(c) : CCFD_OK;
}

extern "C" int ccfd_gpu_implicit_step(ccfd_gpu_ctx* c,
                                      const ccfd_gpu_implicit_cfg* cfg, double t,
                                      double dt, double res[4], int* nNewton,
                                      int* nGMRES) {
  if (!c || !c->hasMesh || !cfg || !nNewton || !nGMRES)
    return fail(CCFD_EINVAL, "bad implicit-step arguments");
  if (cfg->nKdim < 1 || cfg->nKdim > 200 || cfg->nNewtonIter < 1 ||
      cfg->precond < 0 || cfg->precond > 3 || cfg->jacobianOrder < 0 ||
      cfg->jacobianOrder > 2 || cfg->sweeps < 0 || cfg->sweeps > 64 ||
      cfg->cycles < 0 || cfg->cycles > 100)
    return fail(CCFD_EINVAL, "nKdim must be in 1..200, nNewtonIter >= 1, "
                             "precond 0..3, jacobianOrder 0..2, sweeps 0..64, "
                             "cycles 0..100");
  hipSetDevice(c->device);
  int rc = ensure(c, cfg->nKdim);
  if (rc) return rc;
  return c->prm.precision == CCFD_FP64
             ? implicit_step_t<double>(c, cfg, t, dt, nullptr, false, res, nNewton,
                                       nGMRES)
             : implicit_step_t<float>(c, cfg, t, dt, nullptr, false, res, nNewton,
                                      nGMRES);
}

template <typename T>
static int steady_t(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
                    double cfl, int maxSteps, double drop, double res[4],
                    double res0[4], int* steps, int* nNewton, int* nGMRES) {
  ImplicitState* s = c->imp;
  if (!s->dtLoc) {
    s->dtLoc = (double*)imalloc(c, s, (size_t)c->Np * 8);
    if (!s->dtLoc) return fail(CCFD_ENOMEM, "local step allocation failed");
  }
  const Eq<double> q = eqd(c);
  const double gamPrMax = fmax(4.0 / 3.0, q.gam / q.Pr);
  const double mu = c->prm.navierstokes ? q.mu : 0.0;
  // res0[0] > 0: continue a solve, measuring the drop against res0 and
  // carrying on with the CFL number the previous call stopped at
  const bool cont = res0[0] > 0.0;
  double cflk = cont && s->steadyCfl > 0.0 ? fmin(cfl, s->steadyCfl)
                                           : fmin(cfl, 1.0);
  for (int k = 0; k < maxSteps; ++k) {
    // pseudo-time steps from the current state. The CFL number starts at
    // 1 and follows the linear solver: it grows by 20% after a step whose
    // GMRES reduced its residual tenfold, stays after one that reduced it
    // by more than half, and is cut by 30% after one that did not (at large
    // steps the preconditioned Krylov solve stops converging and the Newton
    // update is of no use) or whose update had to be scaled down (more than
    // a 20% density or pressure change somewhere).
    int rc = refresh_prim(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_local_dt<T>, cells_grid(c), dim3(256), 0, c->st, c->N,
                       c->Np, (const T*)c->prim[c->cur], c->area, c->sx, c->sy,
                       q.gam, mu, gamPrMax, cflk, cflk, s->dtLoc);
    double relax = 1.0;
    rc = implicit_step_t<T>(c, cfg, 0.0, 0.0, s->dtLoc, true, res, nNewton,
                            nGMRES, 0.2, &relax);
    if (rc) return rc;
    const double lo = fmin(cfl, 1.0), rel = s->gmresRel;
    if (relax < 1.0 || rel > 0.5)
      cflk = fmax(lo, 0.7 * cflk);
    else if (rel <= 0.1)
      cflk = fmin(cfl, 1.2 * cflk);
    *steps = k + 1;
    s->steadyCfl = cflk;
    if (k == 0 && !cont)
      for (int v = 0; v < 4; ++v) res0[v] = res[v];
    if (!(res[0] == res[0])) return fail(CCFD_ENONFINITE, "non-finite residual");
    if (res[0] <= drop * res0[0]) break;
  }
  return CCFD_OK;
}

extern "C" int ccfd_gpu_implicit_steady(ccfd_gpu_ctx* c,
                                        const ccfd_gpu_implicit_cfg* cfg,
                                        double cfl, int maxSteps, double drop,
                                        double res[4], double res0[4],
                                        int* steps, int* nNewton, int* nGMRES) {
  if (!c || !c->hasMesh || !cfg || !res || !res0 || !steps || !nNewton ||
      !nGMRES)
    return fail(CCFD_EINVAL, "bad steady-solve arguments");
  if (cfg->nKdim < 1 || cfg->nKdim > 200 || cfg->nNewtonIter < 1 ||
      cfg->precond < 0 || cfg->precond > 3 || cfg->jacobianOrder < 0 ||
      cfg->jacobianOrder > 2 || cfg->sweeps < 0 || cfg->sweeps > 64 ||
      cfg->cycles < 0 || cfg->cycles > 100 ||
      !(cfl > 0.0) || maxSteps < 1)
    return fail(CCFD_EINVAL, "bad steady-solve settings");
  hipSetDevice(c->device);
  int rc = ensure(c, cfg->nKdim);
  if (rc) return rc;
  *steps = 0;
  return c->prm.precision == CCFD_FP64
             ? steady_t<double>(c, cfg, cfl, maxSteps, drop, res, res0, steps,
                                nNewton, nGMRES)
             : steady_t<float>(c, cfg, cfl, maxSteps, drop, res, res0, steps,
                               nNewton, nGMRES);
}

extern "C" int ccfd_gpu_gmres(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
                              double tGlobal, double time, double dt,
                              double alpha, const double* B, double normB,
                              double* dX, int* nIter) {
  if (!c || !c->hasMesh || !cfg || !B || !dX || !nIter)
    return fail(CCFD_EINVAL, "bad GMRES arguments");
  if (cfg->nKdim < 1 || cfg->nKdim > 200 || cfg->precond < 0 ||
      cfg->precond > 3 || cfg->jacobianOrder < 0 || cfg->jacobianOrder > 2 ||
      cfg->sweeps < 0 || cfg->sweeps > 64 ||
      cfg->cycles < 0 || cfg->cycles > 100)
    return fail(CCFD_EINVAL, "nKdim must be in 1..200, precond 0..3, "
                             "jacobianOrder 0..2, sweeps 0..64, cycles 0..100");
  hipSetDevice(c->device);
  int rc = ensure(c, cfg->nKdim);
  if (rc) return rc;
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np;
  // linearisation point: the resident state
  hipMemcpyAsync(s->XK, c->U, 4 * (size_t)Np * 8, hipMemcpyDeviceToDevice, c->st);
  if ((rc = upload_vec(c, B, s->B))) return rc;
  if (cfg->precond && (rc = prepare_lusgs(c, jacobian_order(c, cfg)))) return rc;
  if (cfg->precond >= 2 && (rc = prepare_multicolor(c))) return rc;
  if (cfg->precond == 3 && (rc = prepare_mg(c))) return rc;
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_load_state<T>, cells_grid(c), dim3(256), 0, c->st, N, Np,
                       s->XK, c->U, (T*)c->prim[c->cur], eqd(c));
    operator_at<T>(c, time);
    hipLaunchKernelGGL((k_copy_cvt<T, double>), dim3((4 * Np + 255) / 256),
                       dim3(256), 0, c->st, (long)4 * Np, (const T*)c->ut, s->RXK);
    return gmres_any<T>(c, cfg, tGlobal, time, dt, nullptr, alpha, normB, nIter);
  };
  rc = c->prm.precision == CCFD_FP64 ? run(double(0)) : run(float(0));
  if (rc) return rc;
  // restore the primitive state of XK
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_load_state<double>, cells_grid(c), dim3(256), 0, c->st,
                       N, Np, s->XK, c->U, (double*)c->prim[c->cur], eqd(c));
  else
    hipLaunchKernelGGL(k_load_state<float>, cells_grid(c), dim3(256), 0, c->st,
                       N, Np, s->XK, c->U, (float*)c->prim[c->cur], eqd(c));
  return download_vec(c, s->dX, dX);
}

// the preconditioner's blocks (buildMatrix + calcDinv, src/linearSolver.c:
// 186-252) at the resident state, in the reference's element and side
// order: D (N x 4 x 4, I - dt dR/dX for each element) and, per side, the
// block towards the face neighbour (nSlots x 4 x 4, -dt dR_i/dX_j; zero for
// boundary sides). X is the primitive state for precond 1 (the reference's
// perturbation) and the conservative state otherwise. method 0 assembles
// by coloured finite differences, 1 element by element (first-order
// conservative Jacobians only; other cases fall back to method 0).
extern "C" int ccfd_gpu_jacobian_blocks(ccfd_gpu_ctx* c,
                                        const ccfd_gpu_implicit_cfg* cfg,
                                        double time, double dt, int method,
                                        double* D, double* Blk) {
  if (!c || !c->hasMesh || !cfg || !D || !Blk || method < 0 || method > 1 ||
      cfg->precond < 1 || cfg->precond > 3 || cfg->jacobianOrder < 0 ||
      cfg->jacobianOrder > 2)
    return fail(CCFD_EINVAL, "bad Jacobian-block arguments");
  hipSetDevice(c->device);
  int rc = ensure(c, std::max(cfg->nKdim, 1));
  if (rc) return rc;
  ImplicitState* s = c->imp;
  const int N = c->N, Np = c->Np, S = c->S;
  const int jo = jacobian_order(c, cfg);
  if ((rc = prepare_lusgs(c, jo))) return rc;
  if (cfg->precond >= 2 && (rc = prepare_multicolor(c))) return rc;
  hipMemcpyAsync(s->XK, c->U, 4 * (size_t)Np * 8, hipMemcpyDeviceToDevice, c->st);
  const int saved = s->jacMethod;
  s->jacMethod = method;
  const bool cons = cfg->precond >= 2;
  // R(X) at the linearisation point, prim = consPrim(X) as a Newton
  // iteration has it
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_load_state<T>, cells_grid(c), dim3(256), 0, c->st, N,
                       Np, s->XK, c->U, (T*)c->prim[c->cur], eqd(c));
    operator_at<T>(c, time);
    hipLaunchKernelGGL((k_copy_cvt<T, double>), dim3((4 * Np + 255) / 256),
                       dim3(256), 0, c->st, 4 * (long)Np, (const T*)c->ut,
                       s->RXK);
    return build_matrix<T, T>(c, time, dt, nullptr, cons, jo, false);
  };
  rc = c->prm.precision == CCFD_FP64 ? run(double(0)) : run(float(0));
  s->jacMethod = saved;
  // restore the primitive state of the resident U
  if (c->prm.precision == CCFD_FP64)
    hipLaunchKernelGGL(k_load_state<double>, cells_grid(c), dim3(256), 0, c->st,
                       N, Np, s->XK, c->U, (double*)c->prim[c->cur], eqd(c));
  else
    hipLaunchKernelGGL(k_load_state<float>, cells_grid(c), dim3(256), 0, c->st,
                       N, Np, s->XK, c->U, (float*)c->prim[c->cur], eqd(c));
  if (rc) return rc;
  const size_t kb = c->prm.precision == CCFD_FP64 ? 8 : 4;
  std::vector<unsigned char> hd(16 * (size_t)Np * kb), hb(16 * (size_t)S * Np * kb);
  if (hipMemcpyAsync(hd.data(), s->Dg, hd.size(), hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipMemcpyAsync(hb.data(), s->Blk, hb.size(), hipMemcpyDeviceToHost,
                     c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return fail(CCFD_EHIP, "Jacobian block download failed");
  auto at = [&](const std::vector<unsigned char>& v, size_t k) {
    return kb == 8 ? ((const double*)v.data())[k]
                   : (double)((const float*)v.data())[k];
  };
  for (int i = 0; i < N; ++i) {
    const int e = c->perm[i];
    const long s0 = c->slotStart[e], ns = c->slotStart[e + 1] - s0;
    for (int k = 0; k < 16; ++k) {
      D[16 * (size_t)e + k] = at(hd, (size_t)k * Np + i);
      for (int sl = 0; sl < ns; ++sl)
        Blk[16 * (size_t)(s0 + sl) + k] = at(hb, ((size_t)k * S + sl) * Np + i);
    }
  }
  return CCFD_OK;
}
