// operator.hpp - the two kernels of one finite-volume operator evaluation
// (ccfd fvTimeDerivative, src/finiteVolume.c:78-121) on the renumbered
// structure-of-arrays mesh, plus their launch descriptors.
//
//   recon_kernel  (2nd order only): least-squares gradient of the primitive
//                 state (src/reconstruction.c:252-275), ghost cell states
//                 (setBCatBarys, src/boundary.c:320-331), Barth-Jespersen /
//                 Venkatakrishnan limiting (src/reconstruction.c:25-192) and
//                 side states p + GP.grad (src/reconstruction.c:293-311).
//   flux_kernel   one thread per element evaluates the Riemann flux of each
//                 of its sides in the owning side's orientation
//                 (src/fluxCalculation.c:884-997, the partner takes -flux),
//                 ghost side states (setBCatSides, src/boundary.c:300-314),
//                 optional viscous flux, and sums them in the reference's
//                 side order (src/finiteVolume.c:96-120): a deterministic
//                 gather without atomics or a face-flux buffer. Optional
//                 epilogue: the low-storage RK update
//                 (src/timeDiscretization.c:380-397) and per-block partial
//                 sums of the global residual (src/analyze.c:521-539).
//
// Layout: element arrays [var][Np]; side arrays [var][slot][Np] with slot in
// the reference's per-element side order; "link" per slot encodes the
// neighbour (element << 2 | partner slot), a ghost -(g + 1), or INT_MIN
// (no side). flen < 0 marks a side that is not the face owner.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "physics.hpp"

namespace ccfd {

struct BCTable {
  BCDev bc[kMaxBC];
};

template <typename T>
struct OpArgs {
  int N, Np, S, G;
  int order, ns;
  // mesh, slot arrays [c][S][Np] / [S][Np]
  const int* __restrict__ link;
  const T* __restrict__ w;
  const T* __restrict__ gp;
  const T* __restrict__ fn;
  const T* __restrict__ flen;
  const T* __restrict__ bbv;
  const T* __restrict__ bbd;
  // element arrays [Np]
  const T* __restrict__ areaq;
  const T* __restrict__ venk;
  const double* __restrict__ bary;   // [2][Np]
  const double* __restrict__ area;   // residual weight
  // ghosts [G]
  const int* __restrict__ ghostBC;
  const double* __restrict__ ghostBary;  // [2][G]
  const int* __restrict__ ghostCell;     // interior element (renumbered)
  const int* __restrict__ ghostK;        // its slot facing the ghost
  // state
  const T* __restrict__ prim;   // [4][Np] operator input
  T* __restrict__ side;         // [4][S][Np]
  T* __restrict__ ux;           // [4][Np]
  T* __restrict__ uy;
  T* __restrict__ ghostP;       // [4][G] ghost element states (barycentres)
  T* __restrict__ ghostS;       // [4][G] ghost side states
  double* __restrict__ bcIn;    // [4][G] interior side states at the
                                // boundary sides (null: not kept)
  const T* __restrict__ src;    // [4][Np] source term (null: none)
  T* __restrict__ ut;           // [4][Np]
  // time: time = tdev ? tdev[0] + tcoef * tdev[1] : time0
  double time0;
  const double* tdev;
  double tcoef;
  // RK epilogue: U = U0 + (rkc * dt) * ut; prim_out = consPrim(U)
  int rk;
  double rkc;
  const double* dtdev;   // device dt (if null use dt0)
  double dt0;
  const double* __restrict__ U0;
  double* __restrict__ U;
  T* __restrict__ prim_out;
  int write_ut;
  int write_U;          // store U (last stage); intermediate stages do not
  // residual partial sums [4][nblocks] of area * ut^2
  double* res_part;
  Eq<T> eq;
  Eq<double> eqD;   // the same constants in FP64: the FP64 state's consPrim
                    // and the time-step limits
  ExactCfg ex;
  BCTable bct;
  // batched arena (all null for one mesh): per launched block,
  // (arena block << 9) | elements in it; instance of each arena block;
  // per-instance time and dt; per-element local dt; the arena's boundary
  // condition table (ghostBC indexes it instead of bct)
  const int* __restrict__ blockMap;
  const int* __restrict__ blockInst;
  const double* __restrict__ tInst;
  const double* __restrict__ dtInst;
  const double* __restrict__ dtLoc;
  const BCDev* __restrict__ bcp;
};

// XCD-aware block order: consecutive logical blocks (neighbouring cells
// along the space-filling curve) go to the same XCD, which dispatches
// hardware blocks round-robin over 8 XCDs, so neighbour data is shared in
// one L2.
__device__ inline int logical_block() {
  int b = blockIdx.x, nb = gridDim.x;
  int nb8 = (nb / 8) * 8;
  if (b >= nb8) return b;
  return (b % 8) * (nb8 / 8) + b / 8;
}

// can slot s of an element with link l be absent? Only in meshes with
// quadrilaterals: when every element has S = 3 sides all slots exist, and
// the checks compile away
template <int S>
__device__ inline bool no_side(int l) {
  return S != 3 && l == kNoSide;
}

// launched block -> first element and element count
struct BlockPos {
  int lb, base, n;
};
template <typename T>
__device__ inline BlockPos block_pos(const OpArgs<T>& a) {
  const int lb = logical_block();
  if (a.blockMap) {
    const int m = a.blockMap[lb];
    return {lb, (m >> 9) * kBlock, m & 511};
  }
  return {lb, lb * kBlock, min(kBlock, a.N - lb * kBlock)};
}

// stage time of element i (boundary states that depend on time)
template <typename T>
__device__ inline T op_time(const OpArgs<T>& a, int i) {
  if (a.tInst) {
    const int k = a.blockInst[i / kBlock];
    return T(a.tInst[k] + a.tcoef * a.dtInst[k]);
  }
  return a.tdev ? T(a.tdev[0] + a.tcoef * a.tdev[1]) : T(a.time0);
}

// time step of element i in the RK update. The device pointers are read
// with relaxed atomic loads so the compiler cannot fold the choice into one
// load through a selected address (which would put dt0 in private memory).
template <typename T>
__device__ inline double stage_dt(const OpArgs<T>& a, int i) {
  const double* p = a.dtLoc ? a.dtLoc + i
                  : a.dtInst ? a.dtInst + a.blockInst[i / kBlock]
                  : a.dtdev;
  if (p) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return a.dt0;
}

template <typename T>
__device__ inline BCDev op_bc(const OpArgs<T>& a, int id) {
  if (a.bcp) return a.bcp[id];
  return a.bct.bc[id];
}

// ------------------------------------------------------------ recon kernel
// SC: sides per element known at compile time (3: triangles, 4: quads /
// mixed with padding) so all neighbour loads of an element issue together
template <typename T, int LIM, int SC>
__global__ __launch_bounds__(kBlock) void recon_kernel(OpArgs<T> a) {
  const BlockPos bp = block_pos(a);
  if ((int)threadIdx.x >= bp.n) return;
  const int i = bp.base + threadIdx.x;
  const int Np = a.Np;
  constexpr int S = SC;
  T p[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) p[v] = a.prim[v * Np + i];
  int lk[S];
#pragma unroll
  for (int s = 0; s < S; ++s) lk[s] = a.link[s * Np + i];
  T gx[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  T gy[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  T umin[4], umax[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) umin[v] = umax[v] = p[v];
  // least-squares gradient and neighbour extrema, sides in reference order
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (no_side<S>(lk[s])) continue;
    T q[4];
    if (lk[s] >= 0) {
      const int j = lk[s] >> 2;
#pragma unroll
      for (int v = 0; v < 4; ++v) q[v] = a.prim[v * Np + j];
    } else {
      // ghost cell state from setBCatBarys (bc_bary_kernel)
      const int g = -lk[s] - 1;
#pragma unroll
      for (int v = 0; v < 4; ++v) q[v] = a.ghostP[v * a.G + g];
    }
    const T wx = a.w[s * Np + i], wy = a.w[(S + s) * Np + i];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const T d = q[v] - p[v];
      gx[v] += wx * d;
      gy[v] += wy * d;
      umin[v] = tmin(umin[v], q[v]);
      umax[v] = tmax(umax[v], q[v]);
    }
  }
  const T eps = LIM == 2 ? a.venk[i] : T(0.0);
  T GPx[S], GPy[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    GPx[s] = a.gp[s * Np + i];
    GPy[s] = a.gp[(S + s) * Np + i];
  }
  // limiter: one variable at a time bounds the live division temporaries
#pragma unroll 1
  for (int v = 0; v < 4; ++v) {
    const T dmin = umin[v] - p[v], dmax = umax[v] - p[v];
    T phi = T(1.0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (no_side<S>(lk[s])) continue;
      const T d = gx[v] * GPx[s] + gy[v] * GPy[s];
      phi = tmin(phi, limiter_phi<LIM, T>(d, dmax, dmin, dmax * dmax,
                                          dmin * dmin, eps));
    }
    gx[v] *= phi;
    gy[v] *= phi;
  }
  if (a.ux) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      a.ux[v * Np + i] = gx[v];
      a.uy[v * Np + i] = gy[v];
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (no_side<S>(lk[s])) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      a.side[(v * S + s) * Np + i] = p[v] + GPx[s] * gx[v] + GPy[s] * gy[v];
  }
}

// ------------------------------------------------------- boundary kernels
// setBCatBarys (src/boundary.c:320-331): ghost cell states at the mirrored
// barycentres from the interior element state; one thread per ghost side.
template <typename T>
__global__ void bc_bary_kernel(OpArgs<T> a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G || a.ghostBC[g] < 0) return;
  const int i = a.ghostCell[g], k = a.ghostK[g];
  const int Np = a.Np, S = a.S;
  T in[4], out[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) in[v] = a.prim[v * Np + i];
  T x[2] = {T(a.ghostBary[g]), T(a.ghostBary[a.G + g])};
  bc_ghost<T, false>(a.eq, op_bc(a, a.ghostBC[g]), a.ex, a.fn[k * Np + i],
              a.fn[(S + k) * Np + i], op_time(a, i), in, out, x);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.ghostP[v * a.G + g] = out[v];
}

// setBCatSides (src/boundary.c:300-314): ghost side states from the
// reconstructed interior side state (the element state at 1st order), at
// x = side midpoint.
template <typename T>
__global__ void bc_side_kernel(OpArgs<T> a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G || a.ghostBC[g] < 0) return;
  const int i = a.ghostCell[g], k = a.ghostK[g];
  const int Np = a.Np, S = a.S;
  T in[4], out[4];
  if (a.order == 1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) in[v] = a.prim[v * Np + i];
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) in[v] = a.side[(v * S + k) * Np + i];
  }
  if (a.bcIn) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.bcIn[v * a.G + g] = double(in[v]);
  }
  T x[2] = {a.gp[k * Np + i] + T(a.bary[i]),
            a.gp[(S + k) * Np + i] + T(a.bary[Np + i])};
  bc_ghost<T, false>(a.eq, op_bc(a, a.ghostBC[g]), a.ex, a.fn[k * Np + i],
              a.fn[(S + k) * Np + i], op_time(a, i), in, out, x);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.ghostS[v * a.G + g] = out[v];
}

// exact-function ghost states (boundary type 6, src/boundary.c:272-274):
// the common boundary kernels leave them out (they would all carry the
// exact Riemann solver's stack); this kernel runs after them where a mesh
// has such boundaries. side = 0: at the mirrored barycentre (ghostP, as
// bc_bary_kernel), 1: at the side midpoint (ghostS, as bc_side_kernel)
template <typename T>
__global__ void bc_exact_kernel(OpArgs<T> a, int side) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G || a.ghostBC[g] < 0) return;
  const BCDev b = op_bc(a, a.ghostBC[g]);
  if (b.type != 6) return;
  const int i = a.ghostCell[g], k = a.ghostK[g];
  const int Np = a.Np, S = a.S;
  T x[2], out[4];
  if (side) {
    x[0] = a.gp[k * Np + i] + T(a.bary[i]);
    x[1] = a.gp[(S + k) * Np + i] + T(a.bary[Np + i]);
  } else {
    x[0] = T(a.ghostBary[g]);
    x[1] = T(a.ghostBary[a.G + g]);
  }
  bc_ghost_exact<T>(a.eq, a.ex, b.exactFunc, x, op_time(a, i), out);
  T* dst = side ? a.ghostS : a.ghostP;
#pragma unroll
  for (int v = 0; v < 4; ++v) dst[v * a.G + g] = out[v];
}

// ------------------------------------------------------------- flux kernel
template <typename T>
__device__ inline void block_sum4(double v[4], double* res_part, int lb,
                                  int nblocks) {
  __shared__ double red[4][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double x = v[k];
    // fixed butterfly order: deterministic for a given block
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[k][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[threadIdx.x][w];
    res_part[threadIdx.x * nblocks + lb] = s;
  }
}

// dU/dt of element i from its gathered side fluxes (src/finiteVolume.c:
// 106-119), then the optional RK stage update (src/timeDiscretization.c:
// 380-397; the conservative state is stored only where it is needed after
// the stage, its primitive form always) and the residual contribution
// area * ut^2 (src/analyze.c:521-539)
// element data the epilogue reads that does not depend on the fluxes; the
// fused kernels load it before their first barrier so its latency overlaps
// the reconstruction
template <typename T>
struct StagePre {
  double u0[4];
  T aq;
};

template <typename T>
__device__ inline StagePre<T> stage_preload(const OpArgs<T>& a, int i) {
  StagePre<T> p;
  p.aq = a.areaq[i];
#pragma unroll
  for (int v = 0; v < 4; ++v) p.u0[v] = a.rk ? a.U0[v * a.Np + i] : 0.0;
  return p;
}

template <typename T>
__device__ inline void stage_epilogue(const OpArgs<T>& a, int i,
                                      const T acc[4], double r2[4],
                                      const StagePre<T>& pre) {
  const int Np = a.Np;
  const T aq = pre.aq;
  T ut[4];
  if (a.src) {
    // src/finiteVolume.c:116-119 with the source term
#pragma unroll
    for (int v = 0; v < 4; ++v) ut[v] = (a.src[v * Np + i] - acc[v]) * aq;
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) ut[v] = (T(0.0) - acc[v]) * aq;
  }
  if (a.write_ut) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.ut[v * Np + i] = ut[v];
  }
  if (a.rk) {
    const double c = a.rkc * stage_dt(a, i);
    double Uc[4], pn[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) Uc[v] = pre.u0[v] + c * double(ut[v]);
    // intermediate stages only feed the next stage's primitive state
    if (a.write_U) {
#pragma unroll
      for (int v = 0; v < 4; ++v) a.U[v * Np + i] = Uc[v];
    }
    cons_prim<double>(a.eqD, Uc, pn);
#pragma unroll
    for (int v = 0; v < 4; ++v) a.prim_out[v * Np + i] = T(pn[v]);
  }
  if (a.res_part) {
    const double ar = a.area[i];
#pragma unroll
    for (int v = 0; v < 4; ++v) r2[v] = ar * double(ut[v]) * double(ut[v]);
  }
}

template <typename T>
__device__ inline void stage_epilogue(const OpArgs<T>& a, int i,
                                      const T acc[4], double r2[4]) {
  stage_epilogue<T>(a, i, acc, r2, stage_preload(a, i));
}

template <int CODE, typename T, int ORDER, int NS>
__global__ __launch_bounds__(kBlock) void flux_kernel(OpArgs<T> a) {
  const BlockPos bp = block_pos(a);
  const int lb = bp.lb;
  const int i = bp.base + threadIdx.x;
  const bool active = (int)threadIdx.x < bp.n;
  const int Np = a.Np, S = a.S;
  double r2[4] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
    T pc[4];
    if (ORDER == 1 || NS) {
#pragma unroll
      for (int v = 0; v < 4; ++v) pc[v] = a.prim[v * Np + i];
    }
    // one side at a time keeps register pressure (and occupancy) in check
#pragma unroll 1
    for (int s = 0; s < S; ++s) {
      const int lk = a.link[s * Np + i];
      if (lk == kNoSide) continue;
      T own[4], oth[4];
      if (ORDER == 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) own[v] = pc[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) own[v] = a.side[(v * S + s) * Np + i];
      }
      const T nx = a.fn[s * Np + i], ny = a.fn[(S + s) * Np + i];
      T fl = a.flen[s * Np + i];
      const bool owner = fl > T(0.0);
      const T len = owner ? fl : -fl;
      int j = -1, g = -1;
      if (lk >= 0) {
        j = lk >> 2;
        const int ps = lk & 3;
        if (ORDER == 1) {
#pragma unroll
          for (int v = 0; v < 4; ++v) oth[v] = a.prim[v * Np + j];
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) oth[v] = a.side[(v * S + ps) * Np + j];
        }
      } else {
        // ghost side state from setBCatSides (bc_side_kernel)
        g = -lk - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) oth[v] = a.ghostS[v * a.G + g];
      }
      const T* L = owner ? own : oth;
      const T* R = owner ? oth : own;
      T Lr[4] = {L[0], nx * L[1] + ny * L[2], -ny * L[1] + nx * L[2], L[3]};
      T Rr[4] = {R[0], nx * R[1] + ny * R[2], -ny * R[1] + nx * R[2], R[3]};
      T F[4];
      conv_flux_t<CODE, T>(a.eq, Lr, Rr, F);
      T f[4];
      f[0] = F[0];
      f[1] = nx * F[1] - ny * F[2];
      f[2] = ny * F[1] + nx * F[2];
      f[3] = F[3];
      if (NS) {
        // owner element "a" and connection element "b" of the face
        T pa_[4], pb_[4], uxa[4], uya[4], uxb[4], uyb[4];
        const int ea = owner ? i : j;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          pa_[v] = owner ? pc[v] : a.prim[v * Np + ea];
          uxa[v] = ORDER == 2 ? a.ux[v * Np + ea] : T(0.0);
          uya[v] = ORDER == 2 ? a.uy[v * Np + ea] : T(0.0);
        }
        if (g >= 0) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            pb_[v] = a.ghostP[v * a.G + g];
            uxb[v] = T(0.0);
            uyb[v] = T(0.0);
          }
        } else {
          const int eb = owner ? j : i;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            pb_[v] = owner ? a.prim[v * Np + eb] : pc[v];
            uxb[v] = ORDER == 2 ? a.ux[v * Np + eb] : T(0.0);
            uyb[v] = ORDER == 2 ? a.uy[v * Np + eb] : T(0.0);
          }
        }
        T bvx = a.bbv[s * Np + i], bvy = a.bbv[(S + s) * Np + i];
        if (!owner) {
          bvx = -bvx;
          bvy = -bvy;
        }
        const T bd = a.bbd[s * Np + i];
        const T bb0 = bvx / bd, bb1 = bvy / bd;
        T st[4], gxf[4], gyf[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          st[v] = T(0.5) * (R[v] + L[v]);
          T gmx = T(0.5) * (uxa[v] + uxb[v]);
          T gmy = T(0.5) * (uya[v] + uyb[v]);
          T cr = gmx * bb0 + gmy * bb1 - (pb_[v] - pa_[v]) / bd;
          gxf[v] = gmx - cr * bb0;
          gyf[v] = gmy - cr * bb1;
        }
        T fx[4], fy[4];
        diffusion_flux<T>(a.eq, st, gxf, gyf, fx, fy);
#pragma unroll
        for (int v = 0; v < 4; ++v) f[v] -= (fx[v] * nx + fy[v] * ny);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        f[v] *= len;
        if (!owner) f[v] = -f[v];
        acc[v] += f[v];
      }
    }
    stage_epilogue<T>(a, i, acc, r2);
  }
  if (a.res_part) block_sum4<T>(r2, a.res_part, lb, gridDim.x);
}

// ------------------------------------- first-order Jacobian, face by face
//
// The blocks of dR/dU (R the first-order operator, U the conservative
// state) that the coloured finite differences of the preconditioner record
// - the diagonal block of each element and the block of each side towards
// its face neighbour - evaluated element by element in one launch: R_i at
// the element's state, then with each variable of the element (diagonal
// block) or of one neighbour (that side's block) perturbed by reps, each
// recorded as (R_i(perturbed) - R_i) / reps. A first-order R_i depends on
// its own state and its face neighbours' only, so this is exactly what the
// coloured differences record, with the same operations (flux_kernel's
// first-order side flux, summed in side order), hence the same values.
//
// One side's contribution to the flux sum of element i at first order:
// the Riemann flux of the owning orientation (src/fluxCalculation.c:884-
// 997), minus the viscous flux with the neighbour-difference face gradient
// (cell gradients are zero at first order), times the length, negated on
// the side that does not own the face.
template <int CODE, typename T, int NS>
__device__ inline void side_flux_o1(const OpArgs<T>& a, int i, int s, int lk,
                                    const T own[4], const T oth[4],
                                    T f[4]) {
  const int Np = a.Np, S = a.S;
  const T nx = a.fn[s * Np + i], ny = a.fn[(S + s) * Np + i];
  const T fl = a.flen[s * Np + i];
  const bool owner = fl > T(0.0);
  const T len = owner ? fl : -fl;
  const T* L = owner ? own : oth;
  const T* R = owner ? oth : own;
  T Lr[4] = {L[0], nx * L[1] + ny * L[2], -ny * L[1] + nx * L[2], L[3]};
  T Rr[4] = {R[0], nx * R[1] + ny * R[2], -ny * R[1] + nx * R[2], R[3]};
  T F[4];
  conv_flux_t<CODE, T>(a.eq, Lr, Rr, F);
  f[0] = F[0];
  f[1] = nx * F[1] - ny * F[2];
  f[2] = ny * F[1] + nx * F[2];
  f[3] = F[3];
  if (NS) {
    T pa_[4], pb_[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) pa_[v] = owner ? own[v] : oth[v];
    if (lk < 0) {
      const int g = -lk - 1;
#pragma unroll
      for (int v = 0; v < 4; ++v) pb_[v] = a.ghostP[v * a.G + g];
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) pb_[v] = owner ? oth[v] : own[v];
    }
    T bvx = a.bbv[s * Np + i], bvy = a.bbv[(S + s) * Np + i];
    if (!owner) {
      bvx = -bvx;
      bvy = -bvy;
    }
    const T bd = a.bbd[s * Np + i];
    const T bb0 = bvx / bd, bb1 = bvy / bd;
    T st[4], gxf[4], gyf[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      st[v] = T(0.5) * (R[v] + L[v]);
      T gmx = T(0.5) * (T(0.0) + T(0.0));
      T gmy = T(0.5) * (T(0.0) + T(0.0));
      T cr = gmx * bb0 + gmy * bb1 - (pb_[v] - pa_[v]) / bd;
      gxf[v] = gmx - cr * bb0;
      gyf[v] = gmy - cr * bb1;
    }
    T fx[4], fy[4];
    diffusion_flux<T>(a.eq, st, gxf, gyf, fx, fy);
#pragma unroll
    for (int v = 0; v < 4; ++v) f[v] -= (fx[v] * nx + fy[v] * ny);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    f[v] *= len;
    if (!owner) f[v] = -f[v];
  }
}

// state of element j from the FP64 conservative state X, perturbed by reps
// in variable iv (iv < 0: none), in the operator's arithmetic
template <typename T>
__device__ inline void jac_state(const double* __restrict__ X, int Np, int j,
                                 int iv, double reps, const Eq<double>& qd,
                                 T p[4]) {
  double u[4], q[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    u[v] = X[v * Np + j];
    if (v == iv) u[v] += reps;
  }
  cons_prim<double>(qd, u, q);
#pragma unroll
  for (int v = 0; v < 4; ++v) p[v] = T(q[v]);
}

// ghost side state of element i's boundary side at first order
// (setBCatSides on the element state); exact-function boundaries do not
// depend on it and are read from ghostS
template <typename T>
__device__ inline void jac_ghost(const OpArgs<T>& a, int i, int s, int g,
                                 const T in[4], T out[4]) {
  const int Np = a.Np, S = a.S;
  const BCDev b = op_bc(a, a.ghostBC[g]);
  if (b.type == 6) {
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v] = a.ghostS[v * a.G + g];
    return;
  }
  T x[2] = {a.gp[s * Np + i] + T(a.bary[i]),
            a.gp[(S + s) * Np + i] + T(a.bary[Np + i])};
  T inc[4] = {in[0], in[1], in[2], in[3]};
  bc_ghost<T, false>(a.eq, b, a.ex, a.fn[s * Np + i], a.fn[(S + s) * Np + i],
                     op_time(a, i), inc, out, x);
}

template <int CODE, typename T, typename M, int NS>
__global__ __launch_bounds__(kBlock) void jac1_kernel(
    OpArgs<T> a, const double* __restrict__ X, double reps, double srEps,
    Eq<double> qd, M* __restrict__ Dg, M* __restrict__ Blk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const int Np = a.Np, S = a.S;
  const T aq = a.areaq[i];
  T p0[4];
  jac_state<T>(X, Np, i, -1, reps, qd, p0);
  // base side fluxes and R_i
  T fb[kMaxSlots][4], R0[4];
  T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
  for (int s = 0; s < S; ++s) {
    const int lk = a.link[s * Np + i];
    if (lk == kNoSide) continue;
    T oth[4];
    if (lk >= 0) jac_state<T>(X, Np, lk >> 2, -1, reps, qd, oth);
    else jac_ghost<T>(a, i, s, -lk - 1, p0, oth);
    side_flux_o1<CODE, T, NS>(a, i, s, lk, p0, oth, fb[s]);
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += fb[s][v];
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) R0[v] = (T(0.0) - acc[v]) * aq;
  for (int iv = 0; iv < 4; ++iv) {
    // diagonal block: the element's own variable iv perturbed
    T p1[4];
    jac_state<T>(X, Np, i, iv, reps, qd, p1);
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] = T(0.0);
    for (int s = 0; s < S; ++s) {
      const int lk = a.link[s * Np + i];
      if (lk == kNoSide) continue;
      T oth[4], f[4];
      if (lk >= 0) {
        const int j = lk >> 2;
        jac_state<T>(X, Np, j, j == i ? iv : -1, reps, qd, oth);
      } else {
        jac_ghost<T>(a, i, s, -lk - 1, p1, oth);
      }
      side_flux_o1<CODE, T, NS>(a, i, s, lk, p1, oth, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] += f[v];
    }
#pragma unroll
    for (int jv = 0; jv < 4; ++jv) {
      const T r = (T(0.0) - acc[jv]) * aq;
      Dg[(4 * jv + iv) * Np + i] =
          M((double(r) - double(R0[jv])) * srEps);
    }
    // side blocks: the face neighbour's variable iv perturbed
    for (int s = 0; s < S; ++s) {
      const int lk = a.link[s * Np + i];
      if (lk < 0 || (lk >> 2) == i) continue;
      T q1[4], f[4];
      jac_state<T>(X, Np, lk >> 2, iv, reps, qd, q1);
      side_flux_o1<CODE, T, NS>(a, i, s, lk, p0, q1, f);
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[v] = T(0.0);
      for (int s2 = 0; s2 < S; ++s2) {
        const int l2 = a.link[s2 * Np + i];
        if (l2 == kNoSide) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] += s2 == s ? f[v] : fb[s2][v];
      }
#pragma unroll
      for (int jv = 0; jv < 4; ++jv) {
        const T r = (T(0.0) - acc[jv]) * aq;
        Blk[((size_t)(4 * jv + iv) * S + s) * Np + i] =
            M((double(r) - double(R0[jv])) * srEps);
      }
    }
  }
}

// the partials of block_sum4 for a tile of TB threads covering TB / kBlock
// layout blocks from element `base`: one partial per layout block that
// holds elements, summed exactly as block_sum4 sums a block of kBlock
// threads (so the residual does not depend on the tile)
template <int TB>
__device__ inline void block_sum4_tile(double v[4], double* res_part, int base,
                                       int N, int nblocks) {
  __shared__ double red[4][TB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[k][wid] = x;
  }
  __syncthreads();
  constexpr int W = kBlock / 64, H = TB / kBlock;
  if (threadIdx.x < 4 * H) {
    const int v = threadIdx.x & 3, hb = threadIdx.x >> 2;
    if (base + hb * kBlock < N) {
      double s = 0.0;
      for (int w = 0; w < W; ++w) s += red[v][hb * W + w];
      res_part[v * nblocks + base / kBlock + hb] = s;
    }
  }
}

// launch entry of one flux code (defined in op_flux.hip, one object per code)
template <typename T>
void launch_flux_code(int code, const OpArgs<T>& a, int nblocks,
                      hipStream_t st);
// the first-order Jacobian blocks above (M: the matrix type)
template <typename T, typename M>
void launch_jac1_code(int code, const OpArgs<T>& a, const double* X,
                      double reps, double srEps, const Eq<double>& qd, M* Dg,
                      M* Blk, hipStream_t st);

}  // namespace ccfd
