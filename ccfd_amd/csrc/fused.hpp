// fused.hpp - one-kernel 2nd-order operator evaluation with the side states
// held in LDS (Euler; ccfd fvTimeDerivative, src/finiteVolume.c:78-121).
//
// Each block owns kBlock consecutive elements of the space-filling-curve
// order. Phase 1 reconstructs the limited side states (least-squares
// gradient, limiter, p + GP.grad: src/reconstruction.c:198-311) of the owned
// elements and of the block's halo - the elements outside the block that
// share a face with it, listed per block at mesh set-up - into LDS. Phase 2
// evaluates each owned element's side fluxes from LDS in the owning side's
// orientation (src/fluxCalculation.c:884-997), gathers them in the
// reference's side order and runs the optional RK stage / residual epilogue,
// exactly as flux_kernel does. The side states never cross HBM: compared
// with recon_kernel + flux_kernel this drops their write and both reads
// (own and neighbour); halo elements are reconstructed by every block that
// borders them, extra arithmetic the FP64 units absorb.
//
// The arithmetic is the same expression for expression as the two-kernel
// path, so results are bit-identical to it. Ghost element states come from
// bc_bary_kernel and ghost side states from bc_side_recon_kernel, both
// launched before the fused kernel.
#pragma once

#include "operator.hpp"

namespace ccfd {

struct FusedArgs {
  // halo of arena block b: haloCells[haloStart[b] .. haloEnd[b]), ascending
  // element ids (one mesh: haloEnd = haloStart + 1)
  const int* __restrict__ haloStart;
  const int* __restrict__ haloEnd;
  const int* __restrict__ haloCells;
  int tile;                           // LDS columns: kBlock + max halo
};

// limited gradient (gx, gy) of element i, its state p, side offsets GP and
// links (src/reconstruction.c:252-290)
template <typename T, int LIM, int S>
__device__ inline void limited_gradient(const OpArgs<T>& a, int i, T p[4],
                                        int lk[S], T gx[4], T gy[4], T GPx[S],
                                        T GPy[S]) {
  const int Np = a.Np;
#pragma unroll
  for (int v = 0; v < 4; ++v) p[v] = a.prim[v * Np + i];
#pragma unroll
  for (int s = 0; s < S; ++s) lk[s] = a.link[s * Np + i];
  T umin[4], umax[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    gx[v] = gy[v] = T(0.0);
    umin[v] = umax[v] = p[v];
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (no_side<S>(lk[s])) continue;
    T q[4];
    if (lk[s] >= 0) {
      const int j = lk[s] >> 2;
#pragma unroll
      for (int v = 0; v < 4; ++v) q[v] = a.prim[v * Np + j];
    } else {
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
#pragma unroll
  for (int s = 0; s < S; ++s) {
    GPx[s] = a.gp[s * Np + i];
    GPy[s] = a.gp[(S + s) * Np + i];
  }
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
}

// limited side states of element i into LDS column col; returns its links
template <typename T, int LIM, int S>
__device__ inline void fused_recon(const OpArgs<T>& a, int i, int col,
                                   int tile, T* __restrict__ lds, int lk[S]) {
  T p[4], gx[4], gy[4], GPx[S], GPy[S];
  limited_gradient<T, LIM, S>(a, i, p, lk, gx, gy, GPx, GPy);
#pragma unroll
  for (int s = 0; s < S; ++s) {
#pragma unroll
    for (int v = 0; v < 4; ++v)
      lds[(v * S + s) * tile + col] = p[v] + GPx[s] * gx[v] + GPy[s] * gy[v];
  }
}

// ghost side states for the fused operator (setBCatSides,
// src/boundary.c:300-314): one thread per ghost side reconstructs its
// interior element and applies the boundary condition at the side midpoint
template <typename T, int LIM, int S>
__global__ void bc_side_recon_kernel(OpArgs<T> a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.G || a.ghostBC[g] < 0) return;
  const int i = a.ghostCell[g], k = a.ghostK[g];
  const int Np = a.Np;
  T p[4], gx[4], gy[4], GPx[S], GPy[S];
  int lk[S];
  limited_gradient<T, LIM, S>(a, i, p, lk, gx, gy, GPx, GPy);
  T in[4], out[4];
  T gpx = GPx[0], gpy = GPy[0];
#pragma unroll
  for (int s = 1; s < S; ++s)
    if (s == k) { gpx = GPx[s]; gpy = GPy[s]; }
#pragma unroll
  for (int v = 0; v < 4; ++v) in[v] = p[v] + gpx * gx[v] + gpy * gy[v];
  if (a.bcIn) {
#pragma unroll
    for (int v = 0; v < 4; ++v) a.bcIn[v * a.G + g] = double(in[v]);
  }
  T x[2] = {gpx + T(a.bary[i]), gpy + T(a.bary[Np + i])};
  bc_ghost<T, false>(a.eq, op_bc(a, a.ghostBC[g]), a.ex, a.fn[k * Np + i],
              a.fn[(S + k) * Np + i], op_time(a, i), in, out, x);
#pragma unroll
  for (int v = 0; v < 4; ++v) a.ghostS[v * a.G + g] = out[v];
}

// LDS column of neighbour element j: its offset in the tile of TB
// elements, or TB + its position in the tile's ascending halo list
template <int TB = kBlock>
__device__ inline int tile_column(int j, int base, const int* hcell, int nh) {
  int col = j - base;
  if (col < 0 || col >= TB) {
    int lo = 0, hi = nh - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (hcell[mid] < j) lo = mid + 1; else hi = mid;
    }
    col = TB + lo;
  }
  return col;
}

// exclusive prefix sum of x over the block of TB threads; *total gets the
// block sum
template <int TB = kBlock>
__device__ inline int block_scan_excl(int x, int* wsum, int* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int inc = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < TB / 64; ++k) {
    const int v = wsum[k];
    if (k < w) pre += v;
    tot += v;
  }
  *total = tot;
  return pre + inc - x;
}

template <int CODE, typename T, int LIM, int S>
__global__ __launch_bounds__(kBlock) void fused_kernel(OpArgs<T> a,
                                                       FusedArgs f) {
  extern __shared__ __align__(16) unsigned char fused_smem[];
  T* lds = reinterpret_cast<T*>(fused_smem);
  const int tile = f.tile;
  int* hcell = reinterpret_cast<int*>(lds + 4 * S * tile);
  const BlockPos bp = block_pos(a);
  const int lb = bp.lb, base = bp.base;
  const int i = base + threadIdx.x;
  const bool active = (int)threadIdx.x < bp.n;
  const int Np = a.Np;
  const int h0 = f.haloStart[base / kBlock];
  const int nh = f.haloEnd[base / kBlock] - h0;

  StagePre<T> pre{};
  if (active) pre = stage_preload(a, i);

  // phase 1: owned, then halo side states into LDS
  int lk[S];
  if (active) fused_recon<T, LIM, S>(a, i, threadIdx.x, tile, lds, lk);
  // halo elements on the last threads: the first threads take the extra
  // list entries of the flux phase
  for (int h = kBlock - 1 - (int)threadIdx.x; h < nh; h += kBlock) {
    const int e = f.haloCells[h0 + h];
    hcell[h] = e;
    int lkh[S];
    fused_recon<T, LIM, S>(a, e, kBlock + h, tile, lds, lkh);
  }
  __syncthreads();

  // phase 2: side fluxes from LDS, gathered in side order
  double r2[4] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
#pragma unroll 1
    for (int s = 0; s < S; ++s) {
      const int l = lk[s];
      if (no_side<S>(l)) continue;
      const T nx = a.fn[s * Np + i], ny = a.fn[(S + s) * Np + i];
      const T fl = a.flen[s * Np + i];
      const bool owner = fl > T(0.0);
      const T len = owner ? fl : -fl;
      const int c = threadIdx.x;
      T L[4], R[4];   // owning side left
      if (l >= 0) {
        const int col = tile_column(l >> 2, base, hcell, nh), ps = l & 3;
        const int cL = owner ? c : col, sL = owner ? s : ps;
        const int cR = owner ? col : c, sR = owner ? ps : s;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          L[v] = lds[(v * S + sL) * tile + cL];
          R[v] = lds[(v * S + sR) * tile + cR];
        }
      } else {
        // ghost side state from bc_side_recon_kernel; the boundary side
        // owns its face
        const int g = -l - 1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          L[v] = lds[(v * S + s) * tile + c];
          R[v] = a.ghostS[v * a.G + g];
        }
      }
      T Lr[4] = {L[0], nx * L[1] + ny * L[2], -ny * L[1] + nx * L[2], L[3]};
      T Rr[4] = {R[0], nx * R[1] + ny * R[2], -ny * R[1] + nx * R[2], R[3]};
      T F[4];
      conv_flux_t<CODE, T>(a.eq, Lr, Rr, F);
      T fv[4];
      fv[0] = F[0];
      fv[1] = nx * F[1] - ny * F[2];
      fv[2] = ny * F[1] + nx * F[2];
      fv[3] = F[3];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        fv[v] *= len;
        if (!owner) fv[v] = -fv[v];
        acc[v] += fv[v];
      }
    }
    stage_epilogue<T>(a, i, acc, r2, pre);
  }
  if (a.res_part) block_sum4<T>(r2, a.res_part, lb, gridDim.x);
}

// Face-once form of fused_kernel: after phase 1, the block lists the sides
// whose face flux it evaluates - each owning side of an element of the
// block, and each non-owning side whose owner lies outside the block - so
// every face inside the block is evaluated once instead of once per
// element. Each flux (owner orientation, times the face length) overwrites
// the listed side's own LDS slot, which no other list entry reads; phase 3
// gathers them per element in the reference's side order with the sign of
// the side (src/fluxCalculation.c:993-996), the same sums as fused_kernel.
//
// TB: elements (and threads) of a block's tile, a multiple of the layout's
// block size. A wider tile of one mesh (TB = 2 blocks) has a smaller halo
// per element; its residual partials are still written per layout block,
// with the sums of the 256-thread form, so every result is the same.
template <int CODE, typename T, int LIM, int S, int TB>
__global__ __launch_bounds__(TB) void fused_face_kernel(OpArgs<T> a,
                                                        FusedArgs f) {
  extern __shared__ __align__(16) unsigned char fused_smem[];
  __shared__ int wsum[TB / 64];
  T* lds = reinterpret_cast<T*>(fused_smem);
  const int tile = f.tile;
  int* hcell = reinterpret_cast<int*>(lds + 4 * S * tile);
  unsigned short* ent =
      reinterpret_cast<unsigned short*>(hcell + (tile - TB));
  BlockPos bp;
  if (TB == kBlock) {
    bp = block_pos(a);
  } else {   // one mesh (no arena block map)
    const int lbt = logical_block();
    bp = {lbt, lbt * TB, min(TB, a.N - lbt * TB)};
  }
  const int lb = bp.lb, base = bp.base;
  const int i = base + threadIdx.x;
  const bool active = (int)threadIdx.x < bp.n;
  const int Np = a.Np;
  const int h0 = f.haloStart[base / TB];
  const int nh = f.haloEnd[base / TB] - h0;

  // phase 1: side states into LDS; which of the element's sides it owns
  // and which face fluxes this block evaluates for it
  int lk[S];
  unsigned ownMask = 0, faceMask = 0;
  StagePre<T> pre{};
  if (active) {
    pre = stage_preload(a, i);
    fused_recon<T, LIM, S>(a, i, threadIdx.x, tile, lds, lk);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (no_side<S>(lk[s])) continue;
      const bool owner = a.flen[s * Np + i] > T(0.0);
      const int jc = (lk[s] >> 2) - base;
      if (owner) ownMask |= 1u << s;
      if (owner || (lk[s] >= 0 && (jc < 0 || jc >= TB)))
        faceMask |= 1u << s;
    }
  }
  // halo elements on the last threads: the first threads take the extra
  // list entries of the flux phase
  for (int h = TB - 1 - (int)threadIdx.x; h < nh; h += TB) {
    const int e = f.haloCells[h0 + h];
    hcell[h] = e;
    int lkh[S];
    fused_recon<T, LIM, S>(a, e, TB + h, tile, lds, lkh);
  }
  int nf;
  int k = block_scan_excl<TB>(__popc(faceMask), wsum, &nf);
#pragma unroll
  for (int s = 0; s < S; ++s)
    if (faceMask & (1u << s)) ent[k++] = (unsigned short)((threadIdx.x << 2) | s);
  __syncthreads();

  // phase 2: one flux per listed side, into that side's LDS slot
  for (int e = threadIdx.x; e < nf; e += TB) {
    const int c = ent[e] >> 2, s = ent[e] & 3;
    const int ie = base + c;
    const int l = a.link[s * Np + ie];
    const T nx = a.fn[s * Np + ie], ny = a.fn[(S + s) * Np + ie];
    const T fl = a.flen[s * Np + ie];
    const bool owner = fl > T(0.0);
    const T len = owner ? fl : -fl;
    T L[4], R[4];   // owning side left
    if (l >= 0) {
      const int col = tile_column<TB>(l >> 2, base, hcell, nh), ps = l & 3;
      const int cL = owner ? c : col, sL = owner ? s : ps;
      const int cR = owner ? col : c, sR = owner ? ps : s;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        L[v] = lds[(v * S + sL) * tile + cL];
        R[v] = lds[(v * S + sR) * tile + cR];
      }
    } else {
      // a boundary side owns its face; the ghost state is on the right
      const int g = -l - 1;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        L[v] = lds[(v * S + s) * tile + c];
        R[v] = a.ghostS[v * a.G + g];
      }
    }
    T Lr[4] = {L[0], nx * L[1] + ny * L[2], -ny * L[1] + nx * L[2], L[3]};
    T Rr[4] = {R[0], nx * R[1] + ny * R[2], -ny * R[1] + nx * R[2], R[3]};
    T F[4];
    conv_flux_t<CODE, T>(a.eq, Lr, Rr, F);
    T fv[4];
    fv[0] = F[0];
    fv[1] = nx * F[1] - ny * F[2];
    fv[2] = ny * F[1] + nx * F[2];
    fv[3] = F[3];
#pragma unroll
    for (int v = 0; v < 4; ++v) lds[(v * S + s) * tile + c] = fv[v] * len;
  }
  __syncthreads();

  // phase 3: gather in side order
  double r2[4] = {0.0, 0.0, 0.0, 0.0};
  if (active) {
    T acc[4] = {T(0.0), T(0.0), T(0.0), T(0.0)};
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (no_side<S>(lk[s])) continue;
      int slot = s, col = threadIdx.x;
      if (!(faceMask & (1u << s))) {
        // evaluated by the owning side, an element of this block
        slot = lk[s] & 3;
        col = (lk[s] >> 2) - base;
      }
      const bool owner = ownMask & (1u << s);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const T g = lds[(v * S + slot) * tile + col];
        acc[v] += owner ? g : -g;
      }
    }
    stage_epilogue<T>(a, i, acc, r2, pre);
  }
  if (a.res_part) {
    if (TB == kBlock) block_sum4<T>(r2, a.res_part, lb, gridDim.x);
    else block_sum4_tile<TB>(r2, a.res_part, base, a.N, a.Np / kBlock);
  }
}

// dynamic LDS bytes of one fused block of tb elements: side states of the
// tile columns (tb + the largest halo), the halo id list and the face list
inline size_t fused_lds_bytes(size_t tsize, int S, int tile, int tb = kBlock) {
  return 4 * (size_t)S * tile * tsize + (size_t)(tile - tb) * 4 +
         (size_t)S * tb * 2;
}

// launch entry of one flux code (defined in op_flux.hip)
// face: 0 fused_kernel, 1 fused_face_kernel; tb: tile elements (kBlock, or
// 2 kBlock for the face-once kernel of one mesh)
template <typename T>
void launch_fused_code(int code, const OpArgs<T>& a, const FusedArgs& f,
                       int limiter, int face, int nblocks, size_t lds,
                       hipStream_t st, int tb = kBlock);

}  // namespace ccfd
