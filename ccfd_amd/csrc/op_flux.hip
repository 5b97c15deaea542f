// op_flux.hip - instantiates the operator kernels of one Riemann flux.
// Built once per flux code with -DFLUX_CODE=<1..11> so the 11 flux
// specialisations compile in parallel; each object exports
// ccfd_launch_flux_<code>_{d,f} (two-kernel path, FP64 / FP32) and
// ccfd_launch_fused_<code>_{d,f} (fused LDS paths) used by ccfd_gpu.hip,
// and ccfd_launch_jac1_<code>_{d,f} (first-order Jacobian blocks) used by
// implicit.hip.
#include "fused.hpp"

#ifndef FLUX_CODE
#error "compile with -DFLUX_CODE=<1..11>"
#endif

#define CCFD_CAT2(a, b) a##b
#define CCFD_CAT(a, b) CCFD_CAT2(a, b)

namespace ccfd {

template <typename T, int ORDER, int NS>
static void go(const OpArgs<T>& a, int nb, hipStream_t st) {
  hipLaunchKernelGGL((flux_kernel<FLUX_CODE, T, ORDER, NS>), dim3(nb),
                     dim3(kBlock), 0, st, a);
}

template <typename T>
static void dispatch(const OpArgs<T>& a, int nb, hipStream_t st) {
  if (a.order == 1) {
    if (a.ns) go<T, 1, 1>(a, nb, st); else go<T, 1, 0>(a, nb, st);
  } else {
    if (a.ns) go<T, 2, 1>(a, nb, st); else go<T, 2, 0>(a, nb, st);
  }
}

template <typename T, int LIM, int S>
static void go_fused(const OpArgs<T>& a, const FusedArgs& f, int face, int nb,
                     size_t lds, hipStream_t st, int tb) {
  if (face && tb == 2 * kBlock)
    hipLaunchKernelGGL((fused_face_kernel<FLUX_CODE, T, LIM, S, 2 * kBlock>),
                       dim3(nb), dim3(2 * kBlock), lds, st, a, f);
  else if (face)
    hipLaunchKernelGGL((fused_face_kernel<FLUX_CODE, T, LIM, S, kBlock>),
                       dim3(nb), dim3(kBlock), lds, st, a, f);
  else
    hipLaunchKernelGGL((fused_kernel<FLUX_CODE, T, LIM, S>), dim3(nb),
                       dim3(kBlock), lds, st, a, f);
}

template <typename T>
static void dispatch_fused(const OpArgs<T>& a, const FusedArgs& f, int lim,
                           int face, int nb, size_t lds, hipStream_t st,
                           int tb) {
  if (a.S == 3) {
    if (lim == 1) go_fused<T, 1, 3>(a, f, face, nb, lds, st, tb);
    else go_fused<T, 2, 3>(a, f, face, nb, lds, st, tb);
  } else {
    if (lim == 1) go_fused<T, 1, 4>(a, f, face, nb, lds, st, tb);
    else go_fused<T, 2, 4>(a, f, face, nb, lds, st, tb);
  }
}

template <typename T, typename M>
static void dispatch_jac1(const OpArgs<T>& a, const double* X, double reps,
                          double srEps, const Eq<double>& qd, M* Dg, M* Blk,
                          hipStream_t st) {
  const dim3 g((a.N + kBlock - 1) / kBlock), b(kBlock);
  if (a.ns)
    hipLaunchKernelGGL((jac1_kernel<FLUX_CODE, T, M, 1>), g, b, 0, st, a, X,
                       reps, srEps, qd, Dg, Blk);
  else
    hipLaunchKernelGGL((jac1_kernel<FLUX_CODE, T, M, 0>), g, b, 0, st, a, X,
                       reps, srEps, qd, Dg, Blk);
}

}  // namespace ccfd

void CCFD_CAT(CCFD_CAT(ccfd_launch_jac1_, FLUX_CODE), _d)(
    const ccfd::OpArgs<double>& a, const double* X, double reps, double srEps,
    const ccfd::Eq<double>& qd, double* Dg, double* Blk, hipStream_t st) {
  ccfd::dispatch_jac1<double, double>(a, X, reps, srEps, qd, Dg, Blk, st);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_jac1_, FLUX_CODE), _f)(
    const ccfd::OpArgs<float>& a, const double* X, double reps, double srEps,
    const ccfd::Eq<double>& qd, float* Dg, float* Blk, hipStream_t st) {
  ccfd::dispatch_jac1<float, float>(a, X, reps, srEps, qd, Dg, Blk, st);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_flux_, FLUX_CODE), _d)(
    const ccfd::OpArgs<double>& a, int nb, hipStream_t st) {
  ccfd::dispatch<double>(a, nb, st);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_flux_, FLUX_CODE), _f)(
    const ccfd::OpArgs<float>& a, int nb, hipStream_t st) {
  ccfd::dispatch<float>(a, nb, st);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_fused_, FLUX_CODE), _d)(
    const ccfd::OpArgs<double>& a, const ccfd::FusedArgs& f, int lim, int face,
    int nb, size_t lds, hipStream_t st, int tb) {
  ccfd::dispatch_fused<double>(a, f, lim, face, nb, lds, st, tb);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_fused_, FLUX_CODE), _f)(
    const ccfd::OpArgs<float>& a, const ccfd::FusedArgs& f, int lim, int face,
    int nb, size_t lds, hipStream_t st, int tb) {
  ccfd::dispatch_fused<float>(a, f, lim, face, nb, lds, st, tb);
}
