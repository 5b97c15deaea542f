// op_flux.hip - instantiates the operator kernels of one Riemann flux.
// Built once per flux code with -DFLUX_CODE=<1..11> so the 11 flux
// specialisations compile in parallel; each object exports
// ccfd_launch_flux_<code>_{d,f} (FP64 / FP32) used by ccfd_gpu.hip.
#include "operator.hpp"

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

}  // namespace ccfd

void CCFD_CAT(CCFD_CAT(ccfd_launch_flux_, FLUX_CODE), _d)(
    const ccfd::OpArgs<double>& a, int nb, hipStream_t st) {
  ccfd::dispatch<double>(a, nb, st);
}

void CCFD_CAT(CCFD_CAT(ccfd_launch_flux_, FLUX_CODE), _f)(
    const ccfd::OpArgs<float>& a, int nb, hipStream_t st) {
  ccfd::dispatch<float>(a, nb, st);
}
