// implicit.hip - placeholder until the JFNK path lands
#include "ctx.hpp"
void ccfd::implicit_free(ccfd_gpu_ctx* c) { (void)c; }
extern "C" int ccfd_gpu_implicit_step(ccfd_gpu_ctx* c, const ccfd_gpu_implicit_cfg* cfg,
    double t, double dt, double res[4], int* nNewton, int* nGMRES) {
  (void)c; (void)cfg; (void)t; (void)dt; (void)res; (void)nNewton; (void)nGMRES;
  return ccfd::fail(CCFD_ENOTSUP, "implicit path not built");
}
