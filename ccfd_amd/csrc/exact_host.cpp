// exact_host.cpp - the exact functions (src/exactFunction.c:35-183) on the
// host, from the same source the device kernels use (physics.hpp), so that
// initial conditions set from an exact function (setInitialCondition,
// src/initialCondition.c:205-211) use the C library's sin, cos, tanh, exp
// and pow, as the reference does.
#include <math.h>

#include "ccfd_gpu.h"
#include "physics.hpp"

extern "C" int ccfd_exact_func(int exactFunc, const ccfd_gpu_params* p,
                               const ccfd_mesh* m, const double x[2],
                               double time, double pVar[4]) {
  if (!p || !m || !x || !pVar || exactFunc < 1 || exactFunc > 6)
    return CCFD_EINVAL;
  ccfd::Eq<double> q = {p->gam, p->gam - 1.0, 1.0 / (p->gam - 1.0), p->mu, p->Pr};
  ccfd::ExactCfg ex;
  for (int k = 0; k < 2; ++k)
    for (int v = 0; v < 4; ++v) ex.refState[k][v] = p->refState[k][v];
  ex.rp1Dinterface = p->rp1Dinterface;
  // dxRef as initMesh computes it (src/mesh.c:1923)
  ex.dxRef = sqrt(1.0 / (m->totalArea_q * m->nElems));
  ex.xMin = m->xMin; ex.xMax = m->xMax;
  ex.yMin = m->yMin; ex.yMax = m->yMax;
  ccfd::exact_func<double>(q, ex, exactFunc, x, time, pVar);
  return CCFD_OK;
}
