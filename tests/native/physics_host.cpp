// physics_host.cpp - host build of the device physics (ccfd_amd/csrc/
// physics.hpp) for CPU tests: checks the transcription of the fluxes,
// boundary states and limiters against the oracle without a GPU.
// Test infrastructure; built by ccfd_amd/build.py into tests/native/.
#include "physics.hpp"

using namespace ccfd;

extern "C" void phys_riemann(int code, int fp32, double gam, long n,
                             const double* L, const double* R, double* F) {
  for (long i = 0; i < n; ++i) {
    if (fp32) {
      Eq<float> q = {float(gam), float(gam - 1.0), float(1.0 / (gam - 1.0)), 0.f, 0.72f};
      float l[4], r[4], f[4];
      for (int v = 0; v < 4; ++v) { l[v] = float(L[4 * i + v]); r[v] = float(R[4 * i + v]); }
      conv_flux<float>(q, code, l, r, f);
      for (int v = 0; v < 4; ++v) F[4 * i + v] = f[v];
    } else {
      Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), 0.0, 0.72};
      conv_flux<double>(q, code, L + 4 * i, R + 4 * i, F + 4 * i);
    }
  }
}

extern "C" void phys_bc(int type, const double* bcp, double gam, double mu,
                        long n, const double* nrm, const double* in,
                        double* out) {
  Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), mu, 0.72};
  BCDev b;
  b.type = type;
  b.exactFunc = 0;
  for (int v = 0; v < 4; ++v) b.pVar[v] = bcp[v];
  ExactCfg ex = {};
  double x[2] = {0.0, 0.0};
  for (long i = 0; i < n; ++i)
    bc_ghost<double>(q, b, ex, nrm[2 * i], nrm[2 * i + 1], 0.0, in + 4 * i,
                     out + 4 * i, x);
}

// state conversions: mode 0 prim_cons, 1 cons_prim, 2 cons_char (4 -> 3,
// reference state ref), 3 char_cons (3 -> 4; slot 2 left 0)
extern "C" void phys_conv(int mode, double gam, long n, const double* in,
                          double* out, const double* ref) {
  Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), 0.0, 0.72};
  for (long i = 0; i < n; ++i) {
    if (mode == 0) prim_cons<double>(q, in + 4 * i, out + 4 * i);
    else if (mode == 1) cons_prim<double>(q, in + 4 * i, out + 4 * i);
    else if (mode == 2) cons_char<double>(q, in + 4 * i, out + 3 * i, ref + 4 * i);
    else {
      out[4 * i + 2] = 0.0;
      char_cons<double>(q, in + 3 * i, out + 4 * i, ref + 4 * i);
    }
  }
}

// per-side limiter value: lim 1 Barth-Jespersen, 2 Venkatakrishnan
extern "C" void phys_limiter(int lim, long n, const double* d,
                             const double* dMax, const double* dMin,
                             double eps, double* out) {
  for (long i = 0; i < n; ++i)
    out[i] = lim == 1
                 ? limiter_phi<1, double>(d[i], dMax[i], dMin[i],
                                          dMax[i] * dMax[i],
                                          dMin[i] * dMin[i], eps)
                 : limiter_phi<2, double>(d[i], dMax[i], dMin[i],
                                          dMax[i] * dMax[i],
                                          dMin[i] * dMin[i], eps);
}

// viscous flux (f in x, g in y) at a face state with its gradients
extern "C" void phys_visc(double gam, double mu, double Pr, const double* st,
                          const double* gx, const double* gy, double* f,
                          double* g) {
  Eq<double> q = {gam, gam - 1.0, 1.0 / (gam - 1.0), mu, Pr};
  diffusion_flux<double>(q, st, gx, gy, f, g);
}
