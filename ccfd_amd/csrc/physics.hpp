// physics.hpp - pointwise physics of the ccfd finite-volume operator for the
// HIP kernels: equation of state, the eleven Riemann fluxes of the selector,
// boundary ghost states, limiters, viscous flux.
//
// Templated on the arithmetic type T (double: bit-identical drop-in;
// float: FP32-flux mode). With T = double every expression keeps ccfd's
// evaluation order so that, compiled with -ffp-contract=off and IEEE
// division/sqrt, results are bit-identical to the reference:
//   src/equationOfState.c:19-110, src/fluxCalculation.c:30-873,
//   src/boundary.c:147-294, src/reconstruction.c:25-192,
//   src/exactFunction.c:35-183, src/exactRiemann.c:23-214.
// Also compiled for the host (see tests) so the transcription can be checked
// against the CPU oracle without a GPU.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define CCFD_HD __host__ __device__ inline
#define CCFD_HD_NOINLINE __host__ __device__ __attribute__((noinline))
#else
#define CCFD_HD inline
#define CCFD_HD_NOINLINE __attribute__((noinline))
#endif

namespace ccfd {

template <typename T>
struct Eq {
  T gam, gam1, gam1q, mu, Pr;
};

template <typename T> CCFD_HD T tmin(T a, T b) { return fmin(a, b); }
template <typename T> CCFD_HD T tmax(T a, T b) { return fmax(a, b); }
template <typename T> CCFD_HD T tabs(T a) { return fabs(a); }
template <typename T> CCFD_HD T tsqrt(T a) { return sqrt(a); }

// ---------------------------------------------------------------- EOS
template <typename T>
CCFD_HD void prim_cons(const Eq<T>& q, const T p[4], T c[4]) {
  c[0] = p[0];
  c[1] = p[1] * p[0];
  c[2] = p[2] * p[0];
  c[3] = q.gam1q * p[3] + T(0.5) * (c[1] * p[1] + c[2] * p[2]);
}

// consPrim with the reference's positivity clamp (src/equationOfState.c:37-52)
template <typename T>
CCFD_HD void cons_prim(const Eq<T>& q, const T c[4], T p[4]) {
  p[0] = c[0];
  p[1] = c[1] / c[0];
  p[2] = c[2] / c[0];
  p[3] = q.gam1 * (c[3] - T(0.5) * (c[1] * p[1] + c[2] * p[2]));
  if (p[0] < T(1e-19)) p[0] = T(1e-10);
  if (p[3] < T(1e-19)) p[3] = T(1e-10);
}

template <typename T>
CCFD_HD void cons_char(const Eq<T>& q, const T cv[4], T ch[3], const T ref[4]) {
  T c = tsqrt(q.gam * ref[3] / ref[0]);
  T u = ref[1];
  T H = q.gam1q * c * c + T(0.5) * u * u;
  T phi = u * u - T(2.0) * H;
  T a1 = T(1.0) / (T(2.0) * c * phi);
  T a2 = T(1.0) / phi;
  T a3 = u * c;
  T v[3] = {cv[0], cv[1], cv[3]};
  T K[3][3] = {{a1 * u * (phi - a3), -a1 * (phi - T(2.0) * a3), -a2},
               {a2 * (u * u + phi), T(-2.0) * u * a2, T(2.0) * a2},
               {-a1 * u * (a3 + phi), a1 * (phi + T(2.0) * a3), -a2}};
  for (int i = 0; i < 3; ++i) {
    ch[i] = T(0.0);
    for (int j = 0; j < 3; ++j) ch[i] += K[i][j] * v[j];
  }
}

template <typename T>
CCFD_HD void char_cons(const Eq<T>& q, const T ch[3], T cv[4], const T ref[4]) {
  T u = ref[1];
  T c = tsqrt(q.gam * ref[3] / ref[0]);
  T H = q.gam1q * c * c + T(0.5) * u * u;
  T K[3][3] = {{T(1.0), T(1.0), T(1.0)},
               {u - c, u, u + c},
               {H - u * c, T(0.5) * u * u, H + u * c}};
  T v[3] = {T(0.0), T(0.0), T(0.0)};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i] += K[i][j] * ch[j];
  cv[0] = v[0];
  cv[1] = v[1];
  cv[3] = v[2];
}

// ------------------------------------------------- exact Riemann solver
template <typename T>
CCFD_HD void rs_pre(const T G[9], T* f, T* fd, T p, T rk, T pk, T ck) {
  if (p < pk) {
    T pRat = p / pk;
    *f = G[3] * ck * (pow(pRat, G[0]) - T(1.0));
    *fd = (T(1.0) / (rk * ck)) * pow(pRat, -G[1]);
  } else {
    T ak = G[4] / rk;
    T bk = G[5] * pk;
    T qrt = tsqrt(ak / (bk + p));
    *f = (p - pk) * qrt;
    *fd = (T(1.0) - T(0.5) * (p - pk) / (bk + p)) * qrt;
  }
}

// exactRiemann (src/exactRiemann.c:59-214); vacuum returns NaN
template <typename T>
CCFD_HD void exact_riemann(T gam, T rl, T rr, T* rho, T ul, T ur, T* u, T pl,
                           T pr, T* p, T al, T ar, T s) {
  const T tol = T(1e-6);
  const int nIter = 1000;
  T G[9];
  G[0] = (gam - T(1.0)) / (T(2.0) * gam);
  G[1] = (gam + T(1.0)) / (T(2.0) * gam);
  G[2] = T(2.0) * gam / (gam - T(1.0));
  G[3] = T(2.0) / (gam - T(1.0));
  G[4] = T(2.0) / (gam + T(1.0));
  G[5] = (gam - T(1.0)) / (gam + T(1.0));
  G[6] = T(0.5) * (gam - T(1.0));
  G[7] = T(1.0) / gam;
  G[8] = gam - T(1.0);
  T du = ur - ul;
  T duCrit = G[3] * (al + ar) - du;
  if (duCrit < T(0.0)) {
    *rho = *u = *p = T(NAN);
    return;
  }
  T pv = T(0.5) * (pl + pr) - T(0.125) * (ur - ul) * (rl + rr) * (al + ar);
  T pMin = tmin(pl, pr), pMax = tmax(pl, pr);
  T qRat = pMax / pMin;
  if ((qRat < T(2.0)) && (pMin < pv) && (pv < pMax)) {
    *p = tmax(tol, pv);
  } else if (pv < pMin) {
    T pnu = al + ar - G[6] * (ur - ul);
    T pde = al / pow(pl, G[0]) + ar / pow(pr, G[0]);
    *p = pow(pnu / pde, G[2]);
  } else {
    T gel = tsqrt((G[4] / rl) / (G[5] * pl + tmax(tol, pv)));
    T ger = tsqrt((G[4] / rr) / (G[5] * pr + tmax(tol, pv)));
    *p = (gel * pl + ger * pr - (ur - ul)) / (gel + ger);
    *p = tmax(tol, *p);
  }
  T p0 = *p, cha = T(2.0) * tol;
  int KK = 0;
  T fr = T(0.0), fl = T(0.0), frd = T(0.0), fld = T(0.0);
  while ((cha > tol) && (KK < nIter)) {
    rs_pre(G, &fl, &fld, *p, rl, pl, al);
    rs_pre(G, &fr, &frd, *p, rr, pr, ar);
    *p -= (fl + fr + du) / (fld + frd);
    cha = T(2.0) * tabs((*p - p0) / (*p + p0));
    if (*p < T(0.0)) *p = tol;
    p0 = *p;
    KK++;
  }
  *u = T(0.5) * (ul + ur + fr - fl);
  T pm = *p, um = *u;
  if (s < um) {
    if (pm < pl) {
      if (s < ul - al) {
        *rho = rl; *u = ul; *p = pl;
      } else {
        T cml = al * pow(pm / pl, G[0]);
        T stl = um - cml;
        if (s > stl) {
          *rho = rl * pow(pm / pl, G[7]); *u = um; *p = pm;
        } else {
          *u = G[4] * (al + G[6] * ul + s);
          T c = G[4] * (al + G[6] * (ul - s));
          *rho = rl * pow(c / al, G[3]);
          *p = pl * pow(c / al, G[2]);
        }
      }
    } else {
      T pml = pm / pl;
      T sl = ul - al * tsqrt(G[1] * pml + G[0]);
      if (s < sl) {
        *rho = rl; *u = ul; *p = pl;
      } else {
        *rho = rl * (pml + G[5]) / (pml * G[5] + T(1.0)); *u = um; *p = pm;
      }
    }
  } else {
    if (pm > pr) {
      T pmr = pm / pr;
      T sr = ur + ar * tsqrt(G[1] * pmr + G[0]);
      if (s > sr) {
        *rho = rr; *u = ur; *p = pr;
      } else {
        *rho = rr * (pmr + G[5]) / (pmr * G[5] + T(1.0)); *u = um; *p = pm;
      }
    } else {
      T shr = ur + ar;
      if (s > shr) {
        *rho = rr; *u = ur; *p = pr;
      } else {
        T cmr = ar * pow(pm / pr, G[0]);
        T str = um + cmr;
        if (s < str) {
          *rho = rr * pow(pm / pr, G[7]); *u = um; *p = pm;
        } else {
          *u = G[4] * (-ar + G[6] * ur + s);
          T c = G[4] * (ar - G[6] * (ur - s));
          *rho = rr * pow(c / ar, G[3]);
          *p = pr * pow(c / ar, G[2]);
        }
      }
    }
  }
}

// --------------------------------------------------------- Riemann fluxes
// L, R: rotated primitive states (rho, vn, vt, p)

template <typename T>
CCFD_HD void flux_god(const Eq<T>& q, const T* L, const T* R, T* F) {
  T cL = tsqrt(q.gam * L[3] / L[0]);
  T cR = tsqrt(q.gam * R[3] / R[0]);
  T rho, vx, p;
  exact_riemann<T>(q.gam, L[0], R[0], &rho, L[1], R[1], &vx, L[3], R[3], &p,
                   cL, cR, T(0.0));
  T vy = vx > T(0.0) ? L[2] : R[2];
  F[0] = rho * vx;
  F[1] = rho * vx * vx + p;
  F[2] = rho * vx * vy;
  F[3] = vx * (q.gam / q.gam1 * p + T(0.5) * rho * (vx * vx + vy * vy));
}

template <typename T>
CCFD_HD void flux_roe(const Eq<T>& q, const T* L, const T* R, T* F) {
  T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  T mxL = rhoL * vxL, mxR = rhoR * vxR;
  T myL = rhoL * vyL, myR = rhoR * vyR;
  T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  T HR = (eR + pR) / rhoR, HL = (eL + pL) / rhoL;
  T sR = tsqrt(rhoR), sL = tsqrt(rhoL);
  T sQ = T(1.0) / (sL + sR);
  T ub = (sR * vxR + sL * vxL) * sQ;
  T vb = (sR * vyR + sL * vyL) * sQ;
  T Hb = (sR * HR + sL * HL) * sQ;
  T cb = tsqrt(q.gam1 * (Hb - T(0.5) * (ub * ub + vb * vb)));
  T a[4] = {ub - cb, ub, ub, ub + cb};
  T r1[4] = {T(1.0), a[0], vb, Hb - ub * cb};
  T r2[4] = {T(1.0), ub, vb, T(0.5) * (ub * ub + vb * vb)};
  T r3[4] = {T(0.0), T(0.0), T(1.0), vb};
  T r4[4] = {T(1.0), a[3], vb, Hb + ub * cb};
  T dRho = rhoR - rhoL, dMx = mxR - mxL, dMy = myR - myL, dE = eR - eL;
  T dEq = dE - (dMy - vb * dRho) * vb;
  T cbQ = T(1.0) / cb;
  // first wave strength uses the global gamma-1 (src/fluxCalculation.c:115)
  T w2 = -q.gam1 * cbQ * cbQ * (dRho * (ub * ub - Hb) + dEq - dMx * ub);
  T w1 = T(-0.5) * cbQ * (dMx - dRho * (ub + cb)) - T(0.5) * w2;
  T w4 = dRho - w1 - w2;
  T w3 = dMy - vb * dRho;
  T fR[4] = {mxR, mxR * vxR + pR, mxR * vyR, vxR * (eR + pR)};
  T fL[4] = {mxL, mxL * vxL + pL, mxL * vyL, vxL * (eL + pL)};
  T cL = tsqrt(q.gam * pL / rhoL), cR = tsqrt(q.gam * pR / rhoR);
  T al[4] = {vxL - cL, vxL, vxL, vxL + cL};
  T ar[4] = {vxR - cR, vxR, vxR, vxR + cR};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T da = tmax(tmax(T(0.0), a[i] - al[i]), ar[i] - a[i]);
    if (tabs(a[i]) < da)
      a[i] = T(0.5) * (a[i] * a[i] / da + da);
    else
      a[i] = tabs(a[i]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
    F[i] = T(0.5) * (fR[i] + fL[i] - w1 * tabs(a[0]) * r1[i] -
                     w2 * tabs(a[1]) * r2[i] - w3 * tabs(a[2]) * r3[i] -
                     w4 * tabs(a[3]) * r4[i]);
}

template <typename T>
struct HllPre {
  T rhoL, vxL, vyL, pL, rhoR, vxR, vyR, pR;
  T rLq, rRq, sL, sR, sQ, eL, eR, uL[4], uR[4], fL[4], fR[4], cL, cR;
};

template <typename T>
CCFD_HD void hll_prefix(const Eq<T>& q, const T* L, const T* R, HllPre<T>& h) {
  h.rhoL = L[0]; h.vxL = L[1]; h.vyL = L[2]; h.pL = L[3];
  h.rhoR = R[0]; h.vxR = R[1]; h.vyR = R[2]; h.pR = R[3];
  h.rLq = T(1.0) / h.rhoL;
  h.rRq = T(1.0) / h.rhoR;
  h.sL = tsqrt(h.rhoL);
  h.sR = tsqrt(h.rhoR);
  h.sQ = T(1.0) / (h.sL + h.sR);
  h.eR = q.gam1q * h.pR + T(0.5) * h.rhoR * (h.vxR * h.vxR + h.vyR * h.vyR);
  h.eL = q.gam1q * h.pL + T(0.5) * h.rhoL * (h.vxL * h.vxL + h.vyL * h.vyL);
  h.uL[0] = h.rhoL; h.uL[1] = h.rhoL * h.vxL; h.uL[2] = h.rhoL * h.vyL; h.uL[3] = h.eL;
  h.uR[0] = h.rhoR; h.uR[1] = h.rhoR * h.vxR; h.uR[2] = h.rhoR * h.vyR; h.uR[3] = h.eR;
  h.fL[0] = h.uL[1]; h.fL[1] = h.uL[1] * h.vxL + h.pL;
  h.fL[2] = h.uL[1] * h.vyL; h.fL[3] = h.vxL * (h.eL + h.pL);
  h.fR[0] = h.uR[1]; h.fR[1] = h.uR[1] * h.vxR + h.pR;
  h.fR[2] = h.uR[1] * h.vyR; h.fR[3] = h.vxR * (h.eR + h.pR);
  h.cL = tsqrt(q.gam * h.pL * h.rLq);
  h.cR = tsqrt(q.gam * h.pR * h.rRq);
}

template <typename T>
CCFD_HD void hll_blend(const HllPre<T>& h, T arp, T alm, T* F) {
  T qq = T(1.0) / (arp - alm);
  if (alm > T(0.0)) {
    for (int i = 0; i < 4; ++i) F[i] = h.fL[i];
  } else if (arp < T(0.0)) {
    for (int i = 0; i < 4; ++i) F[i] = h.fR[i];
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      F[i] = (arp * h.fL[i] - alm * h.fR[i]) * qq +
             (arp * alm) * qq * (h.uR[i] - h.uL[i]);
  }
}

template <typename T>
CCFD_HD void roe_means(const Eq<T>& q, const HllPre<T>& h, T* uM, T* cM) {
  T HL = (h.eL + h.pL) * h.rLq;
  T HR = (h.eR + h.pR) * h.rRq;
  *uM = (h.sR * h.vxR + h.sL * h.vxL) * h.sQ;
  T vM = (h.sR * h.vyR + h.sL * h.vyL) * h.sQ;
  T HM = (h.sR * HR + h.sL * HL) * h.sQ;
  *cM = tsqrt(q.gam1 * (HM - T(0.5) * (*uM * *uM + vM * vM)));
}

template <typename T>
CCFD_HD void flux_hll(const Eq<T>& q, const T* L, const T* R, T* F) {
  HllPre<T> h;
  hll_prefix(q, L, R, h);
  T uM, cM;
  roe_means(q, h, &uM, &cM);
  T arp = tmax(h.vxR + h.cR, uM + cM);
  T alm = tmin(h.vxL - h.cL, uM - cM);
  hll_blend(h, arp, alm, F);
}

template <typename T>
CCFD_HD void flux_hlle(const Eq<T>& q, const T* L, const T* R, T* F) {
  HllPre<T> h;
  hll_prefix(q, L, R, h);
  T uM = (h.sR * h.vxR + h.sL * h.vxL) * h.sQ;
  T eta2 = T(0.5) * h.sR * h.sL / (h.sR + h.sL) / (h.sR + h.sL);
  T d = tsqrt((h.sR * h.cR * h.cR + h.sL * h.cL * h.cL) * h.sQ +
              eta2 * (h.vxR - h.vxL) * (h.vxR - h.vxL));
  T arp = tmax(h.vxR + h.cR, uM + d);
  T alm = tmin(h.vxL - h.cL, uM - d);
  hll_blend(h, arp, alm, F);
}

// HLLC (src/fluxCalculation.c:315-396). The upwind side K is selected
// first and one state / flux / star state is evaluated for it: the same
// arithmetic as each branch of the reference, with half the live values.
template <typename T>
CCFD_HD void flux_hllc(const Eq<T>& q, const T* L, const T* R, T* F) {
  const T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  const T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  const T rLq = T(1.0) / rhoL;
  const T rRq = T(1.0) / rhoR;
  const T sL = tsqrt(rhoL);
  const T sR = tsqrt(rhoR);
  const T sQ = T(1.0) / (sL + sR);
  const T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  const T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  const T cL = tsqrt(q.gam * pL * rLq);
  const T cR = tsqrt(q.gam * pR * rRq);
  const T HL = (eL + pL) * rLq;
  const T HR = (eR + pR) * rRq;
  const T uM = (sR * vxR + sL * vxL) * sQ;
  const T vM = (sR * vyR + sL * vyL) * sQ;
  const T HM = (sR * HR + sL * HL) * sQ;
  const T cM = tsqrt(q.gam1 * (HM - T(0.5) * (uM * uM + vM * vM)));
  const T arp = tmax(vxR + cR, uM + cM);
  const T alm = tmin(vxL - cL, uM - cM);
  const bool supL = alm > T(0.0);
  const bool supR = !supL && (arp < T(0.0));
  T as = T(0.0);
  bool left = supL;
  if (!supL && !supR) {
    as = (pR - pL + (rhoL * vxL) * (alm - vxL) - (rhoR * vxR) * (arp - vxR)) /
         (rhoL * (alm - vxL) - rhoR * (arp - vxR));
    left = (alm <= T(0.0)) && (as >= T(0.0));
  }
  const T rho = left ? rhoL : rhoR, vx = left ? vxL : vxR;
  const T vy = left ? vyL : vyR, p = left ? pL : pR, e = left ? eL : eR;
  const T m = rho * vx;
  F[0] = m;
  F[1] = m * vx + p;
  F[2] = m * vy;
  F[3] = vx * (e + p);
  if (!supL && !supR) {
    const T Sk = left ? alm : arp;
    const T fac = rho * (Sk - vx) / (Sk - as);
    const T us3 = fac * (e / rho + (as - vx) * (as + p / (rho * (Sk - vx))));
    F[0] = F[0] + Sk * (fac - rho);
    F[1] = F[1] + Sk * (as * fac - m);
    F[2] = F[2] + Sk * (vy * fac - rho * vy);
    F[3] = F[3] + Sk * (us3 - e);
  }
}

template <typename T>
CCFD_HD void flux_lxf(const Eq<T>& q, const T* L, const T* R, T* F) {
  T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  T cL = tsqrt(q.gam * pL / rhoL), cR = tsqrt(q.gam * pR / rhoR);
  T a = tmax(tabs(vxR) + cR, tabs(vxL) + cL);
  T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  T dU[4] = {rhoR - rhoL, rhoR * vxR - rhoL * vxL, rhoR * vyR - rhoL * vyL,
             eR - eL};
  T fL[4] = {rhoL * vxL, rhoL * vxL * vxL + pL, rhoL * vxL * vyL,
             vxL * (eL + pL)};
  T fR[4] = {rhoR * vxR, rhoR * vxR * vxR + pR, rhoR * vxR * vyR,
             vxR * (eR + pR)};
#pragma unroll
  for (int i = 0; i < 4; ++i)
    F[i] = T(0.5) * (fR[i] + fL[i]) - T(0.5) * a * dU[i];
}

template <typename T>
CCFD_HD void stw_half(const Eq<T>& q, T rho, T vx, T vy, T c, const T* ap,
                      T* f) {
  T g2q = T(0.5) / q.gam;
  f[0] = rho * g2q * (T(2.0) * q.gam1 * ap[1] + ap[0] + ap[3]);
  f[1] = f[0] * vx + (ap[3] - ap[0]) * rho * c * g2q;
  f[2] = f[0] * vy;
  f[3] = f[0] * T(0.5) * (vx * vx + vy * vy) +
         (ap[3] - ap[0]) * rho * c * vx * g2q +
         (ap[3] + ap[0]) * rho * c * c * g2q * q.gam1q;
}

template <typename T>
CCFD_HD void flux_stw(const Eq<T>& q, const T* L, const T* R, T* F) {
  T cL = tsqrt(q.gam * L[3] / L[0]), cR = tsqrt(q.gam * R[3] / R[0]);
  T aL[4] = {L[1] - cL, L[1], L[1], L[1] + cL};
  T aR[4] = {R[1] - cR, R[1], R[1], R[1] + cR};
  T ap[4], am[4], fp[4], fm[4];
  for (int i = 0; i < 4; ++i) {
    ap[i] = tmax(aL[i], T(0.0));
    am[i] = tmin(aR[i], T(0.0));
  }
  stw_half(q, L[0], L[1], L[2], cL, ap, fp);
  stw_half(q, R[0], R[1], R[2], cR, am, fm);
  for (int i = 0; i < 4; ++i) F[i] = fp[i] + fm[i];
}

template <typename T>
CCFD_HD void flux_cen(const Eq<T>& q, const T* L, const T* R, T* F) {
  T eL = q.gam1q * L[3] + T(0.5) * L[0] * (L[1] * L[1] + L[2] * L[2]);
  T eR = q.gam1q * R[3] + T(0.5) * R[0] * (R[1] * R[1] + R[2] * R[2]);
  T fL[4], fR[4];
  fL[0] = L[0] * L[1]; fL[1] = fL[0] * L[1] + L[3];
  fL[2] = fL[0] * L[2]; fL[3] = L[1] * (eL + L[3]);
  fR[0] = R[0] * R[1]; fR[1] = fR[0] * R[1] + R[3];
  fR[2] = fR[0] * R[2]; fR[3] = R[1] * (eR + R[3]);
  for (int i = 0; i < 4; ++i) F[i] = T(0.5) * (fL[i] + fR[i]);
}

template <typename T>
CCFD_HD void flux_ausmd(const Eq<T>& q, const T* L, const T* R, T* F) {
  T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  T HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
  T cm = tmax(tsqrt(q.gam * pL / rhoL), tsqrt(q.gam * pR / rhoR));
  T aL = T(2.0) * pL / rhoL / (pL / rhoL + pR / rhoR);
  T aR = T(2.0) * pR / rhoR / (pL / rhoL + pR / rhoR);
  T uP, pP, uM, pM;
  if (tabs(vxL) < cm) {
    uP = T(0.25) * aL * (vxL + cm) * (vxL + cm) / cm +
         T(0.5) * (T(1.0) - aL) * (vxL + tabs(vxL));
    pP = T(0.25) * pL * (vxL + cm) * (vxL + cm) / (cm * cm) *
         (T(2.0) - vxL / cm);
  } else {
    uP = T(0.5) * (vxL + tabs(vxL));
    pP = T(0.5) * pL * (vxL + tabs(vxL)) / vxL;
  }
  if (tabs(vxR) < cm) {
    uM = T(-0.25) * aR * (vxR - cm) * (vxR - cm) / cm +
         T(0.5) * (T(1.0) - aR) * (vxR - tabs(vxR));
    pM = T(0.25) * pR * (vxR - cm) * (vxR - cm) / (cm * cm) *
         (T(2.0) + vxR / cm);
  } else {
    uM = T(0.5) * (vxR - tabs(vxR));
    pM = T(0.5) * pR * (vxR - tabs(vxR)) / vxR;
  }
  T rhoU = uP * rhoL + uM * rhoR;
  F[0] = rhoU;
  F[1] = T(0.5) * (rhoU * (vxR + vxL) - tabs(rhoU) * (vxR - vxL)) + (pP + pM);
  F[2] = T(0.5) * (rhoU * (vyR + vyL) - tabs(rhoU) * (vyR - vyL));
  F[3] = T(0.5) * (rhoU * (HR + HL) - tabs(rhoU) * (HR - HL));
}

template <typename T>
CCFD_HD void flux_ausmdv(const Eq<T>& q, const T* L, const T* R, T* F) {
  T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  T HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
  T cL = tsqrt(q.gam * pL / rhoL), cR = tsqrt(q.gam * pR / rhoR);
  T cm = tmax(cL, cR);
  T aL = T(2.0) * pL / rhoL / (pL / rhoL + pR / rhoR);
  T aR = T(2.0) * pR / rhoR / (pL / rhoL + pR / rhoR);
  T uP, pP, uM, pM;
  if (tabs(vxL) < cm) {
    pP = T(0.25) * pL * (vxL + cm) * (vxL + cm) / (cm * cm) * (T(2.0) - vxL / cm);
    if (vxL > T(0.0))
      uP = vxL + aL * (vxL - cm) * (vxL - cm);
    else
      uP = aL * (vxL + cm) * (vxL + cm);
  } else if (vxL > T(0.0)) {
    uP = vxL; pP = pL;
  } else {
    uP = T(0.0); pP = T(0.0);
  }
  if (tabs(vxR) < cm) {
    pM = T(0.25) * pR * (vxR - cm) * (vxR - cm) / (cm * cm) * (T(2.0) + vxR / cm);
    if (vxR > T(0.0))  // the reference's documented bug: vxL (line 670)
      uM = -aR * (vxL - cm) * (vxL - cm);
    else
      uM = vxR - aR * (vxR + cm) * (vxR + cm);
  } else if (vxR > T(0.0)) {
    uM = T(0.0); pM = T(0.0);
  } else {
    uM = vxR; pM = pR;
  }
  T rhoU = uP * rhoL + uM * rhoR;
  T s = tmin(T(1.0), T(10.0) * tabs(pR - pL) / tmin(pR, pL));
  T rhoUsq = T(0.5) * (T(1.0) + s) * (rhoL * vxL * uP + rhoR * vxR * uM);
  rhoUsq += T(0.25) * (T(1.0) - s) * (rhoU * (vxR + vxL) - tabs(rhoU) * (vxR - vxL));
  F[0] = rhoU;
  F[1] = rhoUsq + (pP + pM);
  F[2] = T(0.5) * (rhoU * (vyR + vyL) - tabs(rhoU) * (vyR - vyL));
  F[3] = T(0.5) * (rhoU * (HR + HL) - tabs(rhoU) * (HR - HL));
  bool ta = (vxL - cL < T(0.0)) && (vxR - cR > T(0.0));
  bool tb = (vxL + cL < T(0.0)) && (vxR + cR > T(0.0));
  T tL[4] = {T(1.0), vxL, vyL, HL}, tR[4] = {T(1.0), vxR, vyR, HR};
  if (ta && !tb)
    for (int i = 0; i < 4; ++i)
      F[i] -= T(0.125) * ((vxR - cR) - (vxL - cL)) * (rhoR * tR[i] - rhoL * tL[i]);
  if (!ta && tb)
    for (int i = 0; i < 4; ++i)
      F[i] -= T(0.125) * ((vxR + cR) - (vxL + cL)) * (rhoR * tR[i] - rhoL * tL[i]);
}

template <typename T>
CCFD_HD void flux_vanleer(const Eq<T>& q, const T* L, const T* R, T* F) {
  T rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
  T rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
  T gam = q.gam;
  T cL = tsqrt(gam * pL / rhoL), cR = tsqrt(gam * pR / rhoR);
  T eL = q.gam1q * pL + T(0.5) * rhoL * (vxL * vxL + vyL * vyL);
  T eR = q.gam1q * pR + T(0.5) * rhoR * (vxR * vxR + vyR * vyR);
  T HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
  T fp[4], fm[4], ML = vxL / cL, MR = vxR / cR;
  if (ML > T(1.0)) {
    fp[0] = rhoL * vxL; fp[1] = fp[0] * vxL + pL;
    fp[2] = fp[0] * vyL; fp[3] = fp[0] * HL;
  } else if ((ML < T(1.0)) && (ML > T(-1.0))) {
    T cx = q.gam1 * vxL + T(2.0) * cL;
    fp[0] = T(0.25) * rhoL * cL * (ML + T(1.0)) * (ML + T(1.0));
    fp[1] = fp[0] * cx / gam;
    fp[2] = fp[0] * vyL;
    fp[3] = T(0.5) * (fp[1] * cx * gam / (gam * gam - T(1.0)) + fp[2] * vyL);
  } else {
    fp[0] = fp[1] = fp[2] = fp[3] = T(0.0);
  }
  if (MR < T(-1.0)) {
    fm[0] = rhoR * vxR; fm[1] = fm[0] * vxR + pR;
    fm[2] = fm[0] * vyR; fm[3] = fm[0] * HR;
  } else if ((MR < T(1.0)) && (MR > T(-1.0))) {
    T cx = q.gam1 * vxR - T(2.0) * cR;
    fm[0] = T(-0.25) * rhoR * cR * (T(1.0) - MR) * (T(1.0) - MR);
    fm[1] = fm[0] * cx / gam;
    fm[2] = fm[0] * vyR;
    fm[3] = T(0.5) * (fm[1] * cx * gam / (gam * gam - T(1.0)) + fm[2] * vyR);
  } else {
    fm[0] = fm[1] = fm[2] = fm[3] = T(0.0);
  }
  for (int i = 0; i < 4; ++i) F[i] = fp[i] + fm[i];
}

// convectiveFlux selector (src/fluxCalculation.c:803-844); the kernels
// instantiate one flux per launch, this runtime switch serves the batch API
template <typename T>
CCFD_HD void conv_flux(const Eq<T>& q, int code, const T* L, const T* R, T* F) {
  switch (code) {
    case 1: flux_god(q, L, R, F); break;
    case 2: flux_roe(q, L, R, F); break;
    case 3: flux_hll(q, L, R, F); break;
    case 4: flux_hlle(q, L, R, F); break;
    case 5: flux_hllc(q, L, R, F); break;
    case 6: flux_lxf(q, L, R, F); break;
    case 7: flux_stw(q, L, R, F); break;
    case 8: flux_cen(q, L, R, F); break;
    case 9: flux_ausmd(q, L, R, F); break;
    case 10: flux_ausmdv(q, L, R, F); break;
    case 11: flux_vanleer(q, L, R, F); break;
    default: F[0] = F[1] = F[2] = F[3] = T(NAN);
  }
}

template <int CODE, typename T>
CCFD_HD void conv_flux_t(const Eq<T>& q, const T* L, const T* R, T* F) {
  if (CODE == 1) flux_god(q, L, R, F);
  else if (CODE == 2) flux_roe(q, L, R, F);
  else if (CODE == 3) flux_hll(q, L, R, F);
  else if (CODE == 4) flux_hlle(q, L, R, F);
  else if (CODE == 5) flux_hllc(q, L, R, F);
  else if (CODE == 6) flux_lxf(q, L, R, F);
  else if (CODE == 7) flux_stw(q, L, R, F);
  else if (CODE == 8) flux_cen(q, L, R, F);
  else if (CODE == 9) flux_ausmd(q, L, R, F);
  else if (CODE == 10) flux_ausmdv(q, L, R, F);
  else flux_vanleer(q, L, R, F);
}

// diffusionFlux (src/fluxCalculation.c:855-873), primitive-gradient form
template <typename T>
CCFD_HD void diffusion_flux(const Eq<T>& q, const T st[4], const T gx[4],
                            const T gy[4], T f[4], T g[4]) {
  T mu = q.mu;
  f[0] = T(0.0);
  f[1] = (T(4.0 / 3.0) * gx[1] - T(2.0 / 3.0) * gy[2]) * mu;
  f[2] = (gy[1] + gx[2]) * mu;
  f[3] = (T(4.0 / 3.0) * st[1] * gx[1] - T(2.0 / 3.0) * st[1] * gy[2] +
          (gx[2] + gy[1]) * st[2] +
          q.gam / (q.gam1 * q.Pr * st[0] * st[0]) *
              (st[0] * gx[3] - st[3] * gx[0])) * mu;
  g[0] = T(0.0);
  g[1] = (gy[1] + gx[2]) * mu;
  g[2] = (T(4.0 / 3.0) * gy[2] - T(2.0 / 3.0) * gx[1]) * mu;
  g[3] = (T(4.0 / 3.0) * st[2] * gy[2] - T(2.0 / 3.0) * st[2] * gx[1] +
          (gy[1] + gx[2]) * st[1] +
          q.gam / (q.gam1 * q.Pr * st[0] * st[0]) *
              (st[0] * gy[3] - st[3] * gy[0])) * mu;
}

// ------------------------------------------------------ boundary ghosts
// One boundary condition, as the device sees it.
struct BCDev {
  int type;
  int exactFunc;
  double pVar[4];
};

// exact functions (src/exactFunction.c:35-183)
struct ExactCfg {
  double refState[2][4];
  double rp1Dinterface;
  double dxRef;
  double xMin, xMax, yMin, yMax;
};

template <typename T>
CCFD_HD void exact_func(const Eq<T>& q, const ExactCfg& ex, int f,
                        const T x[2], T time, T pv[4]) {
  T pi = T(3.141592653589793238462643383279502884);
  T gam1 = q.gam1;
  switch (f) {
    case 1: {
      pv[0] = T(1.0); pv[1] = T(0.0); pv[2] = T(0.0); pv[3] = T(1.0);
      T xLen = T(ex.xMax - ex.xMin), yLen = T(ex.yMax - ex.yMin);
      if (x[0] >= T(0.3) * xLen + T(1.0 / 30.0) * xLen *
                      cos(T(2.0) * pi * T(3.0) / yLen * x[1]))
        pv[0] = T(0.25);
      if ((x[0] <= T(0.1) * xLen) && (x[0] >= T(1.0 / 30.0) * xLen)) {
        pv[0] = T(4.22);
        pv[3] = T(4.9);
      }
      break;
    }
    case 2: {
      pv[0] = T(1.0); pv[1] = T(0.0); pv[2] = T(0.0);
      T hWidth = tmin(T(ex.xMax - ex.xMin), T(ex.yMax - ex.yMin)) / T(100.0) * T(6.0);
      T xPeak = T(ex.xMin) + T(0.5) * T(ex.xMax - ex.xMin);
      T yPeak = T(ex.yMin) + T(0.5) * T(ex.yMax - ex.yMin);
      T expLog = exp(-log(T(2.0)));
      T dr2 = (x[0] - xPeak) * (x[0] - xPeak) + (x[1] - yPeak) * (x[1] - yPeak);
      pv[3] = T(1.0) + T(1.0) * pow(expLog, dr2 / (hWidth * hWidth));
      break;
    }
    case 3: {
      T r[4];
      r[0] = r[1] = r[2] = T(2.0) + T(0.1) * sin(pi * (x[0] + x[1]) - T(2.0) * pi * time);
      r[3] = r[0] * r[0];
      cons_prim(q, r, pv);
      break;
    }
    case 4: {
      T factor = pi / (T(0.2) * T(ex.dxRef));
      const T cSum[4] = {T(9.4), T(57.157676649772950686), T(-33.0), T(566.0)};
      const T cDiff[4] = {T(6.6), T(57.157676649772950686), T(-33.0), T(561.0)};
      T sqrt3q = T(1.0) / tsqrt(T(3.0)), cv[4];
      for (int i = 0; i < 4; ++i)
        cv[i] = T(0.5) * (cSum[i] - cDiff[i] * tanh((x[0] - (T(1.0 / 6.0) +
                (T(20.0) * time + x[1]) * sqrt3q)) * factor));
      pv[0] = cv[0];
      pv[1] = cv[1] / cv[0];
      pv[2] = cv[2] / cv[0];
      pv[3] = gam1 * (cv[3] - T(0.5) * (cv[1] * cv[1] + cv[2] * cv[2]) / cv[0]);
      break;
    }
    case 5: {
      T rl = T(ex.refState[0][0]), ul = T(ex.refState[0][1]), pl = T(ex.refState[0][3]);
      T rr = T(ex.refState[1][0]), ur = T(ex.refState[1][1]), pr = T(ex.refState[1][3]);
      T cl = tsqrt(q.gam * pl / rl), cr = tsqrt(q.gam * pr / rr);
      if (time == T(0.0)) {
        if (x[0] <= T(ex.rp1Dinterface)) {
          pv[0] = rl; pv[1] = ul; pv[3] = pl;
        } else {
          pv[0] = rr; pv[1] = ur; pv[3] = pr;
        }
      } else {
        T s = (x[0] - T(ex.rp1Dinterface)) / time;
        exact_riemann<T>(q.gam, rl, rr, &pv[0], ul, ur, &pv[1], pl, pr, &pv[3],
                         cl, cr, s);
      }
      pv[2] = T(0.0);
      break;
    }
    case 6: {
      T c0v[4] = {T(1.0), T(0.1), T(0.0), T(1.0)}, p0[4];
      cons_prim(q, c0v, p0);
      T c0 = tsqrt(q.gam * p0[3] / p0[0]);
      T H0 = (c0v[3] + p0[3]) / p0[0];
      T Rm[4][4] = {{T(1.0), T(1.0), T(0.0), T(1.0)},
                    {p0[1] - c0, p0[1], T(0.0), p0[1] + c0},
                    {p0[2], p0[2], T(1.0), p0[2]},
                    {H0 - p0[1] * c0, (p0[1] * p0[1] + p0[2] * p0[2]) / T(2.0),
                     p0[2], H0 + p0[1] * c0}};
      T vec[4] = {T(0.0), T(0.0), T(0.0),
                  T(0.00001) * sin(pi * (x[0] - (c0 + p0[1]) * time))};
      T r[4];
      for (int i = 0; i < 4; ++i) {
        r[i] = c0v[i];
        for (int j = 0; j < 4; ++j) r[i] += Rm[i][j] * vec[j];
      }
      cons_prim(q, r, pv);
      break;
    }
    default:
      pv[0] = pv[1] = pv[2] = pv[3] = T(NAN);
  }
}

// source term at one point (evalSource, src/source.c:22-60): function 1 is
// the manufactured sine-wave source; any other number leaves src as it was
// (the reference's switch has no default). ns: the Navier-Stokes variant's
// extra viscous term of the energy equation.
template <typename T>
CCFD_HD void eval_source(const Eq<T>& q, int f, int ns, const T x[2], T time,
                         T src[4]) {
  if (f != 1) return;
  const T pi = T(3.141592653589793238462643383279502884);
  const T amp = T(0.1), om = pi * T(1.0), a = T(2.0) * pi;
  const T ph = om * (x[0] + x[1]) - a * time;
  const T t1 = cos(ph);
  const T t2 = sin(T(2.0) * (om * (x[0] + x[1]) - a * time));
  src[0] = (-a + T(2.0) * om) * t1;
  src[1] = (-a + om * (q.gam * T(3.0) - T(1.0))) * t1 + amp * om * q.gam1 * t2;
  src[2] = src[1];
  src[3] = ((T(2.0) + q.gam * T(6.0)) * om - T(4.0) * a) * t1 +
           amp * (T(2.0) * om * q.gam - a) * t2;
  if (ns)
    src[3] = src[3] + T(2.0) * q.mu * q.gam * om * om / q.Pr * sin(ph);
#pragma unroll
  for (int v = 0; v < 4; ++v) src[v] *= amp;
}

// characteristic ghost state (type 5; src/boundary.c:217-271)
template <typename T>
CCFD_HD void bc_ghost_char(const Eq<T>& q, const BCDev& b, T nx, T ny,
                           const T* in, T* out) {
  T bp[4] = {T(b.pVar[0]), T(b.pVar[1]), T(b.pVar[2]), T(b.pVar[3])};
  T c = tsqrt(q.gam * bp[3] / bp[0]);
  T v = nx * bp[1] + ny * bp[2];
  T il[4] = {in[0], nx * in[1] + ny * in[2], -ny * in[1] + nx * in[2], in[3]};
  T gl[4] = {bp[0], nx * bp[1] + ny * bp[2], -ny * bp[1] + nx * bp[2], bp[3]};
  T ic[4], gc[4], ich[3], gch[3];
  prim_cons(q, il, ic);
  prim_cons(q, gl, gc);
  cons_char(q, ic, ich, in);
  cons_char(q, gc, gch, in);
  if (v + c > T(0.0)) gch[2] = ich[2];
  if (v > T(0.0)) gch[1] = ich[1];
  if (v - c > T(0.0)) gch[0] = ich[0];
  char_cons(q, gch, gc, in);
  if (v > T(0.0)) gc[2] = ic[2];
  cons_prim(q, gc, out);
  T vx = out[1], vy = out[2];
  out[1] = nx * vx - ny * vy;
  out[2] = ny * vx + nx * vy;
}

// exact-function ghost state (type 6; src/boundary.c:272-274), out of line:
// a kernel that contains this call reserves the exact Riemann solver's
// stack at every launch
template <typename T>
CCFD_HD_NOINLINE void bc_ghost_exact(const Eq<T>& q, const ExactCfg& ex, int f,
                                     const T* x, T time, T* out) {
  exact_func(q, ex, f, x, time, out);
}

// boundary() (src/boundary.c:147-294). n: outward normal of the interior side.
// EXACT = false leaves exact-function boundaries (type 6) to a kernel of
// their own (NaN here), so that the common boundary kernels carry no stack.
template <typename T, bool EXACT = true>
CCFD_HD void bc_ghost(const Eq<T>& q, const BCDev& b, const ExactCfg& ex,
                      T nx, T ny, T time, const T in[4], T out[4],
                      const T x[2]) {
  switch (b.type) {
    case 1:
    case 2: {
      T vxX = nx * in[1] + ny * in[2];
      T vyX = -ny * in[1] + nx * in[2];
      T vxY = -vxX;
      T vyY = vyX;
      if (b.type == 2 && q.mu > T(0.0)) vyY = -vyX;
      out[1] = nx * vxY - ny * vyY;
      out[2] = ny * vxY + nx * vyY;
      out[0] = in[0];
      out[3] = in[3];
      break;
    }
    case 3:
      out[0] = T(b.pVar[0]); out[1] = T(b.pVar[1]);
      out[2] = T(b.pVar[2]); out[3] = T(b.pVar[3]);
      break;
    case 4:
      out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
      break;
    case 5:
      bc_ghost_char<T>(q, b, nx, ny, in, out);
      break;
    case 6:
      if constexpr (EXACT)
        bc_ghost_exact<T>(q, ex, b.exactFunc, x, time, out);
      else
        out[0] = out[1] = out[2] = out[3] = T(NAN);
      break;
    case 8: {
      T c = tsqrt(q.gam * in[3] / in[0]);
      T v = nx * in[1] + ny * in[2];
      T p = (v / c < T(1.0)) ? T(b.pVar[3]) : in[3];
      out[0] = in[0] * p / in[3];
      out[1] = in[1];
      out[2] = in[2];
      out[3] = p;
      break;
    }
    default:
      out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
  }
}

// ------------------------------------------------------------ limiters
// per-side limiter value for one variable (src/reconstruction.c:59-69,
// 143-160); d = grad . GP
//
// The two signs of d use the same expression on different bounds, so the
// bound is selected first and the expression evaluated once: the same
// operations on the same operands as the branches of the reference, without
// divergent paths in a wavefront. d == 0 (or NaN) gives 1, as in the
// reference; the expression's value there is discarded.
template <int LIM, typename T>
CCFD_HD T limiter_phi(T d, T dMax, T dMin, T dMax2, T dMin2, T eps) {
  const bool pos = d > T(0.0);
  const T dm = pos ? dMax : dMin;
  T pl;
  if (LIM == 1) {
    pl = tmin(T(1.0), dm / d);
  } else {
    const T dm2 = pos ? dMax2 : dMin2;
    const T d2 = d * d;
    pl = T(1.0) / d * (((dm2 + eps) * d + T(2.0) * d2 * dm) /
                       (dm2 + T(2.0) * d2 + d * dm + eps));
  }
  return (d > T(0.0) || d < T(0.0)) ? pl : T(1.0);
}

}  // namespace ccfd
