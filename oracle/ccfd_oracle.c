/*
 * ccfd_oracle.c - CPU restatement of ccfd's hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation. It is
 * never linked into the product. Each function cites the reference lines it
 * restates; floating-point expressions keep the reference's evaluation order
 * (compiled with -ffp-contract=off, as the reference's -std=c99 build) so
 * results are bit-identical to the reference on the same inputs.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_oracle.h"

struct oracle_ctx {
	const ccfd_mesh *m;
	oracle_cfg c;
	double gam1, gam1q, pi;
	long N, S, G;
	double *prim, *cons, *cons0;   /* 4 / elem */
	double *ux, *uy, *ut;          /* 4 / elem */
	double *sideP, *flux;          /* 4 / slot */
	double *ghostElemP;            /* 4 / ghost: ghost cell state */
	double *ghostSideP;            /* 4 / ghost: ghost side state */
	double *venkEps;
	double *src;                   /* 4 / elem: source term */
	double rk[6];
	/* implicit work arrays */
	double *XK, *RXK, *Q, *FX0, *FXK, *dX;
};

static long slot_owner_elem(const ccfd_mesh *m, long s);

/* ---------- equation of state, src/equationOfState.c:19-110 ---------- */

static void prim_cons(const oracle_ctx *o, const double p[4], double c[4])
{
	c[0] = p[0];
	c[1] = p[1] * p[0];
	c[2] = p[2] * p[0];
	c[3] = o->gam1q * p[3] + 0.5 * (c[1] * p[1] + c[2] * p[2]);
}

static void cons_prim(const oracle_ctx *o, const double c[4], double p[4])
{
	p[0] = c[0];
	p[1] = c[1] / c[0];
	p[2] = c[2] / c[0];
	p[3] = o->gam1 * (c[3] - 0.5 * (c[1] * p[1] + c[2] * p[2]));
	if (p[0] < 1e-19)
		p[0] = 1e-10;
	if (p[3] < 1e-19)
		p[3] = 1e-10;
}

static void cons_char(const oracle_ctx *o, const double cv[4], double ch[3],
		const double ref[4])
{
	double c = sqrt(o->c.gam * ref[3] / ref[0]);
	double u = ref[1];
	double H = o->gam1q * c * c + 0.5 * u * u;
	double phi = u * u - 2.0 * H;
	double a1 = 1.0 / (2.0 * c * phi);
	double a2 = 1.0 / phi;
	double a3 = u * c;
	double v[3] = {cv[0], cv[1], cv[3]};
	double K[3][3] = {
		{a1 * u * (phi - a3), -a1 * (phi - 2.0 * a3), -a2},
		{a2 * (u * u + phi), -2.0 * u * a2, 2.0 * a2},
		{-a1 * u * (a3 + phi), a1 * (phi + 2.0 * a3), -a2}
	};
	for (int i = 0; i < 3; ++i) {
		ch[i] = 0.0;
		for (int j = 0; j < 3; ++j)
			ch[i] += K[i][j] * v[j];
	}
}

static void char_cons(const oracle_ctx *o, const double ch[3], double cv[4],
		const double ref[4])
{
	double u = ref[1];
	double c = sqrt(o->c.gam * ref[3] / ref[0]);
	double H = o->gam1q * c * c + 0.5 * u * u;
	double K[3][3] = {
		{1.0, 1.0, 1.0},
		{u - c, u, u + c},
		{H - u * c, 0.5 * u * u, H + u * c}
	};
	double v[3] = {0.0, 0.0, 0.0};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			v[i] += K[i][j] * ch[j];
	cv[0] = v[0];
	cv[1] = v[1];
	cv[3] = v[2];
}

/* ---------- exact Riemann solver, src/exactRiemann.c:59-214 ---------- */

typedef struct { double G[9]; } rs_consts;

static void rs_pre(const rs_consts *k, double *f, double *fd, double p,
		double rk, double pk, double ck)
{
	if (p < pk) {
		double pRat = p / pk;
		*f = k->G[3] * ck * (pow(pRat, k->G[0]) - 1.0);
		*fd = (1.0 / (rk * ck)) * pow(pRat, -k->G[1]);
	} else {
		double ak = k->G[4] / rk;
		double bk = k->G[5] * pk;
		double qrt = sqrt(ak / (bk + p));
		*f = (p - pk) * qrt;
		*fd = (1.0 - 0.5 * (p - pk) / (bk + p)) * qrt;
	}
}

static void exact_riemann(double gam, double rl, double rr, double *rho,
		double ul, double ur, double *u, double pl, double pr, double *p,
		double al, double ar, double s)
{
	const double tol = 1e-6;
	const int nIter = 1000;
	rs_consts k;
	k.G[0] = (gam - 1.0) / (2.0 * gam);
	k.G[1] = (gam + 1.0) / (2.0 * gam);
	k.G[2] = 2.0 * gam / (gam - 1.0);
	k.G[3] = 2.0 / (gam - 1.0);
	k.G[4] = 2.0 / (gam + 1.0);
	k.G[5] = (gam - 1.0) / (gam + 1.0);
	k.G[6] = 0.5 * (gam - 1.0);
	k.G[7] = 1.0 / gam;
	k.G[8] = gam - 1.0;
	const double *G = k.G;

	double du = ur - ul;
	double duCrit = G[3] * (al + ar) - du;
	if (duCrit < 0.0) {
		*rho = *u = *p = NAN;
		return;
	}
	double qMax = 2.0;
	double pv = 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (al + ar);
	double pMin = fmin(pl, pr), pMax = fmax(pl, pr);
	double qRat = pMax / pMin;
	if ((qRat < qMax) && (pMin < pv) && (pv < pMax)) {
		*p = fmax(tol, pv);
	} else if (pv < pMin) {
		double pnu = al + ar - G[6] * (ur - ul);
		double pde = al / pow(pl, G[0]) + ar / pow(pr, G[0]);
		*p = pow(pnu / pde, G[2]);
	} else {
		double gel = sqrt((G[4] / rl) / (G[5] * pl + fmax(tol, pv)));
		double ger = sqrt((G[4] / rr) / (G[5] * pr + fmax(tol, pv)));
		*p = (gel * pl + ger * pr - (ur - ul)) / (gel + ger);
		*p = fmax(tol, *p);
	}
	double p0 = *p, cha = 2.0 * tol;
	int KK = 0;
	double fr = 0.0, fl = 0.0, frd = 0.0, fld = 0.0;
	while ((cha > tol) && (KK < nIter)) {
		rs_pre(&k, &fl, &fld, *p, rl, pl, al);
		rs_pre(&k, &fr, &frd, *p, rr, pr, ar);
		*p -= (fl + fr + du) / (fld + frd);
		cha = 2.0 * fabs((*p - p0) / (*p + p0));
		if (*p < 0.0)
			*p = tol;
		p0 = *p;
		KK++;
	}
	*u = 0.5 * (ul + ur + fr - fl);
	double pm = *p, um = *u;
	if (s < um) {
		if (pm < pl) {
			if (s < ul - al) {
				*rho = rl; *u = ul; *p = pl;
			} else {
				double cml = al * pow(pm / pl, G[0]);
				double stl = um - cml;
				if (s > stl) {
					*rho = rl * pow(pm / pl, G[7]);
					*u = um; *p = pm;
				} else {
					*u = G[4] * (al + G[6] * ul + s);
					double c = G[4] * (al + G[6] * (ul - s));
					*rho = rl * pow(c / al, G[3]);
					*p = pl * pow(c / al, G[2]);
				}
			}
		} else {
			double pml = pm / pl;
			double sl = ul - al * sqrt(G[1] * pml + G[0]);
			if (s < sl) {
				*rho = rl; *u = ul; *p = pl;
			} else {
				*rho = rl * (pml + G[5]) / (pml * G[5] + 1.0);
				*u = um; *p = pm;
			}
		}
	} else {
		if (pm > pr) {
			double pmr = pm / pr;
			double sr = ur + ar * sqrt(G[1] * pmr + G[0]);
			if (s > sr) {
				*rho = rr; *u = ur; *p = pr;
			} else {
				*rho = rr * (pmr + G[5]) / (pmr * G[5] + 1.0);
				*u = um; *p = pm;
			}
		} else {
			double shr = ur + ar;
			if (s > shr) {
				*rho = rr; *u = ur; *p = pr;
			} else {
				double cmr = ar * pow(pm / pr, G[0]);
				double str = um + cmr;
				if (s < str) {
					*rho = rr * pow(pm / pr, G[7]);
					*u = um; *p = pm;
				} else {
					*u = G[4] * (-ar + G[6] * ur + s);
					double c = G[4] * (ar - G[6] * (ur - s));
					*rho = rr * pow(c / ar, G[3]);
					*p = pr * pow(c / ar, G[2]);
				}
			}
		}
	}
}

/* ---------- exact functions, src/exactFunction.c:35-183 ---------- */

void oracle_exact_func(const oracle_cfg *cfg, const ccfd_mesh *m, int f,
		const double x[2], double time, double pv[4])
{
	double gam = cfg->gam, gam1 = gam - 1.0, pi = acos(-1.0);
	switch (f) {
	case 1: {
		pv[0] = 1.0; pv[1] = 0.0; pv[2] = 0.0; pv[3] = 1.0;
		double xLen = m->xMax - m->xMin, yLen = m->yMax - m->yMin;
		if (x[0] >= 0.3 * xLen + 1.0 / 30.0 * xLen *
				cos(2.0 * pi * 3.0 / yLen * x[1]))
			pv[0] = 0.25;
		if ((x[0] <= 0.1 * xLen) && (x[0] >= 1.0 / 30.0 * xLen)) {
			pv[0] = 4.22;
			pv[3] = 4.9;
		}
		break;
	}
	case 2: {
		pv[0] = 1.0; pv[1] = 0.0; pv[2] = 0.0;
		double hWidth = fmin((m->xMax - m->xMin), (m->yMax - m->yMin)) /
			100.0 * 6.0;
		double xPeak = m->xMin + 0.5 * (m->xMax - m->xMin);
		double yPeak = m->yMin + 0.5 * (m->yMax - m->yMin);
		double expLog = exp(-log(2.0));
		double hWidth2 = hWidth * hWidth;
		double dr2 = (x[0] - xPeak) * (x[0] - xPeak) +
			(x[1] - yPeak) * (x[1] - yPeak);
		pv[3] = 1.0 + 1.0 * pow(expLog, dr2 / hWidth2);
		break;
	}
	case 3: {
		double omega = pi * 1.0, a = 2.0 * pi, r[4];
		r[0] = r[1] = r[2] = 2.0 + 0.1 * sin(omega * (x[0] + x[1]) - a * time);
		r[3] = r[0] * r[0];
		pv[0] = r[0];
		pv[1] = r[1] / r[0];
		pv[2] = r[2] / r[0];
		pv[3] = gam1 * (r[3] - 0.5 * (r[1] * pv[1] + r[2] * pv[2]));
		if (pv[0] < 1e-19) pv[0] = 1e-10;
		if (pv[3] < 1e-19) pv[3] = 1e-10;
		break;
	}
	case 4: {
		double factor = pi / (0.2 * cfg->dxRef);
		double cSum[4] = {9.4, 57.157676649772950686, -33.0, 566.0};
		double cDiff[4] = {6.6, 57.157676649772950686, -33.0, 561.0};
		double sqrt3q = 1 / sqrt(3.0), cv[4];
		for (int i = 0; i < 4; ++i)
			cv[i] = 0.5 * (cSum[i] - cDiff[i] * tanh((x[0] -
				(1.0 / 6.0 + (20.0 * time + x[1]) * sqrt3q)) * factor));
		pv[0] = cv[0];
		pv[1] = cv[1] / cv[0];
		pv[2] = cv[2] / cv[0];
		pv[3] = gam1 * (cv[3] - 0.5 * (cv[1] * cv[1] + cv[2] * cv[2]) / cv[0]);
		break;
	}
	case 5: {
		double rl = cfg->refState[0][0], ul = cfg->refState[0][1];
		double pl = cfg->refState[0][3];
		double rr = cfg->refState[1][0], ur = cfg->refState[1][1];
		double pr = cfg->refState[1][3];
		double cl = sqrt(gam * pl / rl), cr = sqrt(gam * pr / rr);
		if (time == 0.0) {
			if (x[0] <= cfg->rp1Dinterface) {
				pv[0] = rl; pv[1] = ul; pv[3] = pl;
			} else {
				pv[0] = rr; pv[1] = ur; pv[3] = pr;
			}
		} else {
			double s = (x[0] - cfg->rp1Dinterface) / time;
			exact_riemann(gam, rl, rr, &pv[0], ul, ur, &pv[1], pl, pr,
					&pv[3], cl, cr, s);
		}
		pv[2] = 0.0;
		break;
	}
	case 6: {
		double omega = pi * 1.0, amplitude = 0.00001;
		double c0v[4] = {1.0, 0.1, 0.0, 1.0}, p0[4];
		p0[0] = c0v[0];
		p0[1] = c0v[1] / c0v[0];
		p0[2] = c0v[2] / c0v[0];
		p0[3] = gam1 * (c0v[3] - 0.5 * (c0v[1] * p0[1] + c0v[2] * p0[2]));
		double c0 = sqrt(gam * p0[3] / p0[0]);
		double H0 = (c0v[3] + p0[3]) / p0[0];
		double R[4][4] = {
			{1.0, 1.0, 0.0, 1.0},
			{p0[1] - c0, p0[1], 0.0, p0[1] + c0},
			{p0[2], p0[2], 1.0, p0[2]},
			{H0 - p0[1] * c0, (p0[1] * p0[1] + p0[2] * p0[2]) / 2.0,
				p0[2], H0 + p0[1] * c0}
		};
		double vec[4] = {0.0, 0.0, 0.0, amplitude *
			sin(omega * (x[0] - (c0 + p0[1]) * time))};
		double r[4];
		for (int i = 0; i < 4; ++i) {
			r[i] = c0v[i];
			for (int j = 0; j < 4; ++j)
				r[i] += R[i][j] * vec[j];
		}
		pv[0] = r[0];
		pv[1] = r[1] / r[0];
		pv[2] = r[2] / r[0];
		pv[3] = gam1 * (r[3] - 0.5 * (r[1] * pv[1] + r[2] * pv[2]));
		if (pv[0] < 1e-19) pv[0] = 1e-10;
		if (pv[3] < 1e-19) pv[3] = 1e-10;
		break;
	}
	default:
		pv[0] = pv[1] = pv[2] = pv[3] = NAN;
	}
}

/* ---------- Riemann fluxes, src/fluxCalculation.c:30-789 ---------- */

typedef struct { double gam, gam1, gam1q; } eq_t;

static void f_god(const eq_t *q, const double *L, const double *R, double *F)
{
	double cL = sqrt(q->gam * L[3] / L[0]);
	double cR = sqrt(q->gam * R[3] / R[0]);
	double rho, vx, p;
	exact_riemann(q->gam, L[0], R[0], &rho, L[1], R[1], &vx, L[3], R[3], &p,
			cL, cR, 0.0);
	double vy = vx > 0.0 ? L[2] : R[2];
	F[0] = rho * vx;
	F[1] = rho * vx * vx + p;
	F[2] = rho * vx * vy;
	F[3] = vx * (q->gam / q->gam1 * p + 0.5 * rho * (vx * vx + vy * vy));
}

static void f_roe(const eq_t *q, const double *L, const double *R, double *F)
{
	double rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
	double rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
	double mxL = rhoL * vxL, mxR = rhoR * vxR;
	double myL = rhoL * vyL, myR = rhoR * vyR;
	double eR = q->gam1q * pR + 0.5 * rhoR * (vxR * vxR + vyR * vyR);
	double eL = q->gam1q * pL + 0.5 * rhoL * (vxL * vxL + vyL * vyL);
	double HR = (eR + pR) / rhoR, HL = (eL + pL) / rhoL;
	double sR = sqrt(rhoR), sL = sqrt(rhoL);
	double sQ = 1.0 / (sL + sR);
	double ub = (sR * vxR + sL * vxL) * sQ;
	double vb = (sR * vyR + sL * vyL) * sQ;
	double Hb = (sR * HR + sL * HL) * sQ;
	double cb = sqrt(q->gam1 * (Hb - 0.5 * (ub * ub + vb * vb)));
	double a[4] = {ub - cb, ub, ub, ub + cb};
	double r1[4] = {1.0, a[0], vb, Hb - ub * cb};
	double r2[4] = {1.0, ub, vb, 0.5 * (ub * ub + vb * vb)};
	double r3[4] = {0.0, 0.0, 1.0, vb};
	double r4[4] = {1.0, a[3], vb, Hb + ub * cb};
	double dRho = rhoR - rhoL, dMx = mxR - mxL, dMy = myR - myL;
	double dE = eR - eL;
	double dEq = dE - (dMy - vb * dRho) * vb;
	double cbQ = 1.0 / cb;
	/* wave strengths; the first uses the global gamma-1 (line 115) */
	double w2 = -q->gam1 * cbQ * cbQ * (dRho * (ub * ub - Hb) + dEq - dMx * ub);
	double w1 = -0.5 * cbQ * (dMx - dRho * (ub + cb)) - 0.5 * w2;
	double w4 = dRho - w1 - w2;
	double w3 = dMy - vb * dRho;
	double fR[4] = {mxR, mxR * vxR + pR, mxR * vyR, vxR * (eR + pR)};
	double fL[4] = {mxL, mxL * vxL + pL, mxL * vyL, vxL * (eL + pL)};
	double cL = sqrt(q->gam * pL / rhoL), cR = sqrt(q->gam * pR / rhoR);
	double al[4] = {vxL - cL, vxL, vxL, vxL + cL};
	double ar[4] = {vxR - cR, vxR, vxR, vxR + cR};
	for (int i = 0; i < 4; ++i) {
		double da = fmax(fmax(0.0, a[i] - al[i]), ar[i] - a[i]);
		if (fabs(a[i]) < da)
			a[i] = 0.5 * (a[i] * a[i] / da + da);
		else
			a[i] = fabs(a[i]);
	}
	for (int i = 0; i < 4; ++i)
		F[i] = 0.5 * (fR[i] + fL[i] - w1 * fabs(a[0]) * r1[i]
				- w2 * fabs(a[1]) * r2[i] - w3 * fabs(a[2]) * r3[i]
				- w4 * fabs(a[3]) * r4[i]);
}

/* shared prefix of HLL / HLLE / HLLC */
typedef struct {
	double rhoL, vxL, vyL, pL, rhoR, vxR, vyR, pR;
	double rLq, rRq, sL, sR, sQ, eL, eR, uL[4], uR[4], fL[4], fR[4], cL, cR;
} hll_pre;

static void hll_prefix(const eq_t *q, const double *L, const double *R,
		hll_pre *h)
{
	h->rhoL = L[0]; h->vxL = L[1]; h->vyL = L[2]; h->pL = L[3];
	h->rhoR = R[0]; h->vxR = R[1]; h->vyR = R[2]; h->pR = R[3];
	h->rLq = 1.0 / h->rhoL;
	h->rRq = 1.0 / h->rhoR;
	h->sL = sqrt(h->rhoL);
	h->sR = sqrt(h->rhoR);
	h->sQ = 1.0 / (h->sL + h->sR);
	h->eR = q->gam1q * h->pR + 0.5 * h->rhoR * (h->vxR * h->vxR + h->vyR * h->vyR);
	h->eL = q->gam1q * h->pL + 0.5 * h->rhoL * (h->vxL * h->vxL + h->vyL * h->vyL);
	h->uL[0] = h->rhoL; h->uL[1] = h->rhoL * h->vxL;
	h->uL[2] = h->rhoL * h->vyL; h->uL[3] = h->eL;
	h->uR[0] = h->rhoR; h->uR[1] = h->rhoR * h->vxR;
	h->uR[2] = h->rhoR * h->vyR; h->uR[3] = h->eR;
	h->fL[0] = h->uL[1]; h->fL[1] = h->uL[1] * h->vxL + h->pL;
	h->fL[2] = h->uL[1] * h->vyL; h->fL[3] = h->vxL * (h->eL + h->pL);
	h->fR[0] = h->uR[1]; h->fR[1] = h->uR[1] * h->vxR + h->pR;
	h->fR[2] = h->uR[1] * h->vyR; h->fR[3] = h->vxR * (h->eR + h->pR);
	h->cL = sqrt(q->gam * h->pL * h->rLq);
	h->cR = sqrt(q->gam * h->pR * h->rRq);
}

static void hll_blend(const hll_pre *h, double arp, double alm, double *F)
{
	double q = 1.0 / (arp - alm);
	if (alm > 0.0) {
		for (int i = 0; i < 4; ++i) F[i] = h->fL[i];
	} else if (arp < 0.0) {
		for (int i = 0; i < 4; ++i) F[i] = h->fR[i];
	} else {
		for (int i = 0; i < 4; ++i)
			F[i] = (arp * h->fL[i] - alm * h->fR[i]) * q
				+ (arp * alm) * q * (h->uR[i] - h->uL[i]);
	}
}

static void roe_means(const eq_t *q, const hll_pre *h, double *uM, double *cM)
{
	double HL = (h->eL + h->pL) * h->rLq;
	double HR = (h->eR + h->pR) * h->rRq;
	*uM = (h->sR * h->vxR + h->sL * h->vxL) * h->sQ;
	double vM = (h->sR * h->vyR + h->sL * h->vyL) * h->sQ;
	double HM = (h->sR * HR + h->sL * HL) * h->sQ;
	*cM = sqrt(q->gam1 * (HM - 0.5 * (*uM * *uM + vM * vM)));
}

static void f_hll(const eq_t *q, const double *L, const double *R, double *F)
{
	hll_pre h;
	hll_prefix(q, L, R, &h);
	double uM, cM;
	roe_means(q, &h, &uM, &cM);
	double arp = fmax(h.vxR + h.cR, uM + cM);
	double alm = fmin(h.vxL - h.cL, uM - cM);
	hll_blend(&h, arp, alm, F);
}

static void f_hlle(const eq_t *q, const double *L, const double *R, double *F)
{
	hll_pre h;
	hll_prefix(q, L, R, &h);
	double uM = (h.sR * h.vxR + h.sL * h.vxL) * h.sQ;
	double eta2 = 0.5 * h.sR * h.sL / (h.sR + h.sL) / (h.sR + h.sL);
	double d = sqrt((h.sR * h.cR * h.cR + h.sL * h.cL * h.cL) * h.sQ
			+ eta2 * (h.vxR - h.vxL) * (h.vxR - h.vxL));
	double arp = fmax(h.vxR + h.cR, uM + d);
	double alm = fmin(h.vxL - h.cL, uM - d);
	hll_blend(&h, arp, alm, F);
}

static void f_hllc(const eq_t *q, const double *L, const double *R, double *F)
{
	hll_pre h;
	hll_prefix(q, L, R, &h);
	double uM, cM;
	roe_means(q, &h, &uM, &cM);
	double arp = fmax(h.vxR + h.cR, uM + cM);
	double alm = fmin(h.vxL - h.cL, uM - cM);
	if (alm > 0.0) {
		for (int i = 0; i < 4; ++i) F[i] = h.fL[i];
	} else if (arp < 0.0) {
		for (int i = 0; i < 4; ++i) F[i] = h.fR[i];
	} else {
		double as = (h.pR - h.pL + h.uL[1] * (alm - h.vxL) - h.uR[1] * (arp - h.vxR))
			/ (h.rhoL * (alm - h.vxL) - h.rhoR * (arp - h.vxR));
		if ((alm <= 0.0) && (as >= 0.0)) {
			double fac = h.rhoL * (alm - h.vxL) / (alm - as);
			double us[4] = {fac, as * fac, h.vyL * fac,
				fac * (h.eL / h.rhoL + (as - h.vxL) *
					(as + h.pL / (h.rhoL * (alm - h.vxL))))};
			for (int i = 0; i < 4; ++i)
				F[i] = h.fL[i] + alm * (us[i] - h.uL[i]);
		} else {
			double fac = h.rhoR * (arp - h.vxR) / (arp - as);
			double us[4] = {fac, as * fac, h.vyR * fac,
				fac * (h.eR / h.rhoR + (as - h.vxR) *
					(as + h.pR / (h.rhoR * (arp - h.vxR))))};
			for (int i = 0; i < 4; ++i)
				F[i] = h.fR[i] + arp * (us[i] - h.uR[i]);
		}
	}
}

static void f_lxf(const eq_t *q, const double *L, const double *R, double *F)
{
	double rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
	double rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
	double cL = sqrt(q->gam * pL / rhoL), cR = sqrt(q->gam * pR / rhoR);
	double a = fmax(fabs(vxR) + cR, fabs(vxL) + cL);
	double eL = q->gam1q * pL + 0.5 * rhoL * (vxL * vxL + vyL * vyL);
	double eR = q->gam1q * pR + 0.5 * rhoR * (vxR * vxR + vyR * vyR);
	double dU[4] = {rhoR - rhoL, rhoR * vxR - rhoL * vxL,
		rhoR * vyR - rhoL * vyL, eR - eL};
	double fL[4] = {rhoL * vxL, rhoL * vxL * vxL + pL, rhoL * vxL * vyL,
		vxL * (eL + pL)};
	double fR[4] = {rhoR * vxR, rhoR * vxR * vxR + pR, rhoR * vxR * vyR,
		vxR * (eR + pR)};
	for (int i = 0; i < 4; ++i)
		F[i] = 0.5 * (fR[i] + fL[i]) - 0.5 * a * dU[i];
}

static void stw_half(const eq_t *q, double rho, double vx, double vy,
		double c, const double *ap, double *f)
{
	double g2q = 0.5 / q->gam;
	f[0] = rho * g2q * (2.0 * q->gam1 * ap[1] + ap[0] + ap[3]);
	f[1] = f[0] * vx + (ap[3] - ap[0]) * rho * c * g2q;
	f[2] = f[0] * vy;
	f[3] = f[0] * 0.5 * (vx * vx + vy * vy) + (ap[3] - ap[0]) *
		rho * c * vx * g2q + (ap[3] + ap[0]) *
		rho * c * c * g2q * q->gam1q;
}

static void f_stw(const eq_t *q, const double *L, const double *R, double *F)
{
	double cL = sqrt(q->gam * L[3] / L[0]), cR = sqrt(q->gam * R[3] / R[0]);
	double aL[4] = {L[1] - cL, L[1], L[1], L[1] + cL};
	double aR[4] = {R[1] - cR, R[1], R[1], R[1] + cR};
	double ap[4], am[4], fp[4], fm[4];
	for (int i = 0; i < 4; ++i) {
		ap[i] = fmax(aL[i], 0.0);
		am[i] = fmin(aR[i], 0.0);
	}
	stw_half(q, L[0], L[1], L[2], cL, ap, fp);
	stw_half(q, R[0], R[1], R[2], cR, am, fm);
	for (int i = 0; i < 4; ++i)
		F[i] = fp[i] + fm[i];
}

static void f_cen(const eq_t *q, const double *L, const double *R, double *F)
{
	double eL = q->gam1q * L[3] + 0.5 * L[0] * (L[1] * L[1] + L[2] * L[2]);
	double eR = q->gam1q * R[3] + 0.5 * R[0] * (R[1] * R[1] + R[2] * R[2]);
	double fL[4], fR[4];
	fL[0] = L[0] * L[1]; fL[1] = fL[0] * L[1] + L[3];
	fL[2] = fL[0] * L[2]; fL[3] = L[1] * (eL + L[3]);
	fR[0] = R[0] * R[1]; fR[1] = fR[0] * R[1] + R[3];
	fR[2] = fR[0] * R[2]; fR[3] = R[1] * (eR + R[3]);
	for (int i = 0; i < 4; ++i)
		F[i] = 0.5 * (fL[i] + fR[i]);
}

static void f_ausmd(const eq_t *q, const double *L, const double *R, double *F)
{
	double rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
	double rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
	double eL = q->gam1q * pL + 0.5 * rhoL * (vxL * vxL + vyL * vyL);
	double eR = q->gam1q * pR + 0.5 * rhoR * (vxR * vxR + vyR * vyR);
	double HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
	double cm = fmax(sqrt(q->gam * pL / rhoL), sqrt(q->gam * pR / rhoR));
	double aL = 2.0 * pL / rhoL / (pL / rhoL + pR / rhoR);
	double aR = 2.0 * pR / rhoR / (pL / rhoL + pR / rhoR);
	double uP, pP, uM, pM;
	if (fabs(vxL) < cm) {
		uP = 0.25 * aL * (vxL + cm) * (vxL + cm) / cm
			+ 0.5 * (1.0 - aL) * (vxL + fabs(vxL));
		pP = 0.25 * pL * (vxL + cm) * (vxL + cm) / (cm * cm) * (2.0 - vxL / cm);
	} else {
		uP = 0.5 * (vxL + fabs(vxL));
		pP = 0.5 * pL * (vxL + fabs(vxL)) / vxL;
	}
	if (fabs(vxR) < cm) {
		uM = -0.25 * aR * (vxR - cm) * (vxR - cm) / cm
			+ 0.5 * (1.0 - aR) * (vxR - fabs(vxR));
		pM = 0.25 * pR * (vxR - cm) * (vxR - cm) / (cm * cm) * (2.0 + vxR / cm);
	} else {
		uM = 0.5 * (vxR - fabs(vxR));
		pM = 0.5 * pR * (vxR - fabs(vxR)) / vxR;
	}
	double rhoU = uP * rhoL + uM * rhoR;
	F[0] = rhoU;
	F[1] = 0.5 * (rhoU * (vxR + vxL) - fabs(rhoU) * (vxR - vxL)) + (pP + pM);
	F[2] = 0.5 * (rhoU * (vyR + vyL) - fabs(rhoU) * (vyR - vyL));
	F[3] = 0.5 * (rhoU * (HR + HL) - fabs(rhoU) * (HR - HL));
}

static void f_ausmdv(const eq_t *q, const double *L, const double *R, double *F)
{
	double rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
	double rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
	double eL = q->gam1q * pL + 0.5 * rhoL * (vxL * vxL + vyL * vyL);
	double eR = q->gam1q * pR + 0.5 * rhoR * (vxR * vxR + vyR * vyR);
	double HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
	double cL = sqrt(q->gam * pL / rhoL), cR = sqrt(q->gam * pR / rhoR);
	double cm = fmax(cL, cR);
	double aL = 2.0 * pL / rhoL / (pL / rhoL + pR / rhoR);
	double aR = 2.0 * pR / rhoR / (pL / rhoL + pR / rhoR);
	double uP, pP, uM, pM;
	if (fabs(vxL) < cm) {
		pP = 0.25 * pL * (vxL + cm) * (vxL + cm) / (cm * cm) * (2.0 - vxL / cm);
		if (vxL > 0.0)
			uP = vxL + aL * (vxL - cm) * (vxL - cm);
		else
			uP = aL * (vxL + cm) * (vxL + cm);
	} else if (vxL > 0.0) {
		uP = vxL; pP = pL;
	} else {
		uP = 0.0; pP = 0.0;
	}
	if (fabs(vxR) < cm) {
		pM = 0.25 * pR * (vxR - cm) * (vxR - cm) / (cm * cm) * (2.0 + vxR / cm);
		if (vxR > 0.0)   /* the reference's documented bug: vxL (line 670) */
			uM = -aR * (vxL - cm) * (vxL - cm);
		else
			uM = vxR - aR * (vxR + cm) * (vxR + cm);
	} else if (vxR > 0.0) {
		uM = 0.0; pM = 0.0;
	} else {
		uM = vxR; pM = pR;
	}
	double rhoU = uP * rhoL + uM * rhoR;
	double s = fmin(1.0, 10.0 * fabs(pR - pL) / fmin(pR, pL));
	double rhoUsq = 0.5 * (1.0 + s) * (rhoL * vxL * uP + rhoR * vxR * uM);
	rhoUsq += 0.25 * (1.0 - s) * (rhoU * (vxR + vxL) - fabs(rhoU) * (vxR - vxL));
	F[0] = rhoU;
	F[1] = rhoUsq + (pP + pM);
	F[2] = 0.5 * (rhoU * (vyR + vyL) - fabs(rhoU) * (vyR - vyL));
	F[3] = 0.5 * (rhoU * (HR + HL) - fabs(rhoU) * (HR - HL));
	int ta = (vxL - cL < 0.0) && (vxR - cR > 0.0);
	int tb = (vxL + cL < 0.0) && (vxR + cR > 0.0);
	double tL[4] = {1.0, vxL, vyL, HL}, tR[4] = {1.0, vxR, vyR, HR};
	if (ta && !tb)
		for (int i = 0; i < 4; ++i)
			F[i] -= 0.125 * ((vxR - cR) - (vxL - cL)) * (rhoR * tR[i] - rhoL * tL[i]);
	if (!ta && tb)
		for (int i = 0; i < 4; ++i)
			F[i] -= 0.125 * ((vxR + cR) - (vxL + cL)) * (rhoR * tR[i] - rhoL * tL[i]);
}

static void f_vanleer(const eq_t *q, const double *L, const double *R, double *F)
{
	double rhoL = L[0], vxL = L[1], vyL = L[2], pL = L[3];
	double rhoR = R[0], vxR = R[1], vyR = R[2], pR = R[3];
	double gam = q->gam;
	double cL = sqrt(gam * pL / rhoL), cR = sqrt(gam * pR / rhoR);
	double eL = q->gam1q * pL + 0.5 * rhoL * (vxL * vxL + vyL * vyL);
	double eR = q->gam1q * pR + 0.5 * rhoR * (vxR * vxR + vyR * vyR);
	double HL = (eL + pL) / rhoL, HR = (eR + pR) / rhoR;
	double fp[4], fm[4], ML = vxL / cL, MR = vxR / cR;
	if (ML > 1.0) {
		fp[0] = rhoL * vxL; fp[1] = fp[0] * vxL + pL;
		fp[2] = fp[0] * vyL; fp[3] = fp[0] * HL;
	} else if ((ML < 1.0) && (ML > -1.0)) {
		double cx = q->gam1 * vxL + 2.0 * cL;
		fp[0] = 0.25 * rhoL * cL * (ML + 1.0) * (ML + 1.0);
		fp[1] = fp[0] * cx / gam;
		fp[2] = fp[0] * vyL;
		fp[3] = 0.5 * (fp[1] * cx * gam / (gam * gam - 1.0) + fp[2] * vyL);
	} else {
		fp[0] = fp[1] = fp[2] = fp[3] = 0.0;
	}
	if (MR < -1.0) {
		fm[0] = rhoR * vxR; fm[1] = fm[0] * vxR + pR;
		fm[2] = fm[0] * vyR; fm[3] = fm[0] * HR;
	} else if ((MR < 1.0) && (MR > -1.0)) {
		double cx = q->gam1 * vxR - 2.0 * cR;
		fm[0] = -0.25 * rhoR * cR * (1.0 - MR) * (1.0 - MR);
		fm[1] = fm[0] * cx / gam;
		fm[2] = fm[0] * vyR;
		fm[3] = 0.5 * (fm[1] * cx * gam / (gam * gam - 1.0) + fm[2] * vyR);
	} else {
		fm[0] = fm[1] = fm[2] = fm[3] = 0.0;
	}
	for (int i = 0; i < 4; ++i)
		F[i] = fp[i] + fm[i];
}

static void conv_flux(const eq_t *q, int iFlux, const double *L,
		const double *R, double *F)
{
	switch (iFlux) {
	case 1: f_god(q, L, R, F); break;
	case 2: f_roe(q, L, R, F); break;
	case 3: f_hll(q, L, R, F); break;
	case 4: f_hlle(q, L, R, F); break;
	case 5: f_hllc(q, L, R, F); break;
	case 6: f_lxf(q, L, R, F); break;
	case 7: f_stw(q, L, R, F); break;
	case 8: f_cen(q, L, R, F); break;
	case 9: f_ausmd(q, L, R, F); break;
	case 10: f_ausmdv(q, L, R, F); break;
	case 11: f_vanleer(q, L, R, F); break;
	default: F[0] = F[1] = F[2] = F[3] = NAN;
	}
}

void oracle_riemann(const oracle_cfg *cfg, int iFlux, const double L[4],
		const double R[4], double F[4])
{
	eq_t q = {cfg->gam, cfg->gam - 1.0, 1.0 / (cfg->gam - 1.0)};
	conv_flux(&q, iFlux, L, R, F);
}

/* ---------- context ---------- */

oracle_ctx *oracle_create(const ccfd_mesh *m, const oracle_cfg *cfg)
{
	oracle_ctx *o = calloc(1, sizeof(*o));
	if (!o)
		return NULL;
	o->m = m;
	o->c = *cfg;
	o->gam1 = cfg->gam - 1.0;
	o->gam1q = 1.0 / o->gam1;
	o->pi = acos(-1.0);
	o->N = m->nElems;
	o->S = m->nSlots;
	o->G = m->nGhosts;
	long N = o->N, S = o->S, G = o->G;
	o->prim = calloc(4 * N + 4, sizeof(double));
	o->cons = calloc(4 * N + 4, sizeof(double));
	o->cons0 = calloc(4 * N + 4, sizeof(double));
	o->ux = calloc(4 * N + 4, sizeof(double));
	o->uy = calloc(4 * N + 4, sizeof(double));
	o->ut = calloc(4 * N + 4, sizeof(double));
	o->sideP = calloc(4 * S + 4, sizeof(double));
	o->flux = calloc(4 * S + 4, sizeof(double));
	o->ghostElemP = calloc(4 * G + 4, sizeof(double));
	o->ghostSideP = calloc(4 * G + 4, sizeof(double));
	o->venkEps = calloc(N + 1, sizeof(double));
	o->src = calloc(4 * N + 4, sizeof(double));
	if (!o->src || !o->prim || !o->cons || !o->cons0 || !o->ux || !o->uy || !o->ut ||
			!o->sideP || !o->flux || !o->ghostElemP ||
			!o->ghostSideP || !o->venkEps) {
		oracle_destroy(o);
		return NULL;
	}
	/* initFV, src/finiteVolume.c:47-52 */
	for (long e = 0; e < N; ++e)
		o->venkEps[e] = (cfg->venk_k * sqrt(m->area[e])) *
			(cfg->venk_k * sqrt(m->area[e])) *
			(cfg->venk_k * sqrt(m->area[e]));
	if (cfg->nRKstages > 0 && oracle_rk_coeffs(cfg->nRKstages,
				cfg->timeOrder, o->rk) != 0) {
		oracle_destroy(o);
		return NULL;
	}
	return o;
}

void oracle_destroy(oracle_ctx *o)
{
	if (!o)
		return;
	free(o->prim); free(o->cons); free(o->cons0); free(o->ux); free(o->uy);
	free(o->ut); free(o->sideP); free(o->flux); free(o->ghostElemP);
	free(o->ghostSideP); free(o->venkEps); free(o->src);
	free(o->XK); free(o->RXK); free(o->Q); free(o->FX0); free(o->FXK);
	free(o->dX);
	free(o);
}

void oracle_set_prim(oracle_ctx *o, const double *prim)
{
	memcpy(o->prim, prim, sizeof(double) * 4 * o->N);
	for (long e = 0; e < o->N; ++e)
		prim_cons(o, o->prim + 4 * e, o->cons + 4 * e);
}

void oracle_get_prim(const oracle_ctx *o, double *prim)
{
	memcpy(prim, o->prim, sizeof(double) * 4 * o->N);
}

void oracle_get_cons(const oracle_ctx *o, double *cons)
{
	memcpy(cons, o->cons, sizeof(double) * 4 * o->N);
}

/* ---------- boundary(), src/boundary.c:147-294 ---------- */

static void bc_ghost(const oracle_ctx *o, long g, double time,
		const double in[4], double out[4], const double x[2])
{
	const ccfd_mesh *m = o->m;
	const ccfd_bc *b = m->bc + m->ghostBC[g];
	long s = m->ghostSlot[g];
	double nx = m->n[2 * s], ny = m->n[2 * s + 1];
	switch (b->type) {
	case CCFD_BC_SLIPWALL:
	case CCFD_BC_WALL: {
		double vxX = nx * in[1] + ny * in[2];
		double vyX = -ny * in[1] + nx * in[2];
		double vxY = -vxX;
		double vyY = vyX;
		if (b->type == CCFD_BC_WALL && o->c.mu > 0.0)
			vyY = -vyX;
		out[1] = nx * vxY - ny * vyY;
		out[2] = ny * vxY + nx * vyY;
		out[0] = in[0];
		out[3] = in[3];
		break;
	}
	case CCFD_BC_INFLOW:
		out[0] = b->pVar[0]; out[1] = b->pVar[1];
		out[2] = b->pVar[2]; out[3] = b->pVar[3];
		break;
	case CCFD_BC_OUTFLOW:
		out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
		break;
	case CCFD_BC_CHARACTERISTIC: {
		double c = sqrt(o->c.gam * b->pVar[3] / b->pVar[0]);
		double v = nx * b->pVar[1] + ny * b->pVar[2];
		double il[4] = {in[0], nx * in[1] + ny * in[2],
			-ny * in[1] + nx * in[2], in[3]};
		double gl[4] = {b->pVar[0], nx * b->pVar[1] + ny * b->pVar[2],
			-ny * b->pVar[1] + nx * b->pVar[2], b->pVar[3]};
		double ic[4], gc[4], ich[3], gch[3];
		prim_cons(o, il, ic);
		prim_cons(o, gl, gc);
		cons_char(o, ic, ich, in);
		cons_char(o, gc, gch, in);
		if (v + c > 0.0)
			gch[2] = ich[2];
		if (v > 0.0)
			gch[1] = ich[1];
		if (v - c > 0.0)
			gch[0] = ich[0];
		char_cons(o, gch, gc, in);
		if (v > 0.0)
			gc[2] = ic[2];
		cons_prim(o, gc, out);
		double vx = out[1], vy = out[2];
		out[1] = nx * vx - ny * vy;
		out[2] = ny * vx + nx * vy;
		break;
	}
	case CCFD_BC_EXACTSOL:
		oracle_exact_func(&o->c, m, b->exactFunc, x, time, out);
		break;
	case CCFD_BC_PRESSURE_OUT: {
		double c = sqrt(o->c.gam * in[3] / in[0]);
		double v = nx * in[1] + ny * in[2];
		double p = (v / c < 1.0) ? b->pVar[3] : in[3];
		out[0] = in[0] * p / in[3];
		out[1] = in[1];
		out[2] = in[2];
		out[3] = p;
		break;
	}
	default:
		break;
	}
}

/* ---------- limiters, src/reconstruction.c:25-192 ---------- */

static void limit(oracle_ctx *o, long e)
{
	const ccfd_mesh *m = o->m;
	const double *p = o->prim + 4 * e;
	double *ux = o->ux + 4 * e, *uy = o->uy + 4 * e;
	double uMin[4], uMax[4];
	for (int v = 0; v < 4; ++v)
		uMin[v] = uMax[v] = p[v];
	for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
		long c = m->nbr[s];
		const double *q = c >= 0 ? o->prim + 4 * c : o->ghostElemP + 4 * (-c - 1);
		for (int v = 0; v < 4; ++v) {
			uMin[v] = fmin(uMin[v], q[v]);
			uMax[v] = fmax(uMax[v], q[v]);
		}
	}
	double dMin[4], dMax[4], dMin2[4], dMax2[4];
	for (int v = 0; v < 4; ++v) {
		dMin[v] = uMin[v] - p[v];
		dMax[v] = uMax[v] - p[v];
		dMin2[v] = dMin[v] * dMin[v];
		dMax2[v] = dMax[v] * dMax[v];
	}
	double phi[4] = {1.0, 1.0, 1.0, 1.0}, eps = o->venkEps[e];
	for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
		double gx = m->GP[2 * s], gy = m->GP[2 * s + 1];
		for (int v = 0; v < 4; ++v) {
			double pl = 1.0;
			double d = ux[v] * gx + uy[v] * gy;
			if (o->c.limiter == 1) {
				if (d > 0.0)
					pl = fmin(1.0, dMax[v] / d);
				else if (d < 0.0)
					pl = fmin(1.0, dMin[v] / d);
			} else {
				double d2 = d * d;
				if (d > 0.0)
					pl = 1.0 / d * (((dMax2[v] + eps) * d + 2.0 * d2 * dMax[v])
						/ (dMax2[v] + 2.0 * d2 + d * dMax[v] + eps));
				else if (d < 0.0)
					pl = 1.0 / d * (((dMin2[v] + eps) * d + 2.0 * d2 * dMin[v])
						/ (dMin2[v] + 2.0 * d2 + d * dMin[v] + eps));
			}
			phi[v] = fmin(phi[v], pl);
		}
	}
	for (int v = 0; v < 4; ++v) {
		ux[v] *= phi[v];
		uy[v] *= phi[v];
	}
}

/* ---------- spatial operator ---------- */

/* spatialReconstruction (src/reconstruction.c:198-314) + setBCatBarys */
static void reconstruct(oracle_ctx *o, double time)
{
	const ccfd_mesh *m = o->m;
	long N = o->N;
	if (o->c.spatialOrder == 1) {
		for (long e = 0; e < N; ++e) {
			for (int v = 0; v < 4; ++v) {
				o->ux[4 * e + v] = 0.0;
				o->uy[4 * e + v] = 0.0;
			}
			for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s)
				for (int v = 0; v < 4; ++v)
					o->sideP[4 * s + v] = o->prim[4 * e + v];
		}
		return;
	}
	for (long e = 0; e < 4 * N; ++e)
		o->ux[e] = o->uy[e] = 0.0;
	/* setBCatBarys, src/boundary.c:320-331 */
	for (long g = 0; g < o->G; ++g) {
		long el = slot_owner_elem(m, m->ghostSlot[g]);
		bc_ghost(o, g, time, o->prim + 4 * el, o->ghostElemP + 4 * g,
				m->ghostBary + 2 * g);
	}
	/* least-squares gradients, src/reconstruction.c:252-275 */
	for (long e = 0; e < N; ++e) {
		const double *p = o->prim + 4 * e;
		for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
			long c = m->nbr[s];
			const double *q = c >= 0 ? o->prim + 4 * c : o->ghostElemP + 4 * (-c - 1);
			double wx = m->w[2 * s], wy = m->w[2 * s + 1];
			for (int v = 0; v < 4; ++v) {
				double d = q[v] - p[v];
				o->ux[4 * e + v] += wx * d;
				o->uy[4 * e + v] += wy * d;
			}
		}
	}
	/* limiter and side states, src/reconstruction.c:278-312 */
	for (long e = 0; e < N; ++e) {
		limit(o, e);
		const double *p = o->prim + 4 * e;
		for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
			double dx = m->GP[2 * s], dy = m->GP[2 * s + 1];
			for (int v = 0; v < 4; ++v)
				o->sideP[4 * s + v] = p[v] + dx * o->ux[4 * e + v]
					+ dy * o->uy[4 * e + v];
		}
	}
}

static long slot_owner_elem(const ccfd_mesh *m, long s)
{
	long lo = 0, hi = m->nElems - 1;
	while (lo < hi) {
		long mid = (lo + hi + 1) / 2;
		if (m->slotStart[mid] <= s) lo = mid; else hi = mid - 1;
	}
	return lo;
}

/* diffusionFlux, src/fluxCalculation.c:855-873 */
static void diffusion(const oracle_ctx *o, const double st[4],
		const double gx[4], const double gy[4], double f[4], double g[4])
{
	double mu = o->c.mu, gam = o->c.gam, Pr = o->c.Pr;
	f[0] = 0.0;
	f[1] = (4.0 / 3.0 * gx[1] - 2.0 / 3.0 * gy[2]) * mu;
	f[2] = (gy[1] + gx[2]) * mu;
	f[3] = (4.0 / 3.0 * st[1] * gx[1] - 2.0 / 3.0 * st[1] * gy[2]
		+ (gx[2] + gy[1]) * st[2]
		+ gam / (o->gam1 * Pr * st[0] * st[0])
		* (st[0] * gx[3] - st[3] * gx[0])) * mu;
	g[0] = 0.0;
	g[1] = (gy[1] + gx[2]) * mu;
	g[2] = (4.0 / 3.0 * gy[2] - 2.0 / 3.0 * gx[1]) * mu;
	g[3] = (4.0 / 3.0 * st[2] * gy[2] - 2.0 / 3.0 * st[2] * gx[1]
		+ (gy[1] + gx[2]) * st[1]
		+ gam / (o->gam1 * Pr * st[0] * st[0])
		* (st[0] * gy[3] - st[3] * gy[0])) * mu;
}

/* fluxCalculation, src/fluxCalculation.c:882-998 */
static void face_fluxes(oracle_ctx *o)
{
	const ccfd_mesh *m = o->m;
	eq_t q = {o->c.gam, o->gam1, o->gam1q};
	static const double zero4[4] = {0.0, 0.0, 0.0, 0.0};
	for (long f = 0; f < m->nFaces; ++f) {
		long s = m->faceSlot[f];
		long c = m->nbr[s];
		long e = slot_owner_elem(m, s);
		double nx = m->n[2 * s], ny = m->n[2 * s + 1];
		const double *pa = o->sideP + 4 * s;
		const double *pb = c >= 0 ? o->sideP + 4 * m->partner[s]
			: o->ghostSideP + 4 * (-c - 1);
		double L[4] = {pa[0], nx * pa[1] + ny * pa[2],
			-ny * pa[1] + nx * pa[2], pa[3]};
		double R[4] = {pb[0], nx * pb[1] + ny * pb[2],
			-ny * pb[1] + nx * pb[2], pb[3]};
		double fc[4] = {0.0, 0.0, 0.0, 0.0}, fx[4] = {0}, fy[4] = {0};
		if (o->c.navierstokes) {
			const double *ea = o->prim + 4 * e;
			const double *eb = c >= 0 ? o->prim + 4 * c : o->ghostElemP + 4 * (-c - 1);
			const double *uxa = o->ux + 4 * e, *uya = o->uy + 4 * e;
			const double *uxb = c >= 0 ? o->ux + 4 * c : zero4;
			const double *uyb = c >= 0 ? o->uy + 4 * c : zero4;
			double st[4], gmx[4], gmy[4], gx[4], gy[4], cr[4];
			double bd = m->bbd[s];
			double bb[2] = {m->bbv[2 * s] / bd, m->bbv[2 * s + 1] / bd};
			for (int v = 0; v < 4; ++v) {
				st[v] = 0.5 * (pb[v] + pa[v]);
				gmx[v] = 0.5 * (uxa[v] + uxb[v]);
				gmy[v] = 0.5 * (uya[v] + uyb[v]);
			}
			for (int v = 0; v < 4; ++v)
				cr[v] = gmx[v] * bb[0] + gmy[v] * bb[1] - (eb[v] - ea[v]) / bd;
			for (int v = 0; v < 4; ++v) {
				gx[v] = gmx[v] - cr[v] * bb[0];
				gy[v] = gmy[v] - cr[v] * bb[1];
			}
			conv_flux(&q, o->c.iFlux, L, R, fc);
			diffusion(o, st, gx, gy, fx, fy);
		} else {
			conv_flux(&q, o->c.iFlux, L, R, fc);
		}
		double *fl = o->flux + 4 * s;
		fl[0] = fc[0];
		fl[1] = nx * fc[1] - ny * fc[2];
		fl[2] = ny * fc[1] + nx * fc[2];
		fl[3] = fc[3];
		if (o->c.navierstokes)
			for (int v = 0; v < 4; ++v)
				fl[v] -= (fx[v] * nx + fy[v] * ny);
		double len = m->len[s];
		for (int v = 0; v < 4; ++v)
			fl[v] *= len;
		if (c >= 0) {
			double *fp = o->flux + 4 * m->partner[s];
			for (int v = 0; v < 4; ++v)
				fp[v] = -fl[v];
		}
	}
}

/* ---------- quadrature points, src/mesh.c:117-165 ---------- */

int oracle_quadrature(const ccfd_mesh *m, long e, double xGP[5][2],
		double wGP[5])
{
	long s0 = m->slotStart[e];
	if (m->elemType[e] == 3) {
		for (int k = 0; k < 3; ++k) {
			xGP[k][0] = m->bary[2 * e] + m->GP[2 * (s0 + k)];
			xGP[k][1] = m->bary[2 * e + 1] + m->GP[2 * (s0 + k) + 1];
			wGP[k] = m->area[e] / 3.0;
		}
		return 3;
	}
	for (int k = 0; k < 4; ++k) {
		xGP[k][0] = m->bary[2 * e] + m->GP[2 * (s0 + k)];
		xGP[k][1] = m->bary[2 * e + 1] + m->GP[2 * (s0 + k) + 1];
		wGP[k] = m->area[e] / 6.0;
	}
	const double *a = m->xy + 2 * m->elemNode[4 * e];
	const double *c = m->xy + 2 * m->elemNode[4 * e + 2];
	xGP[4][0] = 0.5 * (a[0] + c[0]);
	xGP[4][1] = 0.5 * (a[1] + c[1]);
	wGP[4] = m->area[e] / 3.0;
	return 5;
}

/* ---------- source term, src/source.c ---------- */

/* evalSource (src/source.c:22-60); function 1 is the manufactured sine
 * wave, other numbers leave the values as they were */
static void eval_source(const oracle_ctx *o, int f, const double x[2],
		double time, double src[4])
{
	if (f != 1)
		return;
	double gam = o->c.gam, gam1 = o->gam1, pi = o->pi;
	double freq = 1.0, amp = 0.1, om = pi * freq, a = 2.0 * pi;
	double t1 = cos(om * (x[0] + x[1]) - a * time);
	double t2 = sin(2.0 * (om * (x[0] + x[1]) - a * time));
	src[0] = (-a + 2.0 * om) * t1;
	src[1] = (-a + om * (gam * 3.0 - 1.0)) * t1 + amp * om * gam1 * t2;
	src[2] = src[1];
	if (o->c.navierstokes) {
		double t3 = sin(om * (x[0] + x[1]) - a * time);
		src[3] = ((2.0 + gam * 6.0) * om - 4.0 * a) * t1
			+ amp * (2.0 * om * gam - a) * t2
			+ 2.0 * o->c.mu * gam * om * om / o->c.Pr * t3;
	} else {
		src[3] = ((2.0 + gam * 6.0) * om - 4.0 * a) * t1
			+ amp * (2.0 * om * gam - a) * t2;
	}
	for (int v = 0; v < 4; ++v)
		src[v] *= amp;
}

/* calcSource (src/source.c:66-85): quadrature of evalSource per element;
 * the point values carry over between points as the reference's local
 * array does */
static void calc_source(oracle_ctx *o, double time)
{
	for (long e = 0; e < o->N; ++e) {
		double xg[5][2], wg[5], pt[4] = {0.0, 0.0, 0.0, 0.0};
		double *S = o->src + 4 * e;
		S[0] = S[1] = S[2] = S[3] = 0.0;
		int n = oracle_quadrature(o->m, e, xg, wg);
		for (int k = 0; k < n; ++k) {
			eval_source(o, o->c.sourceFunc, xg[k], time, pt);
			for (int v = 0; v < 4; ++v)
				S[v] += pt[v] * wg[k];
		}
	}
}

void oracle_fvtd(oracle_ctx *o, double time, double *u_t, double *u_x,
		double *u_y, double *sideP, double *flux)
{
	const ccfd_mesh *m = o->m;
	long N = o->N;
	reconstruct(o, time);
	/* setBCatSides, src/boundary.c:300-314 */
	for (long g = 0; g < o->G; ++g) {
		long s = m->ghostSlot[g];
		long e = slot_owner_elem(m, s);
		double x[2] = {m->GP[2 * s] + m->bary[2 * e],
			m->GP[2 * s + 1] + m->bary[2 * e + 1]};
		bc_ghost(o, g, time, o->sideP + 4 * s, o->ghostSideP + 4 * g, x);
	}
	face_fluxes(o);
	/* source term, src/finiteVolume.c:91-93 (zero without calcSource,
	 * src/finiteVolume.c:58-64) */
	if (o->c.sourceFunc)
		calc_source(o, time);
	/* residual gather, src/finiteVolume.c:96-120 */
	for (long e = 0; e < N; ++e) {
		double acc[4] = {0.0, 0.0, 0.0, 0.0};
		for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s)
			for (int v = 0; v < 4; ++v)
				acc[v] += o->flux[4 * s + v];
		for (int v = 0; v < 4; ++v)
			o->ut[4 * e + v] = (o->src[4 * e + v] - acc[v]) * m->areaq[e];
	}
	if (u_t) memcpy(u_t, o->ut, sizeof(double) * 4 * N);
	if (u_x) memcpy(u_x, o->ux, sizeof(double) * 4 * N);
	if (u_y) memcpy(u_y, o->uy, sizeof(double) * 4 * N);
	if (sideP) memcpy(sideP, o->sideP, sizeof(double) * 4 * o->S);
	if (flux) memcpy(flux, o->flux, sizeof(double) * 4 * o->S);
}

/* ---------- time stepping, src/timeDiscretization.c ---------- */

int oracle_rk_coeffs(int nRK, int order, double rk[6])
{
	for (int i = 0; i < 6; ++i)
		rk[i] = 0.0;
	if (nRK == 1 && order == 1) {
		rk[1] = 1.0;
		return 0;
	}
	static const double t3[3][3] = {{0.1418, 0.4, 1.0},
		{0.1918, 0.4929, 1.0}, {0.2884, 0.5010, 1.0}};
	static const double t4[3][4] = {{0.0833, 0.2069, 0.4265, 1.0},
		{0.1084, 0.2602, 0.5052, 1.0}, {0.1666, 0.3027, 0.5275, 1.0}};
	static const double t5[3][5] = {{0.0533, 0.1263, 0.2375, 0.4414, 1.0},
		{0.0695, 0.1602, 0.2898, 0.5060, 1.0},
		{0.1067, 0.1979, 0.3232, 0.5201, 1.0}};
	if (order < 1 || order > 3)
		return -1;
	if (nRK == 3)
		for (int i = 0; i < 3; ++i) rk[i + 1] = t3[order - 1][i];
	else if (nRK == 4)
		for (int i = 0; i < 4; ++i) rk[i + 1] = t4[order - 1][i];
	else if (nRK == 5)
		for (int i = 0; i < 5; ++i) rk[i + 1] = t5[order - 1][i];
	else
		return -1;
	return 0;
}

double oracle_calc_dt(oracle_ctx *o, double t, double pTime, double stopTime,
		int *viscousDominates, int *nonfinite)
{
	const ccfd_mesh *m = o->m;
	double gam = o->c.gam, mu = o->c.mu, dt;
	*viscousDominates = 0;
	*nonfinite = 0;
	if (o->c.timeStep1D) {
		double dtMax = 1e150;
		for (long e = 0; e < o->N; ++e) {
			const double *p = o->prim + 4 * e;
			double a = sqrt(gam * p[3] / p[0]);
			double dtc = o->c.cfl * m->sy[e] / (fabs(p[1]) + a);
			if (!isfinite(dtc))
				*nonfinite = 1;
			dtMax = fmin(dtMax, dtc);
		}
		dt = dtMax;
	} else {
		double gamPrMax = fmax(4.0 / 3.0, gam / o->c.Pr);
		double dtConv = 1e150, dtVisc = 1e150;
		for (long e = 0; e < o->N; ++e) {
			const double *p = o->prim + 4 * e;
			double a = sqrt(gam * p[3] / p[0]);
			double ssr = (fabs(p[1]) + a) * m->sx[e] + (fabs(p[2]) + a) * m->sy[e];
			double dtc = o->c.cfl * m->area[e] / ssr;
			if (!isfinite(dtc))
				*nonfinite = 1;
			dtConv = fmin(dtConv, dtc);
		}
		if (mu > 1e-10) {
			for (long e = 0; e < o->N; ++e) {
				const double *p = o->prim + 4 * e;
				double ssr = gamPrMax * mu * m->sx[e] * m->sx[e]
					+ gamPrMax * mu * m->sy[e] * m->sy[e];
				double dtv = o->c.dfl * p[0] * p[0] * m->area[e] * m->area[e]
					/ (4.0 * ssr);
				if (!isfinite(dtv))
					*nonfinite = 1;
				dtVisc = fmin(dtVisc, dtv);
			}
		}
		dt = fmin(dtConv, dtVisc);
		if (dtVisc < dtConv)
			*viscousDominates = 1;
	}
	if ((t + dt > pTime) || (t + dt > stopTime))
		dt = fmin(pTime, stopTime) - t;
	else if ((t + 1.5 * dt > pTime) || (t + 1.5 * dt > stopTime))
		dt = 0.5 * (fmin(pTime, stopTime) - t);
	return dt;
}

void oracle_global_residual(const oracle_ctx *o, const double *ut, double r[4])
{
	const ccfd_mesh *m = o->m;
	r[0] = r[1] = r[2] = r[3] = 0.0;
	for (long e = 0; e < o->N; ++e)
		for (int v = 0; v < 4; ++v)
			r[v] += m->area[e] * ut[4 * e + v] * ut[4 * e + v];
	for (int v = 0; v < 4; ++v)
		r[v] = sqrt(r[v] * m->totalArea_q);
}

void oracle_explicit_step(oracle_ctx *o, double t, double dt, double res[4])
{
	long N = o->N;
	int nRK = o->c.nRKstages;
	if (nRK == 1 && o->c.timeOrder == 1) {
		oracle_fvtd(o, t, NULL, NULL, NULL, NULL, NULL);
		for (long e = 0; e < N; ++e) {
			for (int v = 0; v < 4; ++v)
				o->cons[4 * e + v] += dt * o->ut[4 * e + v];
			cons_prim(o, o->cons + 4 * e, o->prim + 4 * e);
		}
	} else {
		memcpy(o->cons0, o->cons, sizeof(double) * 4 * N);
		for (int k = 1; k <= nRK; ++k) {
			double dtStage = o->rk[k - 1] * dt;
			oracle_fvtd(o, t + dtStage, NULL, NULL, NULL, NULL, NULL);
			for (long e = 0; e < N; ++e) {
				for (int v = 0; v < 4; ++v)
					o->cons[4 * e + v] = o->cons0[4 * e + v]
						+ o->rk[k] * dt * o->ut[4 * e + v];
				cons_prim(o, o->cons + 4 * e, o->prim + 4 * e);
			}
		}
	}
	if (res)
		oracle_global_residual(o, o->ut, res);
}

/* ---------- implicit path, src/linearSolver.c / src/timeDiscretization.c --- */

double oracle_dot(long N, const double *a, const double *b)
{
	double r = 0.0;
	for (long e = 0; e < N; ++e) {
		r += a[4 * e] * b[4 * e];
		r += a[4 * e + 1] * b[4 * e + 1];
		r += a[4 * e + 2] * b[4 * e + 2];
		r += a[4 * e + 3] * b[4 * e + 3];
	}
	return r;
}

/* LU-SGS block storage: per element the diagonal block and one block per
 * slot (coupling of the element's row to its neighbour's column), i.e. the
 * entries src/linearSolver.c:203-222 records in the dense matrix */
typedef struct {
	double *Dinv;   /* 16 / elem */
	double *Blk;    /* 16 / slot: dRdU[row elem][col nbr] (already I - dt J) */
	double *Dg;     /* 16 / elem */
} lusgs_t;

/* calcDinv, src/linearSolver.c:136-179 */
static int inv4(const double A[4][4], double Ai[4][4])
{
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
	if (fabs(det) <= DBL_EPSILON)
		return 0;
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
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j)
			Ai[i][j] = C[j][i] / det;
	return 1;
}

/* buildMatrix (src/linearSolver.c:186-252): perturb each primitive
 * component of each element by sqrt(eps), record the change of its own
 * and its face neighbours' time derivative. */
static int build_matrix(oracle_ctx *o, lusgs_t *L, double time, double dt)
{
	const ccfd_mesh *m = o->m;
	long N = o->N;
	double rEps0 = sqrt(DBL_EPSILON), srEps0 = 1.0 / rEps0;
	memset(L->Dg, 0, sizeof(double) * 16 * N);
	memset(L->Blk, 0, sizeof(double) * 16 * o->S);
	for (long i = 0; i < N; ++i) {
		for (int iv = 0; iv < 4; ++iv) {
			o->prim[4 * i + iv] += rEps0;
			oracle_fvtd(o, time, NULL, NULL, NULL, NULL, NULL);
			o->prim[4 * i + iv] -= rEps0;
			for (int jv = 0; jv < 4; ++jv)
				L->Dg[16 * i + 4 * jv + iv] +=
					(o->ut[4 * i + jv] - o->RXK[4 * i + jv]) * srEps0;
			for (long s = m->slotStart[i]; s < m->slotStart[i + 1]; ++s) {
				long j = m->nbr[s];
				if (j < 0)
					continue;
				/* dense entry (row j, column i); every slot of j facing
				 * i holds the whole entry, a self-neighbour adds to the
				 * diagonal, as the dense array would */
				for (int jv = 0; jv < 4; ++jv) {
					double d = (o->ut[4 * j + jv] - o->RXK[4 * j + jv]) * srEps0;
					if (j == i) {
						L->Dg[16 * i + 4 * jv + iv] += d;
						continue;
					}
					for (long t = m->slotStart[j]; t < m->slotStart[j + 1]; ++t)
						if (m->nbr[t] == i)
							L->Blk[16 * t + 4 * jv + iv] += d;
				}
			}
		}
	}
	for (long i = 0; i < N; ++i) {
		for (int k = 0; k < 16; ++k)
			L->Dg[16 * i + k] *= -dt;
		for (int k = 0; k < 4; ++k)
			L->Dg[16 * i + 5 * k] += 1.0;
	}
	for (long s = 0; s < o->S; ++s)
		for (int k = 0; k < 16; ++k)
			L->Blk[16 * s + k] *= -dt;
	for (long i = 0; i < N; ++i) {
		double A[4][4], Ai[4][4];
		for (int a = 0; a < 4; ++a)
			for (int b = 0; b < 4; ++b)
				A[a][b] = L->Dg[16 * i + 4 * a + b];
		if (!inv4(A, Ai))
			return -2;
		for (int a = 0; a < 4; ++a)
			for (int b = 0; b < 4; ++b)
				L->Dinv[16 * i + 4 * a + b] = Ai[a][b];
	}
	return 0;
}

/* LUSGS (src/linearSolver.c:261-333) */
static void lusgs_apply(const oracle_ctx *o, const lusgs_t *L, const double *B,
		double *X, double *Xs)
{
	const ccfd_mesh *m = o->m;
	long N = o->N;
	memset(X, 0, sizeof(double) * 4 * N);
	memset(Xs, 0, sizeof(double) * 4 * N);
	for (long i = 0; i < N; ++i) {
		double t1[4] = {0.0, 0.0, 0.0, 0.0};
		for (long s = m->slotStart[i]; s < m->slotStart[i + 1]; ++s) {
			long j = m->nbr[s];
			if (j >= 0 && j < i)
				for (int a = 0; a < 4; ++a)
					for (int b = 0; b < 4; ++b)
						t1[a] += L->Blk[16 * s + 4 * a + b] * Xs[4 * j + b];
		}
		double t2[4] = {0.0, 0.0, 0.0, 0.0};
		for (int a = 0; a < 4; ++a) {
			for (int b = 0; b < 4; ++b)
				t2[a] += L->Dinv[16 * i + 4 * a + b] * (B[4 * i + b] - t1[b]);
			Xs[4 * i + a] = t2[a];
		}
	}
	for (long i = N - 1; i >= 0; --i) {
		double t1[4] = {0.0, 0.0, 0.0, 0.0};
		for (long s = m->slotStart[i]; s < m->slotStart[i + 1]; ++s) {
			long j = m->nbr[s];
			if (j > i && j < N)
				for (int a = 0; a < 4; ++a)
					for (int b = 0; b < 4; ++b)
						t1[a] += L->Blk[16 * s + 4 * a + b] * X[4 * j + b];
		}
		double t2[4] = {0.0, 0.0, 0.0, 0.0};
		for (int a = 0; a < 4; ++a) {
			for (int b = 0; b < 4; ++b)
				t2[a] += L->Dinv[16 * i + 4 * a + b] * t1[b];
			X[4 * i + a] = Xs[4 * i + a] - t2[a];
		}
	}
}

/* matrixVector, src/linearSolver.c:343-373 */
static void mat_vec(oracle_ctx *o, double time, double dt, double alpha,
		const double *v, double *res)
{
	long N = o->N;
	double epsFD = oracle_dot(N, v, v);
	epsFD = sqrt(DBL_EPSILON) / sqrt(epsFD);
	for (long e = 0; e < N; ++e) {
		for (int k = 0; k < 4; ++k)
			o->cons[4 * e + k] = o->XK[4 * e + k] + epsFD * v[4 * e + k];
		cons_prim(o, o->cons + 4 * e, o->prim + 4 * e);
	}
	oracle_fvtd(o, time, NULL, NULL, NULL, NULL, NULL);
	for (long e = 0; e < 4 * N; ++e)
		res[e] = v[e] - alpha * dt * (o->ut[e] - o->RXK[e]) / epsFD;
}

/* GMRES_M, src/linearSolver.c:385-510 (single cycle, returns at nKdim) */
static int gmres(oracle_ctx *o, const oracle_implicit_cfg *ic, double tGlobal,
		double time, double dt, double alpha, const double *B, double normB,
		double *dX, int *nIter)
{
	long N = o->N;
	int nK = ic->nKdim;
	double abortCrit = ic->epsGMRES * normB;
	double *V = calloc((size_t)4 * N * (nK + 1), sizeof(double));
	double *Z = calloc((size_t)4 * N * (nK + 1), sizeof(double));
	double *W = calloc((size_t)4 * N, sizeof(double));
	double *Xs = calloc((size_t)4 * N, sizeof(double));
	double *H = calloc((size_t)(nK + 1) * (nK + 1), sizeof(double));
	double *gm = calloc(nK + 1, sizeof(double));
	double *C = calloc(nK + 1, sizeof(double)), *Sn = calloc(nK + 1, sizeof(double));
	double *alp = calloc(nK + 1, sizeof(double));
	lusgs_t L = {0};
	int rc = 0;
	if (ic->precond) {
		L.Dinv = calloc(16 * N, sizeof(double));
		L.Dg = calloc(16 * N, sizeof(double));
		L.Blk = calloc(16 * o->S, sizeof(double));
	}
	double normR0 = normB;
	for (long e = 0; e < 4 * N; ++e) {
		dX[e] = 0.0;
		V[e] = (-B[e]) / normR0;
	}
	gm[0] = normR0;
	if (ic->precond) {
		/* primitive state at XK is what buildMatrix perturbs */
		rc = build_matrix(o, &L, tGlobal, dt);
		if (rc)
			goto done;
	}
#define Hm(i, j) H[(i) * (nK + 1) + (j)]
	for (int mm = 0; mm < nK; ++mm) {
		(*nIter)++;
		double *Vm = V + (size_t)4 * N * mm, *Zm = Z + (size_t)4 * N * mm;
		if (ic->precond)
			lusgs_apply(o, &L, Vm, Zm, Xs);
		else
			memcpy(Zm, Vm, sizeof(double) * 4 * N);
		mat_vec(o, time, dt, alpha, Zm, W);
		for (int nn = 0; nn <= mm; ++nn) {
			double *Vn = V + (size_t)4 * N * nn;
			Hm(nn, mm) = oracle_dot(N, Vn, W);
			for (long e = 0; e < 4 * N; ++e)
				W[e] -= Hm(nn, mm) * Vn[e];
		}
		double res = oracle_dot(N, W, W);
		Hm(mm + 1, mm) = sqrt(res);
		for (int nn = 0; nn <= mm - 1; ++nn) {
			double tmp = C[nn] * Hm(nn, mm) + Sn[nn] * Hm(nn + 1, mm);
			Hm(nn + 1, mm) = -Sn[nn] * Hm(nn, mm) + C[nn] * Hm(nn + 1, mm);
			Hm(nn, mm) = tmp;
		}
		double bet = sqrt(Hm(mm, mm) * Hm(mm, mm) + Hm(mm + 1, mm) * Hm(mm + 1, mm));
		Sn[mm] = Hm(mm + 1, mm) / bet;
		C[mm] = Hm(mm, mm) / bet;
		Hm(mm, mm) = bet;
		gm[mm + 1] = -Sn[mm] * gm[mm];
		gm[mm] = C[mm] * gm[mm];
		if ((fabs(gm[mm + 1]) <= abortCrit) || (mm == nK - 1)) {
			for (int nn = mm; nn >= 0; --nn) {
				alp[nn] = gm[nn];
				for (int k = nn + 1; k <= mm; ++k)
					alp[nn] -= Hm(nn, k) * alp[k];
				alp[nn] /= Hm(nn, nn);
			}
			for (int nn = 0; nn <= mm; ++nn) {
				double *Zn = Z + (size_t)4 * N * nn;
				for (long e = 0; e < 4 * N; ++e)
					dX[e] += alp[nn] * Zn[e];
			}
			goto done;
		}
		double *Vn1 = V + (size_t)4 * N * (mm + 1);
		for (long e = 0; e < 4 * N; ++e)
			Vn1[e] = W[e] / Hm(mm + 1, mm);
	}
#undef Hm
done:
	free(V); free(Z); free(W); free(Xs); free(H); free(gm); free(C); free(Sn);
	free(alp); free(L.Dinv); free(L.Dg); free(L.Blk);
	return rc;
}

int oracle_implicit_step(oracle_ctx *o, const oracle_implicit_cfg *ic,
		double t, double dt, double res[4], int *nNewton, int *nGMRES)
{
	long N = o->N;
	double alpha = 1.0, beta = 1.0;
	if (!o->XK) {
		o->XK = calloc(4 * N, sizeof(double));
		o->RXK = calloc(4 * N, sizeof(double));
		o->Q = calloc(4 * N, sizeof(double));
		o->FX0 = calloc(4 * N, sizeof(double));
		o->FXK = calloc(4 * N, sizeof(double));
		o->dX = calloc(4 * N, sizeof(double));
	}
	double eps2 = ic->epsNewton * ic->epsNewton;
	double eps2sq = sqrt(eps2);
	for (long e = 0; e < N; ++e) {
		for (int k = 0; k < 4; ++k)
			o->Q[4 * e + k] = o->cons[4 * e + k];
		cons_prim(o, o->cons + 4 * e, o->prim + 4 * e);
	}
	double time = t + beta * dt;
	oracle_fvtd(o, time, NULL, NULL, NULL, NULL, NULL);
	for (long e = 0; e < 4 * N; ++e) {
		o->FX0[e] = o->cons[e] - o->Q[e] - alpha * dt * o->ut[e];
		o->XK[e] = o->cons[e];
		o->RXK[e] = o->ut[e];
		o->FXK[e] = o->FX0[e];
	}
	double n2F0 = oracle_dot(N, o->FX0, o->FX0);
	double eps2loc = fmax(1e-8 * 1e-8 * N * eps2 / n2F0, eps2);
	double n2FK = n2F0;
	int nIn = 0;
	double crit = 0.999, n2old = n2FK, gA, gB;
	while ((n2FK > eps2loc * n2F0) && (nIn < ic->nNewtonIter)) {
		if (nIn == 0) {
			crit = 0.999;
			n2old = n2FK;
		} else {
			gA = ic->gammaEW * n2FK / n2old;
			if (ic->gammaEW * crit * crit < 0.1)
				gB = fmin(0.999, gA);
			else
				gB = fmin(0.999, fmax(gA, ic->gammaEW * crit * crit));
			crit = fmin(0.999, fmax(gB, 0.5 * eps2sq / sqrt(n2FK)));
			n2old = n2FK;
		}
		nIn++;
		/* GMRES_M overwrites the forcing term with epsGMRES*|B|
		 * (src/linearSolver.c:388), so crit only feeds the next iteration */
		int rc = gmres(o, ic, t, time, dt, alpha, o->FXK, sqrt(n2FK), o->dX,
				nGMRES);
		crit = ic->epsGMRES * sqrt(n2FK);
		if (rc)
			return rc;
		for (long e = 0; e < N; ++e) {
			for (int k = 0; k < 4; ++k) {
				o->XK[4 * e + k] += o->dX[4 * e + k];
				o->cons[4 * e + k] = o->XK[4 * e + k];
			}
			cons_prim(o, o->cons + 4 * e, o->prim + 4 * e);
		}
		oracle_fvtd(o, time, NULL, NULL, NULL, NULL, NULL);
		for (long e = 0; e < 4 * N; ++e) {
			o->RXK[e] = o->ut[e];
			o->FXK[e] = o->cons[e] - o->Q[e] - alpha * dt * o->ut[e];
		}
		n2FK = oracle_dot(N, o->FXK, o->FXK);
	}
	*nNewton += nIn;
	if (nIn == ic->nNewtonIter)
		return -1;
	if (res)
		oracle_global_residual(o, o->ut, res);
	return 0;
}

/* buildMatrix (src/linearSolver.c:186-252) on its own, at the current state:
 * R(X) from one operator call, then the blocks of I - dt dR/dpVar and the
 * inverses of the diagonal ones, as GMRES_M's preconditioner set-up makes
 * them. */
int oracle_build_matrix(oracle_ctx *o, double time, double dt, double *Dg,
		double *Blk, double *Dinv)
{
	long N = o->N;
	/* R(X) in a buffer of its own: the implicit step's arrays stay as
	 * they are */
	double *saved = o->RXK, *r = malloc(sizeof(double) * 4 * N);
	if (!r)
		return -3;
	oracle_fvtd(o, time, NULL, NULL, NULL, NULL, NULL);
	memcpy(r, o->ut, sizeof(double) * 4 * N);
	o->RXK = r;
	lusgs_t L = {Dinv, Blk, Dg};
	int rc = build_matrix(o, &L, time, dt);
	o->RXK = saved;
	free(r);
	return rc;
}

/* LUSGS (src/linearSolver.c:261-333) with the blocks of oracle_build_matrix */
int oracle_lusgs(const oracle_ctx *o, const double *Dinv, const double *Blk,
		const double *B, double *X)
{
	double *Xs = malloc(sizeof(double) * 4 * o->N);
	if (!Xs)
		return -3;
	lusgs_t L = {(double *)Dinv, (double *)Blk, NULL};
	lusgs_apply(o, &L, B, X, Xs);
	free(Xs);
	return 0;
}

/* one GMRES_M solve (src/linearSolver.c:385-510) on its own: the Krylov
 * solve of (I - dt dR/dU) dX = -B at the current state, as a Newton
 * iteration makes it (X = the current state, R(X) from one operator call,
 * the operator at time t + dt, buildMatrix at t). The state is restored
 * afterwards. */
int oracle_gmres(oracle_ctx *o, const oracle_implicit_cfg *ic, double t,
		double dt, const double *B, double *dX, int *nIter)
{
	long N = o->N;
	double *saved[2] = {o->XK, o->RXK};
	double *xk = malloc(sizeof(double) * 4 * N);
	double *rxk = malloc(sizeof(double) * 4 * N);
	double *prim = malloc(sizeof(double) * 4 * N);
	if (!xk || !rxk || !prim) {
		free(xk); free(rxk); free(prim);
		return -3;
	}
	memcpy(xk, o->cons, sizeof(double) * 4 * N);
	memcpy(prim, o->prim, sizeof(double) * 4 * N);
	oracle_fvtd(o, t + dt, NULL, NULL, NULL, NULL, NULL);
	memcpy(rxk, o->ut, sizeof(double) * 4 * N);
	o->XK = xk;
	o->RXK = rxk;
	*nIter = 0;
	int rc = gmres(o, ic, t, t + dt, dt, 1.0, B, sqrt(oracle_dot(N, B, B)),
			dX, nIter);
	memcpy(o->cons, xk, sizeof(double) * 4 * N);
	memcpy(o->prim, prim, sizeof(double) * 4 * N);
	o->XK = saved[0];
	o->RXK = saved[1];
	free(xk); free(rxk); free(prim);
	return rc;
}

/* ---------- analysis consumers, src/analyze.c ---------- */

void oracle_calc_errors(oracle_ctx *o, int f, double time, double L1[4],
		double L2[4], double Linf[4])
{
	const ccfd_mesh *m = o->m;
	for (int v = 0; v < 4; ++v)
		L1[v] = L2[v] = Linf[v] = 0.0;
	reconstruct(o, time);
	for (long e = 0; e < o->N; ++e) {
		double xg[5][2], wg[5];
		int n = oracle_quadrature(m, e, xg, wg);
		const double *p = o->prim + 4 * e;
		const double *gx = o->ux + 4 * e, *gy = o->uy + 4 * e;
		for (int k = 0; k < n; ++k) {
			double ex[4];
			oracle_exact_func(&o->c, m, f, xg[k], time, ex);
			double dx[2] = {xg[k][0] - m->bary[2 * e],
				xg[k][1] - m->bary[2 * e + 1]};
			for (int v = 0; v < 4; ++v) {
				double pv = p[v] + dx[0] * gx[v] + dx[1] * gy[v];
				double err = fabs(ex[v] - pv);
				Linf[v] = fmax(Linf[v], err);
				L1[v] += err * wg[k];
				L2[v] += err * err * wg[k];
			}
		}
	}
	for (int v = 0; v < 4; ++v) {
		L1[v] *= m->totalArea_q;
		L2[v] = sqrt(L2[v] * m->totalArea_q);
	}
}

/* the x of a side's midpoint as initWing / calcCoef compute it */
static double side_x(const ccfd_mesh *m, long s, int d)
{
	long e = slot_owner_elem(m, s);
	return m->GP[2 * s + d] + m->bary[2 * e + d];
}

/* insert slot s into a list (stable in x: after entries of equal x), the
 * reference's linked-list insertion (src/analyze.c:119-178) */
static void wing_insert(const ccfd_mesh *m, long *list, long *n, long s)
{
	double xi = side_x(m, s, 0);
	long k = 0;
	while (k < *n && !(xi < side_x(m, list[k], 0)))
		++k;
	for (long j = *n; j > k; --j)
		list[j] = list[j - 1];
	list[k] = s;
	++*n;
}

long oracle_wing_lists(const oracle_ctx *o, int wallCode, long *slots,
		long *nPressure, long *nSuction)
{
	const ccfd_mesh *m = o->m;
	long G = m->nGhosts, nP = 0, nS = 0;
	/* the wing BC: the reference walks its BC list, built by prepending
	 * (src/boundary.c:44-45), and keeps the last match, i.e. the first in
	 * parameter-file order */
	int wing = -1;
	for (int b = m->nBC - 1; b >= 0; --b)
		if (m->bc[b].type * 100 + m->bc[b].id == wallCode)
			wing = b;
	long *pr = malloc(sizeof(long) * (G + 1));
	long *su = malloc(sizeof(long) * (G + 1));
	if (!pr || !su) {
		free(pr);
		free(su);
		*nPressure = *nSuction = 0;
		return -1;
	}
	/* firstBCside is built by prepending (src/mesh.c:1739-1747): walk the
	 * ghosts from the last to the first */
	for (long g = G - 1; g >= 0; --g) {
		if (m->ghostBC[g] != wing)
			continue;
		long s = m->ghostSlot[g];
		/* the ghost side's normal is minus the interior one (src/mesh.c:
		 * 285-286); n_y <= 0 goes to the suction list */
		if (-m->n[2 * s + 1] <= 0.0)
			wing_insert(m, su, &nS, s);
		else
			wing_insert(m, pr, &nP, s);
	}
	for (long k = 0; k < nP; ++k)
		slots[k] = pr[k];
	for (long k = 0; k < nS; ++k)
		slots[nP + k] = su[k];
	free(pr);
	free(su);
	*nPressure = nP;
	*nSuction = nS;
	return nP + nS;
}

void oracle_wing_coef(const oracle_ctx *o, const long *slots, long nPressure,
		long nSuction, double refLength, const double ref[4], double alpha,
		double *cl, double *cd, double *cp, double *xyphi)
{
	const ccfd_mesh *m = o->m;
	double pi = o->pi;
	double c_l = 0.0, c_d = 0.0;
	double v = ref[1] / cos(alpha * pi / 180.0);
	double qInfQ = 1.0 / (ref[0] * 0.5 * v * v);
	double qInfLq = qInfQ / refLength;
	double pInf = ref[3];
	for (long k = 0; k < nPressure + nSuction; ++k) {
		long s = slots[k];
		double p0 = o->sideP[4 * s + 3];
		double nx = m->n[2 * s], ny = m->n[2 * s + 1], len = m->len[s];
		c_l += ny * p0 * len;
		c_d += nx * p0 * len;
		if (cp)
			cp[k] = (p0 - pInf) * qInfQ;
		if (xyphi) {
			double x = side_x(m, s, 0), y = side_x(m, s, 1);
			xyphi[3 * k] = x;
			xyphi[3 * k + 1] = y;
			xyphi[3 * k + 2] = atan2(y, x);
		}
	}
	c_l *= qInfLq;
	c_d *= qInfLq;
	double aL = alpha * pi / 180.0;
	*cl = c_l * cos(aL) - c_d * sin(aL);
	*cd = c_d * cos(aL) + c_l * sin(aL);
}

long oracle_locate(const ccfd_mesh *m, const double x[2])
{
	for (long e = 0; e < m->nElems; ++e) {
		int nv = m->elemType[e], inside = 1;
		for (int i = 0; i < nv; ++i) {
			const double *a = m->xy + 2 * m->elemNode[4 * e + i];
			const double *b = m->xy + 2 * m->elemNode[4 * e + (i + 1) % nv];
			double u[2] = {b[1] - a[1], a[0] - b[0]};
			/* the reference normalises with u_x twice (src/analyze.c:55):
			 * a horizontal edge gives 0 / 0 and never rejects */
			double t = sqrt(u[0] * u[0] + u[0] * u[0]);
			u[0] /= t;
			u[1] /= t;
			double r[2] = {x[0] - a[0], x[1] - a[1]};
			if (u[0] * r[0] + u[1] * r[1] > 0.0)
				inside = 0;
		}
		if (inside)
			return e;
	}
	return -1;
}
