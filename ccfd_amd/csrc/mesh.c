/*
 * mesh.c - host mesh preprocessing (see include/ccfd_mesh.h).
 *
 * Restates the reference's mesh construction (ccfd src/mesh.c) on flat
 * arrays: element and side numbering, face pairing and ownership, periodic
 * hookup and all geometric operator inputs, with the reference's
 * floating-point operation order so the results are bit-identical.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_mesh.h"

#define SETERR(...) do { if (err) snprintf(err, 256, __VA_ARGS__); } while (0)

/* one entry of the side list sorted for pairing (src/mesh.c:1555-1563) */
typedef struct {
	long n0, n1;     /* min / max node id */
	int bc;          /* 1 for a boundary edge */
	long create;     /* creation index: elements' sides, then BC edges */
	long slot;       /* element slot, or -(k + 1) for BC edge k */
} pair_entry;

/* (node0, node1, BC) as compare() at src/mesh.c:362-378; ties keep the
 * creation order, which is what glibc's merge-sort qsort produces there */
static int pair_cmp(const void *a, const void *b)
{
	const pair_entry *A = a, *B = b;
	if (A->n0 != B->n0)
		return A->n0 < B->n0 ? -1 : 1;
	if (A->n1 != B->n1)
		return A->n1 < B->n1 ? -1 : 1;
	if (A->bc != B->bc)
		return A->bc - B->bc;
	return (A->create > B->create) - (A->create < B->create);
}

/* The side list in pair_cmp order without a global comparison sort: a
 * stable counting sort of the entries (given in creation order) by n0,
 * then each n0 bucket ordered by pair_cmp - insertion sort for the few
 * entries a node usually has, qsort for a large bucket. pair_cmp is a total
 * order (creation breaks every tie), so the result is the one a single
 * qsort gives, in time linear in the entries for bounded node degree.
 * Returns the sorted copy (the input is left as is), NULL when out of
 * memory. */
static pair_entry *sort_pairs(const pair_entry *pe, long n, long nVert)
{
	long *start = calloc((size_t)nVert + 1, sizeof(long));
	pair_entry *out = malloc(sizeof(pair_entry) * (n ? n : 1));
	if (!start || !out) {
		free(start);
		free(out);
		return NULL;
	}
	for (long i = 0; i < n; ++i)
		start[pe[i].n0 + 1]++;
	for (long v = 0; v < nVert; ++v)
		start[v + 1] += start[v];
	for (long i = 0; i < n; ++i)
		out[start[pe[i].n0]++] = pe[i];
	/* start[v] is now the end of bucket v */
	long lo = 0;
	for (long v = 0; v < nVert; ++v) {
		long hi = start[v];
		if (hi - lo > 32) {
			qsort(out + lo, hi - lo, sizeof(pair_entry), pair_cmp);
		} else {
			for (long i = lo + 1; i < hi; ++i) {
				pair_entry t = out[i];
				long j = i - 1;
				while (j >= lo && pair_cmp(out + j, &t) > 0) {
					out[j + 1] = out[j];
					--j;
				}
				out[j + 1] = t;
			}
		}
		lo = hi;
	}
	free(start);
	return out;
}

static void *xcalloc(size_t n, size_t s)
{
	void *p = calloc(n ? n : 1, s);
	return p;
}

void ccfd_mesh_free(ccfd_mesh *m)
{
	if (!m)
		return;
	free(m->xy); free(m->elemType); free(m->domain); free(m->elemNode);
	free(m->bary); free(m->area); free(m->areaq); free(m->sx); free(m->sy);
	free(m->slotStart); free(m->nbr); free(m->partner); free(m->owner);
	free(m->n); free(m->len); free(m->fn); free(m->flen); free(m->GP);
	free(m->GPc); free(m->w); free(m->bbv); free(m->bbd); free(m->slotNode);
	free(m->faceSlot); free(m->ghostSlot); free(m->ghostBC);
	free(m->ghostBary); free(m->ghostGP); free(m->bc);
	free(m);
}

/* initRecordPoints' element search, src/analyze.c:43-76 */
long ccfd_mesh_locate(const ccfd_mesh *m, const double x[2])
{
	if (!m || !x)
		return -1;
	for (long e = 0; e < m->nElems; ++e) {
		const int nv = m->elemType[e];
		int inside = 1;
		for (int k = 0; k < nv; ++k) {
			const double *p = m->xy + 2 * m->elemNode[4 * e + k];
			const double *q = m->xy + 2 * m->elemNode[4 * e + (k + 1) % nv];
			double nx = q[1] - p[1], ny = p[0] - q[0];
			/* the reference's norm: sqrt(nx^2 + nx^2) */
			double l = sqrt(nx * nx + nx * nx);
			nx /= l;
			ny /= l;
			if (nx * (x[0] - p[0]) + ny * (x[1] - p[1]) > 0.0)
				inside = 0;
		}
		if (inside)
			return e;
	}
	return -1;
}

/* createElemInfo, src/mesh.c:296-341 */
static void elem_info(ccfd_mesh *m, long e)
{
	int type = m->elemType[e];
	const long *nd = m->elemNode + 4 * e;
	double bx = 0.0, by = 0.0;
	for (int i = 0; i < type; ++i) {
		bx += m->xy[2 * nd[i]] / type;
		by += m->xy[2 * nd[i] + 1] / type;
	}
	m->bary[2 * e] = bx;
	m->bary[2 * e + 1] = by;

	double len[4] = {0}, nx[4] = {0}, ny[4] = {0};
	for (int i = 0; i < type; ++i) {
		int j = (i + 1) % type;
		double xi = m->xy[2 * nd[i]], yi = m->xy[2 * nd[i] + 1];
		double xj = m->xy[2 * nd[j]], yj = m->xy[2 * nd[j] + 1];
		double vx = xj - xi, vy = yj - yi;
		len[i] = sqrt(vx * vx + vy * vy);
		nx[i] = yj - yi;
		ny[i] = xi - xj;
		double t = sqrt(nx[i] * nx[i] + ny[i] * ny[i]);
		nx[i] /= t;
		ny[i] /= t;
	}
	double area = 0.5 * len[0] * len[1] * fabs(nx[0] * ny[1] - nx[1] * ny[0]);
	if (type == 4)
		area += 0.5 * len[2] * len[3] * fabs(nx[2] * ny[3] - nx[3] * ny[2]);
	m->area[e] = area;
	m->areaq[e] = 1.0 / area;

	if (m->totalArea_q != 0.0)
		m->totalArea_q = 1.0 / (1.0 / m->totalArea_q + area);
	else
		m->totalArea_q = m->areaq[e];

	double sx = 0.0, sy = 0.0;
	for (int i = 0; i < type; ++i) {
		sx += fabs(nx[i]) * len[i] / 2.0;
		sy += fabs(ny[i]) * len[i] / 2.0;
	}
	m->sx[e] = sx;
	m->sy[e] = sy;
}


/*
 * createSideInfo (src/mesh.c:250-289) for the list side in slot s. "conn" is
 * either another slot (>= 0) or a ghost -(g + 1). Writes this side's n, len,
 * GP and the connection side's GP, n, len (and the ghost barycentre).
 */
static void side_info(ccfd_mesh *m, long s, long e, double *connN,
		double *connLen, double *connGP)
{
	long a = m->slotNode[2 * s], b = m->slotNode[2 * s + 1];
	double x0 = m->xy[2 * a], y0 = m->xy[2 * a + 1];
	double x1 = m->xy[2 * b], y1 = m->xy[2 * b + 1];
	double nx = y1 - y0, ny = x0 - x1;
	double len = sqrt(nx * nx + ny * ny);
	nx /= len;
	ny /= len;
	double gx = 0.5 * (x0 + x1), gy = 0.5 * (y0 + y1);
	double GPx = gx - m->bary[2 * e], GPy = gy - m->bary[2 * e + 1];
	if ((nx * GPx + ny * GPy) < 0.0) {
		nx *= -1.0;
		ny *= -1.0;
	}
	m->n[2 * s] = nx;
	m->n[2 * s + 1] = ny;
	m->len[s] = len;
	m->GP[2 * s] = GPx;
	m->GP[2 * s + 1] = GPy;

	long c = m->nbr[s];
	double cbx, cby;
	if (c < 0) {
		long g = -c - 1;
		double tmp = 2.0 * fabs(GPx * nx + GPy * ny);
		cbx = m->bary[2 * e] + tmp * nx;
		cby = m->bary[2 * e + 1] + tmp * ny;
		m->ghostBary[2 * g] = cbx;
		m->ghostBary[2 * g + 1] = cby;
	} else {
		cbx = m->bary[2 * c];
		cby = m->bary[2 * c + 1];
	}
	connGP[0] = gx - cbx;
	connGP[1] = gy - cby;
	connN[0] = -nx;
	connN[1] = -ny;
	*connLen = len;
}

/* slotElem: the element of every slot (3 nTri + 4 nQuad entries), filled
 * here */
static ccfd_mesh *mesh_build(long nVert, const double *xy,
		long nTri, const long *tri, long nQuad, const long *quad,
		long nBCedge, const long *bcEdge,
		int nBC, const ccfd_bc *bc, char *err, long *slotElem)
{
	ccfd_mesh *m = xcalloc(1, sizeof(*m));
	if (!m) {
		SETERR("out of memory");
		return NULL;
	}
	long nE = nTri + nQuad;
	m->nNodes = nVert;
	m->nElems = nE;
	m->nTrias = nTri;
	m->nQuads = nQuad;
	m->nBC = nBC;
	m->xy = malloc(sizeof(double) * 2 * (nVert ? nVert : 1));
	m->bc = xcalloc(nBC, sizeof(ccfd_bc));
	if (!m->xy || !m->bc)
		goto oom;
	memcpy(m->xy, xy, sizeof(double) * 2 * nVert);
	if (nBC)
		memcpy(m->bc, bc, sizeof(ccfd_bc) * nBC);

	if (nVert > 0) {
		m->xMin = m->xMax = xy[0];
		m->yMin = m->yMax = xy[1];
		for (long i = nVert - 1; i >= 0; --i) {
			m->xMin = fmin(xy[2 * i], m->xMin);
			m->xMax = fmax(xy[2 * i], m->xMax);
			m->yMin = fmin(xy[2 * i + 1], m->yMin);
			m->yMax = fmax(xy[2 * i + 1], m->yMax);
		}
	}

	m->elemType = malloc(sizeof(int) * (nE ? nE : 1));
	m->domain = malloc(sizeof(int) * (nE ? nE : 1));
	m->elemNode = malloc(sizeof(long) * 4 * (nE ? nE : 1));
	m->slotStart = malloc(sizeof(long) * (nE + 1));
	m->bary = xcalloc(2 * nE, sizeof(double));
	m->area = xcalloc(nE, sizeof(double));
	m->areaq = xcalloc(nE, sizeof(double));
	m->sx = xcalloc(nE, sizeof(double));
	m->sy = xcalloc(nE, sizeof(double));
	if (!m->elemType || !m->domain || !m->elemNode || !m->slotStart ||
			!m->bary || !m->area || !m->areaq || !m->sx || !m->sy)
		goto oom;

	long nSlots = 0;
	for (long e = 0; e < nE; ++e) {
		int type = e < nTri ? 3 : 4;
		const long *src = e < nTri ? tri + 4 * e : quad + 5 * (e - nTri);
		m->elemType[e] = type;
		m->domain[e] = (int)src[type];
		for (int i = 0; i < 4; ++i)
			m->elemNode[4 * e + i] = i < type ? src[i] : -1;
		for (int i = 0; i < type; ++i) {
			if (src[i] < 0 || src[i] >= nVert) {
				SETERR("element %ld references node %ld (of %ld)",
						e, src[i], nVert);
				ccfd_mesh_free(m);
				return NULL;
			}
		}
		m->slotStart[e] = nSlots;
		for (int i = 0; i < type; ++i)
			slotElem[nSlots + i] = e;
		nSlots += type;
	}
	m->slotStart[nE] = nSlots;
	m->nSlots = nSlots;

	if ((nSlots + nBCedge) % 2 != 0) {
		SETERR("side count %ld + %ld BC edges is odd: open mesh",
				nSlots, nBCedge);
		ccfd_mesh_free(m);
		return NULL;
	}

	m->nbr = malloc(sizeof(long) * (nSlots ? nSlots : 1));
	m->partner = malloc(sizeof(long) * (nSlots ? nSlots : 1));
	m->owner = xcalloc(nSlots, 1);
	m->n = xcalloc(2 * nSlots, sizeof(double));
	m->len = xcalloc(nSlots, sizeof(double));
	m->fn = xcalloc(2 * nSlots, sizeof(double));
	m->flen = xcalloc(nSlots, sizeof(double));
	m->GP = xcalloc(2 * nSlots, sizeof(double));
	m->GPc = xcalloc(2 * nSlots, sizeof(double));
	m->w = xcalloc(2 * nSlots, sizeof(double));
	m->bbv = xcalloc(2 * nSlots, sizeof(double));
	m->bbd = xcalloc(nSlots, sizeof(double));
	m->slotNode = malloc(sizeof(long) * 2 * (nSlots ? nSlots : 1));
	if (!m->nbr || !m->partner || !m->owner || !m->n || !m->len ||
			!m->fn || !m->flen || !m->GP || !m->GPc || !m->w ||
			!m->bbv || !m->bbd || !m->slotNode)
		goto oom;

	/* side creation (src/mesh.c:1532-1566): side iSide joins nodes iSide and
	 * iSide+1 and is prepended, so list position = type - 1 - iSide */
	long nEntries = nSlots + nBCedge;
	pair_entry *pe = malloc(sizeof(pair_entry) * (nEntries ? nEntries : 1));
	if (!pe)
		goto oom;
	long create = 0;
	for (long e = 0; e < nE; ++e) {
		int type = m->elemType[e];
		for (int is = 0; is < type; ++is) {
			long s = m->slotStart[e] + (type - 1 - is);
			long a = m->elemNode[4 * e + is];
			long b = m->elemNode[4 * e + (is + 1) % type];
			m->slotNode[2 * s] = a;
			m->slotNode[2 * s + 1] = b;
			pe[create].n0 = a < b ? a : b;
			pe[create].n1 = a < b ? b : a;
			pe[create].bc = 0;
			pe[create].create = create;
			pe[create].slot = s;
			create++;
		}
	}
	int *edgeBC = malloc(sizeof(int) * (nBCedge ? nBCedge : 1));
	if (!edgeBC) {
		free(pe);
		goto oom;
	}
	for (long k = 0; k < nBCedge; ++k) {
		long a = bcEdge[3 * k], b = bcEdge[3 * k + 1];
		long code = bcEdge[3 * k + 2];
		int type = (int)(code / 100), id = (int)(code % 100);
		edgeBC[k] = -1;
		for (int i = 0; i < nBC; ++i) {
			if (bc[i].type == type && bc[i].id == id) {
				edgeBC[k] = i;
				break;
			}
		}
		if (type != 0 && edgeBC[k] < 0) {
			SETERR("BC %ld from mesh file not defined in parameter file",
					code);
			free(pe);
			free(edgeBC);
			ccfd_mesh_free(m);
			return NULL;
		}
		if (a < 0 || a >= nVert || b < 0 || b >= nVert) {
			SETERR("BC edge %ld references a missing node", k);
			free(pe);
			free(edgeBC);
			ccfd_mesh_free(m);
			return NULL;
		}
		pe[create].n0 = a < b ? a : b;
		pe[create].n1 = a < b ? b : a;
		pe[create].bc = 1;
		pe[create].create = create;
		pe[create].slot = -(k + 1);
		create++;
	}
	{
		pair_entry *sorted = sort_pairs(pe, nEntries, nVert);
		free(pe);
		pe = sorted;
		if (!pe) {
			free(edgeBC);
			goto oom;
		}
	}

	/* pairing (src/mesh.c:1713-1755); count ghosts first */
	long nPairs = nEntries / 2, nBCpairs = 0;
	for (long p = 0; p < nPairs; ++p) {
		pair_entry *A = pe + 2 * p, *B = pe + 2 * p + 1;
		if (A->n0 != B->n0 || A->n1 != B->n1) {
			SETERR("problem in connect: unmatched side (%ld,%ld)",
					A->n0, A->n1);
			free(pe);
			free(edgeBC);
			ccfd_mesh_free(m);
			return NULL;
		}
		if (A->bc || B->bc)
			nBCpairs++;
	}
	/* ghost k for every BC pair; periodic ones are dropped after hookup */
	long *bcPairSlot = malloc(sizeof(long) * (nBCpairs ? nBCpairs : 1));
	int *bcPairBC = malloc(sizeof(int) * (nBCpairs ? nBCpairs : 1));
	double *bcGhostBary = xcalloc(2 * nBCpairs, sizeof(double));
	double *bcGhostGP = xcalloc(2 * nBCpairs, sizeof(double));
	double *bcGhostN = xcalloc(2 * nBCpairs, sizeof(double));
	if (!bcPairSlot || !bcPairBC || !bcGhostBary || !bcGhostGP || !bcGhostN) {
		free(pe); free(edgeBC); free(bcPairSlot); free(bcPairBC);
		free(bcGhostBary); free(bcGhostGP); free(bcGhostN);
		goto oom;
	}
	long kb = 0;
	for (long p = 0; p < nPairs; ++p) {
		pair_entry *A = pe + 2 * p, *B = pe + 2 * p + 1;
		if (A->bc || B->bc) {
			if (A->bc) {
				pair_entry *t = A;
				A = B;
				B = t;
			}
			if (A->bc) {
				SETERR("two boundary edges share nodes (%ld,%ld)",
						A->n0, A->n1);
				free(pe); free(edgeBC); free(bcPairSlot);
				free(bcPairBC); free(bcGhostBary); free(bcGhostGP);
				free(bcGhostN);
				ccfd_mesh_free(m);
				return NULL;
			}
			long s = A->slot, k = -B->slot - 1;
			m->nbr[s] = -(kb + 1);
			m->partner[s] = -1;
			m->owner[s] = 1;
			bcPairSlot[kb] = s;
			bcPairBC[kb] = edgeBC[k];
			kb++;
		} else {
			long sa = A->slot, sb = B->slot;
			m->nbr[sa] = slotElem[sb];
			m->nbr[sb] = slotElem[sa];
			m->partner[sa] = sb;
			m->partner[sb] = sa;
			m->owner[sa] = 1;
			m->owner[sb] = 0;
		}
	}
	free(pe);
	free(edgeBC);
	/* temporary ghost barycentres live in bcGhostBary via m->ghostBary */
	m->ghostBary = bcGhostBary;

	/* createElemInfo for all elements in order */
	m->totalArea_q = 0.0;
	for (long e = 0; e < nE; ++e)
		elem_info(m, e);

	/* createSideInfo for every side of the side list (owners, including
	 * both sides of future periodic pairs) */
	for (long s = 0; s < nSlots; ++s) {
		if (!m->owner[s])
			continue;
		double cn[2], cl, cgp[2];
		side_info(m, s, slotElem[s], cn, &cl, cgp);
		long c = m->nbr[s];
		if (c < 0) {
			long g = -c - 1;
			bcGhostGP[2 * g] = cgp[0];
			bcGhostGP[2 * g + 1] = cgp[1];
			bcGhostN[2 * g] = cn[0];
			bcGhostN[2 * g + 1] = cn[1];
		} else {
			long t = m->partner[s];
			m->n[2 * t] = cn[0];
			m->n[2 * t + 1] = cn[1];
			m->len[t] = cl;
			m->GP[2 * t] = cgp[0];
			m->GP[2 * t + 1] = cgp[1];
		}
	}

	/* connectPeriodicBC (src/mesh.c:172-244). GP position of a BC side is
	 * its (ghost) GP plus the ghost barycentre; each x1 side takes the
	 * first x2 side of its group, in boundary-side order, within 1e-7 (L1)
	 * of its position plus the connection vector. A match is unique for
	 * any mesh whose periodic sides are more than 1e-7 apart. */
	long nPer2 = 0;
	for (long k = 0; k < nBCpairs; ++k) {
		int ib = bcPairBC[k];
		if (ib >= 0 && bc[ib].type == CCFD_BC_PERIODIC && bc[ib].id % 10 == 2)
			nPer2++;
	}
	long nPeriodic = 0;
	if (nPer2 > 0) {
		long *cand = malloc(sizeof(long) * nPer2);
		double *key = malloc(sizeof(double) * nPer2);
		unsigned char *used = xcalloc(nPer2, 1);
		if (!cand || !key || !used) {
			free(cand); free(key); free(used);
			free(bcPairSlot); free(bcPairBC); free(bcGhostGP);
			free(bcGhostN);
			goto oom;
		}
		long q = 0;
		for (long k = 0; k < nBCpairs; ++k) {
			int ib = bcPairBC[k];
			if (ib >= 0 && bc[ib].type == CCFD_BC_PERIODIC &&
					bc[ib].id % 10 == 2)
				cand[q++] = k;
		}
		for (long k = 0; k < nBCpairs; ++k) {
			int ib = bcPairBC[k];
			if (!(ib >= 0 && bc[ib].type == CCFD_BC_PERIODIC &&
					bc[ib].id % 10 == 1))
				continue;
			double ax = bcGhostGP[2 * k] + bcGhostBary[2 * k];
			double ay = bcGhostGP[2 * k + 1] + bcGhostBary[2 * k + 1];
			int group = bc[ib].id / 10;
			long hit = -1;
			for (long j = 0; j < nPer2; ++j) {
				long kk = cand[j];
				int jb = bcPairBC[kk];
				if (bc[jb].id / 10 != group)
					continue;
				double sxp = bcGhostGP[2 * kk] + bcGhostBary[2 * kk];
				double syp = bcGhostGP[2 * kk + 1] + bcGhostBary[2 * kk + 1];
				if ((fabs(ax + bc[ib].connection[0] - sxp) +
				     fabs(ay + bc[ib].connection[1] - syp)) <= 1e-7) {
					hit = kk;
					used[j] = 1;
					break;
				}
			}
			if (hit < 0) {
				SETERR("connectPeriodicBC: no connection for side GP %g %g",
						ax, ay);
				free(cand); free(key); free(used);
				free(bcPairSlot); free(bcPairBC); free(bcGhostGP);
				free(bcGhostN);
				ccfd_mesh_free(m);
				return NULL;
			}
			long ur = bcPairSlot[k], tg = bcPairSlot[hit];
			m->nbr[ur] = slotElem[tg];
			m->nbr[tg] = slotElem[ur];
			m->partner[ur] = tg;
			m->partner[tg] = ur;
			m->owner[ur] = 1;
			m->owner[tg] = 0;
			bcPairSlot[k] = -1;
			bcPairSlot[hit] = -1;
			nPeriodic++;
		}
		free(cand);
		free(key);
		free(used);
	}
	(void)nPeriodic;

	/* compact the remaining (non-periodic) ghosts: BCside[] */
	long nG = 0;
	for (long k = 0; k < nBCpairs; ++k)
		if (bcPairSlot[k] >= 0)
			nG++;
	m->nGhosts = nG;
	m->ghostSlot = malloc(sizeof(long) * (nG ? nG : 1));
	m->ghostBC = malloc(sizeof(int) * (nG ? nG : 1));
	m->ghostBary = xcalloc(2 * nG, sizeof(double));
	m->ghostGP = xcalloc(2 * nG, sizeof(double));
	if (!m->ghostSlot || !m->ghostBC || !m->ghostBary || !m->ghostGP) {
		free(bcPairSlot); free(bcPairBC); free(bcGhostGP); free(bcGhostN);
		free(bcGhostBary);
		goto oom;
	}
	long g = 0;
	for (long k = 0; k < nBCpairs; ++k) {
		long s = bcPairSlot[k];
		if (s < 0)
			continue;
		m->ghostSlot[g] = s;
		m->ghostBC[g] = bcPairBC[k];
		m->ghostBary[2 * g] = bcGhostBary[2 * k];
		m->ghostBary[2 * g + 1] = bcGhostBary[2 * k + 1];
		m->ghostGP[2 * g] = bcGhostGP[2 * k];
		m->ghostGP[2 * g + 1] = bcGhostGP[2 * k + 1];
		m->nbr[s] = -(g + 1);
		g++;
	}
	free(bcPairSlot);
	free(bcPairBC);
	free(bcGhostBary);
	free(bcGhostN);

	/* connection GP, flux normal/length, face list */
	long nF = 0;
	for (long s = 0; s < nSlots; ++s)
		nF += m->owner[s];
	m->nFaces = nF;
	m->faceSlot = malloc(sizeof(long) * (nF ? nF : 1));
	if (!m->faceSlot) {
		free(bcGhostGP);
		goto oom;
	}
	long f = 0;
	for (long s = 0; s < nSlots; ++s) {
		long c = m->nbr[s];
		if (c < 0) {
			long gg = -c - 1;
			m->GPc[2 * s] = m->ghostGP[2 * gg];
			m->GPc[2 * s + 1] = m->ghostGP[2 * gg + 1];
		} else {
			long t = m->partner[s];
			m->GPc[2 * s] = m->GP[2 * t];
			m->GPc[2 * s + 1] = m->GP[2 * t + 1];
		}
		long o = (m->owner[s] || c < 0) ? s : m->partner[s];
		m->fn[2 * s] = m->n[2 * o];
		m->fn[2 * s + 1] = m->n[2 * o + 1];
		m->flen[s] = m->len[o];
		if (m->owner[s])
			m->faceSlot[f++] = s;
	}
	free(bcGhostGP);

	/* createReconstructionInfo (src/mesh.c:80-114), per element */
	for (long e = 0; e < nE; ++e) {
		double r11 = 0.0, r12 = 0.0, r22 = 0.0;
		for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
			double bx = m->GP[2 * s] - m->GPc[2 * s];
			double by = m->GP[2 * s + 1] - m->GPc[2 * s + 1];
			m->bbv[2 * s] = bx;
			m->bbv[2 * s + 1] = by;
			m->bbd[s] = sqrt(bx * bx + by * by);
			r11 += bx * bx;
			r12 += bx * by;
			r22 += by * by;
		}
		r11 = sqrt(r11);
		r12 = r12 / r11;
		r22 = sqrt(r22 - r12 * r12);
		for (long s = m->slotStart[e]; s < m->slotStart[e + 1]; ++s) {
			double bx = m->bbv[2 * s], by = m->bbv[2 * s + 1];
			double alpha1 = bx / (r11 * r11);
			double alpha2 = (by - r12 / r11 * bx) / (r22 * r22);
			m->w[2 * s] = alpha1 - r12 / r11 * alpha2;
			m->w[2 * s + 1] = alpha2;
		}
	}
	return m;

oom:
	SETERR("out of memory");
	ccfd_mesh_free(m);
	return NULL;
}

ccfd_mesh *ccfd_mesh_build(long nVert, const double *xy,
		long nTri, const long *tri, long nQuad, const long *quad,
		long nBCedge, const long *bcEdge,
		int nBC, const ccfd_bc *bc, char *err)
{
	if (nTri < 0 || nQuad < 0) {
		SETERR("negative element count");
		return NULL;
	}
	long *slotElem = malloc(sizeof(long) * (3 * nTri + 4 * nQuad + 1));
	if (!slotElem) {
		SETERR("out of memory");
		return NULL;
	}
	ccfd_mesh *m = mesh_build(nVert, xy, nTri, tri, nQuad, quad, nBCedge,
			bcEdge, nBC, bc, err, slotElem);
	free(slotElem);
	return m;
}

void ccfd_mesh_venk_eps(const ccfd_mesh *m, double venk_k, double *eps_sq)
{
	for (long e = 0; e < m->nElems; ++e) {
		eps_sq[e] = (venk_k * sqrt(m->area[e])) *
			(venk_k * sqrt(m->area[e])) *
			(venk_k * sqrt(m->area[e]));
	}
}

int ccfd_cart_connectivity(long iMax, long jMax, double x0, double y0,
		double x1, double y1, const int nSeg[4], const int bcType[4][20],
		const int bcRange[4][20][2],
		long *nVert, double **xyOut, long *nQuad, long **quadOut,
		long *nBCedge, long **bcOut)
{
	long nv = (iMax + 1) * (jMax + 1);
	long nq = iMax * jMax;
	long nb = 0;
	for (int sd = 0; sd < 4; ++sd)
		for (int k = 0; k < nSeg[sd]; ++k)
			nb += bcRange[sd][k][1] - bcRange[sd][k][0] + 1;
	double *xy = malloc(sizeof(double) * 2 * nv);
	long *quad = malloc(sizeof(long) * 5 * (nq ? nq : 1));
	long *bce = malloc(sizeof(long) * 3 * (nb ? nb : 1));
	if (!xy || !quad || !bce) {
		free(xy); free(quad); free(bce);
		return -1;
	}
	double dx = (x1 - x0) / iMax;
	double dy = (y1 - y0) / jMax;
	double y = y0, x;
	long k = 0;
	for (long j = 0; j < jMax + 1; ++j) {
		x = x0;
		for (long i = 0; i < iMax + 1; ++i) {
			xy[2 * k] = x;
			xy[2 * k + 1] = y;
			k++;
			x += dx;
		}
		y += dy;
	}
	k = 0;
	for (long j = 0; j < jMax; ++j) {
		for (long i = 0; i < iMax; ++i) {
			long q0 = k + j;
			quad[5 * k] = q0;
			quad[5 * k + 1] = q0 + 1;
			quad[5 * k + 2] = q0 + 1 + iMax + 1;
			quad[5 * k + 3] = q0 + 1 + iMax + 1 - 1;
			quad[5 * k + 4] = 1;
			k++;
		}
	}
	k = 0;
	for (int s = 0; s < nSeg[0]; ++s)             /* bottom */
		for (long i = bcRange[0][s][0] - 1; i < bcRange[0][s][1]; ++i) {
			bce[3 * k] = i;
			bce[3 * k + 1] = i + 1;
			bce[3 * k + 2] = bcType[0][s];
			k++;
		}
	for (int s = 0; s < nSeg[1]; ++s)             /* right */
		for (long j = bcRange[1][s][0] - 1; j < bcRange[1][s][1]; ++j) {
			bce[3 * k] = (iMax + 1) * (j + 1) - 1;
			bce[3 * k + 1] = (iMax + 1) * (j + 2) - 1;
			bce[3 * k + 2] = bcType[1][s];
			k++;
		}
	for (int s = 0; s < nSeg[2]; ++s)             /* top */
		for (long i = bcRange[2][s][0] - 1; i < bcRange[2][s][1]; ++i) {
			bce[3 * k] = (iMax + 1) * jMax + i;
			bce[3 * k + 1] = bce[3 * k] + 1;
			bce[3 * k + 2] = bcType[2][s];
			k++;
		}
	for (int s = 0; s < nSeg[3]; ++s)             /* left */
		for (long j = bcRange[3][s][0] - 1; j < bcRange[3][s][1]; ++j) {
			bce[3 * k] = j * (iMax + 1);
			bce[3 * k + 1] = bce[3 * k] + iMax + 1;
			bce[3 * k + 2] = bcType[3][s];
			k++;
		}
	*nVert = nv;
	*xyOut = xy;
	*nQuad = nq;
	*quadOut = quad;
	*nBCedge = nb;
	*bcOut = bce;
	return 0;
}

int ccfd_gmsh_write(const char *path, long nVert, const double *xy,
		long nTri, const long *tri, long nQuad, const long *quad,
		long nBCedge, const long *bcEdge)
{
	FILE *f = fopen(path, "w");
	if (!f)
		return -1;
	fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%ld\n", nVert);
	for (long i = 0; i < nVert; ++i)
		fprintf(f, "%ld %.17g %.17g 0\n", i + 1, xy[2 * i], xy[2 * i + 1]);
	fprintf(f, "$EndNodes\n$Elements\n%ld\n", nBCedge + nTri + nQuad);
	long id = 1;
	for (long k = 0; k < nBCedge; ++k, ++id)
		fprintf(f, "%ld 1 2 %ld %ld %ld %ld\n", id, bcEdge[3 * k + 2],
				bcEdge[3 * k + 2], bcEdge[3 * k] + 1,
				bcEdge[3 * k + 1] + 1);
	for (long k = 0; k < nTri; ++k, ++id)
		fprintf(f, "%ld 2 2 %ld %ld %ld %ld %ld\n", id, tri[4 * k + 3],
				tri[4 * k + 3], tri[4 * k] + 1, tri[4 * k + 1] + 1,
				tri[4 * k + 2] + 1);
	for (long k = 0; k < nQuad; ++k, ++id)
		fprintf(f, "%ld 3 2 %ld %ld %ld %ld %ld %ld\n", id,
				quad[5 * k + 4], quad[5 * k + 4], quad[5 * k] + 1,
				quad[5 * k + 1] + 1, quad[5 * k + 2] + 1,
				quad[5 * k + 3] + 1);
	fprintf(f, "$EndElements\n");
	return fclose(f) == 0 ? 0 : -1;
}

/* find the line after `header` (a full line); returns 0 if absent */
static int seek_section(FILE *f, const char *header, char *line, int n)
{
	rewind(f);
	while (fgets(line, n, f))
		if (!strcmp(line, header))
			return fgets(line, n, f) != NULL;
	return 0;
}

/*
 * gmsh format 4 (readGmsh case 4, src/mesh.c:742-975). As the reference:
 * boundary codes are the first physical tag of each curve entity that has
 * one, assigned to the line-element blocks in block order; triangles and
 * quadrilaterals get domain 1; nodes must be numbered 1..n in file order;
 * element blocks of other types are not read past (their element lines
 * are taken as the next block headers, as the reference's loop does).
 */
static int gmsh4_read(FILE *f, long *nVertOut, double **xyOut,
		long *nTriOut, long **triOut, long *nQuadOut, long **quadOut,
		long *nBCedgeOut, long **bcOut, char *err)
{
	char line[4096];
	long nP = 0, nC = 0;
	if (!seek_section(f, "$Entities\n", line, sizeof(line)) ||
			sscanf(line, "%ld %ld", &nP, &nC) != 2 || nC <= 0) {
		SETERR("no curves in mesh file");
		return -1;
	}
	for (long l = 0; l < nP; ++l)
		if (!fgets(line, sizeof(line), f)) {
			SETERR("truncated $Entities section");
			return -1;
		}
	long *pTag = malloc(sizeof(long) * nC), nTag = 0;
	if (!pTag)
		return -1;
	for (long l = 0; l < nC; ++l) {
		if (!fgets(line, sizeof(line), f)) {
			SETERR("truncated $Entities section");
			free(pTag);
			return -1;
		}
		/* tag, bounding box (6), number of physical tags, tags ... */
		char *tok = strtok(line, " ");
		for (int k = 0; k < 7 && tok; ++k)
			tok = strtok(NULL, " ");
		if (!tok)
			continue;
		if (strtol(tok, NULL, 10)) {
			tok = strtok(NULL, " ");
			if (tok)
				pTag[nTag++] = strtol(tok, NULL, 10);
		}
	}
	long nBlocks = 0, nv = 0;
	if (!seek_section(f, "$Nodes\n", line, sizeof(line)) ||
			sscanf(line, "%ld %ld", &nBlocks, &nv) != 2 || nv <= 0) {
		SETERR("no nodes in mesh file");
		free(pTag);
		return -1;
	}
	double *xy = malloc(sizeof(double) * 2 * nv);
	long *id = malloc(sizeof(long) * nv);
	if (!xy || !id) {
		free(pTag); free(xy); free(id);
		return -1;
	}
	long iv = 0;
	for (long b = 0; b < nBlocks; ++b) {
		long nin = 0;
		if (!fgets(line, sizeof(line), f) ||
				sscanf(line, "%*d %*d %*d %ld", &nin) != 1 ||
				nin < 0 || iv + nin > nv) {
			SETERR("bad $Nodes block %ld", b);
			goto fail_nodes;
		}
		for (long i = 0; i < nin; ++i)
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%ld", &id[iv + i]) != 1) {
				SETERR("bad node tag in block %ld", b);
				goto fail_nodes;
			}
		for (long i = 0; i < nin; ++i, ++iv) {
			double x, y;
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%lg %lg", &x, &y) != 2 ||
					id[iv] != iv + 1) {
				SETERR("node %ld does not match its position", iv + 1);
				goto fail_nodes;
			}
			xy[2 * iv] = x;
			xy[2 * iv + 1] = y;
		}
	}
	if (iv != nv) {
		SETERR("$Nodes holds %ld of %ld nodes", iv, nv);
		goto fail_nodes;
	}
	free(id);
	id = NULL;
	long ne = 0;
	if (!seek_section(f, "$Elements\n", line, sizeof(line)) ||
			sscanf(line, "%ld %ld", &nBlocks, &ne) != 2 || ne <= 0) {
		SETERR("no elements in mesh file");
		free(pTag); free(xy);
		return -1;
	}
	long *bce = malloc(sizeof(long) * 3 * ne);
	long *tri = malloc(sizeof(long) * 4 * ne);
	long *quad = malloc(sizeof(long) * 5 * ne);
	long nb = 0, nt = 0, nq = 0;
	if (!bce || !tri || !quad)
		goto fail_elems;
	for (long b = 0; b < nBlocks; ++b) {
		int type = 0;
		long nin = 0;
		if (!fgets(line, sizeof(line), f) ||
				sscanf(line, "%*d %*d %d %ld", &type, &nin) != 2 ||
				nin < 0 || nb + nin > ne || nt + nin > ne ||
				nq + nin > ne) {
			SETERR("bad $Elements block %ld", b);
			goto fail_elems;
		}
		if (type == 1 && b >= nTag) {
			SETERR("line block %ld has no curve with a physical tag", b);
			goto fail_elems;
		}
		for (long i = 0; i < nin && type >= 1 && type <= 3; ++i) {
			long n[4];
			int need = type + 1;
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%*d %ld %ld %ld %ld", &n[0], &n[1],
						&n[2], &n[3]) < need) {
				SETERR("bad element line in block %ld", b);
				goto fail_elems;
			}
			for (int k = 0; k < need; ++k)
				if (n[k] < 1 || n[k] > nv) {
					SETERR("element node %ld out of range", n[k]);
					goto fail_elems;
				}
			if (type == 1) {
				bce[3 * nb] = n[0] - 1;
				bce[3 * nb + 1] = n[1] - 1;
				bce[3 * nb + 2] = pTag[b];
				nb++;
			} else if (type == 2) {
				for (int k = 0; k < 3; ++k)
					tri[4 * nt + k] = n[k] - 1;
				tri[4 * nt + 3] = 1;
				nt++;
			} else {
				for (int k = 0; k < 4; ++k)
					quad[5 * nq + k] = n[k] - 1;
				quad[5 * nq + 4] = 1;
				nq++;
			}
		}
	}
	free(pTag);
	*nVertOut = nv; *xyOut = xy;
	*nTriOut = nt; *triOut = tri;
	*nQuadOut = nq; *quadOut = quad;
	*nBCedgeOut = nb; *bcOut = bce;
	return 0;
fail_elems:
	free(bce); free(tri); free(quad);
	free(pTag); free(xy);
	return -1;
fail_nodes:
	free(pTag); free(xy); free(id);
	return -1;
}

/* first line containing `word` (strstr, as the reference), stopping at a
 * line containing "End"; returns 0 if not found */
static int emc2_seek(FILE *f, const char *word, char *line, int n)
{
	rewind(f);
	while (fgets(line, n, f)) {
		if (strstr(line, word))
			return fgets(line, n, f) != NULL;
		if (strstr(line, "End"))
			return 0;
	}
	return 0;
}

/* EMC2 / medit .mesh (readEMC2, src/mesh.c:1004-1141): vertices, edges with
 * a nonzero reference as boundary edges (their reference is the BC code),
 * triangles and quadrilaterals with their reference as domain */
int ccfd_emc2_read(const char *path, long *nVertOut, double **xyOut,
		long *nTriOut, long **triOut, long *nQuadOut, long **quadOut,
		long *nBCedgeOut, long **bcOut, char *err)
{
	FILE *f = fopen(path, "r");
	if (!f) {
		SETERR("could not open mesh file '%s'", path);
		return -1;
	}
	char line[1024];
	long nv = 0, nE = 0, nt = 0, nq = 0, nb = 0;
	double *xy = NULL;
	long *bce = NULL, *tri = NULL, *quad = NULL;
	if (!emc2_seek(f, "Vertices", line, sizeof(line)) ||
			sscanf(line, "%ld", &nv) != 1 || nv <= 0) {
		SETERR("no vertices in mesh file");
		goto fail;
	}
	xy = malloc(sizeof(double) * 2 * nv);
	if (!xy)
		goto fail;
	for (long i = 0; i < nv; ++i)
		if (!fgets(line, sizeof(line), f) ||
				sscanf(line, "%lg %lg", &xy[2 * i], &xy[2 * i + 1]) != 2) {
			SETERR("bad vertex %ld", i + 1);
			goto fail;
		}
	if (emc2_seek(f, "Edges", line, sizeof(line)) &&
			sscanf(line, "%ld", &nE) == 1 && nE > 0) {
		bce = malloc(sizeof(long) * 3 * nE);
		if (!bce)
			goto fail;
		for (long i = 0; i < nE; ++i) {
			long a, b, r;
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%ld %ld %ld", &a, &b, &r) != 3) {
				SETERR("bad edge %ld", i + 1);
				goto fail;
			}
			if (r == 0)
				continue;
			if (a < 1 || a > nv || b < 1 || b > nv) {
				SETERR("edge node out of range");
				goto fail;
			}
			bce[3 * nb] = a - 1;
			bce[3 * nb + 1] = b - 1;
			bce[3 * nb + 2] = r;
			nb++;
		}
	}
	if (emc2_seek(f, "Triangles", line, sizeof(line)) &&
			sscanf(line, "%ld", &nt) == 1 && nt > 0) {
		tri = malloc(sizeof(long) * 4 * nt);
		if (!tri)
			goto fail;
		for (long i = 0; i < nt; ++i) {
			long n[4];
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%ld %ld %ld %ld", &n[0], &n[1], &n[2],
						&n[3]) != 4) {
				SETERR("bad triangle %ld", i + 1);
				goto fail;
			}
			for (int k = 0; k < 3; ++k) {
				if (n[k] < 1 || n[k] > nv) {
					SETERR("triangle node out of range");
					goto fail;
				}
				tri[4 * i + k] = n[k] - 1;
			}
			tri[4 * i + 3] = n[3];
		}
	} else {
		nt = 0;
	}
	if (emc2_seek(f, "Quadrangles", line, sizeof(line)) &&
			sscanf(line, "%ld", &nq) == 1 && nq > 0) {
		quad = malloc(sizeof(long) * 5 * nq);
		if (!quad)
			goto fail;
		for (long i = 0; i < nq; ++i) {
			long n[5];
			if (!fgets(line, sizeof(line), f) ||
					sscanf(line, "%ld %ld %ld %ld %ld", &n[0], &n[1],
						&n[2], &n[3], &n[4]) != 5) {
				SETERR("bad quadrangle %ld", i + 1);
				goto fail;
			}
			for (int k = 0; k < 4; ++k) {
				if (n[k] < 1 || n[k] > nv) {
					SETERR("quadrangle node out of range");
					goto fail;
				}
				quad[5 * i + k] = n[k] - 1;
			}
			quad[5 * i + 4] = n[4];
		}
	} else {
		nq = 0;
	}
	fclose(f);
	if (!bce)
		bce = malloc(sizeof(long) * 3);
	if (!tri)
		tri = malloc(sizeof(long) * 4);
	if (!quad)
		quad = malloc(sizeof(long) * 5);
	*nVertOut = nv; *xyOut = xy;
	*nTriOut = nt; *triOut = tri;
	*nQuadOut = nq; *quadOut = quad;
	*nBCedgeOut = nb; *bcOut = bce;
	return 0;
fail:
	fclose(f);
	free(xy); free(bce); free(tri); free(quad);
	return -1;
}

int ccfd_gmsh_read(const char *path, long *nVertOut, double **xyOut,
		long *nTriOut, long **triOut, long *nQuadOut, long **quadOut,
		long *nBCedgeOut, long **bcOut, char *err)
{
	FILE *f = fopen(path, "r");
	if (!f) {
		SETERR("could not open mesh file '%s'", path);
		return -1;
	}
	char line[1024];
	int fmt = 0;
	while (fgets(line, sizeof(line), f)) {
		if (!strcmp(line, "$MeshFormat\n")) {
			if (fgets(line, sizeof(line), f))
				sscanf(line, "%d", &fmt);
			break;
		}
	}
	if (fmt == 4) {
		int rc = gmsh4_read(f, nVertOut, xyOut, nTriOut, triOut, nQuadOut,
				quadOut, nBCedgeOut, bcOut, err);
		fclose(f);
		return rc;
	}
	if (fmt != 2) {
		SETERR("gmsh format %d is not supported (2 or 4)", fmt);
		fclose(f);
		return -1;
	}
	long nv = 0;
	rewind(f);
	while (fgets(line, sizeof(line), f)) {
		if (!strcmp(line, "$Nodes\n")) {
			if (fgets(line, sizeof(line), f))
				sscanf(line, "%ld", &nv);
			break;
		}
	}
	if (nv <= 0) {
		SETERR("no nodes in mesh file");
		fclose(f);
		return -1;
	}
	double *xy = malloc(sizeof(double) * 2 * nv);
	if (!xy) {
		fclose(f);
		return -1;
	}
	for (long i = 0; i < nv; ++i) {
		long id;
		double x, y;
		if (!fgets(line, sizeof(line), f) ||
				sscanf(line, "%ld %lg %lg", &id, &x, &y) != 3 ||
				id != i + 1) {
			SETERR("node %ld does not match its position", i + 1);
			free(xy);
			fclose(f);
			return -1;
		}
		xy[2 * i] = x;
		xy[2 * i + 1] = y;
	}
	long ne = 0;
	rewind(f);
	while (fgets(line, sizeof(line), f)) {
		if (!strcmp(line, "$Elements\n")) {
			if (fgets(line, sizeof(line), f))
				sscanf(line, "%ld", &ne);
			break;
		}
	}
	long *bce = malloc(sizeof(long) * 3 * (ne ? ne : 1));
	long *tri = malloc(sizeof(long) * 4 * (ne ? ne : 1));
	long *quad = malloc(sizeof(long) * 5 * (ne ? ne : 1));
	if (!bce || !tri || !quad) {
		free(xy); free(bce); free(tri); free(quad);
		fclose(f);
		return -1;
	}
	long nb = 0, nt = 0, nq = 0;
	for (long i = 0; i < ne; ++i) {
		if (!fgets(line, sizeof(line), f))
			break;
		char *tok = strtok(line, " \n");
		long v[16];
		int nt_ = 0;
		while (tok && nt_ < 16) {
			v[nt_++] = strtol(tok, NULL, 10);
			tok = strtok(NULL, " \n");
		}
		if (nt_ < 4)
			continue;
		long type = v[1], ntags = v[2], pg = v[3];
		long *nodes = v + 3 + ntags;
		if (type == 1 && pg > 100) {
			bce[3 * nb] = nodes[0] - 1;
			bce[3 * nb + 1] = nodes[1] - 1;
			bce[3 * nb + 2] = pg;
			nb++;
		} else if (type == 2) {
			for (int k = 0; k < 3; ++k)
				tri[4 * nt + k] = nodes[k] - 1;
			tri[4 * nt + 3] = pg;
			nt++;
		} else if (type == 3) {
			for (int k = 0; k < 4; ++k)
				quad[5 * nq + k] = nodes[k] - 1;
			quad[5 * nq + 4] = pg;
			nq++;
		}
	}
	fclose(f);
	*nVertOut = nv;
	*xyOut = xy;
	*nTriOut = nt;
	*triOut = tri;
	*nQuadOut = nq;
	*quadOut = quad;
	*nBCedgeOut = nb;
	*bcOut = bce;
	return 0;
}
