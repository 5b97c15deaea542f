/*
 * driver.c - the reference's host loop on the GPU path (include/ccfd_dropin.h).
 *
 * One global case as in the reference (src/main.c:35-130): the ini is read
 * with the reference's consumption order (initOutput, initEquation,
 * initBoundary, initMesh, initInitialCondition, initFV, initTimeDisc,
 * initLinearSolver), the mesh is preprocessed on the host (mesh.c) and the
 * state lives on the device; timeDisc reproduces the reference's loop
 * (src/timeDiscretization.c:559-777) with CSV output.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ccfd_dropin.h"

#define FAILX(...) do { fprintf(stdout, "| ERROR: "); fprintf(stdout, __VA_ARGS__); fprintf(stdout, "\n"); exit(1); } while (0)

static struct {
	int ready;
	ccfd_mesh *mesh;
	ccfd_gpu_ctx *ctx;
	ccfd_gpu_params prm;
	ccfd_gpu_implicit_cfg icfg;
	int implicit;
	int stationary;
	char fileName[256];
	double IOtimeInterval;
	long IOiterInterval;
	int outputFormat;
	long maxIter;
	double stopTime;
	double t;
	long iter;
	double abortResidual;
	int abortVariable;
	long nNewton, nGMRES;
} C;

static double wall(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

ccfd_gpu_ctx *ccfd_case_context(void) { return C.ctx; }
const ccfd_mesh *ccfd_case_mesh(void) { return C.mesh; }
double ccfd_case_time(void) { return C.t; }
long ccfd_case_iterations(void) { return C.iter; }

#define NEED(x, what) do { if ((x) != 0) { snprintf(e, 256, "missing ini key %s", what); goto fail; } } while (0)

int ccfd_init(const char *iniFile, int device, char *err)
{
	char e[256] = "";
	ccfd_ini *ini = NULL;
	double *xy = NULL;
	long *tri = NULL, *quad = NULL, *bce = NULL;
	ccfd_bc bcs[32];
	double *prim = NULL;
	ccfd_finalize();
	memset(&C, 0, sizeof(C));
	ini = ccfd_ini_read(iniFile, e);
	if (!ini)
		goto fail;
	/* main: stationary flag (src/main.c:54) */
	NEED(ccfd_ini_bool(ini, "stationary", "T", &C.stationary), "stationary");
	/* initOutput (src/output.c:36-48) */
	NEED(ccfd_ini_str(ini, "fileName", NULL, C.fileName, sizeof(C.fileName)), "fileName");
	NEED(ccfd_ini_dbl(ini, "IOtimeInterval", NULL, &C.IOtimeInterval), "IOtimeInterval");
	{
		double d;
		NEED(ccfd_ini_dbl(ini, "IOiterInterval", NULL, &d), "IOiterInterval");
		C.IOiterInterval = (long)d;
		long of;
		NEED(ccfd_ini_long(ini, "outputFormat", "1", &of), "outputFormat");
		C.outputFormat = (int)of;
	}
	/* initEquation (src/equation.c:40-124) */
	ccfd_gpu_params *p = &C.prm;
	double Rgas;
	long flux;
	NEED(ccfd_ini_dbl(ini, "gamma", "1.4", &p->gam), "gamma");
	NEED(ccfd_ini_dbl(ini, "R", "287", &Rgas), "R");
	NEED(ccfd_ini_long(ini, "fluxFunction", NULL, &flux), "fluxFunction");
	p->iFlux = (int)flux;
	{
		int src;
		NEED(ccfd_ini_bool(ini, "calcSource", "F", &src), "calcSource");
		if (src) {
			snprintf(e, 256, "source terms are not supported");
			goto fail;
		}
	}
	p->mu = 0.0;
	p->Pr = 0.72;
	/* initBoundary (src/boundary.c:29-137) */
	long nBC;
	NEED(ccfd_ini_long(ini, "nBC", NULL, &nBC), "nBC");
	if (nBC < 0 || nBC > 32) {
		snprintf(e, 256, "nBC = %ld out of range", nBC);
		goto fail;
	}
	for (long b = 0; b < nBC; ++b) {
		long code;
		NEED(ccfd_ini_long(ini, "BCtype", NULL, &code), "BCtype");
		ccfd_bc *bc = bcs + b;
		memset(bc, 0, sizeof(*bc));
		bc->type = (int)(code / 100);
		bc->id = (int)(code % 100);
		double rho, Ma, alpha, pr;
		switch (bc->type) {
		case CCFD_BC_SLIPWALL:
		case CCFD_BC_OUTFLOW:
			break;
		case CCFD_BC_WALL:
			snprintf(e, 256, "BC 2xx needs the Navier-Stokes build (set navierstokes)");
			goto fail;
		case CCFD_BC_INFLOW:
		case CCFD_BC_CHARACTERISTIC: {
			NEED(ccfd_ini_dbl(ini, "rho", NULL, &rho), "rho");
			NEED(ccfd_ini_dbl(ini, "mach", NULL, &Ma), "mach");
			NEED(ccfd_ini_dbl(ini, "alpha", NULL, &alpha), "alpha");
			NEED(ccfd_ini_dbl(ini, "pressure", NULL, &pr), "pressure");
			double pi = acos(-1.0);
			double c = sqrt(p->gam * pr / rho), v = Ma * c;
			bc->pVar[0] = rho;
			bc->pVar[1] = v * cos(alpha * pi / 180.0);
			bc->pVar[2] = v * sin(alpha * pi / 180.0);
			bc->pVar[3] = pr;
			break;
		}
		case CCFD_BC_EXACTSOL: {
			long f;
			NEED(ccfd_ini_long(ini, "BCexactFunc", NULL, &f), "BCexactFunc");
			bc->exactFunc = (int)f;
			break;
		}
		case CCFD_BC_PERIODIC:
			NEED(ccfd_ini_dbl_array(ini, "connection", 2, NULL, bc->connection), "connection");
			break;
		case CCFD_BC_PRESSURE_OUT:
			NEED(ccfd_ini_dbl(ini, "pressure", NULL, &bc->pVar[3]), "pressure");
			break;
		default:
			snprintf(e, 256, "illegal boundary condition %ld", code);
			goto fail;
		}
	}
	/* initMesh (src/mesh.c:1849-1924) */
	long meshType, nv = 0, nt = 0, nq = 0, nb = 0;
	NEED(ccfd_ini_long(ini, "meshType", "1", &meshType), "meshType");
	if (meshType == 1) {
		long iMax, jMax;
		double x0[2], x1[2], nsegd[4];
		int nSeg[4], bcType[4][20] = {{0}}, bcRange[4][20][2] = {{{0}}};
		NEED(ccfd_ini_long(ini, "nElemsX", NULL, &iMax), "nElemsX");
		NEED(ccfd_ini_long(ini, "nElemsY", NULL, &jMax), "nElemsY");
		NEED(ccfd_ini_dbl_array(ini, "X0", 2, NULL, x0), "X0");
		NEED(ccfd_ini_dbl_array(ini, "Xmax", 2, NULL, x1), "Xmax");
		NEED(ccfd_ini_dbl_array(ini, "nBCsegments", 4, NULL, nsegd), "nBCsegments");
		for (int s = 0; s < 4; ++s) {
			nSeg[s] = (int)nsegd[s];
			if (nSeg[s] < 1 || nSeg[s] > 20) {
				snprintf(e, 256, "nBCsegments out of range");
				goto fail;
			}
			for (int k = 0; k < nSeg[s]; ++k) {
				long v;
				NEED(ccfd_ini_long(ini, "meshBCtype", NULL, &v), "meshBCtype");
				bcType[s][k] = (int)v;
				if (nSeg[s] == 1) {
					bcRange[s][k][0] = 1;
					bcRange[s][k][1] = (int)((s % 2 == 0) ? iMax : jMax);
				} else {
					long a, b2;
					NEED(ccfd_ini_long(ini, "BCstart", NULL, &a), "BCstart");
					NEED(ccfd_ini_long(ini, "BCend", NULL, &b2), "BCend");
					bcRange[s][k][0] = (int)a;
					bcRange[s][k][1] = (int)b2;
				}
			}
		}
		if (ccfd_cart_connectivity(iMax, jMax, x0[0], x0[1], x1[0], x1[1],
				nSeg, bcType, bcRange, &nv, &xy, &nq, &quad, &nb, &bce)) {
			snprintf(e, 256, "cartesian mesh generation failed");
			goto fail;
		}
	} else {
		char fmt[64], file[256], path[600];
		NEED(ccfd_ini_str(ini, "meshFormat", NULL, fmt, sizeof(fmt)), "meshFormat");
		NEED(ccfd_ini_str(ini, "meshFile", NULL, file, sizeof(file)), "meshFile");
		if (strncmp(fmt, ".msh", 4) != 0) {
			snprintf(e, 256, "mesh format %s not supported (gmsh v2 only)", fmt);
			goto fail;
		}
		/* relative to the ini file's directory */
		const char *slash = strrchr(iniFile, '/');
		if (file[0] != '/' && slash)
			snprintf(path, sizeof(path), "%.*s/%s%s", (int)(slash - iniFile),
					iniFile, file, fmt);
		else
			snprintf(path, sizeof(path), "%s%s", file, fmt);
		if (ccfd_gmsh_read(path, &nv, &xy, &nt, &tri, &nq, &quad, &nb, &bce, e))
			goto fail;
	}
	C.mesh = ccfd_mesh_build(nv, xy, nt, tri, nq, quad, nb, bce, (int)nBC, bcs, e);
	if (!C.mesh)
		goto fail;
	/* initInitialCondition (src/initialCondition.c:36-90) */
	long icType;
	NEED(ccfd_ini_long(ini, "icType", NULL, &icType), "icType");
	long N = C.mesh->nElems;
	prim = calloc(4 * N, sizeof(double));
	if (icType == 2) {
		long ef;
		NEED(ccfd_ini_long(ini, "exactFunc", NULL, &ef), "exactFunc");
		if (ef != 5) {
			snprintf(e, 256, "initial exact function %ld not supported", ef);
			goto fail;
		}
		NEED(ccfd_ini_dbl(ini, "RP_1D_interface", "0.5", &p->rp1Dinterface), "RP_1D_interface");
		NEED(ccfd_ini_dbl_array(ini, "StateLeft", 4, NULL, p->refState[0]), "StateLeft");
		NEED(ccfd_ini_dbl_array(ini, "StateRight", 4, NULL, p->refState[1]), "StateRight");
		/* exactFunc 5 at t = 0 (src/exactFunction.c:131-150) */
		for (long i = 0; i < N; ++i) {
			const double *s = C.mesh->bary[2 * i] <= p->rp1Dinterface ?
				p->refState[0] : p->refState[1];
			prim[4 * i] = s[0];
			prim[4 * i + 1] = s[1];
			prim[4 * i + 2] = 0.0;
			prim[4 * i + 3] = s[3];
		}
	} else if (icType == 1) {
		long nd;
		NEED(ccfd_ini_long(ini, "nDomains", "1", &nd), "nDomains");
		double st[20][4];
		if (nd < 1 || nd > 20) {
			snprintf(e, 256, "nDomains out of range");
			goto fail;
		}
		for (long d = 0; d < nd; ++d) {
			long id;
			double rho, Ma, alpha, pr;
			NEED(ccfd_ini_long(ini, "domainID", NULL, &id), "domainID");
			NEED(ccfd_ini_dbl(ini, "rho", NULL, &rho), "rho");
			NEED(ccfd_ini_dbl(ini, "mach", NULL, &Ma), "mach");
			NEED(ccfd_ini_dbl(ini, "alpha", NULL, &alpha), "alpha");
			NEED(ccfd_ini_dbl(ini, "pressure", NULL, &pr), "pressure");
			double pi = acos(-1.0), c = sqrt(p->gam * pr / rho), v = Ma * c;
			st[d][0] = rho;
			st[d][1] = v * cos(alpha * pi / 180.0);
			st[d][2] = v * sin(alpha * pi / 180.0);
			st[d][3] = pr;
		}
		for (long i = 0; i < N; ++i) {
			int d = nd == 1 ? 0 : C.mesh->domain[i] - 1;
			if (d < 0 || d >= nd) {
				snprintf(e, 256, "element %ld in unknown domain", i);
				goto fail;
			}
			memcpy(prim + 4 * i, st[d], 4 * sizeof(double));
		}
	} else {
		snprintf(e, 256, "illegal initial condition %ld", icType);
		goto fail;
	}
	/* initFV (src/finiteVolume.c:30-66) */
	long order, lim = 1;
	NEED(ccfd_ini_long(ini, "spatialOrder", "1", &order), "spatialOrder");
	p->spatialOrder = (int)order;
	p->limiter = 1;
	p->venk_k = 1.0;
	if (order == 2) {
		NEED(ccfd_ini_long(ini, "limiter", "1", &lim), "limiter");
		p->limiter = (int)lim;
		if (lim == 2)
			NEED(ccfd_ini_dbl(ini, "venk_k", "1", &p->venk_k), "venk_k");
	}
	/* initTimeDisc (src/timeDiscretization.c:69-240) */
	long torder, nrk;
	NEED(ccfd_ini_dbl(ini, "CFL", "0.9", &p->cfl), "CFL");
	NEED(ccfd_ini_dbl(ini, "DFL", "0.9", &p->dfl), "DFL");
	NEED(ccfd_ini_bool(ini, "timeStep1D", "F", &p->timeStep1D), "timeStep1D");
	NEED(ccfd_ini_long(ini, "timeOrder", "1", &torder), "timeOrder");
	p->timeOrder = (int)torder;
	NEED(ccfd_ini_bool(ini, "implicit", "F", &C.implicit), "implicit");
	NEED(ccfd_ini_long(ini, "nRKstages", "1", &nrk), "nRKstages");
	p->nRKstages = (int)nrk;
	if (C.stationary) {
		long av;
		NEED(ccfd_ini_dbl(ini, "abortResidual", "1e-6", &C.abortResidual), "abortResidual");
		NEED(ccfd_ini_long(ini, "abortVariable", "1", &av), "abortVariable");
		C.abortVariable = (int)av - 1;
	}
	NEED(ccfd_ini_long(ini, "maxIter", "100000", &C.maxIter), "maxIter");
	NEED(ccfd_ini_dbl(ini, "tEnd", NULL, &C.stopTime), "tEnd");
	/* initLinearSolver (src/linearSolver.c:61-107) */
	if (C.implicit) {
		long nk, nn;
		int pc;
		NEED(ccfd_ini_long(ini, "nKdim", "5", &nk), "nKdim");
		NEED(ccfd_ini_dbl(ini, "epsNewton", "0.001", &C.icfg.epsNewton), "epsNewton");
		NEED(ccfd_ini_dbl(ini, "epsGMRES", "0.001", &C.icfg.epsGMRES), "epsGMRES");
		NEED(ccfd_ini_long(ini, "nNewtonIter", "20", &nn), "nNewtonIter");
		NEED(ccfd_ini_dbl(ini, "gammaEW", "0.9", &C.icfg.gammaEW), "gammaEW");
		NEED(ccfd_ini_bool(ini, "precond", "F", &pc), "precond");
		C.icfg.nKdim = (int)nk;
		C.icfg.nNewtonIter = (int)nn;
		C.icfg.precond = pc;
	}
	p->precision = CCFD_FP64;
	/* device */
	int rc = ccfd_gpu_create(&C.ctx, device);
	if (!rc)
		rc = ccfd_gpu_set_mesh(C.ctx, C.mesh, p);
	if (!rc)
		rc = ccfd_gpu_set_prim(C.ctx, prim);
	if (rc) {
		snprintf(e, 256, "%s", ccfd_gpu_last_error());
		goto fail;
	}
	C.t = 0.0;
	C.ready = 1;
	free(prim); free(xy); free(tri); free(quad); free(bce);
	ccfd_ini_free(ini);
	return 0;
fail:
	if (err)
		snprintf(err, 256, "%s", e);
	free(prim); free(xy); free(tri); free(quad); free(bce);
	ccfd_ini_free(ini);
	ccfd_finalize();
	return -1;
}

void ccfd_finalize(void)
{
	if (C.ctx)
		ccfd_gpu_destroy(C.ctx);
	if (C.mesh)
		ccfd_mesh_free(C.mesh);
	C.ctx = NULL;
	C.mesh = NULL;
	C.ready = 0;
}

static void need_ready(const char *fn)
{
	if (!C.ready)
		FAILX("%s called before ccfd_init", fn);
}

void fvTimeDerivative(double time)
{
	need_ready("fvTimeDerivative");
	if (ccfd_gpu_fvtd(C.ctx, time))
		FAILX("fvTimeDerivative: %s", ccfd_gpu_last_error());
}

void convectiveFlux(double rhoL, double rhoR, double vxL, double vxR,
		double vyL, double vyR, double pL, double pR, double fluxLoc[4])
{
	double L[4] = {rhoL, vxL, vyL, pL}, R[4] = {rhoR, vxR, vyR, pR};
	int code = C.ready ? C.prm.iFlux : 5;
	double gam = C.ready ? C.prm.gam : 1.4;
	/* the reference calls this once per face: no allocation per call */
	if (ccfd_gpu_riemann_batch_noalloc(code, CCFD_FP64, gam, 1, L, R, fluxLoc))
		FAILX("convectiveFlux: %s", ccfd_gpu_last_error());
}

void initTimeDisc(void) { need_ready("initTimeDisc"); }
void initLinearSolver(void) { need_ready("initLinearSolver"); }
void freeLinearSolver(void) { }

double vectorDotProduct(double **A, double **B)
{
	need_ready("vectorDotProduct");
	long N = C.mesh->nElems;
	double *a = malloc(sizeof(double) * 4 * N), *b = malloc(sizeof(double) * 4 * N);
	if (!a || !b)
		FAILX("vectorDotProduct: out of memory");
	for (long i = 0; i < N; ++i)
		for (int v = 0; v < 4; ++v) {
			a[4 * i + v] = A[v][i];
			b[4 * i + v] = B[v][i];
		}
	double r = 0.0;
	if (ccfd_gpu_dot(C.ctx, a, b, &r))
		FAILX("vectorDotProduct: %s", ccfd_gpu_last_error());
	free(a);
	free(b);
	return r;
}

void GMRES_M(double time, double dt, double alpha, double **B, double normB,
		double *abortCrit, double **deltaX)
{
	need_ready("GMRES_M");
	long N = C.mesh->nElems;
	double *b = malloc(sizeof(double) * 4 * N), *x = malloc(sizeof(double) * 4 * N);
	if (!b || !x)
		FAILX("GMRES_M: out of memory");
	for (long i = 0; i < N; ++i)
		for (int v = 0; v < 4; ++v)
			b[4 * i + v] = B[v][i];
	int it = 0;
	*abortCrit = C.icfg.epsGMRES * normB;
	if (ccfd_gpu_gmres(C.ctx, &C.icfg, C.t, time, dt, alpha, b, normB, x, &it))
		FAILX("GMRES_M: %s", ccfd_gpu_last_error());
	for (long i = 0; i < N; ++i)
		for (int v = 0; v < 4; ++v)
			deltaX[v][i] = x[4 * i + v];
	C.nGMRES += it;
	free(b);
	free(x);
}

/* csvOutput (src/output.c:56-113): x-sorted rows (stable) of the element
 * barycentre x, density, x-velocity, pressure */
int ccfd_write_csv(const char *path)
{
	if (!C.ready)
		return -1;
	long N = C.mesh->nElems;
	double *prim = malloc(sizeof(double) * 4 * N);
	double *rows = malloc(sizeof(double) * 4 * N);
	if (!prim || !rows || ccfd_gpu_get_prim(C.ctx, prim)) {
		free(prim);
		free(rows);
		return -1;
	}
	for (long i = 0; i < N; ++i) {
		rows[4 * i] = C.mesh->bary[2 * i];
		rows[4 * i + 1] = prim[4 * i];
		rows[4 * i + 2] = prim[4 * i + 1];
		rows[4 * i + 3] = prim[4 * i + 3];
	}
	/* bubble sort by x, as the reference (stable, adjacent swaps) */
	long n = N;
	int swapped = 1;
	while (swapped && n > 0) {
		swapped = 0;
		for (long i = 0; i < n - 1; ++i) {
			if (rows[4 * i] > rows[4 * (i + 1)]) {
				double t[4];
				memcpy(t, rows + 4 * (i + 1), sizeof(t));
				memcpy(rows + 4 * (i + 1), rows + 4 * i, sizeof(t));
				memcpy(rows + 4 * i, t, sizeof(t));
				swapped = 1;
			}
		}
		n--;
	}
	FILE *f = fopen(path, "w");
	if (!f) {
		free(prim);
		free(rows);
		return -1;
	}
	fprintf(f, "CoordinateX, Density, Velocity, Pressure\n");
	for (long i = 0; i < N; ++i)
		fprintf(f, "%15.9f,%15.9f,%15.9f,%15.9f\n", rows[4 * i],
				rows[4 * i + 1], rows[4 * i + 2], rows[4 * i + 3]);
	fclose(f);
	free(prim);
	free(rows);
	return 0;
}

/* dataOutput (src/output.c:333-354): <fileName>_%015.7f.csv (transient) or
 * <fileName>_%09ld.csv (stationary); CSV only on this path */
static void data_output(double time, long iter)
{
	char name[600];
	if (C.stationary)
		snprintf(name, sizeof(name), "%s_%09ld.csv", C.fileName, iter);
	else
		snprintf(name, sizeof(name), "%s_%015.7f.csv", C.fileName, time);
	if (C.outputFormat != 3) {
		static int warned = 0;
		if (!warned)
			printf("| output format %d not available, writing CSV\n",
					C.outputFormat);
		warned = 1;
	}
	if (ccfd_write_csv(name))
		FAILX("could not write %s", name);
}

void timeDisc(void)
{
	need_ready("timeDisc");
	double t = C.t;
	double printTime = (floor(t / C.IOtimeInterval) + 1) * C.IOtimeInterval;
	long printIter = (0 / C.IOiterInterval + 1) * C.IOiterInterval;
	int hasConverged = C.stationary ? 0 : 1;
	(void)hasConverged;
	data_output(t, 0);
	double dt;
	int vd;
	double t0 = wall();
	if (ccfd_gpu_calc_dt(C.ctx, t, printTime, C.stopTime, &dt, &vd))
		FAILX("%s", ccfd_gpu_last_error());
	printf("| Initial Time Step: %.10g\n", dt);
	long iter;
	for (iter = 1; iter <= C.maxIter; ++iter) {
		if (ccfd_gpu_calc_dt(C.ctx, t, printTime, C.stopTime, &dt, &vd))
			FAILX("%s", ccfd_gpu_last_error());
		double res[4];
		if (!C.implicit) {
			if (ccfd_gpu_explicit_step(C.ctx, t, dt, res))
				FAILX("%s", ccfd_gpu_last_error());
		} else {
			int nn = 0, ng = 0;
			int rc = ccfd_gpu_implicit_step(C.ctx, &C.icfg, t, dt, res, &nn, &ng);
			C.nNewton += nn;
			C.nGMRES += ng;
			if (rc)
				FAILX("%s", ccfd_gpu_last_error());
		}
		t += dt;
		C.t = t;
		C.iter = iter;
		if (C.stopTime - t <= 1e-15) {
			printf("\nTime Limit Reached - Computation Complete!\n");
			printf("| Final Time      : %.10g\n", t);
			printf("| Iteration Number: %ld\n", iter);
			data_output(t, iter);
			break;
		}
		if (C.stationary && fabs(res[C.abortVariable]) <= C.abortResidual) {
			printf("\nConverged - Calculation complete\n");
			printf("| Iteration number: %ld\n", iter);
			data_output(t, iter);
			break;
		}
		if ((printTime - t <= 1e-15) || (iter == printIter)) {
			printf("\nData Output at Iteration %ld\n", iter);
			if (C.stationary)
				printf("|            RHO: %20.14e\n", res[0]);
			else
				printf("| Time     : %.10g\n", t);
			data_output(t, iter);
			if (printTime - t <= 1e-15)
				printTime += C.IOtimeInterval;
			if (iter == printIter)
				printIter += C.IOiterInterval;
		}
	}
	double t1 = wall();
	if (iter > C.maxIter) {
		printf("\nERROR - ERROR - ERROR\n");
		printf("| Maximum Iteration Number Reached. Calculation Aborted!\n");
		data_output(t, iter - 1);
	}
	printf("\nComputation Time: %.10g s\n", t1 - t0);
	if (C.implicit) {
		printf("| Newton Iterations: %ld\n", C.nNewton);
		printf("| GMRES Iterations : %ld\n", C.nGMRES);
	}
}

int ccfd_main(const char *iniFile, int device)
{
	char err[256];
	if (ccfd_init(iniFile, device, err)) {
		printf("| ERROR: %s\n", err);
		return 1;
	}
	timeDisc();
	ccfd_finalize();
	return 0;
}
