/*
 * driver.c - the reference's host loop on the GPU path (include/ccfd_dropin.h).
 *
 * One global case as in the reference (src/main.c:35-130): the ini is read
 * with the reference's consumption order (initOutput, initEquation,
 * initBoundary, initMesh, initInitialCondition, initFV, initTimeDisc,
 * initLinearSolver), the mesh is preprocessed on the host (mesh.c) and the
 * state lives on the device; timeDisc reproduces the reference's loop
 * (src/timeDiscretization.c:559-777) with the reference's output formats
 * (CGNS through libccfd_cgns.so, curve, CSV), restarts from a CGNS flow
 * solution, and the analysis consumers of src/analyze.c (residual file,
 * CL / CD / CP, record points, error norms).
 */
#define _GNU_SOURCE   /* dladdr */
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ccfd_cgns.h"
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
	double dt;                 /* the last time step (elem->dt) */
	int sourceFunc;
	int icType, exactFunc;
	double alpha, ref0[4];     /* refState[0] and alpha of the IC */
	int haveRef;
	/* analysis (src/analyze.c, src/timeDiscretization.c:180-213) */
	int hasExactSolution, calcWing;
	double refLength;
	int wallId;
	double clAbortResidual, cdAbortResidual;
	int abortOnCl, abortOnCd;
	char abortName[4];
	long nWingP, nWingS;
	double *wingXY, *cp, cl, cd;
	long nRec;
	long *recElem;
	FILE **recFile;
	FILE *resFile;
	/* restart (src/main.c:64-94) and output bookkeeping (src/output.c) */
	int restart;
	long iniIter;
	double timeOverall;
	char gridFile[600];
	long nOut, capOut;
	double *outTime;
	long *outIter;
	long nBCedge;
	long *bcEdge;
	double printTime0;
} C;

/* libccfd_cgns.so, loaded on the first CGNS file */
static struct {
	void *h;
	int (*write_mesh)(const char *, const ccfd_mesh *, long, const long *, char *);
	int (*write_solution)(const char *, const char *, const ccfd_mesh *,
			const double *, double, double, char *);
	int (*write_master)(const char *, const char *, const ccfd_mesh *, int,
			const double *, const long *, const char *const *, char *);
	int (*read_solution)(const char *, long, double *, double *, double *, char *);
	int (*read_mesh)(const char *, long *, double **, long *, long **, long *,
			long **, long *, long **, char *);
} CG;

static int cgns_load(char *e)
{
	if (CG.h)
		return 0;
	Dl_info di;
	char path[1024];
	const char *dir = ".";
	int n = 1;
	if (dladdr((void *)cgns_load, &di) && di.dli_fname) {
		const char *sl = strrchr(di.dli_fname, '/');
		if (sl) {
			dir = di.dli_fname;
			n = (int)(sl - di.dli_fname);
		}
	}
	snprintf(path, sizeof(path), "%.*s/libccfd_cgns.so", n, dir);
	void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
	if (!h) {
		snprintf(e, 256, "CGNS files need libccfd_cgns.so (HDF5): %s", dlerror());
		return -1;
	}
	*(void **)&CG.write_mesh = dlsym(h, "ccfd_cgns_write_mesh");
	*(void **)&CG.write_solution = dlsym(h, "ccfd_cgns_write_solution");
	*(void **)&CG.write_master = dlsym(h, "ccfd_cgns_write_master");
	*(void **)&CG.read_solution = dlsym(h, "ccfd_cgns_read_solution");
	*(void **)&CG.read_mesh = dlsym(h, "ccfd_cgns_read_mesh");
	if (!CG.write_mesh || !CG.write_solution || !CG.write_master ||
			!CG.read_solution || !CG.read_mesh) {
		snprintf(e, 256, "libccfd_cgns.so lacks the CGNS entry points");
		dlclose(h);
		return -1;
	}
	CG.h = h;
	return 0;
}

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

/* initAnalyze's files (src/analyze.c:200-282): the residual history for
 * stationary or wing runs, gnuplot scripts, one file per record point */
static int open_analysis_files(char *e)
{
	char name[600];
	if (C.calcWing || C.stationary) {
		snprintf(name, sizeof(name), "%s_analysis.csv", C.fileName);
		/* a restart appends to an existing file (src/analyze.c:209-234) */
		FILE *old = C.restart ? fopen(name, "r") : NULL;
		if (old)
			fclose(old);
		C.resFile = fopen(name, old ? "a" : "w");
		if (!C.resFile) {
			snprintf(e, 256, "cannot open %.200s", name);
			return -1;
		}
		if (old)
			;
		else if (C.calcWing)
			fprintf(C.resFile, "Iter, Time, Residual(%s), CL, CD\n", C.abortName);
		else
			fprintf(C.resFile, "Iter, Time, Residual(RHO), Residual(VX), Residual(VY), Residual(E)\n");
		char res[600];
		snprintf(res, sizeof(res), "%s", name);
		snprintf(name, sizeof(name), "%s_residuals.dem", C.fileName);
		FILE *d = fopen(name, "w");
		if (d) {
			fprintf(d, "set title 'Residual Plot'\n");
			fprintf(d, "set logscale y\n");
			if (C.calcWing) {
				fprintf(d, "plot '%s' using 1:3 title '%s' with lines\n", res, C.abortName);
			} else {
				fprintf(d, "plot '%s' using 1:3 title 'RHO' with lines, \\\n", res);
				fprintf(d, "     '%s' using 1:4 title 'MX' with lines, \\\n", res);
				fprintf(d, "     '%s' using 1:5 title 'MY' with lines, \\\n", res);
				fprintf(d, "     '%s' using 1:6 title 'E' with lines\n", res);
			}
			fprintf(d, "pause -1");
			fclose(d);
		}
		if (C.calcWing) {
			snprintf(name, sizeof(name), "%s_CLCD.dem", C.fileName);
			d = fopen(name, "w");
			if (d) {
				fprintf(d, "set title 'CL/CD Plot'\n");
				fprintf(d, "plot '%s' using 1:4 title 'CL' with lines, \\\n", res);
				fprintf(d, "     '%s' using 1:5 title 'CD' with lines\n", res);
				fprintf(d, "pause -1");
				fclose(d);
			}
		}
	}
	for (long k = 0; k < C.nRec; ++k) {
		snprintf(name, sizeof(name), "%s_recordPoint_%ld.csv", C.fileName, k);
		C.recFile[k] = fopen(name, "w");
		if (!C.recFile[k]) {
			snprintf(e, 256, "cannot open %.200s", name);
			return -1;
		}
		fprintf(C.recFile[k], "Time, Density, VelocityX, VelocityY, Pressure\n");
	}
	return 0;
}

static void close_analysis_files(void)
{
	if (C.resFile)
		fclose(C.resFile);
	C.resFile = NULL;
	for (long k = 0; k < C.nRec; ++k)
		if (C.recFile && C.recFile[k]) {
			fclose(C.recFile[k]);
			C.recFile[k] = NULL;
		}
}

#define NEED(x, what) do { if ((x) != 0) { snprintf(e, 256, "missing ini key %s", what); goto fail; } } while (0)

int ccfd_init(const char *iniFile, int device, char *err)
{
	return ccfd_init_restart(iniFile, NULL, device, err);
}

int ccfd_init_restart(const char *iniFile, const char *restartFile, int device,
		char *err)
{
	char e[256] = "";
	ccfd_ini *ini = NULL;
	double *xy = NULL;
	long *tri = NULL, *quad = NULL, *bce = NULL;
	ccfd_bc bcs[32];
	double *prim = NULL;
	double restartTime = 0.0;
	ccfd_finalize();
	memset(&C, 0, sizeof(C));
	ini = ccfd_ini_read(iniFile, e);
	if (!ini)
		goto fail;
	/* main: stationary flag (src/main.c:54) */
	NEED(ccfd_ini_bool(ini, "stationary", "T", &C.stationary), "stationary");
	if (restartFile) {
		/* src/main.c:64-94: the time stamp after the file name's last
		 * '_' is the first iteration (stationary) or the start time */
		FILE *rf = fopen(restartFile, "r");
		if (!rf) {
			snprintf(e, 256, "No restart file found");
			goto fail;
		}
		fclose(rf);
		const char *u = strrchr(restartFile, '_');
		if (!u) {
			snprintf(e, 256, "restart file name %s has no time stamp", restartFile);
			goto fail;
		}
		double stamp = strtod(u + 1, NULL);
		C.restart = 1;
		if (C.stationary) {
			C.iniIter = (long)(int)stamp;
		} else {
			C.iniIter = 0;
			restartTime = stamp;
		}
	}
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
			long sf;
			NEED(ccfd_ini_long(ini, "sourceFunction", NULL, &sf), "sourceFunction");
			C.sourceFunc = (int)sf;
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
		/* src/mesh.c:1419-1440: .msh / .msh2 / .msh4 gmsh, .mesh EMC2,
		 * .cgns CGNS (through libccfd_cgns.so) */
		int gmsh = !strcmp(fmt, ".msh") || !strcmp(fmt, ".msh2") ||
			!strcmp(fmt, ".msh4");
		int cgns = !strcmp(fmt, ".cgns");
		if (!gmsh && !cgns && strcmp(fmt, ".mesh") != 0) {
			snprintf(e, 256, "Unknown Mesh Format %s", fmt);
			goto fail;
		}
		/* relative to the ini file's directory */
		const char *slash = strrchr(iniFile, '/');
		if (file[0] != '/' && slash)
			snprintf(path, sizeof(path), "%.*s/%s%s", (int)(slash - iniFile),
					iniFile, file, fmt);
		else
			snprintf(path, sizeof(path), "%s%s", file, fmt);
		if (cgns) {
			if (cgns_load(e) || CG.read_mesh(path, &nv, &xy, &nt, &tri,
					&nq, &quad, &nb, &bce, e))
				goto fail;
		} else if ((gmsh ? ccfd_gmsh_read : ccfd_emc2_read)(path, &nv, &xy,
				&nt, &tri, &nq, &quad, &nb, &bce, e)) {
			goto fail;
		}
	}
	C.mesh = ccfd_mesh_build(nv, xy, nt, tri, nq, quad, nb, bce, (int)nBC, bcs, e);
	if (!C.mesh)
		goto fail;
	/* src/mesh.c:1918-1922: the grid file of CGNS output, written unless
	 * restarting */
	snprintf(C.gridFile, sizeof(C.gridFile), "%s_mesh.cgns", C.fileName);
	C.nBCedge = nb;
	C.bcEdge = bce;
	bce = NULL;
	if (C.outputFormat == 1 && !C.restart) {
		if (cgns_load(e) || CG.write_mesh(C.gridFile, C.mesh, C.nBCedge,
				C.bcEdge, e))
			goto fail;
	}
	/* initInitialCondition (src/initialCondition.c:36-90) */
	long icType;
	NEED(ccfd_ini_long(ini, "icType", NULL, &icType), "icType");
	long N = C.mesh->nElems;
	prim = calloc(4 * N, sizeof(double));
	C.icType = (int)icType;
	if (icType == 2) {
		long ef;
		NEED(ccfd_ini_long(ini, "exactFunc", NULL, &ef), "exactFunc");
		if (ef < 1 || ef > 6) {
			snprintf(e, 256, "unknown exact function %ld", ef);
			goto fail;
		}
		C.exactFunc = (int)ef;
		if (ef == 5) {
			NEED(ccfd_ini_dbl(ini, "RP_1D_interface", "0.5", &p->rp1Dinterface), "RP_1D_interface");
			NEED(ccfd_ini_dbl_array(ini, "StateLeft", 4, NULL, p->refState[0]), "StateLeft");
			NEED(ccfd_ini_dbl_array(ini, "StateRight", 4, NULL, p->refState[1]), "StateRight");
			memcpy(C.ref0, p->refState[0], sizeof(C.ref0));
			C.haveRef = 1;
		}
		/* setInitialCondition (src/initialCondition.c:205-211): the
		 * exact function at the barycentres, t = 0 */
		for (long i = 0; i < N; ++i)
			if (ccfd_exact_func((int)ef, p, C.mesh, C.mesh->bary + 2 * i,
					0.0, prim + 4 * i)) {
				snprintf(e, 256, "exact function %ld failed", ef);
				goto fail;
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
			C.alpha = alpha;   /* the global keeps the last domain's */
		}
		memcpy(C.ref0, st[0], sizeof(C.ref0));
		C.haveRef = 1;
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
		static const char *names[4] = {"RHO", "MX", "MY", "E"};
		NEED(ccfd_ini_dbl(ini, "abortResidual", "1e-6", &C.abortResidual), "abortResidual");
		NEED(ccfd_ini_long(ini, "abortVariable", "1", &av), "abortVariable");
		if (av < 1 || av > 4) {
			snprintf(e, 256, "abortVariable %ld out of range", av);
			goto fail;
		}
		C.abortVariable = (int)av - 1;
		snprintf(C.abortName, sizeof(C.abortName), "%s", names[av - 1]);
		/* src/timeDiscretization.c:202-217 */
		NEED(ccfd_ini_dbl(ini, "cl_abortResidual", "0.0", &C.clAbortResidual), "cl_abortResidual");
		NEED(ccfd_ini_dbl(ini, "cd_abortResidual", "0.0", &C.cdAbortResidual), "cd_abortResidual");
		if (C.clAbortResidual == 0.0 && C.cdAbortResidual == 0.0) {
		} else if (C.clAbortResidual > 0.0) {
			C.abortOnCl = 1;
		} else if (C.cdAbortResidual > 0.0) {
			C.abortOnCd = 1;
		} else {
			snprintf(e, 256, "wrong definition of abort residual");
			goto fail;
		}
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
	if (!rc && C.sourceFunc)
		rc = ccfd_gpu_set_source(C.ctx, C.sourceFunc);
	if (rc) {
		snprintf(e, 256, "%s", ccfd_gpu_last_error());
		goto fail;
	}
	/* initAnalyze (src/analyze.c:197-282) */
	NEED(ccfd_ini_bool(ini, "exactSolution", "F", &C.hasExactSolution), "exactSolution");
	NEED(ccfd_ini_bool(ini, "calcWing", "F", &C.calcWing), "calcWing");
	if (C.hasExactSolution && C.icType != 2) {
		snprintf(e, 256, "exactSolution needs an exact-function initial condition");
		goto fail;
	}
	if (C.calcWing) {
		long wid;
		NEED(ccfd_ini_dbl(ini, "referenceLength", "1.0", &C.refLength), "referenceLength");
		NEED(ccfd_ini_long(ini, "wall_id", NULL, &wid), "wall_id");
		C.wallId = (int)wid;
		if (!C.haveRef) {
			snprintf(e, 256, "calcWing needs a freestream (icType 1) state");
			goto fail;
		}
	}
	{
		long nr;
		NEED(ccfd_ini_long(ini, "nRecordPoints", "0", &nr), "nRecordPoints");
		if (nr < 0) {
			snprintf(e, 256, "nRecordPoints %ld", nr);
			goto fail;
		}
		C.nRec = nr;
		C.recElem = calloc(nr + 1, sizeof(long));
		C.recFile = calloc(nr + 1, sizeof(FILE *));
		if (!C.recElem || !C.recFile) {
			snprintf(e, 256, "out of memory");
			goto fail;
		}
		for (long k = 0; k < nr; ++k) {
			double x[2];
			NEED(ccfd_ini_dbl_array(ini, "pointCoords", 2, NULL, x), "pointCoords");
			C.recElem[k] = ccfd_mesh_locate(C.mesh, x);
			if (C.recElem[k] < 0) {
				snprintf(e, 256, "Record Point # %ld is not in Domain", k);
				goto fail;
			}
		}
	}
	if (C.calcWing) {
		C.wingXY = malloc(sizeof(double) * 3 * (C.mesh->nGhosts + 1));
		C.cp = malloc(sizeof(double) * (C.mesh->nGhosts + 1));
		if (!C.wingXY || !C.cp) {
			snprintf(e, 256, "out of memory");
			goto fail;
		}
		if (ccfd_gpu_wing_init(C.ctx, C.mesh, C.wallId, C.refLength,
				&C.nWingP, &C.nWingS, C.wingXY)) {
			snprintf(e, 256, "%s", ccfd_gpu_last_error());
			goto fail;
		}
	}
	/* initTimeDisc (src/timeDiscretization.c:226-239), then
	 * setInitialCondition: a restart reads the state, the time and the
	 * overall time from the CGNS file (src/initialCondition.c:94-158) */
	C.t = restartTime;
	C.printTime0 = (floor(restartTime / C.IOtimeInterval) + 1) * C.IOtimeInterval;
	if (C.restart) {
		double tr = 0.0, to = 0.0;
		if (cgns_load(e) || CG.read_solution(restartFile, N, prim, &tr, &to, e))
			goto fail;
		if (ccfd_gpu_set_prim(C.ctx, prim)) {
			snprintf(e, 256, "%s", ccfd_gpu_last_error());
			goto fail;
		}
		C.t = tr;
		C.timeOverall = to;
	}
	if (open_analysis_files(e))
		goto fail;
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
	close_analysis_files();
	free(C.outTime);
	free(C.outIter);
	free(C.bcEdge);
	C.outTime = NULL;
	C.outIter = NULL;
	C.bcEdge = NULL;
	C.nOut = C.capOut = 0;
	free(C.recElem);
	free(C.recFile);
	free(C.wingXY);
	free(C.cp);
	C.recElem = NULL;
	C.recFile = NULL;
	C.wingXY = NULL;
	C.cp = NULL;
	C.nRec = 0;
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

/* the rows (x, rho, u, p) of every element sorted by x with the
 * reference's bubble sort (stable, adjacent swaps; src/output.c:84-101) */
static double *sorted_rows(const double *prim)
{
	long N = C.mesh->nElems;
	double *rows = malloc(sizeof(double) * 4 * (N ? N : 1));
	if (!rows)
		return NULL;
	for (long i = 0; i < N; ++i) {
		rows[4 * i] = C.mesh->bary[2 * i];
		rows[4 * i + 1] = prim[4 * i];
		rows[4 * i + 2] = prim[4 * i + 1];
		rows[4 * i + 3] = prim[4 * i + 3];
	}
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
	return rows;
}

/* csvOutput (src/output.c:56-113) */
static int write_csv_prim(const char *path, const double *prim)
{
	double *rows = sorted_rows(prim);
	FILE *f = rows ? fopen(path, "w") : NULL;
	if (!f) {
		free(rows);
		return -1;
	}
	fprintf(f, "CoordinateX, Density, Velocity, Pressure\n");
	for (long i = 0; i < C.mesh->nElems; ++i)
		fprintf(f, "%15.9f,%15.9f,%15.9f,%15.9f\n", rows[4 * i],
				rows[4 * i + 1], rows[4 * i + 2], rows[4 * i + 3]);
	fclose(f);
	free(rows);
	return 0;
}

/* curveOutput (src/output.c:253-327): density, velocity and pressure
 * against x, one block each */
static int write_curve_prim(const char *path, const double *prim)
{
	double *rows = sorted_rows(prim);
	FILE *f = rows ? fopen(path, "w") : NULL;
	if (!f) {
		free(rows);
		return -1;
	}
	static const char *head[3] = {"#Density\n", "\n#Velocity\n", "\n#Pressure\n"};
	for (int c = 0; c < 3; ++c) {
		fprintf(f, "%s", head[c]);
		for (long i = 0; i < C.mesh->nElems; ++i)
			fprintf(f, "%15.9f,%15.9f\n", rows[4 * i], rows[4 * i + 1 + c]);
	}
	fclose(f);
	free(rows);
	return 0;
}

int ccfd_write_csv(const char *path)
{
	if (!C.ready)
		return -1;
	long N = C.mesh->nElems;
	double *prim = malloc(sizeof(double) * 4 * N);
	int rc = (!prim || ccfd_gpu_get_prim(C.ctx, prim)) ? -1 : write_csv_prim(path, prim);
	free(prim);
	return rc;
}

/* one file of dataOutput in the case's format: the flow (or, with exact,
 * the exact solution) at time / iteration iter (src/output.c:333-395) */
static void write_output(double time, long iter, const double *prim, int exact)
{
	char name[640], e[256] = "";
	const char *ex = exact ? "_ex" : "";
	if (C.stationary)
		snprintf(name, sizeof(name), "%s%s_%09ld", C.fileName, ex, iter);
	else
		snprintf(name, sizeof(name), "%s%s_%015.7f", C.fileName, ex, time);
	int rc;
	switch (C.outputFormat) {
	case 1:
		strcat(name, ".cgns");
		rc = cgns_load(e) || CG.write_solution(name, C.gridFile, C.mesh, prim,
				time, C.timeOverall, e);
		break;
	case 2:
		strcat(name, ".curve");
		rc = write_curve_prim(name, prim);
		break;
	case 3:
		strcat(name, ".csv");
		rc = write_csv_prim(name, prim);
		break;
	default:
		FAILX("Output Format unknown");
	}
	if (rc)
		FAILX("could not write %s %s", name, e);
}

/* dataOutput (src/output.c:333-395): the flow state, the exact solution
 * when the case has one, and the output time for the CGNS master file */
static void data_output(double time, long iter)
{
	if (C.nOut == C.capOut) {
		long cap = C.capOut ? 2 * C.capOut : 16;
		double *t = realloc(C.outTime, sizeof(double) * cap);
		if (t)
			C.outTime = t;
		long *it = realloc(C.outIter, sizeof(long) * cap);
		if (it)
			C.outIter = it;
		if (!t || !it)
			FAILX("out of memory");
		C.capOut = cap;
	}
	C.outTime[C.nOut] = time;
	C.outIter[C.nOut] = iter;
	C.nOut++;
	long N = C.mesh->nElems;
	double *prim = malloc(sizeof(double) * 4 * (N ? N : 1));
	if (!prim || ccfd_gpu_get_prim(C.ctx, prim))
		FAILX("could not read the state: %s", ccfd_gpu_last_error());
	write_output(time, iter, prim, 0);
	if (C.hasExactSolution) {
		for (long i = 0; i < N; ++i)
			if (ccfd_exact_func(C.exactFunc, &C.prm, C.mesh, C.mesh->bary + 2 * i,
					time, prim + 4 * i))
				FAILX("exact function %d failed", C.exactFunc);
		write_output(time, iter, prim, 1);
	}
	free(prim);
}

/* finalizeDataOutput (src/output.c:407-525): the CGNS master file linking
 * every flow solution written so far */
static void finalize_output(void)
{
	if (C.outputFormat != 1)
		return;
	char name[600], e[256] = "";
	snprintf(name, sizeof(name), "%s_Master.cgns", C.fileName);
	char **files = calloc((size_t)(C.nOut ? C.nOut : 1), sizeof(char *));
	if (!files)
		FAILX("out of memory");
	for (long k = 0; k < C.nOut; ++k) {
		files[k] = malloc(640);
		if (!files[k])
			FAILX("out of memory");
		if (C.stationary)
			snprintf(files[k], 640, "%s_%09ld.cgns", C.fileName, C.outIter[k]);
		else
			snprintf(files[k], 640, "%s_%015.7f.cgns", C.fileName, C.outTime[k]);
	}
	int rc = cgns_load(e) || CG.write_master(name, C.gridFile, C.mesh, (int)C.nOut,
			C.outTime, C.outIter, (const char *const *)files, e);
	for (long k = 0; k < C.nOut; ++k)
		free(files[k]);
	free(files);
	if (rc)
		FAILX("could not write %s %s", name, e);
}

/* ---- analysis consumers, src/analyze.c ---- */

void initAnalyze(void) { need_ready("initAnalyze"); }

/* globalResidual (src/analyze.c:521-539) of the last dU/dt; entries 4 and
 * 5 (the CL / CD residuals) are cleared, as the reference's memset does */
void globalResidual(double resIter[6])
{
	need_ready("globalResidual");
	memset(resIter, 0, 6 * sizeof(double));
	if (ccfd_gpu_residual(C.ctx, resIter))
		FAILX("globalResidual: %s", ccfd_gpu_last_error());
}

/* evalRecordPoints (src/analyze.c:393-402) */
static void eval_record_points(double time)
{
	double *pv = malloc(sizeof(double) * 4 * (C.nRec + 1));
	if (!pv)
		FAILX("record points: out of memory");
	if (ccfd_gpu_get_prim_at(C.ctx, (int)C.nRec, C.recElem, pv))
		FAILX("record points: %s", ccfd_gpu_last_error());
	for (long k = 0; k < C.nRec; ++k)
		fprintf(C.recFile[k], "%20.12f,%20.12f,%20.12f,%20.12f,%20.12f\n",
				time + C.dt, pv[4 * k], pv[4 * k + 1], pv[4 * k + 2],
				pv[4 * k + 3]);
	free(pv);
}

static void write_cp(const char *suffix, const char *head, long k0, long n)
{
	char name[600];
	snprintf(name, sizeof(name), "%s_CP_%s.csv", C.fileName, suffix);
	FILE *f = fopen(name, "w");
	if (!f)
		FAILX("Cannot open Output File for CP I/O");
	fprintf(f, "x, y, phi, %s\n", head);
	for (long k = k0; k < k0 + n; ++k)
		fprintf(f, "%15.7f,%15.7f,%15.7f,%15.7f\n", C.wingXY[3 * k],
				C.wingXY[3 * k + 1], C.wingXY[3 * k + 2], C.cp[k]);
	fclose(f);
}

/* calcCoef (src/analyze.c:292-387): CL and CD on the device, CP files */
static void calc_coef(void)
{
	double coef[2];
	if (ccfd_gpu_wing_coef(C.ctx, C.ref0, C.alpha, coef, C.cp))
		FAILX("calcCoef: %s", ccfd_gpu_last_error());
	write_cp("pressureSide", "CP_pressureSide", 0, C.nWingP);
	write_cp("suctionSide", "CP_suctionSide", C.nWingP, C.nWingS);
	C.cl = coef[0];
	C.cd = coef[1];
	char name[600], ps[600], ss[600];
	snprintf(name, sizeof(name), "%s_CP.dem", C.fileName);
	snprintf(ps, sizeof(ps), "%s_CP_pressureSide.csv", C.fileName);
	snprintf(ss, sizeof(ss), "%s_CP_suctionSide.csv", C.fileName);
	FILE *d = fopen(name, "r");
	if (!d) {
		d = fopen(name, "w");
		if (d) {
			fprintf(d, "unset key\n");
			fprintf(d, "set xlabel 'x'\n");
			fprintf(d, "set ylabel 'cp'\n");
			fprintf(d, "set yrange [*:*] reverse\n");
			fprintf(d, "plot '%s' using 1:4 w l lc rgb 'black', \\\n", ps);
			fprintf(d, "     '%s' using 1:4 w l lc rgb 'black'\n", ss);
			fprintf(d, "pause -1");
		}
	}
	if (d)
		fclose(d);
}

/* analyze (src/analyze.c:411-438), after each step at the new time */
void analyze(double time, long iter, double resIter[6])
{
	need_ready("analyze");
	if (C.nRec > 0)
		eval_record_points(time);
	if (C.calcWing) {
		resIter[4] = C.cl;
		resIter[5] = C.cd;
		calc_coef();
		resIter[4] = fabs(resIter[4] - C.cl) / C.dt;
		resIter[5] = fabs(resIter[5] - C.cd) / C.dt;
		fprintf(C.resFile, "%7ld, %13.8f, %15.8e, %15.10f, %15.10f\n", iter,
				time + C.dt, resIter[C.abortVariable], C.cl, C.cd);
	} else if (C.stationary) {
		fprintf(C.resFile, "%7ld, %13.8f, %15.8e, %15.8e, %15.8e, %15.8e\n",
				iter, time + C.dt, resIter[0], resIter[1], resIter[2],
				resIter[3]);
	}
}

/* calcErrors (src/analyze.c:445-514) against the initial condition's exact
 * function */
void calcErrors(double time)
{
	need_ready("calcErrors");
	double L1[4], L2[4], Li[4];
	if (ccfd_gpu_calc_errors(C.ctx, C.exactFunc, time, L1, L2, Li))
		FAILX("calcErrors: %s", ccfd_gpu_last_error());
	printf("\nError Analysis at t = %g\n", time);
	printf("|                RHO           VX           VY            P\n");
	printf("| L1  :  %7.5e  %7.5e  %7.5e  %7.5e\n", L1[0], L1[1], L1[2], L1[3]);
	printf("| L2  :  %7.5e  %7.5e  %7.5e  %7.5e\n", L2[0], L2[1], L2[2], L2[3]);
	printf("| Linf:  %7.5e  %7.5e  %7.5e  %7.5e\n", Li[0], Li[1], Li[2], Li[3]);
}

void freeAnalyze(void) { }

static void print_coefficients(void)
{
	if (C.calcWing) {
		printf("| Final Lift and Drag Coefficients:\n");
		printf("|   CL: %.10g\n", C.cl);
		printf("|   CD: %.10g\n", C.cd);
	}
}

static void calc_dt(double t, double pTime, double *dt, int *vd)
{
	if (ccfd_gpu_calc_dt(C.ctx, t, pTime, C.stopTime, dt, vd))
		FAILX("%s", ccfd_gpu_last_error());
}

void timeDisc(void)
{
	need_ready("timeDisc");
	double t = C.t;
	/* src/timeDiscretization.c:238-239: from the restart file name's time
	 * stamp (0 without a restart) and the first iteration */
	double printTime = C.printTime0;
	long printIter = (C.iniIter / C.IOiterInterval + 1) * C.IOiterInterval;
	int hasConverged = C.stationary ? 0 : 1;
	data_output(t, C.iniIter);
	double dt;
	int vd;
	double t0 = wall(), tIO = t0;
	calc_dt(t, printTime, &dt, &vd);
	printf("| Initial Time Step: %.10g\n", dt);
	if (vd)
		printf("| Viscous Time Step Dominates!\n");
	long iter;
	for (iter = C.iniIter + 1; iter <= C.maxIter; ++iter) {
		calc_dt(t, printTime, &dt, &vd);
		C.dt = dt;
		double res[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
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
		analyze(t, iter, res);
		if (C.stopTime - t <= 1e-15) {
			printf("\nTime Limit Reached - Computation Complete!\n");
			printf("| Final Time      : %.10g\n", t);
			printf("| Iteration Number: %ld\n", iter);
			printf("| Writing Final State to Disk\n");
			data_output(t, iter);
			finalize_output();
			if (C.hasExactSolution)
				calcErrors(t);
			break;
		}
		/* residual abort criteria, src/timeDiscretization.c:629-688 */
		if (C.stationary && fabs(res[C.abortVariable]) <= C.abortResidual) {
			if (C.abortOnCl) {
				if (fabs(res[4]) <= C.clAbortResidual) {
					printf("\nConverged in CL - Calculation complete\n");
					printf("| Iteration number: %ld\n", iter);
					printf("| Writing final state to disk\n");
					print_coefficients();
					hasConverged = 1;
				}
			} else if (C.abortOnCd) {
				if (fabs(res[5]) <= C.cdAbortResidual) {
					printf("\nConverged in CD - Calculation complete\n");
					printf("| Iteration number: %ld\n", iter);
					printf("| Writing final state to disk\n");
					print_coefficients();
					hasConverged = 1;
				}
			} else {
				printf("\nConverged in '%s' - Calculation complete\n", C.abortName);
				printf("| Iteration number: %ld\n", iter);
				printf("| Writing final state to disk\n");
				print_coefficients();
				hasConverged = 1;
			}
		}
		if (C.stationary && hasConverged) {
			data_output(t, iter);
			finalize_output();
			if (C.hasExactSolution)
				calcErrors(t);
			break;
		}
		if ((printTime - t <= 1e-15) || (iter == printIter)) {
			printf("\nData Output at Iteration %ld\n", iter);
			double tNow = wall();
			printf("| Time since last I/O: %.10g s\n", tNow - tIO);
			tIO = tNow;
			if (C.stationary) {
				if (C.calcWing) {
					printf("| Residuals: %s: %20.14e\n", C.abortName, res[C.abortVariable]);
					printf("|            CL : %20.14e\n", res[4]);
					printf("|            CD : %20.14e\n", res[5]);
				} else {
					printf("|            RHO: %20.14e\n", res[0]);
					printf("|            MX : %20.14e\n", res[1]);
					printf("|            MY : %20.14e\n", res[2]);
					printf("|            E  : %20.14e\n", res[3]);
				}
			} else {
				printf("| Time     : %.10g\n", t);
				double dtp;
				calc_dt(t, printTime + 1e150, &dtp, &vd);
				printf("| Time Step: %.10g\n", dtp);
				if (vd)
					printf("| Viscous Time Step Dominates!\n");
			}
			if (C.hasExactSolution)
				calcErrors(t);
			data_output(t, iter);
			finalize_output();
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
		printf("| Final Time %g has not been reached.\n", C.stopTime);
		printf("| Current Time: %.10g\n", t);
		printf("| Final State will be written to disk\n");
		data_output(t, iter - 1);
		finalize_output();
		if (C.hasExactSolution)
			calcErrors(t);
	}
	printf("\nComputation Time: %.10g s\n", t1 - t0);
	if (C.implicit) {
		printf("| Newton Iterations: %ld\n", C.nNewton);
		printf("| GMRES Iterations : %ld\n", C.nGMRES);
	}
	close_analysis_files();
}

int ccfd_main(const char *iniFile, int device)
{
	return ccfd_main_restart(iniFile, NULL, device);
}

int ccfd_main_restart(const char *iniFile, const char *restartFile, int device)
{
	char err[256];
	if (ccfd_init_restart(iniFile, restartFile, device, err)) {
		printf("| ERROR: %s\n", err);
		return 1;
	}
	timeDisc();
	ccfd_finalize();
	return 0;
}
