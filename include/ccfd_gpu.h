/*
 * ccfd_gpu.h - C ABI of the MI355X (gfx950) implementation of ccfd's
 * finite-volume operator and the time / linear-solver loops around it.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * Host arrays are in the reference's element order (as built by
 * ccfd_mesh_build), 4 doubles per element: primitive (rho, vx, vy, p) or
 * conservative (rho, mx, my, E). The device keeps its own renumbered
 * (space-filling-curve) structure-of-arrays copy; state stays resident
 * between calls and is only copied back on request.
 *
 * Every entry point returns CCFD_OK (0) or a negative status; the message of
 * the last failure on the calling thread is ccfd_gpu_last_error().
 * A context is used by one host thread at a time.
 */
#ifndef CCFD_GPU_H
#define CCFD_GPU_H

#include "ccfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

enum ccfd_status {
	CCFD_OK = 0,
	CCFD_EINVAL = -1,       /* bad argument / state */
	CCFD_ENOMEM = -2,       /* device allocation failed */
	CCFD_ENOTSUP = -3,      /* unsupported option */
	CCFD_EHIP = -4,         /* HIP runtime error */
	CCFD_ENONFINITE = -5,   /* non-finite time step (src/timeDiscretization.c:259-262) */
	CCFD_ENOCONV = -6,      /* Newton not converged (src/timeDiscretization.c:544-549) */
	CCFD_ESINGULAR = -7     /* singular LU-SGS diagonal block (src/linearSolver.c:245-249) */
};

/* arithmetic of the operator */
enum ccfd_precision {
	CCFD_FP64 = 0,          /* bit-identical drop-in */
	CCFD_FP32_FLUX = 1      /* FP32 reconstruction / flux / residual and
	                           preconditioner, FP64 state update, FP64
	                           Frechet difference and Arnoldi */
};

/*
 * Physics and discretisation settings; meaning and defaults follow the
 * reference's ini keys (initEquation src/equation.c:40-124, initFV
 * src/finiteVolume.c:30-66, initTimeDisc src/timeDiscretization.c:69-240).
 */
typedef struct ccfd_gpu_params {
	double gam;             /* gamma */
	double mu;              /* dynamic viscosity (Navier-Stokes only) */
	double Pr;              /* Prandtl number */
	int navierstokes;       /* 1: -Dnavierstokes semantics (viscous flux, WALL BC) */
	int iFlux;              /* flux function 1..11, src/main.h:76-88 */
	int spatialOrder;       /* 1 or 2 */
	int limiter;            /* 1 Barth-Jespersen, 2 Venkatakrishnan */
	double venk_k;          /* Venkatakrishnan K */
	double refState[2][4];  /* exact function 5 (1-D Riemann problem) */
	double rp1Dinterface;
	double cfl, dfl;        /* convective / diffusive CFL */
	int timeStep1D;         /* 1-D time step (src/timeDiscretization.c:253-271) */
	int nRKstages;          /* 1, 3, 4, 5 */
	int timeOrder;          /* 1, 2, 3 */
	int precision;          /* ccfd_precision */
} ccfd_gpu_params;

typedef struct ccfd_gpu_ctx ccfd_gpu_ctx;

/* last error message of the calling thread */
const char *ccfd_gpu_last_error(void);

/* create a context on a HIP device (one stream per context) */
int ccfd_gpu_create(ccfd_gpu_ctx **ctx, int device);
int ccfd_gpu_destroy(ccfd_gpu_ctx *ctx);

/*
 * Upload a preprocessed mesh with its parameters (the device side of
 * initMesh + initFV + initTimeDisc, src/main.c:101-108). Renumbers the
 * elements along a Hilbert curve; host I/O stays in reference order.
 */
int ccfd_gpu_set_mesh(ccfd_gpu_ctx *ctx, const ccfd_mesh *mesh,
		const ccfd_gpu_params *params);

/* Change parameters that do not alter the mesh (flux, order, limiter, ...) */
int ccfd_gpu_set_params(ccfd_gpu_ctx *ctx, const ccfd_gpu_params *params);

/* ---- element state, reference order, 4 doubles per element ---- */
/* primitive state; also sets the conservative state (primCons,
 * setInitialCondition src/initialCondition.c:218-222) */
int ccfd_gpu_set_prim(ccfd_gpu_ctx *ctx, const double *prim);
/* conservative state; primitive from consPrim (src/equationOfState.c:37-52) */
int ccfd_gpu_set_cons(ccfd_gpu_ctx *ctx, const double *cons);
int ccfd_gpu_get_prim(ccfd_gpu_ctx *ctx, double *prim);
int ccfd_gpu_get_cons(ccfd_gpu_ctx *ctx, double *cons);
/* dU/dt of the last operator evaluation (elem->u_t) */
int ccfd_gpu_get_ut(ccfd_gpu_ctx *ctx, double *ut);
/*
 * Side effects of the operator that host consumers read
 * (src/analyze.c:314-468): limited gradients u_x, u_y (4 per element) and
 * reconstructed side states (4 per slot, ccfd_mesh slot order) of the
 * current primitive state at the time of the last ccfd_gpu_fvtd call
 * (recomputed on demand: the fused operator keeps them in LDS only). Any
 * pointer may be NULL. At 1st order gradients are zero and side states
 * equal the element state.
 */
int ccfd_gpu_get_recon(ccfd_gpu_ctx *ctx, double *u_x, double *u_y,
		double *sideP);

/* ---- operator ---- */
/* fvTimeDerivative (src/finiteVolume.c:78-121) at the given time on the
 * resident primitive state; dU/dt stays on the device */
int ccfd_gpu_fvtd(ccfd_gpu_ctx *ctx, double time);
/*
 * Source term (calcSource, src/source.c:66-85; the reference's calcSource /
 * sourceFunction keys, src/equation.c:119-123): sourceFunc 0 none, 1 the
 * manufactured sine-wave source. Every later operator evaluation (fvtd,
 * explicit stages, implicit residuals) adds the quadrature of the source
 * at its own time. Survives ccfd_gpu_set_mesh.
 */
int ccfd_gpu_set_source(ccfd_gpu_ctx *ctx, int sourceFunc);

/* exact function exactFunc (1..6, src/exactFunction.c:35-183) at point x
 * and time on the host (the C library's transcendental functions, as the
 * reference): refState / rp1Dinterface come from p, the bounding box and
 * dxRef (src/mesh.c:1923) from the mesh. Returns CCFD_EINVAL for an unknown
 * function. */
int ccfd_exact_func(int exactFunc, const ccfd_gpu_params *p,
		const ccfd_mesh *mesh, const double x[2], double time,
		double pVar[4]);

/* ---- analysis (src/analyze.c) ---- */
/*
 * Interior side states at the boundary: for every boundary side (mesh
 * ghost order), the reconstructed primitive state of the interior side as
 * the last operator evaluation left it (side->pVar of the connection of the
 * reference's BC side, which calcCoef reads, src/analyze.c:314,344).
 * 4 doubles per ghost.
 */
int ccfd_gpu_get_bc_sides(ccfd_gpu_ctx *ctx, double *pVar);
/*
 * initWing (src/analyze.c:97-181, readWing :187-192): the sides of the wall
 * whose BC code (100 * type + id) is wallCode, pressure side first, each
 * list ordered as the reference's insertion orders it. mesh is the one given
 * to ccfd_gpu_set_mesh. nPressure / nSuction (may be NULL) get the list
 * lengths; xyphi (may be NULL, 3 per side) the side midpoint x, y and
 * atan2(y, x) that the reference writes to its CP files.
 */
int ccfd_gpu_wing_init(ccfd_gpu_ctx *ctx, const ccfd_mesh *mesh, int wallCode,
		double refLength, long *nPressure, long *nSuction, double *xyphi);
/*
 * calcCoef (src/analyze.c:292-387) from the wall states of the last operator
 * evaluation: ref = freestream (rho, vx, vy, p) (the reference's
 * refState[0]), alpha in degrees. coef = (CL, CD) rotated by alpha; cp (may
 * be NULL) one pressure coefficient per listed side. Bit for bit the
 * reference's sums (list order, one thread).
 */
int ccfd_gpu_wing_coef(ccfd_gpu_ctx *ctx, const double ref[4], double alpha,
		double coef[2], double *cp);
/*
 * calcErrors (src/analyze.c:445-514): reconstruct the current state at
 * `time` and compare it with exact function exactFunc (1..6,
 * src/exactFunction.c) at the quadrature points. The sums run in a fixed
 * order that differs from the reference's element order (relative
 * difference ~1e-15); Linf is exact. Unlike the reference, dU/dt is left
 * as it was.
 */
int ccfd_gpu_calc_errors(ccfd_gpu_ctx *ctx, int exactFunc, double time,
		double L1[4], double L2[4], double Linf[4]);
/* record points (evalRecordPoints, src/analyze.c:393-402): primitive state
 * of n elements given in reference numbering, 4 doubles each */
int ccfd_gpu_get_prim_at(ccfd_gpu_ctx *ctx, int n, const long *elems,
		double *pVar);

/* ---- time stepping (src/timeDiscretization.c) ---- */
/* calcTimeStep (:249-329): global CFL time step with the output / stop
 * clamp; returns CCFD_ENONFINITE if any element gives a non-finite step */
int ccfd_gpu_calc_dt(ccfd_gpu_ctx *ctx, double t, double pTime,
		double stopTime, double *dt, int *viscousDominates);
/* one explicit Euler / low-storage RK step (:337-401); res (may be NULL)
 * receives globalResidual (src/analyze.c:521-539) */
int ccfd_gpu_explicit_step(ccfd_gpu_ctx *ctx, double t, double dt,
		double res[4]);
/*
 * nSteps explicit steps with the time step computed and kept on the device
 * (calcTimeStep + step, no host round trip per step). *t is advanced; res
 * (may be NULL) gets the last step's residual.
 */
int ccfd_gpu_advance(ccfd_gpu_ctx *ctx, int nSteps, double *t, double pTime,
		double stopTime, double res[4]);
/* globalResidual of the current dU/dt */
int ccfd_gpu_residual(ccfd_gpu_ctx *ctx, double res[4]);

/* ---- implicit backward Euler, JFNK-GMRES with optional LU-SGS ---- */
typedef struct ccfd_gpu_implicit_cfg {
	int nKdim;              /* Krylov dimension (single cycle, as GMRES_M) */
	int nNewtonIter;        /* Newton iteration limit */
	int precond;            /* 1: LU-SGS preconditioner, sweeps in the
	                           reference's element order on buildMatrix's
	                           Jacobian; 2: LU-SGS with multicolour sweeps
	                           (greedy colouring of the face graph, a few
	                           launches per sweep instead of one per
	                           dependency level) on the Jacobian with
	                           respect to the conservative state: a
	                           different, stronger preconditioner;
	                           3: as 2 with a multigrid V-cycle around the
	                           sweeps (coarse levels of agglomerated
	                           elements, Galerkin blocks, Gauss-Seidel
	                           smoothing on every level) */
	double epsNewton;       /* relative Newton tolerance */
	double epsGMRES;        /* relative GMRES tolerance */
	double gammaEW;         /* Eisenstat-Walker gamma */
	int jacobianOrder;      /* 0: the LU-SGS blocks are the Jacobian of the
	                           operator itself (buildMatrix); 1: of the
	                           first-order operator, also for 2nd-order
	                           solves (cheaper to build, diagonally
	                           dominant). Not in the reference. */
	int sweeps;             /* symmetric Gauss-Seidel sweeps, 0 or 1: one
	                           (1..64). precond 2: per preconditioner
	                           application, from zero (one is LU-SGS);
	                           precond 3: before and after the coarse
	                           correction on every level. More than one
	                           pays on a diagonally dominant Jacobian
	                           (jacobianOrder 1); on the truncated
	                           2nd-order one the sweeps can diverge. Not
	                           in the reference. */
	int cycles;             /* GMRES cycles of nKdim iterations, each
	                           restarted from the true residual, until
	                           epsGMRES is met (0 or 1: one, as GMRES_M;
	                           1..100). Not in the reference. */
} ccfd_gpu_implicit_cfg;

/* implicitTimeStep (src/timeDiscretization.c:412-552) with GMRES_M,
 * matrixVector, buildMatrix and LUSGS (src/linearSolver.c:136-510) */
int ccfd_gpu_implicit_step(ccfd_gpu_ctx *ctx, const ccfd_gpu_implicit_cfg *cfg,
		double t, double dt, double res[4], int *nNewton, int *nGMRES);

/*
 * Steady solve by implicit pseudo-time stepping: backward-Euler steps as
 * ccfd_gpu_implicit_step, each element with its own pseudo-time step (the
 * CFL and diffusive limits of calcTimeStep per element), from the resident
 * state until the density residual drops by `drop` relative to the first
 * step, or maxSteps. The Newton system is solved with each row divided by
 * its element's step. The CFL number starts at min(cfl, 1) and follows the
 * linear solver, up to cfl: +20% after a step whose GMRES reduced its
 * residual tenfold, unchanged after one that halved it, -30% after one
 * that did not. Each Newton update is scaled down where needed so that no
 * element's density or pressure changes by more than 20%; after a step
 * that needed this the CFL number is cut by 30% as well. A step stops its
 * Newton loop at nNewtonIter without failing (one Newton iteration per
 * step is the usual choice). res / res0: residual of the last / first
 * step; *steps: steps taken by this call. A solve continues over several
 * calls when res0 comes back holding the first call's values (res0[0] > 0
 * on entry): the drop is measured against it and the CFL number carries
 * on where the previous call left it; res0[0] = 0 starts a new solve.
 * Local pseudo-time steps, the CFL control and the update bound are not
 * part of the reference, which steps with one global time step.
 */
int ccfd_gpu_implicit_steady(ccfd_gpu_ctx *ctx, const ccfd_gpu_implicit_cfg *cfg,
		double cfl, int maxSteps, double drop, double res[4],
		double res0[4], int *steps, int *nNewton, int *nGMRES);

/*
 * GMRES_M (src/linearSolver.c:385-510) on a host right-hand side B (4 doubles
 * per element, reference order) about the resident state: XK is the current
 * conservative state and R_XK its dU/dt at `time`; LU-SGS built at tGlobal
 * when cfg->precond. dX receives the update, *nIter the Krylov iterations.
 */
int ccfd_gpu_gmres(ccfd_gpu_ctx *ctx, const ccfd_gpu_implicit_cfg *cfg,
		double tGlobal, double time, double dt, double alpha,
		const double *B, double normB, double *dX, int *nIter);

/* vectorDotProduct (src/linearSolver.c:115-128) of two host vectors (4
 * doubles per element, reference order), summed on the device in a fixed
 * order (deterministic, not the reference's OpenMP order) */
int ccfd_gpu_dot(ccfd_gpu_ctx *ctx, const double *a, const double *b,
		double *out);
/*
 * The LU-SGS / Gauss-Seidel preconditioner's blocks (buildMatrix + calcDinv,
 * src/linearSolver.c:186-252) at the resident state for cfg's precond
 * (1..3) and jacobianOrder, time and dt, in the reference's element and
 * side order: D (N x 16, the diagonal block I - dt dR/dX row by row) and
 * Blk (nSlots x 16, the block of each side towards its face neighbour,
 * zero at boundary sides). method 0: coloured finite differences; 1:
 * element by element in one launch (first-order conservative Jacobians;
 * otherwise as 0). Both give the same blocks.
 */
int ccfd_gpu_jacobian_blocks(ccfd_gpu_ctx *ctx,
		const ccfd_gpu_implicit_cfg *cfg, double time, double dt,
		int method, double *D, double *Blk);

/* ---- flux functions in isolation ---- */
/*
 * convectiveFlux (src/fluxCalculation.c:803-844) on n rotated state pairs
 * (host arrays, 4 doubles per state: rho, vn, vt, p) on the device; code is
 * the reference's flux id 1..11.
 */
int ccfd_gpu_riemann_batch(int code, int precision, double gam, long n,
		const double *L, const double *R, double *F);
/*
 * The same without allocating device memory (no counterpart in the
 * reference): for callers that cannot allocate, such as a per-face
 * convectiveFlux loop, code under HIP graph capture or a serving loop.
 * L, R and F are staged through a device array the library holds for the
 * life of the process (CCFD_NOALLOC_MAX_STATES states, 1.5 MiB per device);
 * larger n is refused with CCFD_EINVAL. Runs on the current device, whose
 * array every call shares: calls that overlap on one device race.
 */
#define CCFD_NOALLOC_MAX_STATES 16384
int ccfd_gpu_riemann_batch_noalloc(int code, int precision, double gam, long n,
		const double *L, const double *R, double *F);

/* ---- kernel variants ---- */
/*
 * The 2nd-order Euler operator has three implementations with
 * bit-identical results: CCFD_KV_SPLIT (reconstruction kernel writes side
 * states to HBM, flux kernel reads them), CCFD_KV_FUSED (one kernel; each
 * block keeps the side states of its elements and their halo in LDS and
 * evaluates every side's flux) and CCFD_KV_FUSED_FACE (the same, with each
 * face inside a block evaluated once). CCFD_KV_AUTO (default) takes
 * CCFD_KV_FUSED_FACE where a fused kernel applies (2nd order, no viscous
 * flux, LDS tile within the device limit) and the split pair elsewhere.
 * ccfd_gpu_get_variant reports what the next operator call will run.
 */
enum ccfd_kernel_variant {
	CCFD_KV_AUTO = 0,
	CCFD_KV_SPLIT = 1,
	CCFD_KV_FUSED = 2,
	CCFD_KV_FUSED_FACE = 3
};
int ccfd_gpu_set_variant(ccfd_gpu_ctx *ctx, int variant);
int ccfd_gpu_get_variant(ccfd_gpu_ctx *ctx, int *variant);
/*
 * Elements per block of the face-once kernel: 256, 512 (a wider tile, whose
 * halo is a smaller share of the elements it reconstructs, used where its
 * LDS fits the device) or 0 (automatic: 256). Every tile gives the same
 * results.
 * get reports the tile the next operator call runs (256 for the other
 * implementations).
 */
int ccfd_gpu_set_tile(ccfd_gpu_ctx *ctx, int elems);
int ccfd_gpu_get_tile(ccfd_gpu_ctx *ctx, int *elems);
/*
 * Autotuner: time every applicable operator implementation and face-once
 * tile (reps fvTimeDerivative calls each on the resident state, which is
 * not changed) and keep the fastest (a candidate later in the list has to
 * be 2% faster than the one kept). ms[0..3]: time per call of split,
 * fused, face-once 256 and face-once 512 (-1 where one does not apply);
 * variant and tile get the choice.
 */
int ccfd_gpu_autotune(ccfd_gpu_ctx *ctx, int reps, int *variant, int *tile,
		double ms[4]);

/* ---- timing helpers ---- */
/*
 * Run nSteps device-driven steps (as ccfd_gpu_advance from t = 0; with
 * residual != 0 each step also computes globalResidual, as the reference's
 * explicit step does) with HIP events around every kernel on the context's
 * stream. ms[0..3]: average duration of the dt reduction, reconstruction
 * kernel, flux / update kernel and other kernels; ms[4]: wall time per
 * step. launches[0..3] (may be NULL) gets the launch counts; res (may be
 * NULL) the residual of the last step.
 */
int ccfd_gpu_profile_steps(ccfd_gpu_ctx *ctx, int nSteps, double pTime,
		double stopTime, int residual, double *ms, int *launches,
		double res[4]);
/*
 * Streaming copy of `bytes` between two device buffers, repeated `reps`
 * times in each of two launch shapes (grid-stride loop; one 16 KiB tile per
 * block); *gbs = (read + write bytes) / time of the faster shape. Reference
 * for the HBM bandwidth reachable on the device and calibration kernel for
 * traffic counters.
 */
int ccfd_gpu_copy_probe(int device, size_t bytes, int reps, double *gbs);
/* wait for all work queued on the context's stream */
int ccfd_gpu_synchronize(ccfd_gpu_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif
