/*
 * ccfd_oracle.h - CPU restatement of ccfd's finite-volume hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared
 * against; nothing in the product (ccfd_amd/, include/) links or calls it.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * Arrays are in the reference's element / slot order (ccfd_mesh), state as
 * 4 doubles per element (rho, vx, vy, p) or (rho, mx, my, E).
 *
 * How each part is pinned (tests/):
 *   fluxes, fvTimeDerivative   bit for bit against the reference's compiled
 *                              sources: fixtures and, where the reference is
 *                              present, live (test_oracle_golden.py)
 *   calcTimeStep, RK, timeDisc the reference's make check outputs; the
 *                              time step's closed forms on a uniform grid;
 *                              every RK table against the reference's source
 *                              text (test_make_check_oracle.py)
 *   implicit step              backward-Euler solution, explicit-Euler limit,
 *                              fixed point (test_oracle_implicit.py, as below)
 *   buildMatrix + calcDinv     forward differences of R; inverses
 *   LUSGS                      the (D + L) D^-1 (D + U) product
 *   GMRES_M                    dense solve of the finite-difference system
 *   vectorDotProduct           correctly rounded sums
 *   globalResidual             area-weighted RMS (test_make_check_oracle.py)
 *   exact Riemann solution     Toro's published Sod solution, isentropic
 *                              fan (test_oracle_golden.py)
 *   source term (calcSource)   bit for bit against the reference's compiled
 *                              source.c + fvTimeDerivative (harness) and the
 *                              integral of a constant (test_oracle_analyze.py)
 *   calcErrors                 closed forms: a uniform offset, quadrature
 *                              points and weights rebuilt from the nodes
 *   calcCoef / initWing        closed form: a linear pressure field on a
 *                              closed polygon; list order and partition
 *   initRecordPoints           point-in-polygon by brute force, and the
 *                              reference's horizontal-edge behaviour
 */
#ifndef CCFD_ORACLE_H
#define CCFD_ORACLE_H

#include "ccfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_cfg {
	double gam, mu, Pr;        /* gamma, viscosity, Prandtl */
	int navierstokes;          /* 1: -Dnavierstokes semantics */
	int iFlux;                 /* 1..11, src/main.h:76-88 */
	int spatialOrder;          /* 1 or 2 */
	int limiter;               /* 1 BJ, 2 VK */
	double venk_k;
	/* exact functions (src/exactFunction.c) */
	double refState[2][4];
	double rp1Dinterface;
	double dxRef;
	/* time stepping (src/timeDiscretization.c) */
	double cfl, dfl;
	int timeStep1D;
	int nRKstages, timeOrder;
	/* source term (src/source.c, src/equation.c:119-123): 0 none */
	int sourceFunc;
} oracle_cfg;

typedef struct oracle_ctx oracle_ctx;

oracle_ctx *oracle_create(const ccfd_mesh *m, const oracle_cfg *cfg);
void oracle_destroy(oracle_ctx *o);

/* element state */
void oracle_set_prim(oracle_ctx *o, const double *prim);   /* also sets cons */
void oracle_get_prim(const oracle_ctx *o, double *prim);
void oracle_get_cons(const oracle_ctx *o, double *cons);

/*
 * fvTimeDerivative (src/finiteVolume.c:78-121) on the current primitive
 * state. Outputs (any may be NULL): u_t (4/elem), u_x, u_y (4/elem, limited
 * gradients), sideP (4/slot, side states), flux (4/slot, side fluxes).
 */
void oracle_fvtd(oracle_ctx *o, double time, double *u_t, double *u_x,
		double *u_y, double *sideP, double *flux);

/* single flux function on one rotated state pair, src/fluxCalculation.c */
void oracle_riemann(const oracle_cfg *cfg, int iFlux, const double L[4],
		const double R[4], double F[4]);

/* calcTimeStep (src/timeDiscretization.c:249-329); returns dt */
double oracle_calc_dt(oracle_ctx *o, double t, double pTime, double stopTime,
		int *viscousDominates, int *nonfinite);

/* one explicit step: Euler or low-storage RK (src/timeDiscretization.c:
 * 337-401); resIter gets the global residual (src/analyze.c:521-539) */
void oracle_explicit_step(oracle_ctx *o, double t, double dt, double res[4]);

/* area-weighted L2 residual of the last fvTimeDerivative */
void oracle_global_residual(const oracle_ctx *o, const double *u_t,
		double res[4]);

/* dot product of two 4-per-element vectors (src/linearSolver.c:115-128) */
double oracle_dot(long nElems, const double *a, const double *b);

/*
 * Implicit backward-Euler step with JFNK-GMRES and optional LU-SGS
 * (src/timeDiscretization.c:412-552, src/linearSolver.c:136-510).
 * Returns 0, or -1 if Newton did not converge, -2 for a singular D block.
 */
typedef struct oracle_implicit_cfg {
	int nKdim, nNewtonIter, precond;
	double epsNewton, epsGMRES, gammaEW;
} oracle_implicit_cfg;
int oracle_implicit_step(oracle_ctx *o, const oracle_implicit_cfg *ic,
		double t, double dt, double res[4], int *nNewton, int *nGMRES);

/*
 * The Jacobian blocks and the preconditioner on their own (buildMatrix +
 * calcDinv, src/linearSolver.c:136-252; LUSGS, :261-333). Dg and Dinv hold
 * 16 doubles per element, Blk 16 per slot: the block of I - dt dR/dpVar in
 * the row of the slot's element and the column of its neighbour, row-major
 * (row: component of dU/dt, column: primitive variable). oracle_build_matrix
 * perturbs and restores the primitive state as the reference does (so it
 * may move by an ulp) and returns -2 for a singular diagonal block;
 * oracle_lusgs returns X = (D + U)^-1 D (D + L)^-1 B, lower and upper by
 * element index. Both return -3 when out of memory.
 */
int oracle_build_matrix(oracle_ctx *o, double time, double dt, double *Dg,
		double *Blk, double *Dinv);
int oracle_lusgs(const oracle_ctx *o, const double *Dinv, const double *Blk,
		const double *B, double *X);

/* one GMRES_M solve (src/linearSolver.c:385-510) of (I - dt dR/dU) dX = -B
 * at the current state (the operator at t + dt, buildMatrix at t, as a
 * Newton iteration of implicitTimeStep); the state is left as it was.
 * nIter gets the Krylov iterations; returns 0, -2 (singular block) or -3. */
int oracle_gmres(oracle_ctx *o, const oracle_implicit_cfg *ic, double t,
		double dt, const double *B, double *dX, int *nIter);

/* RK coefficient table (src/timeDiscretization.c:94-178); returns -1 for an
 * unsupported (stages, order) pair */
int oracle_rk_coeffs(int nRKstages, int timeOrder, double coeff[6]);

/* exact functions (src/exactFunction.c:35-183) and exact Riemann solver
 * (src/exactRiemann.c:59-214) */
void oracle_exact_func(const oracle_cfg *cfg, const ccfd_mesh *m, int f,
		const double x[2], double time, double pVar[4]);

/* ---- analysis consumers (src/analyze.c) ---- */

/* quadrature points and weights of element e (src/mesh.c:117-165): returns
 * nGP (3 for triangles, 5 for quadrilaterals) */
int oracle_quadrature(const ccfd_mesh *m, long e, double xGP[5][2],
		double wGP[5]);

/* calcErrors (src/analyze.c:445-514): spatialReconstruction(time) on the
 * current state, then the L1, L2 and Linf norms of (exact function f at the
 * quadrature points) - (reconstructed state there), summed in element and
 * point order */
void oracle_calc_errors(oracle_ctx *o, int f, double time, double L1[4],
		double L2[4], double Linf[4]);

/* initWing (src/analyze.c:97-181): the interior slots of the sides of the
 * wall whose BC code (100 * type + id) is wallCode, as the pressure-side
 * list (nPressure entries) followed by the suction-side list, each sorted by
 * x of its midpoint the way the reference inserts them. slots holds
 * nGhosts entries. Returns the number listed. */
long oracle_wing_lists(const oracle_ctx *o, int wallCode, long *slots,
		long *nPressure, long *nSuction);

/* calcCoef (src/analyze.c:292-387) from the side states of the last
 * oracle_fvtd: ref = freestream (rho, vx, vy, p), alpha in degrees. cl, cd
 * get the rotated coefficients; cp (may be NULL) one value per listed side,
 * xyphi (may be NULL) its x, y and atan2(y, x). */
void oracle_wing_coef(const oracle_ctx *o, const long *slots, long nPressure,
		long nSuction, double refLength, const double ref[4], double alpha,
		double *cl, double *cd, double *cp, double *xyphi);

/* initRecordPoints (src/analyze.c:33-87): the first element (in element
 * order) the reference's test accepts for point x; -1 if none */
long oracle_locate(const ccfd_mesh *m, const double x[2]);

#ifdef __cplusplus
}
#endif

#endif
