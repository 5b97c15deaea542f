/*
 * ccfd_dropin.h - the reference's host entry points on the MI355X path.
 *
 * libccfd_host.so exports the functions ccfd's C code calls on this path
 * with the reference's names and signatures; they act on one global case
 * (as the reference's module globals do) whose mesh and state live on the
 * GPU behind ccfd_gpu.h. Failures print a message and exit(1), the
 * reference's behaviour; the ccfd_* helpers return status codes instead.
 *
 *   fvTimeDerivative   src/finiteVolume.h:14
 *   convectiveFlux     src/fluxCalculation.c:803 (selector, iFlux global)
 *   initTimeDisc       src/timeDiscretization.h:56
 *   timeDisc           src/timeDiscretization.h:57
 *   initLinearSolver   src/linearSolver.h:39
 *   vectorDotProduct   src/linearSolver.h:40
 *   GMRES_M            src/linearSolver.h:41-42
 *   freeLinearSolver   src/linearSolver.h:43
 *   initAnalyze        src/analyze.h:50 (the analysis keys are read by
 *                      ccfd_init; this only checks the case is set up)
 *   analyze            src/analyze.h:51
 *   calcErrors         src/analyze.h:52
 *   globalResidual     src/analyze.h:53
 *   freeAnalyze        src/analyze.h:54
 *   main()             src/main.c:35-130 -> ccfd_main
 */
#ifndef CCFD_DROPIN_H
#define CCFD_DROPIN_H

#include <stddef.h>

#include "ccfd_gpu.h"
#include "ccfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- parameter files (readInTools semantics, src/readInTools.c) ---- */
typedef struct ccfd_ini ccfd_ini;
ccfd_ini *ccfd_ini_read(const char *path, char *err);
void ccfd_ini_free(ccfd_ini *ini);
/* each reader consumes the first unread occurrence of key; def (may be
 * NULL: mandatory) is used when absent; returns 0 or -1 if missing */
int ccfd_ini_str(ccfd_ini *ini, const char *key, const char *def, char *out,
		size_t n);
int ccfd_ini_dbl(ccfd_ini *ini, const char *key, const char *def, double *out);
int ccfd_ini_long(ccfd_ini *ini, const char *key, const char *def, long *out);
int ccfd_ini_bool(ccfd_ini *ini, const char *key, const char *def, int *out);
int ccfd_ini_dbl_array(ccfd_ini *ini, const char *key, int n, const char *def,
		double *out);
int ccfd_ini_count(ccfd_ini *ini, const char *key);

/* ---- case set-up (src/main.c:53-121) ---- */
/* read the ini, build mesh, boundary conditions, initial state, parameters
 * and the device context; returns 0 or a negative status (message in err,
 * >= 256 bytes, may be NULL) */
int ccfd_init(const char *iniFile, int device, char *err);
/* ccfd_init for a restart from a CGNS flow solution (src/main.c:64-94,
 * src/initialCondition.c:94-158): the state, time and overall time come
 * from the file, the first iteration (stationary) or the print schedule
 * (transient) from the time stamp after the file name's last '_'; no grid
 * file is written and the analysis file is appended to. restartFile NULL
 * is ccfd_init. CGNS files need libccfd_cgns.so (include/ccfd_cgns.h). */
int ccfd_init_restart(const char *iniFile, const char *restartFile,
		int device, char *err);
/* full run: ccfd_init + timeDisc + ccfd_finalize; returns 0 on success */
int ccfd_main(const char *iniFile, int device);
/* the same from a restart file (`ccfd case.ini restart.cgns`) */
int ccfd_main_restart(const char *iniFile, const char *restartFile, int device);
void ccfd_finalize(void);
/* the global case */
ccfd_gpu_ctx *ccfd_case_context(void);
const ccfd_mesh *ccfd_case_mesh(void);
/* current simulation time and last iteration of timeDisc */
double ccfd_case_time(void);
long ccfd_case_iterations(void);
/* write the CSV output of the current state (csvOutput, src/output.c:56-113)
 * to path */
int ccfd_write_csv(const char *path);

/* ---- reference entry points ---- */
void fvTimeDerivative(double time);
void convectiveFlux(double rhoL, double rhoR, double vxL, double vxR,
		double vyL, double vyR, double pL, double pR, double fluxLoc[4]);
void initTimeDisc(void);
void timeDisc(void);
void initLinearSolver(void);
double vectorDotProduct(double **A, double **B);
void GMRES_M(double time, double dt, double alpha, double **B, double normB,
		double *abortCrit, double **deltaX);
void freeLinearSolver(void);
void initAnalyze(void);
/* record points, CL / CD / CP (calcWing) and the residual file after a step
 * to `time`; resIter[4], resIter[5] receive the CL / CD residuals */
void analyze(double time, long iter, double resIter[6]);
void calcErrors(double time);
void globalResidual(double resIter[6]);
void freeAnalyze(void);

#ifdef __cplusplus
}
#endif

#endif
