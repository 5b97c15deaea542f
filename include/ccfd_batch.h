/*
 * ccfd_batch.h - many independent explicit Euler solves on one GPU.
 *
 * A batch owns an arena in device memory, sized from an HBM budget at
 * creation. Each admitted instance (mesh + state + time-step settings)
 * occupies whole 256-element blocks of it. Every stage of every active
 * instance runs in one kernel launch: the launch walks a block table, so
 * instances of any size share the launch without padding to the largest.
 * Each instance keeps its own time, time step and residual. Steps are
 * captured into a HIP graph that is replayed until the set of instances
 * changes. The operator is the same code as the single-mesh path
 * (ccfd_gpu.h), so an instance's results do not depend on what else is in
 * the batch: they are bit-identical alone and in any mix.
 *
 * This has no counterpart in the reference, which solves one case per
 * process. Its time stepping follows calcTimeStep / explicitTimeStepRK
 * (src/timeDiscretization.c:249-401) per instance. The local time step of
 * steady solves (each element at its own CFL limit) is new: the reference
 * computes a local step (src/finiteVolume.c:81-85) but never uses it.
 */
#ifndef CCFD_BATCH_H
#define CCFD_BATCH_H

#include <stddef.h>

#include "ccfd_gpu.h"
#include "ccfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ccfd_batch ccfd_batch;

typedef struct ccfd_batch_cfg {
	/* scheme shared by all instances: gam, iFlux, spatialOrder, limiter,
	 * venk_k, nRKstages, timeOrder, precision, timeStep1D (navierstokes
	 * must be 0; cfl / dfl are per instance) */
	ccfd_gpu_params scheme;
	int maxSides;           /* 3 (triangles) or 4 (quads or mixed) */
	int localTimeStep;      /* 1: every element advances with its own CFL
	                           step (steady solves) */
	size_t hbmBytes;        /* device memory for the arena; 0: half of the
	                           free memory */
	int maxInstances;       /* instance slots; 0: 4096 */
	int maxHalo;            /* fused-operator halo columns per block (LDS);
	                           0: the most the device's LDS holds (at most
	                           512). While an instance with a larger block
	                           halo is admitted, the batch runs the split
	                           operator (same results). */
} ccfd_batch_cfg;

typedef struct ccfd_batch_case {
	double cfl;             /* CFL number */
	double pTime;           /* output time: global steps are clamped to it */
	double stopTime;        /* end time (src/timeDiscretization.c:316-321) */
} ccfd_batch_case;

typedef struct ccfd_batch_info {
	long nElems;
	long steps;             /* steps taken since admission */
	double t, dt;           /* time and last time step (global stepping) */
	double res[4];          /* globalResidual of the last step */
	double res0[4];         /* ... of the first step */
	int nonfinite;          /* a non-finite time step was met */
} ccfd_batch_info;

int ccfd_batch_create(ccfd_batch **b, int device, const ccfd_batch_cfg *cfg);
int ccfd_batch_destroy(ccfd_batch *b);

/* device bytes an instance of this mesh takes in the arena */
size_t ccfd_batch_bytes(const ccfd_batch *b, const ccfd_mesh *m);
/* does it fit now (contiguous room in every pool)? */
int ccfd_batch_fits(ccfd_batch *b, const ccfd_mesh *m, int *fits);

/*
 * Admit an instance: mesh, per-instance settings and the initial primitive
 * state (4 doubles per element, reference order). Returns CCFD_ENOMEM when
 * the arena has no room now (the caller may wait for a removal).
 */
int ccfd_batch_add(ccfd_batch *b, const ccfd_mesh *m,
		const ccfd_batch_case *cs, const double *prim, int *id);
/*
 * Admit an instance from raw connectivity, preprocessed on the device:
 * side pairing (a stable radix sort of the side keys in the side-list
 * order of src/mesh.c:1532-1566, 1713-1755), element and side geometry
 * (createElemInfo / createSideInfo, src/mesh.c:250-341), ghost states'
 * geometry, periodic hookup (connectPeriodicBC, src/mesh.c:172-244),
 * least-squares weights (createReconstructionInfo, src/mesh.c:80-114), the
 * space-filling-curve order and the block halo lists. Every array equals,
 * bit for bit, the one ccfd_batch_add writes for ccfd_mesh_build(m), and
 * the instance behaves identically. The host only checks the input and
 * resolves boundary codes.
 */
int ccfd_batch_add_raw(ccfd_batch *b, const ccfd_raw_mesh *m,
		const ccfd_batch_case *cs, const double *prim, int *id);
/* ccfd_batch_fits for a raw mesh */
int ccfd_batch_fits_raw(ccfd_batch *b, const ccfd_raw_mesh *m, int *fits);

/* retire an instance and free its arena space */
int ccfd_batch_remove(ccfd_batch *b, int id);

/* nSteps explicit steps of every active instance */
int ccfd_batch_step(ccfd_batch *b, int nSteps);
/* wait for the batch's stream */
int ccfd_batch_sync(ccfd_batch *b);

int ccfd_batch_info_get(ccfd_batch *b, int id, ccfd_batch_info *info);
/* active instance ids (up to maxIds) with their info, one device read */
int ccfd_batch_poll(ccfd_batch *b, int maxIds, int *ids,
		ccfd_batch_info *info, int *n);
int ccfd_batch_get_prim(ccfd_batch *b, int id, double *prim);
int ccfd_batch_get_cons(ccfd_batch *b, int id, double *cons);

/*
 * Operator implementation (enum ccfd_kernel_variant) and whether steps are
 * replayed from a captured HIP graph (default 1).
 */
int ccfd_batch_set_variant(ccfd_batch *b, int variant);
int ccfd_batch_get_variant(ccfd_batch *b, int *variant);
int ccfd_batch_set_graph(ccfd_batch *b, int useGraph);

/*
 * Autotuner: time nSteps steps of the current batch with every applicable
 * operator implementation (split, fused, fused face-once) on a copy of the
 * state, which is restored afterwards, keep the fastest and return it;
 * msPerStep[variant] receives the measured times (0 where not applicable;
 * array of 4, may be NULL).
 */
int ccfd_batch_autotune(ccfd_batch *b, int nSteps, int *variant,
		double *msPerStep);

/* elements / bytes in use and capacity of the arena */
int ccfd_batch_usage(ccfd_batch *b, long *elemsUsed, long *elemsCap,
		int *instances);

#ifdef __cplusplus
}
#endif

#endif
