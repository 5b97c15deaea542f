/*
 * ccfd_serve.h - served steady solves: a request queue with continuous
 * batching over ccfd_batch arenas, behind a C ABI.
 *
 * The caller prepares requests (mesh, initial primitive state, CFL number)
 * and submits them in arrival order. Each round admits queued requests in
 * order while the main arena has room (FIFO: a request that does not fit
 * holds back those behind it), runs one burst of steps on every busy
 * arena, polls the residuals and retires solves that reached the target
 * density residual drop (relative to their first step), the step cap, or a
 * non-finite time step. Requests submitted while others run join at the
 * next round: admission does not wait for the batch to drain.
 *
 * Precision promotion (optional polish arena, e.g. an FP64 arena beside an
 * FP32-flux main arena): a solve whose drop reaches switchDrop, or that
 * stalls within 10x of it, continues from its conservative state
 * (converted with consPrim, src/equationOfState.c:37-52) in the polish
 * arena until the target.
 *
 * This has no counterpart in the reference, which solves one case per
 * process. The arenas are reached through a table of functions
 * (ccfd_serve_backend): ccfd_serve_backend_batch fills one for a
 * ccfd_batch; tests fill one with a stand-in.
 */
#ifndef CCFD_SERVE_H
#define CCFD_SERVE_H

#include "ccfd_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

/* one arena as the server uses it; every function returns CCFD_OK or a
 * negative status. mesh is the pointer given to ccfd_serve_submit. */
typedef struct ccfd_serve_backend {
	void *arena;
	int (*fits)(void *arena, const void *mesh, int *fits);
	int (*add)(void *arena, const void *mesh, const double *prim, double cfl,
			int *id);
	int (*remove)(void *arena, int id);
	int (*step)(void *arena, int nSteps);
	int (*poll)(void *arena, int maxIds, int *ids, ccfd_batch_info *info,
			int *n);
	int (*get_cons)(void *arena, int id, double *cons);
	int (*get_prim)(void *arena, int id, double *prim);
} ccfd_serve_backend;

/* the backend of a ccfd_batch (meshes are const ccfd_mesh *; steps run
 * without an output or end time) */
int ccfd_serve_backend_batch(ccfd_batch *b, ccfd_serve_backend *out);
/* the same with meshes given as const ccfd_raw_mesh *, each preprocessed on
 * the device at admission (ccfd_batch_add_raw) */
int ccfd_serve_backend_batch_raw(ccfd_batch *b, ccfd_serve_backend *out);

typedef struct ccfd_serve_cfg {
	double targetDrop;      /* density residual drop of a solve (1e-6) */
	long maxSteps;          /* step cap per solve, both arenas together */
	int stepsPerRound;      /* steps per burst between polls */
	int maxActive;          /* solves in flight at most (0: no limit; 1: one
	                           at a time) */
	double switchDrop;      /* drop at which a solve moves to the polish
	                           arena (used only with one) */
	int stallRounds;        /* rounds of no progress (best drop not halved)
	                           that count as a stall near switchDrop */
	double gamma;           /* ratio of specific heats for consPrim */
	int keepState;          /* keep each retired solve's primitive state */
	double (*clock)(void *arg);  /* seconds; NULL: CLOCK_MONOTONIC */
	void *clockArg;
} ccfd_serve_cfg;

typedef struct ccfd_serve_result {
	long id;                /* the request id given at submission */
	long nElems;
	long steps;             /* steps taken (main + polish arena) */
	long polishSteps;       /* of which in the polish arena */
	double drop;            /* final density residual / first */
	int converged;          /* drop <= targetDrop */
	int nonfinite;          /* retired on a non-finite time step */
	double admitted, done;  /* clock at admission and at retirement */
} ccfd_serve_result;

typedef struct ccfd_server ccfd_server;

/* polish may be NULL (no promotion) */
int ccfd_serve_create(ccfd_server **s, const ccfd_serve_backend *main,
		const ccfd_serve_backend *polish, const ccfd_serve_cfg *cfg);
int ccfd_serve_destroy(ccfd_server *s);

/*
 * Queue a prepared request. prim (4 doubles per element, reference order)
 * is read until the request is admitted; the mesh must stay valid until
 * the request's result has been returned (the polish arena needs it).
 */
int ccfd_serve_submit(ccfd_server *s, long id, const void *mesh, long nElems,
		const double *prim, double cfl);

/*
 * One round: hand promoted solves to the polish arena, admit queued
 * requests, step the busy arenas, poll and retire. active / polishing /
 * queued (may be NULL) get the solves in each state after the round.
 * A request larger than an empty arena is an error (CCFD_ENOMEM); the
 * server keeps it at the head of the queue.
 */
int ccfd_serve_round(ccfd_server *s, int *active, int *polishing,
		int *queued);

/* retired requests not yet returned, in retirement order (up to max) */
int ccfd_serve_results(ccfd_server *s, int max, ccfd_serve_result *out,
		int *n);

/* the kept primitive state of a retired request (keepState), which is then
 * released; CCFD_EINVAL if there is none */
int ccfd_serve_take_state(ccfd_server *s, long id, double *prim);

#ifdef __cplusplus
}
#endif

#endif
