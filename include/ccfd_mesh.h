/*
 * ccfd_mesh.h - host-side mesh preprocessing for the ccfd finite-volume path.
 *
 * Builds the operator's geometric inputs from raw connectivity (vertices,
 * triangles, quadrilaterals, boundary edges) as flat structure-of-arrays in
 * the reference's element order:
 *   - element numbering: triangles first, then quadrilaterals, in input order
 *     (ccfd src/mesh.c:1495-1647);
 *   - per-element side order: reverse creation order, because sides are
 *     prepended to the element's side list (src/mesh.c:1548-1549);
 *   - face ownership: the side that evaluates a face flux is the earlier-created
 *     side of a pair after the stable (node0, node1, BC) sort
 *     (src/mesh.c:1710-1755); for periodic pairs the side next to the
 *     "...1" boundary keeps the flux (src/mesh.c:172-244);
 *   - geometry: createElemInfo (src/mesh.c:296-341), createSideInfo
 *     (src/mesh.c:250-289), createReconstructionInfo (src/mesh.c:80-114),
 *     with the same floating-point operation order so that every operator
 *     input is bit-identical to the reference's.
 *
 * Plain C99, no device code: the GPU library renumbers and uploads these
 * arrays itself.
 */
#ifndef CCFD_MESH_H
#define CCFD_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary condition types, ccfd src/main.h:44-53 */
enum ccfd_bc_type {
	CCFD_BC_SLIPWALL = 1,
	CCFD_BC_WALL = 2,
	CCFD_BC_INFLOW = 3,
	CCFD_BC_OUTFLOW = 4,
	CCFD_BC_CHARACTERISTIC = 5,
	CCFD_BC_EXACTSOL = 6,
	CCFD_BC_PERIODIC = 7,
	CCFD_BC_PRESSURE_OUT = 8
};

/* One boundary condition, as parsed by initBoundary (src/boundary.c:29-137). */
typedef struct ccfd_bc {
	int type;            /* ccfd_bc_type */
	int id;              /* sub-id: code = 100 * type + id */
	int exactFunc;       /* EXACTSOL only */
	double pVar[4];      /* INFLOW / CHARACTERISTIC state, PRESSURE_OUT p */
	double connection[2];/* PERIODIC translation vector */
} ccfd_bc;

/*
 * Preprocessed mesh. Arrays are owned by the struct (free with
 * ccfd_mesh_free). "slot" = one side of one element; slots of element i are
 * slotStart[i] .. slotStart[i+1]-1 in the reference's per-element order.
 */
typedef struct ccfd_mesh {
	long nNodes;
	double *xy;            /* 2 * nNodes */

	long nElems, nTrias, nQuads;
	int *elemType;         /* 3 or 4 */
	int *domain;           /* physical group / domain id */
	long *elemNode;        /* 4 per element, -1 padded, input node order */
	double *bary;          /* 2 per element */
	double *area, *areaq, *sx, *sy;
	double totalArea_q;    /* 1 / sum(area), accumulated as src/mesh.c:328-332 */

	long nSlots;
	long *slotStart;       /* nElems + 1 */
	long *nbr;             /* >= 0: neighbour element; < 0: ghost -(g + 1) */
	long *partner;         /* slot index of the connected side; -1 for ghosts */
	unsigned char *owner;  /* 1 if this slot evaluates the face flux */
	double *n;             /* 2 per slot: outward normal of this side */
	double *len;           /* side length */
	double *fn;            /* 2 per slot: normal used for the face flux
	                          (the owner side's normal) */
	double *flen;          /* length used for the face flux (owner's) */
	double *GP;            /* 2 per slot: side midpoint - element barycentre */
	double *GPc;           /* 2 per slot: connection side's GP */
	double *w;             /* 2 per slot: least-squares gradient weights */
	double *bbv;           /* 2 per slot: baryBaryVec */
	double *bbd;           /* baryBaryDist */
	long *slotNode;        /* 2 per slot: side nodes (creation orientation) */

	long nFaces;           /* sides in the reference's side[] list */
	long *faceSlot;        /* owner slot of each face */

	long nGhosts;          /* non-periodic boundary sides (nBCsides) */
	long *ghostSlot;       /* interior slot next to the ghost */
	int *ghostBC;          /* index into bc[] */
	double *ghostBary;     /* 2 per ghost: mirrored barycentre */
	double *ghostGP;       /* 2 per ghost: ghost side GP (midpoint - ghost bary) */

	int nBC;
	ccfd_bc *bc;

	double xMin, xMax, yMin, yMax;
} ccfd_mesh;

/*
 * Raw connectivity as a generator or a mesh file gives it (the arguments of
 * ccfd_mesh_build in one struct; the arrays belong to the caller). The
 * batched arena can build a request's device image from it on the GPU
 * (ccfd_batch_add_raw) instead of from a host-preprocessed ccfd_mesh.
 */
typedef struct ccfd_raw_mesh {
	long nVert;
	const double *xy;      /* 2 * nVert */
	long nTri;
	const long *tri;       /* 4 * nTri: 3 node ids + domain */
	long nQuad;
	const long *quad;      /* 5 * nQuad: 4 node ids + domain */
	long nBCedge;
	const long *bcEdge;    /* 3 * nBCedge: 2 node ids + code 100*type+id */
	int nBC;
	const ccfd_bc *bc;
} ccfd_raw_mesh;

/*
 * Build a mesh from raw connectivity.
 *   xy      : 2 * nVert vertex coordinates
 *   tri     : 4 * nTri  (3 zero-based node ids + domain)
 *   quad    : 5 * nQuad (4 zero-based node ids + domain)
 *   bcEdge  : 3 * nBCedge (2 zero-based node ids + BC code 100*type+id)
 *   bc      : nBC boundary conditions (codes looked up by type and id)
 * Returns NULL and fills err (if non-NULL, >= 256 bytes) on failure.
 */
ccfd_mesh *ccfd_mesh_build(long nVert, const double *xy,
		long nTri, const long *tri, long nQuad, const long *quad,
		long nBCedge, const long *bcEdge,
		int nBC, const ccfd_bc *bc, char *err);

/*
 * Cartesian generator, createCartMesh (src/mesh.c:385-486): iMax x jMax
 * quads over [x0, x1] x [y0, y1]; vertex coordinates accumulated by repeated
 * addition exactly as the reference does. bcType[side][k], bcRange[side][k][2]
 * (one-based, inclusive) for sides bottom, right, top, left; nSeg[side]
 * segments per side. Output arrays are malloc'ed and returned through the
 * pointers (free with free()).
 */
int ccfd_cart_connectivity(long iMax, long jMax, double x0, double y0,
		double x1, double y1, const int nSeg[4], const int bcType[4][20],
		const int bcRange[4][20][2],
		long *nVert, double **xy, long *nQuad, long **quad,
		long *nBCedge, long **bcEdge);

/* gmsh ASCII reader (readGmsh, src/mesh.c:498-975), formats 2 and 4.
 * Format 2: lines with physical group > 100 become boundary edges with that
 * code. Format 4: the line-element blocks take, in block order, the first
 * physical tag of the curve entities that have one; elements get domain 1. */
int ccfd_gmsh_read(const char *path, long *nVert, double **xy,
		long *nTri, long **tri, long *nQuad, long **quad,
		long *nBCedge, long **bcEdge, char *err);

/* EMC2 / medit .mesh reader (readEMC2, src/mesh.c:1004-1141): edges with a
 * nonzero reference become boundary edges with that code; element
 * references become domains. Same outputs as ccfd_gmsh_read. */
int ccfd_emc2_read(const char *path, long *nVert, double **xy,
		long *nTri, long **tri, long *nQuad, long **quad,
		long *nBCedge, long **bcEdge, char *err);

/* gmsh v2 ASCII writer, readable by the reference's readGmsh. */
int ccfd_gmsh_write(const char *path, long nVert, const double *xy,
		long nTri, const long *tri, long nQuad, const long *quad,
		long nBCedge, const long *bcEdge);

void ccfd_mesh_free(ccfd_mesh *m);

/*
 * The element a record point lies in, as initRecordPoints finds it
 * (src/analyze.c:33-87): the first element, in element order, for which the
 * point is on the inner side of every edge. The reference normalises each
 * edge normal with the x component twice, so a horizontal edge (0 / 0)
 * never rejects a point; this keeps that behaviour. Returns -1 if no
 * element accepts the point.
 */
long ccfd_mesh_locate(const ccfd_mesh *m, const double x[2]);

/* Venkatakrishnan epsilon^2 = (K sqrt(A))^3 per element, initFV
 * (src/finiteVolume.c:47-52). */
void ccfd_mesh_venk_eps(const ccfd_mesh *m, double venk_k, double *eps_sq);

#ifdef __cplusplus
}
#endif

#endif
