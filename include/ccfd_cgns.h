/*
 * ccfd_cgns.h - CGNS output and restart files of the drop-in
 * (libccfd_cgns.so, loaded by the drop-in only when a case asks for them).
 *
 * The reference writes and reads these files through the CGNS library
 * (src/output.c:122-244 cgnsOutput, 407-517 cgnsFinalizeOutput, 532-668
 * cgnsWriteMesh; src/initialCondition.c:94-158 cgnsReadSolution). That
 * library is not part of this build; the files are written directly as
 * HDF5 files in the CGNS/HDF5 node layout: every node an HDF5 group with
 * the attributes name, label, type and flags, its data in a dataset
 * " data" (dimensions in reverse order), links as groups of type "LK"
 * with " file", " path" and an external link " link". The node tree is the
 * one the reference's calls produce with a 32-bit CGNS build.
 *
 * Every function returns 0, or -1 with a message in err (256 bytes).
 */
#ifndef CCFD_CGNS_H
#define CCFD_CGNS_H

#include "ccfd_mesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * The grid file (cgnsWriteMesh): coordinates, the Triangles and
 * Quadrilaterals sections and, unless a boundary is periodic, the
 * Boundaries section (BAR_2 in the reference's boundary-side order, node
 * order of the mesh file's boundary edges bcEdge[3 * nBCedge]) with one
 * BC_t node per boundary condition.
 */
int ccfd_cgns_write_mesh(const char *path, const ccfd_mesh *m,
		long nBCedge, const long *bcEdge, char *err);

/*
 * One flow solution (cgnsOutput): the zone linked to gridFile, Density,
 * VelocityX, VelocityY, VelocityZ (zero) and Pressure at cell centres from
 * prim (4 doubles per element: rho, u, v, p), and the ConvergenceInfo
 * descriptor "%20.12f %20.12f" of time and timeOverall.
 */
int ccfd_cgns_write_solution(const char *path, const char *gridFile,
		const ccfd_mesh *m, const double *prim, double time,
		double timeOverall, char *err);

/*
 * The master file (cgnsFinalizeOutput): the zone linked to gridFile, one
 * link FlowSolution%09ld per output to solutionFiles[k]:/Base/Zone/
 * FlowSolution, TimeIterValues with TimeValues and IterationValues, and
 * ZoneIterativeData with FlowSolutionPointers. IterationValues holds the
 * first n 32-bit words of the 64-bit iteration numbers, as the reference's
 * cg_array_write(Integer) of a long array stores them.
 */
int ccfd_cgns_write_master(const char *path, const char *gridFile,
		const ccfd_mesh *m, int n, const double *times, const long *iters,
		const char *const *solutionFiles, char *err);

/*
 * Read a flow solution for a restart (cgnsReadSolution): the zone's
 * element count must be nElems; prim gets rho, u, v, p per element from
 * the zone's first FlowSolution_t node, time and timeOverall the two
 * numbers of the base's first Descriptor_t node.
 */
int ccfd_cgns_read_solution(const char *path, long nElems, double *prim,
		double *time, double *timeOverall, char *err);

/*
 * Read an unstructured 2-D mesh (readCGNS, src/mesh.c:1153-1394), with the
 * outputs of ccfd_gmsh_read (ccfd_mesh.h; arrays malloc'ed, the caller
 * frees them). The first base and, of its zones, the first by name (the
 * CGNS library sorts zones); CoordinateX / CoordinateY of the zone's
 * "GridCoordinates"; sections in file order: the last BAR_2 section gives
 * the boundary edges, the last TRI_3 and QUAD_4 sections the cells
 * (domain 1). The ZoneBC_t's BC_t nodes, in file order, hand their names,
 * read as integer codes, to consecutive boundary edges, as many to each as
 * its point set counts (a list its length, a range 2), whatever the set
 * holds; edges left over keep code 0. Refused with a message: MIXED
 * sections (the CGNS 4.1 library the reference vendors reads only
 * fixed-size sections), other element types, a BC name that is not a
 * number, more BC points than boundary edges, node ids out of range.
 */
int ccfd_cgns_read_mesh(const char *path, long *nVert, double **xy,
		long *nTri, long **tri, long *nQuad, long **quad,
		long *nBCedge, long **bcEdge, char *err);

#ifdef __cplusplus
}
#endif

#endif
