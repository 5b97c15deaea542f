/*
 * cgns_hdf5.c - CGNS output and restart files as HDF5 files in the
 * CGNS/HDF5 node layout (include/ccfd_cgns.h). Built into libccfd_cgns.so
 * against the image's HDF5; the drop-in loads it only when a case writes
 * or reads CGNS.
 *
 * Node layout (the published CGNS file mapping for HDF5):
 *  - a node is a group with the attributes "name" and "label" (fixed-length
 *    strings of 33 bytes), "type" (3 bytes: MT, I4, I8, R4, R8, C1, LK) and
 *    "flags" (one int, 1); groups track and index link creation order;
 *  - its data is the dataset " data" of the matching native type, with the
 *    node's dimensions in reverse order (the mapping stores Fortran order);
 *  - a link is a node of type LK with " path" and " file" (the target as
 *    NUL-terminated characters) and an HDF5 external link " link";
 *  - the root group carries name "HDF5 MotherNode", label "Root Node of
 *    HDF5 File", type MT, and the datasets " format" and " hdf5version".
 */
#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_cgns.h"

#define NAMELEN 33          /* 32 characters + NUL */
#define CGNS_FILE_VERSION 4.10f   /* CGNSLibraryVersion of a 4.1 library */
/* element types of the CGNS standard */
#define CG_BAR_2 3
#define CG_TRI_3 5
#define CG_QUAD_4 7
#define CG_MIXED 20

static void seterr(char *err, const char *fmt, ...)
{
	if (!err)
		return;
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(err, 256, fmt, ap);
	va_end(ap);
}

/* errors are reported through err, not by HDF5's own stack printout */
static void quiet(void)
{
	H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
}

static hid_t gcpl(void)
{
	hid_t p = H5Pcreate(H5P_GROUP_CREATE);
	H5Pset_link_creation_order(p, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
	return p;
}

/* a scalar fixed-length string attribute of len + 1 bytes */
static int str_att(hid_t g, const char *name, const char *value, int len)
{
	char buf[NAMELEN + 1];
	memset(buf, 0, sizeof(buf));
	strncpy(buf, value, (size_t)len);
	hid_t s = H5Screate(H5S_SCALAR);
	hid_t t = H5Tcopy(H5T_C_S1);
	H5Tset_size(t, (size_t)len + 1);
	hid_t a = H5Acreate2(g, name, t, s, H5P_DEFAULT, H5P_DEFAULT);
	int rc = (a < 0 || H5Awrite(a, t, buf) < 0) ? -1 : 0;
	if (a >= 0)
		H5Aclose(a);
	H5Tclose(t);
	H5Sclose(s);
	return rc;
}

static int int_att(hid_t g, const char *name, int value)
{
	hsize_t one = 1;
	hid_t s = H5Screate_simple(1, &one, NULL);
	hid_t a = H5Acreate2(g, name, H5T_NATIVE_INT, s, H5P_DEFAULT, H5P_DEFAULT);
	int rc = (a < 0 || H5Awrite(a, H5T_NATIVE_INT, &value) < 0) ? -1 : 0;
	if (a >= 0)
		H5Aclose(a);
	H5Sclose(s);
	return rc;
}

static hid_t mem_type(const char *type)
{
	if (!strcmp(type, "I4"))
		return H5T_NATIVE_INT32;
	if (!strcmp(type, "I8"))
		return H5T_NATIVE_INT64;
	if (!strcmp(type, "R4"))
		return H5T_NATIVE_FLOAT;
	if (!strcmp(type, "R8"))
		return H5T_NATIVE_DOUBLE;
	return H5T_NATIVE_CHAR;   /* C1 */
}

/* characters (with their NUL) as a 1-D char dataset */
static int str_data(hid_t g, const char *name, const char *text, size_t n)
{
	hsize_t d = n;
	hid_t s = H5Screate_simple(1, &d, NULL);
	hid_t ds = H5Dcreate2(g, name, H5T_NATIVE_CHAR, s, H5P_DEFAULT, H5P_DEFAULT,
			H5P_DEFAULT);
	int rc = (ds < 0 || H5Dwrite(ds, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL,
			H5P_DEFAULT, text) < 0) ? -1 : 0;
	if (ds >= 0)
		H5Dclose(ds);
	H5Sclose(s);
	return rc;
}

/*
 * Node `name` under `parent` with `label` and `type`; ndim dimensions dims
 * (in the node's own order) and data unless type is MT. Returns the open
 * group (the caller closes it), or a negative id.
 */
static hid_t node(hid_t parent, const char *name, const char *label,
		const char *type, int ndim, const long *dims, const void *data)
{
	hid_t p = gcpl();
	hid_t g = H5Gcreate2(parent, name, H5P_DEFAULT, p, H5P_DEFAULT);
	H5Pclose(p);
	if (g < 0)
		return g;
	if (str_att(g, "name", name, NAMELEN - 1) || str_att(g, "label", label, NAMELEN - 1) ||
			str_att(g, "type", type, 2) || int_att(g, "flags", 1)) {
		H5Gclose(g);
		return -1;
	}
	if (strcmp(type, "MT") != 0) {
		hsize_t hd[12];
		for (int i = 0; i < ndim; ++i)
			hd[i] = (hsize_t)dims[ndim - 1 - i];
		hid_t s = H5Screate_simple(ndim, hd, NULL);
		hid_t t = mem_type(type);
		hid_t ds = H5Dcreate2(g, " data", t, s, H5P_DEFAULT, H5P_DEFAULT,
				H5P_DEFAULT);
		int rc = (ds < 0 || H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT,
				data) < 0) ? -1 : 0;
		if (ds >= 0)
			H5Dclose(ds);
		H5Sclose(s);
		if (rc) {
			H5Gclose(g);
			return -1;
		}
	}
	return g;
}

static int leaf(hid_t parent, const char *name, const char *label,
		const char *type, int ndim, const long *dims, const void *data)
{
	hid_t g = node(parent, name, label, type, ndim, dims, data);
	if (g < 0)
		return -1;
	H5Gclose(g);
	return 0;
}

static int text_node(hid_t parent, const char *name, const char *label,
		const char *text)
{
	long n = (long)strlen(text);
	return leaf(parent, name, label, "C1", 1, &n, text);
}

/* a link node to file:path */
static int link_node(hid_t parent, const char *name, const char *file,
		const char *path)
{
	hid_t p = gcpl();
	hid_t g = H5Gcreate2(parent, name, H5P_DEFAULT, p, H5P_DEFAULT);
	H5Pclose(p);
	if (g < 0)
		return -1;
	int rc = str_att(g, "name", name, NAMELEN - 1) || str_att(g, "label", "", NAMELEN - 1) ||
		str_att(g, "type", "LK", 2) || int_att(g, "flags", 1) ||
		H5Lcreate_external(file, path, g, " link", H5P_DEFAULT, H5P_DEFAULT) < 0 ||
		str_data(g, " path", path, strlen(path) + 1) ||
		str_data(g, " file", file, strlen(file) + 1);
	H5Gclose(g);
	return rc ? -1 : 0;
}

/* a new file with the root node and the CGNSLibraryVersion node */
static hid_t create_file(const char *path, char *err)
{
	hid_t fc = H5Pcreate(H5P_FILE_CREATE);
	H5Pset_link_creation_order(fc, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
	hid_t fa = H5Pcreate(H5P_FILE_ACCESS);
	H5Pset_libver_bounds(fa, H5F_LIBVER_V18, H5F_LIBVER_V18);
	hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, fc, fa);
	H5Pclose(fc);
	H5Pclose(fa);
	if (f < 0) {
		seterr(err, "cannot create %s", path);
		return -1;
	}
	hid_t r = H5Gopen2(f, "/", H5P_DEFAULT);
	unsigned maj, min, rel;
	H5get_libversion(&maj, &min, &rel);
	char ver[NAMELEN];
	memset(ver, 0, sizeof(ver));
	snprintf(ver, sizeof(ver), "HDF5 Version %u.%u.%u", maj, min, rel);
	const char *fmt = "IEEE_LITTLE_32";
	float fv = CGNS_FILE_VERSION;
	long one = 1;
	int rc = str_att(r, "name", "HDF5 MotherNode", NAMELEN - 1) ||
		str_att(r, "label", "Root Node of HDF5 File", NAMELEN - 1) ||
		str_att(r, "type", "MT", 2) ||
		str_data(r, " format", fmt, strlen(fmt) + 1) ||
		str_data(r, " hdf5version", ver, NAMELEN) ||
		leaf(r, "CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4", 1, &one, &fv);
	H5Gclose(r);
	if (rc) {
		H5Fclose(f);
		seterr(err, "cannot write the root of %s", path);
		return -1;
	}
	return f;
}

/* Base (cell dimension 2, physical dimension 3) and Zone (unstructured,
 * nNodes vertices, nElems cells); returns the open zone, *base the base */
static hid_t base_zone(hid_t f, const ccfd_mesh *m, hid_t *base)
{
	int bd[2] = {2, 3};
	long two = 2;
	*base = node(f, "Base", "CGNSBase_t", "I4", 1, &two, bd);
	if (*base < 0)
		return -1;
	int size[3] = {(int)m->nNodes, (int)m->nElems, 0};
	long zd[2] = {1, 3};
	hid_t z = node(*base, "Zone", "Zone_t", "I4", 2, zd, size);
	if (z < 0 || text_node(z, "ZoneType", "ZoneType_t", "Unstructured")) {
		if (z >= 0)
			H5Gclose(z);
		H5Gclose(*base);
		return -1;
	}
	return z;
}

static int section(hid_t z, const char *name, int type, long first, long last,
		const int *conn, long nconn)
{
	int hdr[2] = {type, 0};
	long two = 2;
	hid_t s = node(z, name, "Elements_t", "I4", 1, &two, hdr);
	if (s < 0)
		return -1;
	int range[2] = {(int)first, (int)last};
	int rc = leaf(s, "ElementRange", "IndexRange_t", "I4", 1, &two, range) ||
		leaf(s, "ElementConnectivity", "DataArray_t", "I4", 1, &nconn, conn);
	H5Gclose(s);
	return rc ? -1 : 0;
}

static int is_periodic(const ccfd_mesh *m)
{
	for (int b = 0; b < m->nBC; ++b)
		if (m->bc[b].type == CCFD_BC_PERIODIC)
			return 1;
	return 0;
}

typedef struct {
	long a, b, k;
} edge_key;

static int edge_cmp(const void *x, const void *y)
{
	const edge_key *p = x, *q = y;
	if (p->a != q->a)
		return p->a < q->a ? -1 : 1;
	if (p->b != q->b)
		return p->b < q->b ? -1 : 1;
	return (p->k > q->k) - (p->k < q->k);
}

int ccfd_cgns_write_mesh(const char *path, const ccfd_mesh *m, long nBCedge,
		const long *bcEdge, char *err)
{
	quiet();
	if (!path || !m || (nBCedge && !bcEdge)) {
		seterr(err, "null argument");
		return -1;
	}
	const long nV = m->nNodes, nT = m->nTrias, nQ = m->nQuads, nE = m->nElems;
	double *xyz = malloc(sizeof(double) * 3 * (nV ? nV : 1));
	int *tri = malloc(sizeof(int) * 3 * (nT ? nT : 1));
	int *quad = malloc(sizeof(int) * 4 * (nQ ? nQ : 1));
	long nG = m->nGhosts;
	int *bars = malloc(sizeof(int) * 2 * (nG ? nG : 1));
	int *points = malloc(sizeof(int) * (nG ? nG : 1));
	edge_key *keys = malloc(sizeof(edge_key) * (nBCedge ? nBCedge : 1));
	int rc = -1;
	hid_t f = -1, base = -1, z = -1;
	if (!xyz || !tri || !quad || !bars || !points || !keys) {
		seterr(err, "out of memory");
		goto done;
	}
	for (long i = 0; i < nV; ++i) {
		xyz[i] = m->xy[2 * i];
		xyz[nV + i] = m->xy[2 * i + 1];
		xyz[2 * nV + i] = 0.0;
	}
	/* cells in element order, node ids from 1 */
	for (long e = 0, it = 0, iq = 0; e < nE; ++e) {
		if (m->elemType[e] == 3)
			for (int k = 0; k < 3; ++k)
				tri[3 * it + k] = (int)m->elemNode[4 * e + k] + 1;
		else
			for (int k = 0; k < 4; ++k)
				quad[4 * iq + k] = (int)m->elemNode[4 * e + k] + 1;
		it += m->elemType[e] == 3;
		iq += m->elemType[e] == 4;
	}
	f = create_file(path, err);
	if (f < 0)
		goto done;
	z = base_zone(f, m, &base);
	if (z < 0) {
		seterr(err, "cannot write the zone of %s", path);
		goto done;
	}
	hid_t gc = node(z, "GridCoordinates", "GridCoordinates_t", "MT", 0, NULL, NULL);
	if (gc < 0)
		goto fail;
	int bad = leaf(gc, "CoordinateX", "DataArray_t", "R8", 1, &nV, xyz) ||
		leaf(gc, "CoordinateY", "DataArray_t", "R8", 1, &nV, xyz + nV) ||
		leaf(gc, "CoordinateZ", "DataArray_t", "R8", 1, &nV, xyz + 2 * nV);
	H5Gclose(gc);
	if (bad)
		goto fail;
	long n3 = 3 * nT, n4 = 4 * nQ;
	if (nT > 0 && section(z, "Triangles", CG_TRI_3, 1, nT, tri, n3))
		goto fail;
	if (nQ > 0 && section(z, "Quadrilaterals", CG_QUAD_4, nT + 1, nT + nQ, quad, n4))
		goto fail;
	if (!is_periodic(m)) {
		/* the boundary sides per boundary condition, each in the
		 * reference's boundary-side list order (the last ghost first),
		 * with the nodes of the mesh file's edge */
		for (long k = 0; k < nBCedge; ++k) {
			long a = bcEdge[3 * k], b = bcEdge[3 * k + 1];
			keys[k].a = a < b ? a : b;
			keys[k].b = a < b ? b : a;
			keys[k].k = k;
		}
		qsort(keys, (size_t)nBCedge, sizeof(edge_key), edge_cmp);
		long nb = 0;
		long *part = malloc(sizeof(long) * ((size_t)m->nBC + 1));
		if (!part)
			goto fail;
		part[0] = 0;
		for (int c = 0; c < m->nBC; ++c) {
			for (long g = nG - 1; g >= 0; --g) {
				if (m->ghostBC[g] != c)
					continue;
				long s = m->ghostSlot[g];
				long a = m->slotNode[2 * s], b = m->slotNode[2 * s + 1];
				edge_key q = {a < b ? a : b, a < b ? b : a, -1};
				long lo = 0, hi = nBCedge;
				while (lo < hi) {      /* first key >= q */
					long mid = (lo + hi) / 2;
					if (edge_cmp(keys + mid, &q) < 0)
						lo = mid + 1;
					else
						hi = mid;
				}
				if (lo < nBCedge && keys[lo].a == q.a && keys[lo].b == q.b) {
					long k = keys[lo].k;
					a = bcEdge[3 * k];
					b = bcEdge[3 * k + 1];
				}
				bars[2 * nb] = (int)a + 1;
				bars[2 * nb + 1] = (int)b + 1;
				nb++;
			}
			part[c + 1] = nb;
		}
		long n2 = 2 * nG;
		if (section(z, "Boundaries", CG_BAR_2, nE + 1, nE + nG, bars, n2)) {
			free(part);
			goto fail;
		}
		hid_t zbc = node(z, "ZoneBC", "ZoneBC_t", "MT", 0, NULL, NULL);
		if (zbc < 0) {
			free(part);
			goto fail;
		}
		for (int c = 0; c < m->nBC && !bad; ++c) {
			char code[NAMELEN];
			snprintf(code, sizeof(code), "%d", m->bc[c].type * 100 + m->bc[c].id);
			long np = part[c + 1] - part[c];
			for (long i = 0; i < np; ++i)
				points[i] = (int)(nE + part[c] + i + 1);
			hid_t bcn = node(zbc, code, "BC_t", "C1", 1, (long[]){9}, "BCGeneral");
			long pd[2] = {1, np};
			/* an empty list is stored with one dimension of zero
			 * entries: the CGNS library refuses such a BC, which the
			 * reference then reports, and no case has one */
			bad = bcn < 0 ||
				(np > 0 && leaf(bcn, "PointList", "IndexArray_t", "I4", 2, pd, points)) ||
				text_node(bcn, "GridLocation", "GridLocation_t", "EdgeCenter");
			if (bcn >= 0)
				H5Gclose(bcn);
		}
		H5Gclose(zbc);
		free(part);
		if (bad)
			goto fail;
	}
	rc = 0;
	goto done;
fail:
	seterr(err, "cannot write %s", path);
done:
	if (z >= 0)
		H5Gclose(z);
	if (base >= 0)
		H5Gclose(base);
	if (f >= 0)
		H5Fclose(f);
	free(xyz);
	free(tri);
	free(quad);
	free(bars);
	free(points);
	free(keys);
	return rc;
}

/* links of the zone to the grid file's coordinates and cell sections */
static int grid_links(hid_t z, const ccfd_mesh *m, const char *gridFile)
{
	if (link_node(z, "GridCoordinates", gridFile, "/Base/Zone/GridCoordinates"))
		return -1;
	if (m->nTrias > 0 && link_node(z, "Triangles", gridFile, "/Base/Zone/Triangles"))
		return -1;
	if (m->nQuads > 0 &&
			link_node(z, "Quadrilaterals", gridFile, "/Base/Zone/Quadrilaterals"))
		return -1;
	return 0;
}

int ccfd_cgns_write_solution(const char *path, const char *gridFile,
		const ccfd_mesh *m, const double *prim, double time,
		double timeOverall, char *err)
{
	quiet();
	if (!path || !gridFile || !m || !prim) {
		seterr(err, "null argument");
		return -1;
	}
	const long N = m->nElems;
	double *a = malloc(sizeof(double) * 5 * (N ? N : 1));
	if (!a) {
		seterr(err, "out of memory");
		return -1;
	}
	for (long i = 0; i < N; ++i) {
		a[i] = prim[4 * i];
		a[N + i] = prim[4 * i + 1];
		a[2 * N + i] = prim[4 * i + 2];
		a[3 * N + i] = 0.0;
		a[4 * N + i] = prim[4 * i + 3];
	}
	hid_t base = -1, z = -1, fs = -1;
	int rc = -1;
	hid_t f = create_file(path, err);
	if (f < 0)
		goto done;
	z = base_zone(f, m, &base);
	if (z < 0 || grid_links(z, m, gridFile))
		goto fail;
	fs = node(z, "FlowSolution", "FlowSolution_t", "MT", 0, NULL, NULL);
	if (fs < 0 || text_node(fs, "GridLocation", "GridLocation_t", "CellCenter"))
		goto fail;
	static const char *names[5] = {"Density", "VelocityX", "VelocityY",
		"VelocityZ", "Pressure"};
	for (int v = 0; v < 5; ++v)
		if (leaf(fs, names[v], "DataArray_t", "R8", 1, &N, a + (size_t)v * N))
			goto fail;
	char text[128];
	snprintf(text, sizeof(text), "%20.12f %20.12f", time, timeOverall);
	if (text_node(base, "ConvergenceInfo", "Descriptor_t", text))
		goto fail;
	rc = 0;
	goto done;
fail:
	seterr(err, "cannot write %s", path);
done:
	if (fs >= 0)
		H5Gclose(fs);
	if (z >= 0)
		H5Gclose(z);
	if (base >= 0)
		H5Gclose(base);
	if (f >= 0)
		H5Fclose(f);
	free(a);
	return rc;
}

int ccfd_cgns_write_master(const char *path, const char *gridFile,
		const ccfd_mesh *m, int n, const double *times, const long *iters,
		const char *const *solutionFiles, char *err)
{
	quiet();
	if (!path || !gridFile || !m || n < 0 || (n && (!times || !iters ||
			!solutionFiles))) {
		seterr(err, "null argument");
		return -1;
	}
	char (*names)[32] = calloc((size_t)(n ? n : 1), 32);
	int *iv = calloc((size_t)(n ? n : 1), sizeof(int));
	hid_t base = -1, z = -1, bi = -1, zi = -1;
	int rc = -1;
	hid_t f = -1;
	if (!names || !iv) {
		seterr(err, "out of memory");
		goto done;
	}
	for (int k = 0; k < n; ++k)
		snprintf(names[k], 32, "FlowSolution%09ld", iters[k]);
	/* the words of the 64-bit iteration numbers, in memory order */
	memcpy(iv, iters, sizeof(int) * (size_t)n);
	f = create_file(path, err);
	if (f < 0)
		goto done;
	z = base_zone(f, m, &base);
	if (z < 0 || grid_links(z, m, gridFile))
		goto fail;
	/* an output written twice (a periodic output at the last iteration,
	 * then the final one) is linked once; TimeValues, IterationValues and
	 * FlowSolutionPointers keep every output, as the reference lists them
	 * (its library refuses the second link of the same name in write mode
	 * and stops the run there) */
	for (int k = 0; k < n; ++k) {
		int seen = 0;
		for (int j = 0; j < k && !seen; ++j)
			seen = !strcmp(names[j], names[k]);
		if (!seen && link_node(z, names[k], solutionFiles[k],
				"/Base/Zone/FlowSolution"))
			goto fail;
	}
	int nn = n;
	long one = 1, ln = n;
	bi = node(base, "TimeIterValues", "BaseIterativeData_t", "I4", 1, &one, &nn);
	if (bi < 0 || (n > 0 && (leaf(bi, "TimeValues", "DataArray_t", "R8", 1, &ln, times) ||
			leaf(bi, "IterationValues", "DataArray_t", "I4", 1, &ln, iv))))
		goto fail;
	zi = node(z, "ZoneIterativeData", "ZoneIterativeData_t", "MT", 0, NULL, NULL);
	long pd[2] = {32, n};
	if (zi < 0 || (n > 0 && leaf(zi, "FlowSolutionPointers", "DataArray_t", "C1",
			2, pd, names)))
		goto fail;
	rc = 0;
	goto done;
fail:
	seterr(err, "cannot write %s", path);
done:
	if (zi >= 0)
		H5Gclose(zi);
	if (bi >= 0)
		H5Gclose(bi);
	if (z >= 0)
		H5Gclose(z);
	if (base >= 0)
		H5Gclose(base);
	if (f >= 0)
		H5Fclose(f);
	free(names);
	free(iv);
	return rc;
}

/* the label attribute of a child group; 0 when read */
static int child_label(hid_t parent, const char *name, char *label)
{
	hid_t g = H5Gopen2(parent, name, H5P_DEFAULT);
	if (g < 0)
		return -1;
	int rc = -1;
	if (H5Aexists(g, "label") > 0) {
		hid_t a = H5Aopen(g, "label", H5P_DEFAULT);
		hid_t t = H5Tcopy(H5T_C_S1);
		H5Tset_size(t, NAMELEN);
		char buf[NAMELEN + 1];
		memset(buf, 0, sizeof(buf));
		if (a >= 0 && H5Aread(a, t, buf) >= 0) {
			strcpy(label, buf);
			rc = 0;
		}
		H5Tclose(t);
		if (a >= 0)
			H5Aclose(a);
	}
	H5Gclose(g);
	return rc;
}

/* name of the first child (creation order) with the label; 0 when found */
static int first_child(hid_t parent, const char *label, char *name)
{
	H5G_info_t info;
	if (H5Gget_info(parent, &info) < 0)
		return -1;
	for (hsize_t k = 0; k < info.nlinks; ++k) {
		char nm[256], lb[NAMELEN + 1];
		if (H5Lget_name_by_idx(parent, ".", H5_INDEX_CRT_ORDER, H5_ITER_INC, k,
				nm, sizeof(nm), H5P_DEFAULT) < 0)
			continue;
		if (nm[0] == ' ')
			continue;
		if (child_label(parent, nm, lb) == 0 && !strcmp(lb, label)) {
			strcpy(name, nm);
			return 0;
		}
	}
	return -1;
}

/* a node's data as doubles or longs (HDF5 converts); count checked */
static int read_data(hid_t parent, const char *name, hid_t mtype, void *out,
		long count)
{
	hid_t g = H5Gopen2(parent, name, H5P_DEFAULT);
	if (g < 0)
		return -1;
	hid_t d = H5Dopen2(g, " data", H5P_DEFAULT);
	int rc = -1;
	if (d >= 0) {
		hid_t s = H5Dget_space(d);
		if (H5Sget_simple_extent_npoints(s) == count &&
				H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0)
			rc = 0;
		H5Sclose(s);
		H5Dclose(d);
	}
	H5Gclose(g);
	return rc;
}

int ccfd_cgns_read_solution(const char *path, long nElems, double *prim,
		double *time, double *timeOverall, char *err)
{
	quiet();
	if (!path || !prim || !time || !timeOverall || nElems <= 0) {
		seterr(err, "null argument");
		return -1;
	}
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) {
		seterr(err, "cannot open %s", path);
		return -1;
	}
	int rc = -1;
	hid_t base = -1, z = -1, fs = -1;
	double *a = malloc(sizeof(double) * 4 * (size_t)nElems);
	char name[256];
	if (!a) {
		seterr(err, "out of memory");
		goto done;
	}
	if (first_child(f, "CGNSBase_t", name) || (base = H5Gopen2(f, name, H5P_DEFAULT)) < 0 ||
			first_child(base, "Zone_t", name) || (z = H5Gopen2(base, name, H5P_DEFAULT)) < 0) {
		seterr(err, "%s: no CGNS base and zone", path);
		goto done;
	}
	{
		long size[3];
		hid_t d = H5Dopen2(z, " data", H5P_DEFAULT);
		int ok = d >= 0 && H5Dread(d, H5T_NATIVE_LONG, H5S_ALL, H5S_ALL,
				H5P_DEFAULT, size) >= 0;
		if (d >= 0)
			H5Dclose(d);
		if (!ok) {
			seterr(err, "%s: unreadable zone size", path);
			goto done;
		}
		if (size[1] != nElems) {
			seterr(err, "Wrong Number of Elements in CGNS flow solution");
			goto done;
		}
	}
	if (first_child(z, "FlowSolution_t", name) || (fs = H5Gopen2(z, name, H5P_DEFAULT)) < 0) {
		seterr(err, "%s: no flow solution", path);
		goto done;
	}
	static const char *names[4] = {"Density", "VelocityX", "VelocityY", "Pressure"};
	for (int v = 0; v < 4; ++v)
		if (read_data(fs, names[v], H5T_NATIVE_DOUBLE, a + (size_t)v * nElems, nElems)) {
			seterr(err, "%s: cannot read %s", path, names[v]);
			goto done;
		}
	if (first_child(base, "Descriptor_t", name)) {
		seterr(err, "%s: no descriptor", path);
		goto done;
	}
	{
		hid_t g = H5Gopen2(base, name, H5P_DEFAULT);
		hid_t d = g >= 0 ? H5Dopen2(g, " data", H5P_DEFAULT) : -1;
		char text[256];
		memset(text, 0, sizeof(text));
		int ok = 0;
		if (d >= 0) {
			hid_t s = H5Dget_space(d);
			hssize_t n = H5Sget_simple_extent_npoints(s);
			H5Sclose(s);
			if (n > 0 && n < (hssize_t)sizeof(text))
				ok = H5Dread(d, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT,
						text) >= 0;
			H5Dclose(d);
		}
		if (g >= 0)
			H5Gclose(g);
		if (!ok || sscanf(text, "%lg %lg", time, timeOverall) != 2) {
			seterr(err, "%s: unreadable descriptor", path);
			goto done;
		}
	}
	for (long i = 0; i < nElems; ++i) {
		prim[4 * i] = a[i];
		prim[4 * i + 1] = a[nElems + i];
		prim[4 * i + 2] = a[2 * nElems + i];
		prim[4 * i + 3] = a[3 * nElems + i];
	}
	rc = 0;
done:
	if (fs >= 0)
		H5Gclose(fs);
	if (z >= 0)
		H5Gclose(z);
	if (base >= 0)
		H5Gclose(base);
	H5Fclose(f);
	free(a);
	return rc;
}

/* ------------------------------------------------------------ mesh input */

/* a fixed-length string attribute; 0 when read */
static int attr_str(hid_t g, const char *name, char *out, size_t size)
{
	out[0] = 0;
	if (H5Aexists(g, name) <= 0)
		return -1;
	hid_t a = H5Aopen(g, name, H5P_DEFAULT);
	if (a < 0)
		return -1;
	hid_t t = H5Tcopy(H5T_C_S1);
	H5Tset_size(t, size - 1);
	char buf[NAMELEN + 1];
	memset(buf, 0, sizeof(buf));
	int rc = size <= sizeof(buf) && H5Aread(a, t, buf) >= 0 ? 0 : -1;
	if (!rc)
		snprintf(out, size, "%s", buf);
	H5Tclose(t);
	H5Aclose(a);
	return rc;
}

/* open a child node; a link node (type LK) is followed to its target */
static hid_t open_node(hid_t parent, const char *name)
{
	hid_t g = H5Gopen2(parent, name, H5P_DEFAULT);
	if (g < 0)
		return -1;
	char type[4];
	if (attr_str(g, "type", type, sizeof(type)) == 0 && !strcmp(type, "LK")) {
		hid_t t = H5Gopen2(g, " link", H5P_DEFAULT);
		H5Gclose(g);
		return t;
	}
	return g;
}

#define MAXKIDS 256
typedef struct {
	int n;
	char name[MAXKIDS][NAMELEN + 1];
} kids_t;

/* the children with a label, in creation order where the file tracks it
 * (else by name, as the CGNS library lists them); 0 when listed */
static int kids(hid_t parent, const char *label, kids_t *out)
{
	H5G_info_t info;
	out->n = 0;
	if (H5Gget_info(parent, &info) < 0)
		return -1;
	H5_index_t idx = H5_INDEX_CRT_ORDER;
	char nm[256];
	if (info.nlinks > 0 && H5Lget_name_by_idx(parent, ".", idx, H5_ITER_INC, 0,
			nm, sizeof(nm), H5P_DEFAULT) < 0)
		idx = H5_INDEX_NAME;
	for (hsize_t k = 0; k < info.nlinks; ++k) {
		if (H5Lget_name_by_idx(parent, ".", idx, H5_ITER_INC, k, nm, sizeof(nm),
				H5P_DEFAULT) < 0)
			return -1;
		if (nm[0] == ' ' || strlen(nm) > NAMELEN)
			continue;
		hid_t g = open_node(parent, nm);
		if (g < 0)
			continue;
		char lb[NAMELEN + 1];
		int match = attr_str(g, "label", lb, sizeof(lb)) == 0 && !strcmp(lb, label);
		H5Gclose(g);
		if (!match)
			continue;
		if (out->n == MAXKIDS)
			return -1;
		strcpy(out->name[out->n++], nm);
	}
	return 0;
}

/* number of values of a node's data and its HDF5 dimensions (reversed
 * from the CGNS order); -1 without data */
static long data_dims(hid_t g, int *nd, hsize_t dims[8])
{
	hid_t d = H5Dopen2(g, " data", H5P_DEFAULT);
	if (d < 0)
		return -1;
	hid_t s = H5Dget_space(d);
	int n = H5Sget_simple_extent_ndims(s);
	long count = -1;
	if (n >= 0 && n <= 8) {
		H5Sget_simple_extent_dims(s, dims, NULL);
		count = (long)H5Sget_simple_extent_npoints(s);
		*nd = n;
	}
	H5Sclose(s);
	H5Dclose(d);
	return count;
}

/* a node's data converted to the memory type; count checked */
static int data_read(hid_t g, hid_t mtype, void *out, long count)
{
	int nd;
	hsize_t dims[8];
	if (data_dims(g, &nd, dims) != count)
		return -1;
	hid_t d = H5Dopen2(g, " data", H5P_DEFAULT);
	int rc = H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0 ? 0 : -1;
	H5Dclose(d);
	return rc;
}

/* the codes of the zone's BC_t nodes on consecutive boundary edges, as
 * many per node as its point set counts (cg_boco_info's npnts) */
static int bc_codes(hid_t z, long *bce, long nb, const char *path, char *err,
		kids_t *k)
{
	if (kids(z, "ZoneBC_t", k)) {
		seterr(err, "%s: unreadable zone", path);
		return -1;
	}
	if (k->n == 0)
		return 0;
	hid_t zbc = open_node(z, k->name[0]);
	if (zbc < 0 || kids(zbc, "BC_t", k)) {
		if (zbc >= 0)
			H5Gclose(zbc);
		seterr(err, "%s: unreadable ZoneBC", path);
		return -1;
	}
	static const char *sets[4] = {"PointList", "PointRange", "ElementList",
		"ElementRange"};
	long e = 0;
	int rc = 0;
	for (int c = 0; c < k->n && !rc; ++c) {
		int code;
		if (sscanf(k->name[c], "%d", &code) != 1) {
			seterr(err, "%s: boundary condition %s is not a boundary code",
					path, k->name[c]);
			rc = -1;
			break;
		}
		hid_t b = open_node(zbc, k->name[c]);
		long np = 0;
		for (int q = 0; q < 4 && b >= 0; ++q) {
			if (H5Lexists(b, sets[q], H5P_DEFAULT) <= 0)
				continue;
			hid_t ps = open_node(b, sets[q]);
			int nd = 0;
			hsize_t dims[8];
			long cnt = ps >= 0 ? data_dims(ps, &nd, dims) : -1;
			if (ps >= 0)
				H5Gclose(ps);
			if (cnt < 0 || nd != 2) {
				seterr(err, "%s: invalid point set of %s", path, k->name[c]);
				rc = -1;
			}
			np = rc ? 0 : (long)dims[0];
			break;
		}
		if (b >= 0)
			H5Gclose(b);
		if (b < 0) {
			seterr(err, "%s: unreadable boundary condition %s", path, k->name[c]);
			rc = -1;
		}
		for (long p = 0; p < np && !rc; ++p) {
			if (e >= nb) {
				seterr(err, "%s: more boundary condition points than boundary "
						"edges (%ld)", path, nb);
				rc = -1;
				break;
			}
			bce[3 * e++ + 2] = code;
		}
	}
	H5Gclose(zbc);
	return rc;
}

int ccfd_cgns_read_mesh(const char *path, long *nVert, double **xy,
		long *nTri, long **tri, long *nQuad, long **quad,
		long *nBCedge, long **bcEdge, char *err)
{
	quiet();
	if (!path || !nVert || !xy || !nTri || !tri || !nQuad || !quad ||
			!nBCedge || !bcEdge) {
		seterr(err, "null argument");
		return -1;
	}
	*nVert = *nTri = *nQuad = *nBCedge = 0;
	*xy = NULL;
	*tri = *quad = *bcEdge = NULL;
	hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0) {
		seterr(err, "cannot open %s", path);
		return -1;
	}
	int rc = -1;
	hid_t base = -1, z = -1, gc = -1;
	long *conn = NULL;
	double *xs = NULL;
	kids_t *k = malloc(sizeof(kids_t));
	kids_t *ks = malloc(sizeof(kids_t));
	long nv = 0, nt = 0, nq = 0, nb = 0;
	if (!k || !ks) {
		seterr(err, "out of memory");
		goto done;
	}
	if (kids(f, "CGNSBase_t", k) || k->n == 0 ||
			(base = open_node(f, k->name[0])) < 0) {
		seterr(err, "%s: no CGNS base", path);
		goto done;
	}
	if (kids(base, "Zone_t", k) || k->n == 0) {
		seterr(err, "%s: no zone", path);
		goto done;
	}
	int zi = 0;
	for (int i = 1; i < k->n; ++i)
		if (strcmp(k->name[i], k->name[zi]) < 0)
			zi = i;
	long size[3];
	if ((z = open_node(base, k->name[zi])) < 0 ||
			data_read(z, H5T_NATIVE_LONG, size, 3)) {
		seterr(err, "%s: zone %s is not an unstructured zone", path, k->name[zi]);
		goto done;
	}
	nv = size[0];
	if (nv <= 0 || nv > (1L << 31)) {
		seterr(err, "%s: %ld nodes", path, nv);
		goto done;
	}
	xs = malloc(sizeof(double) * 2 * (size_t)nv);
	*xy = malloc(sizeof(double) * 2 * (size_t)nv);
	if (!xs || !*xy) {
		seterr(err, "out of memory");
		goto done;
	}
	if ((gc = open_node(z, "GridCoordinates")) < 0 ||
			read_data(gc, "CoordinateX", H5T_NATIVE_DOUBLE, xs, nv) ||
			read_data(gc, "CoordinateY", H5T_NATIVE_DOUBLE, xs + nv, nv)) {
		seterr(err, "%s: cannot read the coordinates of %ld nodes", path, nv);
		goto done;
	}
	for (long i = 0; i < nv; ++i) {
		(*xy)[2 * i] = xs[i];
		(*xy)[2 * i + 1] = xs[nv + i];
	}
	if (kids(z, "Elements_t", ks)) {
		seterr(err, "%s: unreadable sections", path);
		goto done;
	}
	for (int s = 0; s < ks->n; ++s) {
		hid_t g = open_node(z, ks->name[s]);
		int hdr[2];
		hid_t ec = -1;
		int nd;
		hsize_t dims[8];
		long cnt = -1;
		if (g >= 0 && data_read(g, H5T_NATIVE_INT, hdr, 2) == 0 &&
				(ec = open_node(g, "ElementConnectivity")) >= 0)
			cnt = data_dims(ec, &nd, dims);
		free(conn);
		conn = cnt > 0 ? malloc(sizeof(long) * (size_t)cnt) : NULL;
		int ok = cnt >= 0 && (cnt == 0 || (conn &&
				data_read(ec, H5T_NATIVE_LONG, conn, cnt) == 0));
		if (ec >= 0)
			H5Gclose(ec);
		if (g >= 0)
			H5Gclose(g);
		if (!ok) {
			seterr(err, "%s: unreadable section %s", path, ks->name[s]);
			goto done;
		}
		int npe = hdr[0] == CG_BAR_2 ? 2 : hdr[0] == CG_TRI_3 ? 3 :
			hdr[0] == CG_QUAD_4 ? 4 : 0;
		if (hdr[0] == CG_MIXED) {
			seterr(err, "%s: section %s is MIXED; the CGNS library reads "
					"only fixed-size element sections", path, ks->name[s]);
			goto done;
		}
		if (!npe) {
			seterr(err, "Only TRI_3 and QUAD_4 Elements allowed");
			goto done;
		}
		long n = cnt / npe, w = npe == 2 ? 3 : npe + 1;
		long *out = calloc((size_t)(n ? n : 1) * (size_t)w, sizeof(long));
		if (!out) {
			seterr(err, "out of memory");
			goto done;
		}
		for (long i = 0; i < n; ++i) {
			for (int j = 0; j < npe; ++j) {
				long v = conn[npe * i + j] - 1;
				if (v < 0 || v >= nv) {
					seterr(err, "%s: section %s: node %ld out of range",
							path, ks->name[s], v + 1);
					free(out);
					goto done;
				}
				out[w * i + j] = v;
			}
			if (npe > 2)
				out[w * i + npe] = 1;
		}
		if (npe == 2) {
			free(*bcEdge);
			*bcEdge = out;
			nb = n;
			if (bc_codes(z, out, n, path, err, k))
				goto done;
		} else if (npe == 3) {
			free(*tri);
			*tri = out;
			nt = n;
		} else {
			free(*quad);
			*quad = out;
			nq = n;
		}
	}
	*nVert = nv;
	*nTri = nt;
	*nQuad = nq;
	*nBCedge = nb;
	rc = 0;
done:
	if (rc) {
		free(*xy);
		free(*tri);
		free(*quad);
		free(*bcEdge);
		*xy = NULL;
		*tri = *quad = *bcEdge = NULL;
	}
	if (gc >= 0)
		H5Gclose(gc);
	if (z >= 0)
		H5Gclose(z);
	if (base >= 0)
		H5Gclose(base);
	H5Fclose(f);
	free(conn);
	free(xs);
	free(k);
	free(ks);
	return rc;
}
