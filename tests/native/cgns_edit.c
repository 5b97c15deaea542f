/*
 * cgns_edit.c - test helper: small edits of a CGNS/HDF5 file, to make the
 * variants the mesh reader must handle or refuse:
 *   settype FILE NODE VALUE     first int of NODE's data (a section's
 *                               element type)
 *   rename  FILE NODE NEWNAME   rename a node (link and "name" attribute)
 *   range   FILE BCNODE A B     replace the BC's PointList by a PointRange
 *                               node holding A..B
 *   copy    FILE NODE NEWNODE   copy a node (and its subtree)
 *   delete  FILE NODE           remove a node
 * NODE paths are absolute HDF5 paths. Host only; used by
 * tests/test_cgns_host.py.
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int set_str_attr(hid_t g, const char *name, const char *value, int len)
{
	if (H5Aexists(g, name) > 0)
		H5Adelete(g, name);
	hid_t t = H5Tcopy(H5T_C_S1);
	H5Tset_size(t, (size_t)len);
	hid_t s = H5Screate(H5S_SCALAR);
	hid_t a = H5Acreate2(g, name, t, s, H5P_DEFAULT, H5P_DEFAULT);
	char buf[64];
	memset(buf, 0, sizeof(buf));
	snprintf(buf, sizeof(buf), "%s", value);
	int rc = a < 0 || H5Awrite(a, t, buf) < 0;
	if (a >= 0)
		H5Aclose(a);
	H5Sclose(s);
	H5Tclose(t);
	return rc;
}

static const char *base_name(const char *path)
{
	const char *s = strrchr(path, '/');
	return s ? s + 1 : path;
}

int main(int argc, char **argv)
{
	if (argc < 4)
		return 2;
	hid_t f = H5Fopen(argv[2], H5F_ACC_RDWR, H5P_DEFAULT);
	if (f < 0)
		return 1;
	int rc = 1;
	if (!strcmp(argv[1], "settype") && argc == 5) {
		char p[512];
		snprintf(p, sizeof(p), "%s/ data", argv[3]);
		hid_t d = H5Dopen2(f, p, H5P_DEFAULT);
		int v[2];
		if (d >= 0 && H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL,
				H5P_DEFAULT, v) >= 0) {
			v[0] = atoi(argv[4]);
			rc = H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL,
					H5P_DEFAULT, v) < 0;
		}
		if (d >= 0)
			H5Dclose(d);
	} else if (!strcmp(argv[1], "rename") && argc == 5) {
		char to[512];
		snprintf(to, sizeof(to), "%.*s/%s", (int)(base_name(argv[3]) - argv[3] - 1),
				argv[3], argv[4]);
		if (H5Lmove(f, argv[3], f, to, H5P_DEFAULT, H5P_DEFAULT) >= 0) {
			hid_t g = H5Gopen2(f, to, H5P_DEFAULT);
			rc = g < 0 || set_str_attr(g, "name", argv[4], 33);
			if (g >= 0)
				H5Gclose(g);
		}
	} else if (!strcmp(argv[1], "range") && argc == 6) {
		hid_t b = H5Gopen2(f, argv[3], H5P_DEFAULT);
		if (b >= 0) {
			if (H5Lexists(b, "PointList", H5P_DEFAULT) > 0)
				H5Ldelete(b, "PointList", H5P_DEFAULT);
			hid_t cp = H5Pcreate(H5P_GROUP_CREATE);
			H5Pset_link_creation_order(cp, H5P_CRT_ORDER_TRACKED |
					H5P_CRT_ORDER_INDEXED);
			hid_t g = H5Gcreate2(b, "PointRange", H5P_DEFAULT, cp, H5P_DEFAULT);
			H5Pclose(cp);
			int bad = g < 0 || set_str_attr(g, "name", "PointRange", 33) ||
				set_str_attr(g, "label", "IndexRange_t", 33) ||
				set_str_attr(g, "type", "I4", 3);
			if (!bad) {
				hid_t s = H5Screate(H5S_SCALAR);
				hid_t a = H5Acreate2(g, "flags", H5T_NATIVE_INT, s,
						H5P_DEFAULT, H5P_DEFAULT);
				int one = 1;
				bad = a < 0 || H5Awrite(a, H5T_NATIVE_INT, &one) < 0;
				if (a >= 0)
					H5Aclose(a);
				H5Sclose(s);
			}
			if (!bad) {
				hsize_t dims[2] = {2, 1};   /* CGNS (1, 2), reversed */
				hid_t s = H5Screate_simple(2, dims, NULL);
				hid_t d = H5Dcreate2(g, " data", H5T_NATIVE_INT, s,
						H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
				int v[2] = {atoi(argv[4]), atoi(argv[5])};
				bad = d < 0 || H5Dwrite(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL,
						H5P_DEFAULT, v) < 0;
				if (d >= 0)
					H5Dclose(d);
				H5Sclose(s);
			}
			if (g >= 0)
				H5Gclose(g);
			H5Gclose(b);
			rc = bad;
		}
	} else if (!strcmp(argv[1], "copy") && argc == 5) {
		if (H5Ocopy(f, argv[3], f, argv[4], H5P_DEFAULT, H5P_DEFAULT) >= 0) {
			hid_t g = H5Gopen2(f, argv[4], H5P_DEFAULT);
			rc = g < 0 || set_str_attr(g, "name", base_name(argv[4]), 33);
			if (g >= 0)
				H5Gclose(g);
		}
	} else if (!strcmp(argv[1], "delete") && argc == 4) {
		rc = H5Ldelete(f, argv[3], H5P_DEFAULT) < 0;
	}
	H5Fclose(f);
	return rc;
}
