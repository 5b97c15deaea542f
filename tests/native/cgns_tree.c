/*
 * cgns_tree.c - test helper: print the node tree of an HDF5 file in the
 * CGNS/HDF5 layout, one line per node in link-creation order:
 *   path|name|label|type|flags|dims (HDF5 order)|values
 * values: the first 4096 numbers (or the characters) of " data"; links print
 * their " file" and " path" instead. Root datasets are listed as
 * path|@dataset|characters. Host only; used by tests/test_cgns_host.py.
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void att_str(hid_t g, const char *name, char *out, size_t n)
{
	out[0] = 0;
	if (H5Aexists(g, name) <= 0)
		return;
	hid_t a = H5Aopen(g, name, H5P_DEFAULT);
	hid_t t = H5Aget_type(a);
	if (H5Tget_class(t) == H5T_STRING) {
		hid_t m = H5Tcopy(H5T_C_S1);
		H5Tset_size(m, n - 1);
		char buf[256];
		memset(buf, 0, sizeof(buf));
		H5Aread(a, m, buf);
		snprintf(out, n, "%s", buf);
		H5Tclose(m);
	} else {
		int v = 0;
		H5Aread(a, H5T_NATIVE_INT, &v);
		snprintf(out, n, "%d", v);
	}
	H5Tclose(t);
	H5Aclose(a);
}

static void chars(hid_t g, const char *ds, char *out, size_t n)
{
	out[0] = 0;
	if (H5Lexists(g, ds, H5P_DEFAULT) <= 0)
		return;
	hid_t d = H5Dopen2(g, ds, H5P_DEFAULT);
	hid_t s = H5Dget_space(d);
	hssize_t k = H5Sget_simple_extent_npoints(s);
	char *buf = calloc((size_t)k + 1, 1);
	H5Dread(d, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
	for (hssize_t i = 0; i < k; ++i)
		if (buf[i] == 0)
			buf[i] = '~';      /* NULs made visible */
	snprintf(out, n, "%s", buf);
	free(buf);
	H5Sclose(s);
	H5Dclose(d);
}

static void walk(hid_t g, const char *path)
{
	char name[64], label[64], type[8], flags[16];
	att_str(g, "name", name, sizeof(name));
	att_str(g, "label", label, sizeof(label));
	att_str(g, "type", type, sizeof(type));
	att_str(g, "flags", flags, sizeof(flags));
	printf("%s|%s|%s|%s|%s|", path, name, label, type, flags);
	if (!strcmp(type, "LK")) {
		char f[512], p[512];
		chars(g, " file", f, sizeof(f));
		chars(g, " path", p, sizeof(p));
		H5L_info_t li;
		int ext = H5Lget_info(g, " link", &li, H5P_DEFAULT) >= 0 &&
			li.type == H5L_TYPE_EXTERNAL;
		printf("|%s:%s:%s\n", f, p, ext ? "external" : "missing");
		return;
	}
	if (H5Lexists(g, " data", H5P_DEFAULT) > 0) {
		hid_t d = H5Dopen2(g, " data", H5P_DEFAULT);
		hid_t s = H5Dget_space(d);
		hsize_t dims[8];
		int nd = H5Sget_simple_extent_dims(s, dims, NULL);
		for (int i = 0; i < nd; ++i)
			printf(i ? "x%llu" : "%llu", (unsigned long long)dims[i]);
		printf("|");
		hssize_t k = H5Sget_simple_extent_npoints(s);
		if (!strcmp(type, "C1")) {
			char buf[1024];
			chars(g, " data", buf, sizeof(buf));
			printf("%s", buf);
		} else {
			double *v = malloc(sizeof(double) * (size_t)k);
			H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
			for (hssize_t i = 0; i < k && i < 4096; ++i)
				printf(i ? ",%.17g" : "%.17g", v[i]);
			free(v);
		}
		H5Sclose(s);
		H5Dclose(d);
	} else {
		printf("|");
	}
	printf("\n");
	H5G_info_t info;
	H5Gget_info(g, &info);
	for (hsize_t i = 0; i < info.nlinks; ++i) {
		char nm[256];
		H5Lget_name_by_idx(g, ".", H5_INDEX_CRT_ORDER, H5_ITER_INC, i, nm,
				sizeof(nm), H5P_DEFAULT);
		if (nm[0] == ' ')
			continue;
		char sub[1024];
		snprintf(sub, sizeof(sub), "%s/%s", strcmp(path, "/") ? path : "", nm);
		hid_t c = H5Gopen2(g, nm, H5P_DEFAULT);
		walk(c, sub);
		H5Gclose(c);
	}
}

int main(int argc, char **argv)
{
	if (argc != 2)
		return 2;
	hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
	if (f < 0)
		return 1;
	hid_t r = H5Gopen2(f, "/", H5P_DEFAULT);
	char buf[256];
	chars(r, " format", buf, sizeof(buf));
	printf("/|@format|%s\n", buf);
	chars(r, " hdf5version", buf, sizeof(buf));
	printf("/|@hdf5version|%s\n", buf);
	walk(r, "/");
	H5Gclose(r);
	H5Fclose(f);
	return 0;
}
