/*
 * cgns_sanitize.c - test helper: every entry point of libccfd_cgns.so on a
 * small mixed mesh, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer: grid file, two flow solutions, master file,
 * restart read, mesh read (of the grid file and, through its links, of a
 * solution file), and the refusals (missing files, a wrong element count,
 * null arguments). Prints one line per call. Host only; used by
 * tests/test_host_sanitizers.py.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_cgns.h"
#include "ccfd_mesh.h"

static void read_mesh(const char *path)
{
	long nv, nt, nq, nb;
	double *xy;
	long *tri, *quad, *bce;
	char err[256];
	int rc = ccfd_cgns_read_mesh(path, &nv, &xy, &nt, &tri, &nq, &quad, &nb,
			&bce, err);
	if (rc)
		printf("read_mesh %s: %d %s\n", path, rc, err);
	else
		printf("read_mesh %s: %ld nodes %ld tri %ld quad %ld edges\n", path,
				nv, nt, nq, nb);
	free(xy);
	free(tri);
	free(quad);
	free(bce);
}

int main(void)
{
	/* 2 x 1 quads, the right one split into two triangles */
	const double xy[12] = {0, 0, 1, 0, 2, 0, 0, 1, 1, 1, 2, 1};
	const long quad[5] = {0, 1, 4, 3, 1};
	const long tri[8] = {1, 2, 5, 1, 1, 5, 4, 1};
	const long bce[18] = {0, 1, 101, 1, 2, 101, 2, 5, 401, 5, 4, 101,
		4, 3, 101, 3, 0, 401};
	ccfd_bc bc[2];
	memset(bc, 0, sizeof(bc));
	bc[0].type = 1;
	bc[0].id = 1;
	bc[1].type = 4;
	bc[1].id = 1;
	char err[256];
	ccfd_mesh *m = ccfd_mesh_build(6, xy, 2, tri, 1, quad, 6, bce, 2, bc, err);
	if (!m) {
		printf("mesh: %s\n", err);
		return 1;
	}
	double prim[12];
	for (int i = 0; i < 12; ++i)
		prim[i] = 1.0 + 0.1 * i;
	printf("write_mesh: %d\n", ccfd_cgns_write_mesh("g.cgns", m, 6, bce, err));
	printf("write_solution: %d\n", ccfd_cgns_write_solution("s_1.cgns",
			"g.cgns", m, prim, 0.5, 0.0, err));
	printf("write_solution: %d\n", ccfd_cgns_write_solution("s_2.cgns",
			"g.cgns", m, prim, 1.0, 0.0, err));
	const double times[3] = {0.5, 1.0, 1.0};
	const long iters[3] = {1, 2, 2};
	const char *files[3] = {"s_1.cgns", "s_2.cgns", "s_2.cgns"};
	printf("write_master: %d\n", ccfd_cgns_write_master("master.cgns",
			"g.cgns", m, 3, times, iters, files, err));
	double back[12], t, to;
	int rc = ccfd_cgns_read_solution("s_2.cgns", 3, back, &t, &to, err);
	printf("read_solution: %d same %d t %g\n", rc,
			!rc && !memcmp(back, prim, sizeof(prim)), t);
	rc = ccfd_cgns_read_solution("s_2.cgns", 4, back, &t, &to, err);
	printf("read_solution wrong count: %d %s\n", rc, err);
	rc = ccfd_cgns_read_solution("absent.cgns", 3, back, &t, &to, err);
	printf("read_solution absent: %d %s\n", rc, err);
	read_mesh("g.cgns");
	read_mesh("s_1.cgns");
	read_mesh("absent.cgns");
	printf("null: %d\n", ccfd_cgns_read_mesh(NULL, NULL, NULL, NULL, NULL,
			NULL, NULL, NULL, NULL, err));
	ccfd_mesh_free(m);
	return 0;
}
