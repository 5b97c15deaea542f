/*
 * host_sanitize.c - drives the host C library (mesh.c, ini.c, driver.c)
 * through its set-up and error paths; tests/test_host_sanitizers.py builds
 * it with AddressSanitizer, UndefinedBehaviorSanitizer and leak detection.
 * Test infrastructure. Arguments: case files (.ini), then "--", then gmsh
 * files (.msh). No device is needed: ccfd_init stops where it would create
 * the device context when there is none.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccfd_dropin.h"

static void mesh_file(const char *path)
{
	char err[256];
	long nv, nt, nq, nb;
	double *xy = NULL;
	long *tri = NULL, *quad = NULL, *bce = NULL;
	size_t L = strlen(path);
	int emc2 = L > 5 && !strcmp(path + L - 5, ".mesh");
	if ((emc2 ? ccfd_emc2_read : ccfd_gmsh_read)(path, &nv, &xy, &nt, &tri,
			&nq, &quad, &nb, &bce, err)) {
		printf("read %s: %s\n", path, err);
		return;
	}
	/* one boundary condition per code on the file (slip wall values) */
	ccfd_bc bc[32];
	int nbc = 0;
	memset(bc, 0, sizeof(bc));
	for (long e = 0; e < nb; ++e) {
		int code = (int)bce[3 * e + 2], seen = 0;
		for (int k = 0; k < nbc; ++k)
			seen |= bc[k].type * 100 + bc[k].id == code;
		if (!seen && nbc < 32) {
			bc[nbc].type = code / 100;
			bc[nbc].id = code % 100;
			bc[nbc].pVar[0] = bc[nbc].pVar[3] = 1.0;
			nbc++;
		}
	}
	ccfd_mesh *m = ccfd_mesh_build(nv, xy, nt, tri, nq, quad, nb, bce, nbc,
			bc, err);
	printf("mesh %s: %ld elements\n", path, m ? m->nElems : -1L);
	if (m)
		ccfd_mesh_free(m);
	/* an open mesh (one boundary edge short) must be refused */
	if (nb > 1) {
		ccfd_mesh *b = ccfd_mesh_build(nv, xy, nt, tri, nq, quad, nb - 1,
				bce, nbc, bc, err);
		printf("  open mesh: %s\n", b ? "ACCEPTED" : "refused");
		if (b)
			ccfd_mesh_free(b);
	}
	free(xy);
	free(tri);
	free(quad);
	free(bce);
}

int main(int argc, char *argv[])
{
	char err[256];
	int i = 1;
	for (; i < argc && strcmp(argv[i], "--"); ++i) {
		int rc = ccfd_init(argv[i], 0, err);
		printf("init %s: %d %s\n", argv[i], rc, rc ? err : "");
		if (rc == 0)
			ccfd_finalize();
	}
	printf("missing file: %d\n", ccfd_init("/nonexistent/case.ini", 0, err));
	for (++i; i < argc; ++i)
		mesh_file(argv[i]);
	return 0;
}
