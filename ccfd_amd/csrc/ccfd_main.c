/*
 * ccfd_main.c - command-line front of the drop-in: `ccfd case.ini`, as the
 * reference's binary (src/main.c:35-130), on the GPU path; `ccfd case.ini
 * file.cgns` restarts from a CGNS flow solution.
 */
#include <stdio.h>

#include "ccfd_dropin.h"

int main(int argc, char *argv[])
{
	if (argc != 2 && argc != 3) {
		printf("ERROR: Wrong number of arguments, must be 1 or 2\n");
		printf("usage: %s case.ini [restart.cgns]\n", argv[0]);
		return 1;
	}
	return ccfd_main_restart(argv[1], argc == 3 ? argv[2] : NULL, 0);
}
