/*
 * ccfd_main.c - command-line front of the drop-in: `ccfd case.ini`, as the
 * reference's binary (src/main.c:35-130), on the GPU path. Restart files
 * (CGNS) are not supported on this path.
 */
#include <stdio.h>

#include "ccfd_dropin.h"

int main(int argc, char *argv[])
{
	if (argc != 2) {
		printf("usage: %s case.ini\n", argv[0]);
		return 1;
	}
	return ccfd_main(argv[1], 0);
}
