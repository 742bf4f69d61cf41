/*
 * lux-gray -- command-line launcher of the gray job, standing in for
 * `lux gray key=value ...` (README.md:21 of the reference): constructs the
 * job, applies every key=value argument through conf, then init and exec.
 */
#include "../../include/gray_job.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
	int k = 1;
	if (argc > 1 && !strcmp(argv[1], "gray"))
		++k; /* accept `lux-gray gray key=value ...` like `lux gray ...` */
	if (argc > k && (!strcmp(argv[k], "-h") || !strcmp(argv[k], "--help"))) {
		void *ego = gray_mkmod(NULL);
		printf("usage: %s [gray] key=value ...\noptions:\n", argv[0]);
		gray_print_options(ego, stdout);
		gray_rmmod(ego);
		return EXIT_SUCCESS;
	}
	return gray_run(argc - k, (const char *const *)argv + k);
}
