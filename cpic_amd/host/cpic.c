/*
 * cpic.c -- command line entry: `cpic [-q] [-d] <file.conf>` (src/cpic.c:51-191).
 * Reads the configuration, runs sim_init and sim_run on the GPU path.
 * MPI/TAMPI initialisation is not needed on one GPU; -d only prints the pid
 * (the reference waits for a debugger to attach).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>

#include "conf.h"
#include "sim.h"

static void
usage(const char *argv0)
{
	fprintf(stderr, "Usage: %s [-q] [-d] <file.conf>\n", argv0);
}

int
main(int argc, char *argv[])
{
	int opt, quiet = 0, debug = 0, ret;
	cfg_t conf;
	sim_t *sim;

	while((opt = getopt(argc, argv, "qd")) != -1)
	{
		switch(opt)
		{
			case 'q': quiet = 1; break;
			case 'd': debug = 1; break;
			default: usage(argv[0]); return 1;
		}
	}
	if(optind >= argc)
	{
		usage(argv[0]);
		return 1;
	}
	if(debug)
		fprintf(stderr, "pid %d\n", (int) getpid());
	if(cfg_read_file(&conf, argv[optind]))
	{
		fprintf(stderr, "Configuration read failed: %s:%d: %s\n", argv[optind],
				conf.err_line, conf.err);
		return 1;
	}
	sim = sim_init(&conf, quiet);
	if(!sim)
	{
		fprintf(stderr, "sim_init failed\n");
		cfg_destroy(&conf);
		return 1;
	}
	ret = sim_run(sim);
	sim_free(sim);
	cfg_destroy(&conf);
	return ret ? 1 : 0;
}
