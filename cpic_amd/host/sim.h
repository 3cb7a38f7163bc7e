/*
 * sim.h -- conf-driven host driver of the PIC step, in C, mirroring the
 * reference's driver (src/sim.h, src/sim.c) and stage entry points
 * (src/field.h:9-12, src/particle.h:23-26, src/mover.h:4).  The numerical
 * work of every stage runs on the GPU through include/cpic_hip.h.
 */
#ifndef CPIC_HOST_SIM_H
#define CPIC_HOST_SIM_H

#include <stdint.h>

#include "../../include/cpic_hip.h"
#include "conf.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MAX_DIM 3
enum { X = 0, Y = 1, Z = 2 };

/* Welford running statistics of one timer (src/perf.c:87-133). */
typedef struct perf {
	double t0;          /* start, seconds */
	double acc;         /* accumulated since the last reset */
	double mean, m2;
	int n;
} perf_t;

enum timers {
	TIMER_SOLVER,
	TIMER_FIELD_E,
	TIMER_FIELD_RHO,
	TIMER_PARTICLE_E,
	TIMER_PARTICLE_X,
	TIMER_TOTAL,
	TIMER_ITERATION,
	TIMER_OUTPUT_FIELDS,
	MAX_TIMERS
};

/* specie_t (src/def.h:238-252) */
typedef struct specie {
	const char *name;
	double q, m;
	int64_t nparticles;
	const cfg_node_t *conf;
} specie_t;

typedef struct sim {
	const cfg_t *conf;

	/* configuration (sim_read_config, src/sim.c:39-85) */
	int64_t dim, cycles;
	double dt;
	unsigned int seed;
	double C, e0;
	int64_t period_energy, period_field, period_particle;
	double stop_SEM;
	int64_t mode;
	const char *solver_method;
	int64_t fftw_threads, plasma_chunks, pblock_nmax;
	double L[MAX_DIM], B[MAX_DIM];
	int64_t ntpoints[MAX_DIM];
	int nspecies;
	specie_t species[CPIC_MAX_SPECIES];

	/* device layout choices (optional keys simulation.precision,
	 * simulation.tile, simulation.device) */
	int precision;
	int tile[2];
	int device;

	/* derived (sim_prepare, src/sim.c:88-206) */
	double dx[MAX_DIM];
	int64_t chunksize[MAX_DIM];
	double umax[MAX_DIM];

	/* run state */
	int64_t iter;
	int running;
	int sampling;
	double t;
	int error;          /* first CPIC_E* error of a stage, 0 if none */
	int quiet;

	perf_t timers[MAX_TIMERS];
	cpic_dev_t *dev;
	struct output *output;   /* output.c; field output (output.path) */
} sim_t;

/* sim_init (src/sim.c:239-320): reads the configuration, generates the
 * particles, creates the device state and runs sim_pre_step.  Returns NULL
 * (with a message on stderr) on failure. */
sim_t *sim_init(const cfg_t *conf, int quiet);
int sim_step(sim_t *sim);
int sim_run(sim_t *sim);
void sim_free(sim_t *sim);

/* Stages (src/field.h:9-12, src/particle.h:23, src/mover.h:4). */
int stage_field_E(sim_t *sim);
void stage_plasma_E(sim_t *sim);
void stage_plasma_r(sim_t *sim);
void stage_field_rho(sim_t *sim);

/* Configuration reading and particle generation without a device (used by
 * sim_init and by tests).  sim_read_config fills the configuration fields
 * of `sim` and validates them like sim_prepare. */
int sim_read_config(sim_t *sim, const cfg_t *conf);
/* particles_init for every species (src/particle.c:178-213, pset_init
 * src/plasma.c:20-137): x[s], y[s], ux[s], uy[s] are malloc'd arrays of
 * species[s].nparticles entries indexed by the global index. */
int particles_generate(const sim_t *sim, double **x, double **y, double **ux, double **uy);

/* Loads a configuration file and runs sim_init; for bindings. */
sim_t *sim_load(const char *path, int quiet);
/* One particle's state by species and global index (needs CPIC_F_IDS,
 * which sim_init sets when simulation.track_particles is true). */
int sim_particle(sim_t *sim, int species, int64_t index, double out[6]);
int64_t sim_get_iter(const sim_t *sim);
int64_t sim_get_cycles(const sim_t *sim);

/* Reads a configuration file and generates species sp (n particles) into
 * the given arrays without touching a device; for tests. */
int host_generate(const char *path, int sp, int64_t n, double *x, double *y,
		double *ux, double *uy);

/* Welford timers (src/perf.c) */
void perf_reset(perf_t *p);
void perf_start(perf_t *p);
void perf_stop(perf_t *p);
double perf_measure(const perf_t *p);
void perf_record(perf_t *p, double t);
void perf_stats(const perf_t *p, double *mean, double *std, double *sem);

#ifdef __cplusplus
}
#endif

#endif
