/*
 * sim.c -- the host driver: configuration, particle initialisation, the
 * iteration protocol and the stage entry points, in the reference's order
 * (src/sim.c, src/field.c:269-501, src/particle.c:232-248,
 * src/mover.c:331-362).  Each stage enqueues its GPU work through the C ABI
 * of include/cpic_hip.h; the host never holds device pointers.
 */
#define _GNU_SOURCE
#include "sim.h"
#include "output.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>
#include <time.h>

#define err(...) fprintf(stderr, __VA_ARGS__)

/* a device sync every this many steps surfaces latched kernel errors
 * (|u| > umax, NaN) close to where the reference would abort */
#define CHECK_PERIOD 64

/* ---------------------------------------------------------------- perf -- */

static double
now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double) ts.tv_sec + (double) ts.tv_nsec * 1e-9;
}

void perf_reset(perf_t *p) { p->acc = 0.0; }
void perf_start(perf_t *p) { p->t0 = now(); }
void perf_stop(perf_t *p) { p->acc += now() - p->t0; }
double perf_measure(const perf_t *p) { return p->acc; }

/* Welford (src/perf.c:87-108) */
void
perf_record(perf_t *p, double t)
{
	double m0 = p->mean, s0 = p->m2, m1, s1;
	int n = p->n + 1;
	m1 = m0 + (t - m0) / n;
	s1 = s0 + (t - m0) * (t - m1);
	p->mean = m1;
	p->m2 = s1;
	p->n = n;
}

/* src/perf.c:112-133 */
void
perf_stats(const perf_t *p, double *mean, double *std, double *sem)
{
	*mean = p->mean;
	if(p->n >= 2)
	{
		*std = sqrt(p->m2 / (p->n - 1));
		*sem = *std / sqrt(p->n);
	}
	else
	{
		*std = 0;
		*sem = 0;
	}
}

/* -------------------------------------------------------------- config -- */

#define READ_I64(path, dst) do { \
		int64_t v_; \
		if(cfg_lookup_int64(conf, path, &v_)) { \
			err("Failed to read parameter \"%s\"\n", path); \
			return 1; \
		} \
		dst = v_; \
	} while(0)

#define READ_F(path, dst) do { \
		if(cfg_lookup_float(conf, path, &(dst))) { \
			err("Failed to read parameter \"%s\"\n", path); \
			return 1; \
		} \
	} while(0)

#define READ_S(path, dst) do { \
		if(cfg_lookup_string(conf, path, &(dst))) { \
			err("Failed to read parameter \"%s\"\n", path); \
			return 1; \
		} \
	} while(0)

/* sim_read_config + sim_prepare + species_init (src/sim.c:39-206,
 * src/specie.c:15-67) */
int
sim_read_config(sim_t *s, const cfg_t *conf)
{
	const cfg_node_t *list, *n;
	int64_t tmp;
	int d, i;

	s->conf = conf;
	READ_I64("simulation.dimensions", s->dim);
	READ_I64("simulation.cycles", s->cycles);
	READ_F("simulation.time_step", s->dt);
	READ_I64("simulation.random_seed", tmp);
	s->seed = (unsigned int) tmp;
	READ_F("constants.light_speed", s->C);
	READ_F("constants.vacuum_permittivity", s->e0);
	READ_I64("simulation.sampling_period.energy", s->period_energy);
	READ_I64("simulation.sampling_period.field", s->period_field);
	READ_I64("simulation.sampling_period.particle", s->period_particle);
	READ_F("simulation.stop_SEM", s->stop_SEM);
	READ_I64("simulation.realtime_plot", s->mode);
	READ_S("simulation.solver", s->solver_method);
	READ_I64("simulation.enable_fftw_threads", s->fftw_threads);
	READ_I64("simulation.plasma_chunks", s->plasma_chunks);
	READ_I64("simulation.pblock_nmax", s->pblock_nmax);

	if(s->dim != 2)
	{
		err("Only 2 dimensions supported by now...\n");
		return 1;
	}
	if(cfg_lookup_array_float(conf, "simulation.space_length", s->L, s->dim))
	{
		err("Failed to read parameter \"simulation.space_length\"\n");
		return 1;
	}
	for(d = (int) s->dim; d < MAX_DIM; d++)
		s->L[d] = 0;
	if(cfg_lookup_array_float(conf, "field.magnetic", s->B, MAX_DIM))
	{
		err("Failed to read parameter \"field.magnetic\"\n");
		return 1;
	}
	if(cfg_lookup_array_int(conf, "grid.points", s->ntpoints, s->dim))
	{
		err("Failed to read parameter \"grid.points\"\n");
		return 1;
	}

	/* optional keys of this driver */
	s->precision = 64;
	if(cfg_lookup(conf, "simulation.precision"))
	{
		READ_I64("simulation.precision", tmp);
		s->precision = (int) tmp;
	}
	s->tile[0] = s->tile[1] = 0;
	if((n = cfg_lookup(conf, "simulation.tile")))
	{
		int64_t t[2];
		if(cfg_lookup_array_int(conf, "simulation.tile", t, 2))
		{
			err("Failed to read parameter \"simulation.tile\"\n");
			return 1;
		}
		s->tile[0] = (int) t[0];
		s->tile[1] = (int) t[1];
	}
	s->device = 0;
	if(cfg_lookup(conf, "simulation.device"))
	{
		READ_I64("simulation.device", tmp);
		s->device = (int) tmp;
	}

	/* sim_prepare checks (src/sim.c:110-130) */
	if(s->plasma_chunks <= 0 || s->ntpoints[X] % s->plasma_chunks)
	{
		err("The number of grid points in X %ld cannot be divided by the number of plasma chunks %ld\n",
				(long) s->ntpoints[X], (long) s->plasma_chunks);
		return 1;
	}
	for(d = (int) s->dim; d < MAX_DIM; d++)
	{
		s->ntpoints[d] = 1;
		s->dx[d] = 0.0;
		s->chunksize[d] = 1;
	}
	for(d = 0; d < s->dim; d++)
		s->dx[d] = s->L[d] / s->ntpoints[d];
	s->chunksize[X] = s->ntpoints[X] / s->plasma_chunks;
	s->chunksize[Y] = s->ntpoints[Y];
	for(d = 0; d < MAX_DIM; d++)
		s->umax[d] = s->chunksize[d] * s->dx[d] / s->dt;

	if(strcmp(s->solver_method, "MFT") != 0)
	{
		err("Solver method \"%s\" is not available: the GPU path implements \"MFT\"\n",
				s->solver_method);
		return 1;
	}

	list = cfg_lookup(conf, "species");
	s->nspecies = cfg_length(list);
	if(s->nspecies <= 0)
	{
		err("The number of species must be at least 1.\n");
		return 1;
	}
	if(s->nspecies > CPIC_MAX_SPECIES)
	{
		err("At most %d species are supported\n", CPIC_MAX_SPECIES);
		return 1;
	}
	for(i = 0; i < s->nspecies; i++)
	{
		specie_t *sp = &s->species[i];
		const cfg_node_t *cs = cfg_elem(list, i);
		sp->conf = cs;
		sp->name = "";
		cfg_get_string(cfg_member(cs, "name"), &sp->name);
		cfg_get_float(cfg_member(cs, "charge"), &sp->q);
		cfg_get_float(cfg_member(cs, "mass"), &sp->m);
		sp->nparticles = 0;
		cfg_get_int64(cfg_member(cs, "particles"), &sp->nparticles);
		if(sp->nparticles <= 0)
		{
			err("The number of particles must be greater than 0\n");
			return 1;
		}
	}
	s->iter = -2;
	s->running = 1;
	s->sampling = s->stop_SEM > 0.0;
	return 0;
}

/* ------------------------------------------------------------ particles -- */

/* uniform (src/particle.c:17-21); the multiply-add is contracted like the
 * reference's clang build */
static double
uniform(double a, double b)
{
	return fma(rand() / (RAND_MAX + 1.0), b - a, a);
}

/* SplitMix64: counter-based stream for the "two stream" method */
static uint64_t
splitmix64(uint64_t *state)
{
	uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static double
u01(uint64_t *st)   /* [0, 1) with 53 random bits */
{
	return (double) (splitmix64(st) >> 11) * 0x1.0p-53;
}

/* "two stream" (new init method; no reference counterpart): positions
 * uniform in [0, L), u_x = +-v_x (sign alternating with the global index
 * when counter_stream is true) + sigma_x N(0,1), u_y = v_y + sigma_y N(0,1).
 * The stream of particle i of species s starts from
 * seed + s * 2^32 + i, so the state does not depend on the chunking. */
static void
two_stream_one(const sim_t *s, int sp, int64_t i, const double v[2], const double sig[2],
		int counter, double *x, double *y, double *ux, double *uy)
{
	uint64_t st = (uint64_t) s->seed + ((uint64_t) sp << 32) + (uint64_t) i;
	double a, b, r, n0, n1, sign;
	/* decorrelate neighbouring counters */
	st = splitmix64(&st);
	*x = u01(&st) * s->L[X];
	*y = u01(&st) * s->L[Y];
	if(*x >= s->L[X]) *x = 0.0;
	if(*y >= s->L[Y]) *y = 0.0;
	a = 1.0 - u01(&st);          /* (0, 1] */
	b = u01(&st);
	r = sqrt(-2.0 * log(a));
	n0 = r * cos(2.0 * M_PI * b);
	n1 = r * sin(2.0 * M_PI * b);
	sign = (counter && (i & 1)) ? -1.0 : 1.0;
	*ux = sign * v[0] + sig[0] * n0;
	*uy = v[1] + sig[1] * n1;
}

static int
read_vec(const cfg_node_t *cs, const char *key, double *v, int n)
{
	const cfg_node_t *m = cfg_member(cs, key);
	if(!m)
	{
		err("Species setting \"%s\" missing\n", key);
		return 1;
	}
	if(m->type == CFG_FLOAT || m->type == CFG_INT)
	{
		int k;
		double t;
		cfg_get_float(m, &t);
		for(k = 0; k < n; k++)
			v[k] = t;
		return 0;
	}
	return cfg_array_float(m, v, n);
}

int
particles_generate(const sim_t *s, double **x, double **y, double **ux, double **uy)
{
	int sp;
	int64_t c, i;
	const char *method[CPIC_MAX_SPECIES];
	double v[CPIC_MAX_SPECIES][2], r0[CPIC_MAX_SPECIES][2], dr[CPIC_MAX_SPECIES][2];
	double sig[CPIC_MAX_SPECIES][2];
	int counter[CPIC_MAX_SPECIES];

	for(sp = 0; sp < s->nspecies; sp++)
	{
		const cfg_node_t *cs = s->species[sp].conf;
		int64_t n = s->species[sp].nparticles;
		x[sp] = malloc(sizeof(double) * (size_t) n);
		y[sp] = malloc(sizeof(double) * (size_t) n);
		ux[sp] = malloc(sizeof(double) * (size_t) n);
		uy[sp] = malloc(sizeof(double) * (size_t) n);
		if(!x[sp] || !y[sp] || !ux[sp] || !uy[sp])
		{
			err("out of memory generating %ld particles\n", (long) n);
			return 1;
		}
		if(cfg_get_string(cfg_member(cs, "init_method"), &method[sp]))
		{
			err("Particle init method for specie \"%s\" not specified.\n",
					s->species[sp].name);
			return 1;
		}
		if(strcmp(method[sp], "random position") == 0)
		{
			if(read_vec(cs, "drift_velocity", v[sp], 2))
				return 1;
		}
		else if(strcmp(method[sp], "position delta") == 0)
		{
			if(read_vec(cs, "drift_velocity", v[sp], 2) ||
					read_vec(cs, "position_delta", dr[sp], 2) ||
					read_vec(cs, "position_init", r0[sp], 2))
				return 1;
		}
		else if(strcmp(method[sp], "two stream") == 0)
		{
			int cflag = 1;
			if(read_vec(cs, "drift_velocity", v[sp], 2) ||
					read_vec(cs, "thermal_velocity", sig[sp], 2))
				return 1;
			cfg_get_bool(cfg_member(cs, "counter_stream"), &cflag);
			counter[sp] = cflag;
		}
		else
		{
			err("Unknown init method \"%s\", aborting.\n", method[sp]);
			return 1;
		}
	}

	/* srand(seed + rank) (src/sim.c:153), then chunk -> species -> global
	 * index order (src/plasma.c:307-314, 280-287, 62-63) */
	srand(s->seed + 0u);
	for(c = 0; c < s->plasma_chunks; c++)
	{
		for(sp = 0; sp < s->nspecies; sp++)
		{
			int64_t n = s->species[sp].nparticles;
			if(strcmp(method[sp], "random position") == 0)
			{
				/* init_randpos (src/particle.c:24-88) */
				for(i = c; i < n; i += s->plasma_chunks)
				{
					x[sp][i] = uniform(0.0, s->L[X]);
					y[sp][i] = uniform(0.0, s->L[Y]);
					ux[sp][i] = uniform(-v[sp][X], v[sp][X]);
					uy[sp][i] = uniform(-v[sp][Y], v[sp][Y]);
				}
			}
			else if(strcmp(method[sp], "position delta") == 0)
			{
				/* init_position_delta (src/particle.c:91-168): WRAP of
				 * r0 + dr*i (src/mat.h:52) */
				for(i = c; i < n; i += s->plasma_chunks)
				{
					double r;
					r = fmod(fma(dr[sp][X], (double) i, r0[sp][X]), s->L[X]);
					if(r < 0.0) r += s->L[X];
					x[sp][i] = r;
					r = fmod(fma(dr[sp][Y], (double) i, r0[sp][Y]), s->L[Y]);
					if(r < 0.0) r += s->L[Y];
					y[sp][i] = r;
					ux[sp][i] = v[sp][X];
					uy[sp][i] = v[sp][Y];
				}
			}
			else
			{
				for(i = c; i < n; i += s->plasma_chunks)
					two_stream_one(s, sp, i, v[sp], sig[sp], counter[sp], &x[sp][i],
							&y[sp][i], &ux[sp][i], &uy[sp][i]);
			}
		}
	}
	return 0;
}

/* -------------------------------------------------------------- stages -- */

static void
stage_error(sim_t *s, const char *stage, int ret)
{
	if(!s->error)
		s->error = ret;
	err("%s failed: %s\n", stage, cpic_dev_strerror(ret));
}

int
stage_field_E(sim_t *s)
{
	int ret;
	perf_start(&s->timers[TIMER_FIELD_E]);
	if((ret = cpic_dev_field_E(s->dev, s->iter)))
		stage_error(s, "stage_field_E", ret);
	perf_stop(&s->timers[TIMER_FIELD_E]);
	return ret;
}

void
stage_plasma_E(sim_t *s)
{
	int ret;
	perf_start(&s->timers[TIMER_PARTICLE_E]);
	if((ret = cpic_dev_plasma_E(s->dev, s->iter)))
		stage_error(s, "stage_plasma_E", ret);
	perf_stop(&s->timers[TIMER_PARTICLE_E]);
}

void
stage_plasma_r(sim_t *s)
{
	int ret;
	perf_start(&s->timers[TIMER_PARTICLE_X]);
	if((ret = cpic_dev_plasma_r(s->dev, s->iter)))
		stage_error(s, "stage_plasma_r", ret);
	perf_stop(&s->timers[TIMER_PARTICLE_X]);
}

void
stage_field_rho(sim_t *s)
{
	int ret;
	perf_start(&s->timers[TIMER_FIELD_RHO]);
	if((ret = cpic_dev_field_rho(s->dev, s->iter)))
		stage_error(s, "stage_field_rho", ret);
	perf_stop(&s->timers[TIMER_FIELD_RHO]);
}

/* -------------------------------------------------------------- driver -- */

static int
sim_pre_step(sim_t *s)
{
	int ret;
	if(!s->quiet)
		err("begin sim_pre_step\n");
	/* particle_comm_initial: placing every particle in its tile happened
	 * when the particles were uploaded */
	stage_field_rho(s);
	if(s->error)
		return 1;
	/* the solver plans of the reference are created by this first solve */
	if(stage_field_E(s))
		return 1;
	if((ret = cpic_dev_sync(s->dev)))
	{
		stage_error(s, "sim_pre_step", ret);
		return 1;
	}
	if(!s->quiet)
		err("end sim_pre_step\n");
	return 0;
}

sim_t *
sim_init(const cfg_t *conf, int quiet)
{
	sim_t *s;
	cpic_dev_desc_t desc;
	double *x[CPIC_MAX_SPECIES] = {0}, *y[CPIC_MAX_SPECIES] = {0};
	double *ux[CPIC_MAX_SPECIES] = {0}, *uy[CPIC_MAX_SPECIES] = {0};
	int64_t *ids = NULL;
	int sp, ret, ok = 0, track = 0;
	int64_t i, nmax = 0;

	s = calloc(1, sizeof(*s));
	if(!s)
		return NULL;
	s->quiet = quiet;
	if(sim_read_config(s, conf))
		goto out;
	cfg_get_bool(cfg_lookup(conf, "simulation.track_particles"), &track);
	/* output_init (src/sim.c:268-273) */
	s->output = calloc(1, sizeof(*s->output));
	if(!s->output || output_init(s, s->output))
	{
		err("output_init failed\n");
		goto out;
	}

	memset(&desc, 0, sizeof(desc));
	desc.nx = s->ntpoints[X];
	desc.ny = s->ntpoints[Y];
	desc.L[0] = s->L[X];
	desc.L[1] = s->L[Y];
	desc.dt = s->dt;
	desc.e0 = s->e0;
	for(i = 0; i < 3; i++)
		desc.B[i] = s->B[i];
	desc.nspecies = s->nspecies;
	desc.precision = s->precision;
	for(sp = 0; sp < s->nspecies; sp++)
	{
		desc.q[sp] = s->species[sp].q;
		desc.m[sp] = s->species[sp].m;
	}
	desc.nchunks = s->plasma_chunks;
	desc.tile[0] = s->tile[0];
	desc.tile[1] = s->tile[1];
	desc.flags = track ? (CPIC_F_IDS | CPIC_F_PARTICLE_E) : 0;
	desc.device = s->device;
	/* field_init + plasma_init + solver_init (src/sim.c:275-293) */
	if((ret = cpic_dev_create(&desc, &s->dev)))
	{
		err("device setup failed: %s\n", cpic_dev_strerror(ret));
		goto out;
	}
	s->iter++;

	if(particles_generate(s, x, y, ux, uy))
		goto out;
	for(sp = 0; sp < s->nspecies; sp++)
		if(s->species[sp].nparticles > nmax)
			nmax = s->species[sp].nparticles;
	if(track)
	{
		ids = malloc(sizeof(int64_t) * (size_t) nmax);
		if(!ids)
			goto out;
		for(i = 0; i < nmax; i++)
			ids[i] = i;
	}
	for(sp = 0; sp < s->nspecies; sp++)
	{
		if((ret = cpic_dev_upload_particles(s->dev, sp, s->species[sp].nparticles, ids,
						x[sp], y[sp], ux[sp], uy[sp])))
		{
			err("particle upload failed: %s\n", cpic_dev_strerror(ret));
			goto out;
		}
		free(x[sp]); free(y[sp]); free(ux[sp]); free(uy[sp]);
		x[sp] = y[sp] = ux[sp] = uy[sp] = NULL;
	}

	s->iter = -1;
	s->t = s->iter * s->dt;
	if(sim_pre_step(s))
		goto out;
	s->iter = 0;
	s->t = 0.0;
	ok = 1;
out:
	free(ids);
	for(sp = 0; sp < CPIC_MAX_SPECIES; sp++)
	{
		free(x[sp]); free(y[sp]); free(ux[sp]); free(uy[sp]);
	}
	if(!ok)
	{
		sim_free(s);
		return NULL;
	}
	return s;
}

static long
memory_usage(void)
{
	struct rusage usage;
	if(getrusage(RUSAGE_SELF, &usage))
		return -1;
	return usage.ru_maxrss;
}

/* sampling_complete (src/sim.c:441-479) */
static int
sampling_complete(sim_t *s)
{
	double t, t_solver, mean, std, sem, rsem, mean_solver, std_solver, sem_solver;

	t = perf_measure(&s->timers[TIMER_ITERATION]);
	t_solver = perf_measure(&s->timers[TIMER_SOLVER]);
	perf_stats(&s->timers[TIMER_ITERATION], &mean, &std, &sem);
	perf_record(&s->timers[TIMER_ITERATION], t);
	perf_stats(&s->timers[TIMER_SOLVER], &mean_solver, &std_solver, &sem_solver);
	perf_record(&s->timers[TIMER_SOLVER], t_solver);
	rsem = mean != 0.0 ? sem / mean : sem;
	printf("stats iter=%ld last=%e mean=%e std=%e sem=%e rsem=%e mem=%ld solver=%e\n",
			(long) s->iter, t, mean, std, sem, rsem, memory_usage(), mean_solver);
	if(s->iter < 30)
		return 0;
	if(t > mean + std * 5.0)
		printf("Iteration time exceeded 5 sigma! Go fix your program.\n");
	return 1.96 * sem < s->stop_SEM * mean;
}

/* sim_step (src/sim.c:482-581) */
int
sim_step(sim_t *s)
{
	int ret;

	if(s->iter < 0)
		return -1;
	if(s->iter >= s->cycles)
		return -1;

	perf_reset(&s->timers[TIMER_ITERATION]);
	perf_start(&s->timers[TIMER_ITERATION]);

	/* the solver time is measured separately only when sampling, as it
	 * needs the device to drain */
	if(s->sampling)
	{
		perf_reset(&s->timers[TIMER_SOLVER]);
		perf_start(&s->timers[TIMER_SOLVER]);
	}
	stage_field_E(s);
	if(s->sampling)
	{
		cpic_dev_sync(s->dev);
		perf_stop(&s->timers[TIMER_SOLVER]);
	}
	if(output_fields(s))
	{
		err("output_fields failed\n");
		return -1;
	}
	stage_plasma_E(s);
	stage_plasma_r(s);
	stage_field_rho(s);

	if(s->sampling || (s->iter % CHECK_PERIOD) == CHECK_PERIOD - 1)
		if((ret = cpic_dev_sync(s->dev)))
			stage_error(s, "sim_step", ret);
	perf_stop(&s->timers[TIMER_ITERATION]);
	if(s->error)
		return -1;
	if(s->sampling && sampling_complete(s))
	{
		printf("sampling complete\n");
		s->running = 0;
	}
	s->iter += 1;
	s->t = s->iter * s->dt;
	return 0;
}

/* sim_run (src/sim.c:622-654) */
int
sim_run(sim_t *s)
{
	int ret;
	if(s->iter != 0)
		return -1;
	perf_start(&s->timers[TIMER_TOTAL]);
	printf("Simulation runs now\n");
	while(s->running && s->iter < s->cycles)
	{
		if(sim_step(s))
		{
			err("sim_step failed\n");
			return -1;
		}
	}
	s->running = 0;
	if((ret = cpic_dev_sync(s->dev)))
	{
		stage_error(s, "sim_run", ret);
		return -1;
	}
	perf_stop(&s->timers[TIMER_TOTAL]);
	return 0;
}

void
sim_free(sim_t *s)
{
	if(!s)
		return;
	if(s->dev)
		cpic_dev_destroy(s->dev);
	output_free(s->output);
	free(s->output);
	free(s);
}

/* ------------------------------------------------------------ bindings -- */

sim_t *
sim_load(const char *path, int quiet)
{
	cfg_t *c = calloc(1, sizeof(*c));
	sim_t *s;
	if(!c)
		return NULL;
	if(cfg_read_file(c, path))
	{
		err("Configuration read failed: %s:%d: %s\n", path, c->err_line, c->err);
		free(c);
		return NULL;
	}
	s = sim_init(c, quiet);
	if(!s)
	{
		cfg_destroy(c);
		free(c);
	}
	/* the configuration stays allocated with the simulation (the
	 * reference keeps config_t alive as well) */
	return s;
}

int
sim_particle(sim_t *s, int sp, int64_t index, double out[6])
{
	cpic_particles_t p;
	int64_t n = 0, k;
	int ret = -1;
	if(cpic_dev_count(s->dev, sp, &n))
		return -1;
	memset(&p, 0, sizeof(p));
	p.n = n;
	p.id = malloc(sizeof(int64_t) * (size_t) (n ? n : 1));
	p.x = malloc(sizeof(double) * (size_t) (n ? n : 1));
	p.y = malloc(sizeof(double) * (size_t) (n ? n : 1));
	p.ux = malloc(sizeof(double) * (size_t) (n ? n : 1));
	p.uy = malloc(sizeof(double) * (size_t) (n ? n : 1));
	p.Ex = malloc(sizeof(double) * (size_t) (n ? n : 1));
	p.Ey = malloc(sizeof(double) * (size_t) (n ? n : 1));
	if(p.id && p.x && p.y && p.ux && p.uy && p.Ex && p.Ey &&
			cpic_dev_download_particles(s->dev, sp, &p) == 0)
	{
		for(k = 0; k < p.n; k++)
			if(p.id[k] == index)
			{
				out[0] = p.x[k]; out[1] = p.y[k];
				out[2] = p.ux[k]; out[3] = p.uy[k];
				out[4] = p.Ex[k]; out[5] = p.Ey[k];
				ret = 0;
				break;
			}
	}
	free(p.id); free(p.x); free(p.y); free(p.ux); free(p.uy); free(p.Ex); free(p.Ey);
	return ret;
}

int
host_generate(const char *path, int sp, int64_t n, double *x, double *y, double *ux, double *uy)
{
	cfg_t c;
	sim_t *s;
	double *ax[CPIC_MAX_SPECIES] = {0}, *ay[CPIC_MAX_SPECIES] = {0};
	double *aux[CPIC_MAX_SPECIES] = {0}, *auy[CPIC_MAX_SPECIES] = {0};
	int ret = -1, k;

	if(cfg_read_file(&c, path))
	{
		err("Configuration read failed: %s:%d: %s\n", path, c.err_line, c.err);
		return -1;
	}
	s = calloc(1, sizeof(*s));
	if(s && sim_read_config(s, &c) == 0 && sp >= 0 && sp < s->nspecies &&
			s->species[sp].nparticles == n && particles_generate(s, ax, ay, aux, auy) == 0)
	{
		memcpy(x, ax[sp], sizeof(double) * (size_t) n);
		memcpy(y, ay[sp], sizeof(double) * (size_t) n);
		memcpy(ux, aux[sp], sizeof(double) * (size_t) n);
		memcpy(uy, auy[sp], sizeof(double) * (size_t) n);
		ret = 0;
	}
	for(k = 0; k < CPIC_MAX_SPECIES; k++)
	{
		free(ax[k]); free(ay[k]); free(aux[k]); free(auy[k]);
	}
	free(s);
	cfg_destroy(&c);
	return ret;
}

int64_t sim_get_iter(const sim_t *s) { return s->iter; }
int64_t sim_get_cycles(const sim_t *s) { return s->cycles; }
