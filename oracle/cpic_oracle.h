/*
 * cpic_oracle.h -- FP64 CPU restatement of cpic's per-timestep 2D electrostatic
 * PIC loop (rodarima/cpic), used ONLY to check the GPU results.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library.  The product
 * (cpic_amd/) never links, loads or calls it; the HIP path fails loudly when
 * its own library is missing instead of falling back here.
 *
 * Every function states the reference file:line it restates.  Floating-point
 * contraction policy: the library is built with -ffp-contract=off and the
 * multiply-adds that the reference's clang build (-ffp-contract=on, FMA
 * target) contracts inside one statement are written as explicit fma():
 *   relative_position_grid  grid_delta = block_delta - block_start*dx
 *                           (src/interpolate.c:54)
 *   interpolate_f2p         val += w*E (src/interpolate.c:151-153)
 *   boris_rotation          s_denom += t*t, v_minus = u + h*E,
 *                           u = v_plus + h*E (src/mover.c:46,49,68)
 *   move                    r += u*dt (src/mover.c:80)
 * Nothing else is fused.
 */
#ifndef CPIC_ORACLE_H
#define CPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_SPECIES 8

enum orc_err {
	ORC_OK = 0,
	ORC_EINVAL = -1,   /* bad parameters (reference: err()+return 1 in sim_prepare) */
	ORC_ENOMEM = -2,
	ORC_EUMAX = -3,    /* |u| > umax (reference: check_velocity abort, src/mover.c:98-137) */
	ORC_ESTATE = -4,   /* call out of protocol order (iter -1 / 0 asserts) */
};

enum orc_field {
	ORC_RHO = 0,   /* charge field, rows 0..ny-1 after the ghost fold */
	ORC_PHI = 1,   /* potential, rows 0..ny-1 */
	ORC_EX = 2,    /* E_X, rows 0..ny-1 */
	ORC_EY = 3,    /* E_Y, rows 0..ny-1 */
};

typedef struct orc_params {
	int64_t nx, ny;          /* grid.points */
	double L[2];             /* simulation.space_length */
	double dt;               /* simulation.time_step */
	double e0;               /* constants.vacuum_permittivity */
	double B[3];             /* field.magnetic */
	int64_t nchunks;         /* simulation.plasma_chunks */
	int32_t nspecies;
	double q[ORC_MAX_SPECIES];   /* species[].charge */
	double m[ORC_MAX_SPECIES];   /* species[].mass */
} orc_params_t;

typedef struct orc_sim orc_sim_t;

/* sim_prepare + field_init + MFT_init (src/sim.c:88-206, src/field.c:21-161,
 * src/solver.c:208-335).  Returns NULL and sets *err on failure. */
orc_sim_t *orc_create(const orc_params_t *p, int *err);
void orc_destroy(orc_sim_t *s);

/* pset_init: global index i belongs to chunk i mod Nc (src/plasma.c:62-63,
 * 106-128).  Particles are appended to that chunk in the order given. */
int orc_add_particles(orc_sim_t *s, int species, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy);

/* sim_pre_step (src/sim.c:209-236): global exchange, rho, E.  iter -1 -> 0. */
int orc_pre_step(orc_sim_t *s);
/* sim_step (src/sim.c:482-581).  iter n -> n+1. */
int orc_step(orc_sim_t *s);
int orc_run(orc_sim_t *s, int64_t nsteps);

/* Individual stages (src/field.c:451, src/particle.c:233, src/mover.c:332,
 * src/field.c:269).  They use s->iter like the reference. */
int orc_stage_field_E(orc_sim_t *s);
int orc_stage_plasma_E(orc_sim_t *s);
int orc_stage_plasma_r(orc_sim_t *s);
int orc_stage_field_rho(orc_sim_t *s);
int orc_comm_plasma(orc_sim_t *s, int global_exchange);

int64_t orc_iter(const orc_sim_t *s);
void orc_set_iter(orc_sim_t *s, int64_t iter);

/* Particle access.  orc_count: particles of a species over all chunks.
 * orc_get_particles writes them sorted by global index; any pointer may be
 * NULL.  orc_set_particles overwrites the state of existing particles,
 * matched by global index (ids must all exist). */
int64_t orc_count(const orc_sim_t *s, int species);
int orc_get_particles(const orc_sim_t *s, int species, int64_t *id, double *x,
		double *y, double *ux, double *uy, double *Ex, double *Ey);
int orc_set_particles(orc_sim_t *s, int species, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy);
/* Per-chunk counts and ids in list order (storage layout of the reference's
 * plists), for migration checks.  out_n[Nc*nspecies]; ids may be NULL. */
int orc_chunk_counts(const orc_sim_t *s, int64_t *out_n);
int orc_chunk_ids(const orc_sim_t *s, int64_t chunk, int species, int64_t *ids);

/* Fields as dense nx*ny row-major arrays (X fastest). */
int orc_get_field(const orc_sim_t *s, int which, double *out);
int orc_set_field(orc_sim_t *s, int which, const double *in);

/* Electrostatic energy with the unit-mesh convention:
 * W_E = -(e0/2) * sum_cells rho*phi   (rho here is -q/e0 * weight). */
double orc_field_energy(const orc_sim_t *s);
/* Kinetic energy sum_s m_s/2 * sum_p (ux^2 + uy^2) of the stored velocities. */
double orc_kinetic_energy(const orc_sim_t *s);

/* ---- building blocks exposed for pinning tests ---- */

/* relative_position_grid + linear_interpolation (src/interpolate.c:12-66,
 * 78-100) for one particle: w[2][2] indexed [wx][wy], i0[2]. */
void orc_weights(double x, double y, double dx, double dy, double w[2][2],
		int64_t i0[2]);

/* boris_rotation + check-free move for one particle, full 3D form of
 * src/mover.c:23-83 with explicit B.  r[3], u[3] updated in place. */
void orc_boris(double r[3], double u[3], const double E[3], const double B[3],
		double q, double m, double dt, int set_r);

/* MFT_solve (src/solver.c:466-509) on a dense nx*ny rho, writing nx*ny phi. */
int orc_mft_solve(int64_t nx, int64_t ny, const double *rho, double *phi);

/* Complex radix-2 FFT used by the solver (sign -1 forward, +1 inverse,
 * unnormalised), n a power of two; re/im in place. */
int orc_fft(int64_t n, double *re, double *im, int sign);

/* Cell index per particle (floor of the reference's relative_position_grid)
 * for the particles given; cell = iy*nx + ix. */
void orc_cells(const orc_sim_t *s, int64_t n, const double *x, const double *y,
		int64_t *cell);

const char *orc_strerror(int err);

#ifdef __cplusplus
}
#endif

#endif /* CPIC_ORACLE_H */
