/*
 * cpic_hip.h -- C ABI of the MI355X (gfx950) implementation of cpic's
 * per-timestep 2D electrostatic PIC loop.
 *
 * Plain C types only: no HIP, torch or C++ types cross this boundary.  All
 * device memory is owned by an opaque cpic_dev_t; the host exchanges data
 * only through the upload/download calls, which take host pointers.
 *
 * Each entry point names the reference (rodarima/cpic) interface it
 * replaces.  The four stage calls keep the reference's step order and its
 * iteration protocol (iter 0 pushes velocities back half a step without
 * moving particles, src/mover.c:204-214).
 *
 * Errors: every call returns 0 or a negative CPIC_E* code (the reference
 * aborts instead, src/log.h:69-70).  Kernel-side conditions (|u| > umax,
 * non-finite values) are latched on the device and returned by the next
 * call that synchronises (cpic_dev_sync, downloads, cpic_dev_energy).  A
 * particle re-sort that fails part way (CPIC_ENOMEM or CPIC_EHIP from an
 * upload's commit, cpic_dev_rebuild or cpic_dev_sync) leaves the handle
 * unusable: every later call on it but cpic_dev_destroy returns
 * CPIC_ESTATE.
 */
#ifndef CPIC_HIP_H
#define CPIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPIC_MAX_SPECIES 8

#define CPIC_OK       0
#define CPIC_EINVAL  -1   /* bad argument or unsupported configuration */
#define CPIC_ENOMEM  -2   /* device or host allocation failed */
#define CPIC_EUMAX   -3   /* |u| > umax (reference: check_velocity abort) */
#define CPIC_ESTATE  -4   /* call out of protocol order */
#define CPIC_EHIP    -5   /* HIP runtime error */
#define CPIC_ENODEV  -6   /* no HIP device */
#define CPIC_ENAN    -7   /* non-finite particle state detected */
#define CPIC_ESIZE   -8   /* input larger than the fixed workspace of an
                             allocation-free call */

/* desc.flags */
#define CPIC_F_IDS         0x1   /* carry the particle global index (ppack.i, src/def.h:93) */
#define CPIC_F_PARTICLE_E  0x2   /* keep the gathered E per particle (ppack.E, src/def.h:96) */
#define CPIC_F_PROFILE     0x4   /* time every kernel with HIP events */
#define CPIC_F_NO_DIAG     0x8   /* skip the energy/momentum diagnostics */
#define CPIC_F_NO_GRAPH    0x10  /* cpic_dev_run launches every kernel itself
                                    instead of replaying captured steps */

/* fields for cpic_dev_download_field / cpic_dev_upload_field */
#define CPIC_RHO 0
#define CPIC_PHI 1
#define CPIC_EX  2
#define CPIC_EY  3

typedef struct cpic_dev cpic_dev_t;

/* Simulation constants: the keys sim_read_config reads (src/sim.c:46-80)
 * plus the device layout choices. */
typedef struct cpic_dev_desc {
	int64_t nx, ny;            /* grid.points; powers of two, 2..8192 */
	double L[2];               /* simulation.space_length */
	double dt;                 /* simulation.time_step */
	double e0;                 /* constants.vacuum_permittivity */
	double B[3];               /* field.magnetic; only B[2] may be non-zero */
	int32_t nspecies;          /* length of the species list */
	int32_t precision;         /* 64: FP64 particles; 32: FP32 tile-relative particles */
	double q[CPIC_MAX_SPECIES];/* species[].charge */
	double m[CPIC_MAX_SPECIES];/* species[].mass */
	int64_t nchunks;           /* simulation.plasma_chunks: sets umax_x = (nx/nchunks)*dx/dt */
	int32_t tile[2];           /* tile edge in cells (powers of two); 0 = default */
	int32_t flags;             /* CPIC_F_* */
	double slack;              /* spare bucket capacity fraction; <= 0 = default */
	int32_t device;            /* HIP device ordinal */
	int32_t ppt;               /* particles per thread and chunk in the push: 2 or 4; 0 = 2 */
} cpic_dev_desc_t;

/* Output of cpic_dev_download_particles.  n is the capacity of every
 * non-NULL array on input and the particle count on output. */
typedef struct cpic_particles {
	int64_t n;
	int64_t *id;               /* needs CPIC_F_IDS */
	double *x, *y;             /* absolute positions */
	double *ux, *uy;           /* velocities (u_z is identically 0) */
	double *Ex, *Ey;           /* gathered E, needs CPIC_F_PARTICLE_E */
	int64_t *cell;             /* iy*nx + ix of the particle's cell */
} cpic_particles_t;

typedef struct cpic_dev_stats {
	int64_t nparticles;        /* over all species */
	int64_t capacity;          /* particle slots allocated */
	int64_t nbuckets;          /* tiles * species */
	int64_t leavers;           /* particles that changed tile in the last plasma_r */
	int64_t fallbacks;         /* leavers kept in their source tile (destination full) */
	int64_t misplaced;         /* particles outside their tile after the last plasma_r */
	int64_t rebuilds;          /* full re-sorts so far */
	int32_t tile[2];
	int32_t precision;
} cpic_dev_stats_t;

/* Returns a static description of a CPIC_E* code. */
const char *cpic_dev_strerror(int err);

/* sim_init's device part: field_init (src/field.c:21-161), plasma_init
 * (src/plasma.c:293), solver_init with method "MFT" (src/solver.c:552-597).
 * Validates the configuration the way sim_prepare does (src/sim.c:110-130). */
int cpic_dev_create(const cpic_dev_desc_t *desc, cpic_dev_t **out);

/* solver_end + freeing everything (src/solver.c:632, src/sim.c:323). */
int cpic_dev_destroy(cpic_dev_t *dev);

/* particles_init's result for one species (src/particle.c:178-213): replaces
 * every particle of `species`.  Positions must lie in [0, L]. */
int cpic_dev_upload_particles(cpic_dev_t *dev, int32_t species, int64_t n,
		const int64_t *id, const double *x, const double *y,
		const double *ux, const double *uy);

/* Reads back one species (order: tile-major storage order). */
int cpic_dev_download_particles(cpic_dev_t *dev, int32_t species,
		cpic_particles_t *out);

int cpic_dev_count(cpic_dev_t *dev, int32_t species, int64_t *n);

/* sim_pre_step (src/sim.c:209-236): initial rho and E (iter -1). */
int cpic_dev_pre_step(cpic_dev_t *dev);

/* stage_field_E (src/field.c:451-501): MFT solve rho -> phi, E = -grad phi.
 * The E array itself is computed when something reads it (a download,
 * per-particle E output) or before phi is replaced; the push computes the
 * same values from phi.  An uploaded E is what the next push uses. */
int cpic_dev_field_E(cpic_dev_t *dev, int64_t iter);

/* stage_plasma_E (src/particle.c:233-248, interpolate_f2p_E
 * src/interpolate.c:350-404).  The gather is fused into cpic_dev_plasma_r;
 * this call only stores per-particle E when CPIC_F_PARTICLE_E is set. */
int cpic_dev_plasma_E(cpic_dev_t *dev, int64_t iter);

/* stage_plasma_r (src/mover.c:332-362): gather, Boris push, periodic wrap
 * and migration (comm_plasma, src/comm_plasma.c:1123-1142), plus the charge
 * deposit of the moved particles into the next rho buffer. */
int cpic_dev_plasma_r(cpic_dev_t *dev, int64_t iter);

/* stage_field_rho (src/field.c:269-356): makes the deposited rho current
 * (deposits from the stored positions if plasma_r did not). */
int cpic_dev_field_rho(cpic_dev_t *dev, int64_t iter);

/* sim_step's four stages in order (src/sim.c:482-581). */
int cpic_dev_step(cpic_dev_t *dev, int64_t iter);

/* nsteps calls of cpic_dev_step starting at iter0, queued without host
 * synchronisation (sim_run's loop, src/sim.c:594-617, without the per-step
 * host work).  Steps >= 1 are replayed two at a time from a HIP graph
 * captured on first use (not with CPIC_F_PROFILE or CPIC_F_NO_GRAPH); the
 * results are the same as those of cpic_dev_step. */
int cpic_dev_run(cpic_dev_t *dev, int64_t iter0, int64_t nsteps);

/* Waits for queued work and returns any latched device error. */
int cpic_dev_sync(cpic_dev_t *dev);

/* solve_xy with the MFT method (src/solver.c:599-621, MFT_solve :466-509):
 * phi from the current rho, without the E stencil. */
int cpic_dev_solve_xy(cpic_dev_t *dev);

/* Allocation-free variant of cpic_dev_solve_xy.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  phi = the MFT solve of rho for one nx x ny grid (dense,
 * row-major, X fastest; powers of two), with rho and phi device pointers of
 * the current device, queued on `stream` (a hipStream_t; NULL = the null
 * stream); returns once queued.  It calls no allocator: the spectrum and the
 * tables live in a static workspace of CPIC_NOALLOC_WORKSPACE_BYTES that
 * each GPU holds until the process ends (enough for 2048 x 2048; grids that
 * need more return CPIC_ESIZE).  All calls on one GPU share that workspace:
 * calls that overlap on one GPU (other streams, other host threads) race. */
#define CPIC_NOALLOC_WORKSPACE_BYTES 34078720   /* 32.5 MiB */
int cpic_solve_xy_noalloc(int64_t nx, int64_t ny, const double *rho, double *phi,
		void *stream);

/* solver_rho_size (src/solver.c:382-433) for one rank: the FFTW-padded rho
 * shape 2*(nx/2+1) x ny, in doubles.  Returns the element count. */
int64_t cpic_solver_rho_size(int64_t nx, int64_t ny, int64_t *cnx, int64_t *cny);

/* Dense nx*ny row-major (X fastest) copies of rho, phi, E_X, E_Y. */
int cpic_dev_download_field(cpic_dev_t *dev, int32_t which, double *out);
int cpic_dev_upload_field(cpic_dev_t *dev, int32_t which, const double *in);

/* Per-step diagnostics since creation: field energy -(e0/2) sum rho*phi
 * from each solve, time-centred kinetic energy and momentum from each push.
 * Writes up to `max` entries per array (any may be NULL); *n = steps. */
int cpic_dev_energy(cpic_dev_t *dev, int64_t max, double *field,
		double *kinetic, double *px, double *py, int64_t *n);

/* Migration record of the last plasma_r (needs CPIC_F_IDS): particles that
 * left their tile, as (global index, species, source tile, destination
 * tile).  out arrays sized by *n on input; *n = count on output. */
int cpic_dev_migration(cpic_dev_t *dev, int64_t *n, int64_t *id,
		int32_t *species, int32_t *src_tile, int32_t *dst_tile);

/* Particles per grid cell for one species (nx*ny counts). */
int cpic_dev_cell_counts(cpic_dev_t *dev, int32_t species, int64_t *counts);

/* Full re-sort of the particle storage into tiles with fresh capacities. */
int cpic_dev_rebuild(cpic_dev_t *dev);

int cpic_dev_stats(cpic_dev_t *dev, cpic_dev_stats_t *st);

/* Kernel timings recorded with CPIC_F_PROFILE: for kernel k < *n, its name,
 * launch count and total milliseconds (HIP events on the library's stream). */
int cpic_dev_kernel_times(cpic_dev_t *dev, int32_t *n, const char **names,
		int64_t *launches, double *total_ms);
int cpic_dev_reset_kernel_times(cpic_dev_t *dev);

/* Turns kernel timing (CPIC_F_PROFILE) on or off for later calls, like the
 * reference's per-stage perf timers (src/perf.c, src/sim.c:287-293,
 * 492-563).  Timed steps are launched one kernel at a time (no graph
 * replay); times recorded so far are kept. */
int cpic_dev_set_profiling(cpic_dev_t *dev, int32_t on);

/*
 * Batches.  One cpic_dev_t may hold P independent plasmas (an ensemble of
 * runs, e.g. a parameter scan of the reference's configurations): each has
 * its own grid, species, constants and particles, and every stage call runs
 * all of them with one launch per kernel.  The reference runs one plasma per
 * process; these calls have no counterpart there and extend the per-plasma
 * calls above with a plasma index.  cpic_dev_create is a batch of one, and
 * the per-plasma cpic_dev_* calls address plasma 0.  The stage calls
 * (pre_step, field_E, plasma_E, plasma_r, field_rho, step, run, solve_xy,
 * sync, rebuild) and cpic_dev_migration / stats / kernel_times act on the
 * whole batch; migration reports batch-wide tile indices (plasma p's tiles
 * follow those of plasmas 0..p-1).
 */

/* sim_init's device part for nplasmas configurations at once.  They must
 * share precision, flags, device, slack and ppt. */
int cpic_batch_create(const cpic_dev_desc_t *descs, int32_t nplasmas, cpic_dev_t **out);
int cpic_batch_size(cpic_dev_t *dev, int32_t *nplasmas);
int cpic_batch_upload_particles(cpic_dev_t *dev, int32_t plasma, int32_t species,
		int64_t n, const int64_t *id, const double *x, const double *y,
		const double *ux, const double *uy);
int cpic_batch_download_particles(cpic_dev_t *dev, int32_t plasma, int32_t species,
		cpic_particles_t *out);
int cpic_batch_count(cpic_dev_t *dev, int32_t plasma, int32_t species, int64_t *n);
int cpic_batch_download_field(cpic_dev_t *dev, int32_t plasma, int32_t which, double *out);
int cpic_batch_upload_field(cpic_dev_t *dev, int32_t plasma, int32_t which,
		const double *in);
int cpic_batch_energy(cpic_dev_t *dev, int32_t plasma, int64_t max, double *field,
		double *kinetic, double *px, double *py, int64_t *n);
int cpic_batch_cell_counts(cpic_dev_t *dev, int32_t plasma, int32_t species,
		int64_t *counts);

/*
 * Differentiable runs (FP64).  The reference has no gradients; these calls
 * run its loop (sim_step with the iteration-0 half step) on n particles
 * that keep their index, and return reverse-mode gradients of the
 * time-integrated field energy  loss = sum_{it < nsteps} W(it),
 * W = -(e0/2) sum rho*phi  (the value cpic_dev_energy reports per step),
 * with respect to every particle's initial position and velocity.  Forward
 * states are checkpointed every `checkpoint_every` iterations and the
 * reverse sweep recomputes each segment's positions from its checkpoint;
 * the fields of every iteration (rho and E, 24 bytes per node and
 * iteration) are kept by the forward when they fit 8 GiB of device memory
 * (cpic_adj_set_field_history), else recomputed with the positions.  desc
 * must have precision 64; flags, tile and slack are ignored.
 */
typedef struct cpic_adj cpic_adj_t;

/* species[i] (host, n entries) is the species of particle i. */
int cpic_adj_create(const cpic_dev_desc_t *desc, int64_t n, const int32_t *species,
		cpic_adj_t **out);
int cpic_adj_destroy(cpic_adj_t *adj);

/* Runs iterations 0..nsteps-1 from the initial state (host arrays of n):
 * *loss = sum of W; energy (nullable, nsteps entries) = W per iteration. */
int cpic_adj_forward(cpic_adj_t *adj, const double *x, const double *y,
		const double *ux, const double *uy, int64_t nsteps,
		int32_t checkpoint_every, double *loss, double *energy);

/* Gradients of dloss * loss with respect to the initial x, y, ux, uy of the
 * last forward (host arrays of n; any may be NULL). */
int cpic_adj_backward(cpic_adj_t *adj, double dloss, double *gx, double *gy,
		double *gux, double *guy);

/* State after the last forward (host arrays of n; any may be NULL). */
int cpic_adj_state(cpic_adj_t *adj, double *x, double *y, double *ux, double *uy);

/* Device time of the last forward and of the backward that followed it, in
 * ms (HIP events on the library's stream from the end of the uploads to the
 * start of the downloads, so host staging and PCIe transfers are excluded);
 * -1 for one that has not run.  No counterpart in the reference, which has
 * no differentiable run. */
int cpic_adj_device_ms(cpic_adj_t *adj, double *forward_ms, double *backward_ms);

/* on = 0: later forwards keep no field history, and the backward after such
 * a forward recomputes each segment's rho and E (less device memory, more
 * work); on = 1 (the default): keep it when it fits.  The backward follows
 * what its forward did.  No counterpart in the reference, which has no
 * differentiable run. */
int cpic_adj_set_field_history(cpic_adj_t *adj, int on);

#ifdef __cplusplus
}
#endif

#endif /* CPIC_HIP_H */
