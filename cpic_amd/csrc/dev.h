// Internal state of the gfx950 PIC step.  Not part of the C ABI.
//
// Layout in HBM
//   fields   rho[2] (double, nx*ny, double-buffered: the solve reads one while
//            the push deposits into the other), ghat ((nx/2+1)*ny complex,
//            stored kx-major so the column pass reads contiguous columns),
//            phi (double nx*ny), E (double2 or float2 nx*ny, (Ex,Ey) pairs).
//   particles tile-major buckets (bucket b = tile*nspecies + species), SoA
//            arrays x, y, ux, uy (+ optional id, Ex, Ey).  A bucket owns
//            slots [start, start+cap): segment 1 holds n1 particles at
//            [0, n1), segment 2 holds n2 fallback particles at
//            [cap-res, cap-res+n2), where res is the number of particles
//            that left the bucket in the last push (their slots are kept free
//            so a leaver whose destination is full can always return).
//   FP64 mode stores absolute positions (the reference's r); FP32 mode
//   stores positions in cell units relative to the bucket tile's origin.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/cpic_hip.h"

#define CPIC_BLOCK 256
#define CPIC_MARGIN 1            // cells of drift covered by the LDS tile regions
#define CPIC_DIAG_RING 65536     // per-step diagnostic slots
#define CPIC_MAX_KERNELS 16

// device error word bits
#define CPIC_DERR_UMAX 0x1
#define CPIC_DERR_NAN 0x2

struct Geo {
	int nx, ny;          // grid points
	int lgnx, lgny;
	int tx, ty;          // tile edge in cells
	int lgtx, lgty;
	int ntx, nty;        // tiles per direction
	int lgntx;
	int ntiles;
	int W, H;            // LDS region: (tx + 2M + 1) x (ty + 2M + 1) nodes
	int M;
	int ns;              // species
	double dx, dy;       // cell size
	double idx, idy;     // 1/dx, 1/dy
	double L0, L1;       // simulation.space_length
};

struct Species {
	double h[CPIC_MAX_SPECIES];    // 0.5*dt'*q/m
	double tz[CPIC_MAX_SPECIES];   // B_z * h
	double sz[CPIC_MAX_SPECIES];   // 2 tz / (1 + tz^2)
	double vq[CPIC_MAX_SPECIES];   // -q/e0 (deposit constant)
	double m[CPIC_MAX_SPECIES];
};

struct Buckets {
	int64_t *start;
	int32_t *cap;
	int32_t *n1;
	int32_t *res;
	int32_t *n2;
};

struct Parts {
	void *x, *y, *ux, *uy;   // double or float
	int64_t *id;             // nullable
	double *Ex, *Ey;         // nullable
};

struct Leavers {
	void *x, *y, *ux, *uy;
	int64_t *id;
	double *Ex, *Ey;
	int32_t *dst, *src;
	int32_t *count;          // [2]: double-buffered per step parity
	int64_t cap;
};

struct Counters {            // device-side statistics, reset by the host
	unsigned long long leavers;
	unsigned long long fallbacks;
	unsigned long long misplaced;
	unsigned long long pad;
};

struct DiagSlot {
	double field;
	double kinetic;
	double px, py;
};

struct KernelTimer {
	const char *name;
	int64_t launches;
	double total_ms;
};

struct cpic_dev {
	cpic_dev_desc_t desc;
	Geo g;
	int prec;                // 32 or 64
	int ns;
	double dt;
	double umax[2];
	int device;
	hipStream_t stream;

	// fields
	double *rho[2];
	int rho_cur;
	int rho_next_ready;      // rho[1-cur] has been zeroed
	int rho_pending;         // plasma_r deposited into rho[1-cur]
	double2 *ghat;
	double *phi;
	void *E;                 // double2 (FP64) or float2 (FP32)
	double *costab;          // cos(2 pi k/nx) k<=nx/2, then cos(2 pi k/ny) k<ny
	double2 *tw;             // exp(-2 pi i k/N), k <= N/2, N = max(nx, ny)
	int lgN;

	// particles
	Parts p[2];              // [cur] in use, [1-cur] spare for rebuilds
	int pcur;
	int64_t pcap[2];
	Buckets bk;
	int nb;
	Leavers lv;
	int step_parity;
	int64_t count[CPIC_MAX_SPECIES];

	// pending uploads (host copies, applied by the next rebuild)
	int pending;
	int64_t up_n[CPIC_MAX_SPECIES];
	double *up_buf[CPIC_MAX_SPECIES];   // x,y,ux,uy interleaved as 4 arrays
	int64_t *up_id[CPIC_MAX_SPECIES];

	// diagnostics and errors
	DiagSlot *diag;          // device ring
	int *derr;               // device error word
	Counters *ctr;           // device counters
	int64_t last_iter;
	int64_t nrebuilds;
	int pushed_since_rebuild;

	// profiling
	KernelTimer kt[CPIC_MAX_KERNELS];
	int nkt;
	hipEvent_t *ev;          // pairs
	int *ev_kernel;
	int nev, evcap;
};

// helpers implemented in the .hip files
int cpic_fft_init(cpic_dev *d);
int cpic_solve(cpic_dev *d, int64_t iter, int want_energy);
int cpic_efield(cpic_dev *d, int64_t iter);
int cpic_parts_rebuild(cpic_dev *d);
int cpic_parts_push(cpic_dev *d, int64_t iter);
int cpic_parts_deposit(cpic_dev *d, double *rho_out);
int cpic_parts_gather(cpic_dev *d);
int cpic_parts_download(cpic_dev *d, int species, cpic_particles_t *out);
int cpic_parts_cell_counts(cpic_dev *d, int species, int64_t *counts);
int cpic_parts_alloc(cpic_dev *d, int which, int64_t cap);
void cpic_parts_free(cpic_dev *d, int which);
int cpic_zero(cpic_dev *d, double *p, int64_t n);
int cpic_parts_init(cpic_dev *d);

// profiling wrappers (api.hip)
int cpic_kt_begin(cpic_dev *d, const char *name);
void cpic_kt_end(cpic_dev *d, int slot);

// Every failing HIP call is named on stderr (the reference reports fatal
// conditions with err(), src/log.h:69) and returned as CPIC_EHIP.
#define HIPCHK(call) do { \
		hipError_t e_ = (call); \
		if(e_ != hipSuccess) { \
			fprintf(stderr, "cpic_hip: %s failed: %s (%s:%d)\n", #call, \
					hipGetErrorString(e_), __FILE__, __LINE__); \
			return CPIC_EHIP; \
		} \
	} while(0)

// Dynamic LDS limit requested for kernels with large tiles or rows: the
// 160 KB of a gfx950 workgroup minus room for static LDS.
#define CPIC_LDS_LIMIT (160 * 1024 - 8 * 1024)

static inline int ilog2(int64_t n)
{
	int l = 0;
	while(((int64_t) 1 << l) < n)
		l++;
	return l;
}

static inline int is_pow2(int64_t n)
{
	return n > 0 && (n & (n - 1)) == 0;
}
