// Internal state of the gfx950 PIC step.  Not part of the C ABI.
//
// One state object holds P >= 1 independent plasmas (a single simulation is
// a batch of one).  Every kernel runs once per stage for the whole batch over
// a work list of (plasma, item): tiles for the particle kernels, row blocks
// and column blocks for the FFT passes, node blocks for the E stencil.  Block
// b finds its plasma by binary search in a prefix array of item counts and
// reads the plasma's geometry from a device table.
//
// Layout in HBM
//   fields   rho[2] (double, the plasmas' nx*ny grids back to back,
//            double-buffered: the solve reads one while the push deposits into
//            the other), ghat ((nx/2+1)*ny complex per plasma, stored kx-major
//            so the column pass reads contiguous columns), phi (double),
//            E (double2 or float2 (Ex,Ey) pairs).
//   particles buckets, one per (plasma, tile, species), tile-major inside a
//            plasma; state x, y, ux, uy in groups of 64 slots (+ optional id,
//            Ex, Ey arrays).  A bucket
//            owns slots [start, start+cap): segment 1 holds n1 particles at
//            [0, n1), segment 2 holds n2 fallback particles at
//            [cap-res, cap-res+n2), where res is the number of particles that
//            left the bucket in the last push (their slots are kept free so a
//            leaver whose destination is full can always return).
//   FP64 mode stores absolute positions (the reference's r); FP32 mode
//   stores positions in cell units relative to the bucket tile's origin.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/cpic_hip.h"

#define CPIC_BLOCK 256
#define CPIC_MARGIN 1            // cells of drift covered by the LDS tile regions
#define CPIC_BUCKET_MAX (1 << 27) // slots of one (tile, species) bucket: 32-bit byte offsets
#define CPIC_MAX_KERNELS 24

// device error word bits
#define CPIC_DERR_UMAX 0x1
#define CPIC_DERR_NAN 0x2

// Geometry and constants of one plasma (device table entry).
struct Geo {
	int nx, ny;          // grid points (powers of two)
	int lgnx, lgny;
	int tx, ty;          // tile edge in cells
	int lgtx, lgty;
	int ntx, nty;        // tiles per direction
	int lgntx;
	int ntiles;
	int W, H;            // LDS region: (tx + 2M + 1) x (ty + 2M + 1) nodes
	int M;
	int ns;              // species
	int R, C;            // FFT rows per row block, columns per column block
	double dx, dy;       // cell size
	double idx, idy;     // 1/dx, 1/dy
	double L0, L1;       // simulation.space_length
	double dt, e0, Bz;
	double umax_x, umax_y;
	double q[CPIC_MAX_SPECIES], m[CPIC_MAX_SPECIES];
	int64_t node0;       // first node in rho/phi/E
	int64_t ghat0;       // first element of ghat
	int64_t cos0;        // first element of costab
	int tile0;           // first global tile
	int bucket0;         // first global bucket
	int diag0;           // first slot of this plasma's diagnostic ring
	int pad;
};

// Work-list prefix arrays (P+1 entries each), concatenated.
enum { PRE_TILE = 0, PRE_ROW, PRE_COL, PRE_NODEBLK, PRE_BUCKET, PRE_COUNT };

struct Buckets {
	int64_t *start;
	int32_t *cap;
	int32_t *n1;
	int32_t *res;
	int32_t *n2;
};

// Particle state (double or float) in groups of PGROUP particles: group k
// holds x[64] y[64] ux[64] uy[64] of slots 64k..64k+63, so a wavefront's
// four loads of 64 consecutive slots fall in one 1 KiB (FP32) or 2 KiB
// (FP64) run of memory instead of four separate arrays.
#define PGROUP 64
struct Parts {
	void *r;                 // 4 * capacity elements, capacity a multiple of 64
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

// The push's kinetic and momentum sums go to DIAG_SPREAD accumulators per
// plasma, each in a 128-byte line of its own (wavefront w of block b adds to
// accumulator (4b + w) mod DIAG_SPREAD), which k_append adds into the
// plasma's slot: atomics from every block onto one address serialise.
#define DIAG_SPREAD 64
#define DIAG_LINE 16

struct DiagSlot {
	double field;
	double kinetic;
	double px, py;
};

// Device-side step counter (the reference's sim->iter): set by the host at
// pre_step, advanced by the last kernel of every step, read by the kernels
// that record diagnostics, so captured steps replay without new arguments.
struct StepState {
	long long iter;
	long long pad[7];
};

struct KernelTimer {
	const char *name;
	int64_t launches;
	double total_ms;
};

struct cpic_dev {
	int P;                   // plasmas
	cpic_dev_desc_t *desc;   // [P]
	Geo *hgeo;               // [P] host copy
	Geo *geo;                // [P] device
	int *hpre, *pre;         // [PRE_COUNT][P+1] host / device
	int prec;                // 32 or 64
	int ppt;                 // particles per thread and chunk in the push: 2 or 4
	int flags;
	int device;
	int maxns;
	hipStream_t stream;
	int ring;                // diagnostic slots per plasma

	// totals
	int64_t nodes;           // sum nx*ny
	int64_t ghat_n;          // sum (nx/2+1)*ny
	int ntiles;              // sum of tiles
	int nb;                  // sum of buckets
	int rowblocks, colblocks, nodeblocks;
	int fft_nt_row, fft_nt_col;   // threads per block of the FFT row and column passes
	size_t lds_push, lds_rows, lds_cols;

	// fields
	double *rho[2];
	int rho_cur;
	int rho_next_ready;      // rho[1-cur] has been zeroed
	int E_stale;             // phi is newer than the E array (field_E leaves
	                         // E to the push; cpic_efield brings it up to date)
	int rho_pending;         // plasma_r deposited into rho[1-cur]
	double2 *ghat;
	double *phi;
	void *E;                 // double2 (FP64) or float2 (FP32)
	double *costab;          // per plasma: cos(2 pi k/nx) k<=nx/2, then cos(2 pi k/ny) k<ny
	double2 *tw;             // exp(-2 pi i k/N), k <= N/2, N = max over plasmas
	int lgN;

	// particles
	Parts p[2];              // [cur] in use, [1-cur] spare for rebuilds
	int pcur;
	int64_t pcap[2];
	Buckets bk;
	int *order;              // [ntiles] tile taken by each block of the tile kernels
	int *border;             // [nb] bucket taken by each block of the push
	double *dpart;           // [P][DIAG_SPREAD][DIAG_LINE]: the push's kinetic, px, py sums,
	                         // then one line whose first 8 bytes hold their iteration
	Leavers lv;
	int step_parity;
	int64_t *count;          // [P * CPIC_MAX_SPECIES]

	// pending uploads (host copies, applied by the next rebuild)
	int pending;
	int64_t *up_n;           // [P * CPIC_MAX_SPECIES]
	double **up_buf;         // x,y,ux,uy as 4 arrays
	int64_t **up_id;

	// diagnostics and errors
	DiagSlot *diag;          // device ring, P * ring slots
	int *derr;               // device error word
	Counters *ctr;           // device counters
	StepState *st;           // device step counter
	int64_t iter;            // host mirror of the step counter (next step)
	int64_t last_iter;       // last iteration pushed
	int64_t nrebuilds;
	int broken;              // a re-sort failed part way (its error code):
	                         // every later call but destroy returns CPIC_ESTATE
	unsigned long long misplaced_seen;   // ctr->misplaced already acted on
	unsigned long long fallbacks_seen;   // ctr->fallbacks already acted on

	// captured steps: a graph of two consecutive steps (leaver parity and
	// rho buffer alternate each step), valid for steps >= 1
	hipGraphExec_t graph;
	int graph_parity;        // step_parity at capture
	int graph_rho;           // rho_cur at capture
	int graph_pcur;          // pcur at capture
	int graph_ok;
	int capturing;           // inside hipStreamBeginCapture/EndCapture

	// profiling
	KernelTimer kt[CPIC_MAX_KERNELS];
	int nkt;
	hipEvent_t *ev;          // pairs
	int *ev_kernel;
	int nev, evcap;
};

// helpers implemented in the .hip files
int cpic_fft_init(cpic_dev *d);
void cpic_fft_threads(int64_t mmax, int64_t nymax, int *nt_row, int *nt_col);
void cpic_fft_blocking(Geo *g, int nt_row, int nt_col);
void cpic_fft_spread(Geo *g, int P);
int cpic_fft_blocking_ok(const Geo *g, int nt_row, int nt_col);
size_t cpic_push_lds(const cpic_dev *d);
int cpic_set_iter(cpic_dev *d, int64_t iter);
int cpic_solve(cpic_dev *d, int want_energy);
int cpic_solve_step(cpic_dev *d);
int cpic_solve_ptr(cpic_dev *d, const double *rho, double *phi);
// the same, also zeroing the nodes of rho_zero (another buffer) in the
// inverse row pass
int cpic_solve_ptr_zero(cpic_dev *d, const double *rho, double *phi, double *rho_zero);
// the same, also adding each plasma's field energy -(e0/2) sum rho phi to
// energy[p] (summed over the spectrum by the column pass)
int cpic_solve_ptr_energy(cpic_dev *d, const double *rho, double *phi, double *rho_zero,
		double *energy);
int cpic_efield_ptr(cpic_dev *d, const double *phi, void *E);
int cpic_efield(cpic_dev *d);
int cpic_parts_init(cpic_dev *d);
int cpic_parts_order(cpic_dev *d, const int *cnt);
int cpic_parts_rebuild(cpic_dev *d);
int cpic_parts_push(cpic_dev *d, int first_iter);
int cpic_parts_deposit(cpic_dev *d, double *rho_out);
int cpic_parts_gather(cpic_dev *d);
int cpic_parts_download(cpic_dev *d, int plasma, int species, cpic_particles_t *out);
int cpic_parts_cell_counts(cpic_dev *d, int plasma, int species, int64_t *counts);
int cpic_parts_alloc(cpic_dev *d, int which, int64_t cap);
void cpic_parts_free(cpic_dev *d, int which);
int cpic_zero(cpic_dev *d, double *p, int64_t n);

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

// Pointer qualifiers for LDS and global memory.  Kernels spell them out
// where a call site has both an LDS and a global path: through generic
// pointers the compiler may merge the two into one flat instruction, which
// goes through the vector-memory path even for LDS.
#define LDS_AS __attribute__((address_space(3)))
#define GLB_AS __attribute__((address_space(1)))

// Dynamic LDS limit requested for kernels with large tiles or rows: the
// 160 KB of a gfx950 workgroup minus room for static LDS.
#define CPIC_LDS_LIMIT (160 * 1024 - 8 * 1024)
// LDS budgets used to choose FFT columns and rows per block
#define FFT_LDS_BYTES (64 * 1024)
#define FFT_ROW_LDS_BYTES (76 * 1024)
// grid nodes per block of the E stencil
#define EF_NODES 2048
// threads per block of the FFT passes; a block transforms at most
// FFT_MAX_VALUES complex values (the LDS budget above keeps every pass
// within it)
#define FFT_THREADS 1024
#define FFT_MIN_BLOCKS 256       // blocks per FFT pass worth spreading to (one per CU)
#define FFT_MAX_VALUES 8192
// Smaller blocks for the passes whose rows / columns fit them (more blocks
// resident per CU, cheaper barriers): threads per row and column block,
// values per thread in a row block, and their LDS budgets.
#ifndef FFT_ROW_NT
#define FFT_ROW_NT 1024
#endif
#ifndef FFT_ROW_VPT
#define FFT_ROW_VPT 8
#endif
#ifndef FFT_ROW_LDS_SMALL
#define FFT_ROW_LDS_SMALL (76 * 1024)
#endif
#ifndef FFT_COL_NT
#define FFT_COL_NT 256
#endif
#ifndef FFT_COL_LDS_SMALL
#define FFT_COL_LDS_SMALL (32 * 1024)
#endif
// Bytes of an FFT LDS buffer holding n complex values.
static inline size_t fft_lds_bytes(int64_t n)
{
	return 16 * (size_t) (n > 0 ? n : 1);
}

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

// plasma owning work item x: pre[p] <= x < pre[p+1]
static __device__ __forceinline__ int find_plasma(const int *__restrict__ pre, int P, int x)
{
	int lo = 0, hi = P - 1;
	while(lo < hi)
	{
		int mid = (lo + hi + 1) >> 1;
		if(pre[mid] <= x)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}
