// Particle kernels of the gfx950 PIC step.
//
// k_push: one workgroup per (tile, species) bucket.  Stages the tile's E
// region (tile + 1 cell margin + the CIC +1 node) in LDS -- after a field
// solve it computes that region from the potential (E = -grad phi,
// src/field.c:381-406, staged with a one-node halo), so a step never writes
// or reads a global E array -- streams every particle of the bucket once
// (coalesced loads of the state, stored in groups of 64: x[64] y[64] ux[64]
// uy[64]), and for each particle
//   gathers E with CIC weights    (interpolate_f2p_E, src/interpolate.c:350-404)
//   applies the Boris push        (boris_rotation/move, src/mover.c:23-83, 141-227)
//   wraps periodically            (periodic_boundary_ppack, src/comm_plasma.c:726-747)
//   deposits -q/e0 * w into an LDS rho region (interpolate_p2f_rho,
//                                  src/interpolate.c:283-346) with FP64 LDS atomics
//                                  in both modes: on gfx950 ds_add_f64 is cheap while
//                                  ds_add_f32 costs more than the whole HBM stream
//                                  (FP32 push at 2048^2: 5.71 ms per launch with
//                                  ds_add_f32 on an LDS pointer, 1.80 ms with
//                                  ds_add_f64)
//   and writes the tile's stayers to the other particle buffer (the two
//   buffers swap every step; each wavefront claims its stayers' slots with
//   one LDS atomic, so no block-wide barrier is needed per chunk) while
//   leavers (particles whose cell moved to another tile: comm_plasma's
//   collect, src/comm_plasma.c:101-165) wait in an LDS stage and are copied
//   to a leaver buffer with one global atomic per bucket.
// The LDS rho region is flushed once per workgroup with FP64 global atomics.
// k_append then places each leaver at the end of its destination bucket
// (inject_particles, src/comm_plasma.c:358-400) or, when that bucket is full,
// back into a reserved slot of its source bucket, so no particle is ever lost
// whatever the capacities; such particles are handled by the slow global
// path until the next re-sort.
//
// FP64 mode uses the reference's arithmetic bit for bit: positions are
// absolute, cell indices and weights follow relative_position_grid
// (src/interpolate.c:39-66, including dx[X] in the Y term) and the Boris
// push is the reference's 3D form reduced exactly to B = (0,0,Bz)
// (u_z stays 0).  The translation unit is compiled with -ffp-contract=off;
// the multiply-adds the reference's clang build contracts are explicit fma().
#include "dev.h"
#include "picmath.h"

#include <algorithm>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

template <typename T> struct V2;
template <> struct V2<double> { typedef double2 t; };
template <> struct V2<float> { typedef float2 t; };
// the same pairs as clang vectors, which load through address-space pointers
template <typename T> struct N2 { typedef T t __attribute__((ext_vector_type(2))); };
template <typename T, int U> struct NU { typedef T t __attribute__((ext_vector_type(U))); };

// Arguments of the tile kernels.  Block b handles batch tile order[b], finds
// the tile's plasma in pre_tile and takes the plasma's geometry and
// constants from geo.
struct PushArgs {
	const Geo *geo;
	const int *pre_tile;  // [P+1] first global tile of each plasma
	const int *order;     // [ntiles] tile of each block (cpic_parts_order); k_push:
	                      // [nb] bucket of each block
	const int *pre_bucket; // [P+1] first global bucket of each plasma
	int P;
	int ntiles;           // tiles of the whole batch
	Buckets bk;
	Parts p;              // particles read
	Parts q;              // particles written (the other buffer)
	Leavers lv;
	int par;              // leaver counter parity
	int first;            // iteration 0: dt' = -dt/2 and no position update
	const void *E;        // whole batch, plasma p at node0
	const double *phi;    // potential, whole batch (pushes that compute E)
	double *rho;          // deposit target, whole batch
	DiagSlot *diag;       // nullable: per-plasma rings of `ring` slots
	double *dpart;        // with diag: [P][DIAG_SPREAD][DIAG_LINE] accumulators
	long long *diag_iter; // with diag: the iteration the accumulators are for
	int ring;
	const StepState *st;
	int *derr;
	Counters *ctr;
};

// Per-species constants of the push, derived on the device from the
// plasma's table entry with the host's expressions, so they are the same
// doubles the reference computes: dtqm2 = 0.5 * dt' * q / m and the
// rotation factors t, s = 2t/(1+t^2) (src/mover.c:23-83, 204-227).
// For FP32 particles the velocity limits umax are taken as the largest
// floats not above them: for a float u and a double U, |u| > U exactly when
// |u| > that float, so the check_velocity test stays exact in FP32.
template <typename T> struct SpeciesK {
	T h, tz, sz, vq;
	double mh;
	float umx, umy;
};

static __device__ __forceinline__ float f32_at_most(double v)
{
	float f = (float) v;
	return (double) f > v ? nextafterf(f, -INFINITY) : f;
}

template <typename T>
static __device__ __forceinline__ SpeciesK<T> species_k(const Geo &g, int s, double dtp)
{
	SpeciesK<T> k;
	double h = 0.5 * dtp * g.q[s] / g.m[s];
	double tz = g.Bz * h;
	double sden = fma(tz, tz, 1.0);
	k.h = (T) h;
	k.tz = (T) tz;
	k.sz = (T) (2.0 * tz / sden);
	k.vq = (T) (-g.q[s] / g.e0);
	k.mh = 0.5 * g.m[s];
	k.umx = f32_at_most(g.umax_x);
	k.umy = f32_at_most(g.umax_y);
	return k;
}

enum { MODE_PUSH = 0, MODE_DEPOSIT = 1, MODE_GATHER = 2 };

static __device__ __forceinline__ int wrapi(int i, int n)
{
	i = i < 0 ? i + n : i;
	return i >= n ? i - n : i;
}

// Block -> tile map for tiles of similar size: contiguous tile bands go to
// each of the 8 XCDs (workgroups are dispatched round-robin), so
// neighbouring tiles share an L2.
static inline int tile_map(int b, int ntiles)
{
	if((ntiles & 7) == 0)
		return (b & 7) * (ntiles >> 3) + (b >> 3);
	return b;
}

static __device__ __forceinline__ int bucket_n1(const Buckets &bk, int b)
{
	int lim = bk.cap[b] - bk.res[b];
	int n1 = bk.n1[b];
	return n1 < lim ? n1 : lim;
}

static __device__ __forceinline__ int64_t bucket_slot(const Buckets &bk, int b,
		int n1, int j)
{
	int64_t s = bk.start[b];
	return j < n1 ? s + j : s + (bk.cap[b] - bk.res[b]) + (j - n1);
}

// Per-particle state in registers.
template <typename T> struct PState {
	T x, y, ux, uy;
	int64_t id;
	double Ex, Ey;
};

// Element offset of slot s's x in the grouped state; y, ux and uy follow
// at +64, +128 and +192.
static_assert(PGROUP == 64, "pslot assumes groups of 64");
static __device__ __forceinline__ int64_t pslot(int64_t s)
{
	return ((s >> 6) << 8) + (s & 63);
}

// Bucket starts are multiples of PGROUP (capacities are rounded up to it),
// so slot bstart + j sits at element 4 * bstart + pslot(j): the push keeps a
// per-bucket base and does 32-bit index work per particle.
template <typename T>
static __device__ __forceinline__ const GLB_AS T *bucket_base(const Parts &p, int64_t bstart)
{
	return gptr<T>(p.r) + 4 * bstart;
}

// Byte offset of bucket slot j's x from the bucket base.  It fits 32 bits
// (bucket capacities stay below CPIC_BUCKET_MAX slots, k_scan_caps), so the
// particle loads and stores take the scalar-base + 32-bit-vector-offset form.
template <typename T>
static __device__ __forceinline__ uint32_t pslot_byte(int j)
{
	return ((((uint32_t) j >> 6) << 8) | ((uint32_t) j & 63u)) * (uint32_t) sizeof(T);
}

template <typename T>
static __device__ __forceinline__ GLB_AS T *at_byte(const GLB_AS T *base, uint32_t off)
{
	return (GLB_AS T *) ((const GLB_AS char *) base + off);
}

// One DPP row operation (CTRL, rows ROW_MASK) on both halves of a double;
// lanes without a source get 0.
template <int CTRL, int ROW_MASK>
static __device__ __forceinline__ double dpp_f64(double v)
{
	int lo = __double2loint(v), hi = __double2hiint(v);
	lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROW_MASK, 0xf, false);
	hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROW_MASK, 0xf, false);
	return __hiloint2double(hi, lo);
}

// Sum of v over a full wavefront, left in lane 63: an inclusive scan with
// row_shr 1, 2, 4, 8 inside each row of 16 lanes, then row_bcast 15 (rows
// 1 and 3 add the last lane of the row before) and row_bcast 31 (rows 2
// and 3 add lane 31).  No LDS traffic, unlike shuffles through
// ds_bpermute.
static __device__ __forceinline__ double wave_sum_lane63(double v)
{
	v += dpp_f64<0x111, 0xf>(v);
	v += dpp_f64<0x112, 0xf>(v);
	v += dpp_f64<0x114, 0xf>(v);
	v += dpp_f64<0x118, 0xf>(v);
	v += dpp_f64<0x142, 0xa>(v);
	v += dpp_f64<0x143, 0xc>(v);
	return v;
}

// Number of bits of m set below this lane.
static __device__ __forceinline__ int lanes_below(unsigned long long m)
{
	return (int) __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32),
			__builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

// A block-uniform value read from memory the kernel also writes: the
// compiler must use a vector load, so the value would occupy VGPRs.
static __device__ __forceinline__ int uniform(int v)
{
	return __builtin_amdgcn_readfirstlane(v);
}

static __device__ __forceinline__ int64_t uniform(int64_t v)
{
	unsigned lo = (unsigned) __builtin_amdgcn_readfirstlane((int) (uint64_t) v);
	unsigned hi = (unsigned) __builtin_amdgcn_readfirstlane((int) ((uint64_t) v >> 32));
	return (int64_t) (((uint64_t) hi << 32) | lo);
}

template <typename T>
static __device__ __forceinline__ T p_x(const Parts &p, int64_t s)
{
	return gptr<T>(p.r)[pslot(s)];
}

template <typename T>
static __device__ __forceinline__ T p_y(const Parts &p, int64_t s)
{
	return gptr<T>(p.r)[pslot(s) + PGROUP];
}

template <typename T>
static __device__ __forceinline__ void load_p(const Parts &p, int64_t s, PState<T> &q,
		bool ids)
{
	// global-address-space loads: generic (flat) loads would also count
	// against the LDS wait counter and hold up the LDS gather behind them
	const GLB_AS T *r = gptr<T>(p.r) + pslot(s);
	q.x = r[0];
	q.y = r[PGROUP];
	q.ux = r[2 * PGROUP];
	q.uy = r[3 * PGROUP];
	if(ids)
		q.id = gptr<int64_t>(p.id)[s];
}

template <typename T>
static __device__ __forceinline__ void store_p(const Parts &p, int64_t s, const PState<T> &q,
		bool ids, bool pe)
{
	GLB_AS T *r = gptr<T>(p.r) + pslot(s);
	r[0] = q.x;
	r[PGROUP] = q.y;
	r[2 * PGROUP] = q.ux;
	r[3 * PGROUP] = q.uy;
	if(ids)
		p.id[s] = q.id;
	if(pe)
	{
		p.Ex[s] = q.Ex;
		p.Ey[s] = q.Ey;
	}
}

template <typename T>
static __device__ __forceinline__ void store_l(const Leavers &l, int64_t s, const PState<T> &q,
		bool ids, bool pe, int dst, int src)
{
	((T *) l.x)[s] = q.x;
	((T *) l.y)[s] = q.y;
	((T *) l.ux)[s] = q.ux;
	((T *) l.uy)[s] = q.uy;
	if(ids)
		l.id[s] = q.id;
	if(pe)
	{
		l.Ex[s] = q.Ex;
		l.Ey[s] = q.Ey;
	}
	l.dst[s] = dst;
	l.src[s] = src;
}

// Position of a particle as (global cell, in-cell offset).  For FP32 the
// stored coordinates are cell units relative to the bucket tile origin.
struct CellPos {
	int ix, iy;        // global cell
	int lx, ly;        // cell relative to the LDS region origin (may be out of range)
};

// Gather E at one particle: LDS region when the 2x2 stencil is inside it,
// else straight from HBM (the E array, or with EPHI the potential).
// Summation order of interpolate_f2p (src/interpolate.c:151-154): w00, w01,
// w10, w11 with fma accumulation.
template <typename T, bool EPHI>
static __device__ __forceinline__ void gather(const Geo &g, const typename V2<T>::t *sE,
		const typename V2<T>::t *__restrict__ E, const double *__restrict__ phi,
		const CellPos &c, T rx, T ry, T &ex, T &ey)
{
	typedef typename V2<T>::t T2;
	T dxw = (T) 1 - rx, dyw = (T) 1 - ry;
	T w00 = dxw * dyw, w01 = dxw * ry, w10 = rx * dyw, w11 = rx * ry;
	T2 e00, e01, e10, e11;
	// Explicit address spaces: with generic pointers the compiler merges the
	// two branches into one set of flat loads, which go through the vector
	// memory path even for LDS and wait for every outstanding HBM access.
	if(c.lx >= 0 && c.lx < g.W - 1 && c.ly >= 0 && c.ly < g.H - 1)
	{
		typedef typename N2<T>::t P2;
		const LDS_AS P2 *l = (const LDS_AS P2 *) sE;    // (x, y) pairs
		int i = c.ly * g.W + c.lx;
		P2 a00 = l[i], a01 = l[i + g.W], a10 = l[i + 1], a11 = l[i + g.W + 1];
		e00.x = a00.x;
		e00.y = a00.y;
		e01.x = a01.x;
		e01.y = a01.y;
		e10.x = a10.x;
		e10.y = a10.y;
		e11.x = a11.x;
		e11.y = a11.y;
	}
	else if(EPHI)
	{
		// rare (particles left outside their tile's region by a full
		// leaver buffer); one division at a time keeps the register demand
		// of the particle loop down.  E at a node as k_efield computes it
		// (src/field.c:381-406: central differences over 2 dx and 2 dy),
		// same products and summation order as below.
		int i1x = c.ix + 1 == g.nx ? 0 : c.ix + 1;
		int i1y = c.iy + 1 == g.ny ? 0 : c.iy + 1;
		T acc[2];
#pragma unroll 1
		for(int j = 0; j < 8; j++)
		{
			int k = j & 3, comp = j >> 2;
			int ix = k & 2 ? i1x : c.ix, iy = k & 1 ? i1y : c.iy;
			int64_t n0, n1;
			if(comp)
			{
				n0 = (int64_t) (iy == 0 ? g.ny - 1 : iy - 1) * g.nx + ix;
				n1 = (int64_t) (iy == g.ny - 1 ? 0 : iy + 1) * g.nx + ix;
			}
			else
			{
				n0 = (int64_t) iy * g.nx + (ix == 0 ? g.nx - 1 : ix - 1);
				n1 = (int64_t) iy * g.nx + (ix == g.nx - 1 ? 0 : ix + 1);
			}
			T v = (T) ((phi[n0] - phi[n1]) / (comp ? 2 * g.dy : 2 * g.dx));
			T w = k == 0 ? w00 : k == 1 ? w01 : k == 2 ? w10 : w11;
			T a = k == 0 ? w * v : fma(w, v, acc[comp]);
			acc[comp] = a;
		}
		ex = acc[0];
		ey = acc[1];
		return;
	}
	else
	{
		const GLB_AS T *m = (const GLB_AS T *) E;
		int i1x = c.ix + 1 == g.nx ? 0 : c.ix + 1;
		int i1y = c.iy + 1 == g.ny ? 0 : c.iy + 1;
		int64_t n00 = (int64_t) c.iy * g.nx + c.ix, n01 = (int64_t) i1y * g.nx + c.ix;
		int64_t n10 = (int64_t) c.iy * g.nx + i1x, n11 = (int64_t) i1y * g.nx + i1x;
		e00.x = m[2 * n00];
		e00.y = m[2 * n00 + 1];
		e01.x = m[2 * n01];
		e01.y = m[2 * n01 + 1];
		e10.x = m[2 * n10];
		e10.y = m[2 * n10 + 1];
		e11.x = m[2 * n11];
		e11.y = m[2 * n11 + 1];
	}
	ex = w00 * e00.x;
	ex = fma(w01, e01.x, ex);
	ex = fma(w10, e10.x, ex);
	ex = fma(w11, e11.x, ex);
	ey = w00 * e00.y;
	ey = fma(w01, e01.y, ey);
	ey = fma(w10, e10.y, ey);
	ey = fma(w11, e11.y, ey);
}

// Deposit -q/e0 * w at the 4 nodes (order of src/interpolate.c:259-262).
template <typename T>
static __device__ __forceinline__ void deposit(const Geo &g, double *sR, double *__restrict__ rho,
		const CellPos &c, T rx, T ry, T vq)
{
	T dxw = (T) 1 - rx, dyw = (T) 1 - ry;
	T w00 = dxw * dyw, w01 = dxw * ry, w10 = rx * dyw, w11 = rx * ry;
	if(c.lx >= 0 && c.lx < g.W - 1 && c.ly >= 0 && c.ly < g.H - 1)
	{
		LDS_AS double *l = (LDS_AS double *) sR;
		int i = c.ly * g.W + c.lx;
		lds_add(&l[i], (double) (w00 * vq));
		lds_add(&l[i + g.W], (double) (w01 * vq));
		lds_add(&l[i + 1], (double) (w10 * vq));
		lds_add(&l[i + g.W + 1], (double) (w11 * vq));
	}
	else
	{
		GLB_AS double *m = gptr<double>(rho);
		int i1x = c.ix + 1 == g.nx ? 0 : c.ix + 1;
		int i1y = c.iy + 1 == g.ny ? 0 : c.iy + 1;
		glb_add(&m[(int64_t) c.iy * g.nx + c.ix], (double) (w00 * vq));
		glb_add(&m[(int64_t) i1y * g.nx + c.ix], (double) (w01 * vq));
		glb_add(&m[(int64_t) c.iy * g.nx + i1x], (double) (w10 * vq));
		glb_add(&m[(int64_t) i1y * g.nx + i1x], (double) (w11 * vq));
	}
}

// Cell of a particle and its offsets, both representations.
static __device__ __forceinline__ void locate(const Geo &g, int ox, int oy, double x, double y,
		CellPos &c, double &rx, double &ry)
{
	cell64(g, x, y, c.ix, c.iy, rx, ry);
	c.lx = wrapi(c.ix - ox, g.nx);
	c.ly = wrapi(c.iy - oy, g.ny);
}

static __device__ __forceinline__ void locate(const Geo &g, int ox, int oy, float x, float y,
		CellPos &c, float &rx, float &ry)
{
	// ox, oy = tile origin - M; stored coordinates are relative to the tile
	float fx = floorf(x), fy = floorf(y);
	int cx = (int) fx, cy = (int) fy;
	rx = x - fx;
	ry = y - fy;
	c.lx = cx + g.M;
	c.ly = cy + g.M;
	int bx0 = ox + g.M, by0 = oy + g.M;
	c.ix = (bx0 + cx) & (g.nx - 1);
	c.iy = (by0 + cy) & (g.ny - 1);
}

__global__ __launch_bounds__(CPIC_BLOCK)
void k_zero(double *__restrict__ p, int64_t n)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
		p[i] = 0.0;
}

// Loads the tile's E region into LDS (periodic indexing) and zeroes the
// LDS rho region.
template <typename T>
static __device__ __forceinline__ void tile_setup(const Geo &g, int ox, int oy,
		const typename V2<T>::t *__restrict__ E, typename V2<T>::t *sE, double *sR)
{
	for(int i = threadIdx.x; i < g.W * g.H; i += blockDim.x)
	{
		int ly = i / g.W, lx = i - ly * g.W;
		int gx = wrapi(ox + lx, g.nx), gy = wrapi(oy + ly, g.ny);
		if(sE)
			sE[i] = E[(int64_t) gy * g.nx + gx];
		if(sR)
			sR[i] = 0.0;
	}
}

// Computes the tile's E region in LDS from the potential: the region plus a
// one-node halo of phi is staged in the rho area (row stride W + 2), E
// follows with the expression of k_efield (src/field.c:381-406), then the rho
// area is zeroed.
// The grid sides are powers of two, so "& (n - 1)" wraps any offset.
template <typename T>
static __device__ __forceinline__ void tile_setup_phi(const Geo &g, int ox, int oy,
		const double *__restrict__ phi, typename V2<T>::t *sE, double *sR)
{
	typedef typename V2<T>::t T2;
	const int PW = g.W + 2, PH = g.H + 2;
	for(int i = threadIdx.x; i < PW * PH; i += blockDim.x)
	{
		int ly = i / PW, lx = i - ly * PW;
		int gx = (ox - 1 + lx) & (g.nx - 1), gy = (oy - 1 + ly) & (g.ny - 1);
		sR[i] = phi[(int64_t) gy * g.nx + gx];
	}
	__syncthreads();
	for(int i = threadIdx.x; i < g.W * g.H; i += blockDim.x)
	{
		int ly = i / g.W, lx = i - ly * g.W;
		int c = (ly + 1) * PW + lx + 1;
		double ey = (sR[c - PW] - sR[c + PW]) / (2 * g.dy);
		double ex = (sR[c - 1] - sR[c + 1]) / (2 * g.dx);
		T2 v;
		v.x = ex;
		v.y = ey;
		sE[i] = v;
	}
	__syncthreads();
	for(int i = threadIdx.x; i < g.W * g.H; i += blockDim.x)
		sR[i] = 0.0;
}

// Adds the LDS rho region into the global rho with FP64 atomics.
template <typename T>
static __device__ __forceinline__ void tile_flush(const Geo &g, int ox, int oy, const double *sR,
		double *__restrict__ rho)
{
	for(int i = threadIdx.x; i < g.W * g.H; i += blockDim.x)
	{
		double v = sR[i];
		if(v != 0.0)
		{
			int ly = i / g.W, lx = i - ly * g.W;
			int gx = wrapi(ox + lx, g.nx), gy = wrapi(oy + ly, g.ny);
			unsafeAtomicAdd(&rho[(int64_t) gy * g.nx + gx], (double) v);
		}
	}
}

// One particle: gather, Boris push, move + wrap, deposit at the new
// position; returns whether its cell left the tile (and the destination).
template <typename T, bool EPHI>
static __device__ __forceinline__ bool push_one(const Geo &g, const typename V2<T>::t *sE,
		const typename V2<T>::t *__restrict__ E, const double *__restrict__ phi, double *sR,
		double *__restrict__ rho,
		int tile, int ox, int oy, int s, const SpeciesK<T> &k, bool set_r, double dtp,
		float dtdx, float dtdy, PState<T> &q, int &ndst, double &ke, double &mpx,
		double &mpy, bool &bad, bool &nan_)
{
	const T h = k.h, tz = k.tz, sz = k.sz;
	const double mh = k.mh;
	CellPos c;
	T rx, ry, ex, ey;
	locate(g, ox, oy, q.x, q.y, c, rx, ry);
	gather<T, EPHI>(g, sE, E, phi, c, rx, ry, ex, ey);
	q.Ex = (double) ex;
	q.Ey = (double) ey;
	T unx = q.ux, uny = q.uy;
	boris_bz<T>(h, tz, sz, ex, ey, unx, uny);
	// check_velocity (src/mover.c:98-137)
	if(sizeof(T) == 4 ? fabsf((float) unx) > k.umx || fabsf((float) uny) > k.umy
			: fabs((double) unx) > g.umax_x || fabs((double) uny) > g.umax_y)
		bad = true;
	double ux0 = (double) q.ux, uy0 = (double) q.uy;
	double ux1 = (double) unx, uy1 = (double) uny;
	ke += mh * 0.5 * ((ux0 * ux0 + uy0 * uy0) + (ux1 * ux1 + uy1 * uy1));
	mpx += mh * (ux0 + ux1);
	mpy += mh * (uy0 + uy1);
	q.ux = unx;
	q.uy = uny;
	if(set_r)
	{
		if(sizeof(T) == 8)
		{
			double nx_ = fma((double) unx, dtp, (double) q.x);
			double ny_ = fma((double) uny, dtp, (double) q.y);
			q.x = (T) wrap64(nx_, g.L0);
			q.y = (T) wrap64(ny_, g.L1);
		}
		else
		{
			q.x = (T) fmaf((float) unx, dtdx, (float) q.x);
			q.y = (T) fmaf((float) uny, dtdy, (float) q.y);
		}
	}
	if(!(isfinite(q.x) && isfinite(q.y) && isfinite(q.ux) && isfinite(q.uy)))
	{
		nan_ = true;
		q.x = q.y = (T) 0;
	}
	CellPos c2;
	locate(g, ox, oy, q.x, q.y, c2, rx, ry);
	deposit<T>(g, sR, rho, c2, rx, ry, k.vq);
	int ntile = (c2.iy >> g.lgty) * g.ntx + (c2.ix >> g.lgtx);
	if(ntile == tile)
		return false;
	ndst = g.bucket0 + ntile * g.ns + s;
	if(sizeof(T) == 4)
	{
		// re-express relative to the destination tile
		float fx = floorf((float) q.x), fy = floorf((float) q.y);
		q.x = (T) ((float) (c2.ix & (g.tx - 1)) + ((float) q.x - fx));
		q.y = (T) ((float) (c2.iy & (g.ty - 1)) + ((float) q.y - fy));
	}
	return true;
}

// Loads the chunk of U*NT particles starting at bucket index base (NT
// threads): thread tid takes base + U*tid + u with vector loads when U = 4
// and the chunk lies in the bucket's first segment, else base + u*NT + tid
// (coalesced loads either way; every slot of the chunk exactly once).
template <typename T, bool IDS, int U, int NT>
static __device__ __forceinline__ void load_chunk(const Parts &p, int64_t bstart, int n1,
		int seg2, int ntot, int base, int tid, PState<T> *q, bool *act)
{
	const GLB_AS T *pb = bucket_base<T>(p, bstart);
	if(U == 4 && base + U * NT <= n1)
	{
		// the whole chunk lies in the first segment (the usual case): thread
		// tid takes the U consecutive slots base + U*tid ... (U divides the
		// group of 64, so they are contiguous), one vector load per component.
		// Measured with U = 4 only: with U = 2 (two-element vectors) the push
		// was 4 % slower than with the per-slot loads below.
		typedef typename NU<T, U>::t VU;
		const int j0 = base + U * tid;
		const GLB_AS VU *r = (const GLB_AS VU *) at_byte(pb, pslot_byte<T>(j0));
		VU vx = r[0], vy = r[PGROUP / U], vux = r[2 * PGROUP / U], vuy = r[3 * PGROUP / U];
#pragma unroll
		for(int u = 0; u < U; u++)
		{
			act[u] = true;
			q[u].x = vx[u];
			q[u].y = vy[u];
			q[u].ux = vux[u];
			q[u].uy = vuy[u];
			if(IDS)
				q[u].id = gptr<int64_t>(p.id)[bstart + j0 + u];
		}
		return;
	}
#pragma unroll
	for(int u = 0; u < U; u++)
	{
		int j = base + u * NT + tid;
		act[u] = j < ntot;
		int jj = j + (j >= n1 ? seg2 : 0);
		if(act[u])
		{
			// global-address-space loads (see load_p)
			const GLB_AS T *r = at_byte(pb, pslot_byte<T>(jj));
			q[u].x = r[0];
			q[u].y = r[PGROUP];
			q[u].ux = r[2 * PGROUP];
			q[u].uy = r[3 * PGROUP];
			if(IDS)
				q[u].id = gptr<int64_t>(p.id)[bstart + jj];
		}
	}
}

// Fused gather + push + wrap + deposit + tile compaction.  Each thread
// handles U particles per chunk (particle base + u*256 + tid), all loads of
// a chunk issued before any of them is used, so a workgroup keeps U*256
// particles of HBM reads in flight.  Stayers go to the front of the same
// bucket in the other buffer (a.q), each wavefront claiming its slots with
// one LDS atomic (their order within the bucket is not kept); leavers go
// to the leaver buffer with one atomic per wavefront.
//
// Waves per SIMD the register allocation of k_push must allow.  Measured at
// 2048^2, 64 ppc: the FP32 kernel gains ~3 % from 6 waves (U = 2, now 73
// VGPRs; 7 waves gain nothing), the FP64 kernel loses from any bound (its
// state is twice the size); bounding the FP32 U = 4 kernel (then 95 VGPRs,
// now 93) to 6 waves spilled and lost 12 %.
#ifndef PUSH_MIN_WAVES_F32
#define PUSH_MIN_WAVES_F32 6
#endif
// Threads per k_push workgroup.  512 (half the chunks per bucket; the
// compiler then holds the FP32 kernel to 81 VGPRs) measured slower at
// 2048^2: FP32 push 1.89 vs 1.74 ms, FP64 3.78 vs 3.29 ms.
#ifndef PUSH_BLOCK
#define PUSH_BLOCK CPIC_BLOCK
#endif



// Leavers of one bucket wait in LDS and are placed in the leaver buffer with
// one global atomic per bucket: a counter shared by every workgroup, hit once
// per wavefront with a leaver, serialises in one L2 channel and stalls the
// wavefront until it returns.  A wavefront whose leavers no longer fit in the
// stage claims its slots with its own global atomic.
#ifndef LV_STAGE
#define LV_STAGE 256
#endif
template <typename T, bool IDS, bool PE> struct LvStage {
	T x[LV_STAGE], y[LV_STAGE], ux[LV_STAGE], uy[LV_STAGE];
	int dst[LV_STAGE];
	int64_t id[IDS ? LV_STAGE : 1];
	double Ex[PE ? LV_STAGE : 1], Ey[PE ? LV_STAGE : 1];
	int n;        // records claimed (may pass LV_STAGE)
	int cut;      // first record of the wavefront that did not fit
	int base;     // first leaver-buffer slot of the staged records
};

// An FP32 leaver (coordinates relative to destination bucket ndst) put back
// relative to its source tile at origin (ox, oy) + margin.
template <typename T>
static __device__ __forceinline__ void back_to_tile(const Geo &g, int ndst, int ox, int oy,
		PState<T> &q)
{
	if(sizeof(T) != 4)
		return;
	int dt_ = (ndst - g.bucket0) / g.ns;
	int dtx = (dt_ & (g.ntx - 1)) * g.tx - (ox + g.M);
	int dty = (dt_ >> g.lgntx) * g.ty - (oy + g.M);
	if(dtx > g.nx / 2) dtx -= g.nx;
	if(dtx < -g.nx / 2) dtx += g.nx;
	if(dty > g.ny / 2) dty -= g.ny;
	if(dty < -g.ny / 2) dty += g.ny;
	q.x = (T) ((float) q.x + (float) dtx);
	q.y = (T) ((float) q.y + (float) dty);
}

// EPHI: the tile's E comes from the potential (the E array is stale).
template <typename T, bool IDS, bool PE, int U, bool EPHI>
__global__ __launch_bounds__(PUSH_BLOCK, sizeof(T) == 4 && U == 2 && !IDS && !PE ? PUSH_MIN_WAVES_F32 : 1)
void k_push(PushArgs a)
{
	typedef typename V2<T>::t T2;
	extern __shared__ unsigned char smem[];
	__shared__ int s_kcount;
	__shared__ int s_w;
	__shared__ LvStage<T, IDS, PE> s_lv;

	// one (tile, species) bucket per block: the tile's species run as
	// separate blocks (fewer chunks per block than one block per tile: a
	// stream of this access pattern measured 1.60 ms at 8 chunks per block
	// and 1.55 ms at 4; the push 2.4 % faster in FP32, 3.4 % in FP64)
	const int unit = a.order[blockIdx.x];
	const int pl = find_plasma(a.pre_bucket, a.P, unit);
	// the plasma's table entry, read once into registers (a reference into
	// global memory would be re-read after every particle store); the
	// per-species arrays are read through gp
	const Geo *gp = &a.geo[pl];
	const Geo g = *gp;
	T2 *sE = (T2 *) smem;
	double *sR = (double *) (smem + sizeof(T2) * g.W * g.H);

	const int b = unit;
	const int tile = (b - g.bucket0) / g.ns;
	const int s = b - g.bucket0 - tile * g.ns;
	int tix = tile & (g.ntx - 1), tiy = tile >> g.lgntx;
	int ox = tix * g.tx - g.M, oy = tiy * g.ty - g.M;
	int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	const T2 *__restrict__ E = (const T2 *) a.E + g.node0;
	const double *__restrict__ phi = EPHI ? a.phi + g.node0 : (const double *) NULL;
	double *__restrict__ rho = a.rho + g.node0;
	const double dtp = a.first ? -g.dt / 2 : g.dt;
	const bool set_r = !a.first;
	const float dtdx = (float) (dtp / g.dx), dtdy = (float) (dtp / g.dy);
	// read here, used at the end: its load latency then overlaps the work
	const long long iter = a.diag ? a.st->iter : -1;

	if(EPHI)
		tile_setup_phi<T>(g, ox, oy, phi, sE, sR);
	else
		tile_setup<T>(g, ox, oy, E, sE, sR);
	if(tid == 0)
	{
		s_kcount = 0;
		s_w = 0;
		s_lv.n = 0;
		s_lv.cut = LV_STAGE;
	}
	__syncthreads();

	double ke = 0.0, mpx = 0.0, mpy = 0.0;
	bool bad = false, nan_ = false;

	const int64_t bstart = uniform(a.bk.start[b]);
	const int cap = uniform(a.bk.cap[b]), res = uniform(a.bk.res[b]);
	const int n1 = min(uniform(a.bk.n1[b]), cap - res);
	const int ntot = n1 + uniform(a.bk.n2[b]);
	const int seg2 = cap - res - n1;      // slot = j + seg2 for j >= n1
	const SpeciesK<T> sk = species_k<T>(*gp, s, dtp);

	for(int base = 0; base < ntot; base += U * PUSH_BLOCK)
	{
		PState<T> q[U];
		bool act[U], leave[U];
		int ndst[U];
		load_chunk<T, IDS, U, PUSH_BLOCK>(a.p, bstart, n1, seg2, ntot, base, tid, q, act);
#pragma unroll
		for(int u = 0; u < U; u++)
		{
			leave[u] = false;
			ndst[u] = 0;
			if(act[u])
				leave[u] = push_one<T, EPHI>(g, sE, E, phi, sR, rho, tile, ox, oy, s, sk, set_r,
						dtp, dtdx, dtdy, q[u], ndst[u], ke, mpx, mpy, bad, nan_);
		}
		// leavers: staged in LDS (one LDS atomic per wavefront); once
		// the stage is full, placed directly (one global atomic each)
#pragma unroll
		for(int u = 0; u < U; u++)
		{
			unsigned long long lm = __ballot(leave[u]);
			if(lm)
			{
				int lcnt = __popcll(lm);
				int sb = 0;
				if(lane == 0)
					sb = atomicAdd(&s_lv.n, lcnt);
				sb = __builtin_amdgcn_readfirstlane(sb);
				if(sb + lcnt <= LV_STAGE)
				{
					if(leave[u])
					{
						int k = sb + lanes_below(lm);
						s_lv.x[k] = q[u].x;
						s_lv.y[k] = q[u].y;
						s_lv.ux[k] = q[u].ux;
						s_lv.uy[k] = q[u].uy;
						s_lv.dst[k] = ndst[u];
						if(IDS)
							s_lv.id[k] = q[u].id;
						if(PE)
						{
							s_lv.Ex[k] = q[u].Ex;
							s_lv.Ey[k] = q[u].Ey;
						}
					}
				}
				else
				{
					// records [sb, LV_STAGE) stay unused; later
					// wavefronts all start at or beyond LV_STAGE
					if(lane == 0 && sb < LV_STAGE)
						s_lv.cut = sb;
					int lbase = 0;
					if(lane == 0)
						lbase = atomicAdd(&a.lv.count[a.par], lcnt);
					lbase = __builtin_amdgcn_readfirstlane(lbase);
					if(leave[u])
					{
						int64_t ls = (int64_t) lbase + lanes_below(lm);
						if(ls < a.lv.cap)
							store_l<T>(a.lv, ls, q[u], IDS, PE, ndst[u], b);
						else
						{
							// no room: stays in this bucket (misplaced)
							leave[u] = false;
							back_to_tile<T>(g, ndst[u], ox, oy, q[u]);
							atomicAdd(&a.ctr->misplaced, 1ull);
						}
					}
					unsigned long long placed = __ballot(leave[u]);
					if(lane == 0 && placed)
						atomicAdd(&s_kcount, __popcll(placed));
				}
			}
		}
		// stayers: one LDS atomic per wavefront and chunk claims their
		// slots in the other buffer for all U rounds (order within the
		// bucket is not kept)
		unsigned long long smks[U];
		int wbase = 0, wtot = 0;
#pragma unroll
		for(int u = 0; u < U; u++)
		{
			smks[u] = __ballot(act[u] && !leave[u]);
			wtot += __popcll(smks[u]);
		}
		if(wtot)
		{
			if(lane == 0)
				wbase = atomicAdd(&s_w, wtot);
			wbase = __builtin_amdgcn_readfirstlane(wbase);
		}
#pragma unroll
		for(int u = 0; u < U; u++)
		{
			unsigned long long smk = smks[u];
			int wb = wbase;
			wbase += __popcll(smk);
			if(smk)
			{
				if(act[u] && !leave[u])
				{
					int jj = wb + lanes_below(smk);
					GLB_AS T *r = at_byte(bucket_base<T>(a.q, bstart), pslot_byte<T>(jj));
					r[0] = q[u].x;
					r[PGROUP] = q[u].y;
					r[2 * PGROUP] = q[u].ux;
					r[3 * PGROUP] = q[u].uy;
					if(IDS)
						a.q.id[bstart + jj] = q[u].id;
					if(PE)
					{
						a.q.Ex[bstart + jj] = q[u].Ex;
						a.q.Ey[bstart + jj] = q[u].Ey;
					}
				}
			}
		}
	}

	// the bucket's staged leavers: one global atomic claims their slots
	__syncthreads();
	if(tid == 0)
	{
		int nst = min(s_lv.n, s_lv.cut);
		s_lv.n = nst;
		s_lv.base = nst > 0 ? atomicAdd(&a.lv.count[a.par], nst) : 0;
		if(nst > 0)
		{
			int64_t room = a.lv.cap - (int64_t) s_lv.base;
			s_kcount += (int) (room <= 0 ? 0 : room < nst ? room : nst);
		}
	}
	__syncthreads();
	{
		const int nst = s_lv.n;
		const int64_t lb0 = s_lv.base;
		for(int k = tid; k < nst; k += PUSH_BLOCK)
		{
			PState<T> lq;
			lq.x = s_lv.x[k];
			lq.y = s_lv.y[k];
			lq.ux = s_lv.ux[k];
			lq.uy = s_lv.uy[k];
			lq.id = IDS ? s_lv.id[k] : 0;
			lq.Ex = PE ? s_lv.Ex[k] : 0.0;
			lq.Ey = PE ? s_lv.Ey[k] : 0.0;
			int nd = s_lv.dst[k];
			if(lb0 + k < a.lv.cap)
				store_l<T>(a.lv, lb0 + k, lq, IDS, PE, nd, b);
			else
			{
				// no room: stays in this bucket (misplaced)
				back_to_tile<T>(g, nd, ox, oy, lq);
				int w = atomicAdd(&s_w, 1);
				store_p<T>(a.q, bstart + w, lq, IDS, PE);
				atomicAdd(&a.ctr->misplaced, 1ull);
			}
		}
	}
	__syncthreads();
	if(tid == 0)
	{
		a.bk.n1[b] = s_w;
		a.bk.res[b] = s_kcount;
		a.bk.n2[b] = 0;
	}

	if(bad)
		atomicOr(a.derr, CPIC_DERR_UMAX);
	if(nan_)
		atomicOr(a.derr, CPIC_DERR_NAN);

	tile_flush<T>(g, ox, oy, sR, rho);

	if(a.diag && blockIdx.x == 0 && tid == 0)
		*a.diag_iter = iter;
	if(iter >= 0)
	{
		// wavefront sums, then one of the plasma's DIAG_SPREAD accumulators
		// per wavefront (k_append adds them into the plasma's slot):
		// atomics from every block onto the slot itself serialise, 1 % of
		// the push with a block per tile and 60 % with a block per bucket
		ke = wave_sum_lane63(ke);
		mpx = wave_sum_lane63(mpx);
		mpy = wave_sum_lane63(mpy);
		if(lane == 63)
		{
			int slot = (blockIdx.x * (PUSH_BLOCK / 64) + wid) & (DIAG_SPREAD - 1);
			double *acc = a.dpart + ((int64_t) pl * DIAG_SPREAD + slot) * DIAG_LINE;
			unsafeAtomicAdd(&acc[0], ke);
			unsafeAtomicAdd(&acc[1], mpx);
			unsafeAtomicAdd(&acc[2], mpy);
		}
	}
}

// Adds the push's DIAG_SPREAD accumulators of plasma p into its diagnostic
// slot of iteration iter and clears them for the next push (one full
// wavefront, one lane per accumulator).
static __device__ __forceinline__ void diag_fold(const Geo &g, int p, double *__restrict__ dpart,
		DiagSlot *__restrict__ diag, int ring, long long iter, int lane)
{
	static_assert(DIAG_SPREAD == 64, "one lane per accumulator");
	double *acc = dpart + ((int64_t) p * DIAG_SPREAD + lane) * DIAG_LINE;
	double k = acc[0], x = acc[1], y = acc[2];
	acc[0] = acc[1] = acc[2] = 0.0;
	k = wave_sum_lane63(k);
	x = wave_sum_lane63(x);
	y = wave_sum_lane63(y);
	if(lane == 63)
	{
		DiagSlot *ds = &diag[g.diag0 + iter % ring];
		ds->kinetic += k;
		ds->px += x;
		ds->py += y;
	}
}

// stage_field_rho from stored positions: rho_reset + interpolate_p2f_rho
// (src/field.c:164-230, src/interpolate.c:283-346).
template <typename T>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_deposit(PushArgs a)
{
	extern __shared__ unsigned char smem[];
	int gt = a.order[blockIdx.x];
	const Geo *gp = &a.geo[find_plasma(a.pre_tile, a.P, gt)];
	const Geo g = *gp;
	double *sR = (double *) smem;
	double *__restrict__ rho = a.rho + g.node0;
	int tile = gt - g.tile0;
	int ox = (tile & (g.ntx - 1)) * g.tx - g.M, oy = (tile >> g.lgntx) * g.ty - g.M;
	tile_setup<T>(g, ox, oy, (const typename V2<T>::t *) NULL, (typename V2<T>::t *) NULL, sR);
	__syncthreads();
	for(int s = 0; s < g.ns; s++)
	{
		int b = g.bucket0 + tile * g.ns + s;
		int n1 = bucket_n1(a.bk, b), ntot = n1 + a.bk.n2[b];
		T vq = (T) (-gp->q[s] / g.e0);
		for(int j = threadIdx.x; j < ntot; j += CPIC_BLOCK)
		{
			int64_t slot = bucket_slot(a.bk, b, n1, j);
			CellPos c;
			T rx, ry;
			locate(g, ox, oy, p_x<T>(a.p, slot), p_y<T>(a.p, slot), c, rx, ry);
			deposit<T>(g, sR, rho, c, rx, ry, vq);
		}
	}
	__syncthreads();
	tile_flush<T>(g, ox, oy, sR, rho);
}

// interpolate_f2p_E (src/interpolate.c:350-404): gathered E stored per
// particle (only with CPIC_F_PARTICLE_E).
template <typename T>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_gather(PushArgs a)
{
	typedef typename V2<T>::t T2;
	extern __shared__ unsigned char smem[];
	int gt = a.order[blockIdx.x];
	const Geo g = a.geo[find_plasma(a.pre_tile, a.P, gt)];
	T2 *sE = (T2 *) smem;
	const T2 *__restrict__ E = (const T2 *) a.E + g.node0;
	int tile = gt - g.tile0;
	int ox = (tile & (g.ntx - 1)) * g.tx - g.M, oy = (tile >> g.lgntx) * g.ty - g.M;
	tile_setup<T>(g, ox, oy, E, sE, (double *) NULL);
	__syncthreads();
	for(int s = 0; s < g.ns; s++)
	{
		int b = g.bucket0 + tile * g.ns + s;
		int n1 = bucket_n1(a.bk, b), ntot = n1 + a.bk.n2[b];
		for(int j = threadIdx.x; j < ntot; j += CPIC_BLOCK)
		{
			int64_t slot = bucket_slot(a.bk, b, n1, j);
			CellPos c;
			T rx, ry, ex, ey;
			locate(g, ox, oy, p_x<T>(a.p, slot), p_y<T>(a.p, slot), c, rx, ry);
			gather<T, false>(g, sE, E, (const double *) NULL, c, rx, ry, ex, ey);
			a.p.Ex[slot] = (double) ex;
			a.p.Ey[slot] = (double) ey;
		}
	}
}

// particles per thread and chunk in k_push: 2 by default (2 and 4 measure
// within 1-2 % at 2048^2; 2 leaves room for the FP32 occupancy bound above),
// 4 on request (cpic_dev_desc_t.ppt, chosen by the autotuner)

// Places leavers in their destination bucket, or back in the reserved
// slots of their source bucket when the destination is full.  The last
// kernel of a step: it also advances the device step counter, and with diag
// adds the push's kinetic and momentum accumulators into each plasma's
// diagnostic slot (of the iteration the push recorded, so not racing the
// counter's advance).
template <typename T, bool IDS, bool PE>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_append(const Geo *__restrict__ geo, const int *__restrict__ pre_bucket, int P,
		Buckets bk, Parts p, Leavers lv, int par, Counters *ctr, StepState *st,
		double *__restrict__ dpart, const long long *__restrict__ diag_iter,
		DiagSlot *__restrict__ diag, int ring)
{
	if(diag && threadIdx.x < 64)
	{
		const long long it = *diag_iter;
		if(it >= 0)
			for(int pl = blockIdx.x; pl < P; pl += gridDim.x)
				diag_fold(geo[pl], pl, dpart, diag, ring, it, threadIdx.x);
	}
	int64_t n = lv.count[par];
	if(n > lv.cap)
		n = lv.cap;
	if(blockIdx.x == 0 && threadIdx.x == 0)
	{
		lv.count[1 - par] = 0;
		st->iter += 1;
	}
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int d = lv.dst[i], s = lv.src[i];
		PState<T> q;
		q.x = ((const T *) lv.x)[i];
		q.y = ((const T *) lv.y)[i];
		q.ux = ((const T *) lv.ux)[i];
		q.uy = ((const T *) lv.uy)[i];
		if(IDS)
			q.id = lv.id[i];
		if(PE)
		{
			q.Ex = lv.Ex[i];
			q.Ey = lv.Ey[i];
		}
		int slot = atomicAdd(&bk.n1[d], 1);
		int64_t at;
		if(slot < bk.cap[d] - bk.res[d])
			at = bk.start[d] + slot;
		else
		{
			int s2 = atomicAdd(&bk.n2[s], 1);
			at = bk.start[s] + (bk.cap[s] - bk.res[s]) + s2;
			if(sizeof(T) == 4)
			{
				const Geo &g = geo[find_plasma(pre_bucket, P, d)];
				int td = (d - g.bucket0) / g.ns, ts = (s - g.bucket0) / g.ns;
				int dtx = ((td & (g.ntx - 1)) - (ts & (g.ntx - 1))) * g.tx;
				int dty = ((td >> g.lgntx) - (ts >> g.lgntx)) * g.ty;
				if(dtx > g.nx / 2) dtx -= g.nx;
				if(dtx < -g.nx / 2) dtx += g.nx;
				if(dty > g.ny / 2) dty -= g.ny;
				if(dty < -g.ny / 2) dty += g.ny;
				q.x = (T) ((float) q.x + (float) dtx);
				q.y = (T) ((float) q.y + (float) dty);
			}
			atomicAdd(&ctr->fallbacks, 1ull);
		}
		store_p<T>(p, at, q, IDS, PE);
	}
	if(blockIdx.x == 0 && threadIdx.x == 0)
		atomicAdd(&ctr->leavers, (unsigned long long) n);
}

// ------------------------------------------------------------- rebuild ---

// Segment of staged upload element i: segpre[k] <= i < segpre[k+1].
static __device__ __forceinline__ int find_seg(const int64_t *__restrict__ segpre, int nseg,
		int64_t i)
{
	int lo = 0, hi = nseg - 1;
	while(lo < hi)
	{
		int mid = (lo + hi + 1) >> 1;
		if(segpre[mid] <= i)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

// Destination bucket of each particle in the current layout (species kept),
// skipping the species a plasma's bit mask in `skip` flags (they are being
// replaced by an upload).  One block per bucket of the batch.
template <typename T>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_hist_layout(const Geo *__restrict__ geo, const int *__restrict__ pre_bucket, int P,
		Buckets bk, Parts p, const int *__restrict__ skip, int *__restrict__ cnt)
{
	int b = blockIdx.x;
	int pl = find_plasma(pre_bucket, P, b);
	const Geo g = geo[pl];
	int lb = b - g.bucket0;
	int s = lb % g.ns, tile = lb / g.ns;
	if(skip[pl] & (1 << s))
		return;
	int tix = tile & (g.ntx - 1), tiy = tile >> g.lgntx;
	int ox = tix * g.tx - g.M, oy = tiy * g.ty - g.M;
	int n1 = bucket_n1(bk, b), ntot = n1 + bk.n2[b];
	for(int j = threadIdx.x; j < ntot; j += blockDim.x)
	{
		int64_t slot = bucket_slot(bk, b, n1, j);
		CellPos c;
		T rx, ry;
		locate(g, ox, oy, p_x<T>(p, slot), p_y<T>(p, slot), c, rx, ry);
		int nt = (c.iy >> g.lgty) * g.ntx + (c.ix >> g.lgtx);
		atomicAdd(&cnt[g.bucket0 + nt * g.ns + s], 1);
	}
}

// Staged uploads: segment k holds species seg_s[k] of plasma seg_p[k], its
// nk = segpre[k+1] - segpre[k] particles stored as x[nk], y[nk], ux[nk],
// uy[nk] from buf + 4*segpre[k]; ids (nullable) at id + segpre[k].
struct Staged {
	const int64_t *segpre;
	const int *seg_p, *seg_s;
	int nseg;
	int64_t n;
	const double *buf;
	const int64_t *id;
};

struct StagedAt {
	int seg;
	int64_t nk, o;       // segment length, offset in the segment
	const double *b;     // segment base
};

static __device__ __forceinline__ StagedAt staged_at(const Staged &u, int64_t i)
{
	StagedAt r;
	r.seg = find_seg(u.segpre, u.nseg, i);
	int64_t s0 = u.segpre[r.seg];
	r.nk = u.segpre[r.seg + 1] - s0;
	r.o = i - s0;
	r.b = u.buf + 4 * s0;
	return r;
}

__global__ __launch_bounds__(CPIC_BLOCK)
void k_hist_stage(const Geo *__restrict__ geo, Staged u, int *__restrict__ cnt)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < u.n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		StagedAt a = staged_at(u, i);
		const Geo &g = geo[u.seg_p[a.seg]];
		CellPos c;
		double rx, ry;
		locate(g, 0, 0, a.b[a.o], a.b[a.nk + a.o], c, rx, ry);
		int nt = (c.iy >> g.lgty) * g.ntx + (c.ix >> g.lgtx);
		atomicAdd(&cnt[g.bucket0 + nt * g.ns + u.seg_s[a.seg]], 1);
	}
}

// Capacities and exclusive scan of bucket starts (one workgroup).  With cap
// given, a bucket that would need more than CPIC_BUCKET_MAX slots makes the
// total -1.
__global__ __launch_bounds__(1024)
void k_scan_caps(const int *__restrict__ cnt, int nb, double slack, int pad, int align,
		int64_t *__restrict__ start, int *__restrict__ cap, int64_t *__restrict__ total)
{
	__shared__ int64_t s[1024];
	__shared__ int64_t carry;
	__shared__ int big;
	if(threadIdx.x == 0)
	{
		carry = 0;
		big = 0;
	}
	__syncthreads();
	for(int base = 0; base < nb; base += 1024)
	{
		int b = base + threadIdx.x;
		int64_t c = 0;
		if(b < nb)
		{
			int64_t n = cnt[b];
			c = n + (int64_t) ceil((double) n * slack) + pad;
			c = (c + align - 1) / align * align;
			if(cap)
			{
				if(c > CPIC_BUCKET_MAX)
					big = 1;
				cap[b] = (int) (c > CPIC_BUCKET_MAX ? CPIC_BUCKET_MAX : c);
			}
		}
		s[threadIdx.x] = c;
		__syncthreads();
		for(int off = 1; off < 1024; off <<= 1)
		{
			int64_t v = threadIdx.x >= (unsigned) off ? s[threadIdx.x - off] : 0;
			__syncthreads();
			s[threadIdx.x] += v;
			__syncthreads();
		}
		if(b < nb)
			start[b] = carry + s[threadIdx.x] - c;
		__syncthreads();
		if(threadIdx.x == 1023)
			carry += s[1023];
		__syncthreads();
	}
	if(threadIdx.x == 0)
		*total = big ? -1 : carry;
}

template <typename T, bool IDS, bool PE>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_scatter_layout(const Geo *__restrict__ geo, const int *__restrict__ pre_bucket, int P,
		Buckets bk, Parts p, const int *__restrict__ skip, const int64_t *__restrict__ nstart,
		int *__restrict__ fill, Parts np)
{
	int b = blockIdx.x;
	int pl = find_plasma(pre_bucket, P, b);
	const Geo g = geo[pl];
	int lb = b - g.bucket0;
	int s = lb % g.ns, tile = lb / g.ns;
	if(skip[pl] & (1 << s))
		return;
	int tix = tile & (g.ntx - 1), tiy = tile >> g.lgntx;
	int ox = tix * g.tx - g.M, oy = tiy * g.ty - g.M;
	int n1 = bucket_n1(bk, b), ntot = n1 + bk.n2[b];
	for(int j = threadIdx.x; j < ntot; j += blockDim.x)
	{
		int64_t slot = bucket_slot(bk, b, n1, j);
		PState<T> q;
		load_p<T>(p, slot, q, IDS);
		if(PE)
		{
			q.Ex = p.Ex[slot];
			q.Ey = p.Ey[slot];
		}
		CellPos c;
		T rx, ry;
		locate(g, ox, oy, q.x, q.y, c, rx, ry);
		int nt = (c.iy >> g.lgty) * g.ntx + (c.ix >> g.lgtx);
		int nb_ = g.bucket0 + nt * g.ns + s;
		if(sizeof(T) == 4)
		{
			q.x = (T) ((float) (c.ix & (g.tx - 1)) + (float) rx);
			q.y = (T) ((float) (c.iy & (g.ty - 1)) + (float) ry);
		}
		int k = atomicAdd(&fill[nb_], 1);
		store_p<T>(np, nstart[nb_] + k, q, IDS, PE);
	}
}

template <typename T, bool IDS, bool PE>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_scatter_stage(const Geo *__restrict__ geo, Staged u, const int64_t *__restrict__ nstart,
		int *__restrict__ fill, Parts np)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < u.n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		StagedAt a = staged_at(u, i);
		const Geo &g = geo[u.seg_p[a.seg]];
		int s = u.seg_s[a.seg];
		double xi = a.b[a.o], yi = a.b[a.nk + a.o];
		CellPos c;
		double rx, ry;
		locate(g, 0, 0, xi, yi, c, rx, ry);
		int nt = (c.iy >> g.lgty) * g.ntx + (c.ix >> g.lgtx);
		int nb_ = g.bucket0 + nt * g.ns + s;
		PState<T> q;
		if(sizeof(T) == 8)
		{
			q.x = (T) xi;
			q.y = (T) yi;
		}
		else
		{
			q.x = (T) ((float) (c.ix & (g.tx - 1)) + (float) rx);
			q.y = (T) ((float) (c.iy & (g.ty - 1)) + (float) ry);
		}
		q.ux = (T) a.b[2 * a.nk + a.o];
		q.uy = (T) a.b[3 * a.nk + a.o];
		q.id = IDS && u.id ? u.id[i] : 0;
		q.Ex = q.Ey = 0.0;
		int k = atomicAdd(&fill[nb_], 1);
		store_p<T>(np, nstart[nb_] + k, q, IDS, PE);
	}
}

// ------------------------------------------------------------ download ---

__global__ __launch_bounds__(CPIC_BLOCK)
void k_species_counts(Buckets bk, int bucket0, int ns, int s, int ntiles, int *__restrict__ cnt)
{
	for(int t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x)
	{
		int b = bucket0 + t * ns + s;
		cnt[t] = bucket_n1(bk, b) + bk.n2[b];
	}
}

template <typename T, bool IDS, bool PE>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_download(Geo g, Buckets bk, Parts p, int s, const int64_t *__restrict__ off,
		double *__restrict__ out /* 6 arrays of n */, int64_t *__restrict__ oid,
		int64_t *__restrict__ ocell, int64_t n)
{
	int tile = blockIdx.x;
	int b = g.bucket0 + tile * g.ns + s;
	int tix = tile & (g.ntx - 1), tiy = tile >> g.lgntx;
	int ox = tix * g.tx - g.M, oy = tiy * g.ty - g.M;
	int n1 = bucket_n1(bk, b), ntot = n1 + bk.n2[b];
	for(int j = threadIdx.x; j < ntot; j += blockDim.x)
	{
		int64_t slot = bucket_slot(bk, b, n1, j);
		int64_t o = off[tile] + j;
		PState<T> q;
		load_p<T>(p, slot, q, false);
		T x = q.x, y = q.y;
		CellPos c;
		T rx, ry;
		locate(g, ox, oy, x, y, c, rx, ry);
		double ax, ay;
		if(sizeof(T) == 8)
		{
			ax = (double) x;
			ay = (double) y;
		}
		else
		{
			ax = ((double) c.ix + (double) rx) * g.dx;
			ay = ((double) c.iy + (double) ry) * g.dy;
		}
		out[o] = ax;
		out[n + o] = ay;
		out[2 * n + o] = (double) q.ux;
		out[3 * n + o] = (double) q.uy;
		out[4 * n + o] = PE ? p.Ex[slot] : 0.0;
		out[5 * n + o] = PE ? p.Ey[slot] : 0.0;
		oid[o] = IDS ? p.id[slot] : -1;
		ocell[o] = (int64_t) c.iy * g.nx + c.ix;
	}
}

template <typename T>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_cell_counts(Geo g, Buckets bk, Parts p, int s, unsigned long long *__restrict__ counts)
{
	int tile = blockIdx.x;
	int b = g.bucket0 + tile * g.ns + s;
	int tix = tile & (g.ntx - 1), tiy = tile >> g.lgntx;
	int ox = tix * g.tx - g.M, oy = tiy * g.ty - g.M;
	int n1 = bucket_n1(bk, b), ntot = n1 + bk.n2[b];
	for(int j = threadIdx.x; j < ntot; j += blockDim.x)
	{
		int64_t slot = bucket_slot(bk, b, n1, j);
		CellPos c;
		T rx, ry;
		locate(g, ox, oy, p_x<T>(p, slot), p_y<T>(p, slot), c, rx, ry);
		atomicAdd(&counts[(int64_t) c.iy * g.nx + c.ix], 1ull);
	}
}

// ---------------------------------------------------------------- host ---


#define DISPATCH(kern, T, ...) do { \
		bool ids_ = (d->flags & CPIC_F_IDS) != 0; \
		bool pe_ = (d->flags & CPIC_F_PARTICLE_E) != 0; \
		if(ids_ && pe_) hipLaunchKernelGGL((kern<T, true, true>), __VA_ARGS__); \
		else if(ids_) hipLaunchKernelGGL((kern<T, true, false>), __VA_ARGS__); \
		else if(pe_) hipLaunchKernelGGL((kern<T, false, true>), __VA_ARGS__); \
		else hipLaunchKernelGGL((kern<T, false, false>), __VA_ARGS__); \
	} while(0)

#define DISPATCH_UU(kern, T, U, EP, ...) do { \
		bool ids_ = (d->flags & CPIC_F_IDS) != 0; \
		bool pe_ = (d->flags & CPIC_F_PARTICLE_E) != 0; \
		if(ids_ && pe_) hipLaunchKernelGGL((kern<T, true, true, U, EP>), __VA_ARGS__); \
		else if(ids_) hipLaunchKernelGGL((kern<T, true, false, U, EP>), __VA_ARGS__); \
		else if(pe_) hipLaunchKernelGGL((kern<T, false, true, U, EP>), __VA_ARGS__); \
		else hipLaunchKernelGGL((kern<T, false, false, U, EP>), __VA_ARGS__); \
	} while(0)

// particles per thread and chunk: 2 (default) or 4, chosen per handle; E
// from the potential when the E array is stale
#define DISPATCH_U(kern, T, ...) do { \
		if(d->ppt == 4 && d->E_stale) DISPATCH_UU(kern, T, 4, true, __VA_ARGS__); \
		else if(d->ppt == 4) DISPATCH_UU(kern, T, 4, false, __VA_ARGS__); \
		else if(d->E_stale) DISPATCH_UU(kern, T, 2, true, __VA_ARGS__); \
		else DISPATCH_UU(kern, T, 2, false, __VA_ARGS__); \
	} while(0)

#define DISPATCH_P(kern, ...) do { \
		if(d->prec == 64) DISPATCH(kern, double, __VA_ARGS__); \
		else DISPATCH(kern, float, __VA_ARGS__); \
	} while(0)

#define PRE(d, k) ((d)->pre + (k) * ((d)->P + 1))

static size_t
elem(const cpic_dev *d)
{
	return d->prec == 64 ? sizeof(double) : sizeof(float);
}

// largest LDS tile region of the batch, in nodes
static size_t
max_wh(const cpic_dev *d)
{
	size_t m = 0;
	for(int p = 0; p < d->P; p++)
	{
		size_t n = (size_t) d->hgeo[p].W * d->hgeo[p].H;
		if(n > m)
			m = n;
	}
	return m;
}

// E region plus the rho region, which also holds the potential with its
// one-node halo while a push computes E from it
size_t
cpic_push_lds(const cpic_dev *d)
{
	size_t n = 0, h = 0;
	for(int p = 0; p < d->P; p++)
	{
		size_t w = (size_t) d->hgeo[p].W, hh = (size_t) d->hgeo[p].H;
		if(w * hh > n)
			n = w * hh;
		if((w + 2) * (hh + 2) > h)
			h = (w + 2) * (hh + 2);
	}
	return n * (d->prec == 64 ? sizeof(double2) : sizeof(float2)) + h * sizeof(double);
}

void
cpic_parts_free(cpic_dev *d, int which)
{
	Parts *p = &d->p[which];
	hipFree(p->r);
	hipFree(p->id); hipFree(p->Ex); hipFree(p->Ey);
	memset(p, 0, sizeof(*p));
	d->pcap[which] = 0;
}

// Fault injection for the test of a failed re-sort: while the environment
// variable CPIC_FAIL_PARTS_ALLOC is "1", particle-buffer allocations fail as
// if the device were out of memory (after the old buffer is freed, as a real
// failure would leave it).
static int
parts_alloc_fails(void)
{
	const char *e = getenv("CPIC_FAIL_PARTS_ALLOC");
	return e && e[0] == '1' && e[1] == 0;
}

int
cpic_parts_alloc(cpic_dev *d, int which, int64_t cap)
{
	Parts *p = &d->p[which];
	size_t e = elem(d);
	cap = (cap < PGROUP ? PGROUP : cap + PGROUP - 1) / PGROUP * PGROUP;
	cpic_parts_free(d, which);
	if(parts_alloc_fails() || hipMalloc(&p->r, 4 * e * cap) != hipSuccess)
		goto fail;
	if((d->flags & CPIC_F_IDS) && hipMalloc(&p->id, sizeof(int64_t) * cap) != hipSuccess)
		goto fail;
	if((d->flags & CPIC_F_PARTICLE_E) &&
			(hipMalloc(&p->Ex, sizeof(double) * cap) != hipSuccess ||
			 hipMalloc(&p->Ey, sizeof(double) * cap) != hipSuccess))
		goto fail;
	d->pcap[which] = cap;
	return 0;
fail:
	cpic_parts_free(d, which);
	return CPIC_ENOMEM;
}

static PushArgs
push_args(cpic_dev *d, int first, double *rho)
{
	PushArgs a;
	memset(&a, 0, sizeof(a));
	a.geo = d->geo;
	a.pre_tile = PRE(d, PRE_TILE);
	a.pre_bucket = PRE(d, PRE_BUCKET);
	a.order = d->order;
	a.P = d->P;
	a.ntiles = d->ntiles;
	a.bk = d->bk;
	a.p = d->p[d->pcur];
	a.q = d->p[1 - d->pcur];
	a.lv = d->lv;
	a.par = d->step_parity;
	a.first = first;
	a.E = d->E;
	a.phi = d->phi;
	a.rho = rho;
	a.diag = (d->flags & CPIC_F_NO_DIAG) ? (DiagSlot *) NULL : d->diag;
	a.dpart = d->dpart;
	a.diag_iter = (long long *) (d->dpart + (size_t) d->P * DIAG_SPREAD * DIAG_LINE);
	a.ring = d->ring;
	a.st = d->st;
	a.derr = d->derr;
	a.ctr = d->ctr;
	return a;
}

__global__ void k_set_iter(StepState *st, long long iter)
{
	if(blockIdx.x == 0 && threadIdx.x == 0)
		st->iter = iter;
}

// Sets the device step counter with a kernel (also valid inside a capture).
int
cpic_set_iter(cpic_dev *d, int64_t iter)
{
	hipLaunchKernelGGL(k_set_iter, dim3(1), dim3(64), 0, d->stream, d->st, (long long) iter);
	HIPCHK(hipGetLastError());
	d->iter = iter;
	return 0;
}

int
cpic_zero(cpic_dev *d, double *p, int64_t n)
{
	int blocks = (int) ((n + CPIC_BLOCK - 1) / CPIC_BLOCK);
	if(blocks > 4096)
		blocks = 4096;
	int slot = cpic_kt_begin(d, "zero_rho");
	hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(CPIC_BLOCK), 0, d->stream, p, n);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	return 0;
}

int
cpic_parts_push(cpic_dev *d, int first_iter)
{
	PushArgs a = push_args(d, first_iter, d->rho[1 - d->rho_cur]);
	int slot;

	a.order = d->border;
	slot = cpic_kt_begin(d, "push_gather_deposit");
	if(d->prec == 64)
		DISPATCH_U(k_push, double, dim3(d->nb), dim3(PUSH_BLOCK), d->lds_push, d->stream, a);
	else
		DISPATCH_U(k_push, float, dim3(d->nb), dim3(PUSH_BLOCK), d->lds_push, d->stream, a);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	d->pcur = 1 - d->pcur;

	slot = cpic_kt_begin(d, "append_leavers");
	DISPATCH_P(k_append, dim3(512), dim3(CPIC_BLOCK), 0, d->stream, d->geo, PRE(d, PRE_BUCKET),
			d->P, d->bk, d->p[d->pcur], d->lv, d->step_parity, d->ctr, d->st, d->dpart,
			(const long long *) a.diag_iter, a.diag, d->ring);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	d->step_parity ^= 1;
	d->iter++;
	return 0;
}

int
cpic_parts_deposit(cpic_dev *d, double *rho_out)
{
	PushArgs a = push_args(d, 0, rho_out);
	int slot = cpic_kt_begin(d, "deposit");
	size_t lds = max_wh(d) * sizeof(double);
	if(d->prec == 64)
		hipLaunchKernelGGL(k_deposit<double>, dim3(d->ntiles), dim3(CPIC_BLOCK), lds, d->stream, a);
	else
		hipLaunchKernelGGL(k_deposit<float>, dim3(d->ntiles), dim3(CPIC_BLOCK), lds, d->stream, a);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	return 0;
}

int
cpic_parts_gather(cpic_dev *d)
{
	PushArgs a = push_args(d, 0, d->rho[d->rho_cur]);
	int slot = cpic_kt_begin(d, "gather");
	size_t lds = max_wh(d) * (d->prec == 64 ? sizeof(double2) : sizeof(float2));
	if(d->prec == 64)
		hipLaunchKernelGGL(k_gather<double>, dim3(d->ntiles), dim3(CPIC_BLOCK), lds, d->stream, a);
	else
		hipLaunchKernelGGL(k_gather<float>, dim3(d->ntiles), dim3(CPIC_BLOCK), lds, d->stream, a);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	return 0;
}

// Order in which the tile kernels' blocks take the batch's tiles.  With
// tiles of similar particle counts (max/min <= 1.5): the XCD-banded map
// above.  Otherwise largest first (longest-processing-time scheduling:
// the big tiles of a batch start early instead of forming the tail of the
// launch).  cnt[b] = particles of bucket b (host), or NULL for the banded map.
int
cpic_parts_order(cpic_dev *d, const int *cnt)
{
	const int nt = d->ntiles;
	int *ord = (int *) malloc(sizeof(int) * (size_t) nt);
	int64_t *tc = NULL;
	if(!ord)
		return CPIC_ENOMEM;
	for(int b = 0; b < nt; b++)
		ord[b] = tile_map(b, nt);
	if(cnt && (tc = (int64_t *) calloc((size_t) nt, sizeof(int64_t))))
	{
		int64_t lo = INT64_MAX, hi = 0;
		for(int p = 0; p < d->P; p++)
		{
			const Geo &g = d->hgeo[p];
			for(int t = 0; t < g.ntiles; t++)
			{
				int64_t c = 0;
				for(int s = 0; s < g.ns; s++)
					c += cnt[g.bucket0 + t * g.ns + s];
				tc[g.tile0 + t] = c;
				lo = c < lo ? c : lo;
				hi = c > hi ? c : hi;
			}
		}
		if(hi > lo + lo / 2)
			std::stable_sort(ord, ord + nt, [tc](int a, int b) { return tc[a] > tc[b]; });
		free(tc);
	}
	hipError_t e = hipMemcpy(d->order, ord, sizeof(int) * (size_t) nt, hipMemcpyHostToDevice);
	free(ord);
	if(e != hipSuccess)
		return CPIC_EHIP;
	// the same for the buckets (a push by buckets): XCD bands over buckets,
	// so the buckets of one tile run next to each other on one XCD, largest
	// first when the counts are uneven
	const int nb = d->nb;
	int *bord = (int *) malloc(sizeof(int) * (size_t) nb);
	if(!bord)
		return CPIC_ENOMEM;
	for(int b = 0; b < nb; b++)
		bord[b] = tile_map(b, nb);
	if(cnt)
	{
		int lo = INT_MAX, hi = 0;
		for(int b = 0; b < nb; b++)
		{
			lo = cnt[b] < lo ? cnt[b] : lo;
			hi = cnt[b] > hi ? cnt[b] : hi;
		}
		if(hi > lo + lo / 2)
			std::stable_sort(bord, bord + nb, [cnt](int a, int b) { return cnt[a] > cnt[b]; });
	}
	e = hipMemcpy(d->border, bord, sizeof(int) * (size_t) nb, hipMemcpyHostToDevice);
	free(bord);
	return e == hipSuccess ? 0 : CPIC_EHIP;
}

// Full re-sort of the batch: counts per destination bucket (current layout
// + pending uploads), fresh capacities, scatter into the spare particle
// arrays.  Pending uploads replace their (plasma, species) wholesale.
// In the re-sort, a failing HIP call is named on stderr and leaves through
// the cleanup (which frees the temporaries and marks a half-done re-sort).
#define RB_CHK(call) do { \
		hipError_t e_ = (call); \
		if(e_ != hipSuccess) { \
			fprintf(stderr, "cpic_hip: %s failed: %s (%s:%d)\n", #call, \
					hipGetErrorString(e_), __FILE__, __LINE__); \
			ret = CPIC_EHIP; \
			goto out; \
		} \
	} while(0)

int
cpic_parts_rebuild(cpic_dev *d)
{
	const int P = d->P, MS = CPIC_MAX_SPECIES;
	int nb = d->nb, p, s, k;
	int *cnt = NULL, *fill = NULL, *ncap = NULL, *dskip = NULL;
	int64_t *nstart = NULL, *total = NULL, tot = 0;
	double *stage = NULL;
	int64_t *sid = NULL, *dsegpre = NULL;
	int *dseg = NULL;
	int ret = CPIC_ENOMEM, nseg = 0;
	int touched = 0;         // the particle buffers have been changed
	int *skip = (int *) calloc((size_t) P, sizeof(int));
	int64_t *segpre = (int64_t *) malloc(sizeof(int64_t) * (size_t) (P * MS + 1));
	int *seg = (int *) malloc(sizeof(int) * 2 * (size_t) (P * MS));
	Staged u;

	// captured steps name particle arrays that this may free
	d->graph_ok = 0;
	if(!skip || !segpre || !seg)
		goto out;
	segpre[0] = 0;
	for(p = 0; p < P; p++)
		for(s = 0; s < MS; s++)
		{
			if(!d->up_buf[p * MS + s])
				continue;
			skip[p] |= 1 << s;
			if(d->up_n[p * MS + s] == 0)
				continue;
			seg[nseg] = p;
			seg[P * MS + nseg] = s;
			segpre[nseg + 1] = segpre[nseg] + d->up_n[p * MS + s];
			nseg++;
		}

	RB_CHK(hipStreamSynchronize(d->stream));
	if(hipMalloc(&cnt, sizeof(int) * nb) != hipSuccess ||
			hipMalloc(&fill, sizeof(int) * nb) != hipSuccess ||
			hipMalloc(&ncap, sizeof(int) * nb) != hipSuccess ||
			hipMalloc(&nstart, sizeof(int64_t) * nb) != hipSuccess ||
			hipMalloc(&total, sizeof(int64_t)) != hipSuccess ||
			hipMalloc(&dskip, sizeof(int) * P) != hipSuccess)
		goto out;
	RB_CHK(hipMemsetAsync(cnt, 0, sizeof(int) * nb, d->stream));
	RB_CHK(hipMemsetAsync(fill, 0, sizeof(int) * nb, d->stream));
	RB_CHK(hipMemcpyAsync(dskip, skip, sizeof(int) * P, hipMemcpyHostToDevice, d->stream));

	if(d->pcap[d->pcur] > 0)
	{
		if(d->prec == 64)
			hipLaunchKernelGGL(k_hist_layout<double>, dim3(nb), dim3(CPIC_BLOCK), 0, d->stream,
					d->geo, PRE(d, PRE_BUCKET), P, d->bk, d->p[d->pcur], dskip, cnt);
		else
			hipLaunchKernelGGL(k_hist_layout<float>, dim3(nb), dim3(CPIC_BLOCK), 0, d->stream,
					d->geo, PRE(d, PRE_BUCKET), P, d->bk, d->p[d->pcur], dskip, cnt);
	}
	memset(&u, 0, sizeof(u));
	if(nseg > 0)
	{
		int64_t n = segpre[nseg];
		if(hipMalloc(&stage, sizeof(double) * 4 * n) != hipSuccess ||
				hipMalloc(&dsegpre, sizeof(int64_t) * (nseg + 1)) != hipSuccess ||
				hipMalloc(&dseg, sizeof(int) * 2 * nseg) != hipSuccess)
			goto out;
		if((d->flags & CPIC_F_IDS) && hipMalloc(&sid, sizeof(int64_t) * n) != hipSuccess)
			goto out;
		RB_CHK(hipMemcpyAsync(dsegpre, segpre, sizeof(int64_t) * (nseg + 1),
				hipMemcpyHostToDevice, d->stream));
		RB_CHK(hipMemcpyAsync(dseg, seg, sizeof(int) * nseg, hipMemcpyHostToDevice, d->stream));
		RB_CHK(hipMemcpyAsync(dseg + nseg, seg + P * MS, sizeof(int) * nseg,
				hipMemcpyHostToDevice, d->stream));
		for(k = 0; k < nseg; k++)
		{
			int i = seg[k] * MS + seg[P * MS + k];
			int64_t nk = segpre[k + 1] - segpre[k];
			RB_CHK(hipMemcpyAsync(stage + 4 * segpre[k], d->up_buf[i], sizeof(double) * 4 * nk,
					hipMemcpyHostToDevice, d->stream));
			if(sid && d->up_id[i])
				RB_CHK(hipMemcpyAsync(sid + segpre[k], d->up_id[i], sizeof(int64_t) * nk,
						hipMemcpyHostToDevice, d->stream));
		}
		u.segpre = dsegpre;
		u.seg_p = dseg;
		u.seg_s = dseg + nseg;
		u.nseg = nseg;
		u.n = n;
		u.buf = stage;
		u.id = sid;
		int blocks = (int) ((n + CPIC_BLOCK - 1) / CPIC_BLOCK);
		if(blocks > 8192) blocks = 8192;
		hipLaunchKernelGGL(k_hist_stage, dim3(blocks), dim3(CPIC_BLOCK), 0, d->stream, d->geo,
				u, cnt);
	}
	{
		// slack > 0: spare fraction; 0: default; < 0: no spare at all (tight
		// capacities, which exercises the fallback path)
		double slack = d->desc[0].slack > 0 ? d->desc[0].slack : 0.25;
		int pad = 64;
		if(d->desc[0].slack < 0)
		{
			slack = 0.0;
			pad = 0;
		}
		// capacities in whole groups: bucket starts are multiples of PGROUP
		hipLaunchKernelGGL(k_scan_caps, dim3(1), dim3(1024), 0, d->stream, cnt, nb, slack, pad,
				PGROUP, nstart, ncap, total);
		RB_CHK(hipGetLastError());
		RB_CHK(hipMemcpyAsync(&tot, total, sizeof(int64_t), hipMemcpyDeviceToHost, d->stream));
		RB_CHK(hipStreamSynchronize(d->stream));
		if(tot < 0)
		{
			fprintf(stderr, "cpic_hip: a (tile, species) bucket would need more than %d slots\n",
					CPIC_BUCKET_MAX);
			ret = CPIC_ENOMEM;
			goto out;
		}

		int spare = 1 - d->pcur;
		touched = 1;
		if(d->pcap[spare] < tot)
			if((ret = cpic_parts_alloc(d, spare, tot + tot / 8)))
				goto out;
		if(d->pcap[d->pcur] > 0)
			DISPATCH_P(k_scatter_layout, dim3(nb), dim3(CPIC_BLOCK), 0, d->stream,
					d->geo, PRE(d, PRE_BUCKET), P, d->bk, d->p[d->pcur], dskip, nstart, fill,
					d->p[spare]);
		if(nseg > 0)
		{
			int blocks = (int) ((u.n + CPIC_BLOCK - 1) / CPIC_BLOCK);
			if(blocks > 8192) blocks = 8192;
			DISPATCH_P(k_scatter_stage, dim3(blocks), dim3(CPIC_BLOCK), 0, d->stream,
					d->geo, u, nstart, fill, d->p[spare]);
		}
		RB_CHK(hipGetLastError());
		// new metadata: n1 = count, res = n2 = 0
		RB_CHK(hipMemcpyAsync(d->bk.start, nstart, sizeof(int64_t) * nb, hipMemcpyDeviceToDevice, d->stream));
		RB_CHK(hipMemcpyAsync(d->bk.cap, ncap, sizeof(int) * nb, hipMemcpyDeviceToDevice, d->stream));
		RB_CHK(hipMemcpyAsync(d->bk.n1, cnt, sizeof(int) * nb, hipMemcpyDeviceToDevice, d->stream));
		RB_CHK(hipMemsetAsync(d->bk.res, 0, sizeof(int) * nb, d->stream));
		RB_CHK(hipMemsetAsync(d->bk.n2, 0, sizeof(int) * nb, d->stream));
		RB_CHK(hipStreamSynchronize(d->stream));
		d->pcur = spare;
		// the push writes the other buffer: it needs the same capacity
		if(d->pcap[1 - spare] < d->pcap[spare])
			if((ret = cpic_parts_alloc(d, 1 - spare, d->pcap[spare])))
				goto out;
		// block order for the new counts
		int *hc = (int *) malloc(sizeof(int) * (size_t) nb);
		if(!hc)
		{
			ret = CPIC_ENOMEM;
			goto out;
		}
		if(hipMemcpy(hc, d->bk.n1, sizeof(int) * nb, hipMemcpyDeviceToHost) != hipSuccess)
		{
			free(hc);
			ret = CPIC_EHIP;
			goto out;
		}
		ret = cpic_parts_order(d, hc);
		free(hc);
		if(ret)
			goto out;
	}
	for(k = 0; k < P * MS; k++)
	{
		if(d->up_buf[k])
		{
			d->count[k] = d->up_n[k];
			free(d->up_buf[k]);
			free(d->up_id[k]);
			d->up_buf[k] = NULL;
			d->up_id[k] = NULL;
			d->up_n[k] = 0;
		}
	}
	d->pending = 0;
	d->nrebuilds++;
	d->graph_ok = 0;         // captured steps name the old particle arrays
	ret = 0;
out:
	if(ret && touched && !d->broken)
	{
		// the storage is half rewritten, or a buffer the push writes is gone
		fprintf(stderr, "cpic_hip: particle re-sort failed (%s); the handle is unusable\n",
				ret == CPIC_ENOMEM ? "out of device memory" : "HIP error");
		d->broken = ret;
	}
	hipFree(cnt); hipFree(fill); hipFree(ncap); hipFree(nstart); hipFree(total);
	hipFree(dskip); hipFree(stage); hipFree(sid); hipFree(dsegpre); hipFree(dseg);
	free(skip); free(segpre); free(seg);
	return ret;
}

int
cpic_parts_download(cpic_dev *d, int plasma, int s, cpic_particles_t *out)
{
	const Geo &g = d->hgeo[plasma];
	int nt = g.ntiles;
	int *cnt = NULL;
	int64_t *off = NULL, *tot = NULL, n = 0;
	double *buf = NULL;
	int64_t *oid = NULL, *ocell = NULL;
	double *h = NULL;
	int64_t *hid = NULL, *hcell = NULL;
	int ret = CPIC_ENOMEM;

	if(hipMalloc(&cnt, sizeof(int) * nt) != hipSuccess ||
			hipMalloc(&off, sizeof(int64_t) * nt) != hipSuccess ||
			hipMalloc(&tot, sizeof(int64_t)) != hipSuccess)
		goto out;
	hipLaunchKernelGGL(k_species_counts, dim3((nt + CPIC_BLOCK - 1) / CPIC_BLOCK), dim3(CPIC_BLOCK),
			0, d->stream, d->bk, g.bucket0, g.ns, s, nt, cnt);
	hipLaunchKernelGGL(k_scan_caps, dim3(1), dim3(1024), 0, d->stream, cnt, nt, 0.0, 0, 1,
			off, (int *) NULL, tot);
	HIPCHK(hipMemcpyAsync(&n, tot, sizeof(int64_t), hipMemcpyDeviceToHost, d->stream));
	HIPCHK(hipStreamSynchronize(d->stream));
	if(out->n < n && (out->x || out->y || out->ux || out->uy || out->id || out->Ex ||
				out->Ey || out->cell))
	{
		out->n = n;
		ret = CPIC_EINVAL;
		goto out;
	}
	if(n > 0)
	{
		if(hipMalloc(&buf, sizeof(double) * 6 * n) != hipSuccess ||
				hipMalloc(&oid, sizeof(int64_t) * n) != hipSuccess ||
				hipMalloc(&ocell, sizeof(int64_t) * n) != hipSuccess)
			goto out;
		DISPATCH_P(k_download, dim3(nt), dim3(CPIC_BLOCK), 0, d->stream, g, d->bk,
				d->p[d->pcur], s, off, buf, oid, ocell, n);
		HIPCHK(hipGetLastError());
		h = (double *) malloc(sizeof(double) * 6 * n);
		hid = (int64_t *) malloc(sizeof(int64_t) * n);
		hcell = (int64_t *) malloc(sizeof(int64_t) * n);
		if(!h || !hid || !hcell)
			goto out;
		HIPCHK(hipMemcpyAsync(h, buf, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, d->stream));
		HIPCHK(hipMemcpyAsync(hid, oid, sizeof(int64_t) * n, hipMemcpyDeviceToHost, d->stream));
		HIPCHK(hipMemcpyAsync(hcell, ocell, sizeof(int64_t) * n, hipMemcpyDeviceToHost, d->stream));
		HIPCHK(hipStreamSynchronize(d->stream));
		if(out->x) memcpy(out->x, h, sizeof(double) * n);
		if(out->y) memcpy(out->y, h + n, sizeof(double) * n);
		if(out->ux) memcpy(out->ux, h + 2 * n, sizeof(double) * n);
		if(out->uy) memcpy(out->uy, h + 3 * n, sizeof(double) * n);
		if(out->Ex) memcpy(out->Ex, h + 4 * n, sizeof(double) * n);
		if(out->Ey) memcpy(out->Ey, h + 5 * n, sizeof(double) * n);
		if(out->id) memcpy(out->id, hid, sizeof(int64_t) * n);
		if(out->cell) memcpy(out->cell, hcell, sizeof(int64_t) * n);
	}
	out->n = n;
	ret = 0;
out:
	hipFree(cnt); hipFree(off); hipFree(tot); hipFree(buf); hipFree(oid); hipFree(ocell);
	free(h); free(hid); free(hcell);
	return ret;
}

int
cpic_parts_cell_counts(cpic_dev *d, int plasma, int s, int64_t *counts)
{
	const Geo &g = d->hgeo[plasma];
	unsigned long long *c = NULL;
	int64_t n = (int64_t) g.nx * g.ny;
	if(hipMalloc(&c, sizeof(unsigned long long) * n) != hipSuccess)
		return CPIC_ENOMEM;
	hipMemsetAsync(c, 0, sizeof(unsigned long long) * n, d->stream);
	if(d->prec == 64)
		hipLaunchKernelGGL(k_cell_counts<double>, dim3(g.ntiles), dim3(CPIC_BLOCK), 0, d->stream,
				g, d->bk, d->p[d->pcur], s, c);
	else
		hipLaunchKernelGGL(k_cell_counts<float>, dim3(g.ntiles), dim3(CPIC_BLOCK), 0, d->stream,
				g, d->bk, d->p[d->pcur], s, c);
	hipError_t e = hipMemcpyAsync(counts, c, sizeof(int64_t) * n, hipMemcpyDeviceToHost, d->stream);
	hipStreamSynchronize(d->stream);
	hipFree(c);
	return e == hipSuccess ? 0 : CPIC_EHIP;
}

// Tiles up to 64x64 cells need up to 108 KB of LDS in FP64 mode.
int
cpic_parts_init(cpic_dev *d)
{
	(void) d;
	const void *k[] = {
		(const void *) k_push<double, false, false, 2, false>, (const void *) k_push<double, true, false, 2, false>,
		(const void *) k_push<double, false, true, 2, false>, (const void *) k_push<double, true, true, 2, false>,
		(const void *) k_push<float, false, false, 2, false>, (const void *) k_push<float, true, false, 2, false>,
		(const void *) k_push<float, false, true, 2, false>, (const void *) k_push<float, true, true, 2, false>,
		(const void *) k_push<double, false, false, 4, false>, (const void *) k_push<double, true, false, 4, false>,
		(const void *) k_push<double, false, true, 4, false>, (const void *) k_push<double, true, true, 4, false>,
		(const void *) k_push<float, false, false, 4, false>, (const void *) k_push<float, true, false, 4, false>,
		(const void *) k_push<float, false, true, 4, false>, (const void *) k_push<float, true, true, 4, false>,
		(const void *) k_push<double, false, false, 2, true>, (const void *) k_push<double, true, false, 2, true>,
		(const void *) k_push<double, false, true, 2, true>, (const void *) k_push<double, true, true, 2, true>,
		(const void *) k_push<float, false, false, 2, true>, (const void *) k_push<float, true, false, 2, true>,
		(const void *) k_push<float, false, true, 2, true>, (const void *) k_push<float, true, true, 2, true>,
		(const void *) k_push<double, false, false, 4, true>, (const void *) k_push<double, true, false, 4, true>,
		(const void *) k_push<double, false, true, 4, true>, (const void *) k_push<double, true, true, 4, true>,
		(const void *) k_push<float, false, false, 4, true>, (const void *) k_push<float, true, false, 4, true>,
		(const void *) k_push<float, false, true, 4, true>, (const void *) k_push<float, true, true, 4, true>,
		(const void *) k_deposit<double>, (const void *) k_deposit<float>,
		(const void *) k_gather<double>, (const void *) k_gather<float>,
	};
	for(unsigned i = 0; i < sizeof(k) / sizeof(k[0]); i++)
	{
		// dynamic LDS up to the limit, less the kernel's static LDS
		hipFuncAttributes fa;
		HIPCHK(hipFuncGetAttributes(&fa, k[i]));
		size_t lim = CPIC_LDS_LIMIT;
		if(lim + fa.sharedSizeBytes > 160 * 1024)
			lim = 160 * 1024 - fa.sharedSizeBytes;
		HIPCHK(hipFuncSetAttribute(k[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int) lim));
	}
	return 0;
}
