// Differentiable FP64 PIC run on gfx950: reverse-mode gradients of the
// time-integrated field energy L = sum_it W(it), W = -(e0/2) sum_g rho_g phi_g,
// with respect to every particle's initial position and velocity.
//
// The forward is the reference loop (sim_step, src/sim.c:482-581, with the
// iteration-0 back half-step of src/mover.c:204-214) on particles that keep
// their index (no tile sorting): per iteration
//   rho = Dep(r)  (interpolate_p2f_rho)      phi = S rho  (MFT_solve)
//   E = D phi     (field_E_compute)          E_p = Gat(E, r)  (interpolate_f2p_E)
//   u' = M (u + h E_p) + h E_p  (Boris, B = (0,0,Bz))   r' = wrap(r + dt' u')
// with the production kernels' arithmetic (picmath.h, the same FFT and E
// stencil kernels).  The reverse sweep applies the transposed maps, given
// the adjoints (lr', lu') of (r', u'):
//   mu = lu' + dt' lr'                 (move; the wrap has derivative 1)
//   lu = M^T mu,  lE_p = h (M^T mu + mu)          (Boris, linear in u and E)
//   lE = Dep(lE_p)                     (gather^T = deposit with CIC weights)
//   psi = D^T lE - e0 rho              (stencil^T; -e0 rho = dW/d(S rho))
//   lrho = S psi                       (S is symmetric: real, even G)
//   lr = lr' + grad Gat(E)(r) . lE_p + c grad Gat(lrho)(r),  c = -q/e0
// (the last two are derivatives of the CIC interpolant with respect to the
// particle position).  States are checkpointed every K iterations in the
// forward, which also keeps every iteration's rho and E when they fit a
// fixed budget; the reverse recomputes each segment's positions from its
// checkpoint (and, without that history, their rho and E).  No counterpart
// in the reference, which has no gradients.
#include "dev.h"
#include "host_sort.h"
#include "picmath.h"

#include <cmath>
#include <stdlib.h>
#include <string.h>

struct AdjSpecies {
	double h[CPIC_MAX_SPECIES];   // 0.5 dt' q / m
	double tz[CPIC_MAX_SPECIES];  // Bz h
	double sz[CPIC_MAX_SPECIES];  // 2 tz / (1 + tz^2)
	double c[CPIC_MAX_SPECIES];   // -q / e0
};

struct cpic_adj {
	cpic_dev *d;              // grid, FFT tables and stream (a batch of one, no particles)
	Geo g;
	int64_t n;
	int ns;
	AdjSpecies k[2];          // [0]: iteration 0 (dt' = -dt/2), [1]: later iterations
	int32_t *sp;              // species of each particle (device, in sorted order)
	int32_t *hsp;             // species of each particle (host, caller's order)
	int64_t *perm;            // sorted position i holds the caller's particle perm[i]
	double *hbuf;             // host staging, 4n doubles (pinned)
	double *x, *y, *ux, *uy;  // state
	double *rho[2], *phi, *psi, *lrho;
	double2 *E, *lE;          // lE: two buffers of nodes values (reverse iterations alternate)
	double *lx, *ly, *lux, *luy;
	double *ck;               // checkpoints, 4n doubles each
	int64_t ckcap;
	double *seg;              // positions of one segment's iterations after its first, 2n doubles each
	int64_t segcap;
	double *segf;             // rho and E of one segment's iterations, 3 nodes doubles each
	int64_t segfcap;
	double *hist;             // W per iteration
	int64_t histcap;
	// field history of the last forward (when it fits ADJ_FIELD_HISTORY_MAX
	// and is not switched off): rho of iterations 0..T and E of 0..T-1,
	// which the reverse sweep reads instead of recomputing them
	double *frho, *fE;
	int64_t frhocap, fEcap;   // doubles
	int fhist;                // the last forward kept it
	int fhist_off;            // cpic_adj_set_field_history(a, 0)
	double *wk;               // recompute state of the reverse sweep, 4n doubles
	hipEvent_t ev[4];         // device work of the last forward [0,1] and backward [2,3]
	int evok;                 // events created
	int evrec;                // bit 0: forward pair recorded, bit 1: backward pair
	int *derr;
	int64_t nsteps;
	int K;
	int have;                 // a forward has run
};

static __device__ __forceinline__ void cic_nodes(const Geo &g, double x, double y, int &i00,
		int &i01, int &i10, int &i11, double &rx, double &ry)
{
	int ix, iy;
	cell64(g, x, y, ix, iy, rx, ry);
	int i1x = ix + 1 == g.nx ? 0 : ix + 1;
	int i1y = iy + 1 == g.ny ? 0 : iy + 1;
	i00 = iy * g.nx + ix;
	i01 = i1y * g.nx + ix;
	i10 = iy * g.nx + i1x;
	i11 = i1y * g.nx + i1x;
}

// Deposits.  The particles are kept sorted by their initial cell (cells taken
// round-robin within each block's chunk), so the ADJ_PPT*CPIC_BLOCK
// consecutive particles of a block lie in a small patch of cells: each block
// accumulates its deposit in LDS over the bounding box of its cells and
// flushes the non-zero nodes with one global atomic each.
// A block whose box is too large or wraps around the grid deposits with
// global atomics directly.
// 4 particles per thread and 1,024 per block: at 8 the kernels hold 100+
// VGPRs (4 waves per SIMD) and the 256^2 adjoint workload runs 10 % slower
#ifndef ADJ_PPT
#define ADJ_PPT 4
#endif
#define ADJ_CHUNK (ADJ_PPT * CPIC_BLOCK)
// Particles per thread of the reverse sweep's recompute push, which only
// moves particles (no deposit, no LDS box): at 256^2 (4,194,304 particles)
// 4 per thread make ~2.3 rounds of workgroups, the last a third full;
// smaller blocks of work keep the chip busier to the end
#ifndef MOVE_PPT
#define MOVE_PPT 2
#endif
// 20 KB of double: seven depositing blocks per CU (their 72 VGPRs allow
// seven; a 24 KB box held them to six: forward 15.8 -> 15.7 ms at 256^2);
// the boxes there hold at most ~125 nodes
#ifndef ADJ_BOX_NODES
#define ADJ_BOX_NODES 2560
#endif

// LDS accumulation box of a block: nodes [x0, x0 + w) x [y0, y0 + h) in
// coordinates relative to a reference cell (rx, ry), wrapped into
// [-n/2, n/2).  A block's particles then form one compact box also where
// they straddle the periodic boundary: a chunk of the cell-sorted order
// that ends one grid row and begins the next, or particles that crossed
// x = 0.  (With absolute coordinates such a box spans the whole row, and
// the block fell back to global atomics on a few rows of nodes, which
// serialise: 6-9 % of the blocks at 256^2.)
struct Box {
	int x0, y0, w, h;
	int rx, ry;
	bool lds;
};

// grid sides are powers of two
static __device__ __forceinline__ int rel_wrap(int i, int ref, int n)
{
	int r = (i - ref) & (n - 1);
	return r >= n / 2 ? r - n : r;
}

// Bounding box (plus the CIC +1 node) of the block's cells; ix[u] < 0 marks
// no particle.  Thread 0's first particle (the block's first, which always
// exists) is the reference cell.  s6: six ints of LDS.
template <int U>
static __device__ Box block_box(const Geo &g, const int *ix, const int *iy, int *s6, int cap)
{
	if(threadIdx.x == 0)
	{
		s6[0] = 0x7fffffff;
		s6[1] = -0x7fffffff;
		s6[2] = 0x7fffffff;
		s6[3] = -0x7fffffff;
		s6[4] = ix[0];
		s6[5] = iy[0];
	}
	__syncthreads();
	const int rx = s6[4], ry = s6[5];
	int lox = 0x7fffffff, hix = -0x7fffffff, loy = 0x7fffffff, hiy = -0x7fffffff;
#pragma unroll
	for(int u = 0; u < U; u++)
		if(ix[u] >= 0)
		{
			int ax = rel_wrap(ix[u], rx, g.nx), ay = rel_wrap(iy[u], ry, g.ny);
			lox = min(lox, ax);
			hix = max(hix, ax);
			loy = min(loy, ay);
			hiy = max(hiy, ay);
		}
	// wavefront reduction first: four LDS atomics per wavefront
	for(int off = 32; off > 0; off >>= 1)
	{
		lox = min(lox, __shfl_xor(lox, off));
		hix = max(hix, __shfl_xor(hix, off));
		loy = min(loy, __shfl_xor(loy, off));
		hiy = max(hiy, __shfl_xor(hiy, off));
	}
	if((threadIdx.x & 63) == 0 && hix >= lox)
	{
		atomicMin(&s6[0], lox);
		atomicMax(&s6[1], hix);
		atomicMin(&s6[2], loy);
		atomicMax(&s6[3], hiy);
	}
	__syncthreads();
	Box b;
	b.rx = rx;
	b.ry = ry;
	b.x0 = s6[0];
	b.y0 = s6[2];
	b.w = s6[1] - s6[0] + 2;
	b.h = s6[3] - s6[2] + 2;
	b.lds = s6[1] - s6[0] <= g.nx / 2 && s6[3] - s6[2] <= g.ny / 2 && b.w * b.h <= cap;
	__syncthreads();
	return b;
}

static __device__ __forceinline__ void cell_of(const Geo &g, double x, double y, int &ix, int &iy,
		double &rx, double &ry)
{
	cell64(g, x, y, ix, iy, rx, ry);
}

// box node (dx, dy in {0,1} from cell (ix, iy))
static __device__ __forceinline__ int box_node(const Geo &g, const Box &b, int ix, int iy,
		int dx, int dy)
{
	return (rel_wrap(iy, b.ry, g.ny) - b.y0 + dy) * b.w + (rel_wrap(ix, b.rx, g.nx) - b.x0 + dx);
}

static __device__ __forceinline__ int64_t wrap_node(const Geo &g, int x, int y)
{
	x = x >= g.nx ? x - g.nx : x;
	y = y >= g.ny ? y - g.ny : y;
	return (int64_t) y * g.nx + x;
}

// Adds c * (CIC weights of (rx, ry)) at the 4 nodes of cell (ix, iy) into
// the LDS box or, without one, into the global grid (order of
// src/interpolate.c:259-262).
static __device__ __forceinline__ void dep1(const Geo &g, const Box &b, double *sacc,
		double *__restrict__ rho, int ix, int iy, double rx, double ry, double c)
{
	double dxw = 1.0 - rx, dyw = 1.0 - ry;
	double v00 = (dxw * dyw) * c, v01 = (dxw * ry) * c, v10 = (rx * dyw) * c, v11 = (rx * ry) * c;
	if(b.lds)
	{
		LDS_AS double *l = (LDS_AS double *) sacc;
		lds_add(&l[box_node(g, b, ix, iy, 0, 0)], v00);
		lds_add(&l[box_node(g, b, ix, iy, 0, 1)], v01);
		lds_add(&l[box_node(g, b, ix, iy, 1, 0)], v10);
		lds_add(&l[box_node(g, b, ix, iy, 1, 1)], v11);
	}
	else
	{
		GLB_AS double *m = gptr<double>(rho);
		glb_add(&m[wrap_node(g, ix, iy)], v00);
		glb_add(&m[wrap_node(g, ix, iy + 1)], v01);
		glb_add(&m[wrap_node(g, ix + 1, iy)], v10);
		glb_add(&m[wrap_node(g, ix + 1, iy + 1)], v11);
	}
}

static __device__ void box_zero(const Box &b, double *sacc, int ncomp)
{
	if(b.lds)
		for(int i = threadIdx.x; i < b.w * b.h * ncomp; i += blockDim.x)
			sacc[i] = 0.0;
	__syncthreads();
}

// flushes component k of an LDS box holding ncomp interleaved components
static __device__ void box_flush(const Geo &g, const Box &b, const double *sacc, int ncomp,
		double *__restrict__ out, int out_stride)
{
	__syncthreads();
	if(!b.lds)
		return;
	for(int i = threadIdx.x; i < b.w * b.h; i += blockDim.x)
	{
		int ly = i / b.w, lx = i - ly * b.w;
		int gx = (b.rx + b.x0 + lx) & (g.nx - 1), gy = (b.ry + b.y0 + ly) & (g.ny - 1);
		int64_t gnode = (int64_t) gy * g.nx + gx;
		for(int k = 0; k < ncomp; k++)
		{
			double v = ((const LDS_AS double *) sacc)[i * ncomp + k];
			if(v != 0.0)
				glb_add(&gptr<double>(out)[gnode * out_stride + k], v);
		}
	}
}

// rho += Dep(r) for the particles of the block
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_deposit(Geo g, AdjSpecies k, int64_t n, const int32_t *__restrict__ sp,
		const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ rho)
{
	__shared__ double sacc[ADJ_BOX_NODES];
	__shared__ int s6[6];
	int64_t p0 = (int64_t) blockIdx.x * ADJ_CHUNK;
	int ix[ADJ_PPT], iy[ADJ_PPT];
	double rx[ADJ_PPT], ry[ADJ_PPT];
#pragma unroll
	for(int u = 0; u < ADJ_PPT; u++)
	{
		int64_t i = p0 + u * CPIC_BLOCK + threadIdx.x;
		ix[u] = -1;
		if(i < n)
		{
			cell_of(g, x[i], y[i], ix[u], iy[u], rx[u], ry[u]);
		}
	}
	Box b = block_box<ADJ_PPT>(g, ix, iy, s6, ADJ_BOX_NODES);
	box_zero(b, sacc, 1);
#pragma unroll
	for(int u = 0; u < ADJ_PPT; u++)
	{
		int64_t i = p0 + u * CPIC_BLOCK + threadIdx.x;
		if(ix[u] >= 0)
			dep1(g, b, sacc, rho, ix[u], iy[u], rx[u], ry[u], k.c[sp[i]]);
	}
	box_flush(g, b, sacc, 1, rho, 1);
}

// gather + Boris + move + wrap, then (DEP) deposit at the new position into
// rho_next; positions are read from (x, y) and written to (xo, yo), which may
// be the same arrays.  PPT particles per thread: ADJ_PPT with the deposit
// (its LDS box spans one sort chunk), MOVE_PPT without
template <bool DEP, int PPT>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_push(Geo g, AdjSpecies k, int set_r, double dtp, int64_t n,
		const int32_t *__restrict__ sp, const double *x, const double *y, double *xo,
		double *yo, double *__restrict__ ux, double *__restrict__ uy,
		const double2 *__restrict__ E, double *__restrict__ rho_next, int *derr)
{
	__shared__ double sacc[ADJ_BOX_NODES];
	__shared__ int s6[6];
	bool bad = false, nan_ = false;
	int64_t p0 = (int64_t) blockIdx.x * (PPT * CPIC_BLOCK);
	int ix[PPT], iy[PPT];
	double rx[PPT], ry[PPT];
#pragma unroll
	for(int u = 0; u < PPT; u++)
	{
		int64_t i = p0 + u * CPIC_BLOCK + threadIdx.x;
		ix[u] = -1;
		if(i >= n)
			continue;
		int s = sp[i];
		int jx, jy;
		double qx, qy;
		double px = x[i], py = y[i];
		cell_of(g, px, py, jx, jy, qx, qy);
		int j1x = jx + 1 == g.nx ? 0 : jx + 1;
		int j1y = jy + 1 == g.ny ? 0 : jy + 1;
		double dxw = 1.0 - qx, dyw = 1.0 - qy;
		double w00 = dxw * dyw, w01 = dxw * qy, w10 = qx * dyw, w11 = qx * qy;
		double2 e00 = E[(int64_t) jy * g.nx + jx], e01 = E[(int64_t) j1y * g.nx + jx];
		double2 e10 = E[(int64_t) jy * g.nx + j1x], e11 = E[(int64_t) j1y * g.nx + j1x];
		double ex = w00 * e00.x;
		ex = fma(w01, e01.x, ex);
		ex = fma(w10, e10.x, ex);
		ex = fma(w11, e11.x, ex);
		double ey = w00 * e00.y;
		ey = fma(w01, e01.y, ey);
		ey = fma(w10, e10.y, ey);
		ey = fma(w11, e11.y, ey);
		double uu = ux[i], vv = uy[i];
		boris_bz<double>(k.h[s], k.tz[s], k.sz[s], ex, ey, uu, vv);
		if(fabs(uu) > g.umax_x || fabs(vv) > g.umax_y)
			bad = true;
		if(set_r)
		{
			px = wrap64(fma(uu, dtp, px), g.L0);
			py = wrap64(fma(vv, dtp, py), g.L1);
		}
		if(!(isfinite(px) && isfinite(py) && isfinite(uu) && isfinite(vv)))
		{
			nan_ = true;
			px = py = 0.0;
		}
		xo[i] = px;
		yo[i] = py;
		ux[i] = uu;
		uy[i] = vv;
		cell_of(g, px, py, ix[u], iy[u], rx[u], ry[u]);
	}
	if(bad)
		atomicOr(derr, CPIC_DERR_UMAX);
	if(nan_)
		atomicOr(derr, CPIC_DERR_NAN);
	if(!DEP)
		return;
	Box b = block_box<PPT>(g, ix, iy, s6, ADJ_BOX_NODES);
	box_zero(b, sacc, 1);
#pragma unroll
	for(int u = 0; u < PPT; u++)
	{
		int64_t i = p0 + u * CPIC_BLOCK + threadIdx.x;
		if(ix[u] >= 0)
			dep1(g, b, sacc, rho_next, ix[u], iy[u], rx[u], ry[u], k.c[sp[i]]);
	}
	box_flush(g, b, sacc, 1, rho_next, 1);
}

// (ax, ay) += c grad Gat(lrho)(x, y): the derivative of the CIC interpolant
// of lrho at a particle's position, the last term of lr
static __device__ __forceinline__ void rev_b_term(const Geo &g, double c, double x, double y,
		const double *__restrict__ lrho, double &ax, double &ay)
{
	int i00, i01, i10, i11;
	double rx, ry;
	cic_nodes(g, x, y, i00, i01, i10, i11, rx, ry);
	double l00 = lrho[i00], l01 = lrho[i01], l10 = lrho[i10], l11 = lrho[i11];
	double dxw = 1.0 - rx, dyw = 1.0 - ry;
	ax += c * (g.idx * (dyw * (l10 - l00) + ry * (l11 - l01)));
	ay += c * (g.idy * (dxw * (l01 - l00) + rx * (l11 - l10)));
}

// Reverse of move + Boris + gather for one iteration; accumulates the
// gather-weight derivative into (lx, ly) and deposits lE_p into lE.  With
// PEND, first adds the lrho term of the iteration after it (rev_b_term at
// that iteration's positions (bx, by) with its lrho), so that (lx, ly) are
// read and written once for both.
// REV_PPT particles per thread: the kernel keeps six values per particle
// across the box barriers, and 8 per thread would hold it to 3 waves per
// SIMD; 2 rather than 4 make twice the workgroups, so that at 256^2 the last
// of their rounds idles less of the chip (reverse sweep 34.4 -> 33.5 ms).
// A fraction of a round-robin chunk still spans its cells.
#ifndef REV_PPT
#define REV_PPT 2
#endif
#define REV_CHUNK (REV_PPT * CPIC_BLOCK)
static_assert(ADJ_CHUNK % REV_CHUNK == 0, "reverse blocks split the sort's chunks");
#define REV_BOX_NODES 1536          // 2 components: 24 KB
// Device memory the field history of a forward may take (rho and E of every
// iteration, 24 B per node and iteration: 315 MB for 200 iterations at
// 256^2); beyond it the reverse sweep recomputes each segment's fields
#define ADJ_FIELD_HISTORY_MAX ((int64_t) 8 << 30)
template <bool PEND>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_rev_a(Geo g, AdjSpecies k, int set_r, double dtp, int64_t n,
		const int32_t *__restrict__ sp, const double *__restrict__ x,
		const double *__restrict__ y, const double2 *__restrict__ E, double *__restrict__ lx,
		double *__restrict__ ly, double *__restrict__ lux, double *__restrict__ luy,
		double2 *__restrict__ lE, const double *__restrict__ bx, const double *__restrict__ by,
		const double *__restrict__ lrho)
{
	__shared__ double sacc[2 * REV_BOX_NODES];
	__shared__ int s6[6];
	int64_t p0 = (int64_t) blockIdx.x * REV_CHUNK;
	int ix[REV_PPT], iy[REV_PPT];
	double rx[REV_PPT], ry[REV_PPT], lex[REV_PPT], ley[REV_PPT];
#pragma unroll
	for(int u = 0; u < REV_PPT; u++)
	{
		int64_t i = p0 + u * CPIC_BLOCK + threadIdx.x;
		ix[u] = -1;
		if(i >= n)
			continue;
		int s = sp[i];
		double mx = lux[i], my = luy[i];
		double ax = lx[i], ay = ly[i];
		if(PEND)
			rev_b_term(g, k.c[s], bx[i], by[i], lrho, ax, ay);
		if(set_r)
		{
			mx = fma(dtp, ax, mx);
			my = fma(dtp, ay, my);
		}
		double h = k.h[s], tz = k.tz[s], sz = k.sz[s];
		double a_ = 1.0 - tz * sz, b_ = sz;
		double nux = a_ * mx - b_ * my;      // M^T mu
		double nuy = b_ * mx + a_ * my;
		lex[u] = h * (nux + mx);
		ley[u] = h * (nuy + my);
		lux[i] = nux;
		luy[i] = nuy;

		cell_of(g, x[i], y[i], ix[u], iy[u], rx[u], ry[u]);
		int i1x = ix[u] + 1 == g.nx ? 0 : ix[u] + 1;
		int i1y = iy[u] + 1 == g.ny ? 0 : iy[u] + 1;
		double dxw = 1.0 - rx[u], dyw = 1.0 - ry[u];
		double2 e00 = E[(int64_t) iy[u] * g.nx + ix[u]], e01 = E[(int64_t) i1y * g.nx + ix[u]];
		double2 e10 = E[(int64_t) iy[u] * g.nx + i1x], e11 = E[(int64_t) i1y * g.nx + i1x];
		// d/dx and d/dy of the CIC interpolants of Ex and Ey
		double gxx = g.idx * (dyw * (e10.x - e00.x) + ry[u] * (e11.x - e01.x));
		double gyx = g.idy * (dxw * (e01.x - e00.x) + rx[u] * (e11.x - e10.x));
		double gxy = g.idx * (dyw * (e10.y - e00.y) + ry[u] * (e11.y - e01.y));
		double gyy = g.idy * (dxw * (e01.y - e00.y) + rx[u] * (e11.y - e10.y));
		lx[i] = ax + (lex[u] * gxx + ley[u] * gxy);
		ly[i] = ay + (lex[u] * gyx + ley[u] * gyy);
	}
	Box b = block_box<REV_PPT>(g, ix, iy, s6, REV_BOX_NODES);
	box_zero(b, sacc, 2);
#pragma unroll
	for(int u = 0; u < REV_PPT; u++)
	{
		if(ix[u] < 0)
			continue;
		double dxw = 1.0 - rx[u], dyw = 1.0 - ry[u];
		double w[4] = {dxw * dyw, dxw * ry[u], rx[u] * dyw, rx[u] * ry[u]};
		const int ddx[4] = {0, 0, 1, 1}, ddy[4] = {0, 1, 0, 1};
#pragma unroll
		for(int c = 0; c < 4; c++)
		{
			if(b.lds)
			{
				LDS_AS double *l = (LDS_AS double *) sacc;
				int nd = box_node(g, b, ix[u], iy[u], ddx[c], ddy[c]);
				lds_add(&l[2 * nd], w[c] * lex[u]);
				lds_add(&l[2 * nd + 1], w[c] * ley[u]);
			}
			else
			{
				GLB_AS double *m = gptr<double>(lE);
				int64_t nd = wrap_node(g, ix[u] + ddx[c], iy[u] + ddy[c]);
				glb_add(&m[2 * nd], w[c] * lex[u]);
				glb_add(&m[2 * nd + 1], w[c] * ley[u]);
			}
		}
	}
	box_flush(g, b, sacc, 2, (double *) lE, 2);
}

// psi = D^T lE - srho * rho, D the central-difference E stencil
// (E_x[i] = (phi[i-1] - phi[i+1]) / 2dx, so (D_x^T a)[i] = (a[i+1] - a[i-1]) / 2dx);
// zeroes lE_zero (the other lE buffer, which the next reverse iteration
// deposits into)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_field(Geo g, const double2 *__restrict__ lE, const double *__restrict__ rho,
		double srho, double *__restrict__ psi, double2 *__restrict__ lE_zero)
{
	int64_t n = (int64_t) g.nx * g.ny;
	double dx2 = 2 * g.dx, dy2 = 2 * g.dy;
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int iy = (int) (i >> g.lgnx), ix = (int) (i & (g.nx - 1));
		int xm = ix == 0 ? g.nx - 1 : ix - 1, xp = ix == g.nx - 1 ? 0 : ix + 1;
		int ym = iy == 0 ? g.ny - 1 : iy - 1, yp = iy == g.ny - 1 ? 0 : iy + 1;
		int64_t row = (int64_t) iy * g.nx;
		double lphi = (lE[row + xp].x - lE[row + xm].x) / dx2 +
			(lE[(int64_t) yp * g.nx + ix].y - lE[(int64_t) ym * g.nx + ix].y) / dy2;
		psi[i] = lphi - srho * rho[i];
		lE_zero[i] = make_double2(0.0, 0.0);
	}
}

// lr += c * grad Gat(lrho)(r) for the last iteration of the sweep (the
// others are applied by the next k_adj_rev_a)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_rev_b(Geo g, AdjSpecies k, int64_t n, const int32_t *__restrict__ sp,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ lrho, double *__restrict__ lx, double *__restrict__ ly)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		double ax = lx[i], ay = ly[i];
		rev_b_term(g, k.c[sp[i]], x[i], y[i], lrho, ax, ay);
		lx[i] = ax;
		ly[i] = ay;
	}
}

// ---------------------------------------------------------------- host ---

// blocks of the depositing kernels: one per ADJ_CHUNK particles
static int
chunks(int64_t n)
{
	return (int) ((n + ADJ_CHUNK - 1) / ADJ_CHUNK);
}

static int
pblocks(int64_t n)
{
	int64_t b = (n + CPIC_BLOCK - 1) / CPIC_BLOCK;
	if(b > 65536)
		b = 65536;
	return b < 1 ? 1 : (int) b;
}

static void
adj_free(cpic_adj *a)
{
	if(!a)
		return;
	if(a->d)
		hipStreamSynchronize(a->d->stream);
	hipFree(a->sp); hipFree(a->x); hipFree(a->y); hipFree(a->ux); hipFree(a->uy);
	hipFree(a->rho[0]); hipFree(a->rho[1]); hipFree(a->phi); hipFree(a->psi); hipFree(a->lrho);
	hipFree(a->E); hipFree(a->lE);
	hipFree(a->lx); hipFree(a->ly); hipFree(a->lux); hipFree(a->luy);
	hipFree(a->ck); hipFree(a->seg); hipFree(a->segf); hipFree(a->hist);
	hipFree(a->frho); hipFree(a->fE);
	hipFree(a->derr);
	hipFree(a->wk);
	if(a->evok)
		for(int i = 0; i < 4; i++)
			hipEventDestroy(a->ev[i]);
	free(a->hsp);
	free(a->perm);
	if(a->hbuf)
		hipHostFree(a->hbuf);
	if(a->d)
		cpic_dev_destroy(a->d);
	free(a);
}

static void
species_consts(const cpic_dev_desc_t *desc, double dtp, AdjSpecies *k)
{
	memset(k, 0, sizeof(*k));
	for(int s = 0; s < desc->nspecies; s++)
	{
		double h = 0.5 * dtp * desc->q[s] / desc->m[s];
		double tz = desc->B[2] * h;
		double sden = fma(tz, tz, 1.0);
		k->h[s] = h;
		k->tz[s] = tz;
		k->sz[s] = 2.0 * tz / sden;
		k->c[s] = -desc->q[s] / desc->e0;
	}
}

extern "C" int
cpic_adj_create(const cpic_dev_desc_t *desc, int64_t n, const int32_t *species,
		cpic_adj_t **out)
{
	cpic_dev_desc_t dd;
	cpic_adj *a;
	int ret;
	if(!desc || !out || n < 1 || !species || desc->precision != 64)
		return CPIC_EINVAL;
	*out = NULL;
	for(int64_t i = 0; i < n; i++)
		if(species[i] < 0 || species[i] >= desc->nspecies)
			return CPIC_EINVAL;
	dd = *desc;
	dd.flags = CPIC_F_NO_DIAG | CPIC_F_NO_GRAPH;
	a = (cpic_adj *) calloc(1, sizeof(*a));
	if(!a)
		return CPIC_ENOMEM;
	if((ret = cpic_dev_create(&dd, &a->d)))
	{
		free(a);
		return ret;
	}
	a->g = a->d->hgeo[0];
	a->n = n;
	a->ns = desc->nspecies;
	species_consts(desc, -desc->dt / 2, &a->k[0]);
	species_consts(desc, desc->dt, &a->k[1]);
	int64_t nodes = (int64_t) a->g.nx * a->g.ny;
	size_t pn = sizeof(double) * (size_t) n, gn = sizeof(double) * (size_t) nodes;
	if(hipMalloc(&a->sp, sizeof(int32_t) * n) != hipSuccess ||
			hipMalloc(&a->x, pn) != hipSuccess || hipMalloc(&a->y, pn) != hipSuccess ||
			hipMalloc(&a->ux, pn) != hipSuccess || hipMalloc(&a->uy, pn) != hipSuccess ||
			hipMalloc(&a->lx, pn) != hipSuccess || hipMalloc(&a->ly, pn) != hipSuccess ||
			hipMalloc(&a->lux, pn) != hipSuccess || hipMalloc(&a->luy, pn) != hipSuccess ||
			hipMalloc(&a->rho[0], gn) != hipSuccess || hipMalloc(&a->rho[1], gn) != hipSuccess ||
			hipMalloc(&a->phi, gn) != hipSuccess || hipMalloc(&a->psi, gn) != hipSuccess ||
			hipMalloc(&a->lrho, gn) != hipSuccess ||
			hipMalloc(&a->E, 2 * gn) != hipSuccess || hipMalloc(&a->lE, 4 * gn) != hipSuccess ||
			hipMalloc(&a->derr, sizeof(int)) != hipSuccess)
	{
		adj_free(a);
		return CPIC_ENOMEM;
	}
	a->hsp = (int32_t *) malloc(sizeof(int32_t) * (size_t) n);
	a->perm = (int64_t *) malloc(sizeof(int64_t) * (size_t) n);
	if(hipHostMalloc((void **) &a->hbuf, sizeof(double) * 4 * (size_t) n, hipHostMallocDefault)
			!= hipSuccess)
		a->hbuf = NULL;
	if(!a->hsp || !a->perm || !a->hbuf)
	{
		adj_free(a);
		return CPIC_ENOMEM;
	}
	memcpy(a->hsp, species, sizeof(int32_t) * n);
	for(int i = 0; i < 4; i++)
		if(hipEventCreate(&a->ev[i]) != hipSuccess)
		{
			for(int j = 0; j < i; j++)
				hipEventDestroy(a->ev[j]);
			adj_free(a);
			return CPIC_EHIP;
		}
	a->evok = 1;
	*out = a;
	return 0;
}

extern "C" int
cpic_adj_destroy(cpic_adj_t *a)
{
	if(!a)
		return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	adj_free(a);
	return 0;
}

static int
grow(double **p, int64_t *cap, int64_t want)
{
	if(*cap >= want)
		return 0;
	hipFree(*p);
	*p = NULL;
	*cap = 0;
	if(hipMalloc(p, sizeof(double) * (size_t) want) != hipSuccess)
		return CPIC_ENOMEM;
	*cap = want;
	return 0;
}

static int
zero(cpic_adj *a, double *p, int64_t n)
{
	return cpic_zero(a->d, p, n);
}

static int
deposit(cpic_adj *a, const double *x, const double *y, double *rho)
{
	int ret;
	if((ret = zero(a, rho, (int64_t) a->g.nx * a->g.ny)))
		return ret;
	hipLaunchKernelGGL(k_adj_deposit, dim3(chunks(a->n)), dim3(CPIC_BLOCK), 0, a->d->stream,
			a->g, a->k[1], a->n, a->sp, x, y, rho);
	HIPCHK(hipGetLastError());
	return 0;
}

// phi and E (into E) from rho; with w, the field energy of rho,
// -(e0/2) sum rho phi, is added to *w (zero beforehand; the column pass of
// the solve sums it over the spectrum, sum_k G |rho^(k)|^2 / (nx ny), equal
// to sum rho phi by Parseval); with rho_zero, that (other) buffer is zeroed
// by the solve's last pass
static int
fields(cpic_adj *a, const double *rho, double *w, double *rho_zero, double2 *E)
{
	int ret;
	if((ret = cpic_solve_ptr_energy(a->d, rho, a->phi, rho_zero, w)))
		return ret;
	return cpic_efield_ptr(a->d, a->phi, E);
}

// one forward iteration from positions (x, y), velocities (ux, uy) and
// rho_in = Dep(r): E of this iteration into E, the moved positions into
// (xo, yo) (may be (x, y)), Dep(r') of the moved particles into rho_out
static int
forward_step(cpic_adj *a, int64_t it, const double *x, const double *y, double *xo,
		double *yo, const double *rho_in, double *rho_out, double2 *E, double *w)
{
	int ret;
	if((ret = fields(a, rho_in, w, rho_out, E)))
		return ret;
	int first = it == 0;
	double dtp = first ? -a->g.dt / 2 : a->g.dt;
	hipLaunchKernelGGL((k_adj_push<true, ADJ_PPT>), dim3(chunks(a->n)), dim3(CPIC_BLOCK), 0,
			a->d->stream,
			a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, a->n, a->sp, x, y, xo, yo, a->ux,
			a->uy, (const double2 *) E, rho_out, a->derr);
	HIPCHK(hipGetLastError());
	return 0;
}

// The particle half of forward_step, for the reverse sweep's recompute with
// a field history: positions (and the velocities in place) of iteration
// it + 1 from those of it with E of it; no deposit
static int
move_step(cpic_adj *a, int64_t it, const double *x, const double *y, double *xo, double *yo,
		const double2 *E)
{
	int first = it == 0;
	double dtp = first ? -a->g.dt / 2 : a->g.dt;
	hipLaunchKernelGGL((k_adj_push<false, MOVE_PPT>),
			dim3((unsigned) ((a->n + MOVE_PPT * CPIC_BLOCK - 1) / (MOVE_PPT * CPIC_BLOCK))),
			dim3(CPIC_BLOCK), 0, a->d->stream,
			a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, a->n, a->sp, x, y, xo, yo, a->ux,
			a->uy, E, (double *) NULL, a->derr);
	HIPCHK(hipGetLastError());
	return 0;
}

// forward_step in place on the state, on the double-buffered rho[cur] /
// rho[1-cur] and the E array
static int
forward_iter(cpic_adj *a, int64_t it, int *cur, double *w)
{
	int ret, c = *cur;
	if((ret = forward_step(a, it, a->x, a->y, a->x, a->y, a->rho[c], a->rho[1 - c], a->E, w)))
		return ret;
	*cur = 1 - c;
	return 0;
}

static int
adj_derr(cpic_adj *a)
{
	int e = 0;
	HIPCHK(hipMemcpyAsync(&e, a->derr, sizeof(int), hipMemcpyDeviceToHost, a->d->stream));
	HIPCHK(hipStreamSynchronize(a->d->stream));
	if(e & CPIC_DERR_UMAX)
		return CPIC_EUMAX;
	if(e & CPIC_DERR_NAN)
		return CPIC_ENAN;
	return 0;
}

static int
copy_state(cpic_adj *a, double *dst, const double *x, const double *y, const double *ux,
		const double *uy, hipMemcpyKind kind)
{
	size_t pn = sizeof(double) * (size_t) a->n;
	int64_t n = a->n;
	HIPCHK(hipMemcpyAsync(dst, x, pn, kind, a->d->stream));
	HIPCHK(hipMemcpyAsync(dst + n, y, pn, kind, a->d->stream));
	if(ux)
		HIPCHK(hipMemcpyAsync(dst + 2 * n, ux, pn, kind, a->d->stream));
	if(uy)
		HIPCHK(hipMemcpyAsync(dst + 3 * n, uy, pn, kind, a->d->stream));
	return 0;
}

// perm: the particles in the order of their cell (iy*nx + ix), counting
// sort, stable; the deposits' LDS boxes rely on neighbours in this order
// being neighbours on the grid.
static int
sort_by_cell(cpic_adj *a, const double *x, const double *y)
{
	// Within each block's chunk (same cells, so the same LDS box) the cells
	// are taken round-robin: the 64 consecutive particles of a wavefront
	// then lie in different cells, and their deposit atomics spread over the
	// LDS banks instead of piling onto the 4 nodes of one cell.
	return sort_by_cell_rr(a->n, x, y, a->g.idx, a->g.idy, a->g.nx, a->g.ny, ADJ_CHUNK,
			a->perm) ? CPIC_ENOMEM : 0;
}

// Device arrays in sorted order -> host arrays in the caller's order.
static int
download_unsorted(cpic_adj *a, double *const *src, double *const *dst)
{
	const int64_t n = a->n;
	size_t pn = sizeof(double) * (size_t) n;
	for(int c = 0; c < 4; c++)
		if(dst[c])
			HIPCHK(hipMemcpyAsync(a->hbuf + c * n, src[c], pn, hipMemcpyDeviceToHost, a->d->stream));
	HIPCHK(hipStreamSynchronize(a->d->stream));
	for(int c = 0; c < 4; c++)
		if(dst[c])
			scatter_perm(n, a->perm, (const double *) (a->hbuf + c * n), dst[c]);
	return 0;
}

extern "C" int
cpic_adj_forward(cpic_adj_t *a, const double *x, const double *y, const double *ux,
		const double *uy, int64_t nsteps, int32_t checkpoint_every, double *loss,
		double *energy)
{
	int ret, cur = 0;
	if(!a || !x || !y || !ux || !uy || nsteps < 1 || checkpoint_every < 1)
		return CPIC_EINVAL;
	std::atomic<int> invalid(0);
	parallel_ranges(a->n, host_threads(a->n, 1 << 16), [&](int, int64_t lo, int64_t hi) {
		for(int64_t i = lo; i < hi; i++)
			if(!(x[i] >= 0.0 && x[i] <= a->g.L0 && y[i] >= 0.0 && y[i] <= a->g.L1) ||
					!std::isfinite(ux[i]) || !std::isfinite(uy[i]))
			{
				invalid = 1;
				return;
			}
	});
	if(invalid)
		return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	int64_t K = checkpoint_every, nck = (nsteps + K - 1) / K, n = a->n;
	if((ret = grow(&a->ck, &a->ckcap, 4 * n * nck)) || (ret = grow(&a->hist, &a->histcap, nsteps)))
		return ret;
	if((ret = sort_by_cell(a, x, y)))
		return ret;
	size_t pn = sizeof(double) * (size_t) n;
	const double *src[4] = {x, y, ux, uy};
	double *dev[4] = {a->x, a->y, a->ux, a->uy};
	int32_t *ssp = (int32_t *) malloc(sizeof(int32_t) * (size_t) n);
	if(!ssp)
		return CPIC_ENOMEM;
	gather_perm(n, a->perm, (const int32_t *) a->hsp, ssp);
	hipError_t e = hipMemcpy(a->sp, ssp, sizeof(int32_t) * n, hipMemcpyHostToDevice);
	free(ssp);
	if(e != hipSuccess)
		return CPIC_EHIP;
	for(int c = 0; c < 4; c++)
	{
		double *h = a->hbuf + c * n;
		gather_perm(n, a->perm, src[c], h);
		HIPCHK(hipMemcpyAsync(dev[c], h, pn, hipMemcpyHostToDevice, a->d->stream));
	}
	HIPCHK(hipMemsetAsync(a->derr, 0, sizeof(int), a->d->stream));
	a->have = 0;
	a->evrec = 0;             // a backward belongs to the forward before it
	// the field history, when it fits: each iteration's rho and E are
	// written straight into their slots (the solve zeroes the slot the push
	// deposits into), at no extra work in the forward
	const int64_t nodes = (int64_t) a->g.nx * a->g.ny;
	a->fhist = 0;
	if(!a->fhist_off &&
			(double) nodes * (3.0 * (double) nsteps + 1.0) * sizeof(double) <=
			(double) ADJ_FIELD_HISTORY_MAX)
	{
		if(!grow(&a->frho, &a->frhocap, nodes * (nsteps + 1)) &&
				!grow(&a->fE, &a->fEcap, 2 * nodes * nsteps))
			a->fhist = 1;
		else
			(void) hipGetLastError();   // no room: the reverse recomputes instead
	}
	HIPCHK(hipEventRecord(a->ev[0], a->d->stream));
	if((ret = deposit(a, a->x, a->y, a->fhist ? a->frho : a->rho[0])) ||
			(ret = zero(a, a->hist, nsteps)))
		return ret;
	for(int64_t it = 0; it < nsteps; it++)
	{
		if(it % K == 0)
			if((ret = copy_state(a, a->ck + 4 * n * (it / K), a->x, a->y, a->ux, a->uy,
					hipMemcpyDeviceToDevice)))
				return ret;
		if(a->fhist)
			ret = forward_step(a, it, a->x, a->y, a->x, a->y, a->frho + nodes * it,
					a->frho + nodes * (it + 1), (double2 *) a->fE + nodes * it, a->hist + it);
		else
			ret = forward_iter(a, it, &cur, a->hist + it);
		if(ret)
			return ret;
	}
	HIPCHK(hipEventRecord(a->ev[1], a->d->stream));
	a->evrec |= 1;
	double *h = (double *) malloc(sizeof(double) * (size_t) nsteps);
	if(!h)
		return CPIC_ENOMEM;
	if(hipMemcpyAsync(h, a->hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost, a->d->stream)
			!= hipSuccess || hipStreamSynchronize(a->d->stream) != hipSuccess)
	{
		free(h);
		return CPIC_EHIP;
	}
	double L = 0.0;
	for(int64_t it = 0; it < nsteps; it++)
		L += h[it];
	if(energy)
		memcpy(energy, h, sizeof(double) * nsteps);
	free(h);
	if(loss)
		*loss = L;
	if((ret = adj_derr(a)))
		return ret;
	a->nsteps = nsteps;
	a->K = (int) K;
	a->have = 1;
	return 0;
}

extern "C" int
cpic_adj_device_ms(cpic_adj_t *a, double *forward_ms, double *backward_ms)
{
	if(!a || !a->evok)
		return CPIC_EINVAL;
	float f = 0.0f, b = 0.0f;
	hipSetDevice(a->d->device);
	if(a->evrec & 1)
		HIPCHK(hipEventElapsedTime(&f, a->ev[0], a->ev[1]));
	if(a->evrec & 2)
		HIPCHK(hipEventElapsedTime(&b, a->ev[2], a->ev[3]));
	if(forward_ms)
		*forward_ms = (a->evrec & 1) ? (double) f : -1.0;
	if(backward_ms)
		*backward_ms = (a->evrec & 2) ? (double) b : -1.0;
	return 0;
}

extern "C" int
cpic_adj_set_field_history(cpic_adj_t *a, int on)
{
	if(!a)
		return CPIC_EINVAL;
	a->fhist_off = !on;
	return 0;
}

extern "C" int
cpic_adj_state(cpic_adj_t *a, double *x, double *y, double *ux, double *uy)
{
	if(!a || !a->have)
		return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	double *src[4] = {a->x, a->y, a->ux, a->uy};
	double *dst[4] = {x, y, ux, uy};
	return download_unsorted(a, src, dst);
}

extern "C" int
cpic_adj_backward(cpic_adj_t *a, double dloss, double *gx, double *gy, double *gux,
		double *guy)
{
	int ret;
	if(!a || !a->have)
		return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	const int64_t n = a->n, K = a->K, T = a->nsteps, nodes = (int64_t) a->g.nx * a->g.ny;
	const int64_t nck = (T + K - 1) / K;
	// a segment holds at most min(K, T) iterations (the last one T - (nck-1) K),
	// the positions of all but its first in the segment buffer
	const int64_t Kseg = K < T ? K : T;
	const int fh = a->fhist;
	if((ret = grow(&a->seg, &a->segcap, 2 * n * (Kseg - 1))) ||
			(!fh && (ret = grow(&a->segf, &a->segfcap, 3 * nodes * Kseg))))
		return ret;
	if(!a->wk && hipMalloc(&a->wk, sizeof(double) * 4 * (size_t) n) != hipSuccess)
	{
		a->wk = NULL;
		return CPIC_ENOMEM;
	}
	a->evrec &= ~2;
	HIPCHK(hipEventRecord(a->ev[2], a->d->stream));
	// both lE buffers start at zero; each iteration's k_adj_field zeroes the
	// one the next iteration deposits into
	if((ret = zero(a, a->lx, n)) || (ret = zero(a, a->ly, n)) || (ret = zero(a, a->lux, n)) ||
			(ret = zero(a, a->luy, n)) || (ret = zero(a, (double *) a->lE, 4 * nodes)))
		return ret;
	// keep the forward's final state: the recompute runs on copies
	double *fx = a->x, *fy = a->y, *fux = a->ux, *fuy = a->uy;
	double *wk = a->wk;
	a->x = wk;
	a->y = wk + n;
	a->ux = wk + 2 * n;
	a->uy = wk + 3 * n;
	ret = 0;
	// rho and E of iteration it (in the segment starting at it0), read by
	// the reverse iteration: the forward's field history, or recomputed with
	// the segment's positions below (no deposit or solve again in the
	// reverse iteration either way)
	double *const segf = a->segf, *const frho = a->frho;
	double2 *const fE = (double2 *) a->fE;
	auto seg_rho = [=](int64_t it, int64_t it0) {
		return fh ? frho + nodes * it : segf + 3 * nodes * (it - it0);
	};
	auto seg_E = [=](int64_t it, int64_t it0) {
		return fh ? fE + nodes * it : (double2 *) (segf + 3 * nodes * (it - it0) + nodes);
	};
	for(int64_t sg = nck - 1; sg >= 0 && !ret; sg--)
	{
		int64_t it0 = sg * K, it1 = it0 + K < T ? it0 + K : T;
		// positions (x, then y) of iteration it: the checkpoint's for the
		// segment's first, slot it - it0 - 1 of the segment buffer, which the
		// recompute's pushes write, for the others
		const double *c = a->ck + 4 * n * sg;
		auto slot = [&](int64_t it) -> double * { return a->seg + 2 * n * (it - it0 - 1); };
		auto pos = [&](int64_t it) -> const double * { return it == it0 ? c : slot(it); };
		// velocities only for a segment that moves its particles
		if(it1 - it0 > 1 &&
				(hipMemcpyAsync(a->ux, c + 2 * n, sizeof(double) * n, hipMemcpyDeviceToDevice,
						a->d->stream) != hipSuccess ||
				hipMemcpyAsync(a->uy, c + 3 * n, sizeof(double) * n, hipMemcpyDeviceToDevice,
						a->d->stream) != hipSuccess))
		{
			ret = CPIC_EHIP;
			break;
		}
		if(fh)
		{
			// positions only, moved by the history's E
			for(int64_t it = it0; it + 1 < it1 && !ret; it++)
				ret = move_step(a, it, pos(it), pos(it) + n, slot(it + 1), slot(it + 1) + n,
						seg_E(it, it0));
		}
		else
		{
			if((ret = deposit(a, pos(it0), pos(it0) + n, seg_rho(it0, it0))))
				break;
			for(int64_t it = it0; it < it1 && !ret; it++)
			{
				if(it + 1 < it1)
					ret = forward_step(a, it, pos(it), pos(it) + n, slot(it + 1), slot(it + 1) + n,
							seg_rho(it, it0), seg_rho(it + 1, it0), seg_E(it, it0),
							(double *) NULL);
				else
					ret = fields(a, seg_rho(it, it0), (double *) NULL, (double *) NULL,
							seg_E(it, it0));
			}
		}
		for(int64_t it = it1 - 1; it >= it0 && !ret; it--)
		{
			const double *sx = pos(it), *sy = sx + n;
			int first = it == 0;
			double dtp = first ? -a->g.dt / 2 : a->g.dt;
			double2 *lE = a->lE + (it & 1) * nodes, *lE_next = a->lE + ((it + 1) & 1) * nodes;
			dim3 rg((unsigned) ((n + REV_CHUNK - 1) / REV_CHUNK));
			if(it + 1 < T)
			{
				// the lrho term of iteration it + 1, at its positions: the
				// segment's next slot, or the next checkpoint's
				const double *bx = it + 1 < it1 ? pos(it + 1) : a->ck + 4 * n * (sg + 1);
				hipLaunchKernelGGL(k_adj_rev_a<true>, rg, dim3(CPIC_BLOCK), 0, a->d->stream,
						a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, n, a->sp, sx, sy,
						(const double2 *) seg_E(it, it0), a->lx, a->ly, a->lux, a->luy, lE,
						bx, bx + n, (const double *) a->lrho);
			}
			else
				hipLaunchKernelGGL(k_adj_rev_a<false>, rg, dim3(CPIC_BLOCK), 0, a->d->stream,
						a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, n, a->sp, sx, sy,
						(const double2 *) seg_E(it, it0), a->lx, a->ly, a->lux, a->luy, lE,
						(const double *) NULL, (const double *) NULL, (const double *) NULL);
			hipLaunchKernelGGL(k_adj_field, dim3(pblocks(nodes)), dim3(CPIC_BLOCK), 0,
					a->d->stream, a->g, (const double2 *) lE, seg_rho(it, it0),
					dloss * a->g.e0, a->psi, lE_next);
			if(hipGetLastError() != hipSuccess)
			{
				ret = CPIC_EHIP;
				break;
			}
			if((ret = cpic_solve_ptr(a->d, a->psi, a->lrho)))
				break;
		}
	}
	// the lrho term of iteration 0 (positions: the first checkpoint's)
	if(!ret)
	{
		hipLaunchKernelGGL(k_adj_rev_b, dim3(pblocks(n)), dim3(CPIC_BLOCK), 0, a->d->stream,
				a->g, a->k[1], n, a->sp, (const double *) a->ck, (const double *) a->ck + n,
				(const double *) a->lrho, a->lx, a->ly);
		if(hipGetLastError() != hipSuccess)
			ret = CPIC_EHIP;
	}
	if(!ret && hipEventRecord(a->ev[3], a->d->stream) == hipSuccess)
		a->evrec |= 2;
	hipStreamSynchronize(a->d->stream);
	a->x = fx;
	a->y = fy;
	a->ux = fux;
	a->uy = fuy;
	if(ret)
		return ret;
	double *src[4] = {a->lx, a->ly, a->lux, a->luy};
	double *dst[4] = {gx, gy, gux, guy};
	if((ret = download_unsorted(a, src, dst)))
		return ret;
	return adj_derr(a);
}
