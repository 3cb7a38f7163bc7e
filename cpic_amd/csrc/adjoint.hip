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
// forward; the reverse recomputes each segment from its checkpoint, storing
// the positions of its iterations.  No counterpart in the reference, which
// has no gradients.
#include "dev.h"
#include "picmath.h"

#include <cmath>
#include <stdlib.h>
#include <string.h>

#define NRED 256          // partial sums of the field-energy reduction

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
	int32_t *sp;              // species of each particle
	double *x, *y, *ux, *uy;  // state
	double *rho[2], *phi, *psi, *lrho;
	double2 *E, *lE;
	double *lx, *ly, *lux, *luy;
	double *ck;               // checkpoints, 4n doubles each
	int64_t ckcap;
	double *seg;              // positions of one segment's iterations, 2n doubles each
	int64_t segcap;
	double *hist;             // W per iteration
	int64_t histcap;
	double *part;             // NRED partial sums
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

// rho += c * w at the 4 nodes (order of src/interpolate.c:259-262)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_deposit(Geo g, AdjSpecies k, int64_t n, const int32_t *__restrict__ sp,
		const double *__restrict__ x, const double *__restrict__ y, double *__restrict__ rho)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int i00, i01, i10, i11;
		double rx, ry;
		cic_nodes(g, x[i], y[i], i00, i01, i10, i11, rx, ry);
		double c = k.c[sp[i]];
		double dxw = 1.0 - rx, dyw = 1.0 - ry;
		unsafeAtomicAdd(&rho[i00], (dxw * dyw) * c);
		unsafeAtomicAdd(&rho[i01], (dxw * ry) * c);
		unsafeAtomicAdd(&rho[i10], (rx * dyw) * c);
		unsafeAtomicAdd(&rho[i11], (rx * ry) * c);
	}
}

// gather + Boris + move + wrap, then deposit at the new position into rho_next
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_push(Geo g, AdjSpecies k, int set_r, double dtp, int64_t n,
		const int32_t *__restrict__ sp, double *__restrict__ x, double *__restrict__ y,
		double *__restrict__ ux, double *__restrict__ uy, const double2 *__restrict__ E,
		double *__restrict__ rho_next, int *derr)
{
	bool bad = false, nan_ = false;
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int s = sp[i];
		int i00, i01, i10, i11;
		double rx, ry;
		double px = x[i], py = y[i];
		cic_nodes(g, px, py, i00, i01, i10, i11, rx, ry);
		double dxw = 1.0 - rx, dyw = 1.0 - ry;
		double w00 = dxw * dyw, w01 = dxw * ry, w10 = rx * dyw, w11 = rx * ry;
		double2 e00 = E[i00], e01 = E[i01], e10 = E[i10], e11 = E[i11];
		double ex = w00 * e00.x;
		ex = fma(w01, e01.x, ex);
		ex = fma(w10, e10.x, ex);
		ex = fma(w11, e11.x, ex);
		double ey = w00 * e00.y;
		ey = fma(w01, e01.y, ey);
		ey = fma(w10, e10.y, ey);
		ey = fma(w11, e11.y, ey);
		double u = ux[i], v = uy[i];
		boris_bz<double>(k.h[s], k.tz[s], k.sz[s], ex, ey, u, v);
		if(fabs(u) > g.umax_x || fabs(v) > g.umax_y)
			bad = true;
		if(set_r)
		{
			px = wrap64(fma(u, dtp, px), g.L0);
			py = wrap64(fma(v, dtp, py), g.L1);
		}
		if(!(isfinite(px) && isfinite(py) && isfinite(u) && isfinite(v)))
		{
			nan_ = true;
			px = py = 0.0;
		}
		x[i] = px;
		y[i] = py;
		ux[i] = u;
		uy[i] = v;
		cic_nodes(g, px, py, i00, i01, i10, i11, rx, ry);
		double c = k.c[s];
		dxw = 1.0 - rx;
		dyw = 1.0 - ry;
		unsafeAtomicAdd(&rho_next[i00], (dxw * dyw) * c);
		unsafeAtomicAdd(&rho_next[i01], (dxw * ry) * c);
		unsafeAtomicAdd(&rho_next[i10], (rx * dyw) * c);
		unsafeAtomicAdd(&rho_next[i11], (rx * ry) * c);
	}
	if(bad)
		atomicOr(derr, CPIC_DERR_UMAX);
	if(nan_)
		atomicOr(derr, CPIC_DERR_NAN);
}

// part[b] = sum over block b's contiguous range of a*b (fixed order)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_dot(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
		double *__restrict__ part)
{
	__shared__ double red[CPIC_BLOCK];
	int64_t per = (n + gridDim.x - 1) / gridDim.x;
	int64_t i0 = blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
	double s = 0.0;
	for(int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
		s += a[i] * b[i];
	red[threadIdx.x] = s;
	__syncthreads();
	for(int w = CPIC_BLOCK / 2; w > 0; w >>= 1)
	{
		if((int) threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if(threadIdx.x == 0)
		part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(NRED)
void k_adj_sum(const double *__restrict__ part, double scale, double *__restrict__ out)
{
	__shared__ double red[NRED];
	red[threadIdx.x] = part[threadIdx.x];
	__syncthreads();
	for(int w = NRED / 2; w > 0; w >>= 1)
	{
		if((int) threadIdx.x < w)
			red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if(threadIdx.x == 0)
		*out = scale * red[0];
}

// Reverse of move + Boris + gather for one iteration; accumulates the
// gather-weight derivative into (lx, ly) and deposits lE_p into lE.
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_rev_a(Geo g, AdjSpecies k, int set_r, double dtp, int64_t n,
		const int32_t *__restrict__ sp, const double *__restrict__ x,
		const double *__restrict__ y, const double2 *__restrict__ E, double *__restrict__ lx,
		double *__restrict__ ly, double *__restrict__ lux, double *__restrict__ luy,
		double2 *__restrict__ lE)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int s = sp[i];
		double mx = lux[i], my = luy[i];
		double ax = lx[i], ay = ly[i];
		if(set_r)
		{
			mx = fma(dtp, ax, mx);
			my = fma(dtp, ay, my);
		}
		double h = k.h[s], tz = k.tz[s], sz = k.sz[s];
		double a_ = 1.0 - tz * sz, b_ = sz;
		double nux = a_ * mx - b_ * my;      // M^T mu
		double nuy = b_ * mx + a_ * my;
		double lex = h * (nux + mx), ley = h * (nuy + my);
		lux[i] = nux;
		luy[i] = nuy;

		int i00, i01, i10, i11;
		double rx, ry;
		cic_nodes(g, x[i], y[i], i00, i01, i10, i11, rx, ry);
		double dxw = 1.0 - rx, dyw = 1.0 - ry;
		double2 e00 = E[i00], e01 = E[i01], e10 = E[i10], e11 = E[i11];
		// d/dx and d/dy of the CIC interpolants of Ex and Ey
		double gxx = g.idx * (dyw * (e10.x - e00.x) + ry * (e11.x - e01.x));
		double gyx = g.idy * (dxw * (e01.x - e00.x) + rx * (e11.x - e10.x));
		double gxy = g.idx * (dyw * (e10.y - e00.y) + ry * (e11.y - e01.y));
		double gyy = g.idy * (dxw * (e01.y - e00.y) + rx * (e11.y - e10.y));
		lx[i] = ax + (lex * gxx + ley * gxy);
		ly[i] = ay + (lex * gyx + ley * gyy);

		double w00 = dxw * dyw, w01 = dxw * ry, w10 = rx * dyw, w11 = rx * ry;
		unsafeAtomicAdd(&lE[i00].x, w00 * lex);
		unsafeAtomicAdd(&lE[i00].y, w00 * ley);
		unsafeAtomicAdd(&lE[i01].x, w01 * lex);
		unsafeAtomicAdd(&lE[i01].y, w01 * ley);
		unsafeAtomicAdd(&lE[i10].x, w10 * lex);
		unsafeAtomicAdd(&lE[i10].y, w10 * ley);
		unsafeAtomicAdd(&lE[i11].x, w11 * lex);
		unsafeAtomicAdd(&lE[i11].y, w11 * ley);
	}
}

// psi = D^T lE - srho * rho, D the central-difference E stencil
// (E_x[i] = (phi[i-1] - phi[i+1]) / 2dx, so (D_x^T a)[i] = (a[i+1] - a[i-1]) / 2dx)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_field(Geo g, const double2 *__restrict__ lE, const double *__restrict__ rho,
		double srho, double *__restrict__ psi)
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
	}
}

// lr += c * grad Gat(lrho)(r)
__global__ __launch_bounds__(CPIC_BLOCK)
void k_adj_rev_b(Geo g, AdjSpecies k, int64_t n, const int32_t *__restrict__ sp,
		const double *__restrict__ x, const double *__restrict__ y,
		const double *__restrict__ lrho, double *__restrict__ lx, double *__restrict__ ly)
{
	for(int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
			i += (int64_t) gridDim.x * blockDim.x)
	{
		int i00, i01, i10, i11;
		double rx, ry;
		cic_nodes(g, x[i], y[i], i00, i01, i10, i11, rx, ry);
		double c = k.c[sp[i]];
		double l00 = lrho[i00], l01 = lrho[i01], l10 = lrho[i10], l11 = lrho[i11];
		double dxw = 1.0 - rx, dyw = 1.0 - ry;
		lx[i] += c * (g.idx * (dyw * (l10 - l00) + ry * (l11 - l01)));
		ly[i] += c * (g.idy * (dxw * (l01 - l00) + rx * (l11 - l10)));
	}
}

// ---------------------------------------------------------------- host ---

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
	hipFree(a->ck); hipFree(a->seg); hipFree(a->hist); hipFree(a->part); hipFree(a->derr);
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
			hipMalloc(&a->E, 2 * gn) != hipSuccess || hipMalloc(&a->lE, 2 * gn) != hipSuccess ||
			hipMalloc(&a->part, sizeof(double) * NRED) != hipSuccess ||
			hipMalloc(&a->derr, sizeof(int)) != hipSuccess)
	{
		adj_free(a);
		return CPIC_ENOMEM;
	}
	if(hipMemcpy(a->sp, species, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess)
	{
		adj_free(a);
		return CPIC_EHIP;
	}
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
	hipLaunchKernelGGL(k_adj_deposit, dim3(pblocks(a->n)), dim3(CPIC_BLOCK), 0, a->d->stream,
			a->g, a->k[1], a->n, a->sp, x, y, rho);
	HIPCHK(hipGetLastError());
	return 0;
}

// phi and E from rho; with w, the field energy of rho into *w
static int
fields(cpic_adj *a, const double *rho, double *w)
{
	int ret;
	int64_t nodes = (int64_t) a->g.nx * a->g.ny;
	if((ret = cpic_solve_ptr(a->d, rho, a->phi)))
		return ret;
	if(w)
	{
		hipLaunchKernelGGL(k_adj_dot, dim3(NRED), dim3(CPIC_BLOCK), 0, a->d->stream, rho, a->phi,
				nodes, a->part);
		hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(NRED), 0, a->d->stream, a->part,
				-0.5 * a->g.e0, w);
		HIPCHK(hipGetLastError());
	}
	return cpic_efield_ptr(a->d, a->phi, a->E);
}

// one forward iteration from the state in (x, y, ux, uy) and rho[cur] = Dep(r)
static int
forward_iter(cpic_adj *a, int64_t it, int *cur, double *w)
{
	int ret;
	int c = *cur;
	if((ret = fields(a, a->rho[c], w)))
		return ret;
	if((ret = zero(a, a->rho[1 - c], (int64_t) a->g.nx * a->g.ny)))
		return ret;
	int first = it == 0;
	double dtp = first ? -a->g.dt / 2 : a->g.dt;
	hipLaunchKernelGGL(k_adj_push, dim3(pblocks(a->n)), dim3(CPIC_BLOCK), 0, a->d->stream,
			a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, a->n, a->sp, a->x, a->y, a->ux,
			a->uy, (const double2 *) a->E, a->rho[1 - c], a->derr);
	HIPCHK(hipGetLastError());
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

extern "C" int
cpic_adj_forward(cpic_adj_t *a, const double *x, const double *y, const double *ux,
		const double *uy, int64_t nsteps, int32_t checkpoint_every, double *loss,
		double *energy)
{
	int ret, cur = 0;
	if(!a || !x || !y || !ux || !uy || nsteps < 1 || checkpoint_every < 1)
		return CPIC_EINVAL;
	for(int64_t i = 0; i < a->n; i++)
		if(!(x[i] >= 0.0 && x[i] <= a->g.L0 && y[i] >= 0.0 && y[i] <= a->g.L1) ||
				!std::isfinite(ux[i]) || !std::isfinite(uy[i]))
			return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	int64_t K = checkpoint_every, nck = (nsteps + K - 1) / K, n = a->n;
	if((ret = grow(&a->ck, &a->ckcap, 4 * n * nck)) || (ret = grow(&a->hist, &a->histcap, nsteps)))
		return ret;
	size_t pn = sizeof(double) * (size_t) n;
	HIPCHK(hipMemcpyAsync(a->x, x, pn, hipMemcpyHostToDevice, a->d->stream));
	HIPCHK(hipMemcpyAsync(a->y, y, pn, hipMemcpyHostToDevice, a->d->stream));
	HIPCHK(hipMemcpyAsync(a->ux, ux, pn, hipMemcpyHostToDevice, a->d->stream));
	HIPCHK(hipMemcpyAsync(a->uy, uy, pn, hipMemcpyHostToDevice, a->d->stream));
	HIPCHK(hipMemsetAsync(a->derr, 0, sizeof(int), a->d->stream));
	a->have = 0;
	if((ret = deposit(a, a->x, a->y, a->rho[0])))
		return ret;
	for(int64_t it = 0; it < nsteps; it++)
	{
		if(it % K == 0)
			if((ret = copy_state(a, a->ck + 4 * n * (it / K), a->x, a->y, a->ux, a->uy,
					hipMemcpyDeviceToDevice)))
				return ret;
		if((ret = forward_iter(a, it, &cur, a->hist + it)))
			return ret;
	}
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
cpic_adj_state(cpic_adj_t *a, double *x, double *y, double *ux, double *uy)
{
	if(!a || !a->have)
		return CPIC_EINVAL;
	hipSetDevice(a->d->device);
	size_t pn = sizeof(double) * (size_t) a->n;
	double *src[4] = {a->x, a->y, a->ux, a->uy};
	double *dst[4] = {x, y, ux, uy};
	for(int k = 0; k < 4; k++)
		if(dst[k])
			HIPCHK(hipMemcpyAsync(dst[k], src[k], pn, hipMemcpyDeviceToHost, a->d->stream));
	HIPCHK(hipStreamSynchronize(a->d->stream));
	return 0;
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
	if((ret = grow(&a->seg, &a->segcap, 2 * n * K)))
		return ret;
	if((ret = zero(a, a->lx, n)) || (ret = zero(a, a->ly, n)) || (ret = zero(a, a->lux, n)) ||
			(ret = zero(a, a->luy, n)))
		return ret;
	// keep the forward's final state: the recompute runs on copies
	double *fx = a->x, *fy = a->y, *fux = a->ux, *fuy = a->uy;
	double *wk = NULL;
	if(hipMalloc(&wk, sizeof(double) * 4 * (size_t) n) != hipSuccess)
		return CPIC_ENOMEM;
	a->x = wk;
	a->y = wk + n;
	a->ux = wk + 2 * n;
	a->uy = wk + 3 * n;
	ret = 0;
	for(int64_t sg = nck - 1; sg >= 0 && !ret; sg--)
	{
		int64_t it0 = sg * K, it1 = it0 + K < T ? it0 + K : T;
		int cur = 0;
		const double *c = a->ck + 4 * n * sg;
		if((ret = copy_state(a, wk, c, c + n, c + 2 * n, c + 3 * n, hipMemcpyDeviceToDevice)))
			break;
		if((ret = deposit(a, a->x, a->y, a->rho[0])))
			break;
		for(int64_t it = it0; it < it1 && !ret; it++)
		{
			double *sx = a->seg + 2 * n * (it - it0);
			if(hipMemcpyAsync(sx, a->x, sizeof(double) * n, hipMemcpyDeviceToDevice,
						a->d->stream) != hipSuccess ||
					hipMemcpyAsync(sx + n, a->y, sizeof(double) * n, hipMemcpyDeviceToDevice,
						a->d->stream) != hipSuccess)
				ret = CPIC_EHIP;
			else if(it + 1 < it1)
				ret = forward_iter(a, it, &cur, (double *) NULL);
		}
		for(int64_t it = it1 - 1; it >= it0 && !ret; it--)
		{
			const double *sx = a->seg + 2 * n * (it - it0), *sy = sx + n;
			int first = it == 0;
			double dtp = first ? -a->g.dt / 2 : a->g.dt;
			if((ret = deposit(a, sx, sy, a->rho[0])) || (ret = fields(a, a->rho[0], NULL)) ||
					(ret = zero(a, (double *) a->lE, 2 * nodes)))
				break;
			hipLaunchKernelGGL(k_adj_rev_a, dim3(pblocks(n)), dim3(CPIC_BLOCK), 0, a->d->stream,
					a->g, a->k[first ? 0 : 1], first ? 0 : 1, dtp, n, a->sp, sx, sy,
					(const double2 *) a->E, a->lx, a->ly, a->lux, a->luy, a->lE);
			hipLaunchKernelGGL(k_adj_field, dim3(pblocks(nodes)), dim3(CPIC_BLOCK), 0,
					a->d->stream, a->g, (const double2 *) a->lE, a->rho[0], dloss * a->g.e0,
					a->psi);
			if(hipGetLastError() != hipSuccess)
			{
				ret = CPIC_EHIP;
				break;
			}
			if((ret = cpic_solve_ptr(a->d, a->psi, a->lrho)))
				break;
			hipLaunchKernelGGL(k_adj_rev_b, dim3(pblocks(n)), dim3(CPIC_BLOCK), 0, a->d->stream,
					a->g, a->k[1], n, a->sp, sx, sy, a->lrho, a->lx, a->ly);
			if(hipGetLastError() != hipSuccess)
				ret = CPIC_EHIP;
		}
	}
	hipStreamSynchronize(a->d->stream);
	a->x = fx;
	a->y = fy;
	a->ux = fux;
	a->uy = fuy;
	hipFree(wk);
	if(ret)
		return ret;
	size_t pn = sizeof(double) * (size_t) n;
	double *src[4] = {a->lx, a->ly, a->lux, a->luy};
	double *dst[4] = {gx, gy, gux, guy};
	for(int k = 0; k < 4; k++)
		if(dst[k])
			HIPCHK(hipMemcpyAsync(dst[k], src[k], pn, hipMemcpyDeviceToHost, a->d->stream));
	HIPCHK(hipStreamSynchronize(a->d->stream));
	return adj_derr(a);
}
