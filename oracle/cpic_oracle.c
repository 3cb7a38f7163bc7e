/*
 * cpic_oracle.c -- FP64 CPU restatement of cpic's per-timestep PIC loop.
 *
 * TEST INFRASTRUCTURE ONLY (it checks the GPU results): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Never part of
 * the shipped path.  Built with -ffp-contract=off; see cpic_oracle.h for the
 * explicit fma() policy.
 *
 * Storage is plain SoA arrays per (chunk, species) instead of the
 * reference's pblock/ppack lists (src/def.h:88-227); the particle order
 * inside a chunk after migration is a stable compaction followed by the
 * appended neighbour queues, which only changes the summation order of rho
 * (the reference's own commutative chunk tasks leave that order undefined,
 * src/field.c:304).
 *
 * Documented deviation shared with the HIP path: a position that the
 * periodic wrap rounds to exactly L (x in (-ulp, 0) plus L) is mapped to 0,
 * and a cell index equal to n (x*idx rounding up to n) is wrapped to 0.  The
 * reference keeps x == L (src/comm_plasma.c:738-745) and then indexes one
 * past the last column/row (src/interpolate.c:113-125).
 */
#define _GNU_SOURCE
#include "cpic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { X = 0, Y = 1, Z = 2 };

/* ------------------------------------------------------------------ */
/* particle lists                                                       */
/* ------------------------------------------------------------------ */

typedef struct plist {
	int64_t n, cap;
	int64_t *id;
	double *x, *y, *ux, *uy, *Ex, *Ey;
} plist_t;

struct orc_sim {
	orc_params_t p;
	int64_t nx, ny;
	double dx[2];       /* sim->dx (src/sim.c:174) */
	double fL[2];       /* field L = dx*shape (src/field.c:45-46) */
	double umax[3];     /* src/sim.c:198-200 */
	int64_t chunksize;  /* chunksize[X] (src/sim.c:185) */
	int64_t rs;         /* rho/phi row stride 2*(nx/2+1) (src/solver.c:403) */
	double *rho;        /* rows 0..ny (ghost row ny) */
	double *phi_;       /* rows -1..ny+1, i.e. ny+3 rows */
	double *E[2];       /* nx x (ny+1) */
	double *G;          /* (nx/2+1) x ny */
	double *ghat;       /* ny x (nx/2+1) complex, interleaved */
	plist_t *list;      /* [chunk*nspecies + species] */
	plist_t *q0, *q1;   /* exchange queues per (chunk, species) */
	int64_t iter;
	int64_t nlist;
};

static int
plist_reserve(plist_t *l, int64_t need)
{
	int64_t cap;
	if(need <= l->cap)
		return 0;
	cap = l->cap ? l->cap : 16;
	while(cap < need)
		cap *= 2;
#define GROW(f, T) do { \
		T *t = realloc(l->f, (size_t) cap * sizeof(T)); \
		if(!t) \
			return ORC_ENOMEM; \
		l->f = t; \
	} while(0)
	GROW(id, int64_t);
	GROW(x, double);
	GROW(y, double);
	GROW(ux, double);
	GROW(uy, double);
	GROW(Ex, double);
	GROW(Ey, double);
#undef GROW
	l->cap = cap;
	return 0;
}

static void
plist_free(plist_t *l)
{
	free(l->id); free(l->x); free(l->y); free(l->ux); free(l->uy);
	free(l->Ex); free(l->Ey);
	memset(l, 0, sizeof(*l));
}

/* append particle j of src to dst */
static int
plist_push(plist_t *dst, const plist_t *src, int64_t j)
{
	int64_t k;
	if(plist_reserve(dst, dst->n + 1))
		return ORC_ENOMEM;
	k = dst->n++;
	dst->id[k] = src->id[j];
	dst->x[k] = src->x[j];
	dst->y[k] = src->y[j];
	dst->ux[k] = src->ux[j];
	dst->uy[k] = src->uy[j];
	dst->Ex[k] = src->Ex[j];
	dst->Ey[k] = src->Ey[j];
	return 0;
}

/* ------------------------------------------------------------------ */
/* FFT: iterative radix-2, twiddles w_k = exp(-2 pi i k/n) from libm      */
/* ------------------------------------------------------------------ */

typedef struct twcache {
	int64_t n;
	double *c, *s;   /* cos(2 pi k/n), sin(2 pi k/n), k <= n/2 */
} twcache_t;

/* One slot per power of two, filled once and never evicted. */
static twcache_t twc[64];

/* cos/sin(2 pi k/n) by octant reduction, so that the quarter points are
 * exact (0, +-1) and the two halves of each octant are mirror images. */
static void
twiddle(int64_t k, int64_t n, double *c, double *s)
{
	int64_t k8 = 8 * k;
	double a;
	if(k8 <= n)
	{
		a = 2.0 * M_PI * (double) k / (double) n;
		*c = cos(a); *s = sin(a);
	}
	else if(k8 <= 2 * n)        /* (pi/4, pi/2] */
	{
		a = 2.0 * M_PI * (double) (n / 4 - k) / (double) n;
		*c = sin(a); *s = cos(a);
	}
	else if(k8 <= 3 * n)        /* (pi/2, 3pi/4] */
	{
		a = 2.0 * M_PI * (double) (k - n / 4) / (double) n;
		*c = -sin(a); *s = cos(a);
	}
	else                        /* (3pi/4, pi] */
	{
		a = 2.0 * M_PI * (double) (n / 2 - k) / (double) n;
		*c = -cos(a); *s = sin(a);
	}
}

static const twcache_t *
twiddles(int64_t n)
{
	int lg = 0;
	int64_t k;
	twcache_t *t;

	while(((int64_t) 1 << lg) < n)
		lg++;
	t = &twc[lg];
	if(__atomic_load_n(&t->n, __ATOMIC_ACQUIRE) == n)
		return t;
#ifdef _OPENMP
#pragma omp critical(orc_twiddles)
#endif
	{
		if(__atomic_load_n(&t->n, __ATOMIC_ACQUIRE) != n)
		{
			t->c = malloc(sizeof(double) * (size_t) (n / 2 + 1));
			t->s = malloc(sizeof(double) * (size_t) (n / 2 + 1));
			for(k = 0; k <= n / 2; k++)
				twiddle(k, n, &t->c[k], &t->s[k]);
			__atomic_store_n(&t->n, n, __ATOMIC_RELEASE);
		}
	}
	return t;
}

static int
is_pow2(int64_t n)
{
	return n > 0 && (n & (n - 1)) == 0;
}

int
orc_fft(int64_t n, double *re, double *im, int sign)
{
	int64_t i, j, len, half, step, k, bit;
	const twcache_t *tw;

	if(!is_pow2(n))
		return ORC_EINVAL;
	if(n == 1)
		return 0;
	tw = twiddles(n);

	/* bit reversal */
	for(i = 1, j = 0; i < n; i++)
	{
		for(bit = n >> 1; j & bit; bit >>= 1)
			j ^= bit;
		j ^= bit;
		if(i < j)
		{
			double t;
			t = re[i]; re[i] = re[j]; re[j] = t;
			t = im[i]; im[i] = im[j]; im[j] = t;
		}
	}

	for(len = 2; len <= n; len <<= 1)
	{
		half = len >> 1;
		step = n / len;
		for(i = 0; i < n; i += len)
		{
			for(k = 0; k < half; k++)
			{
				/* w = exp(sign * 2 pi i k/len) */
				double wr = tw->c[k * step];
				double wi = (double) sign * tw->s[k * step];
				double ar = re[i + k], ai = im[i + k];
				double br = re[i + k + half], bi = im[i + k + half];
				double vr = br * wr - bi * wi;
				double vi = br * wi + bi * wr;
				re[i + k] = ar + vr;
				im[i + k] = ai + vi;
				re[i + k + half] = ar - vr;
				im[i + k + half] = ai - vi;
			}
		}
	}
	return 0;
}

/* Real forward transform of one row (length n, even): out[k] for k=0..n/2,
 * computed with the half-length complex packing z[j] = x[2j] + i x[2j+1]. */
static void
rfft_row(int64_t n, const double *x, double *ore, double *oim, double *wre,
		double *wim)
{
	int64_t m = n / 2, k, j;
	const twcache_t *tw;

	if(n == 1)
	{
		ore[0] = x[0]; oim[0] = 0.0;
		return;
	}
	for(j = 0; j < m; j++)
	{
		wre[j] = x[2 * j];
		wim[j] = x[2 * j + 1];
	}
	orc_fft(m, wre, wim, -1);
	tw = twiddles(n);
	for(k = 0; k <= m; k++)
	{
		int64_t ka = (k == m) ? 0 : k;
		int64_t kb = (m - k == m) ? 0 : m - k;
		/* E = (Z[k] + conj Z[m-k])/2, O = (Z[k] - conj Z[m-k])/(2i) */
		double zr = wre[ka], zi = wim[ka];
		double cr = wre[kb], ci = -wim[kb];
		double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
		double dr = 0.5 * (zr - cr), di = 0.5 * (zi - ci);
		/* O = d / i = (di, -dr) */
		double or_ = di, oi = -dr;
		/* w = exp(-2 pi i k/n) */
		double wr = tw->c[k], wi = -tw->s[k];
		ore[k] = er + (or_ * wr - oi * wi);
		oim[k] = ei + (or_ * wi + oi * wr);
	}
}

/* Real inverse transform (unnormalised) of one row: in[k] k=0..n/2 with the
 * imaginary parts of k=0 and k=n/2 ignored, out x[0..n-1]. */
static void
irfft_row(int64_t n, const double *ire, const double *iim, double *x,
		double *wre, double *wim)
{
	int64_t m = n / 2, k, j;
	const twcache_t *tw;

	if(n == 1)
	{
		x[0] = ire[0];
		return;
	}
	tw = twiddles(n);
	for(k = 0; k < m; k++)
	{
		double ar = ire[k], ai = (k == 0) ? 0.0 : iim[k];
		double br = ire[m - k], bi = (m - k == m) ? 0.0 : iim[m - k];
		/* conj(X[m-k]) */
		double cr = br, ci = -bi;
		double Ar = ar + cr, Ai = ai + ci;
		double Dr = ar - cr, Di = ai - ci;
		/* B = D * exp(+2 pi i k/n) */
		double wr = tw->c[k], wi = tw->s[k];
		double Br = Dr * wr - Di * wi, Bi = Dr * wi + Di * wr;
		/* Z = A + i B */
		wre[k] = Ar - Bi;
		wim[k] = Ai + Br;
	}
	orc_fft(m, wre, wim, +1);
	for(j = 0; j < m; j++)
	{
		x[2 * j] = wre[j];
		x[2 * j + 1] = wim[j];
	}
}

/* MFT coefficients (src/solver.c:269-280): G[ky][kx], kx = 0..nx/2. */
static void
mft_coefficients(int64_t nx, int64_t ny, double *G)
{
	int64_t ix, iy, hx = nx / 2 + 1;
	double cx = 2.0 * M_PI / nx, cy = 2.0 * M_PI / ny;

	for(iy = 0; iy < ny; iy++)
		for(ix = 0; ix < hx; ix++)
		{
			if(ix == 0 && iy == 0)
				G[iy * hx + ix] = 0.0;
			else
				G[iy * hx + ix] = 1.0 /
					(2.0 * (cos(cx * ix) + cos(cy * iy)) - 4.0);
		}
}

/* MFT_solve (src/solver.c:466-509): r2c 2D, ghat *= G, c2r 2D, /(nx*ny).
 * rho: rows 0..ny-1 with stride rs_in; phi rows with stride rs_out. */
static int
mft_solve_strided(int64_t nx, int64_t ny, const double *rho, int64_t rs_in,
		double *phi, int64_t rs_out, const double *G, double *ghat)
{
	int64_t hx = nx / 2 + 1;
	double N = (double) (int) (nx * ny);
	int64_t iy, kx;

	if(!is_pow2(nx) || !is_pow2(ny))
		return ORC_EINVAL;

	/* rows: r2c along X */
#ifdef _OPENMP
#pragma omp parallel
#endif
	{
		double *wre = malloc(sizeof(double) * (size_t) (nx + ny + 2));
		double *wim = malloc(sizeof(double) * (size_t) (nx + ny + 2));
		double *ore = malloc(sizeof(double) * (size_t) (hx + ny));
		double *oim = malloc(sizeof(double) * (size_t) (hx + ny));
		double *col_re = malloc(sizeof(double) * (size_t) ny);
		double *col_im = malloc(sizeof(double) * (size_t) ny);

#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for(iy = 0; iy < ny; iy++)
		{
			int64_t k;
			rfft_row(nx, rho + iy * rs_in, ore, oim, wre, wim);
			for(k = 0; k < hx; k++)
			{
				ghat[2 * (iy * hx + k)] = ore[k];
				ghat[2 * (iy * hx + k) + 1] = oim[k];
			}
		}

		/* columns: forward along Y, multiply by G, inverse along Y */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for(kx = 0; kx < hx; kx++)
		{
			int64_t y;
			for(y = 0; y < ny; y++)
			{
				col_re[y] = ghat[2 * (y * hx + kx)];
				col_im[y] = ghat[2 * (y * hx + kx) + 1];
			}
			orc_fft(ny, col_re, col_im, -1);
			/* MFT_kernel (src/solver.c:338-363): g *= G */
			for(y = 0; y < ny; y++)
			{
				double g = G[y * hx + kx];
				col_re[y] = col_re[y] * g;
				col_im[y] = col_im[y] * g;
			}
			orc_fft(ny, col_re, col_im, +1);
			for(y = 0; y < ny; y++)
			{
				ghat[2 * (y * hx + kx)] = col_re[y];
				ghat[2 * (y * hx + kx) + 1] = col_im[y];
			}
		}

		/* rows: c2r along X, then MFT_normalize (src/solver.c:366-379) */
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
		for(iy = 0; iy < ny; iy++)
		{
			int64_t k, x;
			double *out = phi + iy * rs_out;
			for(k = 0; k < hx; k++)
			{
				ore[k] = ghat[2 * (iy * hx + k)];
				oim[k] = ghat[2 * (iy * hx + k) + 1];
			}
			irfft_row(nx, ore, oim, out, wre, wim);
			for(x = 0; x < nx; x++)
				out[x] /= N;
		}
		free(wre); free(wim); free(ore); free(oim);
		free(col_re); free(col_im);
	}
	return 0;
}

int
orc_mft_solve(int64_t nx, int64_t ny, const double *rho, double *phi)
{
	int ret;
	int64_t hx = nx / 2 + 1;
	double *G, *ghat;

	if(!is_pow2(nx) || !is_pow2(ny) || nx < 2 || ny < 2)
		return ORC_EINVAL;
	G = malloc(sizeof(double) * (size_t) (hx * ny));
	ghat = malloc(2 * sizeof(double) * (size_t) (hx * ny));
	if(!G || !ghat)
	{
		free(G); free(ghat);
		return ORC_ENOMEM;
	}
	mft_coefficients(nx, ny, G);
	ret = mft_solve_strided(nx, ny, rho, nx, phi, nx, G, ghat);
	free(G);
	free(ghat);
	return ret;
}

/* ------------------------------------------------------------------ */
/* interpolation                                                        */
/* ------------------------------------------------------------------ */

/* relative_position_grid (src/interpolate.c:39-66) with the documented
 * wrap of an index equal to n. */
static inline double
rel_pos(double x0, double x, double dx, double idx, int64_t n, int64_t *i0)
{
	double block_delta = x - x0;
	double block_rel = block_delta * idx;
	double block_start = floor(block_rel);
	double grid_delta = fma(-block_start, dx, block_delta);
	int64_t i = (int64_t) block_start;
	if(i >= n)
		i -= n;
	*i0 = i;
	return grid_delta * idx;
}

/* weights (src/interpolate.c:78-100): Y uses dx[X] in the middle term. */
static inline void
weights(const orc_sim_t *s, double x, double y, double w[2][2], int64_t i0[2])
{
	double relx, rely, delx, dely;
	double idx = 1.0 / s->dx[X], idy = 1.0 / s->dx[Y];

	relx = rel_pos(0.0, x, s->dx[X], idx, s->nx, &i0[X]);
	rely = rel_pos(0.0, y, s->dx[X], idy, s->ny, &i0[Y]);
	/* linear_interpolation (src/interpolate.c:12-32) */
	delx = 1.0 - relx;
	dely = 1.0 - rely;
	w[0][0] = delx * dely;
	w[0][1] = delx * rely;
	w[1][0] = relx * dely;
	w[1][1] = relx * rely;
}

void
orc_weights(double x, double y, double dx, double dy, double w[2][2],
		int64_t i0[2])
{
	double relx, rely, delx, dely;
	relx = rel_pos(0.0, x, dx, 1.0 / dx, INT64_MAX, &i0[X]);
	rely = rel_pos(0.0, y, dx, 1.0 / dy, INT64_MAX, &i0[Y]);
	delx = 1.0 - relx;
	dely = 1.0 - rely;
	w[0][0] = delx * dely;
	w[0][1] = delx * rely;
	w[1][0] = relx * dely;
	w[1][1] = relx * rely;
}

void
orc_cells(const orc_sim_t *s, int64_t n, const double *x, const double *y,
		int64_t *cell)
{
	int64_t i;
	for(i = 0; i < n; i++)
	{
		double w[2][2];
		int64_t i0[2];
		weights(s, x[i], y[i], w, i0);
		cell[i] = i0[Y] * s->nx + i0[X];
	}
}

/* interpolate_f2p (src/interpolate.c:103-155) for one field component:
 * mat is E[d] with stride nx and rows 0..ny. */
static inline double
gather1(const orc_sim_t *s, const double *mat, const double w[2][2],
		const int64_t i0[2])
{
	int64_t nx = s->nx;
	int64_t i1x = i0[X] + 1, i1y = i0[Y] + 1;
	double val;
	if(i1x >= nx)           /* vi64_remod (src/simd_avx2.h:36-46) */
		i1x -= nx;
	val = w[0][0] * mat[i0[Y] * nx + i0[X]];
	val = fma(w[0][1], mat[i1y * nx + i0[X]], val);
	val = fma(w[1][0], mat[i0[Y] * nx + i1x], val);
	val = fma(w[1][1], mat[i1y * nx + i1x], val);
	return val;
}

/* ------------------------------------------------------------------ */
/* Boris push                                                           */
/* ------------------------------------------------------------------ */

static inline void
cross(double r[3], const double a[3], const double b[3])
{
	/* src/mover.c:15-20 as contracted by clang: a*b - c*d ->
	 * fma(a, b, -(c*d)) */
	r[X] = fma(a[Y], b[Z], -(a[Z] * b[Y]));
	r[Y] = fma(a[Z], b[X], -(a[X] * b[Z]));
	r[Z] = fma(a[X], b[Y], -(a[Y] * b[X]));
}

/* boris_rotation (src/mover.c:23-71) */
static inline void
boris(const double u[3], const double E[3], const double B[3], double h,
		double un[3])
{
	double t[3], s[3], vm[3], vp[3], vpl[3], sden;
	int d;
	for(d = 0; d < 3; d++)
	{
		t[d] = B[d] * h;
		sden = fma(t[d], t[d], 1.0);
		vm[d] = fma(h, E[d], u[d]);
		s[d] = 2.0 * t[d] / sden;
	}
	cross(vp, vm, t);
	for(d = 0; d < 3; d++)
		vp[d] = vp[d] + vm[d];
	cross(vpl, vp, s);
	for(d = 0; d < 3; d++)
	{
		vpl[d] = vpl[d] + vm[d];
		un[d] = fma(h, E[d], vpl[d]);
	}
}

void
orc_boris(double r[3], double u[3], const double E[3], const double B[3],
		double q, double m, double dt, int set_r)
{
	double h = 0.5 * dt * q / m, un[3];
	int d;
	boris(u, E, B, h, un);
	if(set_r)
		for(d = 0; d < 3; d++)
			r[d] = fma(un[d], dt, r[d]);
	for(d = 0; d < 3; d++)
		u[d] = un[d];
}

/* ------------------------------------------------------------------ */
/* sim                                                                  */
/* ------------------------------------------------------------------ */

const char *
orc_strerror(int err)
{
	switch(err)
	{
		case ORC_OK: return "ok";
		case ORC_EINVAL: return "invalid parameters";
		case ORC_ENOMEM: return "out of memory";
		case ORC_EUMAX: return "max velocity exceeded";
		case ORC_ESTATE: return "call out of protocol order";
	}
	return "unknown error";
}

orc_sim_t *
orc_create(const orc_params_t *p, int *err)
{
	orc_sim_t *s;
	int64_t nx = p->nx, ny = p->ny, hx;
	int64_t i, nl;
	int e = ORC_EINVAL;

	if(nx < 2 || ny < 2 || !is_pow2(nx) || !is_pow2(ny))
		goto fail0;
	if(p->nchunks <= 0 || nx % p->nchunks)   /* src/sim.c:124-130 */
		goto fail0;
	if(p->nspecies <= 0 || p->nspecies > ORC_MAX_SPECIES)
		goto fail0;
	if(!(p->dt > 0.0) || !(p->L[0] > 0.0) || !(p->L[1] > 0.0))
		goto fail0;

	e = ORC_ENOMEM;
	s = calloc(1, sizeof(*s));
	if(!s)
		goto fail0;
	s->p = *p;
	s->nx = nx;
	s->ny = ny;
	s->dx[X] = p->L[X] / nx;
	s->dx[Y] = p->L[Y] / ny;
	s->fL[X] = s->dx[X] * nx;
	s->fL[Y] = s->dx[Y] * ny;
	s->chunksize = nx / p->nchunks;
	s->umax[X] = (double) s->chunksize * s->dx[X] / p->dt;
	s->umax[Y] = (double) ny * s->dx[Y] / p->dt;
	s->umax[Z] = 1.0 * 0.0 / p->dt;
	hx = nx / 2 + 1;
	s->rs = 2 * hx;
	s->rho = malloc(sizeof(double) * (size_t) (s->rs * (ny + 1)));
	s->phi_ = malloc(sizeof(double) * (size_t) (s->rs * (ny + 3)));
	s->E[X] = malloc(sizeof(double) * (size_t) (nx * (ny + 1)));
	s->E[Y] = malloc(sizeof(double) * (size_t) (nx * (ny + 1)));
	s->G = malloc(sizeof(double) * (size_t) (hx * ny));
	s->ghat = malloc(2 * sizeof(double) * (size_t) (hx * ny));
	nl = p->nchunks * p->nspecies;
	s->nlist = nl;
	s->list = calloc((size_t) nl, sizeof(plist_t));
	s->q0 = calloc((size_t) nl, sizeof(plist_t));
	s->q1 = calloc((size_t) nl, sizeof(plist_t));
	if(!s->rho || !s->phi_ || !s->E[X] || !s->E[Y] || !s->G || !s->ghat ||
			!s->list || !s->q0 || !s->q1)
		goto fail1;

	/* field_init NaN-fills every field (src/field.c:85,118,131) */
	for(i = 0; i < s->rs * (ny + 1); i++)
		s->rho[i] = NAN;
	for(i = 0; i < s->rs * (ny + 3); i++)
		s->phi_[i] = NAN;
	for(i = 0; i < nx * (ny + 1); i++)
		s->E[X][i] = s->E[Y][i] = NAN;

	mft_coefficients(nx, ny, s->G);
	s->iter = -1;
	if(err) *err = 0;
	return s;

fail1:
	orc_destroy(s);
fail0:
	if(err) *err = e;
	return NULL;
}

void
orc_destroy(orc_sim_t *s)
{
	int64_t i;
	if(!s)
		return;
	if(s->list)
		for(i = 0; i < s->nlist; i++)
		{
			plist_free(&s->list[i]);
			plist_free(&s->q0[i]);
			plist_free(&s->q1[i]);
		}
	free(s->list); free(s->q0); free(s->q1);
	free(s->rho); free(s->phi_); free(s->E[X]); free(s->E[Y]);
	free(s->G); free(s->ghat);
	free(s);
}

int64_t orc_iter(const orc_sim_t *s) { return s->iter; }
void orc_set_iter(orc_sim_t *s, int64_t iter) { s->iter = iter; }

static inline plist_t *
L_(const orc_sim_t *s, int64_t c, int sp)
{
	return &s->list[c * s->p.nspecies + sp];
}

int
orc_add_particles(orc_sim_t *s, int sp, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy)
{
	int64_t i;
	if(sp < 0 || sp >= s->p.nspecies || n < 0)
		return ORC_EINVAL;
	for(i = 0; i < n; i++)
	{
		plist_t *l = L_(s, id[i] % s->p.nchunks, sp);
		int64_t k;
		if(id[i] < 0 || !(x[i] >= 0.0 && x[i] <= s->p.L[X]) ||
				!(y[i] >= 0.0 && y[i] <= s->p.L[Y]))
			return ORC_EINVAL;
		if(plist_reserve(l, l->n + 1))
			return ORC_ENOMEM;
		k = l->n++;
		l->id[k] = id[i];
		l->x[k] = x[i];
		l->y[k] = y[i];
		l->ux[k] = ux[i];
		l->uy[k] = uy[i];
		/* particles_init sets E = 0 (src/particle.c:77-79) */
		l->Ex[k] = 0.0;
		l->Ey[k] = 0.0;
	}
	return 0;
}

/* --- stage_field_E (src/field.c:451-501) --- */

static double *
phi_row(const orc_sim_t *s, int64_t iy)  /* iy in -1..ny+1 */
{
	return s->phi_ + (iy + 1) * s->rs;
}

int
orc_stage_field_E(orc_sim_t *s)
{
	int64_t nx = s->nx, ny = s->ny, ix, iy;
	double dx2 = 2 * s->dx[X], dy2 = 2 * s->dx[Y];
	int ret;

	/* field_phi_solve -> solve_xy -> MFT_solve */
	ret = mft_solve_strided(nx, ny, s->rho, s->rs, phi_row(s, 0), s->rs,
			s->G, s->ghat);
	if(ret)
		return ret;

	/* comm_phi_send/recv with one rank (src/comm_field.c:140-201):
	 * phi[-1] = phi[ny-1], phi[ny] = phi[0], phi[ny+1] = phi[1] */
	memcpy(phi_row(s, -1), phi_row(s, ny - 1), sizeof(double) * (size_t) s->rs);
	memcpy(phi_row(s, ny), phi_row(s, 0), sizeof(double) * (size_t) s->rs);
	memcpy(phi_row(s, ny + 1), phi_row(s, 1), sizeof(double) * (size_t) s->rs);

	/* field_E_compute (src/field.c:359-416), rows 0..ny */
#ifdef _OPENMP
#pragma omp parallel for private(ix) schedule(static)
#endif
	for(iy = 0; iy <= ny; iy++)
	{
		const double *pm = phi_row(s, iy - 1), *pc = phi_row(s, iy);
		const double *pp = phi_row(s, iy + 1);
		for(ix = 0; ix < nx; ix++)
		{
			int64_t x0 = (ix + nx - 1) % nx, x1 = (ix + 1) % nx;
			s->E[Y][iy * nx + ix] = (pm[ix] - pp[ix]) / dy2;
			s->E[X][iy * nx + ix] = (pc[x0] - pc[x1]) / dx2;
		}
	}
	return 0;
}

/* --- stage_plasma_E (src/particle.c:233-248, interpolate_f2p_E) --- */

int
orc_stage_plasma_E(orc_sim_t *s)
{
	int64_t li;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
	for(li = 0; li < s->nlist; li++)
	{
		plist_t *l = &s->list[li];
		int64_t j;
		for(j = 0; j < l->n; j++)
		{
			double w[2][2];
			int64_t i0[2];
			weights(s, l->x[j], l->y[j], w, i0);
			l->Ex[j] = gather1(s, s->E[X], w, i0);
			l->Ey[j] = gather1(s, s->E[Y], w, i0);
		}
	}
	return 0;
}

/* --- comm_plasma (src/comm_plasma.c:826-885, 1087-1142) --- */

/* chunk bounds: plasma_chunk_init (src/plasma.c:264-269) */
static void
chunk_bounds(const orc_sim_t *s, int64_t c, double *x0, double *x1)
{
	double cL = s->fL[X] / s->p.nchunks;
	*x0 = 0.0 + (c * cL);
	*x1 = *x0 + cL;
}

/* collect_pset for dim X: lost if x < x0 or x > x1 (src/comm_plasma.c:124-125);
 * returns the number queued. */
static int64_t
collect_x(orc_sim_t *s, int64_t c, int sp)
{
	plist_t *l = L_(s, c, sp);
	plist_t *q0 = &s->q0[c * s->p.nspecies + sp];
	plist_t *q1 = &s->q1[c * s->p.nspecies + sp];
	double x0, x1;
	int64_t j, k = 0, moved = 0;

	chunk_bounds(s, c, &x0, &x1);
	for(j = 0; j < l->n; j++)
	{
		double x = l->x[j];
		if(x < x0)
		{
			plist_push(q0, l, j);
			moved++;
		}
		else if(x > x1)
		{
			plist_push(q1, l, j);
			moved++;
		}
		else
		{
			if(k != j)
			{
				l->id[k] = l->id[j]; l->x[k] = l->x[j]; l->y[k] = l->y[j];
				l->ux[k] = l->ux[j]; l->uy[k] = l->uy[j];
				l->Ex[k] = l->Ex[j]; l->Ey[k] = l->Ey[j];
			}
			k++;
		}
	}
	l->n = k;
	return moved;
}

/* exchange_plasma_x / exchange_pchunk_x (src/comm_plasma.c:403-441,
 * 695-723): chunk c appends q1 of c-1 (all species), then q0 of c+1. */
static void
exchange_x(orc_sim_t *s)
{
	int64_t nc = s->p.nchunks, c, j;
	int sp, ns = s->p.nspecies;

	for(c = 0; c < nc; c++)
	{
		int64_t cp = (c - 1 + nc) % nc, cn = (c + 1) % nc;
		for(sp = 0; sp < ns; sp++)
		{
			plist_t *q = &s->q1[cp * ns + sp];
			for(j = 0; j < q->n; j++)
				plist_push(L_(s, c, sp), q, j);
			q->n = 0;
		}
		for(sp = 0; sp < ns; sp++)
		{
			plist_t *q = &s->q0[cn * ns + sp];
			for(j = 0; j < q->n; j++)
				plist_push(L_(s, c, sp), q, j);
			q->n = 0;
		}
	}
}

/* periodic_boundary_ppack (src/comm_plasma.c:726-747) + the x == L -> 0
 * deviation. */
static inline double
wrap(double r, double L)
{
	if(r >= L)
		r -= L;
	else if(r < 0.0)
	{
		r += L;
		if(r >= L)
			r = 0.0;
	}
	return r;
}

static void
periodic_boundary(orc_sim_t *s, int d)
{
	int64_t li;
	double L = s->p.L[d];
	for(li = 0; li < s->nlist; li++)
	{
		plist_t *l = &s->list[li];
		double *r = d == X ? l->x : l->y;
		int64_t j;
		for(j = 0; j < l->n; j++)
			r[j] = wrap(r[j], L);
	}
}

int
orc_comm_plasma(orc_sim_t *s, int global_exchange)
{
	int64_t c, count;
	int sp;

	if(global_exchange)
	{
		do
		{
			count = 0;
			for(c = 0; c < s->p.nchunks; c++)
				for(sp = 0; sp < s->p.nspecies; sp++)
					count += collect_x(s, c, sp);
			exchange_x(s);
		}
		while(count);
	}
	else
	{
		for(c = 0; c < s->p.nchunks; c++)
			for(sp = 0; sp < s->p.nspecies; sp++)
				collect_x(s, c, sp);
		exchange_x(s);
		periodic_boundary(s, X);
	}
	/* comm_plasma_y with one rank: wrap only (src/comm_plasma.c:1115) */
	periodic_boundary(s, Y);
	return 0;
}

/* --- stage_plasma_r (src/mover.c:141-227, 332-362) --- */

int
orc_stage_plasma_r(orc_sim_t *s)
{
	double dt;
	int set_r, bad = 0;
	int64_t li;

	if(s->iter < 0)
		return ORC_ESTATE;
	if(s->iter == 0)
	{
		dt = -s->p.dt / 2;
		set_r = 0;
	}
	else
	{
		dt = s->p.dt;
		set_r = 1;
	}

#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(|:bad)
#endif
	for(li = 0; li < s->nlist; li++)
	{
		plist_t *l = &s->list[li];
		int sp = (int) (li % s->p.nspecies);
		double h = 0.5 * dt * s->p.q[sp] / s->p.m[sp];
		int64_t j;
		for(j = 0; j < l->n; j++)
		{
			double u[3] = {l->ux[j], l->uy[j], 0.0};
			double E[3] = {l->Ex[j], l->Ey[j], 0.0};
			double un[3];
			int d;
			boris(u, E, s->p.B, h, un);
			/* check_velocity (src/mover.c:98-137) */
			for(d = 0; d < 3; d++)
				if(fabs(un[d]) > s->umax[d])
					bad = 1;
			if(set_r)
			{
				l->x[j] = fma(un[X], dt, l->x[j]);
				l->y[j] = fma(un[Y], dt, l->y[j]);
			}
			l->ux[j] = un[X];
			l->uy[j] = un[Y];
		}
	}
	if(bad)
		return ORC_EUMAX;
	return orc_comm_plasma(s, 0);
}

/* --- stage_field_rho (src/field.c:164-356) --- */

static void
deposit_chunk(orc_sim_t *s, int64_t c)
{
	int sp;
	int64_t nx = s->nx, rs = s->rs;
	for(sp = 0; sp < s->p.nspecies; sp++)
	{
		plist_t *l = L_(s, c, sp);
		double vq = -s->p.q[sp] / s->p.e0;   /* src/interpolate.c:307 */
		int64_t j;
		for(j = 0; j < l->n; j++)
		{
			double w[2][2];
			int64_t i0[2], i1x, i1y;
			weights(s, l->x[j], l->y[j], w, i0);
			i1x = i0[X] + 1;
			if(i1x >= nx)
				i1x -= nx;
			i1y = i0[Y] + 1;
			/* vmat_add_xy order (src/interpolate.c:259-262) */
			s->rho[i0[Y] * rs + i0[X]] += w[0][0] * vq;
			s->rho[i1y * rs + i0[X]] += w[0][1] * vq;
			s->rho[i0[Y] * rs + i1x] += w[1][0] * vq;
			s->rho[i1y * rs + i1x] += w[1][1] * vq;
		}
	}
}

int
orc_stage_field_rho(orc_sim_t *s)
{
	int64_t nx = s->nx, ny = s->ny, rs = s->rs, c, ix, iy;
	int64_t nc = s->p.nchunks;

	/* rho_reset (src/field.c:164-210): rows 0..ny of every chunk column */
	for(iy = 0; iy <= ny; iy++)
		for(ix = 0; ix < nx; ix++)
			s->rho[iy * rs + ix] = 0.0;

	/* rho_update (src/field.c:216-230) per chunk.  Chunk c writes columns
	 * within [c*cs - 1, (c+1)*cs + 1] (one column of rounding slack on each
	 * side), so with an even count and cs >= 4 the even chunks are
	 * disjoint, then the odd ones; the order is fixed regardless of the
	 * thread count. */
	if(nc >= 2 && nc % 2 == 0 && s->chunksize >= 4)
	{
		int parity;
		for(parity = 0; parity < 2; parity++)
		{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
			for(c = parity; c < nc; c += 2)
				deposit_chunk(s, c);
		}
	}
	else
	{
		for(c = 0; c < nc; c++)
			deposit_chunk(s, c);
	}

	/* comm_send/recv_ghost_rho with one rank (src/comm_field.c:52-136) */
	for(ix = 0; ix < nx; ix++)
		s->rho[0 * rs + ix] += s->rho[ny * rs + ix];
	return 0;
}

/* --- driver (src/sim.c:209-236, 482-581) --- */

int
orc_pre_step(orc_sim_t *s)
{
	int ret;
	if(s->iter != -1)
		return ORC_ESTATE;
	if((ret = orc_comm_plasma(s, 1)))
		return ret;
	if((ret = orc_stage_field_rho(s)))
		return ret;
	if((ret = orc_stage_field_E(s)))
		return ret;
	s->iter = 0;
	return 0;
}

int
orc_step(orc_sim_t *s)
{
	int ret;
	if(s->iter < 0)
		return ORC_ESTATE;
	if((ret = orc_stage_field_E(s)))
		return ret;
	if((ret = orc_stage_plasma_E(s)))
		return ret;
	if((ret = orc_stage_plasma_r(s)))
		return ret;
	if((ret = orc_stage_field_rho(s)))
		return ret;
	s->iter++;
	return 0;
}

int
orc_run(orc_sim_t *s, int64_t nsteps)
{
	int64_t i;
	int ret;
	for(i = 0; i < nsteps; i++)
		if((ret = orc_step(s)))
			return ret;
	return 0;
}

/* --- accessors --- */

int64_t
orc_count(const orc_sim_t *s, int sp)
{
	int64_t c, n = 0;
	for(c = 0; c < s->p.nchunks; c++)
		n += L_(s, c, sp)->n;
	return n;
}

typedef struct { int64_t id; int64_t c; int64_t j; } ref_t;

static int
cmp_ref(const void *a, const void *b)
{
	const ref_t *ra = a, *rb = b;
	return (ra->id > rb->id) - (ra->id < rb->id);
}

static ref_t *
sorted_refs(const orc_sim_t *s, int sp, int64_t *n_out)
{
	int64_t n = orc_count(s, sp), c, j, k = 0;
	ref_t *r = malloc(sizeof(ref_t) * (size_t) (n ? n : 1));
	if(!r)
		return NULL;
	for(c = 0; c < s->p.nchunks; c++)
	{
		const plist_t *l = L_(s, c, sp);
		for(j = 0; j < l->n; j++)
		{
			r[k].id = l->id[j]; r[k].c = c; r[k].j = j; k++;
		}
	}
	qsort(r, (size_t) n, sizeof(ref_t), cmp_ref);
	*n_out = n;
	return r;
}

int
orc_get_particles(const orc_sim_t *s, int sp, int64_t *id, double *x,
		double *y, double *ux, double *uy, double *Ex, double *Ey)
{
	int64_t n, k;
	ref_t *r;
	if(sp < 0 || sp >= s->p.nspecies)
		return ORC_EINVAL;
	r = sorted_refs(s, sp, &n);
	if(!r)
		return ORC_ENOMEM;
	for(k = 0; k < n; k++)
	{
		const plist_t *l = L_(s, r[k].c, sp);
		int64_t j = r[k].j;
		if(id) id[k] = l->id[j];
		if(x) x[k] = l->x[j];
		if(y) y[k] = l->y[j];
		if(ux) ux[k] = l->ux[j];
		if(uy) uy[k] = l->uy[j];
		if(Ex) Ex[k] = l->Ex[j];
		if(Ey) Ey[k] = l->Ey[j];
	}
	free(r);
	return 0;
}

static int
cmp_i64(const void *a, const void *b)
{
	int64_t x = *(const int64_t *) a, y = *(const int64_t *) b;
	return (x > y) - (x < y);
}

int
orc_set_particles(orc_sim_t *s, int sp, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy)
{
	int64_t m, k, a;
	ref_t *r;
	if(sp < 0 || sp >= s->p.nspecies)
		return ORC_EINVAL;
	r = sorted_refs(s, sp, &m);
	if(!r)
		return ORC_ENOMEM;
	for(k = 0; k < n; k++)
	{
		/* binary search by id */
		int64_t lo = 0, hi = m - 1, found = -1;
		while(lo <= hi)
		{
			int64_t mid = (lo + hi) / 2;
			if(r[mid].id == id[k]) { found = mid; break; }
			if(r[mid].id < id[k]) lo = mid + 1; else hi = mid - 1;
		}
		if(found < 0)
		{
			free(r);
			return ORC_EINVAL;
		}
		a = found;
		{
			plist_t *l = L_(s, r[a].c, sp);
			int64_t j = r[a].j;
			l->x[j] = x[k]; l->y[j] = y[k];
			l->ux[j] = ux[k]; l->uy[j] = uy[k];
		}
	}
	free(r);
	(void) cmp_i64;
	return 0;
}

int
orc_chunk_counts(const orc_sim_t *s, int64_t *out_n)
{
	int64_t i;
	for(i = 0; i < s->nlist; i++)
		out_n[i] = s->list[i].n;
	return 0;
}

int
orc_chunk_ids(const orc_sim_t *s, int64_t c, int sp, int64_t *ids)
{
	const plist_t *l;
	if(c < 0 || c >= s->p.nchunks || sp < 0 || sp >= s->p.nspecies)
		return ORC_EINVAL;
	l = L_(s, c, sp);
	memcpy(ids, l->id, sizeof(int64_t) * (size_t) l->n);
	return 0;
}

int
orc_get_field(const orc_sim_t *s, int which, double *out)
{
	int64_t nx = s->nx, ny = s->ny, ix, iy;
	for(iy = 0; iy < ny; iy++)
		for(ix = 0; ix < nx; ix++)
		{
			double v;
			switch(which)
			{
				case ORC_RHO: v = s->rho[iy * s->rs + ix]; break;
				case ORC_PHI: v = phi_row(s, iy)[ix]; break;
				case ORC_EX: v = s->E[X][iy * nx + ix]; break;
				case ORC_EY: v = s->E[Y][iy * nx + ix]; break;
				default: return ORC_EINVAL;
			}
			out[iy * nx + ix] = v;
		}
	return 0;
}

int
orc_set_field(orc_sim_t *s, int which, const double *in)
{
	int64_t nx = s->nx, ny = s->ny, ix, iy;
	if(which < ORC_RHO || which > ORC_EY)
		return ORC_EINVAL;
	for(iy = 0; iy < ny; iy++)
		for(ix = 0; ix < nx; ix++)
		{
			double v = in[iy * nx + ix];
			switch(which)
			{
				case ORC_RHO: s->rho[iy * s->rs + ix] = v; break;
				case ORC_PHI: phi_row(s, iy)[ix] = v; break;
				case ORC_EX: s->E[X][iy * nx + ix] = v; break;
				case ORC_EY: s->E[Y][iy * nx + ix] = v; break;
			}
		}
	/* keep the E ghost row ny equal to row 0, as field_E_compute does */
	if(which == ORC_EX || which == ORC_EY)
	{
		double *e = s->E[which == ORC_EX ? X : Y];
		memcpy(e + ny * nx, e, sizeof(double) * (size_t) nx);
	}
	return 0;
}

double
orc_field_energy(const orc_sim_t *s)
{
	/* -(e0/2) sum rho*phi.  The products have both signs and cancel, so the
	 * sum is accumulated as if in twice the working precision (Ogita, Rump
	 * and Oishi's Dot2: exact products by fma, error-free sums), which
	 * makes the oracle's value accurate to the last bits whatever the grid
	 * size. */
	int64_t nx = s->nx, ny = s->ny, ix, iy;
	double sum = 0.0, err = 0.0;
	for(iy = 0; iy < ny; iy++)
	{
		const double *r = &s->rho[iy * s->rs], *f = phi_row(s, iy);
		for(ix = 0; ix < nx; ix++)
		{
			double p = r[ix] * f[ix];
			double pe = fma(r[ix], f[ix], -p);
			double t = sum + p;
			double z = t - sum;
			double te = (sum - (t - z)) + (p - z);
			sum = t;
			err += te + pe;
		}
	}
	return -0.5 * s->p.e0 * (sum + err);
}

double
orc_kinetic_energy(const orc_sim_t *s)
{
	int64_t li;
	double ke = 0.0;
	for(li = 0; li < s->nlist; li++)
	{
		const plist_t *l = &s->list[li];
		int sp = (int) (li % s->p.nspecies);
		double acc = 0.0;
		int64_t j;
		for(j = 0; j < l->n; j++)
			acc += l->ux[j] * l->ux[j] + l->uy[j] * l->uy[j];
		ke += 0.5 * s->p.m[sp] * acc;
	}
	return ke;
}
