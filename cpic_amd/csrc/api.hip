// C ABI of the gfx950 PIC step (declared in include/cpic_hip.h).
//
// Stage protocol (src/sim.c:209-236, 482-581): pre_step deposits rho and
// solves E once (iter -1); each step then runs field_E, plasma_E, plasma_r,
// field_rho.  rho is double-buffered: plasma_r deposits the moved particles
// into rho[1-cur] and field_rho makes it current, so every stage observes
// the same rho as the reference's stage of the same name.
#include "dev.h"

#include <cmath>
#include <stdlib.h>
#include <string.h>

extern "C" {

const char *
cpic_dev_strerror(int err)
{
	switch(err)
	{
		case CPIC_OK: return "ok";
		case CPIC_EINVAL: return "invalid argument or unsupported configuration";
		case CPIC_ENOMEM: return "out of memory";
		case CPIC_EUMAX: return "max velocity exceeded";
		case CPIC_ESTATE: return "call out of protocol order";
		case CPIC_EHIP: return "HIP runtime error";
		case CPIC_ENODEV: return "no HIP device";
		case CPIC_ENAN: return "non-finite particle state";
	}
	return "unknown error";
}

} // extern "C"

// ------------------------------------------------------------- timers ---

int
cpic_kt_begin(cpic_dev *d, const char *name)
{
	int k;
	if(!(d->desc.flags & CPIC_F_PROFILE))
		return -1;
	for(k = 0; k < d->nkt; k++)
		if(strcmp(d->kt[k].name, name) == 0)
			break;
	if(k == d->nkt)
	{
		if(d->nkt == CPIC_MAX_KERNELS)
			return -1;
		d->kt[k].name = name;
		d->kt[k].launches = 0;
		d->kt[k].total_ms = 0;
		d->nkt++;
	}
	if(d->nev + 2 > d->evcap)
	{
		int nc = d->evcap ? 2 * d->evcap : 256;
		hipEvent_t *ev = (hipEvent_t *) realloc(d->ev, sizeof(hipEvent_t) * nc);
		int *ek = (int *) realloc(d->ev_kernel, sizeof(int) * nc);
		if(!ev || !ek)
			return -1;
		for(int i = d->evcap; i < nc; i++)
			hipEventCreate(&ev[i]);
		d->ev = ev;
		d->ev_kernel = ek;
		d->evcap = nc;
	}
	int slot = d->nev;
	d->ev_kernel[slot / 2] = k;
	hipEventRecord(d->ev[slot], d->stream);
	d->nev += 2;
	return slot;
}

void
cpic_kt_end(cpic_dev *d, int slot)
{
	if(slot < 0)
		return;
	hipEventRecord(d->ev[slot + 1], d->stream);
}

static void
kt_collect(cpic_dev *d)
{
	if(!d->nev)
		return;
	hipStreamSynchronize(d->stream);
	for(int i = 0; i + 1 < d->nev; i += 2)
	{
		float ms = 0;
		if(hipEventElapsedTime(&ms, d->ev[i], d->ev[i + 1]) == hipSuccess)
		{
			KernelTimer *t = &d->kt[d->ev_kernel[i / 2]];
			t->launches++;
			t->total_ms += ms;
		}
	}
	d->nev = 0;
}

// -------------------------------------------------------------- setup ---

static int
alloc_zero(void **p, size_t bytes)
{
	if(hipMalloc(p, bytes ? bytes : 8) != hipSuccess)
		return CPIC_ENOMEM;
	if(hipMemset(*p, 0, bytes ? bytes : 8) != hipSuccess)
		return CPIC_EHIP;
	return 0;
}

static void
dev_free(cpic_dev *d)
{
	if(!d)
		return;
	if(d->stream)
		hipStreamSynchronize(d->stream);
	hipFree(d->rho[0]); hipFree(d->rho[1]); hipFree(d->ghat); hipFree(d->phi);
	hipFree(d->E); hipFree(d->costab); hipFree(d->tw);
	cpic_parts_free(d, 0);
	cpic_parts_free(d, 1);
	hipFree(d->bk.start); hipFree(d->bk.cap); hipFree(d->bk.n1); hipFree(d->bk.res);
	hipFree(d->bk.n2);
	hipFree(d->lv.x); hipFree(d->lv.y); hipFree(d->lv.ux); hipFree(d->lv.uy);
	hipFree(d->lv.id); hipFree(d->lv.Ex); hipFree(d->lv.Ey); hipFree(d->lv.dst);
	hipFree(d->lv.src); hipFree(d->lv.count);
	hipFree(d->diag); hipFree(d->derr); hipFree(d->ctr);
	for(int s = 0; s < CPIC_MAX_SPECIES; s++)
	{
		free(d->up_buf[s]);
		free(d->up_id[s]);
	}
	for(int i = 0; i < d->evcap; i++)
		hipEventDestroy(d->ev[i]);
	free(d->ev);
	free(d->ev_kernel);
	if(d->stream)
		hipStreamDestroy(d->stream);
	free(d);
}

static int
leavers_alloc(cpic_dev *d, int64_t cap)
{
	Leavers *l = &d->lv;
	size_t e = d->prec == 64 ? sizeof(double) : sizeof(float);
	hipFree(l->x); hipFree(l->y); hipFree(l->ux); hipFree(l->uy);
	hipFree(l->id); hipFree(l->Ex); hipFree(l->Ey); hipFree(l->dst); hipFree(l->src);
	l->x = l->y = l->ux = l->uy = NULL;
	l->id = NULL;
	l->Ex = l->Ey = NULL;
	l->dst = l->src = NULL;
	l->cap = 0;
	if(hipMalloc(&l->x, e * cap) != hipSuccess || hipMalloc(&l->y, e * cap) != hipSuccess ||
			hipMalloc(&l->ux, e * cap) != hipSuccess ||
			hipMalloc(&l->uy, e * cap) != hipSuccess ||
			hipMalloc(&l->dst, sizeof(int32_t) * cap) != hipSuccess ||
			hipMalloc(&l->src, sizeof(int32_t) * cap) != hipSuccess)
		return CPIC_ENOMEM;
	if((d->desc.flags & CPIC_F_IDS) && hipMalloc(&l->id, sizeof(int64_t) * cap) != hipSuccess)
		return CPIC_ENOMEM;
	if((d->desc.flags & CPIC_F_PARTICLE_E) &&
			(hipMalloc(&l->Ex, sizeof(double) * cap) != hipSuccess ||
			 hipMalloc(&l->Ey, sizeof(double) * cap) != hipSuccess))
		return CPIC_ENOMEM;
	l->cap = cap;
	return 0;
}

extern "C" int
cpic_dev_create(const cpic_dev_desc_t *desc, cpic_dev_t **out)
{
	cpic_dev *d;
	int ndev = 0, ret, s;
	int64_t nx, ny, tx, ty;

	if(!desc || !out)
		return CPIC_EINVAL;
	*out = NULL;
	nx = desc->nx;
	ny = desc->ny;
	if(!is_pow2(nx) || !is_pow2(ny) || nx < 2 || ny < 2 || nx > 8192 || ny > 8192)
		return CPIC_EINVAL;
	if(!(desc->L[0] > 0) || !(desc->L[1] > 0) || !(desc->dt > 0) || desc->e0 == 0.0)
		return CPIC_EINVAL;
	if(desc->nspecies < 1 || desc->nspecies > CPIC_MAX_SPECIES)
		return CPIC_EINVAL;
	if(desc->precision != 32 && desc->precision != 64)
		return CPIC_EINVAL;
	// In-plane B makes u_z non-zero, which the reference rejects through
	// umax_Z = 0 (src/sim.c:200, src/mover.c:113).
	if(desc->B[0] != 0.0 || desc->B[1] != 0.0)
		return CPIC_EINVAL;
	if(desc->nchunks < 1 || nx % desc->nchunks)   // src/sim.c:124-130
		return CPIC_EINVAL;
	for(s = 0; s < desc->nspecies; s++)
		if(!(desc->m[s] != 0.0) || !std::isfinite(desc->q[s]))
			return CPIC_EINVAL;
	tx = desc->tile[0] > 0 ? desc->tile[0] : 16;
	ty = desc->tile[1] > 0 ? desc->tile[1] : 16;
	if(tx > nx) tx = nx;
	if(ty > ny) ty = ny;
	if(!is_pow2(tx) || !is_pow2(ty) || tx > 64 || ty > 64)
		return CPIC_EINVAL;

	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
		return CPIC_ENODEV;
	if(desc->device < 0 || desc->device >= ndev)
		return CPIC_EINVAL;
	if(hipSetDevice(desc->device) != hipSuccess)
		return CPIC_EHIP;

	d = (cpic_dev *) calloc(1, sizeof(*d));
	if(!d)
		return CPIC_ENOMEM;
	d->desc = *desc;
	d->prec = desc->precision;
	d->ns = desc->nspecies;
	d->dt = desc->dt;
	d->device = desc->device;

	Geo &g = d->g;
	g.nx = (int) nx;
	g.ny = (int) ny;
	g.lgnx = ilog2(nx);
	g.lgny = ilog2(ny);
	g.tx = (int) tx;
	g.ty = (int) ty;
	g.lgtx = ilog2(tx);
	g.lgty = ilog2(ty);
	g.ntx = (int) (nx / tx);
	g.nty = (int) (ny / ty);
	g.lgntx = ilog2(g.ntx);
	g.ntiles = g.ntx * g.nty;
	g.M = CPIC_MARGIN;
	g.W = g.tx + 2 * g.M + 1;
	g.H = g.ty + 2 * g.M + 1;
	g.ns = d->ns;
	g.dx = desc->L[0] / nx;      // src/sim.c:174
	g.dy = desc->L[1] / ny;
	g.idx = 1.0 / g.dx;          // src/interpolate.c:365-366
	g.idy = 1.0 / g.dy;
	g.L0 = desc->L[0];
	g.L1 = desc->L[1];
	d->nb = g.ntiles * d->ns;
	// umax = chunksize * dx / dt (src/sim.c:198-200)
	d->umax[0] = (double) (nx / desc->nchunks) * g.dx / desc->dt;
	d->umax[1] = (double) ny * g.dy / desc->dt;

	if(hipStreamCreate(&d->stream) != hipSuccess)
	{
		free(d);
		return CPIC_EHIP;
	}
	size_t ng = (size_t) nx * ny;
	size_t esz = d->prec == 64 ? sizeof(double2) : sizeof(float2);
	if((ret = alloc_zero((void **) &d->rho[0], sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->rho[1], sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->ghat, sizeof(double2) * (size_t) (nx / 2 + 1) * ny)) ||
			(ret = alloc_zero((void **) &d->phi, sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->E, esz * ng)) ||
			(ret = alloc_zero((void **) &d->diag, sizeof(DiagSlot) * CPIC_DIAG_RING)) ||
			(ret = alloc_zero((void **) &d->derr, sizeof(int))) ||
			(ret = alloc_zero((void **) &d->ctr, sizeof(Counters))) ||
			(ret = alloc_zero((void **) &d->bk.start, sizeof(int64_t) * d->nb)) ||
			(ret = alloc_zero((void **) &d->bk.cap, sizeof(int32_t) * d->nb)) ||
			(ret = alloc_zero((void **) &d->bk.n1, sizeof(int32_t) * d->nb)) ||
			(ret = alloc_zero((void **) &d->bk.res, sizeof(int32_t) * d->nb)) ||
			(ret = alloc_zero((void **) &d->bk.n2, sizeof(int32_t) * d->nb)) ||
			(ret = alloc_zero((void **) &d->lv.count, sizeof(int32_t) * 2)) ||
			(ret = cpic_fft_init(d)) ||
			(ret = cpic_parts_init(d)))
	{
		dev_free(d);
		return ret;
	}
	d->rho_cur = 0;
	d->rho_next_ready = 1;   // both buffers start zeroed
	d->last_iter = -1;
	*out = d;
	return 0;
}

extern "C" int
cpic_dev_destroy(cpic_dev_t *d)
{
	if(!d)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	dev_free(d);
	return 0;
}

// ---------------------------------------------------------- particles ---

extern "C" int
cpic_dev_upload_particles(cpic_dev_t *d, int32_t s, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy)
{
	if(!d || s < 0 || s >= d->ns || n < 0 || (n > 0 && (!x || !y || !ux || !uy)))
		return CPIC_EINVAL;
	if((d->desc.flags & CPIC_F_IDS) && n > 0 && !id)
		return CPIC_EINVAL;
	for(int64_t i = 0; i < n; i++)
		if(!(x[i] >= 0.0 && x[i] <= d->g.L0 && y[i] >= 0.0 && y[i] <= d->g.L1) ||
				!std::isfinite(ux[i]) || !std::isfinite(uy[i]))
			return CPIC_EINVAL;
	free(d->up_buf[s]);
	free(d->up_id[s]);
	d->up_buf[s] = (double *) malloc(sizeof(double) * 4 * (n ? n : 1));
	d->up_id[s] = NULL;
	if(!d->up_buf[s])
		return CPIC_ENOMEM;
	memcpy(d->up_buf[s], x, sizeof(double) * n);
	memcpy(d->up_buf[s] + n, y, sizeof(double) * n);
	memcpy(d->up_buf[s] + 2 * n, ux, sizeof(double) * n);
	memcpy(d->up_buf[s] + 3 * n, uy, sizeof(double) * n);
	if(d->desc.flags & CPIC_F_IDS)
	{
		d->up_id[s] = (int64_t *) malloc(sizeof(int64_t) * (n ? n : 1));
		if(!d->up_id[s])
			return CPIC_ENOMEM;
		memcpy(d->up_id[s], id, sizeof(int64_t) * n);
	}
	d->up_n[s] = n;
	d->pending = 1;
	return 0;
}

static int
commit(cpic_dev *d)
{
	int ret;
	int64_t tot = 0;
	if(!d->pending)
		return 0;
	hipSetDevice(d->device);
	for(int s = 0; s < d->ns; s++)
		tot += d->up_buf[s] ? d->up_n[s] : d->count[s];
	int64_t want = tot / 32;
	if(want < 16384)
		want = 16384;
	if(d->lv.cap < want)
		if((ret = leavers_alloc(d, want)))
			return ret;
	if((ret = cpic_parts_rebuild(d)))
		return ret;
	HIPCHK(hipMemsetAsync(d->ctr, 0, sizeof(Counters), d->stream));
	return 0;
}

static int
derr_check(cpic_dev *d)
{
	int e = 0;
	HIPCHK(hipMemcpyAsync(&e, d->derr, sizeof(int), hipMemcpyDeviceToHost, d->stream));
	HIPCHK(hipStreamSynchronize(d->stream));
	if(e & CPIC_DERR_UMAX)
		return CPIC_EUMAX;
	if(e & CPIC_DERR_NAN)
		return CPIC_ENAN;
	return 0;
}

extern "C" int
cpic_dev_sync(cpic_dev_t *d)
{
	if(!d)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	HIPCHK(hipStreamSynchronize(d->stream));
	return derr_check(d);
}

extern "C" int
cpic_dev_count(cpic_dev_t *d, int32_t s, int64_t *n)
{
	if(!d || s < 0 || s >= d->ns || !n)
		return CPIC_EINVAL;
	*n = d->up_buf[s] ? d->up_n[s] : d->count[s];
	return 0;
}

extern "C" int
cpic_dev_download_particles(cpic_dev_t *d, int32_t s, cpic_particles_t *out)
{
	int ret;
	if(!d || !out || s < 0 || s >= d->ns)
		return CPIC_EINVAL;
	if(out->id && !(d->desc.flags & CPIC_F_IDS))
		return CPIC_EINVAL;
	if((out->Ex || out->Ey) && !(d->desc.flags & CPIC_F_PARTICLE_E))
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	hipSetDevice(d->device);
	if((ret = cpic_parts_download(d, s, out)))
		return ret;
	return derr_check(d);
}

extern "C" int
cpic_dev_cell_counts(cpic_dev_t *d, int32_t s, int64_t *counts)
{
	int ret;
	if(!d || !counts || s < 0 || s >= d->ns)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	return cpic_parts_cell_counts(d, s, counts);
}

extern "C" int
cpic_dev_rebuild(cpic_dev_t *d)
{
	int ret;
	if(!d)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	if(d->pending)
		return commit(d);
	if((ret = cpic_parts_rebuild(d)))
		return ret;
	HIPCHK(hipMemsetAsync(d->ctr, 0, sizeof(Counters), d->stream));
	return 0;
}

// ------------------------------------------------------------- stages ---

extern "C" int
cpic_dev_field_E(cpic_dev_t *d, int64_t iter)
{
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	if((ret = cpic_solve(d, iter, 1)))
		return ret;
	return cpic_efield(d, iter);
}

extern "C" int
cpic_dev_solve_xy(cpic_dev_t *d)
{
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	return cpic_solve(d, -1, 0);
}

extern "C" int
cpic_dev_plasma_E(cpic_dev_t *d, int64_t iter)
{
	int ret;
	(void) iter;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	if(d->desc.flags & CPIC_F_PARTICLE_E)
		return cpic_parts_gather(d);
	return 0;
}

static int
zero_rho(cpic_dev *d, double *r)
{
	return cpic_zero(d, r, (int64_t) d->g.nx * d->g.ny);
}

extern "C" int
cpic_dev_plasma_r(cpic_dev_t *d, int64_t iter)
{
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if(iter < 0)
		return CPIC_ESTATE;
	if((ret = commit(d)))
		return ret;
	if(d->rho_pending)
		return CPIC_ESTATE;    // plasma_r twice without field_rho
	if(!d->rho_next_ready)
		if((ret = zero_rho(d, d->rho[1 - d->rho_cur])))
			return ret;
	if((ret = cpic_parts_push(d, iter)))
		return ret;
	d->rho_pending = 1;
	d->rho_next_ready = 0;
	d->last_iter = iter;
	return 0;
}

extern "C" int
cpic_dev_field_rho(cpic_dev_t *d, int64_t iter)
{
	int ret;
	(void) iter;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	if(d->rho_pending)
	{
		d->rho_cur = 1 - d->rho_cur;
		d->rho_pending = 0;
		d->rho_next_ready = 0;
		return 0;
	}
	// rho_reset + rho_update + ghost fold from the stored positions
	if((ret = zero_rho(d, d->rho[d->rho_cur])))
		return ret;
	return cpic_parts_deposit(d, d->rho[d->rho_cur]);
}

extern "C" int
cpic_dev_pre_step(cpic_dev_t *d)
{
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if((ret = cpic_dev_field_rho(d, -1)))
		return ret;
	return cpic_dev_field_E(d, -1);
}

extern "C" int
cpic_dev_step(cpic_dev_t *d, int64_t iter)
{
	int ret;
	if((ret = cpic_dev_field_E(d, iter)))
		return ret;
	if((ret = cpic_dev_plasma_E(d, iter)))
		return ret;
	if((ret = cpic_dev_plasma_r(d, iter)))
		return ret;
	return cpic_dev_field_rho(d, iter);
}

extern "C" int
cpic_dev_run(cpic_dev_t *d, int64_t iter0, int64_t nsteps)
{
	int ret;
	for(int64_t i = 0; i < nsteps; i++)
		if((ret = cpic_dev_step(d, iter0 + i)))
			return ret;
	return 0;
}

extern "C" int64_t
cpic_solver_rho_size(int64_t nx, int64_t ny, int64_t *cnx, int64_t *cny)
{
	// one rank: local_size = ny * (nx/2+1) complex (src/solver.c:392-432)
	int64_t padded = 2 * (nx / 2 + 1);
	if(cnx)
		*cnx = padded;
	if(cny)
		*cny = ny;
	return padded * ny;
}

// ------------------------------------------------------------- fields ---

extern "C" int
cpic_dev_download_field(cpic_dev_t *d, int32_t which, double *out)
{
	int64_t n;
	int ret;
	if(!d || !out)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	n = (int64_t) d->g.nx * d->g.ny;
	hipSetDevice(d->device);
	if(which == CPIC_RHO)
		HIPCHK(hipMemcpyAsync(out, d->rho[d->rho_cur], sizeof(double) * n,
				hipMemcpyDeviceToHost, d->stream));
	else if(which == CPIC_PHI)
		HIPCHK(hipMemcpyAsync(out, d->phi, sizeof(double) * n, hipMemcpyDeviceToHost,
				d->stream));
	else if(which == CPIC_EX || which == CPIC_EY)
	{
		int c = which == CPIC_EX ? 0 : 1;
		if(d->prec == 64)
		{
			double2 *h = (double2 *) malloc(sizeof(double2) * n);
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpyAsync(h, d->E, sizeof(double2) * n, hipMemcpyDeviceToHost, d->stream);
			hipStreamSynchronize(d->stream);
			for(int64_t i = 0; i < n; i++)
				out[i] = c ? h[i].y : h[i].x;
			free(h);
		}
		else
		{
			float2 *h = (float2 *) malloc(sizeof(float2) * n);
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpyAsync(h, d->E, sizeof(float2) * n, hipMemcpyDeviceToHost, d->stream);
			hipStreamSynchronize(d->stream);
			for(int64_t i = 0; i < n; i++)
				out[i] = c ? h[i].y : h[i].x;
			free(h);
		}
	}
	else
		return CPIC_EINVAL;
	HIPCHK(hipStreamSynchronize(d->stream));
	return 0;
}

extern "C" int
cpic_dev_upload_field(cpic_dev_t *d, int32_t which, const double *in)
{
	int64_t n;
	if(!d || !in)
		return CPIC_EINVAL;
	n = (int64_t) d->g.nx * d->g.ny;
	hipSetDevice(d->device);
	HIPCHK(hipStreamSynchronize(d->stream));
	if(which == CPIC_RHO)
		HIPCHK(hipMemcpy(d->rho[d->rho_cur], in, sizeof(double) * n, hipMemcpyHostToDevice));
	else if(which == CPIC_PHI)
		HIPCHK(hipMemcpy(d->phi, in, sizeof(double) * n, hipMemcpyHostToDevice));
	else if(which == CPIC_EX || which == CPIC_EY)
	{
		int c = which == CPIC_EX ? 0 : 1;
		if(d->prec == 64)
		{
			double2 *h = (double2 *) malloc(sizeof(double2) * n);
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpy(h, d->E, sizeof(double2) * n, hipMemcpyDeviceToHost);
			for(int64_t i = 0; i < n; i++)
				if(c) h[i].y = in[i]; else h[i].x = in[i];
			hipMemcpy(d->E, h, sizeof(double2) * n, hipMemcpyHostToDevice);
			free(h);
		}
		else
		{
			float2 *h = (float2 *) malloc(sizeof(float2) * n);
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpy(h, d->E, sizeof(float2) * n, hipMemcpyDeviceToHost);
			for(int64_t i = 0; i < n; i++)
				if(c) h[i].y = (float) in[i]; else h[i].x = (float) in[i];
			hipMemcpy(d->E, h, sizeof(float2) * n, hipMemcpyHostToDevice);
			free(h);
		}
	}
	else
		return CPIC_EINVAL;
	return 0;
}

// -------------------------------------------------------- diagnostics ---

extern "C" int
cpic_dev_energy(cpic_dev_t *d, int64_t max, double *field, double *kinetic,
		double *px, double *py, int64_t *n)
{
	int64_t steps, k;
	DiagSlot *h;
	if(!d || max < 0)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	HIPCHK(hipStreamSynchronize(d->stream));
	steps = d->last_iter + 1;
	if(steps > CPIC_DIAG_RING)
		steps = CPIC_DIAG_RING;
	if(n)
		*n = steps;
	k = steps < max ? steps : max;
	if(k <= 0)
		return derr_check(d);
	h = (DiagSlot *) malloc(sizeof(DiagSlot) * k);
	if(!h)
		return CPIC_ENOMEM;
	HIPCHK(hipMemcpy(h, d->diag, sizeof(DiagSlot) * k, hipMemcpyDeviceToHost));
	for(int64_t i = 0; i < k; i++)
	{
		if(field) field[i] = h[i].field;
		if(kinetic) kinetic[i] = h[i].kinetic;
		if(px) px[i] = h[i].px;
		if(py) py[i] = h[i].py;
	}
	free(h);
	return derr_check(d);
}

extern "C" int
cpic_dev_migration(cpic_dev_t *d, int64_t *n, int64_t *id, int32_t *species,
		int32_t *src_tile, int32_t *dst_tile)
{
	int32_t cnt = 0;
	int64_t k;
	if(!d || !n)
		return CPIC_EINVAL;
	if(id && !(d->desc.flags & CPIC_F_IDS))
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	HIPCHK(hipStreamSynchronize(d->stream));
	HIPCHK(hipMemcpy(&cnt, d->lv.count + (1 - d->step_parity), sizeof(int32_t),
			hipMemcpyDeviceToHost));
	k = cnt < d->lv.cap ? cnt : d->lv.cap;
	if(*n < k)
	{
		*n = k;
		return CPIC_EINVAL;
	}
	*n = k;
	if(k == 0)
		return 0;
	int32_t *src = (int32_t *) malloc(sizeof(int32_t) * k);
	int32_t *dst = (int32_t *) malloc(sizeof(int32_t) * k);
	if(!src || !dst)
	{
		free(src); free(dst);
		return CPIC_ENOMEM;
	}
	hipMemcpy(src, d->lv.src, sizeof(int32_t) * k, hipMemcpyDeviceToHost);
	hipMemcpy(dst, d->lv.dst, sizeof(int32_t) * k, hipMemcpyDeviceToHost);
	if(id)
		hipMemcpy(id, d->lv.id, sizeof(int64_t) * k, hipMemcpyDeviceToHost);
	for(int64_t i = 0; i < k; i++)
	{
		if(species) species[i] = src[i] % d->ns;
		if(src_tile) src_tile[i] = src[i] / d->ns;
		if(dst_tile) dst_tile[i] = dst[i] / d->ns;
	}
	free(src);
	free(dst);
	return 0;
}

extern "C" int
cpic_dev_stats(cpic_dev_t *d, cpic_dev_stats_t *st)
{
	Counters c;
	if(!d || !st)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	HIPCHK(hipStreamSynchronize(d->stream));
	HIPCHK(hipMemcpy(&c, d->ctr, sizeof(c), hipMemcpyDeviceToHost));
	memset(st, 0, sizeof(*st));
	for(int s = 0; s < d->ns; s++)
		st->nparticles += d->up_buf[s] ? d->up_n[s] : d->count[s];
	st->capacity = d->pcap[d->pcur];
	st->nbuckets = d->nb;
	st->leavers = (int64_t) c.leavers;
	st->fallbacks = (int64_t) c.fallbacks;
	st->misplaced = (int64_t) c.misplaced;
	st->rebuilds = d->nrebuilds;
	st->tile[0] = d->g.tx;
	st->tile[1] = d->g.ty;
	st->precision = d->prec;
	return 0;
}

extern "C" int
cpic_dev_kernel_times(cpic_dev_t *d, int32_t *n, const char **names, int64_t *launches,
		double *total_ms)
{
	if(!d || !n)
		return CPIC_EINVAL;
	hipSetDevice(d->device);
	kt_collect(d);
	int k = d->nkt < *n ? d->nkt : *n;
	for(int i = 0; i < k; i++)
	{
		if(names) names[i] = d->kt[i].name;
		if(launches) launches[i] = d->kt[i].launches;
		if(total_ms) total_ms[i] = d->kt[i].total_ms;
	}
	*n = d->nkt;
	return 0;
}

extern "C" int
cpic_dev_reset_kernel_times(cpic_dev_t *d)
{
	if(!d)
		return CPIC_EINVAL;
	kt_collect(d);
	for(int i = 0; i < d->nkt; i++)
	{
		d->kt[i].launches = 0;
		d->kt[i].total_ms = 0;
	}
	return 0;
}
