// C ABI of the gfx950 PIC step (declared in include/cpic_hip.h).
//
// Stage protocol (src/sim.c:209-236, 482-581): pre_step deposits rho and
// solves E once (iter -1); each step then runs field_E, plasma_E, plasma_r,
// field_rho.  rho is double-buffered: plasma_r deposits the moved particles
// into rho[1-cur] and field_rho makes it current, so every stage observes
// the same rho as the reference's stage of the same name.
//
// A state object holds a batch of P independent plasmas; every stage runs
// all of them with one launch per kernel.  cpic_dev_create makes a batch of
// one, and the cpic_dev_* calls that address one plasma act on plasma 0.
//
// cpic_dev_run replays steps >= 1 two at a time from a captured HIP graph
// (the leaver parity and the rho buffers alternate each step, so two steps
// return every pointer to where it was).  The step counter lives on the
// device, so a replay needs no new kernel arguments.
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
		case CPIC_ESIZE: return "input larger than the fixed workspace of an allocation-free call";
	}
	return "unknown error";
}

} // extern "C"

// ------------------------------------------------------------- timers ---

int
cpic_kt_begin(cpic_dev *d, const char *name)
{
	int k;
	if(!(d->flags & CPIC_F_PROFILE))
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
		if(!ev)
			return -1;
		d->ev = ev;
		int *ek = (int *) realloc(d->ev_kernel, sizeof(int) * nc);
		if(!ek)
			return -1;
		d->ev_kernel = ek;
		for(int i = d->evcap; i < nc; i++)
			hipEventCreate(&ev[i]);
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

// After a re-sort that failed part way the particle storage is gone (or a
// buffer the push writes was freed): the handle refuses further work.
#define BROKEN_CHECK(d) do { \
		if((d) && (d)->broken) \
			return CPIC_ESTATE; \
	} while(0)

// Makes the handle's device current (not while a step is being captured:
// the stream is then recording and the device is already current).
static void
set_device(cpic_dev *d)
{
	if(!d->capturing)
		hipSetDevice(d->device);
}

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
graph_drop(cpic_dev *d)
{
	if(d->graph)
		hipGraphExecDestroy(d->graph);
	d->graph = NULL;
	d->graph_ok = 0;
}

static void
dev_free(cpic_dev *d)
{
	if(!d)
		return;
	if(d->stream)
		hipStreamSynchronize(d->stream);
	graph_drop(d);
	hipFree(d->rho[0]); hipFree(d->rho[1]); hipFree(d->ghat); hipFree(d->phi);
	hipFree(d->E); hipFree(d->costab); hipFree(d->tw);
	cpic_parts_free(d, 0);
	cpic_parts_free(d, 1);
	hipFree(d->bk.start); hipFree(d->bk.cap); hipFree(d->bk.n1); hipFree(d->bk.res);
	hipFree(d->bk.n2);
	hipFree(d->order);
	hipFree(d->border);
	hipFree(d->dpart);
	hipFree(d->lv.x); hipFree(d->lv.y); hipFree(d->lv.ux); hipFree(d->lv.uy);
	hipFree(d->lv.id); hipFree(d->lv.Ex); hipFree(d->lv.Ey); hipFree(d->lv.dst);
	hipFree(d->lv.src); hipFree(d->lv.count);
	hipFree(d->diag); hipFree(d->derr); hipFree(d->ctr); hipFree(d->st);
	hipFree(d->geo); hipFree(d->pre);
	if(d->up_buf)
		for(int k = 0; k < d->P * CPIC_MAX_SPECIES; k++)
		{
			free(d->up_buf[k]);
			free(d->up_id[k]);
		}
	free(d->up_buf);
	free(d->up_id);
	free(d->up_n);
	free(d->count);
	free(d->desc);
	free(d->hgeo);
	free(d->hpre);
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
	graph_drop(d);
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
	if((d->flags & CPIC_F_IDS) && hipMalloc(&l->id, sizeof(int64_t) * cap) != hipSuccess)
		return CPIC_ENOMEM;
	if((d->flags & CPIC_F_PARTICLE_E) &&
			(hipMalloc(&l->Ex, sizeof(double) * cap) != hipSuccess ||
			 hipMalloc(&l->Ey, sizeof(double) * cap) != hipSuccess))
		return CPIC_ENOMEM;
	l->cap = cap;
	return 0;
}

// sim_prepare's checks (src/sim.c:110-130) and this implementation's limits.
static int
desc_check(const cpic_dev_desc_t *desc)
{
	int64_t nx = desc->nx, ny = desc->ny, tx, ty;
	if(!is_pow2(nx) || !is_pow2(ny) || nx < 2 || ny < 2 || nx > 8192 || ny > 8192)
		return CPIC_EINVAL;
	if(!(desc->L[0] > 0) || !(desc->L[1] > 0) || !(desc->dt > 0) || desc->e0 == 0.0 ||
			!std::isfinite(desc->L[0]) || !std::isfinite(desc->L[1]) ||
			!std::isfinite(desc->dt) || !std::isfinite(desc->e0) || !std::isfinite(desc->B[2]))
		return CPIC_EINVAL;
	if(desc->nspecies < 1 || desc->nspecies > CPIC_MAX_SPECIES)
		return CPIC_EINVAL;
	if(desc->precision != 32 && desc->precision != 64)
		return CPIC_EINVAL;
	if(desc->ppt != 0 && desc->ppt != 2 && desc->ppt != 4)
		return CPIC_EINVAL;
	// In-plane B makes u_z non-zero, which the reference rejects through
	// umax_Z = 0 (src/sim.c:200, src/mover.c:113).
	if(desc->B[0] != 0.0 || desc->B[1] != 0.0)
		return CPIC_EINVAL;
	if(desc->nchunks < 1 || nx % desc->nchunks)   // src/sim.c:124-130
		return CPIC_EINVAL;
	for(int s = 0; s < desc->nspecies; s++)
		if(!(desc->m[s] != 0.0) || !std::isfinite(desc->m[s]) || !std::isfinite(desc->q[s]))
			return CPIC_EINVAL;
	tx = desc->tile[0] > 0 ? desc->tile[0] : 16;
	ty = desc->tile[1] > 0 ? desc->tile[1] : 16;
	if(tx > nx) tx = nx;
	if(ty > ny) ty = ny;
	if(!is_pow2(tx) || !is_pow2(ty) || tx > 64 || ty > 64)
		return CPIC_EINVAL;
	return 0;
}

// Table entry of one plasma (offsets are filled in by the caller).
static void
geo_fill(const cpic_dev_desc_t *desc, Geo *g)
{
	int64_t nx = desc->nx, ny = desc->ny;
	int64_t tx = desc->tile[0] > 0 ? desc->tile[0] : 16;
	int64_t ty = desc->tile[1] > 0 ? desc->tile[1] : 16;
	if(tx > nx) tx = nx;
	if(ty > ny) ty = ny;
	memset(g, 0, sizeof(*g));
	g->nx = (int) nx;
	g->ny = (int) ny;
	g->lgnx = ilog2(nx);
	g->lgny = ilog2(ny);
	g->tx = (int) tx;
	g->ty = (int) ty;
	g->lgtx = ilog2(tx);
	g->lgty = ilog2(ty);
	g->ntx = (int) (nx / tx);
	g->nty = (int) (ny / ty);
	g->lgntx = ilog2(g->ntx);
	g->ntiles = g->ntx * g->nty;
	g->M = CPIC_MARGIN;
	g->W = g->tx + 2 * g->M + 1;
	g->H = g->ty + 2 * g->M + 1;
	g->ns = desc->nspecies;
	g->dx = desc->L[0] / nx;      // src/sim.c:174
	g->dy = desc->L[1] / ny;
	g->idx = 1.0 / g->dx;         // src/interpolate.c:365-366
	g->idy = 1.0 / g->dy;
	g->L0 = desc->L[0];
	g->L1 = desc->L[1];
	g->dt = desc->dt;
	g->e0 = desc->e0;
	g->Bz = desc->B[2];
	// umax = chunksize * dx / dt (src/sim.c:198-200)
	g->umax_x = (double) (nx / desc->nchunks) * g->dx / desc->dt;
	g->umax_y = (double) ny * g->dy / desc->dt;
	for(int s = 0; s < desc->nspecies; s++)
	{
		g->q[s] = desc->q[s];
		g->m[s] = desc->m[s];
	}
}

static int
batch_create(const cpic_dev_desc_t *descs, int P, cpic_dev_t **out)
{
	cpic_dev *d;
	int ndev = 0, ret, p, k;
	const int MS = CPIC_MAX_SPECIES;

	if(!descs || !out || P < 1 || P > 65536)
		return CPIC_EINVAL;
	*out = NULL;
	for(p = 0; p < P; p++)
	{
		if((ret = desc_check(&descs[p])))
			return ret;
		// one precision, flag set, device and slack for the whole batch
		if(descs[p].precision != descs[0].precision || descs[p].flags != descs[0].flags ||
				descs[p].device != descs[0].device || descs[p].slack != descs[0].slack ||
				descs[p].ppt != descs[0].ppt)
			return CPIC_EINVAL;
	}
	if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
		return CPIC_ENODEV;
	if(descs[0].device < 0 || descs[0].device >= ndev)
		return CPIC_EINVAL;
	if(hipSetDevice(descs[0].device) != hipSuccess)
		return CPIC_EHIP;

	d = (cpic_dev *) calloc(1, sizeof(*d));
	if(!d)
		return CPIC_ENOMEM;
	d->P = P;
	d->prec = descs[0].precision;
	d->ppt = descs[0].ppt == 4 ? 4 : 2;
	d->flags = descs[0].flags;
	d->device = descs[0].device;
	d->ring = P <= 4 ? 65536 : (262144 / P > 4096 ? 262144 / P : 4096);
	d->desc = (cpic_dev_desc_t *) malloc(sizeof(cpic_dev_desc_t) * P);
	d->hgeo = (Geo *) calloc((size_t) P, sizeof(Geo));
	d->hpre = (int *) calloc((size_t) PRE_COUNT * (P + 1), sizeof(int));
	d->count = (int64_t *) calloc((size_t) P * MS, sizeof(int64_t));
	d->up_n = (int64_t *) calloc((size_t) P * MS, sizeof(int64_t));
	d->up_buf = (double **) calloc((size_t) P * MS, sizeof(double *));
	d->up_id = (int64_t **) calloc((size_t) P * MS, sizeof(int64_t *));
	if(!d->desc || !d->hgeo || !d->hpre || !d->count || !d->up_n || !d->up_buf || !d->up_id)
	{
		dev_free(d);
		return CPIC_ENOMEM;
	}
	memcpy(d->desc, descs, sizeof(cpic_dev_desc_t) * P);

	// table entries, work-list prefixes and totals
	int64_t nodes = 0, ghat = 0, cosn = 0, tiles = 0, buckets = 0;
	int64_t rows = 0, cols = 0, nblk = 0;
	size_t lds_rows = 0, lds_cols = 0;
	int *pt = d->hpre + PRE_TILE * (P + 1), *pr = d->hpre + PRE_ROW * (P + 1);
	int *pc = d->hpre + PRE_COL * (P + 1), *pn = d->hpre + PRE_NODEBLK * (P + 1);
	int *pb = d->hpre + PRE_BUCKET * (P + 1);
	int64_t mmax = 1, nymax = 1;
	for(p = 0; p < P; p++)
	{
		geo_fill(&descs[p], &d->hgeo[p]);
		if(d->hgeo[p].nx / 2 > mmax) mmax = d->hgeo[p].nx / 2;
		if(d->hgeo[p].ny > nymax) nymax = d->hgeo[p].ny;
	}
	// one block size per FFT pass for the whole batch
	cpic_fft_threads(mmax, nymax, &d->fft_nt_row, &d->fft_nt_col);
	for(p = 0; p < P; p++)
	{
		cpic_fft_blocking(&d->hgeo[p], d->fft_nt_row, d->fft_nt_col);
		if(!cpic_fft_blocking_ok(&d->hgeo[p], d->fft_nt_row, d->fft_nt_col))
		{
			dev_free(d);
			return CPIC_EINVAL;
		}
	}
	cpic_fft_spread(d->hgeo, P);
	for(p = 0; p < P; p++)
	{
		Geo *g = &d->hgeo[p];
		int64_t n = (int64_t) g->nx * g->ny, hx = g->nx / 2 + 1;
		g->node0 = nodes;
		g->ghat0 = ghat;
		g->cos0 = cosn;
		g->tile0 = (int) tiles;
		g->bucket0 = (int) buckets;
		g->diag0 = p * d->ring;
		pt[p] = (int) tiles;
		pr[p] = (int) rows;
		pc[p] = (int) cols;
		pn[p] = (int) nblk;
		pb[p] = (int) buckets;
		nodes += n;
		ghat += hx * g->ny;
		cosn += hx + g->ny;
		tiles += g->ntiles;
		buckets += (int64_t) g->ntiles * g->ns;
		rows += g->ny / g->R;
		cols += (g->nx / 2 + g->C - 1) / g->C;   // kx = 0 and nx/2 share a column
		nblk += (n + EF_NODES - 1) / EF_NODES;
		size_t lr = fft_lds_bytes((int64_t) g->R * hx);
		size_t lc = fft_lds_bytes((int64_t) g->C * g->ny);
		if(lr > lds_rows) lds_rows = lr;
		if(lc > lds_cols) lds_cols = lc;
		if(tiles > (1 << 30) || buckets > (1 << 30) || rows > (1 << 30) || cols > (1 << 30) ||
				nblk > (1 << 30))
		{
			dev_free(d);
			return CPIC_EINVAL;
		}
	}
	pt[P] = (int) tiles;
	pr[P] = (int) rows;
	pc[P] = (int) cols;
	pn[P] = (int) nblk;
	pb[P] = (int) buckets;
	d->nodes = nodes;
	d->ghat_n = ghat;
	d->ntiles = (int) tiles;
	d->nb = (int) buckets;
	d->rowblocks = (int) rows;
	d->colblocks = (int) cols;
	d->nodeblocks = (int) nblk;
	d->lds_rows = lds_rows;
	d->lds_cols = lds_cols;
	d->lds_push = cpic_push_lds(d);
	for(p = 0; p < P; p++)
		if(d->hgeo[p].ns > d->maxns)
			d->maxns = d->hgeo[p].ns;

	if(hipStreamCreate(&d->stream) != hipSuccess)
	{
		free(d);
		return CPIC_EHIP;
	}
	size_t esz = d->prec == 64 ? sizeof(double2) : sizeof(float2);
	size_t ng = (size_t) nodes;
	if((ret = alloc_zero((void **) &d->rho[0], sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->rho[1], sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->ghat, sizeof(double2) * (size_t) ghat)) ||
			(ret = alloc_zero((void **) &d->phi, sizeof(double) * ng)) ||
			(ret = alloc_zero((void **) &d->E, esz * ng)) ||
			(ret = alloc_zero((void **) &d->diag, sizeof(DiagSlot) * (size_t) P * d->ring)) ||
			(ret = alloc_zero((void **) &d->derr, sizeof(int))) ||
			(ret = alloc_zero((void **) &d->ctr, sizeof(Counters))) ||
			(ret = alloc_zero((void **) &d->st, sizeof(StepState))) ||
			(ret = alloc_zero((void **) &d->geo, sizeof(Geo) * P)) ||
			(ret = alloc_zero((void **) &d->pre, sizeof(int) * PRE_COUNT * (P + 1))) ||
			(ret = alloc_zero((void **) &d->order, sizeof(int) * d->ntiles)) ||
			(ret = alloc_zero((void **) &d->border, sizeof(int) * d->nb)) ||
			(ret = alloc_zero((void **) &d->dpart,
					sizeof(double) * (DIAG_SPREAD * DIAG_LINE * (size_t) P + DIAG_LINE))) ||
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
	if(hipMemcpy(d->geo, d->hgeo, sizeof(Geo) * P, hipMemcpyHostToDevice) != hipSuccess ||
			hipMemcpy(d->pre, d->hpre, sizeof(int) * PRE_COUNT * (P + 1),
				hipMemcpyHostToDevice) != hipSuccess)
	{
		dev_free(d);
		return CPIC_EHIP;
	}
	(void) k;
	if((ret = cpic_parts_order(d, NULL)))
	{
		dev_free(d);
		return ret;
	}
	d->rho_cur = 0;
	d->rho_next_ready = 1;   // both buffers start zeroed
	d->last_iter = -1;
	d->iter = INT64_MIN;     // device counter not set yet
	*out = d;
	return 0;
}

extern "C" int
cpic_dev_create(const cpic_dev_desc_t *desc, cpic_dev_t **out)
{
	return batch_create(desc, 1, out);
}

extern "C" int
cpic_batch_create(const cpic_dev_desc_t *descs, int32_t nplasmas, cpic_dev_t **out)
{
	return batch_create(descs, nplasmas, out);
}

extern "C" int
cpic_batch_size(cpic_dev_t *d, int32_t *nplasmas)
{
	if(!d || !nplasmas)
		return CPIC_EINVAL;
	*nplasmas = d->P;
	return 0;
}

extern "C" int
cpic_dev_destroy(cpic_dev_t *d)
{
	if(!d)
		return CPIC_EINVAL;
	set_device(d);
	dev_free(d);
	return 0;
}

// ---------------------------------------------------------- particles ---

static int
plasma_ok(const cpic_dev *d, int p)
{
	return d && p >= 0 && p < d->P;
}

extern "C" int
cpic_batch_upload_particles(cpic_dev_t *d, int32_t p, int32_t s, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy)
{
	BROKEN_CHECK(d);
	if(!plasma_ok(d, p) || s < 0 || s >= d->hgeo[p].ns || n < 0 ||
			(n > 0 && (!x || !y || !ux || !uy)))
		return CPIC_EINVAL;
	if((d->flags & CPIC_F_IDS) && n > 0 && !id)
		return CPIC_EINVAL;
	const Geo &g = d->hgeo[p];
	for(int64_t i = 0; i < n; i++)
		if(!(x[i] >= 0.0 && x[i] <= g.L0 && y[i] >= 0.0 && y[i] <= g.L1) ||
				!std::isfinite(ux[i]) || !std::isfinite(uy[i]))
			return CPIC_EINVAL;
	int k = p * CPIC_MAX_SPECIES + s;
	free(d->up_buf[k]);
	free(d->up_id[k]);
	d->up_buf[k] = (double *) malloc(sizeof(double) * 4 * (n ? n : 1));
	d->up_id[k] = NULL;
	if(!d->up_buf[k])
		return CPIC_ENOMEM;
	memcpy(d->up_buf[k], x, sizeof(double) * n);
	memcpy(d->up_buf[k] + n, y, sizeof(double) * n);
	memcpy(d->up_buf[k] + 2 * n, ux, sizeof(double) * n);
	memcpy(d->up_buf[k] + 3 * n, uy, sizeof(double) * n);
	if(d->flags & CPIC_F_IDS)
	{
		d->up_id[k] = (int64_t *) malloc(sizeof(int64_t) * (n ? n : 1));
		if(!d->up_id[k])
			return CPIC_ENOMEM;
		memcpy(d->up_id[k], id, sizeof(int64_t) * n);
	}
	d->up_n[k] = n;
	d->pending = 1;
	return 0;
}

extern "C" int
cpic_dev_upload_particles(cpic_dev_t *d, int32_t s, int64_t n, const int64_t *id,
		const double *x, const double *y, const double *ux, const double *uy)
{
	return cpic_batch_upload_particles(d, 0, s, n, id, x, y, ux, uy);
}

static int64_t
species_count(const cpic_dev *d, int p, int s)
{
	int k = p * CPIC_MAX_SPECIES + s;
	return d->up_buf[k] ? d->up_n[k] : d->count[k];
}

static int
commit(cpic_dev *d)
{
	int ret;
	int64_t tot = 0;
	if(!d->pending)
		return 0;
	set_device(d);
	for(int p = 0; p < d->P; p++)
		for(int s = 0; s < d->hgeo[p].ns; s++)
			tot += species_count(d, p, s);
	int64_t want = tot / 32;
	if(want < 16384)
		want = 16384;
	if(d->lv.cap < want)
		if((ret = leavers_alloc(d, want)))
			return ret;
	if((ret = cpic_parts_rebuild(d)))
		return ret;
	HIPCHK(hipMemsetAsync(d->ctr, 0, sizeof(Counters), d->stream));
	d->misplaced_seen = 0;
	d->fallbacks_seen = 0;
	return 0;
}

// A leaver buffer that overflowed left particles in buckets whose tile they
// had left: still correct, but gathered and deposited through the global
// path, and in FP32 with coordinates drifting away from their tile; and they
// try to leave again every step, so the demand snowballs.  At a sync point
// (between stages, no leaver in flight), when that has happened since the
// last check, the buffer grows to 1.25x the last step's demand (at least
// double, at most every particle) and the particles are re-sorted.  Leavers
// that found their destination bucket full went back to their source bucket
// (fallbacks) and retry likewise: a re-sort sizes the buckets for the
// current counts.  The counters stay cumulative.
static int
heal_overflow(cpic_dev *d)
{
	Counters c;
	int ret;
	HIPCHK(hipMemcpy(&c, d->ctr, sizeof(c), hipMemcpyDeviceToHost));
	if(c.misplaced <= d->misplaced_seen)
	{
		if(c.fallbacks <= d->fallbacks_seen)
			return 0;
		d->fallbacks_seen = c.fallbacks;
		return cpic_parts_rebuild(d);
	}
	d->misplaced_seen = c.misplaced;
	d->fallbacks_seen = c.fallbacks;
	int64_t tot = 0;
	for(int p = 0; p < d->P; p++)
		for(int s = 0; s < d->hgeo[p].ns; s++)
			tot += species_count(d, p, s);
	int32_t demand = 0;
	HIPCHK(hipMemcpy(&demand, d->lv.count + (1 - d->step_parity), sizeof(int32_t),
			hipMemcpyDeviceToHost));
	int64_t want = 2 * d->lv.cap;
	if(want < (int64_t) demand + demand / 4)
		want = (int64_t) demand + demand / 4;
	if(want > tot + 16384)
		want = tot + 16384;
	if(want > d->lv.cap && (ret = leavers_alloc(d, want)))
		return ret;
	return cpic_parts_rebuild(d);
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
	BROKEN_CHECK(d);
	if(!d)
		return CPIC_EINVAL;
	set_device(d);
	HIPCHK(hipStreamSynchronize(d->stream));
	int ret = derr_check(d);
	if(ret || d->pending)
		return ret;
	if((ret = heal_overflow(d)))
		return ret;
	HIPCHK(hipStreamSynchronize(d->stream));
	return 0;
}

extern "C" int
cpic_batch_count(cpic_dev_t *d, int32_t p, int32_t s, int64_t *n)
{
	if(!plasma_ok(d, p) || s < 0 || s >= d->hgeo[p].ns || !n)
		return CPIC_EINVAL;
	*n = species_count(d, p, s);
	return 0;
}

extern "C" int
cpic_dev_count(cpic_dev_t *d, int32_t s, int64_t *n)
{
	return cpic_batch_count(d, 0, s, n);
}

extern "C" int
cpic_batch_download_particles(cpic_dev_t *d, int32_t p, int32_t s, cpic_particles_t *out)
{
	BROKEN_CHECK(d);
	int ret;
	if(!plasma_ok(d, p) || !out || s < 0 || s >= d->hgeo[p].ns)
		return CPIC_EINVAL;
	if(out->id && !(d->flags & CPIC_F_IDS))
		return CPIC_EINVAL;
	if((out->Ex || out->Ey) && !(d->flags & CPIC_F_PARTICLE_E))
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	set_device(d);
	if((ret = cpic_parts_download(d, p, s, out)))
		return ret;
	return derr_check(d);
}

extern "C" int
cpic_dev_download_particles(cpic_dev_t *d, int32_t s, cpic_particles_t *out)
{
	return cpic_batch_download_particles(d, 0, s, out);
}

extern "C" int
cpic_batch_cell_counts(cpic_dev_t *d, int32_t p, int32_t s, int64_t *counts)
{
	BROKEN_CHECK(d);
	int ret;
	if(!plasma_ok(d, p) || !counts || s < 0 || s >= d->hgeo[p].ns)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	set_device(d);
	return cpic_parts_cell_counts(d, p, s, counts);
}

extern "C" int
cpic_dev_cell_counts(cpic_dev_t *d, int32_t s, int64_t *counts)
{
	return cpic_batch_cell_counts(d, 0, s, counts);
}

extern "C" int
cpic_dev_rebuild(cpic_dev_t *d)
{
	BROKEN_CHECK(d);
	int ret;
	if(!d)
		return CPIC_EINVAL;
	set_device(d);
	if(d->pending)
		return commit(d);
	if((ret = cpic_parts_rebuild(d)))
		return ret;
	HIPCHK(hipMemsetAsync(d->ctr, 0, sizeof(Counters), d->stream));
	d->misplaced_seen = 0;
	d->fallbacks_seen = 0;
	return 0;
}

// ------------------------------------------------------------- stages ---

// Makes the device step counter equal to iter before work that reads it.
static int
want_iter(cpic_dev *d, int64_t iter)
{
	if(d->iter == iter)
		return 0;
	return cpic_set_iter(d, iter);
}

// The E array is brought up to date with phi only when something reads or
// writes it, or before phi changes: field_E leaves it stale and the push
// computes the E of its tiles from phi with the same expression.
static int
ensure_E(cpic_dev *d)
{
	int ret;
	if(!d->E_stale)
		return 0;
	set_device(d);
	if((ret = cpic_efield(d)))
		return ret;
	d->E_stale = 0;
	return 0;
}

extern "C" int
cpic_dev_field_E(cpic_dev_t *d, int64_t iter)
{
	BROKEN_CHECK(d);
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	set_device(d);
	if((ret = want_iter(d, iter)))
		return ret;
	if((ret = cpic_solve_step(d)))
		return ret;
	d->E_stale = 1;
	return 0;
}

extern "C" int
cpic_dev_solve_xy(cpic_dev_t *d)
{
	BROKEN_CHECK(d);
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	if((ret = ensure_E(d)))
		return ret;
	set_device(d);
	return cpic_solve(d, 0);
}

extern "C" int
cpic_dev_plasma_E(cpic_dev_t *d, int64_t iter)
{
	BROKEN_CHECK(d);
	int ret;
	(void) iter;
	if(!d)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	if(d->flags & CPIC_F_PARTICLE_E)
	{
		if((ret = ensure_E(d)))
			return ret;
		return cpic_parts_gather(d);
	}
	return 0;
}

static int
zero_rho(cpic_dev *d, double *r)
{
	return cpic_zero(d, r, d->nodes);
}

extern "C" int
cpic_dev_plasma_r(cpic_dev_t *d, int64_t iter)
{
	BROKEN_CHECK(d);
	int ret;
	if(!d)
		return CPIC_EINVAL;
	if(iter < 0)
		return CPIC_ESTATE;
	if((ret = commit(d)))
		return ret;
	if(d->rho_pending)
		return CPIC_ESTATE;    // plasma_r twice without field_rho
	set_device(d);
	if(!d->rho_next_ready)
		if((ret = zero_rho(d, d->rho[1 - d->rho_cur])))
			return ret;
	if((ret = want_iter(d, iter)))
		return ret;
	if((ret = cpic_parts_push(d, iter == 0)))
		return ret;
	d->rho_pending = 1;
	d->rho_next_ready = 0;
	d->last_iter = iter;
	return 0;
}

extern "C" int
cpic_dev_field_rho(cpic_dev_t *d, int64_t iter)
{
	BROKEN_CHECK(d);
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
	set_device(d);
	if((ret = zero_rho(d, d->rho[d->rho_cur])))
		return ret;
	return cpic_parts_deposit(d, d->rho[d->rho_cur]);
}

extern "C" int
cpic_dev_pre_step(cpic_dev_t *d)
{
	BROKEN_CHECK(d);
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
	BROKEN_CHECK(d);
	int ret;
	if((ret = cpic_dev_field_E(d, iter)))
		return ret;
	if((ret = cpic_dev_plasma_E(d, iter)))
		return ret;
	if((ret = cpic_dev_plasma_r(d, iter)))
		return ret;
	return cpic_dev_field_rho(d, iter);
}

// Steps iter and iter+1 from the captured graph, capturing it first if
// needed.  Only called in the steady state: no pending upload, rho_next not
// zeroed yet, no deposit pending, device counter == iter, iter >= 1.
static int
graph_steps(cpic_dev *d, int64_t iter)
{
	int ret;
	if(d->graph_ok && (d->graph_parity != d->step_parity || d->graph_rho != d->rho_cur ||
				d->graph_pcur != d->pcur))
		return 1;   // out of phase: caller runs one step eagerly
	if(!d->graph_ok)
	{
		hipGraph_t g = NULL;
		graph_drop(d);
		d->graph_parity = d->step_parity;
		d->graph_rho = d->rho_cur;
		d->graph_pcur = d->pcur;
		HIPCHK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
		d->capturing = 1;
		ret = cpic_dev_step(d, iter);
		if(!ret)
			ret = cpic_dev_step(d, iter + 1);
		d->capturing = 0;
		hipError_t e = hipStreamEndCapture(d->stream, &g);
		if(ret || e != hipSuccess)
		{
			if(g)
				hipGraphDestroy(g);
			return ret ? ret : CPIC_EHIP;
		}
		e = hipGraphInstantiate(&d->graph, g, NULL, NULL, 0);
		hipGraphDestroy(g);
		if(e != hipSuccess)
		{
			d->graph = NULL;
			fprintf(stderr, "cpic_hip: hipGraphInstantiate failed: %s\n",
					hipGetErrorString(e));
			return CPIC_EHIP;
		}
		d->graph_ok = 1;
		// the capture advanced the host state by two steps; the launch
		// below performs them
	}
	else
	{
		d->iter += 2;
		d->last_iter = iter + 1;
	}
	HIPCHK(hipGraphLaunch(d->graph, d->stream));
	return 0;
}

extern "C" int
cpic_dev_run(cpic_dev_t *d, int64_t iter0, int64_t nsteps)
{
	BROKEN_CHECK(d);
	int ret;
	if(!d || nsteps < 0)
		return CPIC_EINVAL;
	int graphs = !(d->flags & (CPIC_F_NO_GRAPH | CPIC_F_PROFILE));
	for(int64_t i = 0; i < nsteps; )
	{
		int64_t it = iter0 + i;
		if(graphs && it >= 1 && nsteps - i >= 2 && !d->pending && !d->rho_pending &&
				!d->rho_next_ready && d->iter == it)
		{
			set_device(d);
			ret = graph_steps(d, it);
			if(ret < 0)
				return ret;
			if(ret == 0)
			{
				i += 2;
				continue;
			}
		}
		if((ret = cpic_dev_step(d, it)))
			return ret;
		i++;
	}
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
cpic_batch_download_field(cpic_dev_t *d, int32_t p, int32_t which, double *out)
{
	BROKEN_CHECK(d);
	int64_t n, o;
	int ret;
	if(!plasma_ok(d, p) || !out)
		return CPIC_EINVAL;
	if((ret = commit(d)))
		return ret;
	n = (int64_t) d->hgeo[p].nx * d->hgeo[p].ny;
	o = d->hgeo[p].node0;
	set_device(d);
	if(which == CPIC_RHO)
		HIPCHK(hipMemcpyAsync(out, d->rho[d->rho_cur] + o, sizeof(double) * n,
				hipMemcpyDeviceToHost, d->stream));
	else if(which == CPIC_PHI)
		HIPCHK(hipMemcpyAsync(out, d->phi + o, sizeof(double) * n, hipMemcpyDeviceToHost,
				d->stream));
	else if(which == CPIC_EX || which == CPIC_EY)
	{
		int c = which == CPIC_EX ? 0 : 1;
		int ret = ensure_E(d);
		if(ret)
			return ret;
		if(d->prec == 64)
		{
			double2 *h = (double2 *) malloc(sizeof(double2) * n);
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpyAsync(h, (double2 *) d->E + o, sizeof(double2) * n, hipMemcpyDeviceToHost,
					d->stream);
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
			hipMemcpyAsync(h, (float2 *) d->E + o, sizeof(float2) * n, hipMemcpyDeviceToHost,
					d->stream);
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
cpic_dev_download_field(cpic_dev_t *d, int32_t which, double *out)
{
	return cpic_batch_download_field(d, 0, which, out);
}

extern "C" int
cpic_batch_upload_field(cpic_dev_t *d, int32_t p, int32_t which, const double *in)
{
	BROKEN_CHECK(d);
	int64_t n, o;
	if(!plasma_ok(d, p) || !in)
		return CPIC_EINVAL;
	n = (int64_t) d->hgeo[p].nx * d->hgeo[p].ny;
	o = d->hgeo[p].node0;
	set_device(d);
	// phi and E are about to diverge: E keeps the value field_E gave it
	if(which == CPIC_PHI || which == CPIC_EX || which == CPIC_EY)
	{
		int ret = ensure_E(d);
		if(ret)
			return ret;
	}
	HIPCHK(hipStreamSynchronize(d->stream));
	if(which == CPIC_RHO)
		HIPCHK(hipMemcpy(d->rho[d->rho_cur] + o, in, sizeof(double) * n, hipMemcpyHostToDevice));
	else if(which == CPIC_PHI)
		HIPCHK(hipMemcpy(d->phi + o, in, sizeof(double) * n, hipMemcpyHostToDevice));
	else if(which == CPIC_EX || which == CPIC_EY)
	{
		int c = which == CPIC_EX ? 0 : 1;
		if(d->prec == 64)
		{
			double2 *h = (double2 *) malloc(sizeof(double2) * n);
			double2 *E = (double2 *) d->E + o;
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpy(h, E, sizeof(double2) * n, hipMemcpyDeviceToHost);
			for(int64_t i = 0; i < n; i++)
				if(c) h[i].y = in[i]; else h[i].x = in[i];
			hipMemcpy(E, h, sizeof(double2) * n, hipMemcpyHostToDevice);
			free(h);
		}
		else
		{
			float2 *h = (float2 *) malloc(sizeof(float2) * n);
			float2 *E = (float2 *) d->E + o;
			if(!h)
				return CPIC_ENOMEM;
			hipMemcpy(h, E, sizeof(float2) * n, hipMemcpyDeviceToHost);
			for(int64_t i = 0; i < n; i++)
				if(c) h[i].y = (float) in[i]; else h[i].x = (float) in[i];
			hipMemcpy(E, h, sizeof(float2) * n, hipMemcpyHostToDevice);
			free(h);
		}
	}
	else
		return CPIC_EINVAL;
	return 0;
}

extern "C" int
cpic_dev_upload_field(cpic_dev_t *d, int32_t which, const double *in)
{
	return cpic_batch_upload_field(d, 0, which, in);
}

// -------------------------------------------------------- diagnostics ---

extern "C" int
cpic_batch_energy(cpic_dev_t *d, int32_t p, int64_t max, double *field, double *kinetic,
		double *px, double *py, int64_t *n)
{
	BROKEN_CHECK(d);
	int64_t steps, k;
	DiagSlot *h;
	if(!plasma_ok(d, p) || max < 0)
		return CPIC_EINVAL;
	set_device(d);
	HIPCHK(hipStreamSynchronize(d->stream));
	steps = d->last_iter + 1;
	if(steps > d->ring)
		steps = d->ring;
	if(n)
		*n = steps;
	k = steps < max ? steps : max;
	if(k <= 0)
		return derr_check(d);
	h = (DiagSlot *) malloc(sizeof(DiagSlot) * k);
	if(!h)
		return CPIC_ENOMEM;
	if(hipMemcpy(h, d->diag + d->hgeo[p].diag0, sizeof(DiagSlot) * k, hipMemcpyDeviceToHost)
			!= hipSuccess)
	{
		free(h);
		return CPIC_EHIP;
	}
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
cpic_dev_energy(cpic_dev_t *d, int64_t max, double *field, double *kinetic,
		double *px, double *py, int64_t *n)
{
	return cpic_batch_energy(d, 0, max, field, kinetic, px, py, n);
}

// plasma owning global bucket b (host copy of the prefix array)
static int
bucket_plasma(const cpic_dev *d, int b)
{
	const int *pb = d->hpre + PRE_BUCKET * (d->P + 1);
	int lo = 0, hi = d->P - 1;
	while(lo < hi)
	{
		int mid = (lo + hi + 1) >> 1;
		if(pb[mid] <= b)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

extern "C" int
cpic_dev_migration(cpic_dev_t *d, int64_t *n, int64_t *id, int32_t *species,
		int32_t *src_tile, int32_t *dst_tile)
{
	BROKEN_CHECK(d);
	int32_t cnt = 0;
	int64_t k;
	if(!d || !n)
		return CPIC_EINVAL;
	if(id && !(d->flags & CPIC_F_IDS))
		return CPIC_EINVAL;
	set_device(d);
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
		const Geo &gs = d->hgeo[bucket_plasma(d, src[i])];
		const Geo &gd = d->hgeo[bucket_plasma(d, dst[i])];
		if(species) species[i] = (src[i] - gs.bucket0) % gs.ns;
		if(src_tile) src_tile[i] = gs.tile0 + (src[i] - gs.bucket0) / gs.ns;
		if(dst_tile) dst_tile[i] = gd.tile0 + (dst[i] - gd.bucket0) / gd.ns;
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
	set_device(d);
	HIPCHK(hipStreamSynchronize(d->stream));
	HIPCHK(hipMemcpy(&c, d->ctr, sizeof(c), hipMemcpyDeviceToHost));
	memset(st, 0, sizeof(*st));
	for(int p = 0; p < d->P; p++)
		for(int s = 0; s < d->hgeo[p].ns; s++)
			st->nparticles += species_count(d, p, s);
	st->capacity = d->pcap[d->pcur];
	st->nbuckets = d->nb;
	st->leavers = (int64_t) c.leavers;
	st->fallbacks = (int64_t) c.fallbacks;
	st->misplaced = (int64_t) c.misplaced;
	st->rebuilds = d->nrebuilds;
	st->tile[0] = d->hgeo[0].tx;
	st->tile[1] = d->hgeo[0].ty;
	st->precision = d->prec;
	return 0;
}

extern "C" int
cpic_dev_kernel_times(cpic_dev_t *d, int32_t *n, const char **names, int64_t *launches,
		double *total_ms)
{
	if(!d || !n)
		return CPIC_EINVAL;
	set_device(d);
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
cpic_dev_set_profiling(cpic_dev_t *d, int32_t on)
{
	if(!d)
		return CPIC_EINVAL;
	if(on)
		d->flags |= CPIC_F_PROFILE;
	else
	{
		kt_collect(d);
		d->flags &= ~CPIC_F_PROFILE;
	}
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
