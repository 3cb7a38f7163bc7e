// MFT Poisson solver on gfx950: rho -> phi -> E.
//
// Restates MFT_solve (src/solver.c:466-509): a 2D real-to-complex FFT of rho
// (ny x nx, X fastest), multiplication by the Green's function
// G(kx,ky) = 1/(2(cos 2pi kx/nx + cos 2pi ky/ny) - 4), G(0,0) = 0
// (src/solver.c:269-280), the complex-to-real inverse and the division by
// nx*ny (src/solver.c:366-379), then E = -grad(phi) by periodic central
// differences (field_E_compute, src/field.c:359-416).
//
// Three LDS passes, all HBM-bound:
//   1. k_rows_r2c   R rows per workgroup: half-length complex FFT of each
//                   row + real post-processing; writes ghat kx-major
//                   (transposed) so that R consecutive ky are contiguous.
//   2. k_cols       one or more kx columns per workgroup: forward FFT along
//                   ky, multiply by G/(nx*ny) (the normalisation is folded in;
//                   nx*ny is a power of two so the scaling is exact) and the
//                   field energy partial sum (as the last forward stage writes
//                   LDS when that stage fills the workgroup, else in a sweep
//                   of their own), inverse FFT along ky.
//   3. k_rows_c2r   R rows per workgroup: Hermitian pre-processing + inverse
//                   half-length FFT; writes phi rows, and zeroes the same rows
//                   of the rho buffer the next push deposits into and the next
//                   step's diagnostic slot.
// E = -grad(phi) is computed by the push for its tiles from phi (in LDS);
// k_efield writes the E array only when something reads it.
//
// The transforms are Stockham autosort FFTs in LDS (natural order in and
// out): radix-8 stages after one radix-2 or radix-4 stage for lg n mod 3,
// each thread holding up to FFT_MAXV complex values in registers per stage.
// Twiddles: each workgroup copies two small factor tables into LDS from one
// half table of exp(+2 pi i k/N), N = the largest grid edge of the batch,
// built with an octant reduction so the quarter points are exact; a
// butterfly's twiddle w is then the product of two LDS entries and its
// other powers w^r are products of w.  The block size of each
// pass is chosen per batch (cpic_fft_threads).  Each pass runs once for all
// plasmas of a batch: block b takes the b-th item of the pass's work list
// (row block or column block) and finds its plasma by binary search.
#include "dev.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static __device__ __forceinline__ double2 cmul(double2 a, double2 w)
{
	return make_double2(fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x));
}

// Twiddles of one block, in LDS: w(i) = exp(+2 pi i i/N) = hi[i >> a] *
// lo[i & (2^a - 1)], a = ceil(lg N / 2), filled from the half table tw[k] =
// exp(+2 pi i k/N), k <= N/2 (second half -tw[k - N/2]).  Two LDS reads and
// a complex product per twiddle instead of a global load (L2 latency in
// every Stockham stage); the product is within ~2 ulp of the table value.
#define TW_LDS 128     // 2^a and 2^(lg N - a) for N <= 8192
struct TwLds {
	double2 lo[TW_LDS];
	double2 hi[TW_LDS];
};
struct TwRef {
	const LDS_AS double *lo, *hi;   // (re, im) pairs
	int a;
};

static_assert((1 << 7) <= TW_LDS, "twiddle factors for N = 8192");

// Fills t (all threads of the block; the caller's next barrier publishes it).
static __device__ __forceinline__ TwRef tw_fill(TwLds &t, const double2 *__restrict__ tw,
		int lgN)
{
	int a = (lgN + 1) >> 1, b = lgN - a, half = 1 << (lgN - 1);
	for(int i = threadIdx.x; i < (1 << a); i += blockDim.x)
		t.lo[i] = tw[i];
	for(int i = threadIdx.x; i < (1 << b); i += blockDim.x)
	{
		int idx = i << a;
		double2 w;
		if(idx < half)
			w = tw[idx];
		else
		{
			w = tw[idx - half];
			w.x = -w.x;
			w.y = -w.y;
		}
		t.hi[i] = w;
	}
	TwRef r;
	r.lo = (const LDS_AS double *) t.lo;
	r.hi = (const LDS_AS double *) t.hi;
	r.a = a;
	return r;
}

// w = exp(sign * 2 pi i idx / N), 0 <= idx < N
static __device__ __forceinline__ double2 twiddle_at(const TwRef &t, int idx, double sign)
{
	int ih = 2 * (idx >> t.a), il = 2 * (idx & ((1 << t.a) - 1));
	double2 w = cmul(make_double2(t.hi[ih], t.hi[ih + 1]), make_double2(t.lo[il], t.lo[il + 1]));
	w.y *= sign;
	return w;
}

// In-register R-point DFT (R = 2, 4 or 8), exp(sign 2 pi i / R) roots.
template <int R>
static __device__ __forceinline__ void dft_small(double2 *v, double sign);

// 8 points: two 4-point DFTs of the even and odd points, combined with
// w^k, w = exp(sign 2 pi i / 8): w^1 = (c, sign c), w^2 = (0, sign),
// w^3 = (-c, sign c), c = sqrt(1/2).
static __device__ __forceinline__ void dft8(double2 *v, double sign)
{
	double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
	dft_small<4>(e, sign);
	dft_small<4>(o, sign);
	const double c = 0.70710678118654752440;
	double2 t[4];
	t[0] = o[0];
	// w^1 o1 = c (o1.x - sign o1.y, o1.y + sign o1.x)
	t[1] = make_double2(c * (o[1].x - sign * o[1].y), c * (o[1].y + sign * o[1].x));
	// w^2 o2 = sign i o2 = (-sign o2.y, sign o2.x)
	t[2] = make_double2(-sign * o[2].y, sign * o[2].x);
	// w^3 o3 = c (-o3.x - sign o3.y, -o3.y + sign o3.x)
	t[3] = make_double2(c * (-o[3].x - sign * o[3].y), c * (-o[3].y + sign * o[3].x));
#pragma unroll
	for(int k = 0; k < 4; k++)
	{
		v[k] = make_double2(e[k].x + t[k].x, e[k].y + t[k].y);
		v[k + 4] = make_double2(e[k].x - t[k].x, e[k].y - t[k].y);
	}
}

template <int R>
static __device__ __forceinline__ void dft_small(double2 *v, double sign)
{
	if(R == 8)
		dft8(v, sign);
	else if(R == 2)
	{
		double2 a = v[0], b = v[1];
		v[0] = make_double2(a.x + b.x, a.y + b.y);
		v[1] = make_double2(a.x - b.x, a.y - b.y);
	}
	else
	{
		double2 s02 = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
		double2 d02 = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
		double2 s13 = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
		double2 d13 = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
		// (i sign) * d13
		double2 j13 = make_double2(-sign * d13.y, sign * d13.x);
		v[0] = make_double2(s02.x + s13.x, s02.y + s13.y);
		v[1] = make_double2(d02.x + j13.x, d02.y + j13.y);
		v[2] = make_double2(s02.x - s13.x, s02.y - s13.y);
		v[3] = make_double2(d02.x - j13.x, d02.y - j13.y);
	}
}

// complex values a thread holds during one FFT stage: a block of NT threads
// transforms at most FFT_MAXV * NT values
#define FFT_MAXV 8
// at most 64 VGPRs per thread (eight waves per SIMD)
#define FFT_MIN_WAVES 8
static_assert(FFT_MAXV * FFT_THREADS >= FFT_MAX_VALUES, "FFT stage capacity");

// Multiplies v[r], r = 1 .. R-1, by w^r for a butterfly's twiddle w: one
// table lookup per butterfly instead of R-1, the powers by products at
// most three deep (w^2, w^4 = (w^2)^2, w^3 = w^2 w, w^5 = w^4 w,
// w^6 = w^4 w^2, w^7 = w^4 w^3), each within a few ulp of the exact root.
template <int R>
static __device__ __forceinline__ void apply_twiddles(double2 *v, double2 w1)
{
	v[1] = cmul(v[1], w1);
	if(R >= 4)
	{
		double2 w2 = cmul(w1, w1);
		v[2] = cmul(v[2], w2);
		double2 w3 = cmul(w2, w1);
		v[3] = cmul(v[3], w3);
		if(R == 8)
		{
			double2 w4 = cmul(w2, w2);
			v[4] = cmul(v[4], w4);
			v[5] = cmul(v[5], cmul(w4, w1));
			v[6] = cmul(v[6], cmul(w4, w2));
			v[7] = cmul(v[7], cmul(w4, w3));
		}
	}
}

struct NoPost {
	__device__ double2 operator()(int, int, double2 v) const { return v; }
};

// One Stockham stage of radix R over nbat transforms of length n = 2^lgn,
// current span p = 2^lgp: butterfly i of a transform reads x[i + t n/R],
// applies w^(t k) with w = exp(sign 2 pi i/(R p)), k = i mod p, and writes
// the R-point DFT to y[(i - k) R + k + t p].  x comes from the LDS buffer
// (transform b at b*stride, IN_LDS) or from in(b, j); y goes to the LDS
// buffer (OUT_LDS, as post(b, j, value)) or to out(b, j, value).  In place
// in LDS, all reads precede a barrier and all writes follow it, so one
// buffer holds both; a stage that writes LDS ends with a barrier.  Needs
// (n/R)*nbat <= (FFT_MAXV/R) * NT.
template <int R, int NT, bool IN_LDS, bool OUT_LDS, class In, class Out, class Post = NoPost>
static __device__ __forceinline__ void stockham_stage(double2 *buf, int stride, int lgn,
		int nbat, int lgp, const TwRef &tw, int lgN, double sign, In in, Out out,
		Post post = Post())
{
	constexpr int LGR = R == 8 ? 3 : R == 4 ? 2 : 1;
	constexpr int MAXB = FFT_MAXV / R;
	const int lgnb = lgn - LGR, nb = 1 << lgnb, p = 1 << lgp;
	const int total = nb * nbat;
	const int tsh = lgN - LGR - lgp;       // table index of w^(t k) = t k 2^tsh
	double2 v[MAXB][R];
	int jo[MAXB], bo[MAXB];
#pragma unroll
	for(int m = 0; m < MAXB; m++)
	{
		int t_ = threadIdx.x + m * NT;
		jo[m] = -1;
		bo[m] = 0;
		if(t_ < total)
		{
			int bat = t_ >> lgnb, i = t_ & (nb - 1);
			int k = i & (p - 1);
#pragma unroll
			for(int r = 0; r < R; r++)
				v[m][r] = IN_LDS ? buf[bat * stride + i + r * nb] : in(bat, i + r * nb);
			if(k)
			{
				apply_twiddles<R>(v[m], twiddle_at(tw, k << tsh, sign));
			}
			dft_small<R>(v[m], sign);
			jo[m] = ((i - k) << LGR) + k;
			bo[m] = bat;
		}
	}
	if(IN_LDS && OUT_LDS)
		__syncthreads();
#pragma unroll
	for(int m = 0; m < MAXB; m++)
		if(jo[m] >= 0)
		{
#pragma unroll
			for(int r = 0; r < R; r++)
			{
				if(OUT_LDS)
					buf[bo[m] * stride + jo[m] + r * p] = post(bo[m], jo[m] + r * p, v[m][r]);
				else
					out(bo[m], jo[m] + r * p, v[m][r]);
			}
		}
	if(OUT_LDS)
		__syncthreads();
}

struct NoIn {
	__device__ double2 operator()(int, int) const { return make_double2(0.0, 0.0); }
};
struct NoOut {
	__device__ void operator()(int, int, double2) const {}
};

// Radix of the first stage of a length-2^lgn transform: radix-8 stages after
// one radix-2 or radix-4 stage for lg n mod 3.
static __device__ __forceinline__ int first_lgr(int lgn)
{
	return lgn % 3 == 1 ? 1 : lgn % 3 == 2 ? 2 : 3;
}

// FFT of nbat transforms of length n = 2^lgn (Stockham autosort, natural
// order in and out), w = exp(sign 2 pi i/len), reading element j of
// transform b through in(b, j) and leaving post(b, j, X_j) in the LDS
// buffer (transform b at b*stride).  Ends at a barrier.
template <int NT, class In, class Post = NoPost>
static __device__ void fft_from(double2 *buf, int stride, int lgn, int nbat, const TwRef &tw,
		int lgN, double sign, In in, Post post = Post())
{
	NoOut none;
	if(lgn == 0)
	{
		for(int b = threadIdx.x; b < nbat; b += NT)
			buf[b * stride] = post(b, 0, in(b, 0));
		__syncthreads();
		return;
	}
	int l0 = first_lgr(lgn);
	if(l0 == lgn)
	{
		if(l0 == 1)
			stockham_stage<2, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none,
					post);
		else if(l0 == 2)
			stockham_stage<4, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none,
					post);
		else
			stockham_stage<8, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none,
					post);
		return;
	}
	if(l0 == 1)
		stockham_stage<2, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none);
	else if(l0 == 2)
		stockham_stage<4, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none);
	else
		stockham_stage<8, NT, false, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, in, none);
	int lgp = l0;
	for(; lgp + 3 < lgn; lgp += 3)
		stockham_stage<8, NT, true, true>(buf, stride, lgn, nbat, lgp, tw, lgN, sign,
				NoIn(), none);
	stockham_stage<8, NT, true, true>(buf, stride, lgn, nbat, lgp, tw, lgN, sign, NoIn(), none,
			post);
}

// The same transform from the LDS buffer (written before a barrier), its
// result written element by element through out(b, j, value).
template <int NT, class Out>
static __device__ void fft_to(double2 *buf, int stride, int lgn, int nbat, const TwRef &tw,
		int lgN, double sign, Out out)
{
	NoIn none;
	if(lgn == 0)
	{
		for(int b = threadIdx.x; b < nbat; b += NT)
			out(b, 0, buf[b * stride]);
		return;
	}
	int l0 = first_lgr(lgn);
	if(l0 == lgn)
	{
		if(l0 == 1)
			stockham_stage<2, NT, true, false>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, out);
		else if(l0 == 2)
			stockham_stage<4, NT, true, false>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, out);
		else
			stockham_stage<8, NT, true, false>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, out);
		return;
	}
	NoOut nout;
	if(l0 == 1)
		stockham_stage<2, NT, true, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, nout);
	else if(l0 == 2)
		stockham_stage<4, NT, true, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, nout);
	else
		stockham_stage<8, NT, true, true>(buf, stride, lgn, nbat, 0, tw, lgN, sign, none, nout);
	int lgp = l0;
	for(; lgp + 3 < lgn; lgp += 3)
		stockham_stage<8, NT, true, true>(buf, stride, lgn, nbat, lgp, tw, lgN, sign, none, nout);
	stockham_stage<8, NT, true, false>(buf, stride, lgn, nbat, lgp, tw, lgN, sign, none, out);
}

// Row block of workgroup b in the inverse row pass: contiguous bands of row
// blocks per XCD (workgroups are dispatched to the 8 XCDs round-robin), so
// the row blocks whose transposed loads (runs of R values of 16 B) share
// 128-byte lines of the spectrum run on one XCD and find them in its L2:
// k_rows_c2r 27.2 -> 24.0 us at 2048^2.  The forward pass keeps the plain
// order (its transposed stores measured 26.8 -> 28.1 us with the bands).
static __device__ __forceinline__ int row_block(int b, int n)
{
	return (n & 7) == 0 ? (b & 7) * (n >> 3) + (b >> 3) : b;
}

// Pass 1: R rows per block of NT threads; row y real of length nx = 2m.
template <int NT>
__global__ __launch_bounds__(NT, FFT_MIN_WAVES)
void k_rows_r2c(const Geo *__restrict__ geo, const int *__restrict__ pre, int P,
		const double *__restrict__ rho_all, double2 *__restrict__ ghat_all,
		const double2 *__restrict__ tw, int lgN)
{
	extern __shared__ double2 sh[];
	const int rb = blockIdx.x;
	int p = find_plasma(pre, P, rb);
	const Geo g = geo[p];
	int nx = g.nx, ny = g.ny, R = g.R;
	const double *__restrict__ rho = rho_all + g.node0;
	double2 *__restrict__ ghat = ghat_all + g.ghat0;
	int m = nx >> 1, lgm = g.lgnx - 1;
	int stride = m + 1;
	int y0 = (rb - pre[p]) * R;
	__shared__ TwLds s_tw;
	const TwRef twr = tw_fill(s_tw, tw, lgN);

	// the first stage reads the rows straight from rho (as complex pairs)
	const double2 *__restrict__ rows2 = (const double2 *) (rho + (int64_t) y0 * nx);
	fft_from<NT>(sh, stride, lgm, R, twr, lgN, -1.0,
			[=](int r, int j) { return rows2[(int64_t) r * m + j]; });

	// X[k] = E + W^k O, X[m-k] = conj(E - W^k O), W = exp(-2 pi i/nx)
	int npair = (m >> 1) + 1;
	int tsh = lgN - (lgm + 1);
	// fused with the transposed store ghat[k*ny + y]: consecutive threads
	// take consecutive rows of one k, so each k writes R contiguous values
	for(int i = threadIdx.x; i < R * npair; i += blockDim.x)
	{
		int k = i / R, r = i - k * R;
		const int b = r * stride;
		double2 xk, xm;
		if(k == 0)
		{
			double2 z = sh[b];
			xk = make_double2(z.x + z.y, 0.0);
			xm = make_double2(z.x - z.y, 0.0);
		}
		else if(2 * k == m)
		{
			double2 z = sh[b + k];
			xk = xm = make_double2(z.x, -z.y);
		}
		else
		{
			double2 zk = sh[b + k], zm = sh[b + m - k];
			// E = (zk + conj zm)/2, O = (zk - conj zm)/(2i)
			double er = 0.5 * (zk.x + zm.x), ei = 0.5 * (zk.y - zm.y);
			double dr = 0.5 * (zk.x - zm.x), di = 0.5 * (zk.y + zm.y);
			double2 O = make_double2(di, -dr);
			double2 wo = cmul(O, twiddle_at(twr, k << tsh, -1.0));
			xk = make_double2(er + wo.x, ei + wo.y);
			xm = make_double2(er - wo.x, -(ei - wo.y));
		}
		ghat[(int64_t) k * ny + y0 + r] = xk;
		if(2 * k != m)
			ghat[(int64_t) (m - k) * ny + y0 + r] = xm;
	}
}

// Pass 2: C columns (kx) per block of NT threads, each ny long and contiguous.
// Columns kx = 0 and kx = nx/2 of a real field are real sequences along y:
// the first block of a plasma transforms them together as one complex
// column (kx = 0 in the real part, kx = nx/2 in the imaginary part) and
// separates their spectra by symmetry, so a plasma has nx/2 columns to
// transform, not nx/2 + 1 (1024 columns at 2048^2: the blocks fill the chip
// in one round instead of leaving one block for a second).
template <int NT>
__global__ __launch_bounds__(NT, FFT_MIN_WAVES)
void k_cols(const Geo *__restrict__ geo, const int *__restrict__ pre, int P,
		double2 *__restrict__ ghat_all, const double *__restrict__ costab_all,
		const double2 *__restrict__ tw, int lgN, DiagSlot *__restrict__ diag,
		int ring, const StepState *__restrict__ st, double *__restrict__ eout)
{
	extern __shared__ double2 sh[];
	__shared__ double red[NT / 64];
	int p = find_plasma(pre, P, blockIdx.x);
	const Geo g = geo[p];
	int nx = g.nx, ny = g.ny, C = g.C;
	double2 *__restrict__ ghat = ghat_all + g.ghat0;
	int m = nx >> 1;
	int lgny = g.lgny;
	int k0 = (blockIdx.x - pre[p]) * C;
	int ncol = min(C, m - k0);
	const bool packed = k0 == 0;   // column 0 carries kx = 0 and kx = m
	const double *cosx = costab_all + g.cos0, *cosy = cosx + (m + 1);
	double invN = 1.0 / ((double) nx * (double) ny);
	long long iter = st->iter;
	// the field energy goes to eout[p] when given, else to the step's slot
	double *energy = eout ? eout + p :
		(diag && iter >= 0) ? &diag[g.diag0 + iter % ring].field : (double *) NULL;
	double escale = -0.5 * g.e0 * invN;
	__shared__ TwLds s_tw;
	const TwRef twr = tw_fill(s_tw, tw, lgN);

	// the first stage reads the columns straight from ghat
	auto load = [=](int c, int y) {
		return (packed && c == 0) ? make_double2(ghat[y].x, ghat[(int64_t) m * ny + y].x) :
			ghat[(int64_t) (k0 + c) * ny + y];
	};
	// G exactly as MFT_init computes it (src/solver.c:269-280)
	auto green = [&](int kx, int ky) {
		return (kx == 0 && ky == 0) ? 0.0 : 1.0 / (2.0 * (cosx[kx] + cosy[ky]) - 4.0);
	};
	double e = 0.0;
	// butterflies of the last forward stage (radix 8 from 8 points up)
	const int lastb = ncol * (ny >= 8 ? ny / 8 : 1);
	if(!packed && lastb >= NT)
	{
		// G/(nx ny) applied and the energy summed as the last forward stage
		// writes each value (one LDS round trip and barrier fewer); only
		// when that stage keeps every thread busy, as the Green's function
		// costs a division per value
		fft_from<NT>(sh, ny, lgny, ncol, twr, lgN, -1.0, load, [&](int c, int ky, double2 v) {
			double G = green(k0 + c, ky);
			e += 2.0 * G * (v.x * v.x + v.y * v.y);
			double gs = G * invN;
			return make_double2(v.x * gs, v.y * gs);
		});
	}
	else
	{
		fft_from<NT>(sh, ny, lgny, ncol, twr, lgN, -1.0, load);
		// the packed column's ky pairs (ky, ny - ky) first (one thread per pair),
		// then the other columns' values
		const int npair = packed ? (ny >> 1) + 1 : 0;
		for(int j = threadIdx.x; j < npair + (ncol - (packed ? 1 : 0)) * ny; j += blockDim.x)
		{
			if(j < npair)
			{
				// Z = A + iB with A, B the spectra of the real columns kx = 0
				// and kx = m: A[k] = (Z[k] + conj Z[-k])/2, B[k] = (Z[k] -
				// conj Z[-k])/(2i), A[-k] = conj A[k], B[-k] = conj B[k]
				int k = j, kn = (ny - j) & (ny - 1);
				double2 zk = sh[k], zn = sh[kn];
				double2 A = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
				double2 B = make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
				double a2 = A.x * A.x + A.y * A.y, b2 = B.x * B.x + B.y * B.y;
				double ga = green(0, k), gb = green(m, k);
				e += ga * a2 + gb * b2;
				ga *= invN;
				gb *= invN;
				// Z'[k] = ga A + i gb B
				sh[k] = make_double2(ga * A.x - gb * B.y, ga * A.y + gb * B.x);
				if(kn != k)
				{
					double gan = green(0, kn), gbn = green(m, kn);
					e += gan * a2 + gbn * b2;
					gan *= invN;
					gbn *= invN;
					// Z'[-k] = gan conj(A) + i gbn conj(B)
					sh[kn] = make_double2(gan * A.x + gbn * B.y, -gan * A.y + gbn * B.x);
				}
				continue;
			}
			int i = j - npair + (packed ? ny : 0);
			int c = i >> lgny, ky = i & (ny - 1);
			int kx = k0 + c;
			double G = green(kx, ky);
			double2 v = sh[i];
			e += 2.0 * G * (v.x * v.x + v.y * v.y);
			double g = G * invN;
			sh[i] = make_double2(v.x * g, v.y * g);
		}
		__syncthreads();
	}
	if(energy)
	{
		// wavefront sums, then one barrier
		for(int off = 32; off > 0; off >>= 1)
			e += __shfl_down(e, off);
		if((threadIdx.x & 63) == 0)
			red[threadIdx.x >> 6] = e;
		__syncthreads();
		if(threadIdx.x == 0)
		{
			double t = 0.0;
			for(int w = 0; w < NT / 64; w++)
				t += red[w];
			unsafeAtomicAdd(energy, t * escale);
		}
	}
	// the last stage writes the columns straight to ghat
	fft_to<NT>(sh, ny, lgny, ncol, twr, lgN, +1.0, [=](int c, int y, double2 v) {
		if(packed && c == 0)
		{
			ghat[y] = make_double2(v.x, 0.0);
			ghat[(int64_t) m * ny + y] = make_double2(v.y, 0.0);
		}
		else
			ghat[(int64_t) (k0 + c) * ny + y] = v;
	});
}

// Pass 3: R rows per block, c2r along X (imaginary parts of X[0] and
// X[nx/2] ignored, as for a Hermitian input).
template <int NT>
__global__ __launch_bounds__(NT, FFT_MIN_WAVES)
void k_rows_c2r(const Geo *__restrict__ geo, const int *__restrict__ pre, int P,
		const double2 *__restrict__ ghat_all, double *__restrict__ phi_all,
		const double2 *__restrict__ tw, int lgN, double *__restrict__ rho_zero_all,
		DiagSlot *__restrict__ diag, int ring, const StepState *__restrict__ st)
{
	extern __shared__ double2 sh[];
	const int rb = row_block(blockIdx.x, gridDim.x);
	int p = find_plasma(pre, P, rb);
	const Geo g = geo[p];
	int nx = g.nx, ny = g.ny, R = g.R;
	const double2 *__restrict__ ghat = ghat_all + g.ghat0;
	double *__restrict__ phi = phi_all + g.node0;
	int m = nx >> 1, lgm = g.lgnx - 1;
	int stride = m + 1;
	int y0 = (rb - pre[p]) * R;
	__shared__ TwLds s_tw;
	const TwRef twr = tw_fill(s_tw, tw, lgN);

	// Z[k] = A + iB, A = X[k] + conj X[m-k], B = (X[k] - conj X[m-k]) e^{+2 pi i k/nx}
	int npair = (m >> 1) + 1;
	int tsh = lgN - (lgm + 1);
	// fused with the transposed load of X[k] and X[m-k] (consecutive
	// threads take consecutive rows of one k); the twiddle factors are read
	// right away, so their fill is published first
	__syncthreads();
	for(int i = threadIdx.x; i < R * npair; i += blockDim.x)
	{
		int k = i / R, r = i - k * R;
		const int b = r * stride;
		double2 xk = ghat[(int64_t) k * ny + y0 + r];
		double2 xm = ghat[(int64_t) (m - k) * ny + y0 + r];
		if(k == 0)
			sh[b] = make_double2(xk.x + xm.x, xk.x - xm.x);
		else if(2 * k == m)
			sh[b + k] = make_double2(2.0 * xk.x, -2.0 * xk.y);
		else
		{
			double2 wk = twiddle_at(twr, k << tsh, 1.0);         // e^{+2 pi i k/nx}
			double2 wm = twiddle_at(twr, (m - k) << tsh, 1.0);   // e^{+2 pi i (m-k)/nx}
			double2 A = make_double2(xk.x + xm.x, xk.y - xm.y);
			double2 D = make_double2(xk.x - xm.x, xk.y + xm.y);
			double2 B = cmul(D, wk);
			double2 A2 = make_double2(xm.x + xk.x, xm.y - xk.y);
			double2 D2 = make_double2(xm.x - xk.x, xm.y + xk.y);
			double2 B2 = cmul(D2, wm);
			sh[b + k] = make_double2(A.x - B.y, A.y + B.x);
			sh[b + m - k] = make_double2(A2.x - B2.y, A2.y + B2.x);
		}
	}
	__syncthreads();
	// the last stage writes the rows straight to phi (as complex pairs)
	double2 *__restrict__ prow = (double2 *) (phi + (int64_t) y0 * nx);
	fft_to<NT>(sh, stride, lgm, R, twr, lgN, +1.0,
			[=](int r, int j, double2 v) { prow[(int64_t) r * m + j] = v; });
	// the same rows of the next rho buffer, which the coming push deposits
	// into, and the next step's diagnostic slot start at zero
	if(rho_zero_all)
	{
		double *__restrict__ rz = rho_zero_all + g.node0 + (int64_t) y0 * nx;
		for(int i = threadIdx.x; i < R * nx; i += blockDim.x)
			rz[i] = 0.0;
	}
	if(diag && rb == pre[p] && threadIdx.x == 0)
	{
		DiagSlot *z = &diag[g.diag0 + (st->iter + 1) % ring];
		z->field = 0.0;
		z->kinetic = 0.0;
		z->px = 0.0;
		z->py = 0.0;
	}
}

// E = -grad(phi) (src/field.c:381-406), zeroing of the next rho buffer and
// of the next diagnostic slot.  Work items: blocks of EF_NODES nodes.
template <typename T2>
__global__ __launch_bounds__(CPIC_BLOCK)
void k_efield(const Geo *__restrict__ geo, const int *__restrict__ pre, int P,
		const double *__restrict__ phi_all, T2 *__restrict__ E_all,
		double *__restrict__ rho_zero, DiagSlot *__restrict__ diag, int ring,
		const StepState *__restrict__ st)
{
	int p = find_plasma(pre, P, blockIdx.x);
	const Geo g = geo[p];
	int nx = g.nx, ny = g.ny;
	int blk = blockIdx.x - pre[p];
	const double *__restrict__ phi = phi_all + g.node0;
	T2 *__restrict__ E = E_all + g.node0;
	double dx2 = 2 * g.dx, dy2 = 2 * g.dy;
	int64_t n = (int64_t) nx * ny;
	int64_t i0 = (int64_t) blk * EF_NODES;
	int64_t i1 = i0 + EF_NODES < n ? i0 + EF_NODES : n;
	for(int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
	{
		int iy = (int) (i >> g.lgnx), ix = (int) (i & (nx - 1));
		int x0 = ix == 0 ? nx - 1 : ix - 1, x1 = ix == nx - 1 ? 0 : ix + 1;
		int y0 = iy == 0 ? ny - 1 : iy - 1, y1 = iy == ny - 1 ? 0 : iy + 1;
		const double *row = phi + (int64_t) iy * nx;
		double ey = (phi[(int64_t) y0 * nx + ix] - phi[(int64_t) y1 * nx + ix]) / dy2;
		double ex = (row[x0] - row[x1]) / dx2;
		T2 v;
		v.x = ex;
		v.y = ey;
		E[i] = v;
		if(rho_zero)
			rho_zero[g.node0 + i] = 0.0;
	}
	if(diag && blk == 0 && threadIdx.x == 0)
	{
		DiagSlot *z = &diag[g.diag0 + (st->iter + 1) % ring];
		z->field = 0.0;
		z->kinetic = 0.0;
		z->px = 0.0;
		z->py = 0.0;
	}
}

// ---------------------------------------------------------------- host ---

// exp(+2 pi i k/n), 0 <= k <= n/2, by octant reduction so that the
// quarter points are exact (host tables; the allocation-free solve builds
// its table on the device with the same reduction)
static __host__ __device__ void
twiddle(int64_t k, int64_t n, double *c, double *s)
{
	int64_t k8 = 8 * k;
	double a;
	if(k8 <= n)
	{
		a = 2.0 * M_PI * (double) k / (double) n;
		*c = cos(a); *s = sin(a);
	}
	else if(k8 <= 2 * n)
	{
		a = 2.0 * M_PI * (double) (n / 4 - k) / (double) n;
		*c = sin(a); *s = cos(a);
	}
	else if(k8 <= 3 * n)
	{
		a = 2.0 * M_PI * (double) (k - n / 4) / (double) n;
		*c = -sin(a); *s = cos(a);
	}
	else
	{
		a = 2.0 * M_PI * (double) (n / 2 - k) / (double) n;
		*c = -cos(a); *s = sin(a);
	}
}

// One row of 8192 points needs 64 KB + 16 B of LDS, a column 128 KB: the
// passes' dynamic LDS limit is raised, leaving room for each kernel's
// static LDS.
static int
fft_lds_attrs(void)
{
	const void *kern[] = {
		(const void *) k_rows_r2c<256>, (const void *) k_rows_c2r<256>, (const void *) k_cols<256>,
		(const void *) k_rows_r2c<512>, (const void *) k_rows_c2r<512>, (const void *) k_cols<512>,
		(const void *) k_rows_r2c<1024>, (const void *) k_rows_c2r<1024>,
		(const void *) k_cols<1024>};
	for(unsigned i = 0; i < sizeof(kern) / sizeof(kern[0]); i++)
	{
		hipFuncAttributes fa;
		HIPCHK(hipFuncGetAttributes(&fa, kern[i]));
		size_t lim = CPIC_LDS_LIMIT;
		if(lim + fa.sharedSizeBytes > 160 * 1024)
			lim = 160 * 1024 - fa.sharedSizeBytes;
		HIPCHK(hipFuncSetAttribute(kern[i], hipFuncAttributeMaxDynamicSharedMemorySize,
				(int) lim));
	}
	return 0;
}

int
cpic_fft_init(cpic_dev *d)
{
	int64_t N = 1, k, ncos = 0;
	double *ct;
	double2 *tw;
	int p;

	for(p = 0; p < d->P; p++)
	{
		const Geo *g = &d->hgeo[p];
		if(g->nx > N) N = g->nx;
		if(g->ny > N) N = g->ny;
		ncos += g->nx / 2 + 1 + g->ny;
	}
	d->lgN = ilog2(N);
	ct = (double *) malloc(sizeof(double) * (size_t) ncos);
	tw = (double2 *) malloc(sizeof(double2) * (size_t) (N / 2 + 1));
	if(!ct || !tw)
	{
		free(ct); free(tw);
		return CPIC_ENOMEM;
	}
	// MFT_init's cosines (src/solver.c:258-280), bit-identical
	for(p = 0; p < d->P; p++)
	{
		const Geo *g = &d->hgeo[p];
		int64_t hx = g->nx / 2 + 1;
		double cx = 2.0 * M_PI / g->nx, cy = 2.0 * M_PI / g->ny;
		for(k = 0; k < hx; k++)
			ct[g->cos0 + k] = cos(cx * k);
		for(k = 0; k < g->ny; k++)
			ct[g->cos0 + hx + k] = cos(cy * k);
	}
	for(k = 0; k <= N / 2; k++)
		twiddle(k, N, &tw[k].x, &tw[k].y);

	if(hipMalloc(&d->costab, sizeof(double) * (size_t) ncos) != hipSuccess ||
			hipMalloc(&d->tw, sizeof(double2) * (size_t) (N / 2 + 1)) != hipSuccess)
	{
		free(ct); free(tw);
		return CPIC_ENOMEM;
	}
	hipError_t e1 = hipMemcpy(d->costab, ct, sizeof(double) * (size_t) ncos,
			hipMemcpyHostToDevice);
	hipError_t e2 = hipMemcpy(d->tw, tw, sizeof(double2) * (size_t) (N / 2 + 1),
			hipMemcpyHostToDevice);
	free(ct);
	free(tw);
	if(e1 != hipSuccess || e2 != hipSuccess)
		return CPIC_EHIP;
	return fft_lds_attrs();
}

// Threads per block of the row and column passes for a batch whose largest
// row holds mmax = nx/2 complex values and whose longest column ny = nymax:
// FFT_ROW_NT / FFT_COL_NT threads when one row / column fits such a block
// (FFT_MAXV values per thread), else FFT_THREADS.
void
cpic_fft_threads(int64_t mmax, int64_t nymax, int *nt_row, int *nt_col)
{
	*nt_row = mmax <= (int64_t) FFT_MAXV * FFT_ROW_NT ? FFT_ROW_NT : FFT_THREADS;
	*nt_col = nymax <= (int64_t) FFT_MAXV * FFT_COL_NT ? FFT_COL_NT : FFT_THREADS;
}

// FFT blocking of one plasma for row blocks of nt_row threads and column
// blocks of nt_col threads: R rows per row block (a power of two dividing
// ny), C columns per column block, each within the pass's values per
// thread and LDS budget (FFT_THREADS blocks: 4 values per thread, 76 KB and
// 64 KB, so that two share a CU).
void
cpic_fft_blocking(Geo *g, int nt_row, int nt_col)
{
	int64_t m = g->nx / 2 > 0 ? g->nx / 2 : 1;
	int64_t per = (int64_t) (g->nx / 2 + 1) * (int64_t) sizeof(double2);
	int64_t vr = nt_row == FFT_THREADS ? 4 * FFT_THREADS : (int64_t) FFT_ROW_VPT * nt_row;
	int64_t lr = nt_row == FFT_THREADS ? FFT_ROW_LDS_BYTES : FFT_ROW_LDS_SMALL;
	int64_t R = lr / per, r = 1;
	if(R * m > vr)
		R = vr / m;
	if(R < 1)
		R = 1;
	if(R > g->ny)
		R = g->ny;
	while(r * 2 <= R)
		r *= 2;
	while(r > 1 && (int64_t) fft_lds_bytes(r * (g->nx / 2 + 1)) > lr)
		r /= 2;
	g->R = (int) r;
	int64_t vc = nt_col == FFT_THREADS ? 4 * FFT_THREADS : (int64_t) FFT_MAXV * nt_col;
	int64_t lc = nt_col == FFT_THREADS ? FFT_LDS_BYTES : FFT_COL_LDS_SMALL;
	int64_t C = lc / ((int64_t) g->ny * (int64_t) sizeof(double2));
	if(C * g->ny > vc)
		C = vc / g->ny;
	if(C < 1)
		C = 1;
	if(C > g->nx / 2)
		C = g->nx / 2;
	if(C > 16)
		C = 16;
	while(C > 1 && (int64_t) fft_lds_bytes(C * g->ny) > lc)
		C--;
	g->C = (int) C;
}

// Fewer blocks per FFT pass than CUs leave most of the chip idle (a single
// 256^2 grid has 8 row blocks of 32 rows): halve the rows (columns) per
// block of the plasma with the most, down to one, until the batch has
// FFT_MIN_BLOCKS blocks per pass.  Large grids and big batches keep their
// blocking.
void
cpic_fft_spread(Geo *g, int P)
{
	for(;;)
	{
		int64_t rows = 0;
		int best = -1;
		for(int p = 0; p < P; p++)
		{
			rows += g[p].ny / g[p].R;
			if(g[p].R > 1 && (best < 0 || g[p].R > g[best].R))
				best = p;
		}
		if(rows >= FFT_MIN_BLOCKS || best < 0)
			break;
		g[best].R /= 2;
	}
	for(;;)
	{
		int64_t cols = 0;
		int best = -1;
		for(int p = 0; p < P; p++)
		{
			int64_t hx = g[p].nx / 2;     // columns to transform (kx = 0 and nx/2 packed)
			cols += (hx + g[p].C - 1) / g[p].C;
			if(g[p].C > 1 && (best < 0 || g[p].C > g[best].C))
				best = p;
		}
		if(cols >= FFT_MIN_BLOCKS || best < 0)
			break;
		g[best].C /= 2;
	}
}

// Whether the blocking keeps every FFT pass within its blocks' values
// (FFT_MAXV per thread) and the LDS limit per block.
int
cpic_fft_blocking_ok(const Geo *g, int nt_row, int nt_col)
{
	// dynamic LDS, leaving 16 KB for the kernels' static LDS (twiddle
	// factors, reduction partials)
	const int64_t lim = 144 * 1024;
	return (int64_t) fft_lds_bytes((int64_t) g->R * (g->nx / 2 + 1)) <= lim &&
		(int64_t) fft_lds_bytes((int64_t) g->C * g->ny) <= lim &&
		(int64_t) g->R * (g->nx / 2) <= (int64_t) FFT_MAXV * nt_row &&
		(int64_t) g->C * g->ny <= (int64_t) FFT_MAXV * nt_col;
}

// Launch of an FFT pass with blocks of nt threads (the instantiations for
// the thread counts cpic_fft_threads chooses: 256, 512 or 1024).
#define FFT_LAUNCH(kern, nt, grid, lds, stream, ...) do { \
		if((nt) == 256) \
			hipLaunchKernelGGL(kern<256>, dim3(grid), dim3(256), lds, stream, __VA_ARGS__); \
		else if((nt) == 512) \
			hipLaunchKernelGGL(kern<512>, dim3(grid), dim3(512), lds, stream, __VA_ARGS__); \
		else \
			hipLaunchKernelGGL(kern<1024>, dim3(grid), dim3(1024), lds, stream, __VA_ARGS__); \
	} while(0)

// phi = S rho for every plasma of the batch (rho and phi laid out like the
// state's own buffers); with diag, the field energy of each plasma goes to
// its slot of the current step, with eout, it is added to eout[p].
static int
solve(cpic_dev *d, const double *rho, double *phi, DiagSlot *diag, double *rho_zero,
		DiagSlot *diag_zero, double *eout = NULL)
{
	int slot;
	const int P = d->P;
	const int *pre = d->pre;

	slot = cpic_kt_begin(d, "fft_rows_r2c");
	FFT_LAUNCH(k_rows_r2c, d->fft_nt_row, d->rowblocks, d->lds_rows, d->stream,
			d->geo, pre + PRE_ROW * (P + 1), P, rho, d->ghat, d->tw, d->lgN);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());

	slot = cpic_kt_begin(d, "fft_cols_green");
	FFT_LAUNCH(k_cols, d->fft_nt_col, d->colblocks, d->lds_cols, d->stream,
			d->geo, pre + PRE_COL * (P + 1), P, d->ghat, d->costab, d->tw, d->lgN, diag,
			d->ring, d->st, eout);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());

	slot = cpic_kt_begin(d, "fft_rows_c2r");
	FFT_LAUNCH(k_rows_c2r, d->fft_nt_row, d->rowblocks, d->lds_rows, d->stream,
			d->geo, pre + PRE_ROW * (P + 1), P, d->ghat, phi, d->tw, d->lgN, rho_zero,
			diag_zero, d->ring, d->st);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	return 0;
}

int
cpic_solve(cpic_dev *d, int want_energy)
{
	DiagSlot *diag = (want_energy && !(d->flags & CPIC_F_NO_DIAG)) ? d->diag : (DiagSlot *) NULL;
	return solve(d, d->rho[d->rho_cur], d->phi, diag, (double *) NULL, (DiagSlot *) NULL);
}

// field_E's solve: also zeroes the rho buffer the next push deposits into
// (unless already zero or holding a pending deposit) and the next step's
// diagnostic slot, in the inverse row pass.  E is then computed from phi by
// the push itself (or by cpic_efield when something reads E).
int
cpic_solve_step(cpic_dev *d)
{
	double *rz = NULL;
	DiagSlot *diag = (d->flags & CPIC_F_NO_DIAG) ? (DiagSlot *) NULL : d->diag;
	if(!d->rho_next_ready && !d->rho_pending)
		rz = d->rho[1 - d->rho_cur];
	int ret = solve(d, d->rho[d->rho_cur], d->phi, diag, rz, diag);
	if(!ret && rz)
		d->rho_next_ready = 1;
	return ret;
}

int
cpic_solve_ptr(cpic_dev *d, const double *rho, double *phi)
{
	return solve(d, rho, phi, (DiagSlot *) NULL, (double *) NULL, (DiagSlot *) NULL);
}

int
cpic_solve_ptr_zero(cpic_dev *d, const double *rho, double *phi, double *rho_zero)
{
	return solve(d, rho, phi, (DiagSlot *) NULL, rho_zero, (DiagSlot *) NULL);
}

int
cpic_solve_ptr_energy(cpic_dev *d, const double *rho, double *phi, double *rho_zero,
		double *energy)
{
	return solve(d, rho, phi, (DiagSlot *) NULL, rho_zero, (DiagSlot *) NULL, energy);
}

int
cpic_efield_ptr(cpic_dev *d, const double *phi, void *E)
{
	const int P = d->P;
	if(d->prec == 64)
		hipLaunchKernelGGL(k_efield<double2>, dim3(d->nodeblocks), dim3(CPIC_BLOCK), 0, d->stream,
				d->geo, d->pre + PRE_NODEBLK * (P + 1), P, phi, (double2 *) E,
				(double *) NULL, (DiagSlot *) NULL, d->ring, d->st);
	else
		hipLaunchKernelGGL(k_efield<float2>, dim3(d->nodeblocks), dim3(CPIC_BLOCK), 0, d->stream,
				d->geo, d->pre + PRE_NODEBLK * (P + 1), P, phi, (float2 *) E,
				(double *) NULL, (DiagSlot *) NULL, d->ring, d->st);
	HIPCHK(hipGetLastError());
	return 0;
}

// E from phi into the E array (for the callers that read E; the step's push
// computes its tiles' E from phi itself).
int
cpic_efield(cpic_dev *d)
{
	double *rz = NULL;
	DiagSlot *dz = NULL;
	int slot;
	const int P = d->P;

	slot = cpic_kt_begin(d, "efield");
	if(d->prec == 64)
		hipLaunchKernelGGL(k_efield<double2>, dim3(d->nodeblocks), dim3(CPIC_BLOCK), 0, d->stream,
				d->geo, d->pre + PRE_NODEBLK * (P + 1), P, d->phi, (double2 *) d->E, rz, dz,
				d->ring, d->st);
	else
		hipLaunchKernelGGL(k_efield<float2>, dim3(d->nodeblocks), dim3(CPIC_BLOCK), 0, d->stream,
				d->geo, d->pre + PRE_NODEBLK * (P + 1), P, d->phi, (float2 *) d->E, rz, dz,
				d->ring, d->st);
	cpic_kt_end(d, slot);
	HIPCHK(hipGetLastError());
	return 0;
}

// ------------------------------------------------- allocation-free solve ---
//
// cpic_solve_xy_noalloc runs the three passes above on one grid for callers
// that cannot allocate device memory.  Its spectrum, cosine and twiddle
// tables and the passes' one-entry work lists live in `scratch`, a static
// device array that each GPU holds from the library's load on that GPU to
// the end of the process.  In host code the name `scratch` stands for a
// host-side shadow, not a device address: the device address is looked up
// once per GPU with hipGetSymbolAddress and kept in ws_addr[], and only that
// address is handed to kernels.

#define WS_GEO 0            // Geo of the grid
#define WS_PRE 512          // row work list {0, rows}, column work list {0, cols}
#define WS_STEP 768         // StepState (iter -1: the passes record no energy)
#define WS_TABLES 1024      // cosines, twiddles, spectrum, each 256-aligned
#define WS_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[CPIC_NOALLOC_WORKSPACE_BYTES];

static_assert(sizeof(Geo) <= WS_PRE - WS_GEO, "Geo fits its workspace slot");
static_assert(sizeof(StepState) <= WS_TABLES - WS_STEP, "StepState fits its slot");

static pthread_mutex_t ws_lock = PTHREAD_MUTEX_INITIALIZER;
static unsigned char *ws_addr[WS_MAX_DEVICES];

// Writes the work lists and tables of one solve into the workspace: every
// entry the passes read is rewritten, so what an earlier call left there
// does not matter.  The tables are the host's expressions (cos(2 pi k/n) as
// MFT_init computes them, src/solver.c:258-280; the octant-reduced
// twiddles) evaluated with the device's cos and sin.
__global__ __launch_bounds__(CPIC_BLOCK)
void k_noalloc_setup(Geo g, int rows, int cols, int lgN, unsigned char *__restrict__ ws,
		double *__restrict__ costab, double2 *__restrict__ tw)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int stride = gridDim.x * blockDim.x;
	if(i == 0)
	{
		*(Geo *) (ws + WS_GEO) = g;
		int *pre = (int *) (ws + WS_PRE);
		pre[0] = 0;
		pre[1] = rows;
		pre[2] = 0;
		pre[3] = cols;
		StepState *st = (StepState *) (ws + WS_STEP);
		st->iter = -1;
	}
	int hx = g.nx / 2 + 1, N = 1 << lgN;
	double cx = 2.0 * M_PI / g.nx, cy = 2.0 * M_PI / g.ny;
	for(int k = i; k < hx; k += stride)
		costab[k] = cos(cx * k);
	for(int k = i; k < g.ny; k += stride)
		costab[hx + k] = cos(cy * k);
	for(int k = i; k <= N / 2; k += stride)
		twiddle(k, N, &tw[k].x, &tw[k].y);
}

// Device address of this GPU's workspace (the current device), looked up
// on the first call on each GPU.
static int
ws_get(unsigned char **out)
{
	int dev = 0, ret = 0;
	*out = NULL;
	HIPCHK(hipGetDevice(&dev));
	if(dev < 0 || dev >= WS_MAX_DEVICES)
		return CPIC_EINVAL;
	pthread_mutex_lock(&ws_lock);
	if(!ws_addr[dev])
	{
		void *p = NULL;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if(e != hipSuccess || !p)
		{
			fprintf(stderr, "cpic_hip: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
					dev, hipGetErrorString(e));
			ret = CPIC_EHIP;
		}
		else if(!(ret = fft_lds_attrs()))
			ws_addr[dev] = (unsigned char *) p;
	}
	*out = ws_addr[dev];
	pthread_mutex_unlock(&ws_lock);
	return ret;
}

static size_t
align256(size_t n)
{
	return (n + 255) & ~(size_t) 255;
}

extern "C" int
cpic_solve_xy_noalloc(int64_t nx, int64_t ny, const double *rho, double *phi, void *stream)
{
	int ret;
	if(!is_pow2(nx) || !is_pow2(ny) || nx < 2 || ny < 2 || nx > 8192 || ny > 8192 ||
			!rho || !phi)
		return CPIC_EINVAL;
	// the blocking of a batch of one (batch_create)
	Geo g;
	int nt_row, nt_col;
	memset(&g, 0, sizeof(g));
	g.nx = (int) nx;
	g.ny = (int) ny;
	g.lgnx = ilog2(nx);
	g.lgny = ilog2(ny);
	cpic_fft_threads(nx / 2, ny, &nt_row, &nt_col);
	cpic_fft_blocking(&g, nt_row, nt_col);
	if(!cpic_fft_blocking_ok(&g, nt_row, nt_col))
		return CPIC_EINVAL;
	cpic_fft_spread(&g, 1);

	int64_t hx = nx / 2 + 1, N = nx > ny ? nx : ny;
	size_t o_cos = WS_TABLES;
	size_t o_tw = align256(o_cos + sizeof(double) * (size_t) (hx + ny));
	size_t o_ghat = align256(o_tw + sizeof(double2) * (size_t) (N / 2 + 1));
	size_t need = o_ghat + sizeof(double2) * (size_t) hx * (size_t) ny;
	if(need > (size_t) CPIC_NOALLOC_WORKSPACE_BYTES)
	{
		fprintf(stderr, "cpic_solve_xy_noalloc: a %lld x %lld grid needs %zu bytes of "
				"workspace, more than the fixed %zu\n", (long long) nx, (long long) ny,
				need, (size_t) CPIC_NOALLOC_WORKSPACE_BYTES);
		return CPIC_ESIZE;
	}
	unsigned char *ws;
	if((ret = ws_get(&ws)))
		return ret;

	hipStream_t s = (hipStream_t) stream;
	int lgN = ilog2(N);
	int rows = (int) (ny / g.R), cols = (int) ((nx / 2 + g.C - 1) / g.C);
	size_t lds_rows = fft_lds_bytes((int64_t) g.R * hx);
	size_t lds_cols = fft_lds_bytes((int64_t) g.C * ny);
	const Geo *geo = (const Geo *) (ws + WS_GEO);
	const int *pre = (const int *) (ws + WS_PRE);
	const StepState *st = (const StepState *) (ws + WS_STEP);
	double *costab = (double *) (ws + o_cos);
	double2 *tw = (double2 *) (ws + o_tw);
	double2 *ghat = (double2 *) (ws + o_ghat);
	int64_t nset = hx > ny ? hx : ny;
	if(N / 2 + 1 > nset)
		nset = N / 2 + 1;
	nset = (nset + CPIC_BLOCK - 1) / CPIC_BLOCK;

	hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned) nset), dim3(CPIC_BLOCK), 0, s, g,
			rows, cols, lgN, ws, costab, tw);
	HIPCHK(hipGetLastError());
	FFT_LAUNCH(k_rows_r2c, nt_row, rows, lds_rows, s,
			geo, pre, 1, rho, ghat, (const double2 *) tw, lgN);
	HIPCHK(hipGetLastError());
	FFT_LAUNCH(k_cols, nt_col, cols, lds_cols, s,
			geo, pre + 2, 1, ghat, (const double *) costab, (const double2 *) tw, lgN,
			(DiagSlot *) NULL, 1, st, (double *) NULL);
	HIPCHK(hipGetLastError());
	FFT_LAUNCH(k_rows_c2r, nt_row, rows, lds_rows, s,
			geo, pre, 1, (const double2 *) ghat, phi, (const double2 *) tw, lgN,
			(double *) NULL, (DiagSlot *) NULL, 1, st);
	HIPCHK(hipGetLastError());
	return 0;
}
