// Host side of the differentiable path's particle order, split over host
// threads: a stable counting sort of the particles by cell with the cells of
// each block chunk taken round-robin, and the gathers and scatters through
// the resulting permutation.  Plain C++ (no HIP), so a host-only program can
// check it against a sequential restatement.
#pragma once

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

// Threads for n items, at least min_per_thread items each, at most 16 (a
// GPU's share of a host).
static inline int host_threads(int64_t n, int64_t min_per_thread)
{
	unsigned hw = std::thread::hardware_concurrency();
	int64_t t = hw == 0 ? 1 : (hw < 16 ? hw : 16);
	int64_t by_n = n / (min_per_thread > 0 ? min_per_thread : 1);
	if(by_n < t)
		t = by_n > 1 ? by_n : 1;
	return (int) t;
}

// f(t, lo, hi) for the nt ranges [n t / nt, n (t+1) / nt); a range whose
// thread cannot be started runs on the calling thread.
template <typename F>
static void parallel_ranges(int64_t n, int nt, F f)
{
	if(nt <= 1)
	{
		f(0, (int64_t) 0, n);
		return;
	}
	std::vector<std::thread> th;
	th.reserve((size_t) nt);
	for(int t = 1; t < nt; t++)
	{
		int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
		try
		{
			th.emplace_back(f, t, lo, hi);
		}
		catch(...)
		{
			f(t, lo, hi);
		}
	}
	f(0, (int64_t) 0, n / nt);
	for(auto &x : th)
		x.join();
}

// perm[j] = the particle at sorted position j: ordered by cell iy*nx + ix
// (stable), then within each run [k chunk, (k+1) chunk) the cells are taken
// round-robin (one particle of every cell that still has one, per round),
// so consecutive positions lie in different cells.  Cells are clamped to the
// grid.  Returns 0, or -1 when memory runs out.
static int sort_by_cell_rr(int64_t n, const double *x, const double *y, double idx, double idy,
		int nx, int ny, int64_t chunk, int64_t *perm)
{
	const int64_t nc = (int64_t) nx * ny;
	int nt = host_threads(n, 1 << 16);
	while(nt > 1 && (int64_t) nt * nc * (int64_t) sizeof(int64_t) > ((int64_t) 64 << 20))
		nt--;
	int64_t *cell = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n > 0 ? n : 1));
	int64_t *scell = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n > 0 ? n : 1));
	int64_t *cnt = (int64_t *) calloc((size_t) nt * (size_t) nc, sizeof(int64_t));
	if(!cell || !scell || !cnt)
	{
		free(cell);
		free(scell);
		free(cnt);
		return -1;
	}
	parallel_ranges(n, nt, [&](int t, int64_t lo, int64_t hi) {
		int64_t *c = cnt + (int64_t) t * nc;
		for(int64_t i = lo; i < hi; i++)
		{
			int64_t ix = (int64_t) (x[i] * idx), iy = (int64_t) (y[i] * idy);
			ix = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);
			iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
			cell[i] = iy * nx + ix;
			c[cell[i]]++;
		}
	});
	// exclusive offsets, cell-major then thread: thread t's particles of a
	// cell follow those of threads < t, so the scatter is stable
	int64_t off = 0;
	for(int64_t c = 0; c < nc; c++)
		for(int t = 0; t < nt; t++)
		{
			int64_t v = cnt[(int64_t) t * nc + c];
			cnt[(int64_t) t * nc + c] = off;
			off += v;
		}
	parallel_ranges(n, nt, [&](int t, int64_t lo, int64_t hi) {
		int64_t *c = cnt + (int64_t) t * nc;
		for(int64_t i = lo; i < hi; i++)
		{
			int64_t p = c[cell[i]]++;
			perm[p] = i;
			scell[p] = cell[i];
		}
	});
	free(cnt);
	free(cell);
	const int64_t nchunks = (n + chunk - 1) / chunk;
	std::atomic<int> fail(0);
	parallel_ranges(nchunks, host_threads(nchunks, 64), [&](int, int64_t k0, int64_t k1) {
		int64_t *tmp = (int64_t *) malloc(sizeof(int64_t) * 4 * (size_t) (chunk + 1));
		if(!tmp)
		{
			fail = 1;
			return;
		}
		int64_t *rs = tmp + (chunk + 1), *pos = rs + (chunk + 1), *act = pos + (chunk + 1);
		for(int64_t k = k0; k < k1; k++)
		{
			int64_t c0 = k * chunk, c1 = c0 + chunk < n ? c0 + chunk : n, nr = 0, w = 0;
			for(int64_t j = c0; j < c1; j++)
				if(j == c0 || scell[j] != scell[j - 1])
					rs[nr++] = j;
			rs[nr] = c1;
			for(int64_t q = 0; q < nr; q++)
			{
				act[q] = q;
				pos[q] = rs[q];
			}
			// one particle of every cell that still has one, per round
			for(int64_t na = nr; na > 0; )
			{
				int64_t m = 0;
				for(int64_t q = 0; q < na; q++)
				{
					int64_t r = act[q];
					tmp[w++] = perm[pos[r]++];
					if(pos[r] < rs[r + 1])
						act[m++] = r;
				}
				na = m;
			}
			memcpy(perm + c0, tmp, sizeof(int64_t) * (size_t) (c1 - c0));
		}
		free(tmp);
	});
	free(scell);
	return fail ? -1 : 0;
}

// dst[j] = src[perm[j]]
template <typename T>
static void gather_perm(int64_t n, const int64_t *perm, const T *src, T *dst)
{
	parallel_ranges(n, host_threads(n, 1 << 16), [&](int, int64_t lo, int64_t hi) {
		for(int64_t j = lo; j < hi; j++)
			dst[j] = src[perm[j]];
	});
}

// dst[perm[j]] = src[j]
template <typename T>
static void scatter_perm(int64_t n, const int64_t *perm, const T *src, T *dst)
{
	parallel_ranges(n, host_threads(n, 1 << 16), [&](int, int64_t lo, int64_t hi) {
		for(int64_t j = lo; j < hi; j++)
			dst[perm[j]] = src[j];
	});
}
