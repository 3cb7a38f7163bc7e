// Host-only check of cpic_amd/csrc/host_sort.h: the threaded stable
// counting sort by cell with the per-chunk round-robin interleave must give
// the permutation of the plain sequential restatement below, on grids and
// particle counts that cover empty input, a single particle, partial chunks,
// clamped edge positions and the 256^2 / 4,194,304-particle differentiable
// workload; the gathers and scatters through a permutation must round-trip.
#include "host_sort.h"
#include <cstdio>
#include <random>
#include <chrono>
// sequential restatement: stable counting sort, then per chunk one
// particle of every cell that still has one, per round
static void seq(int64_t n, const double *x, const double *y, double idx, double idy, int nx, int ny, int64_t CH, int64_t *perm)
{
	const int64_t nc = (int64_t) nx * ny;
	int64_t *start = (int64_t *) calloc((size_t) nc + 1, sizeof(int64_t));
	int64_t *cell = (int64_t *) malloc(sizeof(int64_t) * (size_t) (n+1));
	for(int64_t i = 0; i < n; i++) {
		int64_t ix = (int64_t) (x[i] * idx), iy = (int64_t) (y[i] * idy);
		ix = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);
		iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
		cell[i] = iy * nx + ix; start[cell[i] + 1]++; }
	for(int64_t c = 0; c < nc; c++) start[c + 1] += start[c];
	for(int64_t i = 0; i < n; i++) perm[start[cell[i]]++] = i;
	free(start);
	int64_t *tmp = (int64_t *) malloc(sizeof(int64_t) * 4 * (CH + 1));
	int64_t *rs = tmp + (CH + 1), *pos = rs + (CH + 1), *act = pos + (CH + 1);
	for(int64_t c0 = 0; c0 < n; c0 += CH) {
		int64_t c1 = c0 + CH < n ? c0 + CH : n, nr = 0, k = 0;
		for(int64_t j = c0; j < c1; j++) if(j == c0 || cell[perm[j]] != cell[perm[j - 1]]) rs[nr++] = j;
		rs[nr] = c1;
		for(int64_t q = 0; q < nr; q++) { act[q] = q; pos[q] = rs[q]; }
		for(int64_t na = nr; na > 0; ) { int64_t w = 0;
			for(int64_t q = 0; q < na; q++) { int64_t r = act[q]; tmp[k++] = perm[pos[r]++]; if(pos[r] < rs[r + 1]) act[w++] = r; }
			na = w; }
		memcpy(perm + c0, tmp, sizeof(int64_t) * (size_t) (c1 - c0)); }
	free(tmp); free(cell);
}
int main()
{
	std::mt19937_64 rng(5);
	struct C { int64_t n; int nx, ny; double L0, L1; } cs[] = {
		{0, 16, 16, 16, 16}, {1, 16, 16, 16, 16}, {1000, 8, 4, 8, 2}, {1024, 64, 64, 64, 64},
		{1025, 64, 64, 64, 64}, {262144, 64, 64, 64, 64}, {300001, 32, 128, 32, 64}, {4194304, 256, 256, 256, 256}};
	int bad = 0;
	for(auto c : cs) {
		std::vector<double> x(c.n + 1), y(c.n + 1);
		std::uniform_real_distribution<double> ux(0, c.L0), uy(0, c.L1);
		for(int64_t i = 0; i < c.n; i++) { x[i] = ux(rng); y[i] = uy(rng); }
		if(c.n > 10) { x[3] = c.L0; y[5] = c.L1; x[7] = 0.0; }   // edges clamp
		std::vector<int64_t> a(c.n + 1), b(c.n + 1);
		auto t0 = std::chrono::steady_clock::now();
		seq(c.n, x.data(), y.data(), c.nx / c.L0, c.ny / c.L1, c.nx, c.ny, 1024, a.data());
		auto t1 = std::chrono::steady_clock::now();
		int r = sort_by_cell_rr(c.n, x.data(), y.data(), c.nx / c.L0, c.ny / c.L1, c.nx, c.ny, 1024, b.data());
		auto t2 = std::chrono::steady_clock::now();
		bool same = r == 0 && std::equal(a.begin(), a.begin() + c.n, b.begin());
		bad += !same;
		printf("n %lld grid %dx%d: %s  seq %.1f ms par %.1f ms\n", (long long) c.n, c.nx, c.ny, same ? "same" : "DIFFERENT",
			std::chrono::duration<double, std::milli>(t1 - t0).count(), std::chrono::duration<double, std::milli>(t2 - t1).count());
	}
	// gathers/scatters round trip
	int64_t n = 1000003; std::vector<int64_t> p(n); for(int64_t i = 0; i < n; i++) p[i] = (i * 7919) % n;
	std::vector<double> s(n), g(n), back(n); for(int64_t i = 0; i < n; i++) s[i] = i * 0.5;
	gather_perm(n, p.data(), s.data(), g.data()); scatter_perm(n, p.data(), g.data(), back.data());
	bool rt = back == s && g[10] == s[p[10]];
	printf("gather/scatter round trip: %s\n", rt ? "ok" : "BAD");
	return bad || !rt;
}
