// HBM streaming ceiling of the device, for the measurement next to the
// push's achieved bandwidth: a kernel that reads one buffer and writes
// another, 16 B per element (one dwordx4 load and one dwordx4 store), with
// U loads in flight per thread before its first store -- the access shape
// of the push's particle stream (state read once, written once) without its
// arithmetic, LDS work or atomics.  Not part of the PIC library's C ABI:
// built as lib/libcpic_probe.so and loaded only by bench.py.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define PROBE_BLOCK 256

template <int U>
__global__ __launch_bounds__(PROBE_BLOCK)
void k_stream(const float4 *__restrict__ a, float4 *__restrict__ b, int64_t n)
{
	int64_t base = (int64_t) blockIdx.x * U * PROBE_BLOCK + threadIdx.x;
	float4 v[U];
#pragma unroll
	for(int u = 0; u < U; u++)
	{
		int64_t i = base + (int64_t) u * PROBE_BLOCK;
		if(i < n)
			v[u] = a[i];
	}
#pragma unroll
	for(int u = 0; u < U; u++)
	{
		int64_t i = base + (int64_t) u * PROBE_BLOCK;
		if(i < n)
			b[i] = v[u];
	}
}

template <int U>
static void launch(const float4 *a, float4 *b, int64_t n, hipStream_t s)
{
	int64_t blocks = (n + (int64_t) U * PROBE_BLOCK - 1) / ((int64_t) U * PROBE_BLOCK);
	hipLaunchKernelGGL(k_stream<U>, dim3((unsigned) blocks), dim3(PROBE_BLOCK), 0, s, a, b, n);
}

// The push's own access pattern without its arithmetic, LDS work or
// atomics: FP32 particle state in groups of 64 slots (x[64] y[64] ux[64]
// uy[64], 1 KiB per group), 1024 slots per workgroup.  Thread t loads slots
// 4t..4t+3 of its chunk with one dwordx4 load per component, then writes
// its u-th particle to output slot u*256 + t with one dword store per
// component (consecutive lanes write consecutive slots, as the push's
// stayer stores do).
template <bool VSTORE>
__global__ __launch_bounds__(PROBE_BLOCK)
void k_stream_groups(const float *__restrict__ a, float *__restrict__ b, int64_t nslots)
{
	int64_t base = (int64_t) blockIdx.x * 1024;
	int t = threadIdx.x;
	int64_t j0 = base + 4 * t;
	if(j0 + 4 > nslots)
		return;
	// element of slot s, component c: (s >> 6) * 256 + c * 64 + (s & 63)
	const float4 *r = (const float4 *) (a + (j0 >> 6) * 256 + (j0 & 63));
	float4 x = r[0], y = r[16], ux = r[32], uy = r[48];
	if(VSTORE)
	{
		// the same 4 slots written back with one dwordx4 store per component
		float4 *w = (float4 *) (b + (j0 >> 6) * 256 + (j0 & 63));
		w[0] = x;
		w[16] = y;
		w[32] = ux;
		w[48] = uy;
		return;
	}
	float vx[4] = {x.x, x.y, x.z, x.w}, vy[4] = {y.x, y.y, y.z, y.w};
	float vux[4] = {ux.x, ux.y, ux.z, ux.w}, vuy[4] = {uy.x, uy.y, uy.z, uy.w};
#pragma unroll
	for(int u = 0; u < 4; u++)
	{
		int64_t s = base + u * PROBE_BLOCK + t;
		float *w = b + (s >> 6) * 256 + (s & 63);
		w[0] = vx[u];
		w[64] = vy[u];
		w[128] = vux[u];
		w[192] = vuy[u];
	}
}

// The same pattern walked the way k_push walks its buckets: each workgroup
// streams ts slots (one bucket's particles) in chunks of 1024, and
// workgroup b takes bucket (b & 7) * (n / 8) + (b >> 3), contiguous bands
// per XCD as k_push's order does.  Streaming 8 chunks in sequence per
// workgroup instead of one costs about 8 % by itself, 4 chunks about 4 %.
__global__ __launch_bounds__(PROBE_BLOCK)
void k_stream_tiles(const float *__restrict__ a, float *__restrict__ b, int ts, int ntiles)
{
	int tile = blockIdx.x;
	if((ntiles & 7) == 0)
		tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
	const int t = threadIdx.x;
	const int64_t t0 = (int64_t) tile * ts;
	for(int c = 0; c < ts; c += 1024)
	{
		int64_t base = t0 + c, j0 = base + 4 * t;
		const float4 *r = (const float4 *) (a + (j0 >> 6) * 256 + (j0 & 63));
		float4 x = r[0], y = r[16], ux = r[32], uy = r[48];
		float vx[4] = {x.x, x.y, x.z, x.w}, vy[4] = {y.x, y.y, y.z, y.w};
		float vux[4] = {ux.x, ux.y, ux.z, ux.w}, vuy[4] = {uy.x, uy.y, uy.z, uy.w};
#pragma unroll
		for(int u = 0; u < 4; u++)
		{
			int64_t sl = base + u * PROBE_BLOCK + t;
			float *w = b + (sl >> 6) * 256 + (sl & 63);
			w[0] = vx[u];
			w[64] = vy[u];
			w[128] = vux[u];
			w[192] = vuy[u];
		}
	}
}

// Average milliseconds per launch of k_stream_tiles over two buffers of
// `bytes` bytes each, whole tiles of tile_slots slots (a multiple of 1024).
extern "C" int
cpic_probe_stream_tiles(int64_t bytes, int tile_slots, int reps, double *ms_avg)
{
	if(tile_slots < 1024 || tile_slots % 1024 || reps < 1 || !ms_avg)
		return -1;
	int64_t ntiles = bytes / 16 / tile_slots;
	if(ntiles < 1 || ntiles > (1 << 30))
		return -1;
	int64_t nslots = ntiles * tile_slots;
	float *a = NULL, *b = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	float ms = 0.0f;
	int ret = -1;
	if(hipMalloc(&a, 16 * (size_t) nslots) != hipSuccess ||
			hipMalloc(&b, 16 * (size_t) nslots) != hipSuccess ||
			hipMemset(a, 1, 16 * (size_t) nslots) != hipSuccess ||
			hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
			hipEventCreate(&e1) != hipSuccess)
		goto out;
	for(int r = 0; r <= reps; r++)
	{
		if(r == 1 && hipEventRecord(e0, s) != hipSuccess)
			goto out;
		hipLaunchKernelGGL(k_stream_tiles, dim3((unsigned) ntiles), dim3(PROBE_BLOCK), 0, s, a, b,
				tile_slots, (int) ntiles);
	}
	if(hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
			hipEventSynchronize(e1) != hipSuccess ||
			hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
		goto out;
	*ms_avg = (double) ms / reps;
	ret = 0;
out:
	if(ret)
		fprintf(stderr, "cpic_probe_stream_tiles: HIP error: %s\n",
				hipGetErrorString(hipGetLastError()));
	if(e0) hipEventDestroy(e0);
	if(e1) hipEventDestroy(e1);
	if(s) hipStreamDestroy(s);
	hipFree(a);
	hipFree(b);
	return ret;
}

// Average milliseconds per launch of k_stream_groups over two buffers of
// `bytes` bytes each (whole chunks of 1024 slots; 16 B per slot), as for
// cpic_probe_stream; vstore = 1 writes with dwordx4 stores instead of the
// push's per-slot stores.
extern "C" int
cpic_probe_stream_groups(int64_t bytes, int vstore, int reps, double *ms_avg)
{
	if(bytes < 16384 || reps < 1 || !ms_avg)
		return -1;
	int64_t nslots = bytes / 16 / 1024 * 1024;
	float *a = NULL, *b = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	float ms = 0.0f;
	int ret = -1;
	if(hipMalloc(&a, 16 * (size_t) nslots) != hipSuccess ||
			hipMalloc(&b, 16 * (size_t) nslots) != hipSuccess ||
			hipMemset(a, 1, 16 * (size_t) nslots) != hipSuccess ||
			hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
			hipEventCreate(&e1) != hipSuccess)
		goto out;
	for(int r = 0; r <= reps; r++)
	{
		if(r == 1 && hipEventRecord(e0, s) != hipSuccess)
			goto out;
		if(vstore)
			hipLaunchKernelGGL(k_stream_groups<true>, dim3((unsigned) (nslots / 1024)),
					dim3(PROBE_BLOCK), 0, s, a, b, nslots);
		else
			hipLaunchKernelGGL(k_stream_groups<false>, dim3((unsigned) (nslots / 1024)),
					dim3(PROBE_BLOCK), 0, s, a, b, nslots);
	}
	if(hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
			hipEventSynchronize(e1) != hipSuccess ||
			hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
		goto out;
	*ms_avg = (double) ms / reps;
	ret = 0;
out:
	if(ret)
		fprintf(stderr, "cpic_probe_stream_groups: HIP error: %s\n",
				hipGetErrorString(hipGetLastError()));
	if(e0) hipEventDestroy(e0);
	if(e1) hipEventDestroy(e1);
	if(s) hipStreamDestroy(s);
	hipFree(a);
	hipFree(b);
	return ret;
}

// Average milliseconds per launch over `reps` launches (after one warm-up)
// of the U = `unroll` kernel over two buffers of `bytes` bytes each (HIP
// events on a stream of its own).  Returns 0, or -1 on a bad argument or a
// HIP error (named on stderr).
extern "C" int
cpic_probe_stream(int64_t bytes, int unroll, int reps, double *ms_avg)
{
	if(bytes < 16 || reps < 1 || !ms_avg || (unroll != 1 && unroll != 2 && unroll != 4 &&
				unroll != 8))
		return -1;
	int64_t n = bytes / 16;
	float4 *a = NULL, *b = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	float ms = 0.0f;
	int ret = -1;
	if(hipMalloc(&a, 16 * (size_t) n) != hipSuccess || hipMalloc(&b, 16 * (size_t) n) != hipSuccess ||
			hipMemset(a, 1, 16 * (size_t) n) != hipSuccess ||
			hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
			hipEventCreate(&e1) != hipSuccess)
		goto out;
	for(int r = 0; r <= reps; r++)
	{
		if(r == 1 && hipEventRecord(e0, s) != hipSuccess)
			goto out;
		switch(unroll)
		{
			case 1: launch<1>(a, b, n, s); break;
			case 2: launch<2>(a, b, n, s); break;
			case 4: launch<4>(a, b, n, s); break;
			default: launch<8>(a, b, n, s); break;
		}
	}
	if(hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
			hipEventSynchronize(e1) != hipSuccess ||
			hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
		goto out;
	*ms_avg = (double) ms / reps;
	ret = 0;
out:
	if(ret)
		fprintf(stderr, "cpic_probe_stream: HIP error: %s\n", hipGetErrorString(hipGetLastError()));
	if(e0) hipEventDestroy(e0);
	if(e1) hipEventDestroy(e1);
	if(s) hipStreamDestroy(s);
	hipFree(a);
	hipFree(b);
	return ret;
}
