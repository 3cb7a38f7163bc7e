// Per-particle arithmetic shared by the production push (particles.hip) and
// the differentiable FP64 path (adjoint.hip), so both perform the same
// operations in the same order.  Compiled with -ffp-contract=off: the
// multiply-adds the reference's clang build contracts are explicit fma().
#pragma once

#include "dev.h"

// A pointer into device memory as a global-address-space pointer.
template <typename T>
static __device__ __forceinline__ GLB_AS T *gptr(const void *p)
{
	return (GLB_AS T *) (T *) p;
}

// FP64 atomic adds into LDS and into global memory.
static __device__ __forceinline__ void lds_add(LDS_AS double *p, double v)
{
	__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

static __device__ __forceinline__ void glb_add(GLB_AS double *p, double v)
{
	[[clang::atomic(ignore_denormal_mode)]]
	{
		__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
}

// FP64 cell index and relative offset (src/interpolate.c:39-66).  The Y
// middle term uses dx (the reference passes dx[X], src/interpolate.c:88).
static __device__ __forceinline__ void cell64(const Geo &g, double x, double y,
		int &ix, int &iy, double &rx, double &ry)
{
	double bx = x * g.idx;
	double sx = floor(bx);
	double gx = fma(-sx, g.dx, x);
	rx = gx * g.idx;
	ix = (int) sx;
	if(ix >= g.nx)
		ix -= g.nx;
	double by = y * g.idy;
	double sy = floor(by);
	double gy = fma(-sy, g.dx, y);
	ry = gy * g.idy;
	iy = (int) sy;
	if(iy >= g.ny)
		iy -= g.ny;
}

// periodic_boundary_ppack with the x == L -> 0 rule.  A displacement larger
// than L (only possible after |u| > umax, which is flagged) is reduced with
// floor so that every index stays inside the grid.
static __device__ __forceinline__ double wrap64(double r, double L)
{
	if(r >= L)
		r -= L;
	else if(r < 0.0)
	{
		r += L;
		if(r >= L)
			r = 0.0;
	}
	if(!(r >= 0.0 && r < L))
	{
		r = r - floor(r / L) * L;
		if(!(r >= 0.0 && r < L))
			r = 0.0;
	}
	return r;
}

// boris_rotation + the velocity half of move (src/mover.c:23-83) for
// B = (0,0,Bz): t = (0,0,tz), s = (0,0,sz) with tz = Bz*h, sz = 2tz/(1+tz^2),
// h = dtqm2 = 0.5*dt'*q/m.  u_z stays 0, so the 3D cross products reduce to
// the 2D forms below.  As a linear map: u' = M (u + hE) + hE with
// M = [[1 - tz sz, sz], [-sz, 1 - tz sz]].
template <typename T>
static __device__ __forceinline__ void boris_bz(T h, T tz, T sz, T ex, T ey, T &ux, T &uy)
{
	T vmx = fma(h, ex, ux);
	T vmy = fma(h, ey, uy);
	T t1 = vmy * tz;
	T vpx = t1 + vmx;
	T t2 = vmx * tz;
	T vpy = vmy - t2;
	T t3 = vpy * sz;
	T vplx = t3 + vmx;
	T t4 = vpx * sz;
	T vply = vmy - t4;
	ux = fma(h, ex, vplx);
	uy = fma(h, ey, vply);
}
