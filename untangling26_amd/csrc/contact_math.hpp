// Closest-point coefficients, the IPC log barrier and additive CCD, written
// once for FP32 and FP64 (T = float or double).
//
// distance.hpp:36-330 (point-edge, point-triangle and edge-edge coefficients
// with region classification), cipc_kernel.hpp:15-36 (barrier derivatives),
// accd.hpp:29-204 (additive CCD of ppf-contact-solver).  Evaluation order is
// kept identical to the CPU restatement so that FP64 results agree bit for bit
// wherever no transcendental function is involved.
#pragma once

#include "ut_common.hpp"

namespace ut
{

template <typename T>
struct v3
{
	T x, y, z;
};
template <typename T>
UT_HD v3<T> V3m(T x, T y, T z)
{
	return v3<T>{ x, y, z };
}
template <typename T>
UT_HD v3<T> operator+(v3<T> a, v3<T> b)
{
	return v3<T>{ a.x + b.x, a.y + b.y, a.z + b.z };
}
template <typename T>
UT_HD v3<T> operator-(v3<T> a, v3<T> b)
{
	return v3<T>{ a.x - b.x, a.y - b.y, a.z - b.z };
}
template <typename T>
UT_HD v3<T> operator*(T s, v3<T> a)
{
	return v3<T>{ s * a.x, s * a.y, s * a.z };
}
template <typename T>
UT_HD T dot(v3<T> a, v3<T> b)
{
	return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T>
UT_HD v3<T> cross(v3<T> a, v3<T> b)
{
	return v3<T>{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
template <typename T>
UT_HD v3<T> load_v3(const float* __restrict__ p, uint32_t i)
{
	return v3<T>{ (T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2] };
}
template <typename T>
UT_HD v3<T> load_v3(const double* __restrict__ p, uint32_t i)
{
	return v3<T>{ (T)p[3 * i], (T)p[3 * i + 1], (T)p[3 * i + 2] };
}

// 2x2 symmetric solve (distance.hpp:36-54): false when the determinant is 0.
template <typename T>
UT_HD bool solve2(T a00, T a01, T a10, T a11, T b0, T b1, T* x0, T* x1)
{
	const T det = a00 * a11 - a10 * a01;
	if (det == (T)0)
		return false;
	*x0 = (a11 / det) * b0 + (-a01 / det) * b1;
	*x1 = (-a10 / det) * b0 + (a00 / det) * b1;
	return true;
}

// point_edge_distance_coeff (distance.hpp:55-67); Epsilon = FLT_EPSILON (Core/constant_value.h:12)
template <typename T>
UT_HD void pe_coeff(v3<T> p, v3<T> e0, v3<T> e1, T* c0, T* c1)
{
	const v3<T> r = e1 - e0;
	const T		d = dot(r, r);
	*c0 = (T)0.5;
	*c1 = (T)0.5;
	if (d > (T)1.19209290E-7)
	{
		const T x = dot(r, p - e0) / d;
		*c0 = (T)1 - x;
		*c1 = x;
	}
}

// point_triangle_distance_coeff_unclassified (distance.hpp:69-160)
template <typename T>
UT_HD v3<T> pt_coeff(v3<T> p, v3<T> t0, v3<T> t1, v3<T> t2)
{
	const v3<T> r0 = t1 - t0, r1 = t2 - t0, q = p - t0;
	T			c0, c1;
	if (!solve2(dot(r0, r0), dot(r0, r1), dot(r1, r0), dot(r1, r1), dot(r0, q), dot(r1, q), &c0, &c1))
	{
		c0 = (T)1 / (T)3;
		c1 = (T)1 / (T)3;
	}
	const v3<T> c{ (T)1 - c0 - c1, c0, c1 };
	const bool	inside = c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x <= 1 && c.y <= 1 && c.z <= 1;
	if (inside)
		return c;
	T a, b;
	if (c.x < (T)0)
	{
		pe_coeff(p, t1, t2, &a, &b);
		if (a >= 0 && a <= 1)
			return v3<T>{ 0, a, b };
		return a > 1 ? v3<T>{ 0, 1, 0 } : v3<T>{ 0, 0, 1 };
	}
	if (c.y < (T)0)
	{
		pe_coeff(p, t0, t2, &a, &b);
		if (a >= 0 && a <= 1)
			return v3<T>{ a, 0, b };
		return a > 1 ? v3<T>{ 1, 0, 0 } : v3<T>{ 0, 0, 1 };
	}
	pe_coeff(p, t0, t1, &a, &b);
	if (a >= 0 && a <= 1)
		return v3<T>{ a, b, 0 };
	return a > 1 ? v3<T>{ 1, 0, 0 } : v3<T>{ 0, 1, 0 };
}

template <typename T>
struct c4
{
	T a, b, c, d;
};

template <typename T>
UT_HD T clamp01(T v)
{
	return v < (T)0 ? (T)0 : (v > (T)1 ? (T)1 : v);
}

// edge_edge_distance_coeff_unclassified (distance.hpp:255-310)
template <typename T>
UT_HD c4<T> ee_coeff(v3<T> a0, v3<T> a1, v3<T> b0, v3<T> b1)
{
	const v3<T> r0 = a1 - a0, r1 = b1 - b0;
	const T		len0 = dot(r0, r0), len1 = dot(r1, r1);
	const v3<T> cr = cross(r0, r1);
	const T		parallel = dot(cr, cr) / (len0 * len1 + (T)1e-8);
	// edge_edge_distance_coeff: A = [r0, -r1], solve A^T A x = A^T (b0 - a0)
	const v3<T> nr1 = V3m(-r1.x, -r1.y, -r1.z);
	const v3<T> q = b0 - a0;
	T			x0, x1;
	c4<T>		res{ (T)0.5, (T)0.5, (T)0.5, (T)0.5 };
	if (solve2(dot(r0, r0), dot(r0, nr1), dot(nr1, r0), dot(nr1, nr1), dot(r0, q), dot(nr1, q), &x0, &x1))
		res = c4<T>{ (T)1 - x0, x0, (T)1 - x1, x1 };
	const bool valid = res.a >= 0 && res.b >= 0 && res.c >= 0 && res.d >= 0 && res.a <= 1 && res.b <= 1
		&& res.c <= 1 && res.d <= 1;
	if (parallel < (T)1e-4 || !valid)
	{
		T c1a, c1b, c2a, c2b, c3a, c3b, c4a, c4b;
		pe_coeff(a0, b0, b1, &c1a, &c1b);
		pe_coeff(a1, b0, b1, &c2a, &c2b);
		pe_coeff(b0, a0, a1, &c3a, &c3b);
		pe_coeff(b1, a0, a1, &c4a, &c4b);
		const c4<T> cand[4] = { { 1, 0, clamp01(c1a), clamp01(c1b) }, { 0, 1, clamp01(c2a), clamp01(c2b) },
			{ clamp01(c3a), clamp01(c3b), 1, 0 }, { clamp01(c4a), clamp01(c4b), 0, 1 } };
		T	  best = (T)3.402823466e+38;
		int	  idx = 0;
		for (int i = 0; i < 4; ++i)
		{
			const v3<T> xa = cand[i].a * a0 + cand[i].b * a1;
			const v3<T> xb = cand[i].c * b0 + cand[i].d * b1;
			const v3<T> d = xa - xb;
			const T		dd = dot(d, d);
			if (dd < best)
			{
				best = dd;
				idx = i;
			}
		}
		res = cand[idx];
	}
	return res;
}

template <typename T>
UT_HD T pt_dist2(v3<T> p, v3<T> t0, v3<T> t1, v3<T> t2)
{
	const v3<T> c = pt_coeff(p, t0, t1, t2);
	const v3<T> x = c.x * (t0 - p) + c.y * (t1 - p) + c.z * (t2 - p);
	return dot(x, x);
}
template <typename T>
UT_HD T ee_dist2(v3<T> a0, v3<T> a1, v3<T> b0, v3<T> b1)
{
	const c4<T> c = ee_coeff(a0, a1, b0, b1);
	const v3<T> xa = c.a * a0 + c.b * a1;
	const v3<T> xb = c.c * b0 + c.d * b1;
	const v3<T> d = xb - xa;
	return dot(d, d);
}

// Log barrier b(d) = -(d - dhat)^2 ln(d / dhat) and derivatives, cipc_kernel.hpp:15-36.
template <typename T>
UT_HD T barrier(T d, T dhat)
{
	const T m = d - dhat;
	return -m * m * log(d / dhat);
}
template <typename T>
UT_HD T barrier_d1(T d, T dhat)
{
	return (dhat - d) * ((T)2 * log(d / dhat) - dhat / d + (T)1);
}
template <typename T>
UT_HD T barrier_d2(T d, T dhat)
{
	const T r = dhat / d;
	return (r + (T)2) * r - (T)2 * log(d / dhat) - (T)3;
}

// ---- additive CCD (accd.hpp:29-204) --------------------------------------------
constexpr double kCcdReduction = 0.01;
constexpr double kLineSearchMaxT = 1.0;
constexpr int	 kCcdMaxIter = 10000;

template <typename T>
UT_HD void centerize(v3<T>* x)
{
	v3<T> m{ 0, 0, 0 };
	for (int k = 0; k < 4; ++k)
		m = m + (T)0.25 * x[k];
	for (int k = 0; k < 4; ++k)
		x[k] = x[k] - m;
}
template <typename T>
UT_HD T max_relative_u(const v3<T>* u)
{
	T m = 0;
	for (int i = 0; i < 4; ++i)
		for (int j = i + 1; j < 4; ++j)
		{
			const v3<T> d = u[i] - u[j];
			const T		l = dot(d, d);
			m = m > l ? m : l;
		}
	return sqrt(m);
}

// is_ee selects the edge-edge or point-triangle squared distance.
template <typename T>
UT_HD T ccd_dist2(bool is_ee, const v3<T>* x)
{
	return is_ee ? ee_dist2(x[0], x[1], x[2], x[3]) : pt_dist2(x[0], x[1], x[2], x[3]);
}

// ccd_helper (accd.hpp:48-110): conservative advancement.  Returns the
// sentinel 0.001 when the iteration cap is hit, as the reference does.
//
// gmin (device only, may be null): the running minimum over the pairs of the
// post-processed TOI (toi_postprocess).  The advanced toi only grows, and the
// post-processed value of a pair whose toi has reached r is at least
// max(r, 3e-3) (0.5 below 3e-3, 1 for the sentinel), so once that bound
// reaches gmin the pair cannot lower the minimum and returns max_t, which the
// post-processing keeps at 1.  The minimum is unchanged, only the work is cut.
template <typename T>
UT_HD T accd_toi(bool is_ee, v3<T>* x0, v3<T>* dx, T offset, const float* gmin = nullptr)
{
	centerize(x0);
	centerize(dx);
	const T u_max = max_relative_u(dx);
	if (u_max == (T)0)
		return (T)kLineSearchMaxT;
	T		toi = 0;
	const T max_t = (T)kLineSearchMaxT;
	const T eps = (T)kCcdReduction * (sqrt(ccd_dist2(is_ee, x0)) - offset);
	const T target = eps + offset;
	const T eps_sqr = eps * eps;
	const T inv_u_max = (T)1 / u_max;
	for (int iter = 0;; ++iter)
	{
		if (iter > kCcdMaxIter)
			return (T)0.001;
#if defined(__HIP_DEVICE_COMPILE__)
		if (gmin && (iter & 15) == 15)
		{
			const float g = __hip_atomic_load(gmin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if ((T)g <= (toi > (T)3e-3 ? toi : (T)3e-3))
				return max_t;
		}
#endif
		v3<T> xt[4];
		for (int k = 0; k < 4; ++k)
			xt[k] = x0[k] + toi * dx[k];
		const T d2 = ccd_dist2(is_ee, xt);
		const T dmt = (d2 - target * target) / (sqrt(d2) + target);
		if ((max_t - toi) * u_max < dmt - eps)
			return max_t;
		if (toi > (T)0 && dmt * dmt < eps_sqr)
			return toi;
		const T next = toi + dmt * inv_u_max;
		if (next == toi)
			return toi;
		toi = next;
		if (toi > max_t)
			return max_t;
	}
}

// TOI post-processing of fn_narrow_phase_{vf,ee}_ccd_query (narrow_phase.cpp:473-522).
template <typename T>
UT_HD T toi_postprocess(T toi)
{
	if (toi < (T)0 || toi > (T)kLineSearchMaxT || toi == (T)0.001)
		toi = (T)1;
	if (toi < (T)3e-3)
		toi = (T)0.5;
	return toi;
}

} // namespace ut
