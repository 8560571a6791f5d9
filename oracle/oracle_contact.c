/*
 * oracle_contact.c -- FP64 restatement of the proximity DCD (VF/EE with the
 * IPC barrier) and of additive CCD.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows distance.hpp:36-330 (closest-point coefficients), cipc_kernel.hpp:15-36
 * (barrier), narrow_phase.cpp:980-1346 (DCD pair test, stiffness) and
 * accd.hpp:29-204 + narrow_phase.cpp:412-800 (CCD, TOI post-processing).
 *
 * Pinning (tests/test_oracle.py):
 *  - distances: closed forms (point over a triangle interior, beyond an edge,
 *    beyond a vertex; skew and parallel segments) and brute-force minimisation
 *    over a fine barycentric grid;
 *  - barrier: b'(d), b''(d) against FP64 central differences of b(d) (h = 1e-7
 *    relative) and the closed form b(dhat) = 0, b'(dhat) = 0;
 *  - ACCD: a point moving at constant velocity towards a plane has the
 *    analytic contact time (d0 - offset)/v; ACCD returns a time <= it, and the
 *    gap at the returned time is >= offset; zero relative motion returns 1.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

typedef struct
{
	double x, y, z;
} d3;
static d3 D3(double x, double y, double z)
{
	d3 r = { x, y, z };
	return r;
}
static d3 ld(const float* p, uint32_t i) { return D3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static d3 add(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 scl(double s, d3 a) { return D3(s * a.x, s * a.y, s * a.z); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 crs(d3 a, d3 b) { return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

static int solve2(double a00, double a01, double a10, double a11, double b0, double b1, double* x0, double* x1)
{
	const double det = a00 * a11 - a10 * a01;
	if (det == 0.0)
		return 0;
	*x0 = (a11 / det) * b0 + (-a01 / det) * b1;
	*x1 = (-a10 / det) * b0 + (a00 / det) * b1;
	return 1;
}

static void pe_coeff(d3 p, d3 e0, d3 e1, double* c0, double* c1)
{
	d3	   r = sub(e1, e0);
	double d = dot(r, r);
	*c0 = 0.5, *c1 = 0.5;
	if (d > 1.19209290E-7)
	{
		double x = dot(r, sub(p, e0)) / d;
		*c0 = 1.0 - x, *c1 = x;
	}
}

static d3 pt_coeff(d3 p, d3 t0, d3 t1, d3 t2)
{
	d3	   r0 = sub(t1, t0), r1 = sub(t2, t0), q = sub(p, t0);
	double c0, c1, a, b;
	if (!solve2(dot(r0, r0), dot(r0, r1), dot(r1, r0), dot(r1, r1), dot(r0, q), dot(r1, q), &c0, &c1))
		c0 = 1.0 / 3.0, c1 = 1.0 / 3.0;
	d3 c = D3(1.0 - c0 - c1, c0, c1);
	if (c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x <= 1 && c.y <= 1 && c.z <= 1)
		return c;
	if (c.x < 0)
	{
		pe_coeff(p, t1, t2, &a, &b);
		if (a >= 0 && a <= 1)
			return D3(0, a, b);
		return a > 1 ? D3(0, 1, 0) : D3(0, 0, 1);
	}
	if (c.y < 0)
	{
		pe_coeff(p, t0, t2, &a, &b);
		if (a >= 0 && a <= 1)
			return D3(a, 0, b);
		return a > 1 ? D3(1, 0, 0) : D3(0, 0, 1);
	}
	pe_coeff(p, t0, t1, &a, &b);
	if (a >= 0 && a <= 1)
		return D3(a, b, 0);
	return a > 1 ? D3(1, 0, 0) : D3(0, 1, 0);
}

static double clamp01(double v) { return v < 0 ? 0 : (v > 1 ? 1 : v); }

static void ee_coeff(d3 a0, d3 a1, d3 b0, d3 b1, double* c)
{
	d3	   r0 = sub(a1, a0), r1 = sub(b1, b0);
	double len0 = dot(r0, r0), len1 = dot(r1, r1);
	d3	   cr = crs(r0, r1);
	double par = dot(cr, cr) / (len0 * len1 + 1e-8);
	d3	   nr1 = D3(-r1.x, -r1.y, -r1.z), q = sub(b0, a0);
	double x0, x1;
	c[0] = c[1] = c[2] = c[3] = 0.5;
	if (solve2(dot(r0, r0), dot(r0, nr1), dot(nr1, r0), dot(nr1, nr1), dot(r0, q), dot(nr1, q), &x0, &x1))
		c[0] = 1.0 - x0, c[1] = x0, c[2] = 1.0 - x1, c[3] = x1;
	int valid = 1;
	for (int i = 0; i < 4; ++i)
		if (!(c[i] >= 0 && c[i] <= 1))
			valid = 0;
	if (par < 1e-4 || !valid)
	{
		double p1[2], p2[2], p3[2], p4[2];
		pe_coeff(a0, b0, b1, &p1[0], &p1[1]);
		pe_coeff(a1, b0, b1, &p2[0], &p2[1]);
		pe_coeff(b0, a0, a1, &p3[0], &p3[1]);
		pe_coeff(b1, a0, a1, &p4[0], &p4[1]);
		double cand[4][4] = { { 1, 0, clamp01(p1[0]), clamp01(p1[1]) }, { 0, 1, clamp01(p2[0]), clamp01(p2[1]) },
			{ clamp01(p3[0]), clamp01(p3[1]), 1, 0 }, { clamp01(p4[0]), clamp01(p4[1]), 0, 1 } };
		double best = 3.402823466e+38;
		int	   idx = 0;
		for (int i = 0; i < 4; ++i)
		{
			d3	   xa = add(scl(cand[i][0], a0), scl(cand[i][1], a1));
			d3	   xb = add(scl(cand[i][2], b0), scl(cand[i][3], b1));
			d3	   d = sub(xa, xb);
			double dd = dot(d, d);
			if (dd < best)
				best = dd, idx = i;
		}
		memcpy(c, cand[idx], sizeof(double) * 4);
	}
}

static double pt_dist2(d3 p, d3 t0, d3 t1, d3 t2)
{
	d3 c = pt_coeff(p, t0, t1, t2);
	d3 x = add(add(scl(c.x, sub(t0, p)), scl(c.y, sub(t1, p))), scl(c.z, sub(t2, p)));
	return dot(x, x);
}
static double ee_dist2(d3 a0, d3 a1, d3 b0, d3 b1)
{
	double c[4];
	ee_coeff(a0, a1, b0, b1, c);
	d3 xa = add(scl(c[0], a0), scl(c[1], a1));
	d3 xb = add(scl(c[2], b0), scl(c[3], b1));
	d3 d = sub(xb, xa);
	return dot(d, d);
}

/* ---- exported distance / barrier helpers (for pinning tests) ---- */
void or_pt_coeff(const double* p, const double* t0, const double* t1, const double* t2, double* out)
{
	d3 c = pt_coeff(D3(p[0], p[1], p[2]), D3(t0[0], t0[1], t0[2]), D3(t1[0], t1[1], t1[2]), D3(t2[0], t2[1], t2[2]));
	out[0] = c.x, out[1] = c.y, out[2] = c.z;
}
void or_ee_coeff(const double* a0, const double* a1, const double* b0, const double* b1, double* out)
{
	ee_coeff(D3(a0[0], a0[1], a0[2]), D3(a1[0], a1[1], a1[2]), D3(b0[0], b0[1], b0[2]), D3(b1[0], b1[1], b1[2]), out);
}
double or_barrier(double d, double dhat, int deriv)
{
	if (deriv == 0)
	{
		double m = d - dhat;
		return -m * m * log(d / dhat);
	}
	if (deriv == 1)
		return (dhat - d) * (2.0 * log(d / dhat) - dhat / d + 1.0);
	double r = dhat / d;
	return (r + 2.0) * r - 2.0 * log(d / dhat) - 3.0;
}

/* ---- DCD pair tests (narrow_phase.cpp:980-1346) ---- */
static int invalid_combo(const uint32_t* vprop, uint32_t a, uint32_t b)
{
	if (!vprop)
		return 0;
	uint32_t pa = vprop[a], pb = vprop[b];
	return ((pa & 1u) && (pb & 1u)) || ((pa & 2u) && (pb & 2u) && (pa >> 8) == (pb >> 8));
}

static int finish(double d2, double th, double d_hat, int energy, double area, double kappa, d3 diff, uint32_t* rec)
{
	const double kMin = 5e-6, kMinSqr = kMin * kMin;
	if (!(d2 < (th + d_hat) * (th + d_hat) && (d2 > kMinSqr || energy == 0)))
		return 0;
	double d = sqrt(d2);
	d3	   n = D3(diff.x / d, diff.y / d, diff.z / d);
	if (d < kMinSqr)
		n = D3(0, 1, 0);
	double k1, k2;
	if (energy == 0)
	{
		double C = (d - th) - d_hat, stiff = area * kappa;
		k1 = stiff * C, k2 = stiff;
	}
	else
	{
		k1 = area * kappa * or_barrier(d - th, d_hat, 1);
		k2 = area * kappa * or_barrier(d - th, d_hat, 2);
	}
	if (!isfinite(k1) || !isfinite(k2))
		return 0;
	float f[8] = { (float)n.x, (float)n.y, (float)n.z, (float)area, (float)k1, (float)k2, 0, 0 };
	memcpy(rec + 4, f, 24); /* normal, area, k1, k2; barycentrics filled by the caller */
	return 1;
}

/* rec: 16 words of CollisionPairTemplate. Returns 1 when emitted. */
int or_vf_dcd(uint32_t vid, uint32_t fid, const float* x, const uint32_t* faces, const uint32_t* vprop,
	const float* dh, const float* off, const float* rva, const float* rfa, double kappa, int energy, uint32_t* rec)
{
	uint32_t f0 = faces[3 * fid], f1 = faces[3 * fid + 1], f2 = faces[3 * fid + 2];
	if (vid == f0 || vid == f1 || vid == f2 || invalid_combo(vprop, vid, f0))
		return 0;
	d3	   p = ld(x, vid), t0 = ld(x, f0), t1 = ld(x, f1), t2 = ld(x, f2);
	d3	   b = pt_coeff(p, t0, t1, t2);
	d3	   diff = add(add(scl(b.x, sub(p, t0)), scl(b.y, sub(p, t1))), scl(b.z, sub(p, t2)));
	double d2 = dot(diff, diff);
	double d_hat = ((double)dh[vid] + (double)dh[f0]) * 0.5;
	double th = (double)off[vid] + (double)off[f0];
	double area = 0.5 * ((double)rva[vid] + (double)rfa[fid]);
	memset(rec, 0, 64);
	if (!finish(d2, th, d_hat, energy, area, kappa, diff, rec))
		return 0;
	rec[0] = vid, rec[1] = f0, rec[2] = f1 | 0x80000000u, rec[3] = f2;
	float bb[2] = { (float)b.x, (float)b.y };
	memcpy(rec + 10, bb, 8);
	return 1;
}

int or_ee_dcd(uint32_t left, uint32_t right, const float* x, const uint32_t* edges, const uint32_t* vprop,
	const float* dh, const float* off, const float* rea, double kappa, int energy, uint32_t* rec)
{
	uint32_t l0 = edges[2 * left], l1 = edges[2 * left + 1], r0 = edges[2 * right], r1 = edges[2 * right + 1];
	if (l0 == r0 || l0 == r1 || l1 == r0 || l1 == r1 || invalid_combo(vprop, l0, r0))
		return 0;
	d3	   a0 = ld(x, l0), a1 = ld(x, l1), b0 = ld(x, r0), b1 = ld(x, r1);
	double c[4];
	ee_coeff(a0, a1, b0, b1, c);
	d3	   x0 = add(scl(c[0], a0), scl(c[1], a1));
	d3	   x1 = add(scl(c[2], b0), scl(c[3], b1));
	d3	   diff = sub(x0, x1);
	double d2 = dot(diff, diff);
	double d_hat = ((double)dh[l0] + (double)dh[r0]) * 0.5;
	double th = (double)off[l0] + (double)off[r0];
	double area = 0.5 * ((double)rea[left] + (double)rea[right]);
	memset(rec, 0, 64);
	if (!finish(d2, th, d_hat, energy, area, kappa, diff, rec))
		return 0;
	rec[0] = l0, rec[1] = l1, rec[2] = r0, rec[3] = r1 | 0x80000000u;
	float bb[2] = { (float)c[0], (float)c[2] };
	memcpy(rec + 10, bb, 8);
	return 1;
}

/* ---- ACCD (accd.hpp:29-204) ---- */
static double ccd_d2(int is_ee, const d3* x) { return is_ee ? ee_dist2(x[0], x[1], x[2], x[3]) : pt_dist2(x[0], x[1], x[2], x[3]); }

double or_accd(int is_ee, const double* x0_in, const double* dx_in, double offset)
{
	d3 x0[4], dx[4], m0 = D3(0, 0, 0), m1 = D3(0, 0, 0);
	for (int k = 0; k < 4; ++k)
	{
		x0[k] = D3(x0_in[3 * k], x0_in[3 * k + 1], x0_in[3 * k + 2]);
		dx[k] = D3(dx_in[3 * k], dx_in[3 * k + 1], dx_in[3 * k + 2]);
	}
	for (int k = 0; k < 4; ++k)
		m0 = add(m0, scl(0.25, x0[k])), m1 = add(m1, scl(0.25, dx[k]));
	for (int k = 0; k < 4; ++k)
		x0[k] = sub(x0[k], m0), dx[k] = sub(dx[k], m1);
	double um = 0;
	for (int i = 0; i < 4; ++i)
		for (int j = i + 1; j < 4; ++j)
		{
			d3	   d = sub(dx[i], dx[j]);
			double l = dot(d, d);
			um = um > l ? um : l;
		}
	um = sqrt(um);
	if (um == 0.0)
		return 1.0;
	double toi = 0, max_t = 1.0;
	double eps = 0.01 * (sqrt(ccd_d2(is_ee, x0)) - offset);
	double target = eps + offset, eps_sqr = eps * eps, inv = 1.0 / um;
	for (int iter = 0;; ++iter)
	{
		if (iter > 10000)
			return 0.001;
		d3 xt[4];
		for (int k = 0; k < 4; ++k)
			xt[k] = add(x0[k], scl(toi, dx[k]));
		double d2 = ccd_d2(is_ee, xt);
		double dmt = (d2 - target * target) / (sqrt(d2) + target);
		if ((max_t - toi) * um < dmt - eps)
			return max_t;
		if (toi > 0 && dmt * dmt < eps_sqr)
			return toi;
		double next = toi + dmt * inv;
		if (next == toi)
			return toi;
		toi = next;
		if (toi > max_t)
			return max_t;
	}
}

double or_toi_postprocess(double toi)
{
	if (toi < 0 || toi > 1.0 || toi == 0.001)
		toi = 1.0;
	if (toi < 3e-3)
		toi = 0.5;
	return toi;
}

double or_vf_ccd(uint32_t vid, uint32_t fid, const float* xs, const float* xe, const uint32_t* faces,
	const uint32_t* vprop, const float* off)
{
	uint32_t v[4] = { vid, faces[3 * fid], faces[3 * fid + 1], faces[3 * fid + 2] };
	if (vid == v[1] || vid == v[2] || vid == v[3] || invalid_combo(vprop, vid, v[1]))
		return 1.0;
	double x0[12], dx[12];
	for (int k = 0; k < 4; ++k)
		for (int a = 0; a < 3; ++a)
		{
			x0[3 * k + a] = xs[3 * v[k] + a];
			dx[3 * k + a] = (double)xe[3 * v[k] + a] - (double)xs[3 * v[k] + a];
		}
	return or_toi_postprocess(or_accd(0, x0, dx, (double)off[vid] + (double)off[v[1]]));
}

double or_ee_ccd(uint32_t left, uint32_t right, const float* xs, const float* xe, const uint32_t* edges,
	const uint32_t* vprop, const float* off)
{
	uint32_t v[4] = { edges[2 * left], edges[2 * left + 1], edges[2 * right], edges[2 * right + 1] };
	if (v[0] == v[2] || v[0] == v[3] || v[1] == v[2] || v[1] == v[3] || invalid_combo(vprop, v[0], v[2]))
		return 1.0;
	double x0[12], dx[12];
	for (int k = 0; k < 4; ++k)
		for (int a = 0; a < 3; ++a)
		{
			x0[3 * k + a] = xs[3 * v[k] + a];
			dx[3 * k + a] = (double)xe[3 * v[k] + a] - (double)xs[3 * v[k] + a];
		}
	return or_toi_postprocess(or_accd(1, x0, dx, (double)off[v[0]] + (double)off[v[3]]));
}

/* Loops over broad-phase pairs; returns the number emitted (records are 16 words). */
size_t or_dcd_list(int kind, const uint32_t* broad, size_t nb, const float* x, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* vprop, const float* dh, const float* off, const float* rva,
	const float* rea, const float* rfa, double kappa, int energy, uint32_t* recs, uint32_t* ids, size_t cap)
{
	size_t n = 0;
	for (size_t i = 0; i < nb; ++i)
	{
		uint32_t rec[16];
		int		 hit = kind == 0 ? or_vf_dcd(broad[2 * i], broad[2 * i + 1], x, faces, vprop, dh, off, rva, rfa, kappa, energy, rec)
								 : or_ee_dcd(broad[2 * i], broad[2 * i + 1], x, edges, vprop, dh, off, rea, kappa, energy, rec);
		if (!hit)
			continue;
		if (n < cap)
		{
			memcpy(recs + 16 * n, rec, 64);
			ids[2 * n] = broad[2 * i], ids[2 * n + 1] = broad[2 * i + 1];
		}
		n++;
	}
	return n;
}

double or_ccd_list(int kind, const uint32_t* broad, size_t nb, const float* xs, const float* xe,
	const uint32_t* edges, const uint32_t* faces, const uint32_t* vprop, const float* off)
{
	double m = 1.0;
	/* the minimum does not depend on the order, so threads (or_set_threads) change nothing */
#pragma omp parallel for reduction(min : m) schedule(dynamic, 4096) num_threads(or_nthreads) if (or_nthreads > 1)
	for (size_t i = 0; i < nb; ++i)
	{
		double t = kind == 2 ? or_vf_ccd(broad[2 * i], broad[2 * i + 1], xs, xe, faces, vprop, off)
							 : or_ee_ccd(broad[2 * i], broad[2 * i + 1], xs, xe, edges, vprop, off);
		if (t != 1.0 && (float)t < m)
			m = (float)t;
	}
	return m;
}
