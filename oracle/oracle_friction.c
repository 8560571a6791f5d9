/*
 * oracle_friction.c -- FP64 restatement of the lagged IPC friction of a contact
 * pair and of the contact potential at given positions.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 *   or_friction_eval      Friction::ipc_barrier::compute_friction_energy and
 *                         compute_friction_gradient_hessian (friction_kernel.hpp:137-301)
 *   or_pair_block         the pair's Hessian block k2 n n^T + friction Hessian
 *                         (narrow_phase.cpp:2643-2654)
 *   or_friction_init      fn_process_collision_pair_friction, first iteration
 *                         (narrow_phase.cpp:1479-1491, get_friction_mu :357-363)
 *   or_friction_lagged    the same kernel in later iterations (:1492-1506)
 *   or_contact_potential  barrier / quadratic repulsion re-solved at x and the
 *                         lagged friction, energies and gradient
 *                         (fn_compute_repulsion_energy, narrow_phase.cpp:3432-3530)
 *
 * Pinning (tests/test_oracle_friction.py): closed forms of f0 / f1 / f2 on both
 * sides of E and the sliding regime (D = mu_lambda (s - E/3), gradient
 * mu_lambda u / s, Hessian mu_lambda (P - u u^T / s^2) / s); the normal
 * component of the displacement leaves D unchanged; FP64 central differences of
 * the energy against the gradient and of the gradient against the Hessian;
 * the Hessian is positive semidefinite; or_contact_potential's gradient against
 * central differences of its energy on the two-sheet proximity records.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static const double kE = 1e-4;		/* Friction::ando_barrier::friction_eps */
static const double kSmallS = 1e-6; /* Friction::ipc_barrier::small_s */

static double f0(double y, double E) { return y < E ? (y * y) / E - (y * y * y) / (3.0 * E * E) : y - E / 3.0; }
static double f1s(double y, double E) { return y < E ? 2.0 / E - y / (E * E) : 1.0 / y; }
static double f2(double y, double E) { return y < E ? 2.0 / E - (2.0 * y) / (E * E) : 0.0; }

double or_friction_eval(double mu_lambda, const double* n, const double* rel, double E, double* grad, double* hess)
{
	const double nr = n[0] * rel[0] + n[1] * rel[1] + n[2] * rel[2];
	double		 u[3];
	for (int k = 0; k < 3; ++k)
		u[k] = rel[k] - nr * n[k];
	const double s2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], s = sqrt(s2);
	const int	 tiny = s2 < kSmallS * kSmallS;
	const double a = f1s(s, E);
	if (grad)
		for (int k = 0; k < 3; ++k)
			grad[k] = tiny ? 0.0 : mu_lambda * a * u[k];
	if (hess)
		for (int r = 0; r < 3; ++r)
			for (int c = 0; c < 3; ++c)
			{
				const double P = (r == c ? 1.0 : 0.0) - n[r] * n[c];
				double		 h = mu_lambda * a * P;
				if (!tiny)
					h += mu_lambda * (f2(s, E) - a) * (u[r] / s) * (u[c] / s);
				hess[3 * r + c] = h;
			}
	return tiny ? 0.0 : mu_lambda * f0(s, E);
}

static void rec_fields(const uint32_t* r, float f[12]) { memcpy(f, r + 4, 48); }

void or_record_weights(const uint32_t* r, double* w)
{
	float f[12];
	rec_fields(r, f);
	const double b0 = f[6], b1 = f[7];
	if (r[2] & 0x80000000u)
		w[0] = 1.0, w[1] = -b0, w[2] = -b1, w[3] = b0 + b1 - 1.0;
	else
		w[0] = b0, w[1] = 1.0 - b0, w[2] = -b1, w[3] = b1 - 1.0;
}

void or_pair_block(const uint32_t* r, double* K, double* g)
{
	float f[12];
	rec_fields(r, f);
	const double n[3] = { f[0], f[1], f[2] }, k1 = f[4], k2 = f[5], lam = f[11];
	const double rel[3] = { f[8], f[9], f[10] };
	double		 fg[3] = { 0, 0, 0 }, fh[9] = { 0 };
	if (lam != 0.0)
		or_friction_eval(lam, n, rel, kE, fg, fh);
	for (int a = 0; a < 3; ++a)
	{
		if (g)
			g[a] = k1 * n[a] + fg[a];
		if (K)
			for (int b = 0; b < 3; ++b)
				K[3 * a + b] = k2 * n[a] * n[b] + fh[3 * a + b];
	}
}

void or_friction_init(uint32_t* recs, uint32_t n, const float* mu)
{
	for (uint32_t i = 0; i < n; ++i)
	{
		uint32_t* r = recs + 16 * i;
		float	  f[12];
		rec_fields(r, f);
		const float m = 0.5f * (mu[r[0] & 0x7FFFFFFFu] + mu[r[2] & 0x7FFFFFFFu]);
		f[11] = -m * f[4];
		f[8] = f[9] = f[10] = 0.0f;
		memcpy(r + 4, f, 48);
	}
}

void or_friction_lagged(uint32_t* recs, uint32_t n, const double* x, const double* x_start)
{
	for (uint32_t i = 0; i < n; ++i)
	{
		uint32_t* r = recs + 16 * i;
		double	  w[4], rel[3] = { 0, 0, 0 };
		or_record_weights(r, w);
		for (int k = 0; k < 4; ++k)
		{
			const uint32_t v = r[k] & 0x7FFFFFFFu;
			for (int c = 0; c < 3; ++c)
				rel[c] += w[k] * (x[3 * v + c] - x_start[3 * v + c]);
		}
		float f[12];
		rec_fields(r, f);
		f[8] = (float)rel[0], f[9] = (float)rel[1], f[10] = (float)rel[2];
		f[4] = f[5] = 0.0f;
		memcpy(r + 4, f, 48);
	}
}

void or_contact_potential(const uint32_t* recs, uint32_t n, const double* x, const double* x_start, const float* d_hat,
	const float* offset, double kappa, int energy, int terms, double* grad, double* out)
{
	double eb = 0.0, ef = 0.0;
	for (uint32_t i = 0; i < n; ++i)
	{
		const uint32_t* r = recs + 16 * i;
		float			f[12];
		rec_fields(r, f);
		uint32_t v[4];
		for (int k = 0; k < 4; ++k)
			v[k] = r[k] & 0x7FFFFFFFu;
		const double* p[4] = { x + 3 * v[0], x + 3 * v[1], x + 3 * v[2], x + 3 * v[3] };
		const int	  vf = (r[2] & 0x80000000u) != 0;
		if ((terms & 1) && f[5] != 0.0f)
		{
			double w[4], diff[3], c[4];
			if (vf)
			{
				or_pt_coeff(p[0], p[1], p[2], p[3], c);
				w[0] = c[0] + c[1] + c[2], w[1] = -c[0], w[2] = -c[1], w[3] = -c[2];
				for (int k = 0; k < 3; ++k)
					diff[k] = c[0] * (p[0][k] - p[1][k]) + c[1] * (p[0][k] - p[2][k]) + c[2] * (p[0][k] - p[3][k]);
			}
			else
			{
				or_ee_coeff(p[0], p[1], p[2], p[3], c);
				w[0] = c[0], w[1] = c[1], w[2] = -c[2], w[3] = -c[3];
				for (int k = 0; k < 3; ++k)
					diff[k] = (c[0] * p[0][k] + c[1] * p[1][k]) - (c[2] * p[2][k] + c[3] * p[3][k]);
			}
			const uint32_t ka = v[0], kb = vf ? v[1] : v[2];
			const double   d = sqrt(diff[0] * diff[0] + diff[1] * diff[1] + diff[2] * diff[2]);
			const double   dh = ((double)d_hat[ka] + (double)d_hat[kb]) * 0.5;
			const double   th = (double)offset[ka] + (double)offset[kb];
			const double   stiff = (double)f[3] * kappa;
			double		   e, de;
			if (energy == 0)
			{
				const double C = (d - th) - dh;
				e = 0.5 * stiff * C * C, de = stiff * C;
			}
			else
			{
				e = stiff * or_barrier(d - th, dh, 0);
				de = stiff * or_barrier(d - th, dh, 1);
			}
			eb += e;
			if (grad && d > 0.0)
				for (int k = 0; k < 4; ++k)
					for (int c3 = 0; c3 < 3; ++c3)
						grad[3 * v[k] + c3] += de * w[k] * diff[c3] / d;
		}
		if ((terms & 2) && f[11] != 0.0f && x_start)
		{
			double w[4], rel[3] = { 0, 0, 0 }, g[3];
			or_record_weights(r, w);
			for (int k = 0; k < 4; ++k)
				for (int c3 = 0; c3 < 3; ++c3)
					rel[c3] += w[k] * (p[k][c3] - x_start[3 * v[k] + c3]);
			const double nn[3] = { f[0], f[1], f[2] };
			ef += or_friction_eval(f[11], nn, rel, kE, g, NULL);
			if (grad)
				for (int k = 0; k < 4; ++k)
					for (int c3 = 0; c3 < 3; ++c3)
						grad[3 * v[k] + c3] += w[k] * g[c3];
		}
	}
	out[0] = eb, out[1] = ef;
}
