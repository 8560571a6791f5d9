/*
 * oracle_energy.c -- FP64 restatement of the cloth material terms of the
 * incremental potential.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * BW98 stretch + shear with projected Hessians: Energy/stretch_energy.h:200-520
 * and Energies/stretch_face_energy.cpp; dihedral bending with the
 * Gauss-Newton Hessian: Energy/bending_energy.h:164-260 and
 * Energies/bending_energy_kernel.cpp:51-100.
 *
 * Pinning (tests/test_oracle.py): gradients against FP64 central differences
 * of the energies (h = 1e-7 relative to the edge length); the stretch Hessian
 * against central differences of the gradient where its projection is
 * inactive (all fibres stretched); every projected Hessian symmetric positive
 * semi-definite; zero energy and gradient in the rest state.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

static double sgn(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void	  cross3(const double* a, const double* b, double* o)
{
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}

/* energy, dE/dF (6) and projected d2E/dF2 (6x6 row-major) */
static double bw98(const double* Fu, const double* Fv, double mu, double lam, double* g, double* H)
{
	const double I5u = dot3(Fu, Fu), I5v = dot3(Fv, Fv), I6 = dot3(Fu, Fv);
	const double su = sqrt(I5u), sv = sqrt(I5v), isu = 1.0 / su, isv = 1.0 / sv;
	const double e = 0.5 * mu * ((su - 1) * (su - 1) + (sv - 1) * (sv - 1)) + 0.5 * lam * I6 * I6;
	const double gu = mu * (su - 1) * isu, gv = mu * (sv - 1) * isv;
	for (int k = 0; k < 3; ++k)
	{
		g[k] = gu * Fu[k] + lam * I6 * Fv[k];
		g[3 + k] = gv * Fv[k] + lam * I6 * Fu[k];
	}
	memset(H, 0, 36 * sizeof(double));
	const double du = fmax(0.0, 1.0 - isu), dv = fmax(0.0, 1.0 - isv);
	const double cu = isu < 1 ? isu : 1, cv = isv < 1 ? isv : 1;
	for (int i = 0; i < 3; ++i)
	{
		H[6 * i + i] += mu * du;
		H[6 * (3 + i) + 3 + i] += mu * dv;
		for (int j = 0; j < 3; ++j)
		{
			H[6 * i + j] += mu * cu * (Fu[i] * isu) * (Fu[j] * isu);
			H[6 * (3 + i) + 3 + j] += mu * cv * (Fv[i] * isv) * (Fv[j] * isv);
		}
	}
	/* shear (stretch_energy.h:291-323) */
	const double s6 = sgn(I6);
	const double gg[6] = { Fv[0], Fv[1], Fv[2], Fu[0], Fu[1], Fu[2] };
	const double I2 = I5u + I5v;
	const double l0 = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
	double		 q[6], qn = 0, Tq[6], tn = 0;
	for (int i = 0; i < 6; ++i)
	{
		const double Sg = i < 3 ? gg[i + 3] : gg[i - 3];
		q[i] = I6 * Sg + l0 * gg[i];
		qn += q[i] * q[i];
	}
	qn = sqrt(qn);
	for (int i = 0; i < 6; ++i)
		q[i] = qn > 0 ? q[i] / qn : 0.0;
	for (int i = 0; i < 6; ++i)
	{
		const double Sq = i < 3 ? q[i + 3] : q[i - 3];
		Tq[i] = 0.5 * (q[i] + s6 * Sq);
		tn += Tq[i] * Tq[i];
	}
	for (int i = 0; i < 6; ++i)
		for (int j = 0; j < 6; ++j)
		{
			const double Sij = ((i < 3 && j == i + 3) || (i >= 3 && j == i - 3)) ? 1.0 : 0.0;
			const double Tij = 0.5 * ((i == j ? 1.0 : 0.0) + s6 * Sij);
			const double proj = tn > 0 ? Tq[i] * Tq[j] / tn : 0.0;
			H[6 * i + j] += lam * (fabs(I6) * (Tij - proj) + l0 * q[i] * q[j]);
		}
	return e;
}

/* One face: energy; grad[9] (vertex a, component k at 3a+k); hess[81] with
 * block (a, b) entry (i, j) at 27a + 9b + 3i + j. */
double or_bw98_face(const double* x, const float* dm, double area, double mu, double lam, double* grad, double* hess)
{
	const double d0 = dm[0], d1 = dm[1], d2 = dm[2], d3 = dm[3];
	double		 r1[3], r2[3], Fu[3], Fv[3], g[6], H[36];
	for (int k = 0; k < 3; ++k)
	{
		r1[k] = x[3 + k] - x[k];
		r2[k] = x[6 + k] - x[k];
		Fu[k] = d0 * r1[k] + d1 * r2[k];
		Fv[k] = d2 * r1[k] + d3 * r2[k];
	}
	const double e = bw98(Fu, Fv, mu, lam, g, H);
	const double c[3][2] = { { -(d0 + d1), -(d2 + d3) }, { d0, d2 }, { d1, d3 } };
	if (grad)
		for (int a = 0; a < 3; ++a)
			for (int k = 0; k < 3; ++k)
				grad[3 * a + k] = area * (c[a][0] * g[k] + c[a][1] * g[3 + k]);
	if (hess)
		for (int a = 0; a < 3; ++a)
			for (int b = 0; b < 3; ++b)
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j)
					{
						double h = 0;
						for (int p = 0; p < 2; ++p)
							for (int qq = 0; qq < 2; ++qq)
								h += c[a][p] * c[b][qq] * H[6 * (3 * p + i) + 3 * qq + j];
						hess[27 * a + 9 * b + 3 * i + j] = area * h;
					}
	return area * e;
}

/* Hinge (e0, e1, o0, o1) in x[12]: returns the energy 1/2 k (theta - theta0)^2;
 * grad[12]; hess[144] Gauss-Newton, block (a, b) entry (i, j) at 36a + 9b + 3i + j;
 * *theta_out the dihedral angle. */
double or_bending_hinge(const double* x, double theta0, double k, double* grad, double* hess, double* theta_out)
{
	const double *e0 = x, *e1 = x + 3, *o0 = x + 6, *o1 = x + 9;
	double		  a[3], b[3], n1a[3], n2a[3];
	for (int i = 0; i < 3; ++i)
		a[i] = e1[i] - o0[i], b[i] = e0[i] - o0[i];
	cross3(a, b, n1a);
	for (int i = 0; i < 3; ++i)
		a[i] = e0[i] - o1[i], b[i] = e1[i] - o1[i];
	cross3(a, b, n2a);
	const double den = fmax(sqrt(dot3(n1a, n1a) * dot3(n2a, n2a)), 1e-12);
	double		 c = dot3(n1a, n2a) / den;
	c = fmax(-1.0, fmin(1.0, c));
	double cr[3], d[3];
	cross3(n2a, n1a, cr);
	for (int i = 0; i < 3; ++i)
		d[i] = e1[i] - e0[i];
	const double theta = sgn(dot3(cr, d)) * acos(c);
	/* gradient (face_dihedral_angle_grad with v2 = o0, v0 = e1, v1 = e0, v3 = o1) */
	double E0[3], E1[3], E2[3], E3[3], E4[3], n1[3], n2[3];
	for (int i = 0; i < 3; ++i)
	{
		E0[i] = e0[i] - e1[i];
		E1[i] = o0[i] - e1[i];
		E2[i] = o1[i] - e1[i];
		E3[i] = o0[i] - e0[i];
		E4[i] = o1[i] - e0[i];
	}
	cross3(E0, E1, n1);
	cross3(E2, E0, n2);
	const double n1s = fmax(dot3(n1, n1), 1e-12), n2s = fmax(dot3(n2, n2), 1e-12);
	const double e0n = fmax(sqrt(dot3(E0, E0)), 1e-6);
	double		 G[4][3];
	for (int i = 0; i < 3; ++i)
	{
		G[2][i] = (-e0n / n1s) * n1[i];
		G[1][i] = (-dot3(E0, E3) / (e0n * n1s)) * n1[i] - (dot3(E0, E4) / (e0n * n2s)) * n2[i];
		G[0][i] = (dot3(E0, E1) / (e0n * n1s)) * n1[i] + (dot3(E0, E2) / (e0n * n2s)) * n2[i];
		G[3][i] = (-e0n / n2s) * n2[i];
	}
	const double dth = theta - theta0;
	if (grad)
		for (int v = 0; v < 4; ++v)
			for (int i = 0; i < 3; ++i)
				grad[3 * v + i] = k * dth * G[v][i];
	if (hess)
		for (int va = 0; va < 4; ++va)
			for (int vb = 0; vb < 4; ++vb)
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j)
						hess[36 * va + 9 * vb + 3 * i + j] = k * G[va][i] * G[vb][j];
	if (theta_out)
		*theta_out = theta;
	return 0.5 * k * dth * dth;
}

/* Whole assembly in the SELL slot numbering of the GPU path: grad[3V] = dE/dx,
 * vals[9 * nslots] planar.  Inertia m/h^2 (x - x_tilde), BW98 faces, bending
 * hinges, contacts (16-word CollisionPairTemplate) gradient + diagonal blocks. */
void or_assemble(uint32_t nv, uint32_t nf, uint32_t nh, const float* x, const double* x_tilde, const float* mass,
	double inv_h2, const uint32_t* faces, const float* dm, const float* area, const float* mu_lambda,
	const uint32_t* face_slots, const uint32_t* hinges, const float* rest_angle, const float* bend,
	const uint32_t* hinge_slots, const uint32_t* contacts, uint32_t ncontacts, const uint32_t* diag_slot,
	uint32_t nslots, double* grad, double* vals, double* energy)
{
	memset(grad, 0, sizeof(double) * 3 * nv);
	memset(vals, 0, sizeof(double) * 9 * (size_t)nslots);
	double e_in = 0, e_st = 0, e_be = 0;
	for (uint32_t v = 0; v < nv; ++v)
	{
		const double k = (double)mass[v] * inv_h2;
		double		 dd = 0;
		for (int a = 0; a < 3; ++a)
		{
			const double d = (double)x[3 * v + a] - x_tilde[3 * v + a];
			grad[3 * v + a] += k * d;
			dd += d * d;
		}
		e_in += 0.5 * k * dd;
		vals[0 * (size_t)nslots + diag_slot[v]] += k;
		vals[4 * (size_t)nslots + diag_slot[v]] += k;
		vals[8 * (size_t)nslots + diag_slot[v]] += k;
	}
	for (uint32_t f = 0; f < nf; ++f)
	{
		double xx[9], g[9], H[81];
		for (int a = 0; a < 3; ++a)
			for (int k = 0; k < 3; ++k)
				xx[3 * a + k] = x[3 * faces[3 * f + a] + k];
		e_st += or_bw98_face(xx, dm + 4 * f, area[f], mu_lambda[2 * f], mu_lambda[2 * f + 1], g, H);
		for (int a = 0; a < 3; ++a)
		{
			for (int k = 0; k < 3; ++k)
				grad[3 * faces[3 * f + a] + k] += g[3 * a + k];
			for (int b = 0; b < 3; ++b)
				for (int c = 0; c < 9; ++c)
					vals[(size_t)c * nslots + face_slots[9 * f + 3 * a + b]] += H[27 * a + 9 * b + c];
		}
	}
	for (uint32_t h = 0; h < nh; ++h)
	{
		double xx[12], g[12], H[144];
		for (int a = 0; a < 4; ++a)
			for (int k = 0; k < 3; ++k)
				xx[3 * a + k] = x[3 * hinges[4 * h + a] + k];
		e_be += or_bending_hinge(xx, rest_angle[h], bend[h], g, H, NULL);
		for (int a = 0; a < 4; ++a)
		{
			for (int k = 0; k < 3; ++k)
				grad[3 * hinges[4 * h + a] + k] += g[3 * a + k];
			for (int b = 0; b < 4; ++b)
				for (int c = 0; c < 9; ++c)
					vals[(size_t)c * nslots + hinge_slots[16 * h + 4 * a + b]] += H[36 * a + 9 * b + c];
		}
	}
	for (uint32_t i = 0; i < ncontacts; ++i)
	{
		/* gradient w_a (k1 n + friction gradient), diagonal blocks w_a^2 (k2 n n^T + friction
		 * Hessian): fn_perPair_assemble_gradient_hessian (narrow_phase.cpp:2630-2700) */
		const uint32_t* r = contacts + 16 * i;
		double			w[4], K[9], g[3];
		or_record_weights(r, w);
		or_pair_block(r, K, g);
		for (int a = 0; a < 4; ++a)
		{
			const uint32_t v = r[a] & 0x7FFFFFFFu;
			for (int k = 0; k < 3; ++k)
				grad[3 * v + k] += w[a] * g[k];
			for (int c = 0; c < 9; ++c)
				vals[(size_t)c * nslots + diag_slot[v]] += w[a] * w[a] * K[c];
		}
	}
	if (energy)
		energy[0] = e_in, energy[1] = e_st, energy[2] = e_be;
}
