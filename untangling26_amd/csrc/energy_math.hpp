// Cloth material energies in FP64: Baraff-Witkin 98 stretch + shear with the
// projected Hessians of Energy/stretch_energy.h:200-520, and the dihedral
// bending term of Energy/bending_energy.h:164-260 + Energies/bending_energy_kernel.cpp
// (Gauss-Newton Hessian).  Host/device so the CPU restatement can share the
// exact algebra when built from the same header (it is not: the oracle restates
// it in C), and so that unit tests can exercise it on the host.
#pragma once

#include "contact_math.hpp"

namespace ut
{

// 6x6 symmetric work matrix, row-major
struct M6
{
	double a[36];
	UT_HD double& operator()(int r, int c) { return a[6 * r + c]; }
	UT_HD double  operator()(int r, int c) const { return a[6 * r + c]; }
};

// F = [x1 - x0, x2 - x0] Dm with Dm = [[d0, d2], [d1, d3]] (row-major); the
// per-vertex coefficients of dF/dx are c[a] = {dFu/dxa, dFv/dxa}:
//   x0: (-(d0 + d1), -(d2 + d3)),  x1: (d0, d2),  x2: (d1, d3)   (get_dFdx, stretch_energy.h:60-90)
struct Bw98
{
	double energy;
	double dedF[6]; // [dE/dFu ; dE/dFv]
	M6	   H;		// d2E/dF2, projected to PSD
};

UT_HD double sgn(double v) { return v > 0 ? 1.0 : (v < 0 ? -1.0 : 0.0); }

UT_HD Bw98 bw98_eval(const v3<double>& Fu, const v3<double>& Fv, double mu, double lambda)
{
	Bw98 r;
	for (int i = 0; i < 36; ++i)
		r.H.a[i] = 0;
	const double I5u = dot(Fu, Fu), I5v = dot(Fv, Fv);
	const double su = sqrt(I5u), sv = sqrt(I5v);
	const double isu = 1.0 / su, isv = 1.0 / sv;
	const double I6 = dot(Fu, Fv);
	// energies (stretch_energy.h:215-234)
	r.energy = 0.5 * mu * ((su - 1.0) * (su - 1.0) + (sv - 1.0) * (sv - 1.0)) + 0.5 * lambda * I6 * I6;
	// gradient: stretch (:236-253) + shear (:254-262)
	const double gu = mu * (su - 1.0) * isu, gv = mu * (sv - 1.0) * isv;
	r.dedF[0] = gu * Fu.x + lambda * I6 * Fv.x;
	r.dedF[1] = gu * Fu.y + lambda * I6 * Fv.y;
	r.dedF[2] = gu * Fu.z + lambda * I6 * Fv.z;
	r.dedF[3] = gv * Fv.x + lambda * I6 * Fu.x;
	r.dedF[4] = gv * Fv.y + lambda * I6 * Fu.y;
	r.dedF[5] = gv * Fv.z + lambda * I6 * Fu.z;
	// stretch Hessian (:264-290)
	{
		const double du = fmax(0.0, 1.0 - isu), dv = fmax(0.0, 1.0 - isv);
		const double cu = isu < 1.0 ? isu : 1.0, cv = isv < 1.0 ? isv : 1.0;
		const double fu[3] = { Fu.x * isu, Fu.y * isu, Fu.z * isu };
		const double fv[3] = { Fv.x * isv, Fv.y * isv, Fv.z * isv };
		for (int i = 0; i < 3; ++i)
		{
			r.H(i, i) += mu * du;
			r.H(3 + i, 3 + i) += mu * dv;
			for (int j = 0; j < 3; ++j)
			{
				r.H(i, j) += mu * cu * fu[i] * fu[j];
				r.H(3 + i, 3 + j) += mu * cv * fv[i] * fv[j];
			}
		}
	}
	// shear Hessian (:291-323): S swaps the two column blocks
	{
		const double s6 = sgn(I6);
		const double g[6] = { Fv.x, Fv.y, Fv.z, Fu.x, Fu.y, Fu.z }; // flatten(F [[0,1],[1,0]])
		const double I2 = I5u + I5v;
		const double l0 = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
		double		 q[6];
		double		 qn = 0;
		for (int i = 0; i < 6; ++i)
		{
			const double Sg = i < 3 ? g[i + 3] : g[i - 3];
			q[i] = I6 * Sg + l0 * g[i];
			qn += q[i] * q[i];
		}
		qn = sqrt(qn);
		for (int i = 0; i < 6; ++i)
			q[i] = qn > 0 ? q[i] / qn : 0.0;
		// T = (I + s6 S) / 2, Tq = T q
		double Tq[6], tn = 0;
		for (int i = 0; i < 6; ++i)
		{
			const double Sq = i < 3 ? q[i + 3] : q[i - 3];
			Tq[i] = 0.5 * (q[i] + s6 * Sq);
			tn += Tq[i] * Tq[i];
		}
		const double aI6 = fabs(I6);
		for (int i = 0; i < 6; ++i)
			for (int j = 0; j < 6; ++j)
			{
				const double Sij = ((i < 3 && j == i + 3) || (i >= 3 && j == i - 3)) ? 1.0 : 0.0;
				const double Tij = 0.5 * ((i == j ? 1.0 : 0.0) + s6 * Sij);
				const double proj = tn > 0 ? Tq[i] * Tq[j] / tn : 0.0;
				r.H(i, j) += lambda * (aI6 * (Tij - proj) + l0 * q[i] * q[j]);
			}
	}
	return r;
}

// Dihedral angle and its gradient for hinge (e0, e1, o0, o1): the two faces
// are (o0, e1, e0) and (o1, e0, e1); face_dihedral_angle / _grad of
// bending_energy.h:166-236 with v0 = o0, v1 = e1, v2 = e0, v3 = o1 for the
// angle and (v2, v0, v1, v3) = (o0, e1, e0, o1) for the gradient.
UT_HD double dihedral(const v3<double>& e0, const v3<double>& e1, const v3<double>& o0, const v3<double>& o1,
	v3<double> grad[4])
{
	// angle
	const v3<double> n1a = cross(e1 - o0, e0 - o0);
	const v3<double> n2a = cross(e0 - o1, e1 - o1);
	const double	 den = fmax(sqrt(dot(n1a, n1a) * dot(n2a, n2a)), 1e-12);
	double			 c = dot(n1a, n2a) / den;
	c = fmax(-1.0, fmin(1.0, c));
	const double theta = sgn(dot(cross(n2a, n1a), e1 - e0)) * acos(c);
	// gradient (v2 = o0, v0 = e1, v1 = e0, v3 = o1)
	const v3<double> E0 = e0 - e1, E1 = o0 - e1, E2 = o1 - e1, E3 = o0 - e0, E4 = o1 - e0;
	const v3<double> n1 = cross(E0, E1), n2 = cross(E2, E0);
	const double	 n1s = fmax(dot(n1, n1), 1e-12), n2s = fmax(dot(n2, n2), 1e-12);
	const double	 e0n = fmax(sqrt(dot(E0, E0)), 1e-6);
	const v3<double> g_o0 = (-e0n / n1s) * n1;
	const v3<double> g_e1 = (-dot(E0, E3) / (e0n * n1s)) * n1 - (dot(E0, E4) / (e0n * n2s)) * n2;
	const v3<double> g_e0 = (dot(E0, E1) / (e0n * n1s)) * n1 + (dot(E0, E2) / (e0n * n2s)) * n2;
	const v3<double> g_o1 = (-e0n / n2s) * n2;
	grad[0] = g_e0;
	grad[1] = g_e1;
	grad[2] = g_o0;
	grad[3] = g_o1;
	return theta;
}

} // namespace ut
