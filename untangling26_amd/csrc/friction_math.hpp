// Lagged IPC friction of one contact pair, and the pair's whole 3x3 Hessian
// block about its weighted point.
//
// Friction::ipc_barrier (friction_kernel.hpp:137-301): with the pair's lagged
// normal n, weights w and multiplier mu_lambda = -mu k1 (fixed in the first
// Newton iteration of a step, narrow_phase.cpp:1455-1521) and the tangential
// displacement u = (I - n n^T) sum_k w_k (x_k - x_k^start), s = |u|,
//   D = mu_lambda f0(s),  dD/d(rel) = mu_lambda f1(s)/s u,
//   d2D/d(rel)2 = mu_lambda [(f2(s) - f1(s)/s) (u/s)(u/s)^T + f1(s)/s (I - n n^T)]
// with the C1 smoothing f0, f1, f2 of width E = 1e-4 (friction_eps), and
// gradient / energy switched off below s = 1e-6 (small_s).  Evaluated in FP64.
#pragma once

#include "contact_math.hpp"
#include "contact_rec.hpp"

namespace ut
{

constexpr double kFrictionEps = 1e-4;	 // Friction::ando_barrier::friction_eps
constexpr double kFrictionSmallS = 1e-6; // Friction::ipc_barrier::small_s

// symmetric 3x3: xx, xy, xz, yy, yz, zz
struct Sym3
{
	double m[6];
};

UT_HD Sym3 sym_zero() { return Sym3{ { 0, 0, 0, 0, 0, 0 } }; }
UT_HD v3<double> sym_mul(const Sym3& K, v3<double> v)
{
	return v3<double>{ K.m[0] * v.x + K.m[1] * v.y + K.m[2] * v.z, K.m[1] * v.x + K.m[3] * v.y + K.m[4] * v.z,
		K.m[2] * v.x + K.m[4] * v.y + K.m[5] * v.z };
}
UT_HD double sym_at(const Sym3& K, int r, int c)
{
	const int i = r < c ? r : c, j = r < c ? c : r;
	return K.m[i == 0 ? j : (i == 1 ? 2 + j : 5)];
}
// K += a u u^T + b I
UT_HD void sym_add_outer(Sym3& K, double a, v3<double> u)
{
	K.m[0] += a * u.x * u.x, K.m[1] += a * u.x * u.y, K.m[2] += a * u.x * u.z;
	K.m[3] += a * u.y * u.y, K.m[4] += a * u.y * u.z, K.m[5] += a * u.z * u.z;
}
UT_HD void sym_add_diag(Sym3& K, double b) { K.m[0] += b, K.m[3] += b, K.m[5] += b; }

// f0, f1 / s, f2 of friction_kernel.hpp:142-185
UT_HD double fr_f0(double y, double E) { return y < E ? (y * y) / E - (y * y * y) / (3.0 * E * E) : y - E / 3.0; }
UT_HD double fr_f1_div_s(double y, double E) { return y < E ? 2.0 / E - y / (E * E) : 1.0 / y; }
UT_HD double fr_f2(double y, double E) { return y < E ? 2.0 / E - (2.0 * y) / (E * E) : 0.0; }

struct FrictionEval
{
	double	   energy;
	v3<double> grad; // with respect to the pair's relative displacement
	Sym3	   hess;
};

// compute_friction_energy + compute_friction_gradient_hessian (friction_kernel.hpp:187-299)
UT_HD FrictionEval friction_eval(double mu_lambda, v3<double> n, v3<double> rel, double E)
{
	FrictionEval r;
	const v3<double> u = rel - dot(n, rel) * n;
	const double	 s2 = dot(u, u), s = sqrt(s2);
	const double	 f1s = fr_f1_div_s(s, E);
	const bool		 tiny = s2 < kFrictionSmallS * kFrictionSmallS;
	r.energy = tiny ? 0.0 : mu_lambda * fr_f0(s, E);
	r.grad = tiny ? v3<double>{ 0, 0, 0 } : (mu_lambda * f1s) * u;
	// (f1/s) P, P = I - n n^T
	r.hess = sym_zero();
	sym_add_diag(r.hess, mu_lambda * f1s);
	sym_add_outer(r.hess, -mu_lambda * f1s, n);
	if (!tiny)
	{
		const v3<double> ub = (1.0 / s) * u;
		sym_add_outer(r.hess, mu_lambda * (fr_f2(s, E) - f1s), ub);
	}
	return r;
}

// The pair's Hessian block about its weighted point: the repulsion k2 n n^T
// plus the friction Hessian of the record's lagged values (narrow_phase.cpp:2643-2654).
// Block (a, b) of the pair's vertices is w_a w_b times this.
UT_HD Sym3 pair_block(const ContactRec& c)
{
	const v3<double> n{ (double)c.normal[0], (double)c.normal[1], (double)c.normal[2] };
	Sym3			 K = sym_zero();
	sym_add_outer(K, (double)c.k2, n);
	if (c.lambda != 0.0f)
	{
		const FrictionEval f =
			friction_eval((double)c.lambda, n, v3<double>{ (double)c.dv[0], (double)c.dv[1], (double)c.dv[2] },
				kFrictionEps);
		for (int k = 0; k < 6; ++k)
			K.m[k] += f.hess.m[k];
	}
	return K;
}

// The pair's gradient about its weighted point: k1 n plus the friction gradient.
UT_HD v3<double> pair_gradient(const ContactRec& c)
{
	const v3<double> n{ (double)c.normal[0], (double)c.normal[1], (double)c.normal[2] };
	v3<double>		 g = (double)c.k1 * n;
	if (c.lambda != 0.0f)
	{
		const FrictionEval f =
			friction_eval((double)c.lambda, n, v3<double>{ (double)c.dv[0], (double)c.dv[1], (double)c.dv[2] },
				kFrictionEps);
		g = g + f.grad;
	}
	return g;
}

// CollisionPairTemplate weights (collision_data.h:131-136)
UT_HD void record_weights(const ContactRec& c, double w[4])
{
	const double b0 = c.bary[0], b1 = c.bary[1];
	if (c.idx[2] & kTypeBit) // VF
	{
		w[0] = 1.0, w[1] = -b0, w[2] = -b1, w[3] = b0 + b1 - 1.0;
	}
	else // EE
	{
		w[0] = b0, w[1] = 1.0 - b0, w[2] = -b1, w[3] = b1 - 1.0;
	}
}

} // namespace ut
