// Affine bodies (ABD): 12 DOF per body, q_b = (p, a1, a2, a3) with the vertex
// positions x_v = p + A xbar_v, A = [a1 a2 a3] (fn_apply_q_to_x_template,
// newton_solver.cpp:1882-1907).  The Newton system lives in q-space: per body a
// dense 12x12 block (inertia, orthogonality and every contact side on that
// body), and the coupling of a contact's two bodies applied matrix-free through
// the weighted model points of its two sides.
#pragma once

#include "contact_rec.hpp"
#include "ut_runtime.hpp"

namespace ut
{

struct AbdDesc
{
	uint32_t		nb = 0, nv = 0;
	const uint32_t* vert_off = nullptr; // [nb + 1]: the vertices of body b are [vert_off[b], vert_off[b + 1])
	const uint32_t* body = nullptr;		// [nv]: body of each vertex
	const float*	xbar = nullptr;		// [3 nv]: model position of each vertex
	const double*	mass4 = nullptr;	// [16 nb]: compressed 4x4 mass matrix (row-major)
	const double*	volume = nullptr;	// [nb]
	const double*	kappa = nullptr;	// [nb]: orthogonality stiffness
};

struct AbdAssembleArgs
{
	const double* q = nullptr;		 // [12 nb]
	const double* q_tilde = nullptr; // [12 nb] inertia target (null: no inertia)
	double		  inv_h2 = 0.0;
	double		  inertia_grad_scale = 1.0; // 0: quasi-static (no inertia gradient / energy)
	double		  inertia_hess_scale = 1.0; // 0: no inertia Hessian either
	const ContactRec* contacts = nullptr;
	const uint32_t*	  contact_count = nullptr;
	uint32_t		  contact_cap = 0;
	double*			  grad = nullptr;	// [12 nb] dE/dq
	double*			  blocks = nullptr; // [144 nb] row-major 12x12 per body (optional)
	double*			  energy = nullptr; // [2]: inertia, orthogonality (optional)
};

struct AbdPcgReport
{
	uint32_t iterations = 0, refinements = 0;
	double	 relative_residual = 0.0;
	int		 converged = 0, failure = 0;
};

// x = q-to-x of every vertex (FP32 or FP64 output)
void abd_q_to_x(Context* ctx, const AbdDesc& d, const double* q, float* x32, double* x64);
// gradient (and blocks) of inertia + orthogonality + contacts in q-space; grad / blocks are overwritten
void abd_assemble(Context* ctx, const AbdDesc& d, const AbdAssembleArgs& a);
// gq (+)= J^T gx: a vertex-space gradient gathered onto the bodies
void abd_gather(Context* ctx, const AbdDesc& d, const double* gx, double* gq, bool accumulate);
// y = A p: the blocks plus the contacts between bodies (matrix-free)
void abd_apply(Context* ctx, const AbdDesc& d, const double* blocks, const ContactRec* contacts, const uint32_t* count,
	uint32_t cap, const double* p, double* y);

// Block-Jacobi PCG in q-space in one workgroup (the system has 12 nb unknowns):
// precision kPcgFp64, or kPcgMixed (FP32 inner solves with FP64 residual refinement).
struct AbdPcg
{
	Context*	   ctx = nullptr;
	DevBuf<double> minv64, work64, status;
	DevBuf<float>  blocks32, minv32, work32;
	uint32_t	   nb = 0;
	void init(Context* c, uint32_t num_bodies);
	AbdPcgReport solve(const AbdDesc& d, const double* blocks, const ContactRec* contacts, const uint32_t* count,
		uint32_t cap, const double* b, double* x, int precision, uint32_t max_iters, double rel_tol);
};

} // namespace ut
