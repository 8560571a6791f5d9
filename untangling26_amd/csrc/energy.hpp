#pragma once

#include <vector>
#include "ut_runtime.hpp"
#include "contact.hpp"

namespace ut
{

struct AssembleArgs
{
	uint32_t nv = 0, nf = 0, nh = 0, nslots = 0;
	const float*  x = nullptr;	 // current positions float[3V]
	const double* x64 = nullptr; // FP64 positions double[3V]; when set they replace x
	// inertia: m / h^2 (x - x_tilde)
	const double* x_tilde = nullptr; // double[3V]
	const float*  mass = nullptr;	 // per vertex
	double		  inv_h2 = 0.0;
	double		  inertia_grad_scale = 1.0; // 0 in quasi-static mode (no inertia gradient or energy)
	double		  inertia_hess_scale = 1.0; // 0 when the quasi-static inertia Hessian is off
	// BW98 stretch / shear per face
	const uint32_t* faces = nullptr;
	const float*	dm = nullptr;		 // float[4F]: d0, d1, d2, d3 with Dm = [[d0, d2], [d1, d3]]
	const float*	face_area = nullptr; // rest area
	const float*	mu_lambda = nullptr; // float[2F]
	const uint32_t* face_slots = nullptr; // uint32[9F]: slot of block (a, b)
	// dihedral bending per hinge (e0, e1, o0, o1)
	const uint32_t* hinges = nullptr;
	const float*	rest_angle = nullptr;
	const float*	bend_stiff_area = nullptr;
	const uint32_t* hinge_slots = nullptr; // uint32[16H]
	// contacts (diagonal blocks + gradient)
	const ContactRec* contacts = nullptr;
	const uint32_t*	  contact_count = nullptr;
	uint32_t		  contact_cap = 0;
	// Deterministic assembly (optional).  Without it every term is added with FP64
	// atomics, so the rounding of a sum depends on the order the hardware applies
	// them in.  With it: faces and hinges in colour classes whose members share no
	// vertex, applied class after class with plain read-modify-writes (each slot
	// and gradient entry gets its terms in a fixed order), and the contacts
	// gathered per vertex from entry lists sorted by vertex (ct_off[v] ..
	// ct_off[v + 1] of ct_entries, entry e = 4 c + a: corner a of contact c).
	const uint32_t* face_order = nullptr;	   // [nf] faces grouped by colour
	const uint32_t* face_color_off = nullptr;  // HOST [face_colors + 1]
	uint32_t		face_colors = 0;
	const uint32_t* hinge_order = nullptr;	   // [nh]
	const uint32_t* hinge_color_off = nullptr; // HOST [hinge_colors + 1]
	uint32_t		hinge_colors = 0;
	const uint32_t* ct_off = nullptr;	  // [nv + 1]
	const uint32_t* ct_entries = nullptr; // sorted entries
	// outputs
	const uint32_t* diag_slot = nullptr;
	double*			grad = nullptr;	  // dE/dx, double[3V] (the Newton rhs is -grad)
	double*			vals = nullptr;	  // double[9 * nslots]
	float*			vals32 = nullptr; // optional FP32 copy of vals
	double*			energy = nullptr; // optional double[4]: inertia, stretch, bending, (unused)
};

void assemble(Context* ctx, const AssembleArgs& a);

// Greedy colouring of elements of k vertices each (elem[k * i + j]) so that no two
// elements of one colour share a vertex: order = elements by colour (ascending
// index within a colour), color_off = colour offsets.  Returns false when more
// than 256 colours would be needed.
bool color_elements(const uint32_t* elem, uint32_t n, int k, uint32_t nv, std::vector<uint32_t>& order,
	std::vector<uint32_t>& color_off);

} // namespace ut
