// Preconditioned conjugate gradients on a 3x3-block sparse matrix.
//
// Replaces ConjugateGradientSolver::device_solve (Solver/LinearSolver/
// precond_cg.cpp:706-914) and its SpMV callback (newton_solver.cpp:1517-1555).
//
// Matrix layout: SELL-64 block rows.  Rows are grouped in slices of 64 (one
// wavefront); slice s has width w_s = its longest block row and owns the slots
// slice_off[s] .. slice_off[s] + 64 * w_s - 1, slot = slice_off[s] + k*64 + lane
// for the k-th block of row 64*s + lane (short rows are padded with zero
// blocks whose column is the row itself).  Block values are stored planar,
// vals[c * nslots + slot] for entry c = 3*i + j of the block, so each of the 9
// loads of a block is one contiguous 256 B (FP32) or 512 B (FP64) wave access.
//
// Precision modes:
//   kPcgFp64   matrix, vectors and dots in FP64
//   kPcgMixed  FP32 matrix and inner PCG vectors, FP64 dots, FP64 residual
//              b - A x (FP32 values, FP64 accumulation) between inner solves
//   kPcgFp32   FP32 matrix and vectors, FP64 dots (reference-like)
#pragma once

#include "ut_runtime.hpp"
#include "contact.hpp"

namespace ut
{

enum PcgPrecision : int
{
	kPcgFp64 = 0,
	kPcgMixed = 1,
	kPcgFp32 = 2,
};
// Krylov recurrence of the iteration (the iterates are the same in exact arithmetic)
enum PcgRecurrence : int
{
	kPcgClassic = 0,		 // precond_cg.cpp's loop: two reductions, three launches per iteration
	kPcgSingleReduction = 1, // Chronopoulos-Gear: one reduction point, two launches per iteration
};
enum PcgPreconditioner : int
{
	kPrecBlockJacobi = 0, // inverse of each 3x3 diagonal block
	kPrecPointJacobi = 1, // make_pcg_jacobi_inverse, precond_cg.cpp:93-111
};

struct PcgParams
{
	int		 precision = kPcgMixed;
	int		 preconditioner = kPrecBlockJacobi;
	uint32_t max_iters = 100;	   // total PCG iterations (all inner solves together)
	double	 rel_tol = 1e-6;	   // ||b - A x|| / ||b||
	double	 inner_tol = 1e-2;	   // mixed: relative tolerance of each inner solve
	uint32_t max_refine = 20;	   // mixed: refinement steps
	double	 regularization = 0.0; // lambda: q += lambda * mean(diag) * p, precond_cg.cpp:286-339
	uint32_t check_every = 8;	   // iterations launched between host convergence checks
	int		 recurrence = kPcgClassic;
};

struct PcgReport
{
	double	 relative_residual = 0.0; // final ||b - A x|| / ||b|| (recomputed in FP64 for mixed)
	double	 infinity_norm = 0.0;	  // max_i |x_i| (block 2-norm)
	uint32_t iterations = 0;
	uint32_t refinements = 0;
	int		 converged = 0;
	int		 failure = 0; // 0 none, 1 non-finite rhs, 2 non-finite during iteration, 3 iteration budget 0
};

struct PcgSolver
{
	Context*		ctx = nullptr;
	uint32_t		n = 0, nslices = 0, nslots = 0;
	const uint32_t* slice_off = nullptr; // [nslices+1] device
	const uint32_t* col = nullptr;		 // [nslots]
	const uint32_t* diag_slot = nullptr; // [n]

	// Optional contact pairs whose off-diagonal Hessian blocks w_i w_j k2 n n^T
	// (i != j) are applied matrix-free in every product with A; their diagonal
	// blocks must already be in the block values.
	const ContactRec* contacts = nullptr;
	const uint32_t*	  contact_count = nullptr;
	uint32_t		  contact_cap = 0;

	// Per-vertex contact lists of the current solve: entry e = 4 c + a is corner a
	// of contact c; entries sorted by vertex (stable, so contact order within a
	// vertex), ct_off[v] .. ct_off[v+1] the entries of vertex v, ct_pos[e] the
	// sorted position of entry e.  The iteration writes each entry's force
	// w_a k2 (S - w_a s_a) n into ct_contrib[ct_pos[e]] and the SpMV sums a
	// vertex's entries in order: no atomics, fixed summation order.
	uint32_t		 ct_entries = 0;
	DevBuf<uint32_t> ct_keys[2], ct_vals[2], ct_hist, ct_off, ct_pos;
	const uint32_t*	 ct_sorted = nullptr; // the entries e in sorted order (into ct_vals)
	DevBuf<double>	 ct_contrib; // [4 * entries] (x, y, z, 0), FP32 or FP64 by the solve's precision
	// Builds the lists for the current contacts.  solve() builds them itself unless
	// the caller did since the contacts last changed (contacts_prepared, cleared
	// by the solve), e.g. to share them with the assembly.
	void			 prepare_contacts();
	bool			 contacts_prepared = false;

	DevBuf<uint8_t> work; // vectors (sized for FP64)
	DevBuf<double>	partials;
	DevBuf<uint32_t> flags; // [0] done, [1] iterations, [2] non-finite
	uint32_t*		h_flags = nullptr; // pinned host mirror
	double*			h_norm = nullptr;

	void init(Context* c, uint32_t n, uint32_t nslices, uint32_t nslots, const uint32_t* slice_off, const uint32_t* col,
		const uint32_t* diag_slot);
	~PcgSolver();
	// vals64 (mixed mode only): FP64 copy of the matrix for the refinement residual,
	// so that the refined solution is that of the FP64 system.
	PcgReport solve(const PcgParams& p, const void* vals, const double* b, double* x, const double* vals64 = nullptr);
	// y = A x (FP64 vectors; vals FP32 when vals_fp32, else FP64)
	void spmv64(const void* vals, bool vals_fp32, const double* x, double* y);
	// Average duration (ms, HIP events on the solver's stream) of `reps`
	// launches of the PCG iteration's SpMV kernel on vals, over the work vectors.
	float profile_spmv(const void* vals, bool vals_fp32, uint32_t reps);
};

} // namespace ut
