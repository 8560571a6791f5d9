// Gradient and Hessian assembly of the incremental potential on gfx950.
//
// Material terms (inertia, BW98 stretch/shear, dihedral bending) follow
// Energies/{soft_inertia_energy,stretch_face_energy,bending_energy_kernel}.cpp;
// contact terms follow fn_perPair_assemble_gradient_hessian
// (narrow_phase.cpp:2619-2746): gradient w_i (k1 n + friction gradient),
// diagonal blocks w_i^2 (k2 n n^T + friction Hessian) with the lagged friction
// values the record carries (friction_math.hpp).  The contacts' off-diagonal
// blocks (w_i w_j times the same 3x3 block) are not stored: the PCG applies
// them matrix-free (pcg.hip), so the block pattern stays the static cloth
// pattern.
//
// Positions are FP32 (the solver's storage) or FP64 (the differentiable
// potential's finite-difference checks): the kernels are templated on the
// position type P.
//
// Everything is evaluated in FP64 and accumulated into the gradient
// (double[3V]) and the planar block values (double[9][nslots]); the
// mixed-precision solve rounds the blocks to FP32 afterwards.  Accumulation is
// either by FP64 atomics, or deterministic (AssembleArgs: colour classes of
// elements with plain read-modify-writes, contacts gathered per vertex), in
// which case every sum has a fixed order and the result is bit-reproducible.
#include "ut_common.hpp"
#include "energy.hpp"
#include "energy_math.hpp"
#include "friction_math.hpp"

namespace ut
{

// p += v: an atomic, or a plain read-modify-write when no other thread of the
// launch touches p (colour classes, one thread per vertex)
template <bool Atomic>
__device__ __forceinline__ void acc(double* p, double v)
{
	if (Atomic)
		atomicAdd(p, v);
	else
		*p += v;
}

template <bool Atomic>
__device__ __forceinline__ void add3(double* g, uint32_t v, const v3<double>& a)
{
	acc<Atomic>(&g[3 * v + 0], a.x);
	acc<Atomic>(&g[3 * v + 1], a.y);
	acc<Atomic>(&g[3 * v + 2], a.z);
}

__global__ void k_zero64(double* p, size_t n)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0.0;
}

// E = m / (2 h^2) |x - x_tilde|^2 (soft_inertia_energy.cpp:25-110)
template <typename P, bool Atomic>
__global__ void k_asm_inertia(uint32_t nv, const P* __restrict__ x, const double* __restrict__ x_tilde,
	const float* __restrict__ mass, double inv_h2, double grad_scale, double hess_scale,
	const uint32_t* __restrict__ diag_slot, uint32_t nslots, double* __restrict__ grad, double* __restrict__ vals,
	double* __restrict__ energy_part)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	double		   e = 0.0;
	if (v < nv)
	{
		const double k = (double)mass[v] * inv_h2;
		double		 d[3];
		for (int a = 0; a < 3; ++a)
		{
			d[a] = (double)x[3 * v + a] - x_tilde[3 * v + a];
			if (grad_scale != 0.0)
				acc<Atomic>(&grad[3 * v + a], grad_scale * k * d[a]);
		}
		e = grad_scale * 0.5 * k * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
		const uint32_t s = diag_slot[v];
		if (hess_scale != 0.0)
		{
			acc<Atomic>(&vals[0 * (size_t)nslots + s], hess_scale * k);
			acc<Atomic>(&vals[4 * (size_t)nslots + s], hess_scale * k);
			acc<Atomic>(&vals[8 * (size_t)nslots + s], hess_scale * k);
		}
	}
	if (energy_part)
	{
		e = wave_sum(e);
		if (__lane_id() == 0)
			atomicAdd(energy_part, e);
	}
}

// Faces order[i] for i in [0, n) (order null: face i); Atomic = false needs
// the n faces to share no vertex.
template <typename P, bool Atomic>
__global__ void k_asm_stretch(uint32_t n, const uint32_t* __restrict__ order, const P* __restrict__ x,
	const uint32_t* __restrict__ faces, const float* __restrict__ dm, const float* __restrict__ area,
	const float* __restrict__ mu_lambda, const uint32_t* __restrict__ face_slots, uint32_t nslots,
	double* __restrict__ grad, double* __restrict__ vals, double* __restrict__ energy_part)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	double		   e = 0.0;
	if (i < n)
	{
		const uint32_t f = order ? order[i] : i;
		const uint32_t vi[3] = { faces[3 * f], faces[3 * f + 1], faces[3 * f + 2] };
		const v3<double> x0 = load_v3<double>(x, vi[0]), x1 = load_v3<double>(x, vi[1]), x2 = load_v3<double>(x, vi[2]);
		const double	 d0 = dm[4 * f], d1 = dm[4 * f + 1], d2 = dm[4 * f + 2], d3 = dm[4 * f + 3];
		const v3<double> r1 = x1 - x0, r2 = x2 - x0;
		const v3<double> Fu = d0 * r1 + d1 * r2, Fv = d2 * r1 + d3 * r2;
		const double	 A = area[f];
		const Bw98		 b = bw98_eval(Fu, Fv, (double)mu_lambda[2 * f], (double)mu_lambda[2 * f + 1]);
		e = A * b.energy;
		const double c[3][2] = { { -(d0 + d1), -(d2 + d3) }, { d0, d2 }, { d1, d3 } };
		for (int a = 0; a < 3; ++a)
		{
			v3<double> g{ 0, 0, 0 };
			for (int k = 0; k < 3; ++k)
				(&g.x)[k] = A * (c[a][0] * b.dedF[k] + c[a][1] * b.dedF[3 + k]);
			add3<Atomic>(grad, vi[a], g);
		}
		for (int a = 0; a < 3; ++a)
			for (int bb = 0; bb < 3; ++bb)
			{
				const uint32_t s = face_slots[9 * f + 3 * a + bb];
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j)
					{
						double h = 0.0;
						for (int p = 0; p < 2; ++p)
							for (int q = 0; q < 2; ++q)
								h += c[a][p] * c[bb][q] * b.H(3 * p + i, 3 * q + j);
						acc<Atomic>(&vals[(size_t)(3 * i + j) * nslots + s], A * h);
					}
			}
	}
	if (energy_part)
	{
		e = wave_sum(e);
		if (__lane_id() == 0)
			atomicAdd(energy_part, e);
	}
}

// E = 1/2 k A (theta - theta0)^2, Gauss-Newton Hessian k A grad grad^T
template <typename P, bool Atomic>
__global__ void k_asm_bending(uint32_t n, const uint32_t* __restrict__ order, const P* __restrict__ x,
	const uint32_t* __restrict__ hinges, const float* __restrict__ rest_angle, const float* __restrict__ stiff_area,
	const uint32_t* __restrict__ hinge_slots, uint32_t nslots, double* __restrict__ grad, double* __restrict__ vals,
	double* __restrict__ energy_part)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	double		   e = 0.0;
	if (t < n)
	{
		const uint32_t h = order ? order[t] : t;
		uint32_t   vi[4];
		v3<double> p[4];
		for (int k = 0; k < 4; ++k)
		{
			vi[k] = hinges[4 * h + k];
			p[k] = load_v3<double>(x, vi[k]);
		}
		v3<double>	 g[4];
		const double theta = dihedral(p[0], p[1], p[2], p[3], g);
		const double ks = stiff_area[h];
		const double dth = theta - (double)rest_angle[h];
		e = 0.5 * ks * dth * dth;
		for (int a = 0; a < 4; ++a)
			add3<Atomic>(grad, vi[a], (ks * dth) * g[a]);
		for (int a = 0; a < 4; ++a)
			for (int b = 0; b < 4; ++b)
			{
				const uint32_t s = hinge_slots[16 * h + 4 * a + b];
				const double   ga[3] = { g[a].x, g[a].y, g[a].z }, gb[3] = { g[b].x, g[b].y, g[b].z };
				for (int i = 0; i < 3; ++i)
					for (int j = 0; j < 3; ++j)
						acc<Atomic>(&vals[(size_t)(3 * i + j) * nslots + s], ks * ga[i] * gb[j]);
			}
	}
	if (energy_part)
	{
		e = wave_sum(e);
		if (__lane_id() == 0)
			atomicAdd(energy_part, e);
	}
}

__global__ void k_asm_contact(const ContactRec* __restrict__ pairs, const uint32_t* __restrict__ count, uint32_t cap,
	const uint32_t* __restrict__ diag_slot, uint32_t nslots, double* __restrict__ grad, double* __restrict__ vals)
{
	const uint32_t n = min(*count, cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const ContactRec c = pairs[i];
		double			 w[4];
		record_weights(c, w);
		const v3<double> g = pair_gradient(c);
		const Sym3		 K = pair_block(c);
		for (int a = 0; a < 4; ++a)
		{
			const uint32_t v = c.idx[a] & kIndexMask;
			add3<true>(grad, v, w[a] * g);
			const uint32_t s = diag_slot[v];
			for (int r = 0; r < 3; ++r)
				for (int q = 0; q < 3; ++q)
					atomicAdd(&vals[(size_t)(3 * r + q) * nslots + s], w[a] * w[a] * sym_at(K, r, q));
		}
	}
}

// The contacts' gradient and diagonal blocks of vertex v, summed over v's entries in
// list order and added once (one thread per vertex, no atomics).
__global__ void k_asm_contact_gather(uint32_t nv, const ContactRec* __restrict__ pairs,
	const uint32_t* __restrict__ ct_off, const uint32_t* __restrict__ entries, const uint32_t* __restrict__ diag_slot,
	uint32_t nslots, double* __restrict__ grad, double* __restrict__ vals)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	const uint32_t b = ct_off[v], e = ct_off[v + 1];
	if (b == e)
		return;
	v3<double> g{ 0, 0, 0 };
	double	   D[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
	for (uint32_t k = b; k < e; ++k)
	{
		const uint32_t	 ent = entries[k];
		const ContactRec c = pairs[ent >> 2];
		const uint32_t	 a = ent & 3u;
		double			 w[4];
		record_weights(c, w);
		g = g + w[a] * pair_gradient(c);
		const Sym3 K = pair_block(c);
		for (int r = 0; r < 3; ++r)
			for (int q = 0; q < 3; ++q)
				D[3 * r + q] += w[a] * w[a] * sym_at(K, r, q);
	}
	grad[3 * v + 0] += g.x;
	grad[3 * v + 1] += g.y;
	grad[3 * v + 2] += g.z;
	const uint32_t s = diag_slot[v];
	for (int c = 0; c < 9; ++c)
		vals[(size_t)c * nslots + s] += D[c];
}

__global__ void k_round_to32(size_t n, const double* __restrict__ in, float* __restrict__ out)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = (float)in[i];
}

template <typename P>
static void material_terms(hipStream_t s, const AssembleArgs& a, const P* x)
{
	const bool det = a.face_order && a.face_color_off && a.hinge_order && a.hinge_color_off;
	if (a.mass && a.x_tilde)
	{
		if (det)
			k_asm_inertia<P, false><<<div_up(a.nv, 256), 256, 0, s>>>(a.nv, x, a.x_tilde, a.mass, a.inv_h2,
				a.inertia_grad_scale, a.inertia_hess_scale, a.diag_slot, a.nslots, a.grad, a.vals,
				a.energy ? a.energy + 0 : nullptr);
		else
			k_asm_inertia<P, true><<<div_up(a.nv, 256), 256, 0, s>>>(a.nv, x, a.x_tilde, a.mass, a.inv_h2,
				a.inertia_grad_scale, a.inertia_hess_scale, a.diag_slot, a.nslots, a.grad, a.vals,
				a.energy ? a.energy + 0 : nullptr);
	}
	if (a.nf && a.faces && a.dm)
	{
		if (det)
			for (uint32_t c = 0; c < a.face_colors; ++c)
			{
				const uint32_t b = a.face_color_off[c], n = a.face_color_off[c + 1] - b;
				if (n)
					k_asm_stretch<P, false><<<div_up(n, 256), 256, 0, s>>>(n, a.face_order + b, x, a.faces, a.dm,
						a.face_area, a.mu_lambda, a.face_slots, a.nslots, a.grad, a.vals, a.energy ? a.energy + 1 : nullptr);
			}
		else
			k_asm_stretch<P, true><<<div_up(a.nf, 256), 256, 0, s>>>(a.nf, nullptr, x, a.faces, a.dm, a.face_area,
				a.mu_lambda, a.face_slots, a.nslots, a.grad, a.vals, a.energy ? a.energy + 1 : nullptr);
	}
	if (a.nh && a.hinges)
	{
		if (det)
			for (uint32_t c = 0; c < a.hinge_colors; ++c)
			{
				const uint32_t b = a.hinge_color_off[c], n = a.hinge_color_off[c + 1] - b;
				if (n)
					k_asm_bending<P, false><<<div_up(n, 256), 256, 0, s>>>(n, a.hinge_order + b, x, a.hinges,
						a.rest_angle, a.bend_stiff_area, a.hinge_slots, a.nslots, a.grad, a.vals,
						a.energy ? a.energy + 2 : nullptr);
			}
		else
			k_asm_bending<P, true><<<div_up(a.nh, 256), 256, 0, s>>>(a.nh, nullptr, x, a.hinges, a.rest_angle,
				a.bend_stiff_area, a.hinge_slots, a.nslots, a.grad, a.vals, a.energy ? a.energy + 2 : nullptr);
	}
}

void assemble(Context* ctx, const AssembleArgs& a)
{
	UT_REQUIRE((a.x || a.x64) && a.grad && a.vals && a.diag_slot, "assemble: null input");
	hipStream_t	   s = ctx->stream;
	const size_t   nv3 = 3 * (size_t)a.nv, nval = 9 * (size_t)a.nslots;
	k_zero64<<<div_up((uint32_t)nv3, 256), 256, 0, s>>>(a.grad, nv3);
	k_zero64<<<(uint32_t)((nval + 255) / 256), 256, 0, s>>>(a.vals, nval);
	if (a.energy)
		k_zero64<<<1, 64, 0, s>>>(a.energy, 4);
	if (a.x64)
		material_terms(s, a, a.x64);
	else
		material_terms(s, a, a.x);
	if (a.contacts && a.contact_count && a.contact_cap)
	{
		if (a.ct_off && a.ct_entries)
			k_asm_contact_gather<<<div_up(a.nv, 256), 256, 0, s>>>(a.nv, a.contacts, a.ct_off, a.ct_entries, a.diag_slot,
				a.nslots, a.grad, a.vals);
		else
			k_asm_contact<<<1024, 256, 0, s>>>(a.contacts, a.contact_count, a.contact_cap, a.diag_slot, a.nslots, a.grad,
				a.vals);
	}
	if (a.vals32)
		k_round_to32<<<(uint32_t)((nval + 255) / 256), 256, 0, s>>>(nval, a.vals, a.vals32);
	UT_LAUNCH_CHECK();
}

bool color_elements(const uint32_t* elem, uint32_t n, int k, uint32_t nv, std::vector<uint32_t>& order,
	std::vector<uint32_t>& color_off)
{
	constexpr int		  kMaxColors = 256, kWords = kMaxColors / 64;
	std::vector<uint64_t> used((size_t)nv * kWords, 0ull); // colours taken at each vertex
	std::vector<uint8_t>  color(n);
	std::vector<uint32_t> count(kMaxColors, 0u);
	uint32_t			  ncol = 0;
	for (uint32_t i = 0; i < n; ++i)
	{
		uint64_t m[kWords] = {};
		for (int j = 0; j < k; ++j)
		{
			const uint32_t v = elem[(size_t)k * i + j];
			if (v >= nv)
				return false;
			for (int w = 0; w < kWords; ++w)
				m[w] |= used[(size_t)v * kWords + w];
		}
		int c = -1;
		for (int w = 0; w < kWords && c < 0; ++w)
			if (~m[w])
				c = 64 * w + __builtin_ctzll(~m[w]);
		if (c < 0)
			return false;
		color[i] = (uint8_t)c;
		count[c]++;
		ncol = std::max(ncol, (uint32_t)c + 1);
		for (int j = 0; j < k; ++j)
			used[(size_t)elem[(size_t)k * i + j] * kWords + c / 64] |= 1ull << (c % 64);
	}
	color_off.assign(ncol + 1, 0u);
	for (uint32_t c = 0; c < ncol; ++c)
		color_off[c + 1] = color_off[c] + count[c];
	order.resize(n);
	std::vector<uint32_t> fill(color_off.begin(), color_off.end() - 1);
	for (uint32_t i = 0; i < n; ++i)
		order[fill[color[i]]++] = i;
	return true;
}

} // namespace ut
