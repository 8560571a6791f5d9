// Affine bodies on gfx950: q-to-x, the q-space gradient and 12x12 body blocks
// (inertia, orthogonality, contacts), the matrix-free operator, and a
// block-Jacobi PCG that runs whole in one workgroup.
//
//   inertia        E = sum_ij M4[i][j] (q_i - qt_i).(q_j - qt_j) / (2 h^2)
//                  (Energies/abd_inertia_energy.cpp:25-140; M4 the compressed
//                  mass matrix of init_sim_data.cpp:960-1024)
//   orthogonality  E = kappa V sum_ij (a_i.a_j - delta_ij)^2 on the columns of A,
//                  gradient 4 kappa V (sum_j (a_i.a_j) a_j - a_i), Hessian blocks
//                  4 kappa V [a_i a_i^T + (a_i.a_i - 1) I + sum_k a_k a_k^T] and
//                  4 kappa V [a_j a_i^T + (a_i.a_j) I] (Energies/abd_ortho_energy.cpp:25-140)
//   contacts       a record's side s (its vertex, or its face / edge) moves with
//                  one body; its weighted point sum_k w_k x_k is W_s q_b with the
//                  4-vector W_s = sum_k w_k (1, xbar_k) (add_to_rigid_body,
//                  narrow_phase.cpp:2679-2698), so the record adds W_s (x) g to
//                  the gradient and W_s W_t^T (x) K to the blocks (s, t the sides)
//
// The 12 x nb system is small (6,144 unknowns for 512 bodies): one workgroup of
// 1024 threads runs every PCG iteration of a solve with the direction, residual
// and product vectors in LDS and the contacts between bodies applied from the
// records each iteration; no host round trip until the solve ends.
#include "ut_common.hpp"
#include "abd.hpp"
#include "friction_math.hpp"
#include "pcg.hpp"

#include <cmath>
#include <vector>

namespace ut
{
namespace
{

struct Side
{
	uint32_t body;
	double	 w[4];
};

// the two sides of a record: VF (vertex | face), EE (edge | edge)
__device__ __forceinline__ void record_sides(const ContactRec& c, const uint32_t* __restrict__ body,
	const float* __restrict__ xbar, Side& A, Side& B)
{
	double w[4];
	record_weights(c, w);
	const int na = (c.idx[2] & kTypeBit) ? 1 : 2;
	for (int k = 0; k < 4; ++k)
		A.w[k] = B.w[k] = 0.0;
	A.body = B.body = 0;
	for (int k = 0; k < 4; ++k)
	{
		const uint32_t v = c.idx[k] & kIndexMask;
		Side&		   s = k < na ? A : B;
		s.body = body[v];
		s.w[0] += w[k];
		s.w[1] += w[k] * (double)xbar[3 * v];
		s.w[2] += w[k] * (double)xbar[3 * v + 1];
		s.w[3] += w[k] * (double)xbar[3 * v + 2];
	}
}

template <typename P>
__global__ void k_abd_q_to_x(uint32_t nv, const uint32_t* __restrict__ body, const float* __restrict__ xbar,
	const double* __restrict__ q, P* __restrict__ x)
{
	const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
	if (v >= nv)
		return;
	const double* qb = q + 12 * (size_t)body[v];
	const double  u = xbar[3 * v], w = xbar[3 * v + 1], s = xbar[3 * v + 2];
	for (int r = 0; r < 3; ++r)
		x[3 * v + r] = (P)(qb[r] + qb[3 + r] * u + qb[6 + r] * w + qb[9 + r] * s);
}

// per body: inertia + orthogonality energy, gradient and 12x12 block (plain stores)
__global__ void k_abd_bodies(uint32_t nb, const double* __restrict__ mass4, const double* __restrict__ volume,
	const double* __restrict__ kappa, AbdAssembleArgs a, double* __restrict__ grad, double* __restrict__ blocks,
	double* __restrict__ energy)
{
	const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
	double		   e_in = 0.0, e_or = 0.0;
	if (b < nb)
	{
		const double* q = a.q + 12 * (size_t)b;
		double		  g[12] = {};
		double		  H[144] = {};
		if (a.q_tilde)
		{
			const double* qt = a.q_tilde + 12 * (size_t)b;
			const double* M = mass4 + 16 * (size_t)b;
			double		  d[12];
			for (int k = 0; k < 12; ++k)
				d[k] = q[k] - qt[k];
			for (int i = 0; i < 4; ++i)
				for (int j = 0; j < 4; ++j)
				{
					const double m = M[4 * i + j] * a.inv_h2;
					double		 dd = 0.0;
					for (int r = 0; r < 3; ++r)
					{
						g[3 * i + r] += a.inertia_grad_scale * m * d[3 * j + r];
						dd += d[3 * i + r] * d[3 * j + r];
						H[12 * (3 * i + r) + 3 * j + r] += a.inertia_hess_scale * m;
					}
					e_in += a.inertia_grad_scale * 0.5 * m * dd;
				}
		}
		const double s = kappa[b] * volume[b];
		const double* A = q + 3; // columns a_0, a_1, a_2
		double		  dotm[3][3];
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
				dotm[i][j] = A[3 * i] * A[3 * j] + A[3 * i + 1] * A[3 * j + 1] + A[3 * i + 2] * A[3 * j + 2];
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
			{
				const double t = dotm[i][j] - (i == j ? 1.0 : 0.0);
				e_or += s * t * t;
			}
		for (int i = 0; i < 3; ++i)
			for (int r = 0; r < 3; ++r)
			{
				double gi = -A[3 * i + r];
				for (int j = 0; j < 3; ++j)
					gi += dotm[i][j] * A[3 * j + r];
				g[3 + 3 * i + r] += 4.0 * s * gi;
			}
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
				for (int r = 0; r < 3; ++r)
					for (int c = 0; c < 3; ++c)
					{
						double h;
						if (i == j)
						{
							h = A[3 * i + r] * A[3 * i + c] + (r == c ? dotm[i][i] - 1.0 : 0.0);
							for (int k = 0; k < 3; ++k)
								h += A[3 * k + r] * A[3 * k + c];
						}
						else
							h = A[3 * j + r] * A[3 * i + c] + (r == c ? dotm[i][j] : 0.0);
						H[12 * (3 + 3 * i + r) + 3 + 3 * j + c] += 4.0 * s * h;
					}
		for (int k = 0; k < 12; ++k)
			grad[12 * (size_t)b + k] = g[k];
		if (blocks)
			for (int k = 0; k < 144; ++k)
				blocks[144 * (size_t)b + k] = H[k];
	}
	if (energy)
	{
		e_in = wave_sum(e_in);
		e_or = wave_sum(e_or);
		if (__lane_id() == 0)
		{
			atomicAdd(&energy[0], e_in);
			atomicAdd(&energy[1], e_or);
		}
	}
}

__device__ __forceinline__ void add_side_block(double* __restrict__ blocks, uint32_t b, const double* wa,
	const double* wb, const Sym3& K)
{
	double* H = blocks + 144 * (size_t)b;
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j)
		{
			const double ww = wa[i] * wb[j];
			if (ww == 0.0)
				continue;
			for (int r = 0; r < 3; ++r)
				for (int c = 0; c < 3; ++c)
					atomicAdd(&H[12 * (3 * i + r) + 3 * j + c], ww * sym_at(K, r, c));
		}
}

__global__ void k_abd_contacts(const ContactRec* __restrict__ pairs, const uint32_t* __restrict__ count, uint32_t cap,
	const uint32_t* __restrict__ body, const float* __restrict__ xbar, double* __restrict__ grad,
	double* __restrict__ blocks)
{
	const uint32_t n = min(*count, cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
	{
		const ContactRec c = pairs[i];
		Side			 A, B;
		record_sides(c, body, xbar, A, B);
		if (A.body == B.body)
		{
			for (int k = 0; k < 4; ++k)
				A.w[k] += B.w[k], B.w[k] = 0.0;
		}
		const v3<double> g = pair_gradient(c);
		for (int s = 0; s < 2; ++s)
		{
			const Side& S = s == 0 ? A : B;
			if (s == 1 && A.body == B.body)
				break;
			for (int k = 0; k < 4; ++k)
			{
				atomicAdd(&grad[12 * (size_t)S.body + 3 * k + 0], S.w[k] * g.x);
				atomicAdd(&grad[12 * (size_t)S.body + 3 * k + 1], S.w[k] * g.y);
				atomicAdd(&grad[12 * (size_t)S.body + 3 * k + 2], S.w[k] * g.z);
			}
		}
		if (blocks)
		{
			const Sym3 K = pair_block(c);
			add_side_block(blocks, A.body, A.w, A.w, K);
			if (A.body != B.body)
				add_side_block(blocks, B.body, B.w, B.w, K);
		}
	}
}

__global__ void k_abd_apply_blocks(uint32_t nb, const double* __restrict__ blocks, const double* __restrict__ p,
	double* __restrict__ y)
{
	const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
	if (row >= 12 * nb)
		return;
	const uint32_t b = row / 12, r = row % 12;
	const double*  H = blocks + 144 * (size_t)b + 12 * r;
	const double*  pb = p + 12 * (size_t)b;
	double		   s = 0.0;
	for (int k = 0; k < 12; ++k)
		s += H[k] * pb[k];
	y[row] = s;
}

// y_A += W_A (x) K u_B and y_B += W_B (x) K u_A, u_s = sum_k W_s[k] p_{s,k}, for records between two bodies
template <typename T, typename Y>
__device__ __forceinline__ void cross_product(const ContactRec& c, const uint32_t* __restrict__ body,
	const float* __restrict__ xbar, const T* __restrict__ p, Y* __restrict__ y)
{
	Side A, B;
	record_sides(c, body, xbar, A, B);
	if (A.body == B.body)
		return;
	const Sym3 K = pair_block(c);
	v3<double> u[2];
	for (int s = 0; s < 2; ++s)
	{
		const Side& S = s == 0 ? A : B;
		v3<double>	acc{ 0, 0, 0 };
		for (int k = 0; k < 4; ++k)
		{
			const T* pk = p + 12 * (size_t)S.body + 3 * k;
			acc = acc + S.w[k] * v3<double>{ (double)pk[0], (double)pk[1], (double)pk[2] };
		}
		u[s] = acc;
	}
	const v3<double> KuB = sym_mul(K, u[1]), KuA = sym_mul(K, u[0]);
	for (int k = 0; k < 4; ++k)
	{
		atomicAdd(&y[12 * (size_t)A.body + 3 * k + 0], (Y)(A.w[k] * KuB.x));
		atomicAdd(&y[12 * (size_t)A.body + 3 * k + 1], (Y)(A.w[k] * KuB.y));
		atomicAdd(&y[12 * (size_t)A.body + 3 * k + 2], (Y)(A.w[k] * KuB.z));
		atomicAdd(&y[12 * (size_t)B.body + 3 * k + 0], (Y)(B.w[k] * KuA.x));
		atomicAdd(&y[12 * (size_t)B.body + 3 * k + 1], (Y)(B.w[k] * KuA.y));
		atomicAdd(&y[12 * (size_t)B.body + 3 * k + 2], (Y)(B.w[k] * KuA.z));
	}
}

__global__ void k_abd_apply_cross(const ContactRec* __restrict__ pairs, const uint32_t* __restrict__ count,
	uint32_t cap, const uint32_t* __restrict__ body, const float* __restrict__ xbar, const double* __restrict__ p,
	double* __restrict__ y)
{
	const uint32_t n = min(*count, cap);
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
		cross_product<double, double>(pairs[i], body, xbar, p, y);
}

// inverse of every body's 12x12 block (Gauss-Jordan with partial pivoting, FP64);
// a singular block falls back to the inverse of its diagonal
__global__ void k_abd_block_inverse(uint32_t nb, const double* __restrict__ blocks, double* __restrict__ minv64,
	float* __restrict__ minv32, float* __restrict__ blocks32)
{
	const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= nb)
		return;
	double		  M[12][24];
	const double* H = blocks + 144 * (size_t)b;
	for (int r = 0; r < 12; ++r)
		for (int c = 0; c < 12; ++c)
		{
			M[r][c] = H[12 * r + c];
			M[r][12 + c] = r == c ? 1.0 : 0.0;
		}
	bool ok = true;
	for (int c = 0; c < 12 && ok; ++c)
	{
		int	   piv = c;
		double best = fabs(M[c][c]);
		for (int r = c + 1; r < 12; ++r)
			if (fabs(M[r][c]) > best)
				best = fabs(M[r][c]), piv = r;
		if (!(best > 1e-300))
		{
			ok = false;
			break;
		}
		if (piv != c)
			for (int k = 0; k < 24; ++k)
			{
				const double t = M[c][k];
				M[c][k] = M[piv][k];
				M[piv][k] = t;
			}
		const double inv = 1.0 / M[c][c];
		for (int k = 0; k < 24; ++k)
			M[c][k] *= inv;
		for (int r = 0; r < 12; ++r)
			if (r != c && M[r][c] != 0.0)
			{
				const double f = M[r][c];
				for (int k = 0; k < 24; ++k)
					M[r][k] -= f * M[c][k];
			}
	}
	double* out = minv64 + 144 * (size_t)b;
	for (int r = 0; r < 12; ++r)
		for (int c = 0; c < 12; ++c)
		{
			const double d = H[12 * r + r];
			const double v = ok ? M[r][12 + c] : (r == c && d > 0.0 ? 1.0 / d : 0.0);
			out[12 * r + c] = v;
			if (minv32)
				minv32[144 * (size_t)b + 12 * r + c] = (float)v;
			if (blocks32)
				blocks32[144 * (size_t)b + 12 * r + c] = (float)H[12 * r + c];
		}
}

constexpr int kPcgThreads = 1024;

template <typename T>
__device__ __forceinline__ double block_sum2(double a, double b, double* red, double* out_b)
{
	a = wave_sum(a);
	b = wave_sum(b);
	const int w = threadIdx.x / kWave;
	__syncthreads();
	if (__lane_id() == 0)
		red[w] = a, red[16 + w] = b;
	__syncthreads();
	double sa = 0.0, sb = 0.0;
	for (int k = 0; k < kPcgThreads / kWave; ++k)
		sa += red[k], sb += red[16 + k];
	*out_b = sb;
	return sa;
}

// One block-Jacobi PCG solve of A x = b in one workgroup; T is the arithmetic of
// the vectors, products and preconditioner (dots in FP64).  LDS holds p, r and
// the product y; x and z stay with the thread that owns their rows.
// status[0] iterations, [1] relative residual ||r|| / ||b||, [2] converged, [3] failure.
template <typename T>
__global__ __launch_bounds__(kPcgThreads) void k_abd_pcg(uint32_t nb, const T* __restrict__ blocks,
	const T* __restrict__ minv, const ContactRec* __restrict__ pairs, const uint32_t* __restrict__ count,
	uint32_t cap, const uint32_t* __restrict__ body, const float* __restrict__ xbar, const double* __restrict__ b,
	T* __restrict__ x, T* __restrict__ z, uint32_t max_iters, double tol2, double* __restrict__ status)
{
	extern __shared__ unsigned char smem[];
	const uint32_t n = 12 * nb;
	T*			   p = reinterpret_cast<T*>(smem);
	T*			   r = p + n;
	T*			   y = r + n;
	__shared__ double red[32];
	const uint32_t nc = (count && cap) ? min(*count, cap) : 0u;
	const uint32_t t = threadIdx.x;
	double		   bb_l = 0.0;
	for (uint32_t i = t; i < n; i += kPcgThreads)
	{
		r[i] = (T)b[i];
		x[i] = (T)0;
		bb_l += b[i] * b[i];
	}
	__syncthreads();
	double rz_l = 0.0, dummy;
	for (uint32_t i = t; i < n; i += kPcgThreads)
	{
		const uint32_t bo = i / 12, row = i % 12;
		const T*	   M = minv + 144 * (size_t)bo + 12 * row;
		double		   s = 0.0;
		for (int k = 0; k < 12; ++k)
			s += (double)M[k] * (double)r[12 * bo + k];
		z[i] = (T)s;
		rz_l += (double)r[i] * s;
	}
	double		   bb;
	double		   rz = block_sum2<T>(rz_l, bb_l, red, &bb);
	for (uint32_t i = t; i < n; i += kPcgThreads)
		p[i] = z[i];
	uint32_t it = 0;
	int		 conv = bb == 0.0, fail = !isfinite(bb) ? 1 : 0;
	double	 rr = bb;
	while (!conv && !fail && it < max_iters)
	{
		__syncthreads();
		// y = A p: blocks, then the contacts between bodies
		for (uint32_t i = t; i < n; i += kPcgThreads)
		{
			const uint32_t bo = i / 12, row = i % 12;
			const T*	   H = blocks + 144 * (size_t)bo + 12 * row;
			double		   s = 0.0;
			for (int k = 0; k < 12; ++k)
				s += (double)H[k] * (double)p[12 * bo + k];
			y[i] = (T)s;
		}
		__syncthreads();
		for (uint32_t i = t; i < nc; i += kPcgThreads)
			cross_product<T, T>(pairs[i], body, xbar, p, y);
		__syncthreads();
		double pq_l = 0.0;
		for (uint32_t i = t; i < n; i += kPcgThreads)
			pq_l += (double)p[i] * (double)y[i];
		const double pq = block_sum2<T>(pq_l, 0.0, red, &dummy);
		const double alpha = rz / pq;
		if (!(pq > 0.0) || !isfinite(alpha))
		{
			fail = 2;
			break;
		}
		for (uint32_t i = t; i < n; i += kPcgThreads)
		{
			x[i] = (T)((double)x[i] + alpha * (double)p[i]);
			r[i] = (T)((double)r[i] - alpha * (double)y[i]);
		}
		__syncthreads();
		double rzn_l = 0.0, rr_l = 0.0;
		for (uint32_t i = t; i < n; i += kPcgThreads)
		{
			const uint32_t bo = i / 12, row = i % 12;
			const T*	   M = minv + 144 * (size_t)bo + 12 * row;
			double		   s = 0.0;
			for (int k = 0; k < 12; ++k)
				s += (double)M[k] * (double)r[12 * bo + k];
			z[i] = (T)s;
			rzn_l += (double)r[i] * s;
			rr_l += (double)r[i] * (double)r[i];
		}
		const double rzn = block_sum2<T>(rzn_l, rr_l, red, &rr);
		++it;
		if (!isfinite(rr))
		{
			fail = 2;
			break;
		}
		if (rr <= tol2 * bb)
		{
			conv = 1;
			break;
		}
		const double beta = rzn / rz;
		rz = rzn;
		__syncthreads();
		for (uint32_t i = t; i < n; i += kPcgThreads)
			p[i] = (T)((double)z[i] + beta * (double)p[i]);
	}
	if (t == 0)
	{
		status[0] = it;
		status[1] = bb > 0.0 ? sqrt(rr / bb) : 0.0;
		status[2] = conv;
		status[3] = fail;
	}
}

__global__ void k_abd_residual(uint32_t n, const double* __restrict__ b, const double* __restrict__ ax,
	double* __restrict__ r, double* __restrict__ out)
{
	__shared__ double red[32];
	double			  rr = 0.0, bb = 0.0;
	for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
	{
		r[i] = b[i] - ax[i];
		rr += r[i] * r[i];
		bb += b[i] * b[i];
	}
	rr = wave_sum(rr);
	bb = wave_sum(bb);
	if (__lane_id() == 0)
		red[threadIdx.x / kWave] = rr, red[16 + threadIdx.x / kWave] = bb;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		double a = 0.0, c = 0.0;
		for (uint32_t k = 0; k < blockDim.x / kWave; ++k)
			a += red[k], c += red[16 + k];
		out[0] = a, out[1] = c;
	}
}

__global__ void k_abd_axpy(uint32_t n, const float* __restrict__ d, double* __restrict__ x)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		x[i] += (double)d[i];
}

// J^T g: per body, sum over its vertices of (1, xbar_v) (x) g_v (one workgroup per
// body, vertices [vert_off[b], vert_off[b + 1]), fixed-order wave and block sums)
__global__ __launch_bounds__(256) void k_abd_gather(const uint32_t* __restrict__ vert_off,
	const float* __restrict__ xbar, const double* __restrict__ gx, double* __restrict__ gq, int accumulate)
{
	const uint32_t b = blockIdx.x;
	const uint32_t v0 = vert_off[b], v1 = vert_off[b + 1];
	double		   acc[12] = {};
	for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x)
	{
		const double w[4] = { 1.0, (double)xbar[3 * v], (double)xbar[3 * v + 1], (double)xbar[3 * v + 2] };
		for (int k = 0; k < 4; ++k)
			for (int r = 0; r < 3; ++r)
				acc[3 * k + r] += w[k] * gx[3 * (size_t)v + r];
	}
	__shared__ double part[4][12];
	for (int k = 0; k < 12; ++k)
	{
		const double t = wave_sum(acc[k]);
		if (__lane_id() == 0)
			part[threadIdx.x / kWave][k] = t;
	}
	__syncthreads();
	if (threadIdx.x < 12)
	{
		const double t = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] + part[3][threadIdx.x];
		double&		 o = gq[12 * (size_t)b + threadIdx.x];
		o = accumulate ? o + t : t;
	}
}

__global__ void k_zero_d(size_t n, double* p)
{
	const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = 0.0;
}

} // namespace

void abd_q_to_x(Context* ctx, const AbdDesc& d, const double* q, float* x32, double* x64)
{
	UT_REQUIRE(d.body && d.xbar && q && (x32 || x64), "abd q_to_x: null input");
	if (x32)
		k_abd_q_to_x<float><<<div_up(d.nv, 256), 256, 0, ctx->stream>>>(d.nv, d.body, d.xbar, q, x32);
	if (x64)
		k_abd_q_to_x<double><<<div_up(d.nv, 256), 256, 0, ctx->stream>>>(d.nv, d.body, d.xbar, q, x64);
	UT_LAUNCH_CHECK();
}

void abd_assemble(Context* ctx, const AbdDesc& d, const AbdAssembleArgs& a)
{
	UT_REQUIRE(d.nb && d.mass4 && d.volume && d.kappa && a.q && a.grad, "abd assemble: null input");
	hipStream_t s = ctx->stream;
	if (a.energy)
		k_zero_d<<<1, 64, 0, s>>>(2, a.energy);
	k_abd_bodies<<<div_up(d.nb, 64), 64, 0, s>>>(d.nb, d.mass4, d.volume, d.kappa, a, a.grad, a.blocks, a.energy);
	if (a.contacts && a.contact_count && a.contact_cap)
		k_abd_contacts<<<512, 256, 0, s>>>(a.contacts, a.contact_count, a.contact_cap, d.body, d.xbar, a.grad,
			a.blocks);
	UT_LAUNCH_CHECK();
}

void abd_apply(Context* ctx, const AbdDesc& d, const double* blocks, const ContactRec* contacts, const uint32_t* count,
	uint32_t cap, const double* p, double* y)
{
	hipStream_t s = ctx->stream;
	k_abd_apply_blocks<<<div_up(12 * d.nb, 256), 256, 0, s>>>(d.nb, blocks, p, y);
	if (contacts && count && cap)
		k_abd_apply_cross<<<512, 256, 0, s>>>(contacts, count, cap, d.body, d.xbar, p, y);
	UT_LAUNCH_CHECK();
}

void abd_gather(Context* ctx, const AbdDesc& d, const double* gx, double* gq, bool accumulate)
{
	UT_REQUIRE(d.vert_off && d.xbar && gx && gq, "abd gather: null input");
	k_abd_gather<<<d.nb, 256, 0, ctx->stream>>>(d.vert_off, d.xbar, gx, gq, accumulate ? 1 : 0);
	UT_LAUNCH_CHECK();
}

void AbdPcg::init(Context* c, uint32_t num_bodies)
{
	ctx = c;
	nb = num_bodies;
	UT_REQUIRE(nb > 0 && 3 * 12 * (size_t)nb * sizeof(double) <= 150 * 1024,
		"abd pcg: at most 1600 bodies (the solve keeps three vectors in one workgroup's LDS)");
	minv64.reserve(144 * (size_t)nb);
	blocks32.reserve(144 * (size_t)nb);
	minv32.reserve(144 * (size_t)nb);
	work64.reserve(4 * 12 * (size_t)nb + 8);
	work32.reserve(2 * 12 * (size_t)nb);
	status.reserve(8);
	static bool attr = false;
	if (!attr)
	{
		UT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_abd_pcg<double>),
			hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
		UT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_abd_pcg<float>),
			hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
		attr = true;
	}
}

AbdPcgReport AbdPcg::solve(const AbdDesc& d, const double* blocks, const ContactRec* contacts, const uint32_t* count,
	uint32_t cap, const double* b, double* x, int precision, uint32_t max_iters, double rel_tol)
{
	UT_REQUIRE(d.nb == nb, "abd pcg: body count changed");
	hipStream_t	   s = ctx->stream;
	const uint32_t n = 12 * nb;
	AbdPcgReport   rep;
	// without contacts the kernels read no count (null pointer, capacity 0)
	const uint32_t* cnt = contacts ? count : nullptr;
	const uint32_t	cap_eff = contacts && count ? cap : 0u;
	k_abd_block_inverse<<<div_up(nb, 64), 64, 0, s>>>(nb, blocks, minv64.get(), minv32.get(), blocks32.get());
	UT_LAUNCH_CHECK();
	double h[4];
	if (precision != kPcgMixed)
	{
		k_abd_pcg<double><<<1, kPcgThreads, 3 * n * sizeof(double), s>>>(nb, blocks, minv64.get(), contacts, cnt,
			cap_eff, d.body, d.xbar, b, x, work64.get(), max_iters, rel_tol * rel_tol, status.get());
		UT_LAUNCH_CHECK();
		UT_HIP_CHECK(hipMemcpyAsync(h, status.get(), sizeof(h), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		rep.iterations = (uint32_t)h[0];
		rep.relative_residual = h[1];
		rep.converged = h[2] != 0.0;
		rep.failure = (int)h[3];
		return rep;
	}
	// FP32 inner solves, FP64 residual b - A x between them
	double* r64 = work64.get();
	double* ax = r64 + n;
	double* sums = ax + 2 * n;
	float*	d32 = work32.get();
	float*	z32 = d32 + n;
	k_zero_d<<<div_up(n, 256), 256, 0, s>>>(n, x);
	for (int ref = 0; ref < 30 && rep.iterations < max_iters; ++ref)
	{
		abd_apply(ctx, d, blocks, contacts, cnt, cap_eff, x, ax);
		k_abd_residual<<<1, 1024, 0, s>>>(n, b, ax, r64, sums);
		double rs[2];
		UT_HIP_CHECK(hipMemcpyAsync(rs, sums, sizeof(rs), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		rep.relative_residual = rs[1] > 0.0 ? std::sqrt(rs[0] / rs[1]) : 0.0;
		if (!std::isfinite(rs[0]))
		{
			rep.failure = 2;
			break;
		}
		if (rs[1] == 0.0 || rep.relative_residual <= rel_tol)
		{
			rep.converged = 1;
			break;
		}
		// inner tolerance: enough to gain a factor 1e-3, not below what FP32 resolves
		const double inner = std::max(1e-3, 0.5 * rel_tol / rep.relative_residual);
		k_abd_pcg<float><<<1, kPcgThreads, 3 * n * sizeof(float), s>>>(nb, blocks32.get(), minv32.get(), contacts,
			cnt, cap_eff, d.body, d.xbar, r64, d32, z32, max_iters - rep.iterations, inner * inner, status.get());
		k_abd_axpy<<<div_up(n, 256), 256, 0, s>>>(n, d32, x);
		UT_LAUNCH_CHECK();
		UT_HIP_CHECK(hipMemcpyAsync(h, status.get(), sizeof(h), hipMemcpyDeviceToHost, s));
		UT_HIP_CHECK(hipStreamSynchronize(s));
		rep.iterations += (uint32_t)h[0];
		rep.refinements++;
		if (h[3] != 0.0)
		{
			rep.failure = (int)h[3];
			break;
		}
	}
	return rep;
}

} // namespace ut
