/*
 * oracle_pcg.c -- FP64 restatement of ConjugateGradientSolver::host_solve_core
 * (Solver/LinearSolver/precond_cg.cpp:367-693) on a 3x3-block matrix given in
 * block-CSR form.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Pinning (tests/test_oracle.py): a direct FP64 solve (scipy) of small SPD
 * systems; a block-diagonal system solved exactly in one iteration by the
 * block-Jacobi preconditioner; zero right-hand side -> x = 0, converged;
 * iteration budget 0 -> skipped; non-finite right-hand side -> failure.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int or_nthreads = 1;
void or_set_threads(int n) { or_nthreads = n < 1 ? 1 : n; }

static const uint32_t* g_contacts = NULL;
static uint32_t		   g_ncontacts = 0;

/* y += w_a K (P - w_a x_a) for contact c at each of its vertices a */
static void contact_apply(const uint32_t* r, const double* x, double* y)
{
	double	 w[4], K[9], P[3] = { 0, 0, 0 };
	uint32_t v[4];
	or_record_weights(r, w);
	or_pair_block(r, K, NULL);
	for (int a = 0; a < 4; ++a)
	{
		v[a] = r[a] & 0x7FFFFFFFu;
		for (int k = 0; k < 3; ++k)
			P[k] += w[a] * x[3 * v[a] + k];
	}
	for (int a = 0; a < 4; ++a)
	{
		double q[3];
		for (int k = 0; k < 3; ++k)
			q[k] = P[k] - w[a] * x[3 * v[a] + k];
		for (int k = 0; k < 3; ++k)
			y[3 * v[a] + k] += w[a] * (K[3 * k] * q[0] + K[3 * k + 1] * q[1] + K[3 * k + 2] * q[2]);
	}
}

/* y += off-diagonal contact blocks w_i w_j K (i != j) times x, K the pair's block
 * (k2 n n^T + lagged friction Hessian, or_pair_block).  With several threads each
 * sums its share of the contacts into a private vector (then added in thread order). */
static void contact_mv(uint32_t n, const double* x, double* y)
{
	if (or_nthreads <= 1 || g_ncontacts < 4096)
	{
		for (uint32_t c = 0; c < g_ncontacts; ++c)
			contact_apply(g_contacts + 16 * c, x, y);
		return;
	}
	const size_t m = 3 * (size_t)n;
	const int	 T = or_nthreads;
	double*		 buf = (double*)calloc((size_t)T * m, sizeof(double));
#pragma omp parallel num_threads(T)
	{
		const int t = omp_get_thread_num();
		double*	  yb = buf + (size_t)t * m;
#pragma omp for schedule(static)
		for (uint32_t c = 0; c < g_ncontacts; ++c)
			contact_apply(g_contacts + 16 * c, x, yb);
#pragma omp for schedule(static)
		for (size_t k = 0; k < m; ++k)
			for (int u = 0; u < T; ++u)
				y[k] += buf[(size_t)u * m + k];
	}
	free(buf);
}

static void bsr_mv(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* bl, const double* x, double* y)
{
#pragma omp parallel for schedule(static) num_threads(or_nthreads) if (or_nthreads > 1)
	for (uint32_t i = 0; i < n; ++i)
	{
		double a[3] = { 0, 0, 0 };
		for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
		{
			const double* B = bl + 9 * k;
			const double* xv = x + 3 * ci[k];
			for (int r = 0; r < 3; ++r)
				a[r] += B[3 * r] * xv[0] + B[3 * r + 1] * xv[1] + B[3 * r + 2] * xv[2];
		}
		y[3 * i] = a[0], y[3 * i + 1] = a[1], y[3 * i + 2] = a[2];
	}
	contact_mv(n, x, y);
}

/* serial sum, or an OpenMP reduction with several threads */
static double dotn(const double* a, const double* b, size_t m)
{
	double s = 0;
	if (or_nthreads <= 1)
	{
		for (size_t i = 0; i < m; ++i)
			s += a[i] * b[i];
		return s;
	}
#pragma omp parallel for reduction(+ : s) schedule(static) num_threads(or_nthreads)
	for (size_t i = 0; i < m; ++i)
		s += a[i] * b[i];
	return s;
}

/* precond: 0 block Jacobi (3x3 inverse), 1 point Jacobi (make_pcg_jacobi_inverse,
 * precond_cg.cpp:93-111).  Returns iterations; out[0] = true relative residual,
 * out[1] = converged, out[2] = failure code (0, 1 non-finite rhs, 2 breakdown, 3 budget 0). */
uint32_t or_pcg(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* bl, const double* b, double* x,
	int precond, uint32_t max_iters, double rel_tol, double* out)
{
	const size_t m = 3 * (size_t)n;
	double*		 Minv = (double*)calloc(9 * (size_t)(n ? n : 1), sizeof(double));
	double*		 r = (double*)malloc(sizeof(double) * (m ? m : 1));
	double*		 z = (double*)malloc(sizeof(double) * (m ? m : 1));
	double*		 p = (double*)malloc(sizeof(double) * (m ? m : 1));
	double*		 q = (double*)malloc(sizeof(double) * (m ? m : 1));
	uint32_t	 it = 0;
	out[0] = 0, out[1] = 0, out[2] = 0;
	memset(x, 0, sizeof(double) * m);
	for (uint32_t i = 0; i < n; ++i)
	{
		const double* D = NULL;
		for (uint32_t k = rp[i]; k < rp[i + 1]; ++k)
			if (ci[k] == i)
				D = bl + 9 * k;
		double* M = Minv + 9 * i;
		if (!D)
			continue;
		int done = 0;
		if (precond == 0)
		{
			double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8], c02 = D[3] * D[7] - D[4] * D[6];
			double det = D[0] * c00 + D[1] * c01 + D[2] * c02;
			double mag = fabs(D[0]) + fabs(D[4]) + fabs(D[8]);
			if (isfinite(det) && mag > 0 && fabs(det) > 1e-30 * mag * mag * mag)
			{
				double id = 1.0 / det;
				M[0] = c00 * id, M[1] = (D[2] * D[7] - D[1] * D[8]) * id, M[2] = (D[1] * D[5] - D[2] * D[4]) * id;
				M[3] = c01 * id, M[4] = (D[0] * D[8] - D[2] * D[6]) * id, M[5] = (D[2] * D[3] - D[0] * D[5]) * id;
				M[6] = c02 * id, M[7] = (D[1] * D[6] - D[0] * D[7]) * id, M[8] = (D[0] * D[4] - D[1] * D[3]) * id;
				done = 1;
			}
		}
		if (!done)
		{
			double s = fmax(fmax(fabs(D[0]), fabs(D[4])), fabs(D[8]));
			if (s >= 1e-10)
			{
				double mn = fmax(1e-6, 1e-6 * s);
				M[0] = D[0] > mn ? 1.0 / D[0] : 0.0;
				M[4] = D[4] > mn ? 1.0 / D[4] : 0.0;
				M[8] = D[8] > mn ? 1.0 / D[8] : 0.0;
			}
		}
	}
	const double bb = dotn(b, b, m);
	if (!isfinite(bb))
	{
		out[2] = 1;
		goto done;
	}
	if (bb == 0.0)
	{
		out[1] = 1;
		goto done;
	}
	if (max_iters == 0)
	{
		out[0] = 1.0;
		out[2] = 3;
		goto done;
	}
	memcpy(r, b, sizeof(double) * m);
	for (uint32_t i = 0; i < n; ++i)
		for (int a = 0; a < 3; ++a)
			z[3 * i + a] = Minv[9 * i + 3 * a] * r[3 * i] + Minv[9 * i + 3 * a + 1] * r[3 * i + 1]
				+ Minv[9 * i + 3 * a + 2] * r[3 * i + 2];
	memcpy(p, z, sizeof(double) * m);
	double rz = dotn(r, z, m);
	for (it = 0; it < max_iters;)
	{
		bsr_mv(n, rp, ci, bl, p, q);
		const double pq = dotn(p, q, m);
		const double alpha = rz / pq;
		if (!(pq > 0) || !isfinite(alpha))
		{
			out[2] = 2;
			break;
		}
#pragma omp parallel for schedule(static) num_threads(or_nthreads) if (or_nthreads > 1)
		for (size_t k = 0; k < m; ++k)
		{
			x[k] += alpha * p[k];
			r[k] -= alpha * q[k];
		}
#pragma omp parallel for schedule(static) num_threads(or_nthreads) if (or_nthreads > 1)
		for (uint32_t i = 0; i < n; ++i)
			for (int a = 0; a < 3; ++a)
				z[3 * i + a] = Minv[9 * i + 3 * a] * r[3 * i] + Minv[9 * i + 3 * a + 1] * r[3 * i + 1]
					+ Minv[9 * i + 3 * a + 2] * r[3 * i + 2];
		const double rzn = dotn(r, z, m), rr = dotn(r, r, m);
		++it;
		if (rr <= rel_tol * rel_tol * bb)
			break;
		const double beta = rzn / rz;
		rz = rzn;
#pragma omp parallel for schedule(static) num_threads(or_nthreads) if (or_nthreads > 1)
		for (size_t k = 0; k < m; ++k)
			p[k] = z[k] + beta * p[k];
	}
	/* true residual */
	bsr_mv(n, rp, ci, bl, x, q);
	{
		double rr = 0;
		for (size_t k = 0; k < m; ++k)
			rr += (b[k] - q[k]) * (b[k] - q[k]);
		out[0] = sqrt(rr / bb);
		out[1] = (out[2] == 0 && out[0] <= rel_tol) ? 1 : 0;
	}
done:
	free(q);
	free(p);
	free(z);
	free(r);
	free(Minv);
	return it;
}

uint32_t or_pcg_contacts(uint32_t n, const uint32_t* rp, const uint32_t* ci, const double* bl, const uint32_t* contacts,
	uint32_t ncontacts, const double* b, double* x, int precond, uint32_t max_iters, double rel_tol, double* out)
{
	g_contacts = contacts;
	g_ncontacts = ncontacts;
	uint32_t it = or_pcg(n, rp, ci, bl, b, x, precond, max_iters, rel_tol, out);
	g_contacts = NULL;
	g_ncontacts = 0;
	return it;
}
