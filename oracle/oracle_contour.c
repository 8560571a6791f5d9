/*
 * oracle_contour.c -- CPU restatement of mesh adjacency, EF preprocessing and
 * contour construction.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Pinning:
 *  - or_graph_bfs restates cpu_reference of
 *    UnitTest/test_parallel_contour_construction.cpp:155-198 and is checked
 *    against that test's small_fixed known answer (:613-620): components
 *    {0,1,2},{3,4,5},{6},{7,8,9,10},{11,12,13}, inconsistent (odd triangle).
 *  - or_contours_dfs (contour_flood_filling_template) agrees with or_graph_bfs
 *    on every consistent component (same sets, same parity from the same start).
 *  - or_mesh_topology: closed forms on a grid (interior edges have two faces,
 *    boundary edges one; face edges follow face[j] -> face[j+1]).
 *  - or_ef_preprocess: adjacency symmetric; on two crossing sheets every
 *    interior EF pair has exactly two neighbours (a contour is a 1-manifold
 *    curve); loop pairs only where a mesh crosses itself at a shared vertex.
 *    Degenerate filter: hand-built lists where the kept set is known (a
 *    degenerate twin of a contour pair stays, an isolated degenerate pair
 *    goes, an all-degenerate list stays whole).
 *  - or_ef_from_proximity: hand-derived outputs of a vertex lying on a face
 *    and of two crossing edges (every generated pair listed by hand from the
 *    mesh), the 1e-10 distance threshold on both sides, and on a sheet whose
 *    vertex row lies in another sheet's plane every vertex of that row yields
 *    one pair per edge around it.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NONE 0xFFFFFFFFu

/* init_mesh_data.cpp:419-500 */
int or_mesh_topology(uint32_t ne, uint32_t nf, const uint32_t* edges, const uint32_t* faces, uint32_t* face_adj_edges,
	uint32_t* edge_adj_faces)
{
	/* edge lookup by sorted (min, max) keys */
	uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (ne ? ne : 1));
	uint32_t* ord = (uint32_t*)malloc(sizeof(uint32_t) * (ne ? ne : 1));
	for (uint32_t e = 0; e < ne; ++e)
	{
		uint32_t a = edges[2 * e], b = edges[2 * e + 1];
		if (a > b)
		{
			uint32_t t = a;
			a = b;
			b = t;
		}
		key[e] = ((uint64_t)a << 32) | b;
		ord[e] = e;
	}
	/* insertion into a sorted index via simple merge sort on (key, id) */
	for (uint32_t w = 1; w < ne; w *= 2)
	{
		uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * ne);
		for (uint32_t lo = 0; lo < ne; lo += 2 * w)
		{
			uint32_t mid = lo + w < ne ? lo + w : ne, hi = lo + 2 * w < ne ? lo + 2 * w : ne;
			uint32_t i = lo, j = mid, k = lo;
			while (i < mid && j < hi)
				tmp[k++] = key[ord[i]] <= key[ord[j]] ? ord[i++] : ord[j++];
			while (i < mid)
				tmp[k++] = ord[i++];
			while (j < hi)
				tmp[k++] = ord[j++];
		}
		memcpy(ord, tmp, sizeof(uint32_t) * ne);
		free(tmp);
	}
	int ok = 1;
	for (uint32_t e = 0; e < 2 * ne; ++e)
		edge_adj_faces[e] = NONE;
	for (uint32_t f = 0; f < nf && ok; ++f)
	{
		for (int j = 0; j < 3; ++j)
		{
			uint32_t a = faces[3 * f + j], b = faces[3 * f + (j + 1) % 3];
			if (a > b)
			{
				uint32_t t = a;
				a = b;
				b = t;
			}
			const uint64_t k = ((uint64_t)a << 32) | b;
			uint32_t	   lo = 0, hi = ne;
			while (lo < hi)
			{
				uint32_t mid = (lo + hi) / 2;
				if (key[ord[mid]] < k)
					lo = mid + 1;
				else
					hi = mid;
			}
			if (lo >= ne || key[ord[lo]] != k)
			{
				ok = 0;
				break;
			}
			const uint32_t e = ord[lo];
			face_adj_edges[3 * f + j] = e;
			if (edge_adj_faces[2 * e] == NONE)
				edge_adj_faces[2 * e] = f;
			else if (edge_adj_faces[2 * e + 1] == NONE)
				edge_adj_faces[2 * e + 1] = f;
			else
				ok = 0; /* non-manifold edge */
		}
	}
	free(ord);
	free(key);
	return ok;
}

static int cmp_ef(const void* a, const void* b)
{
	const uint32_t* x = (const uint32_t*)a;
	const uint32_t* y = (const uint32_t*)b;
	if (x[0] != y[0])
		return x[0] < y[0] ? -1 : 1;
	if (x[1] != y[1])
		return x[1] < y[1] ? -1 : 1;
	return memcmp(x + 2, y + 2, 48); /* payload tie-break, intersection_resolver.cpp:712-716 */
}

/* First index with (eid, fid) >= (e, f) in the sorted ids. */
static uint32_t lower_pair(const uint32_t* ids, uint32_t P, uint32_t e, uint32_t f)
{
	uint32_t lo = 0, hi = P;
	while (lo < hi)
	{
		uint32_t mid = (lo + hi) / 2;
		if (ids[2 * mid] < e || (ids[2 * mid] == e && ids[2 * mid + 1] < f))
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

typedef struct
{
	uint32_t* v;
	uint32_t  n, cap;
} u32vec;

static void push(u32vec* a, uint32_t x)
{
	if (a->n == a->cap)
	{
		a->cap = a->cap ? 2 * a->cap : 64;
		a->v = (uint32_t*)realloc(a->v, sizeof(uint32_t) * a->cap);
	}
	a->v[a->n++] = x;
}

/* every pair with key (e, f) other than p, ascending */
static void push_key(u32vec* out, const uint32_t* ids, uint32_t P, uint32_t e, uint32_t f, uint32_t p)
{
	for (uint32_t q = lower_pair(ids, P, e, f); q < P && ids[2 * q] == e && ids[2 * q + 1] == f; ++q)
		if (q != p)
			push(out, q);
}

/* compute_adj_pairs (intersection_resolver.cpp:503-558) of pair p among the
 * first P sorted pairs: pairs on the other edges of the faces around e with the
 * same face f, then pairs of an edge of f cutting a face around e.  Every pair
 * of a key is listed (duplicate keys occur), in index order. */
static void ef_adjacent(uint32_t p, uint32_t P, const uint32_t* ids, const uint32_t* face_adj_edges,
	const uint32_t* edge_adj_faces, u32vec* out)
{
	const uint32_t e = ids[2 * p], f = ids[2 * p + 1];
	for (int k = 0; k < 2; ++k)
	{
		const uint32_t T = edge_adj_faces[2 * e + k];
		if (T == NONE)
			continue;
		const uint32_t* fae = face_adj_edges + 3 * T;
		uint32_t		o[2];
		if (fae[0] == e)
			o[0] = fae[1], o[1] = fae[2];
		else if (fae[1] == e)
			o[0] = fae[0], o[1] = fae[2];
		else
			o[0] = fae[0], o[1] = fae[1];
		for (int j = 0; j < 2; ++j)
			push_key(out, ids, P, o[j], f, p);
	}
	for (int ii = 0; ii < 3; ++ii)
	{
		const uint32_t fe = face_adj_edges[3 * f + ii];
		if (fe == e)
			continue;
		/* pairs of edge fe are sorted by face, and edge_adj_faces is ascending */
		for (int k = 0; k < 2; ++k)
		{
			const uint32_t T = edge_adj_faces[2 * e + k];
			if (T != NONE)
				push_key(out, ids, P, fe, T, p);
		}
	}
}

/* is_degenerate_ef_pair (intersection_resolver.cpp:561-569): a barycentric
 * coordinate of the edge point or of the face point below 1e-4, the third face
 * coordinate and the second edge coordinate recomputed from the stored ones. */
int or_ef_degenerate(const uint32_t* r)
{
	float b0, b1, b2;
	memcpy(&b0, r + 5, 4);
	memcpy(&b1, r + 6, 4);
	memcpy(&b2, r + 7, 4);
	const float e1 = 1.0f - b0;
	const float f2 = 1.0f - b1 - b2;
	const float emin = e1 < b0 ? e1 : b0;
	float		fmin = b1;
	if (b2 < fmin)
		fmin = b2;
	if (f2 < fmin)
		fmin = f2;
	return emin < 1e-4f || fmin < 1e-4f;
}

/* host_preprocess_ef_pairs (intersection_resolver.cpp:683-1096):
 *  1. canonical order: (eid, fid), ties by the 48 payload bytes (:695-726);
 *  2. with filter_degenerate (the reference's !ignore_near_zero_dist_pairs,
 *     :841-897): a degenerate pair stays only when one of its adjacent pairs is
 *     not degenerate; when no pair would stay, all stay;
 *  3. loop-vertex pairs (find_loop_vertex :431-501), appended in (pair, face
 *     around the edge) order, each adjacent to its source pair, and appended to
 *     its source pair's adjacency (:1009-1018).
 * Adjacency is packed: the neighbours of pair i are adj[off_i .. off_i + adj_n[i])
 * with off_i = adj_n[0] + ... + adj_n[i-1].
 * out_ids uint32[2*3P], out_pairs uint32[12*3P], adj_n uint32[3P], *adj is
 * malloc'ed here (free with or_free).  Returns the total pair count (kept DCD
 * pairs + loop pairs); *kept gets the pair count after the filter. */
long or_ef_preprocess(uint32_t P, const uint32_t* in_ids, const uint32_t* in_pairs, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, int filter_degenerate,
	uint32_t* out_ids, uint32_t* out_pairs, uint32_t** adj, uint32_t* adj_n, uint32_t* kept)
{
	uint32_t* rec = (uint32_t*)malloc(sizeof(uint32_t) * 14 * (P ? P : 1));
	for (uint32_t i = 0; i < P; ++i)
	{
		rec[14 * i] = in_ids[2 * i];
		rec[14 * i + 1] = in_ids[2 * i + 1];
		memcpy(rec + 14 * i + 2, in_pairs + 12 * i, 48);
	}
	qsort(rec, P, 14 * sizeof(uint32_t), cmp_ef);
	for (uint32_t i = 0; i < P; ++i)
	{
		out_ids[2 * i] = rec[14 * i];
		out_ids[2 * i + 1] = rec[14 * i + 1];
		memcpy(out_pairs + 12 * i, rec + 14 * i + 2, 48);
	}
	free(rec);
	u32vec nb = { 0, 0, 0 };
	if (filter_degenerate && P)
	{
		uint8_t* deg = (uint8_t*)malloc(P);
		uint8_t* keep = (uint8_t*)malloc(P);
		uint32_t ndeg = 0, nkeep = 0;
		for (uint32_t p = 0; p < P; ++p)
			ndeg += (deg[p] = (uint8_t)or_ef_degenerate(out_pairs + 12 * p));
		for (uint32_t p = 0; p < P; ++p)
		{
			keep[p] = !deg[p];
			if (!keep[p])
			{
				nb.n = 0;
				ef_adjacent(p, P, out_ids, face_adj_edges, edge_adj_faces, &nb);
				for (uint32_t j = 0; j < nb.n && !keep[p]; ++j)
					keep[p] = !deg[nb.v[j]];
			}
			nkeep += keep[p];
		}
		if (ndeg && nkeep && nkeep < P)
		{
			uint32_t w = 0;
			for (uint32_t p = 0; p < P; ++p)
			{
				if (!keep[p])
					continue;
				out_ids[2 * w] = out_ids[2 * p];
				out_ids[2 * w + 1] = out_ids[2 * p + 1];
				memmove(out_pairs + 12 * w, out_pairs + 12 * p, 48);
				++w;
			}
			P = w;
		}
		free(deg);
		free(keep);
	}
	*kept = P;
	/* loop vertices, in (pair, adjacent face) order */
	uint32_t  N = P;
	uint32_t* loops_of = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (P ? P : 1));
	uint32_t* loop_src = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (P ? P : 1));
	uint32_t* nloops = (uint32_t*)calloc(P ? P : 1, sizeof(uint32_t));
	for (uint32_t p = 0; p < P; ++p)
	{
		const uint32_t e = out_ids[2 * p], f = out_ids[2 * p + 1];
		const uint32_t a = edges[2 * e], b = edges[2 * e + 1];
		for (int k = 0; k < 2; ++k)
		{
			const uint32_t T = edge_adj_faces[2 * e + k];
			if (T == NONE)
				continue;
			const uint32_t v = faces[3 * T] + faces[3 * T + 1] + faces[3 * T + 2] - a - b;
			if (!(v == faces[3 * f] || v == faces[3 * f + 1] || v == faces[3 * f + 2]))
				continue;
			const uint32_t* fae = face_adj_edges + 3 * T;
			uint32_t		le;
#define HAS(ed) (edges[2 * (ed)] == v || edges[2 * (ed) + 1] == v)
			if (e == fae[0])
				le = HAS(fae[1]) ? fae[1] : fae[2];
			else if (e == fae[1])
				le = HAS(fae[0]) ? fae[0] : fae[2];
			else
				le = HAS(fae[0]) ? fae[0] : fae[1];
#undef HAS
			out_ids[2 * N] = le;
			out_ids[2 * N + 1] = f;
			uint32_t*	r = out_pairs + 12 * N;
			const float one = 1.0f, zero = 0.0f;
			for (int q = 0; q < 5; ++q)
				r[q] = v;
			memcpy(r + 5, &one, 4);
			memcpy(r + 6, &one, 4);
			memcpy(r + 7, &zero, 4);
			for (int q = 8; q < 12; ++q)
				memcpy(r + q, &zero, 4);
			loop_src[N - P] = p;
			loops_of[2 * p + nloops[p]++] = N;
			++N;
		}
	}
	u32vec all = { 0, 0, 0 };
	for (uint32_t p = 0; p < P; ++p)
	{
		nb.n = 0;
		ef_adjacent(p, P, out_ids, face_adj_edges, edge_adj_faces, &nb);
		for (uint32_t j = 0; j < nloops[p]; ++j)
			push(&nb, loops_of[2 * p + j]);
		adj_n[p] = nb.n;
		for (uint32_t j = 0; j < nb.n; ++j)
			push(&all, nb.v[j]);
	}
	for (uint32_t i = P; i < N; ++i)
	{
		adj_n[i] = 1;
		push(&all, loop_src[i - P]);
	}
	*adj = all.v ? all.v : (uint32_t*)malloc(4);
	free(nb.v);
	free(nloops);
	free(loop_src);
	free(loops_of);
	return N;
}

void or_free(void* p) { free(p); }

/* fn_add_ef_from_proximity_pairs (intersection_resolver.cpp:199-363).  A
 * proximity pair (CollisionPairTemplate, 16 words; VF or EE by the type bit)
 * whose weighted point difference d = sum_k w_k x_k has |d|^2 < 1e-10 (FP32,
 * the thickness term is 0 there) becomes EF pairs with zero gradient and
 * stiffness:
 *  VF (v, f): one per edge ae around v (vert_adj CSR order): edge bary
 *    (1, 0) when v is the edge's first vertex else (0, 1), face bary |w_1..3|;
 *  EE (e1, e2): for each face af around e2 holding no vertex of e1: (e1, af)
 *    with e1's bary, the face corners on e2 weighted by e1's bary (e2.x gets
 *    bary.x, e2.y bary.y); then for each face af around e1 holding no vertex
 *    of e2: (e2, af) with e2's bary, its corners on e1 weighted by e1's bary
 *    (the reference uses edge1_bary in both loops).
 * ids: (v, f) for VF, (e1, e2) for EE.  Writes at most cap pairs; returns the
 * number generated. */
size_t or_ef_from_proximity(size_t n, const uint32_t* recs, const uint32_t* ids, const float* x, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* vae_off, const uint32_t* vae, const uint32_t* edge_adj_faces,
	uint32_t* out_pairs, uint32_t* out_ids, size_t cap)
{
	const uint32_t M = 0x7FFFFFFFu, T = 0x80000000u;
	size_t		   cnt = 0;
	for (size_t i = 0; i < n; ++i)
	{
		const uint32_t* r = recs + 16 * i;
		const int		is_vf = (r[2] & T) != 0, is_ee = !is_vf && (r[3] & T) != 0;
		if (!is_vf && !is_ee)
			continue;
		const uint32_t q[4] = { r[0] & M, r[1] & M, r[2] & M, r[3] & M };
		float		   b0, b1;
		memcpy(&b0, r + 10, 4);
		memcpy(&b1, r + 11, 4);
		float w[4];
		if (is_vf)
			w[0] = 1.0f, w[1] = -b0, w[2] = -b1, w[3] = b0 + b1 - 1.0f;
		else
			w[0] = b0, w[1] = 1.0f - b0, w[2] = -b1, w[3] = b1 - 1.0f;
		float d[3];
		for (int c = 0; c < 3; ++c)
			d[c] = w[0] * x[3 * q[0] + c] + w[1] * x[3 * q[1] + c] + w[2] * x[3 * q[2] + c] + w[3] * x[3 * q[3] + c];
		const float d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
		if (d2 >= 1e-10f)
			continue;
		const float zero = 0.0f;
		if (is_vf)
		{
			const uint32_t v = q[0], f = ids[2 * i + 1];
			const float	   fb0 = fabsf(w[1]), fb1 = fabsf(w[2]);
			for (uint32_t k = vae_off[v]; k < vae_off[v + 1]; ++k)
			{
				const uint32_t ae = vae[k];
				if (cnt < cap)
				{
					uint32_t*	o = out_pairs + 12 * cnt;
					const float eb = edges[2 * ae] == v ? 1.0f : 0.0f;
					o[0] = edges[2 * ae], o[1] = edges[2 * ae + 1];
					o[2] = faces[3 * f], o[3] = faces[3 * f + 1], o[4] = faces[3 * f + 2];
					memcpy(o + 5, &eb, 4);
					memcpy(o + 6, &fb0, 4);
					memcpy(o + 7, &fb1, 4);
					for (int z = 8; z < 12; ++z)
						memcpy(o + z, &zero, 4);
					out_ids[2 * cnt] = ae, out_ids[2 * cnt + 1] = f;
				}
				++cnt;
			}
			continue;
		}
		const uint32_t e1 = ids[2 * i], e2 = ids[2 * i + 1];
		const uint32_t a0 = edges[2 * e1], a1 = edges[2 * e1 + 1], c0 = edges[2 * e2], c1 = edges[2 * e2 + 1];
		const float	   eb1[2] = { fabsf(w[0]), fabsf(w[1]) }, eb2[2] = { fabsf(w[2]), fabsf(w[3]) };
		for (int side = 0; side < 2; ++side)
		{
			/* side 0: faces around e2, the EF edge is e1; side 1: faces around e1, edge e2 */
			const uint32_t fe = side ? e1 : e2, ee = side ? e2 : e1;
			const uint32_t s0 = side ? a0 : c0, s1 = side ? a1 : c1; /* corners shared with the face */
			const uint32_t x0 = side ? c0 : a0, x1 = side ? c1 : a1; /* must not be in the face */
			const float*   eb = side ? eb2 : eb1;
			for (int k = 0; k < 2; ++k)
			{
				const uint32_t af = edge_adj_faces[2 * fe + k];
				if (af == NONE)
					continue;
				const uint32_t* F = faces + 3 * af;
				if (F[0] == x0 || F[1] == x0 || F[2] == x0 || F[0] == x1 || F[1] == x1 || F[2] == x1)
					continue;
				float fb[3] = { 0.0f, 0.0f, 0.0f };
				int	  matched = 0;
				for (int j = 0; j < 3; ++j)
				{
					if (F[j] == s0)
						fb[j] = eb1[0], ++matched;
					else if (F[j] == s1)
						fb[j] = eb1[1], ++matched;
				}
				if (matched != 2)
					continue;
				if (cnt < cap)
				{
					uint32_t* o = out_pairs + 12 * cnt;
					o[0] = edges[2 * ee], o[1] = edges[2 * ee + 1];
					o[2] = F[0], o[3] = F[1], o[4] = F[2];
					memcpy(o + 5, &eb[0], 4);
					memcpy(o + 6, &fb[0], 4);
					memcpy(o + 7, &fb[1], 4);
					for (int z = 8; z < 12; ++z)
						memcpy(o + z, &zero, 4);
					out_ids[2 * cnt] = ee, out_ids[2 * cnt + 1] = af;
				}
				++cnt;
			}
		}
	}
	return cnt;
}

/* intersection_resolver2.cpp:266-283 (flip weight: 0 when the intersected
 * face triples are equal) and intersection_resolver_helper.cpp:518-575
 * (stack DFS from each unvisited pair in index order, parity = path sum % 2). */
uint32_t or_contours_dfs(uint32_t N, const uint32_t* pairs, const uint32_t* adj, const uint32_t* adj_n,
	uint32_t* contour_index, uint32_t* mesh_index)
{
	uint8_t*  visited = (uint8_t*)calloc(N ? N : 1, 1);
	uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (N ? N : 1));
	size_t*	  off = (size_t*)malloc(sizeof(size_t) * (N + 1));
	off[0] = 0;
	for (uint32_t i = 0; i < N; ++i)
		off[i + 1] = off[i] + adj_n[i];
	uint32_t  nc = 0;
	for (uint32_t s = 0; s < N; ++s)
	{
		if (visited[s])
			continue;
		size_t sp = 0;
		stack[2 * sp] = s;
		stack[2 * sp + 1] = 0;
		sp++;
		visited[s] = 1;
		contour_index[s] = nc;
		mesh_index[s] = 0;
		while (sp)
		{
			--sp;
			const uint32_t u = stack[2 * sp], sum = stack[2 * sp + 1];
			for (uint32_t j = 0; j < adj_n[u]; ++j)
			{
				const uint32_t v = adj[off[u] + j];
				if (visited[v])
					continue;
				const uint32_t* fu = pairs + 12 * u + 2;
				const uint32_t* fv = pairs + 12 * v + 2;
				const uint32_t	w = (fu[0] == fv[0] && fu[1] == fv[1] && fu[2] == fv[2]) ? 0u : 1u;
				visited[v] = 1;
				contour_index[v] = nc;
				mesh_index[v] = (sum + w) % 2;
				stack[2 * sp] = v;
				stack[2 * sp + 1] = sum + w;
				sp++;
			}
		}
		nc++;
	}
	free(off);
	free(stack);
	free(visited);
	return nc;
}

/* cpu_reference, UnitTest/test_parallel_contour_construction.cpp:155-198:
 * BFS components, parity, and the odd-cycle flag.  comp_min[v] = smallest
 * vertex of v's component.  Returns 1 when any component is inconsistent. */
int or_graph_bfs(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t* comp_min,
	uint32_t* parity, uint32_t* comp_inconsistent)
{
	uint8_t*  visited = (uint8_t*)calloc(n ? n : 1, 1);
	uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
	int		  any = 0;
	for (uint32_t v = 0; v < n; ++v)
		comp_inconsistent[v] = 0;
	for (uint32_t s = 0; s < n; ++s)
	{
		if (visited[s])
			continue;
		size_t head = 0, tail = 0;
		q[tail++] = s;
		visited[s] = 1;
		parity[s] = 0;
		comp_min[s] = s;
		while (head < tail)
		{
			const uint32_t u = q[head++];
			for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e)
			{
				const uint32_t v = col[e], want = parity[u] ^ (w[e] & 1u);
				if (!visited[v])
				{
					visited[v] = 1;
					parity[v] = want;
					comp_min[v] = s;
					q[tail++] = v;
				}
				else if (parity[v] != want)
				{
					comp_inconsistent[s] = 1;
					any = 1;
				}
			}
		}
	}
	free(q);
	free(visited);
	return any;
}

/* classify_contour_type (intersection_resolver2.cpp:311-375) and the mesh_is_closed
 * roles (:2823-2841), contour by contour over its pairs in index order.  pairs:
 * EfPair records (12 words, edge vertices in words 0 and 1: a loop-vertex pair has
 * them equal, collision_data.h:199); is_boundary[p]: the pair's edge has at most
 * one adjacent face (:2814-2818); adj_n[p]: its EF-adjacency degree.  type[nc]
 * receives the ContourType (0 undefined, 1 closed, 2 eight, 3 LL, 4 BLI, 5 cross,
 * 6 BIBI, 7 BBII), closed[2 nc] the roles' mesh_is_closed. */
void or_classify_contours(uint32_t N, const uint32_t* pairs, const uint8_t* is_boundary, const uint32_t* adj_n,
	const uint32_t* contour_index, const uint32_t* mesh_index, uint32_t nc, uint32_t* type, uint32_t* closed)
{
	for (uint32_t c = 0; c < nc; ++c)
	{
		uint32_t nb = 0, nl = 0, front = 0, back = 0;
		int		 is_closed = 1;
		int		 has_role[2] = { 0, 0 }, nonunit[2] = { 1, 1 };
		for (uint32_t p = 0; p < N; ++p)
		{
			if (contour_index[p] != c)
				continue;
			const int loop = pairs[12 * (size_t)p] == pairs[12 * (size_t)p + 1];
			if (loop)
				++nl;
			if (is_boundary[p])
			{
				if (nb == 0)
					front = p;
				back = p;
				++nb;
				is_closed = 0;
			}
			if (!loop && adj_n[p] == 1)
				is_closed = 0;
			const uint32_t r = mesh_index[p] & 1u;
			has_role[r] = 1;
			if (!loop && adj_n[p] == 1)
				nonunit[r] = 0;
		}
		uint32_t t = 0;
		if (nb == 0)
			t = !is_closed ? 0 : nl == 0 ? 1 : nl == 1 ? 2 : nl == 2 ? 3 : 0;
		else if (nl == 0)
			t = nb == 2 ? (mesh_index[front] == mesh_index[back] ? 7 : 6) : 0;
		else if (nl == 1)
			t = nb == 1 ? 4 : nb == 2 ? 5 : 0;
		type[c] = t;
		closed[2 * c] = has_role[0] && nonunit[0];
		closed[2 * c + 1] = has_role[1] && nonunit[1];
	}
}
