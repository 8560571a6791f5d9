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
 */
#include "oracle.h"

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

static int find_pair(const uint32_t* ids, uint32_t P, uint32_t e, uint32_t f)
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
	return (lo < P && ids[2 * lo] == e && ids[2 * lo + 1] == f) ? (int)lo : -1;
}

/* host_preprocess_ef_pairs without the degenerate filter (DCD pairs are never
 * degenerate: the DCD rejects barycentrics below the same 1e-4), with
 * compute_adj_pairs (:503-558) and find_loop_vertex (:431-501).
 * out_ids uint32[2*3P], out_pairs uint32[12*3P], adj uint32[12*3P], adj_n uint32[3P].
 * Returns the total pair count (DCD + loop pairs) or -1 on adjacency overflow. */
long or_ef_preprocess(uint32_t P, const uint32_t* in_ids, const uint32_t* in_pairs, const uint32_t* edges,
	const uint32_t* faces, const uint32_t* face_adj_edges, const uint32_t* edge_adj_faces, uint32_t* out_ids,
	uint32_t* out_pairs, uint32_t* adj, uint32_t* adj_n)
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
	/* loop vertices, in (pair, adjacent face) order */
	uint32_t N = P;
	uint32_t* loops_of = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (P ? P : 1));
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
			uint32_t* r = out_pairs + 12 * N;
			const float one = 1.0f, zero = 0.0f;
			for (int q = 0; q < 5; ++q)
				r[q] = v;
			memcpy(r + 5, &one, 4);
			memcpy(r + 6, &one, 4);
			memcpy(r + 7, &zero, 4);
			for (int q = 8; q < 12; ++q)
				memcpy(r + q, &zero, 4);
			adj[12 * N] = p;
			adj_n[N] = 1;
			loops_of[2 * p + nloops[p]++] = N;
			++N;
		}
	}
	long result = N;
	for (uint32_t p = 0; p < P; ++p)
	{
		const uint32_t e = out_ids[2 * p], f = out_ids[2 * p + 1];
		uint32_t	   n = 0;
		uint32_t	   tmp[32];
		for (int k = 0; k < 2; ++k) /* via edge's adjacent edges */
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
			{
				int q = find_pair(out_ids, P, o[j], f);
				if (q >= 0 && (uint32_t)q != p)
					tmp[n++] = (uint32_t)q;
			}
		}
		for (int ii = 0; ii < 3; ++ii) /* via edge's adjacent faces */
		{
			const uint32_t fe = face_adj_edges[3 * f + ii];
			for (int k = 0; k < 2; ++k)
			{
				const uint32_t T = edge_adj_faces[2 * e + k];
				if (T == NONE || fe == e)
					continue;
				int q = find_pair(out_ids, P, fe, T);
				if (q >= 0 && (uint32_t)q != p)
					tmp[n++] = (uint32_t)q;
			}
		}
		for (uint32_t j = 0; j < nloops[p]; ++j)
			tmp[n++] = loops_of[2 * p + j];
		if (n > 12)
			result = -1;
		adj_n[p] = n > 12 ? 12 : n;
		memcpy(adj + 12 * p, tmp, sizeof(uint32_t) * adj_n[p]);
	}
	free(nloops);
	free(loops_of);
	return result;
}

/* intersection_resolver2.cpp:266-283 (flip weight: 0 when the intersected
 * face triples are equal) and intersection_resolver_helper.cpp:518-575
 * (stack DFS from each unvisited pair in index order, parity = path sum % 2). */
uint32_t or_contours_dfs(uint32_t N, const uint32_t* pairs, const uint32_t* adj, const uint32_t* adj_n,
	uint32_t* contour_index, uint32_t* mesh_index)
{
	uint8_t*  visited = (uint8_t*)calloc(N ? N : 1, 1);
	uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (N ? N : 1));
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
				const uint32_t v = adj[12 * u + j];
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
