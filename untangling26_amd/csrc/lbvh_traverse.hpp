// Wave-cooperative LBVH traversal with staged candidate batches.
//
// Each lane traverses the tree for one query primitive (stack in LDS, spill to
// global memory past kStackLds entries).  Leaf hits are not written out one by
// one: the wave collects them in a per-wave LDS stage and, whenever 64 are
// pending, runs `narrow(query, leaf)` on all 64 at once, one per lane.  The
// narrow functor decides what is emitted (the plain broad-phase pair, or only
// pairs that pass an exact test), and emits through wave-aggregated appends,
// so the global counter sees one atomic per batch rather than one per pair.
#pragma once

#include "ut_common.hpp"
#include "lbvh.hpp"

namespace ut
{

constexpr int kTravThreads = 256;
constexpr int kTravWaves = kTravThreads / kWave;
constexpr int kStackLds = 32;
constexpr int kStackTotal = 128;
constexpr int kStageCap = 192; // < 64 pending + up to 128 new per step

// LDS part of the traversal stacks (LdsDepth entries per lane; deeper entries
// spill to the global stack) and the per-wave candidate stage.
template <int LdsDepth>
struct TravSharedT
{
	uint32_t stack[LdsDepth * kTravThreads];
	uint2	 stage[kTravWaves][kStageCap];
};
using TravShared = TravSharedT<kStackLds>;

__device__ __forceinline__ box3 query_box(int kind,
	uint32_t									q,
	const float* __restrict__ xb,
	const float* __restrict__ xe,
	const uint32_t* __restrict__ edges,
	const float* __restrict__ d_hat,
	const float* __restrict__ thickness)
{
	box3 qb;
	if (kind == kQueryVerts)
	{
		qb = box_add(box_of(ld3(xb, q)), ld3(xe, q));
		qb = box_inflate(qb, thickness[q] + d_hat[q]);
	}
	else
	{
		const uint32_t e0 = edges[2 * q], e1 = edges[2 * q + 1];
		qb = box_add(box_add(box_add(box_of(ld3(xb, e0)), ld3(xb, e1)), ld3(xe, e0)), ld3(xe, e1));
		qb = box_inflate(qb, thickness[e0] + d_hat[e0]);
	}
	return qb;
}

// Wave totals of a traversal (wave-uniform after the call).
struct TravCount
{
	unsigned long long visits = 0, tests = 0;
	unsigned long long steps = 0; // loop steps of the wave (node visits / (64 steps) = lane utilisation)
};

// Runs the traversal for query q (valid when active).  narrow(valid, q, leaf)
// is called by all 64 lanes together.  test(box) decides whether a child box
// is entered; spill_slot selects this lane's region of the global stack.
// When count is given, the wave's node visits and leaf tests are added to it
// (registers; the caller flushes them once).
template <int LdsDepth, typename NodeTest, typename Narrow>
__device__ __forceinline__ void traverse_staged_test(TravSharedT<LdsDepth>& sh,
	bool							 active,
	uint32_t						 q,
	uint32_t						 spill_slot,
	const NodeTest&					 test,
	bool							 cull_lt,
	const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ status,
	uint32_t* __restrict__ spill,
	uint32_t max_visits,
	Narrow&	 narrow,
	TravCount* count = nullptr)
{
	const uint32_t w = threadIdx.x / kWave;
	const uint32_t lane = __lane_id();
	uint2*		   stage = sh.stage[w];
	uint32_t	   pending = 0; // wave-uniform
	uint32_t	   tests = 0;	// wave-uniform: leaf tests run
	const uint64_t lt = lanemask_lt();

	auto run_batch = [&](uint32_t take)
	{
		// lanes < take process stage[pending - take + lane]
		const bool	   valid = lane < take;
		const uint2	   c = valid ? stage[pending - take + lane] : make_uint2(0u, 0u);
		pending -= take;
		tests += take;
		narrow(valid, c.x, c.y);
	};

	int		 sp = 0;
	uint32_t node = 0;
	bool	 go = active;
	uint32_t visits = 0;
	uint32_t steps = 0; // wave-uniform
	while (__ballot(go) != 0ull)
	{
		++steps;
		bool	 e0 = false, e1 = false;
		uint32_t o0 = 0, o1 = 0;
		if (go && ++visits > max_visits)
		{
			status[0] = 0u; // a valid tree is never revisited; stop on corruption
			go = false;
		}
		if (go)
		{
			const BvhNode* rec = &nodes[node];
			const float4   a = rec->a, b = rec->b, c = rec->c;
			const uint2	   refs = make_uint2(rec->d.x, rec->d.y);
			const box3	   b0{ mk3(a.x, a.y, a.z), mk3(a.w, b.x, b.y) };
			const box3	   b1{ mk3(b.z, b.w, c.x), mk3(c.y, c.z, c.w) };
			const bool	   h0 = test(b0);
			const bool	   h1 = test(b1);
			const bool	   l0 = (refs.x & kLeafBit) != 0u, l1 = (refs.y & kLeafBit) != 0u;
			o0 = refs.x & ~kLeafBit;
			o1 = refs.y & ~kLeafBit;
			e0 = h0 && l0 && (!cull_lt || q < o0);
			e1 = h1 && l1 && (!cull_lt || q < o1);
			const bool t0 = h0 && !l0, t1 = h1 && !l1;
			if (t0 && t1)
			{
				if (sp < LdsDepth)
					sh.stack[sp * kTravThreads + threadIdx.x] = refs.y;
				else if (sp < kStackTotal)
					spill[(size_t)spill_slot * (kStackTotal - LdsDepth) + (sp - LdsDepth)] = refs.y;
				if (sp < kStackTotal)
				{
					++sp;
					node = refs.x;
				}
				else
				{
					atomicAdd(&status[1], 1u);
					go = false;
				}
			}
			else if (t0)
				node = refs.x;
			else if (t1)
				node = refs.y;
			else if (sp > 0)
			{
				--sp;
				node = sp < LdsDepth ? sh.stack[sp * kTravThreads + threadIdx.x]
									 : spill[(size_t)spill_slot * (kStackTotal - LdsDepth) + (sp - LdsDepth)];
			}
			else
				go = false;
		}
		// stage this step's hits: lane order, child 0 before child 1
		const uint64_t m0 = __ballot(e0);
		const uint64_t m1 = __ballot(e1);
		const uint32_t n0 = (uint32_t)__popcll(m0);
		if (e0)
			stage[pending + (uint32_t)__popcll(m0 & lt)] = make_uint2(q, o0);
		if (e1)
			stage[pending + n0 + (uint32_t)__popcll(m1 & lt)] = make_uint2(q, o1);
		pending += n0 + (uint32_t)__popcll(m1);
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		while (pending >= (uint32_t)kWave)
			run_batch(kWave);
		__builtin_amdgcn_wave_barrier();
	}
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	if (pending > 0)
		run_batch(pending);
	if (count)
	{
		uint32_t v = visits;
		for (int o = 32; o > 0; o >>= 1)
			v += __shfl_xor(v, o, kWave);
		count->visits += v;
		count->tests += tests;
		count->steps += steps;
	}
}

// Box-overlap traversal of a query box (broad phase); the query index is its spill slot.
template <typename Narrow>
__device__ __forceinline__ void traverse_staged(TravShared& sh,
	bool							 active,
	uint32_t						 q,
	const box3&						 qb,
	bool							 cull_lt,
	const BvhNode* __restrict__ nodes,
	uint32_t* __restrict__ status,
	uint32_t* __restrict__ spill,
	uint32_t max_visits,
	Narrow&	 narrow)
{
	auto test = [&](const box3& b) { return box_overlap(b, qb); };
	traverse_staged_test(sh, active, q, q, test, cull_lt, nodes, status, spill, max_visits, narrow);
}

// Narrow functor of the plain broad phase: write (query, leaf) pairs.
struct EmitPairs
{
	uint32_t* count;
	uint2*	  list;
	uint32_t  max_count;
	__device__ __forceinline__ void operator()(bool valid, uint32_t q, uint32_t leaf)
	{
		const uint32_t s = wave_append(count, valid);
		if (valid && s < max_count)
			list[s] = make_uint2(q, leaf);
	}
};

} // namespace ut
