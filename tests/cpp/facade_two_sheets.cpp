// The C++ facade (include/untangle.hpp) end to end on one mesh: face and edge
// LBVHs, EF intersection detection, EF preprocessing and contours, the PRP
// resolve, and a proximity DCD pass -- the calls the reference's NewtonSolver
// makes on lcs::LBVH, lcs::IntersectionResolver and lcs::NarrowPhasesDetector.
//
// usage: facade_two_sheets DIR NV NE NF
// DIR holds raw little-endian arrays written by the test: x.f32 [3NV],
// edges.u32 [2NE], faces.u32 [3NF], fae.u32 [3NF], eaf.u32 [2NE], va.f32 [NV],
// ea.f32 [NE], fa.f32 [NF], mass.f32 [NV], dhat.f32 [NV], th.f32 [NV].
// Prints one JSON line with the counts.
#include "untangle.hpp"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace
{
template <typename T>
std::vector<T> load(const std::string& path, size_t n)
{
	std::vector<T> v(n);
	std::ifstream  f(path, std::ios::binary);
	if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T))))
		throw std::runtime_error("cannot read " + path);
	return v;
}

template <typename T>
T* upload(const std::vector<T>& h)
{
	T* d = nullptr;
	if (hipMalloc((void**)&d, (h.size() ? h.size() : 1) * sizeof(T)) != hipSuccess)
		throw std::runtime_error("hipMalloc");
	if (h.size() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
		throw std::runtime_error("hipMemcpy");
	return d;
}

template <typename T>
T* zeros(size_t n)
{
	return upload(std::vector<T>(n, T(0)));
}

uint32_t read_u32(const uint32_t* d)
{
	uint32_t h = 0;
	if (hipMemcpy(&h, d, 4, hipMemcpyDeviceToHost) != hipSuccess)
		throw std::runtime_error("hipMemcpy");
	return h;
}
} // namespace

int main(int argc, char** argv)
{
	if (argc != 5)
	{
		std::fprintf(stderr, "usage: %s DIR NV NE NF\n", argv[0]);
		return 2;
	}
	const std::string dir = argv[1];
	const uint32_t	  nv = std::stoul(argv[2]), ne = std::stoul(argv[3]), nf = std::stoul(argv[4]);
	try
	{
		lcs::Device dev(0);
		float*		x = upload(load<float>(dir + "/x.f32", 3 * (size_t)nv));
		uint32_t*	edges = upload(load<uint32_t>(dir + "/edges.u32", 2 * (size_t)ne));
		uint32_t*	faces = upload(load<uint32_t>(dir + "/faces.u32", 3 * (size_t)nf));
		uint32_t*	fae = upload(load<uint32_t>(dir + "/fae.u32", 3 * (size_t)nf));
		uint32_t*	eaf = upload(load<uint32_t>(dir + "/eaf.u32", 2 * (size_t)ne));
		float*		va = upload(load<float>(dir + "/va.f32", nv));
		float*		ea = upload(load<float>(dir + "/ea.f32", ne));
		float*		fa = upload(load<float>(dir + "/fa.f32", nf));
		float*		mass = upload(load<float>(dir + "/mass.f32", nv));
		float*		dhat = upload(load<float>(dir + "/dhat.f32", nv));
		float*		th = upload(load<float>(dir + "/th.f32", nv));

		lcs::LBVH face_tree(dev, lcs::TreeType::Face, nf), edge_tree(dev, lcs::TreeType::Edge, ne);
		face_tree.reduce_face_tree_aabb(x, faces);
		face_tree.construct_tree();
		face_tree.update_face_tree_leave_aabb(th, x, x, faces);
		face_tree.refit();
		edge_tree.reduce_edge_tree_aabb(x, edges);
		edge_tree.construct_tree();
		edge_tree.update_edge_tree_leave_aabb(th, x, x, edges);
		edge_tree.refit();

		lcs::ResolverMesh rm;
		rm.num_verts = nv, rm.num_edges = ne, rm.num_faces = nf, rm.edges = edges, rm.faces = faces;
		rm.face_adj_edges = fae, rm.edge_adj_faces = eaf, rm.rest_x = x, rm.rest_vert_area = va;
		rm.rest_edge_area = ea, rm.rest_face_area = fa, rm.mass = mass;
		ut_prp_params pp{ 256, 10, 32, 1.0f, 1e8f, 0.25f };
		lcs::IntersectionResolver resolver(dev, rm, pp);

		const uint32_t cap = 4 * ne + 64;
		uint32_t*	   count = zeros<uint32_t>(1);
		uint32_t*	   pair_ids = zeros<uint32_t>(2 * (size_t)cap);
		uint32_t*	   pairs = zeros<uint32_t>(12 * (size_t)cap);
		resolver.ef_dcd_intersection_query(face_tree, x, dhat, th, count, pairs, pair_ids, cap);
		const uint32_t nef = read_u32(count);
		const auto	   pc = resolver.host_preprocess_ef_pairs(nef, pairs, pair_ids);
		std::vector<ut_prp_choice> choices;
		ut_prp_stats			   st{};
		uint32_t*				   resp = zeros<uint32_t>(16 * (size_t)4 * cap);
		const uint32_t			   nresp = resolver.host_resolve(x, face_tree, edge_tree, resp, 4 * cap, &choices, &st);

		lcs::ContactMesh cm;
		cm.num_verts = nv, cm.num_edges = ne, cm.edges = edges, cm.faces = faces, cm.d_hat = dhat, cm.offset = th;
		cm.rest_vert_area = va, cm.rest_edge_area = ea, cm.rest_face_area = fa;
		lcs::NarrowPhasesDetector narrow(dev, cm);
		uint32_t*				  ccount = zeros<uint32_t>(1);
		const uint32_t			  ccap = 8 * nv + 64;
		uint32_t*				  crec = zeros<uint32_t>(16 * (size_t)ccap);
		uint32_t*				  cids = zeros<uint32_t>(2 * (size_t)ccap);
		narrow.vf_dcd_query_repulsion(x, ccount, crec, cids, ccap, face_tree);
		const uint32_t nvf = read_u32(ccount);
		narrow.ee_dcd_query_repulsion(x, ccount, crec, cids, ccap, edge_tree);
		const uint32_t nprox = read_u32(ccount);
		dev.synchronize();
		int chosen = 0;
		for (const auto& c : choices)
			chosen += c.combo >= 0;
		std::printf("{\"ef_pairs\": %u, \"ef_total\": %u, \"contours\": %u, \"responses\": %u, \"chosen\": %d, "
					"\"vf_pairs\": %u, \"proximity_pairs\": %u}\n",
			nef, pc.first, pc.second, nresp, chosen, nvf, nprox);
		return 0;
	}
	catch (const lcs::Error& e)
	{
		std::printf("{\"error\": %d, \"message\": \"%s\"}\n", e.code, e.what());
		return 3;
	}
	catch (const std::exception& e)
	{
		std::printf("{\"error\": -100, \"message\": \"%s\"}\n", e.what());
		return 4;
	}
}
