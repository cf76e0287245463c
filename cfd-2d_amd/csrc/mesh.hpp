// Host-side mesh: topology and geometry of a 2-D triangulation.
//
// The builder reproduces the cell/edge numbering, normal orientation, Gauss
// points and Heron areas that src/mesh/MeshReaderSalomeUnv.cpp:61-264 gives
// the reference's Grid, so a mesh passed here as arrays and the same mesh
// read by the reference from its UNV file have identical edge indices and
// bit-identical geometry.  Extensions: periodic pairing of two boundary
// groups, concatenation of several meshes into one index space (batching).
#pragma once

#include <string>
#include <vector>

namespace cfd2d {

struct Mesh {
	int nn = 0, nc = 0, ne = 0;
	std::vector<double> xy;          // 2 per node
	std::vector<int> cnodes;         // 3 per cell
	std::vector<int> cedges;         // 3 per cell, ascending edge index
	std::vector<double> S, cx, cy;   // per cell
	std::vector<int> n1, n2, c1, c2; // per edge; c2 = -1 on a boundary
	std::vector<int> etype;          // 0 inner, 0x100 named boundary
	std::vector<int> egroup;         // edge group index of a boundary edge, -1 inner
	std::vector<double> nx, ny, l;   // unit normal out of c1, length
	std::vector<double> gp;          // 6 per edge: midpoint, Gauss point 1, Gauss point 2
	std::vector<double> shx, shy;    // periodic shift of c2 (0 if none)
	std::vector<std::string> edge_groups, cell_groups;
	std::vector<int> cell_group;     // per cell, -1 if in no group
	std::vector<int> cell_case;      // per cell case index (0 unless concatenated)
	std::vector<int> case_cell_offset;  // n_cases + 1 entries
	std::vector<int> case_edge_offset;  // n_cases + 1 entries
	bool numeric_groups = false;     // groups named by Triangle boundary markers / region attributes
	// what the builder was given, kept so the mesh can be rebuilt in another cell order
	std::vector<int> bnodes, bgroup; // boundary elements: 2 nodes each, group
	int missing_group = -2;          // group of a boundary edge with no element (-2: an error)
	struct Periodic {
		int ga, gb;
		double sx, sy;
	};
	std::vector<Periodic> periodic;  // pairings applied, in order
};

// cell orders for renumber_mesh
enum { ORDER_NONE = 0, ORDER_HILBERT = 1, ORDER_SHUFFLE = 2 };

// Error codes (the reference's Exception types, src/global.h:241-248, where one exists)
enum {
	MESH_OK = 0,
	MESH_ERR_FILE = 301,            // FILE_OPENING_ERROR
	MESH_ERR_UNKNOWN_ELEMENT = 221, // TYPE_MESH_UNV_UNKNOWN_ELEMENT
	MESH_ERR_BND_EDGE = 222,        // TYPE_MESH_UNV_NOT_DEFINED_BND_EDGE
	MESH_ERR_INPUT = 230,
	MESH_ERR_PERIODIC = 302,
};

int build_mesh(int nn, const double* xy, int nc, const int* tri, int nb, const int* bedges,
               const int* bgroup, Mesh& m, std::string& err);
int read_unv(const std::string& path, Mesh& m, std::string& err);
int read_triangle(const std::string& base, Mesh& m, std::string& err);
int pair_periodic(Mesh& m, int group_a, int group_b, double sx, double sy, std::string& err);
int concat_meshes(const std::vector<const Mesh*>& parts, Mesh& out, std::string& err);
// Hilbert-curve position of an integer point of [0, 2^bits)^2
unsigned long long hilbert_index(unsigned long long x, unsigned long long y, int bits);
// The cell permutation of an order (perm[k] = old index of new cell k) and
// the mesh rebuilt in it (nodes numbered by first use, groups and periodic
// pairings carried over); one case only.
int renumber_mesh(const Mesh& in, int order, unsigned long long seed, std::vector<int>& perm, Mesh& out,
                  std::string& err);
// adjoints of the geometry the kernels read -> adjoints of the node
// coordinates (dxy: 2 per node); layouts at the definition
int geometry_vjp(const Mesh& m, bool lsq, const double* Ge, const double* Gc, const double* lny, double* dxy);

}  // namespace cfd2d
