// C ABI: errors and host meshes.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cfd2d.h"
#include "common.hpp"

namespace cfd2d {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
	g_last_error = msg;
	return code;
}

void clear_error() { g_last_error.clear(); }

unsigned long long next_mesh_id()
{
	static std::atomic<unsigned long long> next{1};
	return next.fetch_add(1);
}

}  // namespace cfd2d

using cfd2d::fail;

extern "C" {

const char* cfd2d_error_string(int code)
{
	switch (code) {
	case CFD2D_OK: return "ok";
	case CFD2D_ERR_ARG: return "invalid argument";
	case CFD2D_ERR_HIP: return "HIP runtime error";
	case CFD2D_ERR_NEWTON: return "Riemann solver Newton iteration hit its cap";
	case CFD2D_ERR_NONFINITE: return "non-finite state";
	case CFD2D_ERR_CONFIG: return "configuration error";
	case CFD2D_ERR_BOUNDARY: return "edge without a boundary condition";
	case CFD2D_ERR_MESH_NAME: return "wrong mesh type";
	case CFD2D_ERR_UNV_ELEMENT: return "unknown UNV element type";
	case CFD2D_ERR_UNV_BND_EDGE: return "boundary edge not defined in mesh file";
	case CFD2D_ERR_MESH: return "malformed mesh";
	case CFD2D_ERR_FILE: return "file opening error";
	case CFD2D_ERR_PERIODIC: return "periodic pairing failed";
	default: return "unknown error";
	}
}

const char* cfd2d_last_error(void) { return cfd2d::g_last_error.c_str(); }

int cfd2d_mesh_create(int n_nodes, const double* xy, int n_cells, const int* tri, int n_bedges,
                      const int* bedge_nodes, const int* bedge_group, cfd2d_mesh** out)
{
	if (!out || !xy || !tri || n_nodes <= 0 || n_cells <= 0 || n_bedges < 0 || (n_bedges > 0 && !bedge_nodes))
		return fail(CFD2D_ERR_ARG, "cfd2d_mesh_create: bad arguments");
	auto* h = new (std::nothrow) cfd2d_mesh;
	if (!h) return fail(CFD2D_ERR_ARG, "out of host memory");
	std::string err;
	int rc = cfd2d::build_mesh(n_nodes, xy, n_cells, tri, n_bedges, bedge_nodes, bedge_group, h->m, err);
	if (rc) {
		delete h;
		return fail(rc, err);
	}
	*out = h;
	return CFD2D_OK;
}

int cfd2d_mesh_read(const char* files_type, const char* name, cfd2d_mesh** out)
{
	if (!files_type || !name || !out) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_read: bad arguments");
	auto* h = new (std::nothrow) cfd2d_mesh;
	if (!h) return fail(CFD2D_ERR_ARG, "out of host memory");
	std::string err;
	int rc;
	if (!strcmp(files_type, "salome_unv")) {
		rc = cfd2d::read_unv(name, h->m, err);
	} else if (!strcmp(files_type, "berkeley_triangle")) {
		rc = cfd2d::read_triangle(name, h->m, err);
	} else {
		delete h;
		return fail(CFD2D_ERR_MESH_NAME, "Wrong mesh type.");
	}
	if (rc) {
		delete h;
		return fail(rc, err);
	}
	*out = h;
	return CFD2D_OK;
}

int cfd2d_mesh_set_groups(cfd2d_mesh* m, int kind, int n, const char* const* names)
{
	if (!m || n < 0 || (n > 0 && !names) || (kind != 0 && kind != 1))
		return fail(CFD2D_ERR_ARG, "cfd2d_mesh_set_groups: bad arguments");
	std::vector<std::string>& v = kind == 0 ? m->m.edge_groups : m->m.cell_groups;
	v.clear();
	for (int i = 0; i < n; i++) v.push_back(names[i] ? names[i] : "");
	return CFD2D_OK;
}

int cfd2d_mesh_set_cell_groups(cfd2d_mesh* m, const int* cell_group)
{
	if (!m || !cell_group) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_set_cell_groups: bad arguments");
	m->m.cell_group.assign(cell_group, cell_group + m->m.nc);
	return CFD2D_OK;
}

int cfd2d_mesh_periodic(cfd2d_mesh* m, int ga, int gb, double sx, double sy)
{
	if (!m) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_periodic: null mesh");
	std::string err;
	int rc = cfd2d::pair_periodic(m->m, ga, gb, sx, sy, err);
	m->id = cfd2d::next_mesh_id();  // c2 and the shifts may have changed
	return rc ? fail(CFD2D_ERR_PERIODIC, err) : CFD2D_OK;
}

int cfd2d_mesh_renumber(const cfd2d_mesh* m, int order, int seed, int* cell_perm, cfd2d_mesh** out)
{
	if (!m || !out) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_renumber: bad arguments");
	auto* h = new (std::nothrow) cfd2d_mesh;
	if (!h) return fail(CFD2D_ERR_ARG, "out of host memory");
	std::string err;
	std::vector<int> perm;
	int rc = cfd2d::renumber_mesh(m->m, order, (unsigned long long)(unsigned)seed, perm, h->m, err);
	if (rc) {
		delete h;
		return fail(rc == cfd2d::MESH_ERR_PERIODIC ? CFD2D_ERR_PERIODIC : CFD2D_ERR_MESH, err);
	}
	if (cell_perm) std::copy(perm.begin(), perm.end(), cell_perm);
	*out = h;
	return CFD2D_OK;
}

int cfd2d_mesh_geometry_vjp(const cfd2d_mesh* m, int lsq, const double* Ge, const double* Gc, const double* lny,
                            double* dxy)
{
	if (!m || !dxy) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_geometry_vjp: bad arguments");
	cfd2d::geometry_vjp(m->m, lsq != 0, Ge, Gc, lny, dxy);
	return CFD2D_OK;
}

int cfd2d_mesh_concat(int n, const cfd2d_mesh* const* parts, cfd2d_mesh** out)
{
	if (n <= 0 || !parts || !out) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_concat: bad arguments");
	std::vector<const cfd2d::Mesh*> v;
	for (int i = 0; i < n; i++) {
		if (!parts[i]) return fail(CFD2D_ERR_ARG, "cfd2d_mesh_concat: null part");
		v.push_back(&parts[i]->m);
	}
	auto* h = new (std::nothrow) cfd2d_mesh;
	if (!h) return fail(CFD2D_ERR_ARG, "out of host memory");
	std::string err;
	int rc = cfd2d::concat_meshes(v, h->m, err);
	if (rc) {
		delete h;
		return fail(rc, err);
	}
	*out = h;
	return CFD2D_OK;
}

int cfd2d_mesh_counts(const cfd2d_mesh* m, int* nn, int* nc, int* ne)
{
	if (!m) return fail(CFD2D_ERR_ARG, "null mesh");
	if (nn) *nn = m->m.nn;
	if (nc) *nc = m->m.nc;
	if (ne) *ne = m->m.ne;
	return CFD2D_OK;
}

int cfd2d_mesh_nodes(const cfd2d_mesh* m, double* xy)
{
	if (!m || !xy) return fail(CFD2D_ERR_ARG, "bad arguments");
	memcpy(xy, m->m.xy.data(), sizeof(double) * m->m.xy.size());
	return CFD2D_OK;
}

int cfd2d_mesh_group_count(const cfd2d_mesh* m, int kind, int* n)
{
	if (!m || !n) return fail(CFD2D_ERR_ARG, "bad arguments");
	*n = (int)(kind == 0 ? m->m.edge_groups.size() : m->m.cell_groups.size());
	return CFD2D_OK;
}

int cfd2d_mesh_group_name(const cfd2d_mesh* m, int kind, int i, char* buf, int len)
{
	if (!m || !buf || len <= 0) return fail(CFD2D_ERR_ARG, "bad arguments");
	const auto& v = kind == 0 ? m->m.edge_groups : m->m.cell_groups;
	if (i < 0 || i >= (int)v.size()) return fail(CFD2D_ERR_ARG, "group index out of range");
	strncpy(buf, v[i].c_str(), (size_t)len - 1);
	buf[len - 1] = 0;
	return CFD2D_OK;
}

int cfd2d_mesh_get_cell_groups(const cfd2d_mesh* m, int* cg)
{
	if (!m || !cg) return fail(CFD2D_ERR_ARG, "bad arguments");
	memcpy(cg, m->m.cell_group.data(), sizeof(int) * m->m.nc);
	return CFD2D_OK;
}

int cfd2d_mesh_case_offsets(const cfd2d_mesh* m, int* n_cases, int* co, int* eo)
{
	if (!m) return fail(CFD2D_ERR_ARG, "null mesh");
	int n = (int)m->m.case_cell_offset.size() - 1;
	if (n_cases) *n_cases = n;
	if (co) memcpy(co, m->m.case_cell_offset.data(), sizeof(int) * (n + 1));
	if (eo) memcpy(eo, m->m.case_edge_offset.data(), sizeof(int) * (n + 1));
	return CFD2D_OK;
}

int cfd2d_mesh_export(const cfd2d_mesh* h, double* cell_geom, int* cell_nodes, int* cell_edges,
                      double* edge_geom, int* edge_int, int* edge_group, double* edge_shift)
{
	if (!h) return fail(CFD2D_ERR_ARG, "null mesh");
	const cfd2d::Mesh& m = h->m;
	for (int i = 0; i < m.nc; i++) {
		if (cell_geom) {
			cell_geom[4 * (size_t)i] = m.S[i];
			cell_geom[4 * (size_t)i + 1] = m.cx[i];
			cell_geom[4 * (size_t)i + 2] = m.cy[i];
			cell_geom[4 * (size_t)i + 3] = 0.0;
		}
		for (int k = 0; k < 3; k++) {
			if (cell_nodes) cell_nodes[3 * (size_t)i + k] = m.cnodes[3 * (size_t)i + k];
			if (cell_edges) cell_edges[3 * (size_t)i + k] = m.cedges[3 * (size_t)i + k];
		}
	}
	for (int e = 0; e < m.ne; e++) {
		if (edge_geom) {
			double* g = &edge_geom[9 * (size_t)e];
			g[0] = m.nx[e];
			g[1] = m.ny[e];
			g[2] = m.l[e];
			for (int q = 0; q < 6; q++) g[3 + q] = m.gp[6 * (size_t)e + q];
		}
		if (edge_int) {
			int* t = &edge_int[5 * (size_t)e];
			t[0] = m.n1[e];
			t[1] = m.n2[e];
			t[2] = m.c1[e];
			t[3] = m.c2[e];
			t[4] = m.etype[e];
		}
		if (edge_group) edge_group[e] = m.egroup[e];
		if (edge_shift) {
			edge_shift[2 * (size_t)e] = m.shx[e];
			edge_shift[2 * (size_t)e + 1] = m.shy[e];
		}
	}
	return CFD2D_OK;
}

void cfd2d_mesh_destroy(cfd2d_mesh* m) { delete m; }

}  // extern "C"
