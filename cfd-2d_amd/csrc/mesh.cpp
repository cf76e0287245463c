#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <tuple>

namespace cfd2d {

namespace {

struct HalfEdge {
	int a, b, cell, slot;
	bool operator<(const HalfEdge& o) const
	{
		return std::tie(a, b, cell, slot) < std::tie(o.a, o.b, o.cell, o.slot);
	}
};

struct BKey {
	int a, b, idx;
	bool operator<(const BKey& o) const { return std::tie(a, b, idx) < std::tie(o.a, o.b, o.idx); }
};

}  // namespace

/*
 * Cell/edge construction in the order of src/mesh/MeshReaderSalomeUnv.cpp:
 *  - centroid ((x0 + x1) + x2) / 3 (89-90);
 *  - neighbour across local edge k = (node k, node k+1): the last cell, in
 *    ascending index, other than this one that holds both nodes (100-115);
 *  - an edge is created by the first cell (ascending) that meets it, local
 *    edges in order 0,1,2; a cell records its edges as they are created, so
 *    its list is ascending (118-217);
 *  - n1 = node j, n2 = node j+1 of c1; normal (y2-y1, x1-x2)/l, flipped to
 *    point away from c1's centroid (170-187);
 *  - points: midpoint, midpoint -/+ (n2-n1)/(2 sqrt 3) (160-169);
 *  - Heron area over the cell's edge lengths in list order (250-255).
 * A boundary edge takes the group of the first listed boundary element with
 * the same two nodes (find_edge, 14-22).
 */
int build_mesh_ex(int nn, const double* xy, int nc, const int* tri, int nb, const int* bedges,
                  const int* bgroup, Mesh& m, std::string& err, int missing_group)
{
	if (nn <= 0 || nc <= 0 || !xy || !tri) {
		err = "empty mesh";
		return MESH_ERR_INPUT;
	}
	for (size_t k = 0; k < 3 * (size_t)nc; k++)
		if (tri[k] < 0 || tri[k] >= nn) {
			err = "triangle node index out of range";
			return MESH_ERR_INPUT;
		}
	for (size_t k = 0; k < 2 * (size_t)nb; k++)
		if (bedges[k] < 0 || bedges[k] >= nn) {
			err = "boundary edge node index out of range";
			return MESH_ERR_INPUT;
		}
	m.nn = nn;
	m.nc = nc;
	m.xy.assign(xy, xy + 2 * (size_t)nn);
	m.cnodes.assign(tri, tri + 3 * (size_t)nc);
	m.S.resize(nc);
	m.cx.resize(nc);
	m.cy.resize(nc);
	for (int i = 0; i < nc; i++) {
		const int* t = tri + 3 * (size_t)i;
		m.cx[i] = (xy[2 * t[0]] + xy[2 * t[1]] + xy[2 * t[2]]) / 3.0;
		m.cy[i] = (xy[2 * t[0] + 1] + xy[2 * t[1] + 1] + xy[2 * t[2] + 1]) / 3.0;
	}
	size_t nh = 3 * (size_t)nc;
	std::vector<HalfEdge> h(nh);
	for (int i = 0; i < nc; i++)
		for (int k = 0; k < 3; k++) {
			int a = tri[3 * i + k], b = tri[3 * i + (k + 1) % 3];
			h[3 * (size_t)i + k] = HalfEdge{std::min(a, b), std::max(a, b), i, k};
		}
	std::sort(h.begin(), h.end());
	std::vector<int> neigh(nh);
	for (size_t s = 0; s < nh;) {
		size_t t = s;
		while (t < nh && h[t].a == h[s].a && h[t].b == h[s].b) t++;
		for (size_t q = s; q < t; q++) {
			int self = h[q].cell, other = -2;
			for (size_t r = s; r < t; r++)
				if (h[r].cell != self) other = h[r].cell;
			neigh[3 * (size_t)self + h[q].slot] = other;
		}
		s = t;
	}
	h.clear();
	h.shrink_to_fit();
	int ne = 0;
	for (int i = 0; i < nc; i++)
		for (int j = 0; j < 3; j++) {
			int p = neigh[3 * (size_t)i + j];
			if (p > -1) {
				for (int k = 0; k < 3; k++)
					if (neigh[3 * (size_t)p + k] == i) neigh[3 * (size_t)p + k] = -1;
				ne++;
			}
			if (p == -2) ne++;
		}
	m.ne = ne;
	m.n1.resize(ne);
	m.n2.resize(ne);
	m.c1.resize(ne);
	m.c2.resize(ne);
	m.etype.resize(ne);
	m.egroup.assign(ne, -1);
	m.nx.resize(ne);
	m.ny.resize(ne);
	m.l.resize(ne);
	m.gp.resize(6 * (size_t)ne);
	m.shx.assign(ne, 0.0);
	m.shy.assign(ne, 0.0);
	m.cedges.assign(3 * (size_t)nc, -1);
	std::vector<int> cfi(nc, 0);
	std::vector<BKey> bk(nb);
	for (int k = 0; k < nb; k++) {
		int a = bedges[2 * k], b = bedges[2 * k + 1];
		bk[k] = BKey{std::min(a, b), std::max(a, b), k};
	}
	std::sort(bk.begin(), bk.end());
	const double sq3 = 1.0 / sqrt(3.0);
	int ie = 0, rc = MESH_OK;
	for (int i = 0; i < nc; i++)
		for (int j = 0; j < 3; j++) {
			int p = neigh[3 * (size_t)i + j];
			if (p == -1) continue;
			int a = tri[3 * i + j], b = tri[3 * i + (j + 1) % 3];
			double x1 = xy[2 * a], y1 = xy[2 * a + 1], x2 = xy[2 * b], y2 = xy[2 * b + 1];
			m.n1[ie] = a;
			m.n2[ie] = b;
			double* g = &m.gp[6 * (size_t)ie];
			g[0] = (x1 + x2) / 2.0;
			g[1] = (y1 + y2) / 2.0;
			g[2] = (x1 + x2) / 2.0 - sq3 * (x2 - x1) / 2.0;
			g[3] = (y1 + y2) / 2.0 - sq3 * (y2 - y1) / 2.0;
			g[4] = (x1 + x2) / 2.0 + sq3 * (x2 - x1) / 2.0;
			g[5] = (y1 + y2) / 2.0 + sq3 * (y2 - y1) / 2.0;
			double nx = y2 - y1, ny = x1 - x2;
			double l = sqrt(nx * nx + ny * ny);
			nx /= l;
			ny /= l;
			double vx = m.cx[i] - g[0], vy = m.cy[i] - g[1];
			if (vx * nx + vy * ny > 0) {
				nx *= -1;
				ny *= -1;
			}
			m.nx[ie] = nx;
			m.ny[ie] = ny;
			m.l[ie] = l;
			m.c1[ie] = i;
			m.cedges[3 * (size_t)i + cfi[i]++] = ie;
			if (p > -1) {
				m.c2[ie] = p;
				m.cedges[3 * (size_t)p + cfi[p]++] = ie;
				m.etype[ie] = 0;
			} else {
				m.c2[ie] = -1;
				m.etype[ie] = 0x100;
				BKey key{std::min(a, b), std::max(a, b), -1};
				auto it = std::lower_bound(bk.begin(), bk.end(), key);
				if (it != bk.end() && it->a == key.a && it->b == key.b) {
					m.egroup[ie] = bgroup ? bgroup[it->idx] : -1;
				} else if (missing_group > -2) {
					m.egroup[ie] = missing_group;
				} else {
					char buf[96];
					snprintf(buf, sizeof(buf), "Boundary edge #%d not defined in mesh file.", ie);
					err = buf;
					rc = MESH_ERR_BND_EDGE;
				}
			}
			ie++;
		}
	for (int i = 0; i < nc; i++) {
		if (cfi[i] != 3) {
			err = "non-manifold mesh: a cell does not have three edges";
			return MESH_ERR_INPUT;
		}
		double a = m.l[m.cedges[3 * (size_t)i]], b = m.l[m.cedges[3 * (size_t)i + 1]],
		       c = m.l[m.cedges[3 * (size_t)i + 2]];
		double p = (a + b + c) / 2.0;
		m.S[i] = sqrt(p * (p - a) * (p - b) * (p - c));
	}
	m.cell_group.assign(nc, -1);
	m.cell_case.assign(nc, 0);
	m.case_cell_offset = {0, nc};
	m.case_edge_offset = {0, ne};
	m.bnodes.assign(bedges, bedges + 2 * (size_t)nb);
	m.bgroup.assign(nb, -1);
	if (bgroup) m.bgroup.assign(bgroup, bgroup + nb);
	m.missing_group = missing_group;
	m.periodic.clear();
	return rc;
}

int build_mesh(int nn, const double* xy, int nc, const int* tri, int nb, const int* bedges,
               const int* bgroup, Mesh& m, std::string& err)
{
	return build_mesh_ex(nn, xy, nc, tri, nb, bedges, bgroup, m, err, -2);
}

// ---------------------------------------------------------------- UNV

namespace {

// A line closes/opens a block when the first "-1" in it sits at its end
// (MeshReaderSalomeUnv::read_block, MeshReaderSalomeUnv.cpp:267-286).
bool is_delim(const std::string& s)
{
	size_t pos = s.find("-1");
	return pos != std::string::npos && pos == s.size() - 2;
}

}  // namespace

/*
 * Salome UNV reader with the semantics of src/mesh/MeshReaderSalomeUnv.cpp:
 * datasets 2411 (nodes, in order), 2412 (type 11 = boundary edge,
 * type 41 = triangle, anything else is error 221) and 2467 (named groups of
 * element labels).  Groups are applied in name order, so an element in
 * several groups keeps the last name (the std::map walk at 220-248).
 */
int read_unv(const std::string& path, Mesh& m, std::string& err)
{
	std::ifstream f(path);
	if (!f.is_open()) {
		err = "File '" + path + "' opening error.";
		return MESH_ERR_FILE;
	}
	std::vector<double> xy;
	std::vector<int> tri, tri_label, bed, bed_label;
	std::map<std::string, std::vector<int>> groups;
	std::vector<std::string> block;
	int delims = 0;
	std::string line;
	auto parse = [&](const std::vector<std::string>& sl) -> int {
		if (sl.empty()) return 0;
		int type = atoi(sl[0].c_str());
		size_t i = 1;
		if (type == 2411) {
			xy.clear();
			while (i + 1 < sl.size()) {
				double x = 0, y = 0, z = 0;
				sscanf(sl[i + 1].c_str(), "%lf %lf %lf", &x, &y, &z);
				xy.push_back(x);
				xy.push_back(y);
				i += 2;
			}
		} else if (type == 2412) {
			tri.clear(); tri_label.clear(); bed.clear(); bed_label.clear();
			while (i < sl.size()) {
				int t[6] = {0, 0, 0, 0, 0, 0};
				sscanf(sl[i].c_str(), "%d %d %d %d %d %d", &t[0], &t[1], &t[2], &t[3], &t[4], &t[5]);
				i++;
				if (t[1] == 11) {
					int a = 0, b = 0;
					if (i + 1 >= sl.size()) break;
					i++;  // beam record
					sscanf(sl[i].c_str(), "%d %d", &a, &b);
					i++;
					bed.push_back(a - 1);
					bed.push_back(b - 1);
					bed_label.push_back(t[0] - 1);
				} else if (t[1] == 41) {
					int a = 0, b = 0, c = 0;
					if (i >= sl.size()) break;
					sscanf(sl[i].c_str(), "%d %d %d", &a, &b, &c);
					i++;
					tri.push_back(a - 1);
					tri.push_back(b - 1);
					tri.push_back(c - 1);
					tri_label.push_back(t[0] - 1);
				} else {
					char buf[64];
					snprintf(buf, sizeof(buf), "Unknown element type '%d'.", t[1]);
					err = buf;
					return MESH_ERR_UNKNOWN_ELEMENT;
				}
			}
		} else if (type == 2467) {
			groups.clear();
			while (i < sl.size()) {
				int t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
				sscanf(sl[i].c_str(), "%d %d %d %d %d %d %d %d", &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6], &t[7]);
				i++;
				int n = t[7];
				char name[128] = {0};
				if (i < sl.size()) sscanf(sl[i].c_str(), "%127s", name);
				i++;
				std::vector<int> labels;
				for (int k = 0; k < n / 2 && i < sl.size(); k++, i++) {
					int u[8] = {0, 0, 0, 0, 0, 0, 0, 0};
					sscanf(sl[i].c_str(), "%d %d %d %d %d %d %d %d", &u[0], &u[1], &u[2], &u[3], &u[4], &u[5], &u[6], &u[7]);
					labels.push_back(u[1] - 1);
					labels.push_back(u[5] - 1);
				}
				if (n % 2 == 1 && i < sl.size()) {
					int u[4] = {0, 0, 0, 0};
					sscanf(sl[i].c_str(), "%d %d %d %d", &u[0], &u[1], &u[2], &u[3]);
					labels.push_back(u[1] - 1);
					i++;
				}
				groups[std::string(name)] = labels;
			}
		}
		return 0;
	};
	while (std::getline(f, line)) {
		if (!line.empty() && line.back() == '\r') line.pop_back();
		if (is_delim(line)) {
			if (++delims == 2) {
				int rc = parse(block);
				if (rc) return rc;
				block.clear();
				delims = 0;
			}
			continue;
		}
		block.push_back(line);
	}
	if (!block.empty()) {
		int rc = parse(block);
		if (rc) return rc;
	}
	int nn = (int)(xy.size() / 2), nc = (int)tri_label.size(), nb = (int)bed_label.size();
	// element label -> group index (groups walked in name order, last wins)
	int maxlab = 0;
	for (int v : tri_label) maxlab = std::max(maxlab, v);
	for (int v : bed_label) maxlab = std::max(maxlab, v);
	std::vector<int> label_group(maxlab + 1, -1);
	std::vector<std::string> names;
	for (auto& g : groups) {
		int gi = (int)names.size();
		names.push_back(g.first);
		for (int lab : g.second)
			if (lab >= 0 && lab <= maxlab) label_group[lab] = gi;
	}
	std::vector<int> bgroup(nb), cgroup(nc);
	for (int k = 0; k < nb; k++) bgroup[k] = bed_label[k] >= 0 ? label_group[bed_label[k]] : -1;
	for (int k = 0; k < nc; k++) cgroup[k] = tri_label[k] >= 0 ? label_group[tri_label[k]] : -1;
	int rc = build_mesh(nn, xy.data(), nc, tri.data(), nb, bed.data(), bgroup.data(), m, err);
	if (rc) return rc;
	m.edge_groups = names;
	m.cell_groups = names;
	m.cell_group = cgroup;
	return MESH_OK;
}

// ---------------------------------------------------------------- Triangle

/*
 * Berkeley Triangle files (.node, .ele, .poly), as read by
 * src/mesh/MeshReaderBerkleyTriangle.cpp:3-160: 1-based node indices, the
 * region attribute of each triangle names its cell group, and the boundary
 * marker of each .poly segment names its edge group (edges with no segment
 * get marker -1).  Deviation: edges here carry the two Gauss points of the
 * UNV reader; the reference's Berkeley reader stores only the midpoint, so
 * its explicit method computes no flux on such meshes (fvm_tvd.cpp:341).
 */
int read_triangle(const std::string& base, Mesh& m, std::string& err)
{
	auto open = [&](const char* ext, std::ifstream& f) {
		f.open(base + ext);
		if (!f.is_open()) {
			err = "File '" + base + ext + "' opening error.";
			return false;
		}
		return true;
	};
	auto next_line = [](std::ifstream& f, std::string& s) {
		while (std::getline(f, s)) {
			size_t h = s.find('#');
			if (h != std::string::npos) s = s.substr(0, h);
			if (s.find_first_not_of(" \t\r") != std::string::npos) return true;
		}
		return false;
	};
	std::ifstream fn, fe, fp;
	if (!open(".node", fn) || !open(".ele", fe) || !open(".poly", fp)) return MESH_ERR_FILE;
	std::string s;
	int nn = 0, dim = 0, na = 0, nbm = 0;
	if (!next_line(fn, s) || sscanf(s.c_str(), "%d %d %d %d", &nn, &dim, &na, &nbm) < 1 || nn <= 0) {
		err = "bad .node header";
		return MESH_ERR_INPUT;
	}
	std::vector<double> xy(2 * (size_t)nn);
	int first_index = 1;
	for (int i = 0; i < nn; i++) {
		int idx = 0;
		if (!next_line(fn, s) || sscanf(s.c_str(), "%d %lf %lf", &idx, &xy[2 * i], &xy[2 * i + 1]) < 3) {
			err = "bad .node record";
			return MESH_ERR_INPUT;
		}
		if (i == 0) first_index = idx;
	}
	int nc = 0, npt = 0, nat = 0;
	if (!next_line(fe, s) || sscanf(s.c_str(), "%d %d %d", &nc, &npt, &nat) < 1 || nc <= 0) {
		err = "bad .ele header";
		return MESH_ERR_INPUT;
	}
	std::vector<int> tri(3 * (size_t)nc), attr(nc, 0);
	for (int i = 0; i < nc; i++) {
		int idx, a, b, c, t = 0;
		if (!next_line(fe, s) || sscanf(s.c_str(), "%d %d %d %d %d", &idx, &a, &b, &c, &t) < 4) {
			err = "bad .ele record";
			return MESH_ERR_INPUT;
		}
		tri[3 * i] = a - first_index;
		tri[3 * i + 1] = b - first_index;
		tri[3 * i + 2] = c - first_index;
		attr[i] = t;
	}
	int t0, t1, t2, t3;
	if (!next_line(fp, s)) { err = "bad .poly header"; return MESH_ERR_INPUT; }
	sscanf(s.c_str(), "%d %d %d %d", &t0, &t1, &t2, &t3);
	int nseg = 0, segbm = 0;
	if (!next_line(fp, s) || sscanf(s.c_str(), "%d %d", &nseg, &segbm) < 1) {
		err = "bad .poly segment header";
		return MESH_ERR_INPUT;
	}
	std::vector<int> bed, mark;
	for (int i = 0; i < nseg; i++) {
		int idx, a, b, mk = 0;
		if (!next_line(fp, s) || sscanf(s.c_str(), "%d %d %d %d", &idx, &a, &b, &mk) < 3) {
			err = "bad .poly segment";
			return MESH_ERR_INPUT;
		}
		bed.push_back(a - first_index);
		bed.push_back(b - first_index);
		mark.push_back(mk);
	}
	// group names: decimal markers / attributes, -1 for unlisted boundary edges
	std::vector<int> mvals(mark);
	mvals.push_back(-1);
	std::sort(mvals.begin(), mvals.end());
	mvals.erase(std::unique(mvals.begin(), mvals.end()), mvals.end());
	std::vector<int> bgroup(mark.size());
	for (size_t k = 0; k < mark.size(); k++)
		bgroup[k] = (int)(std::lower_bound(mvals.begin(), mvals.end(), mark[k]) - mvals.begin());
	int missing = (int)(std::lower_bound(mvals.begin(), mvals.end(), -1) - mvals.begin());
	int rc = build_mesh_ex(nn, xy.data(), nc, tri.data(), (int)mark.size(), bed.data(), bgroup.data(), m, err,
	                       missing);
	if (rc) return rc;
	m.edge_groups.clear();
	for (int v : mvals) m.edge_groups.push_back(std::to_string(v));
	std::vector<int> avals(attr);
	std::sort(avals.begin(), avals.end());
	avals.erase(std::unique(avals.begin(), avals.end()), avals.end());
	m.cell_groups.clear();
	for (int v : avals) m.cell_groups.push_back(std::to_string(v));
	for (int i = 0; i < nc; i++)
		m.cell_group[i] = (int)(std::lower_bound(avals.begin(), avals.end(), attr[i]) - avals.begin());
	m.numeric_groups = true;
	return MESH_OK;
}

// ---------------------------------------------------------------- periodic

/*
 * Extension (no reference code): joins each boundary edge of group_a with
 * the boundary edge of group_b whose midpoint plus (sx, sy) coincides with
 * its own (to 1e-9).  The joined edge keeps group_a's index, c1 and normal;
 * its c2 is group_b's cell, whose centroid is seen shifted by (sx, sy);
 * group_b's edge is removed and the remaining edges keep their order.
 */
int pair_periodic(Mesh& m, int ga, int gb, double sx, double sy, std::string& err)
{
	struct K {
		long long x, y;
		int e;
		bool operator<(const K& o) const { return std::tie(x, y, e) < std::tie(o.x, o.y, o.e); }
	};
	std::vector<K> A, B;
	for (int e = 0; e < m.ne; e++) {
		if (m.c2[e] >= 0) continue;
		const double* g = &m.gp[6 * (size_t)e];
		if (m.egroup[e] == ga) A.push_back(K{llround(g[0] * 1e9), llround(g[1] * 1e9), e});
		else if (m.egroup[e] == gb) B.push_back(K{llround((g[0] + sx) * 1e9), llround((g[1] + sy) * 1e9), e});
	}
	if (A.empty() || A.size() != B.size()) {
		err = "periodic groups have different edge counts";
		return MESH_ERR_PERIODIC;
	}
	std::sort(A.begin(), A.end());
	std::sort(B.begin(), B.end());
	std::vector<int> target(m.ne, -1);  // removed edge -> surviving edge
	for (size_t k = 0; k < A.size(); k++) {
		if (A[k].x != B[k].x || A[k].y != B[k].y) {
			err = "periodic groups do not match under the given shift";
			return MESH_ERR_PERIODIC;
		}
		int ea = A[k].e, eb = B[k].e;
		m.c2[ea] = m.c1[eb];
		m.shx[ea] = sx;
		m.shy[ea] = sy;
		m.etype[ea] = 0;
		m.egroup[ea] = -1;
		target[eb] = ea;
	}
	std::vector<int> newid(m.ne, -1);
	int w = 0;
	for (int e = 0; e < m.ne; e++)
		if (target[e] < 0) newid[e] = w++;
	for (int e = 0; e < m.ne; e++)
		if (target[e] >= 0) newid[e] = newid[target[e]];
	int ne2 = w;
	w = 0;
	for (int e = 0; e < m.ne; e++) {
		if (target[e] >= 0) continue;
		m.n1[w] = m.n1[e]; m.n2[w] = m.n2[e]; m.c1[w] = m.c1[e]; m.c2[w] = m.c2[e];
		m.etype[w] = m.etype[e]; m.egroup[w] = m.egroup[e];
		m.nx[w] = m.nx[e]; m.ny[w] = m.ny[e]; m.l[w] = m.l[e];
		for (int q = 0; q < 6; q++) m.gp[6 * (size_t)w + q] = m.gp[6 * (size_t)e + q];
		m.shx[w] = m.shx[e]; m.shy[w] = m.shy[e];
		w++;
	}
	m.ne = ne2;
	for (auto* v : {&m.n1, &m.n2, &m.c1, &m.c2, &m.etype, &m.egroup}) v->resize(ne2);
	for (auto* v : {&m.nx, &m.ny, &m.l, &m.shx, &m.shy}) v->resize(ne2);
	m.gp.resize(6 * (size_t)ne2);
	for (int i = 0; i < m.nc; i++) {
		int* ce = &m.cedges[3 * (size_t)i];
		for (int k = 0; k < 3; k++) ce[k] = newid[ce[k]];
		std::sort(ce, ce + 3);
	}
	if (m.case_edge_offset.size() == 2) m.case_edge_offset[1] = ne2;
	m.periodic.push_back(Mesh::Periodic{ga, gb, sx, sy});
	return MESH_OK;
}

// ---------------------------------------------------------------- renumbering

/*
 * Extension: the cell order decides how far apart in memory the cells and
 * edges that one kernel gathers together lie (the mesh builder numbers edges
 * in cell-discovery order).  Meshes from external generators come in the
 * generator's order; the reference partitions its dual graph with METIS for
 * MPI (src/methods/decomp.cpp:86-104).  For locality on one device the
 * cells are sorted here along a Hilbert curve through their centroids
 * (quantised to a 2^16 x 2^16 grid over their bounding box; ties keep the
 * old order) and the nodes are numbered by first use in that order, as
 * cfd2d/renumber.py does for generator meshes.  ORDER_SHUFFLE sorts by a
 * seeded SplitMix64 hash of the old index instead, to measure what an
 * unordered mesh costs.
 */
unsigned long long hilbert_index(unsigned long long x, unsigned long long y, int bits)
{
	unsigned long long n = 1ull << bits, d = 0;
	for (unsigned long long s = n >> 1; s > 0; s >>= 1) {
		unsigned long long rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
		d += s * s * ((3 * rx) ^ ry);
		if (!ry) {  // rotate the quadrant so the curve stays continuous
			if (rx) {
				x = n - 1 - x;
				y = n - 1 - y;
			}
			std::swap(x, y);
		}
	}
	return d;
}

namespace {

unsigned long long splitmix64(unsigned long long z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

}  // namespace

int renumber_mesh(const Mesh& in, int order, unsigned long long seed, std::vector<int>& perm, Mesh& out,
                  std::string& err)
{
	if (in.case_cell_offset.size() > 2) {
		err = "renumbering a concatenated mesh: renumber its parts before concatenating";
		return MESH_ERR_INPUT;
	}
	if (order < ORDER_NONE || order > ORDER_SHUFFLE) {
		err = "unknown cell order";
		return MESH_ERR_INPUT;
	}
	int nc = in.nc, nn = in.nn;
	std::vector<unsigned long long> key(nc, 0);
	if (order == ORDER_HILBERT) {
		double lo[2] = {in.cx[0], in.cy[0]}, hi[2] = {in.cx[0], in.cy[0]};
		for (int i = 0; i < nc; i++) {
			lo[0] = std::min(lo[0], in.cx[i]);
			hi[0] = std::max(hi[0], in.cx[i]);
			lo[1] = std::min(lo[1], in.cy[i]);
			hi[1] = std::max(hi[1], in.cy[i]);
		}
		double span[2] = {hi[0] > lo[0] ? hi[0] - lo[0] : 1.0, hi[1] > lo[1] ? hi[1] - lo[1] : 1.0};
		for (int i = 0; i < nc; i++) {
			long long qx = std::min((long long)((in.cx[i] - lo[0]) / span[0] * 65535.0), 65535LL);
			long long qy = std::min((long long)((in.cy[i] - lo[1]) / span[1] * 65535.0), 65535LL);
			key[i] = hilbert_index((unsigned long long)qx, (unsigned long long)qy, 16);
		}
	} else if (order == ORDER_SHUFFLE) {
		unsigned long long base = splitmix64(seed);
		for (int i = 0; i < nc; i++) key[i] = splitmix64(base + (unsigned long long)i);
	}
	perm.resize(nc);
	for (int i = 0; i < nc; i++) perm[i] = i;
	std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
	// nodes by first use in the new cell order (unused nodes keep their order, last)
	std::vector<int> nid(nn, -1);
	int w = 0;
	for (int k = 0; k < nc; k++)
		for (int q = 0; q < 3; q++) {
			int v = in.cnodes[3 * (size_t)perm[k] + q];
			if (nid[v] < 0) nid[v] = w++;
		}
	for (int v = 0; v < nn; v++)
		if (nid[v] < 0) nid[v] = w++;
	std::vector<double> xy(2 * (size_t)nn);
	for (int v = 0; v < nn; v++) {
		xy[2 * (size_t)nid[v]] = in.xy[2 * (size_t)v];
		xy[2 * (size_t)nid[v] + 1] = in.xy[2 * (size_t)v + 1];
	}
	std::vector<int> tri(3 * (size_t)nc), bn(in.bnodes.size());
	for (int k = 0; k < nc; k++)
		for (int q = 0; q < 3; q++) tri[3 * (size_t)k + q] = nid[in.cnodes[3 * (size_t)perm[k] + q]];
	for (size_t k = 0; k < bn.size(); k++) bn[k] = nid[in.bnodes[k]];
	int nb = (int)(bn.size() / 2);
	out = Mesh();
	int rc = build_mesh_ex(nn, xy.data(), nc, tri.data(), nb, bn.data(), in.bgroup.data(), out, err,
	                       in.missing_group);
	if (rc) return rc;
	out.edge_groups = in.edge_groups;
	out.cell_groups = in.cell_groups;
	out.numeric_groups = in.numeric_groups;
	for (int k = 0; k < nc; k++) out.cell_group[k] = in.cell_group[perm[k]];
	for (const Mesh::Periodic& p : in.periodic)
		if ((rc = pair_periodic(out, p.ga, p.gb, p.sx, p.sy, err))) return rc;
	return MESH_OK;
}

// ---------------------------------------------------------------- batching

/*
 * Extension: several independent meshes in one index space.  Nodes, cells
 * and edges of part k follow those of part k-1; group names become
 * "<k>:<name>"; cell_case records the part of each cell.
 */
int concat_meshes(const std::vector<const Mesh*>& parts, Mesh& o, std::string& err)
{
	if (parts.empty()) {
		err = "no meshes";
		return MESH_ERR_INPUT;
	}
	o = Mesh();
	o.case_cell_offset.push_back(0);
	o.case_edge_offset.push_back(0);
	int k = 0;
	for (const Mesh* p : parts) {
		int nodeoff = o.nn, celloff = o.nc, edgeoff = o.ne;
		int egoff = (int)o.edge_groups.size(), cgoff = (int)o.cell_groups.size();
		o.xy.insert(o.xy.end(), p->xy.begin(), p->xy.end());
		for (int v : p->cnodes) o.cnodes.push_back(v + nodeoff);
		for (int v : p->cedges) o.cedges.push_back(v + edgeoff);
		o.S.insert(o.S.end(), p->S.begin(), p->S.end());
		o.cx.insert(o.cx.end(), p->cx.begin(), p->cx.end());
		o.cy.insert(o.cy.end(), p->cy.begin(), p->cy.end());
		for (int e = 0; e < p->ne; e++) {
			o.n1.push_back(p->n1[e] + nodeoff);
			o.n2.push_back(p->n2[e] + nodeoff);
			o.c1.push_back(p->c1[e] + celloff);
			o.c2.push_back(p->c2[e] >= 0 ? p->c2[e] + celloff : -1);
			o.etype.push_back(p->etype[e]);
			o.egroup.push_back(p->egroup[e] >= 0 ? p->egroup[e] + egoff : -1);
		}
		o.nx.insert(o.nx.end(), p->nx.begin(), p->nx.end());
		o.ny.insert(o.ny.end(), p->ny.begin(), p->ny.end());
		o.l.insert(o.l.end(), p->l.begin(), p->l.end());
		o.gp.insert(o.gp.end(), p->gp.begin(), p->gp.end());
		o.shx.insert(o.shx.end(), p->shx.begin(), p->shx.end());
		o.shy.insert(o.shy.end(), p->shy.begin(), p->shy.end());
		for (auto& g : p->edge_groups) o.edge_groups.push_back(std::to_string(k) + ":" + g);
		for (auto& g : p->cell_groups) o.cell_groups.push_back(std::to_string(k) + ":" + g);
		for (int v : p->cell_group) o.cell_group.push_back(v >= 0 ? v + cgoff : -1);
		for (int i = 0; i < p->nc; i++) o.cell_case.push_back(k);
		o.nn += p->nn;
		o.nc += p->nc;
		o.ne += p->ne;
		o.case_cell_offset.push_back(o.nc);
		o.case_edge_offset.push_back(o.ne);
		k++;
	}
	return MESH_OK;
}

}  // namespace cfd2d

// ---------------------------------------------------------------- geometry adjoint

namespace cfd2d {

/*
 * Extension: the reverse of the geometry above (centroids, edge normals,
 * lengths and Gauss points, Heron areas, the explicit method's Green-Gauss
 * weights n l and least-squares weights A^-1 d_k) -- given the adjoints of
 * the quantities the device kernels read, the adjoint of every node
 * coordinate.  Inputs (any pointer may be NULL):
 *   Ge [13][ne]  per edge: normal adjoint (x, y) at Gauss points 1 and 2
 *                (summed), the adjoint of side s's offset to Gauss point g
 *                (x, y) at 4 + 4 s + 2 g, the length adjoint at 12
 *   Gc [7][nc]   per cell: area adjoint, then per face k the adjoint of its
 *                Green-Gauss weight sign * n l (x, y) or LSQ weight (x, y)
 *   lny [ne]     adjoint of l * n_y (the wall-force objective's edge term)
 * The normal's orientation (out of c1) and the periodic shifts are held
 * fixed: derivatives are those of the smooth pieces the builder evaluates.
 */
int geometry_vjp(const Mesh& m, bool lsq, const double* Ge, const double* Gc, const double* lny, double* dxy)
{
	const int nc = m.nc, ne = m.ne, nn = m.nn;
	const size_t NE = (size_t)ne, NC = (size_t)nc;
	std::vector<double> nb(2 * NE, 0.0), lb(NE, 0.0), gpb(4 * NE, 0.0), midb(2 * NE, 0.0), cb(2 * NC, 0.0);
	std::vector<double> Sb(NC, 0.0);
	for (int e = 0; e < ne; e++) {
		if (Ge) {
			nb[2 * e] = Ge[0 * NE + e] + Ge[2 * NE + e];
			nb[2 * e + 1] = Ge[1 * NE + e] + Ge[3 * NE + e];
			lb[e] = Ge[12 * NE + e];
			for (int s = 0; s < 2; s++) {
				int cell = s ? m.c2[e] : m.c1[e];
				for (int g = 0; g < 2; g++) {
					double bx = Ge[(size_t)(4 + 4 * s + 2 * g) * NE + e], by = Ge[(size_t)(5 + 4 * s + 2 * g) * NE + e];
					if (cell < 0) continue;
					gpb[4 * e + 2 * g] += bx;  // offset = Gauss point - (centroid + shift)
					gpb[4 * e + 2 * g + 1] += by;
					cb[2 * cell] -= bx;
					cb[2 * cell + 1] -= by;
				}
			}
		}
		if (lny) {
			lb[e] += lny[e] * m.ny[e];
			nb[2 * e + 1] += lny[e] * m.l[e];
		}
	}
	if (Gc) {
		for (int i = 0; i < nc; i++) {
			Sb[i] += Gc[i];
			if (lsq) {
				// w_k = A^-1 d_k, A = sum d d^T: d-bar_k = u_k + (A-bar + A-bar^T) d_k with
				// u_k = A^-1 w-bar_k and A-bar = -sum_k u_k w_k^T
				double d[3][2], mid[3][2];
				int es[3], side[3];
				for (int k = 0; k < 3; k++) {
					int e = m.cedges[3 * (size_t)i + k];
					es[k] = e;
					side[k] = m.c1[e] == i ? 0 : 1;
					int c1 = m.c1[e], c2 = m.c2[e];
					mid[k][0] = m.gp[6 * (size_t)e];
					mid[k][1] = m.gp[6 * (size_t)e + 1];
					if (c2 < 0) {
						double dn = (mid[k][0] - m.cx[i]) * m.nx[e] + (mid[k][1] - m.cy[i]) * m.ny[e];
						d[k][0] = 2.0 * dn * m.nx[e];
						d[k][1] = 2.0 * dn * m.ny[e];
					} else if (i == c1) {
						d[k][0] = (m.cx[c2] + m.shx[e]) - m.cx[c1];
						d[k][1] = (m.cy[c2] + m.shy[e]) - m.cy[c1];
					} else {
						d[k][0] = m.cx[c1] - (m.cx[c2] + m.shx[e]);
						d[k][1] = m.cy[c1] - (m.cy[c2] + m.shy[e]);
					}
				}
				double a11 = 0, a12 = 0, a22 = 0;
				for (int k = 0; k < 3; k++) {
					a11 += d[k][0] * d[k][0];
					a12 += d[k][0] * d[k][1];
					a22 += d[k][1] * d[k][1];
				}
				double det = a11 * a22 - a12 * a12;
				double w[3][2], u[3][2];
				double B11 = 0, B12 = 0, B21 = 0, B22 = 0;  // A-bar
				for (int k = 0; k < 3; k++) {
					w[k][0] = (a22 * d[k][0] - a12 * d[k][1]) / det;
					w[k][1] = (a11 * d[k][1] - a12 * d[k][0]) / det;
					double wbx = Gc[(size_t)(1 + 2 * k) * NC + i], wby = Gc[(size_t)(2 + 2 * k) * NC + i];
					u[k][0] = (a22 * wbx - a12 * wby) / det;
					u[k][1] = (a11 * wby - a12 * wbx) / det;
					B11 -= u[k][0] * w[k][0];
					B12 -= u[k][0] * w[k][1];
					B21 -= u[k][1] * w[k][0];
					B22 -= u[k][1] * w[k][1];
				}
				for (int k = 0; k < 3; k++) {
					double dbx = u[k][0] + (2.0 * B11) * d[k][0] + (B12 + B21) * d[k][1];
					double dby = u[k][1] + (B12 + B21) * d[k][0] + (2.0 * B22) * d[k][1];
					int e = es[k];
					int c1 = m.c1[e], c2 = m.c2[e];
					if (c2 < 0) {
						// d = 2 dn n, dn = (mid - c) . n
						double nx = m.nx[e], ny = m.ny[e];
						double dn = (mid[k][0] - m.cx[i]) * nx + (mid[k][1] - m.cy[i]) * ny;
						double dnb = 2.0 * (nx * dbx + ny * dby);
						nb[2 * e] += 2.0 * dn * dbx + dnb * (mid[k][0] - m.cx[i]);
						nb[2 * e + 1] += 2.0 * dn * dby + dnb * (mid[k][1] - m.cy[i]);
						midb[2 * e] += dnb * nx;
						midb[2 * e + 1] += dnb * ny;
						cb[2 * i] -= dnb * nx;
						cb[2 * i + 1] -= dnb * ny;
					} else {
						int plus = side[k] == 0 ? c2 : c1, minus = side[k] == 0 ? c1 : c2;
						cb[2 * plus] += dbx;
						cb[2 * plus + 1] += dby;
						cb[2 * minus] -= dbx;
						cb[2 * minus + 1] -= dby;
					}
				}
			} else {
				for (int k = 0; k < 3; k++) {
					int e = m.cedges[3 * (size_t)i + k];
					double sg = m.c1[e] == i ? 1.0 : -1.0;  // weight = sg n l
					double wbx = Gc[(size_t)(1 + 2 * k) * NC + i], wby = Gc[(size_t)(2 + 2 * k) * NC + i];
					nb[2 * e] += sg * m.l[e] * wbx;
					nb[2 * e + 1] += sg * m.l[e] * wby;
					lb[e] += sg * (m.nx[e] * wbx + m.ny[e] * wby);
				}
			}
		}
	}
	// Heron: S = sqrt(Q), Q = p (p - a)(p - b)(p - c), p = (a + b + c)/2, over the cell's edges
	for (int i = 0; i < nc; i++) {
		if (Sb[i] == 0.0) continue;
		int e[3] = {m.cedges[3 * (size_t)i], m.cedges[3 * (size_t)i + 1], m.cedges[3 * (size_t)i + 2]};
		double x[3] = {m.l[e[0]], m.l[e[1]], m.l[e[2]]};
		double p = (x[0] + x[1] + x[2]) / 2.0;
		double f[3] = {p - x[0], p - x[1], p - x[2]};
		double S = m.S[i];
		for (int k = 0; k < 3; k++) {
			// dQ/dx_k = dQ/dp / 2 - p (product of the other two factors)
			double dp = f[0] * f[1] * f[2] + p * (f[1] * f[2] + f[0] * f[2] + f[0] * f[1]);
			double dq = 0.5 * dp - p * (f[(k + 1) % 3] * f[(k + 2) % 3]);
			lb[e[k]] += Sb[i] * dq / (2.0 * S);
		}
	}
	std::fill(dxy, dxy + 2 * (size_t)nn, 0.0);
	for (int i = 0; i < nc; i++)
		for (int q = 0; q < 3; q++) {
			int v = m.cnodes[3 * (size_t)i + q];
			dxy[2 * (size_t)v] += cb[2 * i] / 3.0;
			dxy[2 * (size_t)v + 1] += cb[2 * i + 1] / 3.0;
		}
	const double sq3 = 1.0 / sqrt(3.0);
	for (int e = 0; e < ne; e++) {
		int a = m.n1[e], b = m.n2[e];
		double dx = m.xy[2 * (size_t)b] - m.xy[2 * (size_t)a], dy = m.xy[2 * (size_t)b + 1] - m.xy[2 * (size_t)a + 1];
		double l = m.l[e];
		// n = sigma (dy, -dx) / l; sigma fixed by the orientation out of c1
		double hx = dy / l, hy = -dx / l;
		double sigma = hx * m.nx[e] + hy * m.ny[e] >= 0.0 ? 1.0 : -1.0;
		double nbx = nb[2 * e], nby = nb[2 * e + 1];
		double proj = hx * nbx + hy * nby;
		double vbx = sigma * (nbx - proj * hx) / l, vby = sigma * (nby - proj * hy) / l;
		double ddx = -vby + lb[e] * dx / l, ddy = vbx + lb[e] * dy / l;
		// Gauss points mid -/+ (1/sqrt 3) d / 2, mid = (P1 + P2) / 2
		double mbx = midb[2 * e] + gpb[4 * e] + gpb[4 * e + 2], mby = midb[2 * e + 1] + gpb[4 * e + 1] + gpb[4 * e + 3];
		ddx += 0.5 * sq3 * (gpb[4 * e + 2] - gpb[4 * e]);
		ddy += 0.5 * sq3 * (gpb[4 * e + 3] - gpb[4 * e + 1]);
		dxy[2 * (size_t)a] += 0.5 * mbx - ddx;
		dxy[2 * (size_t)a + 1] += 0.5 * mby - ddy;
		dxy[2 * (size_t)b] += 0.5 * mbx + ddx;
		dxy[2 * (size_t)b + 1] += 0.5 * mby + ddy;
	}
	return MESH_OK;
}

}  // namespace cfd2d
