// Reverse of the mesh geometry on the device: from the adjoints of the
// per-edge and per-cell quantities the kernels read (accumulated by the
// adjoint sweep's geometry terms, layouts as cfd2d::geometry_vjp's in
// mesh.cpp) to the adjoint of every node coordinate.  The arithmetic is
// geometry_vjp's; each of its scatters (edges into cells, cells into edges,
// cells and edges into nodes) is a gather here that adds its terms in the
// order the host loop adds them, and contraction into FMAs is off, so the
// two give the same bits.  Included at the end of solver.hip.

namespace {

// Device copy of the mesh pieces the reverse reads, and its work arrays, in
// one allocation per context (rebuilt when a different mesh is passed).
struct GeoDev {
	int nn, nc, ne;
	const double* xy;   // [nn][2]
	const double* cxy;  // [nc][2] centroids
	const double* S;    // [nc]
	const double* eg;   // [7][ne]: nx, ny, l, midpoint x, y, periodic shift x, y
	const int* ce;      // [nc][3] edges, ascending
	const int* c12;     // [ne][2] c1, c2 (-1 on a boundary)
	const int* n12;     // [ne][2] n1, n2
	const int* noff;    // [nn + 1] start of each node's list in nitem
	const int* nitem;   // per node: 4 i (cell i), 4 e + 1 (edge e's n1), 4 e + 2 (its n2),
	                    // cells ascending then edges ascending (the host's order)
	double* ck;    // [4][3 nc] per (cell, face): LSQ d-bar (or boundary dn-bar n) /
	               // Green-Gauss n-bar contribution (x, y), GG l-bar contribution, Heron l-bar
	double* eb;    // [4][ne] LSQ boundary edges: n-bar and midpoint-bar contributions; then
	               // each edge's node terms (n1 x, y, n2 x, y)
	double* cb;    // [2][nc] centroid adjoints
	double* lny;   // [ne]
	double* dxy;   // [nn][2]
};

__device__ __forceinline__ int face_slot(const GeoDev& g, int cell, int e)
{
	const int* c = g.ce + 3 * (size_t)cell;
	return c[0] == e ? 0 : (c[1] == e ? 1 : 2);
}

// per cell: its contributions to its faces' edges and to its neighbours'
// centroids (mesh.cpp geometry_vjp, the Gc loop and the Heron loop)
__global__ __launch_bounds__(BLOCK) void k_geo_cell(GeoDev g, const double* Gc, int lsq)
{
#pragma clang fp contract(off)
	int i = blockIdx.x * BLOCK + threadIdx.x;
	if (i >= g.nc) return;
	const size_t NC = g.nc, NE = g.ne, N3 = 3 * NC;
	double cx = g.cxy[2 * (size_t)i], cy = g.cxy[2 * (size_t)i + 1];
	double Sb = 0.0 + Gc[i];
	int es[3];
	for (int k = 0; k < 3; k++) es[k] = g.ce[3 * (size_t)i + k];
	double v[3][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}}, lk[3] = {0.0, 0.0, 0.0};
	if (lsq) {
		// w_k = A^-1 d_k, A = sum d d^T: d-bar_k = u_k + (A-bar + A-bar^T) d_k,
		// u_k = A^-1 w-bar_k, A-bar = -sum_k u_k w_k^T
		double d[3][2], mid[3][2];
		for (int k = 0; k < 3; k++) {
			int e = es[k], c1 = g.c12[2 * (size_t)e], c2 = g.c12[2 * (size_t)e + 1];
			double nx = g.eg[e], ny = g.eg[NE + e];
			mid[k][0] = g.eg[3 * NE + e];
			mid[k][1] = g.eg[4 * NE + e];
			if (c2 < 0) {
				double dn = (mid[k][0] - cx) * nx + (mid[k][1] - cy) * ny;
				d[k][0] = 2.0 * dn * nx;
				d[k][1] = 2.0 * dn * ny;
			} else {
				double sx = g.eg[5 * NE + e], sy = g.eg[6 * NE + e];
				double x1 = g.cxy[2 * (size_t)c1], y1 = g.cxy[2 * (size_t)c1 + 1];
				double x2 = g.cxy[2 * (size_t)c2], y2 = g.cxy[2 * (size_t)c2 + 1];
				if (i == c1) {
					d[k][0] = (x2 + sx) - x1;
					d[k][1] = (y2 + sy) - y1;
				} else {
					d[k][0] = x1 - (x2 + sx);
					d[k][1] = y1 - (y2 + sy);
				}
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
		double B11 = 0, B12 = 0, B21 = 0, B22 = 0;
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
			if (g.c12[2 * (size_t)e + 1] < 0) {
				// d = 2 dn n, dn = (mid - c) . n; this cell is the edge's only one
				double nx = g.eg[e], ny = g.eg[NE + e];
				double dn = (mid[k][0] - cx) * nx + (mid[k][1] - cy) * ny;
				double dnb = 2.0 * (nx * dbx + ny * dby);
				g.eb[e] = 2.0 * dn * dbx + dnb * (mid[k][0] - cx);
				g.eb[NE + e] = 2.0 * dn * dby + dnb * (mid[k][1] - cy);
				g.eb[2 * NE + e] = dnb * nx;
				g.eb[3 * NE + e] = dnb * ny;
				v[k][0] = dnb * nx;  // this centroid: -v
				v[k][1] = dnb * ny;
			} else {
				v[k][0] = dbx;  // this centroid -v, the neighbour's +v
				v[k][1] = dby;
			}
		}
	} else {
		for (int k = 0; k < 3; k++) {
			int e = es[k];
			double sg = g.c12[2 * (size_t)e] == i ? 1.0 : -1.0;  // weight = sg n l
			double wbx = Gc[(size_t)(1 + 2 * k) * NC + i], wby = Gc[(size_t)(2 + 2 * k) * NC + i];
			double l = g.eg[2 * NE + e];
			v[k][0] = sg * l * wbx;
			v[k][1] = sg * l * wby;
			lk[k] = sg * (g.eg[e] * wbx + g.eg[NE + e] * wby);
		}
	}
	// Heron: S = sqrt(Q), Q = p (p - a)(p - b)(p - c), p = (a + b + c)/2
	double hk[3] = {0.0, 0.0, 0.0};
	if (Sb != 0.0) {
		double x[3] = {g.eg[2 * NE + es[0]], g.eg[2 * NE + es[1]], g.eg[2 * NE + es[2]]};
		double p = (x[0] + x[1] + x[2]) / 2.0;
		double f[3] = {p - x[0], p - x[1], p - x[2]};
		double S = g.S[i];
		for (int k = 0; k < 3; k++) {
			double dp = f[0] * f[1] * f[2] + p * (f[1] * f[2] + f[0] * f[2] + f[0] * f[1]);
			double dq = 0.5 * dp - p * (f[(k + 1) % 3] * f[(k + 2) % 3]);
			hk[k] = Sb * dq / (2.0 * S);
		}
	}
	for (int k = 0; k < 3; k++) {
		size_t s = 3 * (size_t)i + k;
		g.ck[s] = v[k][0];
		g.ck[N3 + s] = v[k][1];
		g.ck[2 * N3 + s] = lk[k];
		g.ck[3 * N3 + s] = hk[k];
	}
}

// per cell: the centroid adjoint -- the Gauss-point offsets of its edges in
// ascending edge order, then (least squares) the d-bar terms of every cell
// that names this one, in ascending (cell, face) order; on one face the host
// adds the other cell's + term before its own cell's - term
__global__ __launch_bounds__(BLOCK) void k_geo_centroid(GeoDev g, const double* Ge, int lsq)
{
#pragma clang fp contract(off)
	int j = blockIdx.x * BLOCK + threadIdx.x;
	if (j >= g.nc) return;
	const size_t NE = g.ne, N3 = 3 * (size_t)g.nc;
	double bx = 0.0, by = 0.0;
	for (int k = 0; k < 3; k++) {
		int e = g.ce[3 * (size_t)j + k];
		for (int s = 0; s < 2; s++) {
			if (g.c12[2 * (size_t)e + s] != j) continue;
			for (int q = 0; q < 2; q++) {
				bx -= Ge[(size_t)(4 + 4 * s + 2 * q) * NE + e];
				by -= Ge[(size_t)(5 + 4 * s + 2 * q) * NE + e];
			}
		}
	}
	if (lsq) {
		// keys 8 (contributing cell) + 2 (its face) + (0: +, 1: -) with the
		// signed terms, sorted
		long long key[6];
		double tx[6], ty[6];
		int n = 0;
		for (int k = 0; k < 3; k++) {
			int e = g.ce[3 * (size_t)j + k];
			size_t s = 3 * (size_t)j + k;
			key[n] = 8ll * j + 2 * k + 1;
			tx[n] = -g.ck[s];
			ty[n] = -g.ck[N3 + s];
			n++;
			int c1 = g.c12[2 * (size_t)e], c2 = g.c12[2 * (size_t)e + 1];
			if (c2 < 0) continue;
			int o = c1 == j ? c2 : c1;  // o == j: a periodic face of the cell with itself
			int so = o == j ? k : face_slot(g, o, e);
			size_t t = 3 * (size_t)o + so;
			key[n] = 8ll * o + 2 * so;
			tx[n] = g.ck[t];
			ty[n] = g.ck[N3 + t];
			n++;
		}
		for (int a = 1; a < n; a++)
			for (int b = a; b > 0 && key[b - 1] > key[b]; b--) {
				long long kk = key[b];
				key[b] = key[b - 1];
				key[b - 1] = kk;
				double x = tx[b];
				tx[b] = tx[b - 1];
				tx[b - 1] = x;
				x = ty[b];
				ty[b] = ty[b - 1];
				ty[b - 1] = x;
			}
		for (int a = 0; a < n; a++) {
			bx += tx[a];
			by += ty[a];
		}
	}
	g.cb[j] = bx;
	g.cb[g.nc + (size_t)j] = by;
}

// per edge: normal, length, midpoint and Gauss-point adjoints, then its two
// nodes' terms (into eb, which held this edge's boundary terms)
__global__ __launch_bounds__(BLOCK) void k_geo_edge(GeoDev g, const double* Ge, const double* Gc, int lsq, double sq3)
{
#pragma clang fp contract(off)
	int e = blockIdx.x * BLOCK + threadIdx.x;
	if (e >= g.ne) return;
	const size_t NE = g.ne, N3 = 3 * (size_t)g.nc;
	int c1 = g.c12[2 * (size_t)e], c2 = g.c12[2 * (size_t)e + 1];
	double nbx = Ge[e] + Ge[2 * NE + e], nby = Ge[NE + e] + Ge[3 * NE + e];
	double lb = Ge[12 * NE + e];
	double gpb[4] = {0.0, 0.0, 0.0, 0.0};
	for (int s = 0; s < 2; s++) {
		int cell = s ? c2 : c1;
		for (int q = 0; q < 2; q++) {
			double bx = Ge[(size_t)(4 + 4 * s + 2 * q) * NE + e], by = Ge[(size_t)(5 + 4 * s + 2 * q) * NE + e];
			if (cell < 0) continue;
			gpb[2 * q] += bx;
			gpb[2 * q + 1] += by;
		}
	}
	double nx = g.eg[e], ny = g.eg[NE + e], l = g.eg[2 * NE + e];
	lb += g.lny[e] * ny;
	nby += g.lny[e] * l;
	// the cells' terms in ascending cell order: Green-Gauss weights, then Heron
	int cs[2] = {c1, c2};
	int nc2 = c2 >= 0 && c2 != c1 ? 2 : 1;
	if (nc2 == 2 && c2 < c1) {
		cs[0] = c2;
		cs[1] = c1;
	}
	double mbx = 0.0, mby = 0.0;
	if (!lsq) {
		for (int a = 0; a < nc2; a++) {
			size_t t = 3 * (size_t)cs[a] + face_slot(g, cs[a], e);
			nbx += g.ck[t];
			nby += g.ck[N3 + t];
			lb += g.ck[2 * N3 + t];
		}
	} else if (c2 < 0) {
		nbx += g.eb[e];
		nby += g.eb[NE + e];
		mbx += g.eb[2 * NE + e];
		mby += g.eb[3 * NE + e];
	}
	for (int a = 0; a < nc2; a++)
		if (0.0 + Gc[cs[a]] != 0.0) lb += g.ck[3 * N3 + 3 * (size_t)cs[a] + face_slot(g, cs[a], e)];
	// n = sigma (dy, -dx) / l with sigma fixed by the orientation out of c1;
	// Gauss points mid -/+ (1/sqrt 3) d / 2, mid = (P1 + P2) / 2
	int na = g.n12[2 * (size_t)e], nb = g.n12[2 * (size_t)e + 1];
	double dx = g.xy[2 * (size_t)nb] - g.xy[2 * (size_t)na], dy = g.xy[2 * (size_t)nb + 1] - g.xy[2 * (size_t)na + 1];
	double hx = dy / l, hy = -dx / l;
	double sigma = hx * nx + hy * ny >= 0.0 ? 1.0 : -1.0;
	double proj = hx * nbx + hy * nby;
	double vbx = sigma * (nbx - proj * hx) / l, vby = sigma * (nby - proj * hy) / l;
	double ddx = -vby + lb * dx / l, ddy = vbx + lb * dy / l;
	double mx = mbx + gpb[0] + gpb[2], my = mby + gpb[1] + gpb[3];
	ddx += 0.5 * sq3 * (gpb[2] - gpb[0]);
	ddy += 0.5 * sq3 * (gpb[3] - gpb[1]);
	g.eb[e] = 0.5 * mx - ddx;
	g.eb[NE + e] = 0.5 * my - ddy;
	g.eb[2 * NE + e] = 0.5 * mx + ddx;
	g.eb[3 * NE + e] = 0.5 * my + ddy;
}

// per node: the centroid thirds of its cells, then its edges' terms
__global__ __launch_bounds__(BLOCK) void k_geo_node(GeoDev g)
{
#pragma clang fp contract(off)
	int v = blockIdx.x * BLOCK + threadIdx.x;
	if (v >= g.nn) return;
	const size_t NE = g.ne, NC = g.nc;
	double x = 0.0, y = 0.0;
	for (int k = g.noff[v]; k < g.noff[v + 1]; k++) {
		int it = g.nitem[k], idx = it >> 2, tag = it & 3;
		if (tag == 0) {
			x += g.cb[idx] / 3.0;
			y += g.cb[NC + idx] / 3.0;
		} else {
			size_t o = tag == 1 ? 0 : 2 * NE;
			x += g.eb[o + idx];
			y += g.eb[o + NE + idx];
		}
	}
	g.dxy[2 * (size_t)v] = x;
	g.dxy[2 * (size_t)v + 1] = y;
}

// the device copy of mesh m for context c (kept while the same mesh is passed)
int geo_upload(cfd2d_ctx* c, const cfd2d_mesh* mh, GeoDev& g)
{
	const cfd2d::Mesh& m = mh->m;
	int nn = m.nn, nc = m.nc, ne = m.ne;
	size_t nitems = 3 * (size_t)nc + 2 * (size_t)ne;
	if ((size_t)ne * 4 + 3 > (size_t)0x7fffffff || nitems > (size_t)0x7fffffff)
		return fail(CFD2D_ERR_ARG, "adjoint geometry: mesh too large for the node lists");
	size_t nd = 2 * (size_t)nn + 2 * (size_t)nc + nc + 7 * (size_t)ne + 12 * (size_t)nc + 4 * (size_t)ne +
	            2 * (size_t)nc + ne + 2 * (size_t)nn;
	size_t ni = 3 * (size_t)nc + 2 * (size_t)ne + 2 * (size_t)ne + (size_t)nn + 1 + nitems;
	size_t bytes = nd * sizeof(double) + ni * sizeof(int);
	if (c->geo_id != mh->id || c->geo_ws_bytes < bytes) {
		if (c->geo_ws_bytes < bytes) {
			if (c->geo_ws) HIPCHK(hipFree(c->geo_ws));
			c->geo_ws = nullptr;
			c->geo_ws_bytes = 0;
			HIPCHK(hipMalloc(&c->geo_ws, bytes));
			c->geo_ws_bytes = bytes;
		}
		c->geo_id = 0;
	}
	double* d = (double*)c->geo_ws;
	double* xy = d;
	double* cxy = xy + 2 * (size_t)nn;
	double* S = cxy + 2 * (size_t)nc;
	double* eg = S + nc;
	g.ck = eg + 7 * (size_t)ne;
	g.eb = g.ck + 12 * (size_t)nc;
	g.cb = g.eb + 4 * (size_t)ne;
	g.lny = g.cb + 2 * (size_t)nc;
	g.dxy = g.lny + ne;
	int* ip = (int*)(g.dxy + 2 * (size_t)nn);
	int* ce = ip;
	int* c12 = ce + 3 * (size_t)nc;
	int* n12 = c12 + 2 * (size_t)ne;
	int* noff = n12 + 2 * (size_t)ne;
	int* nitem = noff + nn + 1;
	g.nn = nn;
	g.nc = nc;
	g.ne = ne;
	g.xy = xy;
	g.cxy = cxy;
	g.S = S;
	g.eg = eg;
	g.ce = ce;
	g.c12 = c12;
	g.n12 = n12;
	g.noff = noff;
	g.nitem = nitem;
	if (c->geo_id == mh->id) return 0;
	std::vector<double> hd(2 * (size_t)nn + 3 * (size_t)nc + 7 * (size_t)ne);
	std::copy(m.xy.begin(), m.xy.end(), hd.begin());
	double* h = hd.data() + 2 * (size_t)nn;
	for (int i = 0; i < nc; i++) {
		h[2 * (size_t)i] = m.cx[i];
		h[2 * (size_t)i + 1] = m.cy[i];
		h[2 * (size_t)nc + i] = m.S[i];
	}
	h += 3 * (size_t)nc;
	for (int e = 0; e < ne; e++) {
		h[e] = m.nx[e];
		h[(size_t)ne + e] = m.ny[e];
		h[2 * (size_t)ne + e] = m.l[e];
		h[3 * (size_t)ne + e] = m.gp[6 * (size_t)e];
		h[4 * (size_t)ne + e] = m.gp[6 * (size_t)e + 1];
		h[5 * (size_t)ne + e] = m.shx[e];
		h[6 * (size_t)ne + e] = m.shy[e];
	}
	std::vector<int> hi(ni);
	std::copy(m.cedges.begin(), m.cedges.end(), hi.begin());
	int* q = hi.data() + 3 * (size_t)nc;
	for (int e = 0; e < ne; e++) {
		q[2 * (size_t)e] = m.c1[e];
		q[2 * (size_t)e + 1] = m.c2[e];
		q[2 * (size_t)ne + 2 * (size_t)e] = m.n1[e];
		q[2 * (size_t)ne + 2 * (size_t)e + 1] = m.n2[e];
	}
	// node lists: cells ascending (each of its nodes), then edges ascending
	int* off = q + 4 * (size_t)ne;
	int* items = off + nn + 1;
	std::fill(off, off + nn + 1, 0);
	for (size_t k = 0; k < 3 * (size_t)nc; k++) off[m.cnodes[k] + 1]++;
	for (int e = 0; e < ne; e++) {
		off[m.n1[e] + 1]++;
		off[m.n2[e] + 1]++;
	}
	for (int v = 0; v < nn; v++) off[v + 1] += off[v];
	std::vector<int> pos(off, off + nn);
	for (int i = 0; i < nc; i++)
		for (int k = 0; k < 3; k++) items[pos[m.cnodes[3 * (size_t)i + k]]++] = 4 * i;
	for (int e = 0; e < ne; e++) {
		items[pos[m.n1[e]]++] = 4 * e + 1;
		items[pos[m.n2[e]]++] = 4 * e + 2;
	}
	HIPCHK(hipMemcpyAsync(xy, hd.data(), 2 * (size_t)nn * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipMemcpyAsync(cxy, hd.data() + 2 * (size_t)nn, (3 * (size_t)nc + 7 * (size_t)ne) * sizeof(double),
	                      hipMemcpyHostToDevice, c->st));
	HIPCHK(hipMemcpyAsync(ce, hi.data(), ni * sizeof(int), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));  // the host vectors go out of scope
	c->geo_id = mh->id;
	return 0;
}

// dxy (host, [nn][2]) from the device adjoints Ge [13][ne], Gc [7][nc] and
// lny [ne] of the context's mesh
int geo_vjp_device(cfd2d_ctx* c, GeoDev& g, bool lsq, const double* Ge, const double* Gc, double* dxy)
{
	const double sq3 = 1.0 / sqrt(3.0);
	hipLaunchKernelGGL(k_geo_cell, dim3(grid_of(g.nc)), dim3(BLOCK), 0, c->st, g, Gc, (int)lsq);
	hipLaunchKernelGGL(k_geo_centroid, dim3(grid_of(g.nc)), dim3(BLOCK), 0, c->st, g, Ge, (int)lsq);
	hipLaunchKernelGGL(k_geo_edge, dim3(grid_of(g.ne)), dim3(BLOCK), 0, c->st, g, Ge, Gc, (int)lsq, sq3);
	hipLaunchKernelGGL(k_geo_node, dim3(grid_of(g.nn)), dim3(BLOCK), 0, c->st, g);
	HIPCHK(hipMemcpyAsync(dxy, g.dxy, 2 * (size_t)g.nn * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return check_launch("adjoint geometry reverse");
}

}  // namespace

extern "C" int cfd2d_geometry_vjp_device(cfd2d_ctx* c, const cfd2d_mesh* m, int lsq, const double* Ge,
                                         const double* Gc, const double* lny, double* dxy)
{
	if (!c || !m || !Ge || !Gc || !dxy) return fail(CFD2D_ERR_ARG, "cfd2d_geometry_vjp_device: bad arguments");
	if (m->m.nc != c->nc || m->m.ne != c->ne)
		return fail(CFD2D_ERR_ARG, "cfd2d_geometry_vjp_device: the mesh is not the context's (cell or edge count)");
	GeoDev g;
	int rc = geo_upload(c, m, g);
	if (rc) return rc;
	// the caller's adjoints in device arrays of their own for this call
	size_t ne = c->ne, nc = c->nc;
	double *dGe = nullptr, *dGc = nullptr;
	HIPCHK(hipMalloc(&dGe, 13 * ne * sizeof(double)));
	if (hipMalloc(&dGc, 7 * nc * sizeof(double)) != hipSuccess) {
		(void)hipFree(dGe);
		return fail(CFD2D_ERR_HIP, "hipMalloc");
	}
	hipError_t e1 = hipMemcpyAsync(dGe, Ge, 13 * ne * sizeof(double), hipMemcpyHostToDevice, c->st);
	hipError_t e2 = hipMemcpyAsync(dGc, Gc, 7 * nc * sizeof(double), hipMemcpyHostToDevice, c->st);
	hipError_t e3 = hipSuccess;
	if (lny) e3 = hipMemcpyAsync(g.lny, lny, ne * sizeof(double), hipMemcpyHostToDevice, c->st);
	else launch_fill(c, g.lny, ne);
	if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) rc = geo_vjp_device(c, g, lsq != 0, dGe, dGc, dxy);
	else rc = fail(CFD2D_ERR_HIP, "cfd2d_geometry_vjp_device: copy");
	(void)hipStreamSynchronize(c->st);
	(void)hipFree(dGe);
	(void)hipFree(dGc);
	return rc;
}

extern "C" int cfd2d_adjoint_geometry(cfd2d_ctx* c, const cfd2d_mesh* m, int nsteps, int ckpt_every,
                                      const int* edge_mask, double* J, double* dJdU0, double* dJdbc, double* dJdxy)
{
	if (!c || !m || nsteps < 1 || !edge_mask || !dJdxy) return fail(CFD2D_ERR_ARG, "cfd2d_adjoint_geometry: bad arguments");
	if (m->m.nc != c->nc || m->m.ne != c->ne)
		return fail(CFD2D_ERR_ARG, "cfd2d_adjoint_geometry: the mesh is not the context's (cell or edge count)");
	if (c->p.flux == CFD2D_FLUX_ROE || c->p.limiter != CFD2D_LIMITER_NONE || c->p.viscous || c->p.steady)
		return fail(CFD2D_ERR_CONFIG, "adjoint: needs FLUX GODUNOV, LAX or HLLC, no limiter, inviscid, unsteady");
	// the mesh's device copy first, so the adjoint arena is sized around it
	GeoDev g;
	int rc = geo_upload(c, m, g);
	if (rc) return rc;
	bool timing = c->timing;
	c->timing = false;
	GeoOut geo;
	rc = c->mm ? adjoint_run<true>(c, nsteps, ckpt_every, edge_mask, J, dJdU0, dJdbc, &geo)
	           : adjoint_run<false>(c, nsteps, ckpt_every, edge_mask, J, dJdU0, dJdbc, &geo);
	c->timing = timing;
	if (rc) return rc;
	// the geometry adjoints stay on the device: the node-coordinate reverse
	// runs there and only dJ/dxy comes back
	HIPCHK(hipMemcpyAsync(g.lny, geo.lny.data(), (size_t)c->ne * sizeof(double), hipMemcpyHostToDevice, c->st));
	return geo_vjp_device(c, g, c->p.gradient == CFD2D_GRAD_LSQ, geo.Ge, geo.Gc, dJdxy);
}
