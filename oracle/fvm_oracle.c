/*
 * CPU restatement of the reference's explicit finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker the HIP path is compared with and
 * the CPU baseline that bench.py times.  The product library never links,
 * loads or calls this file.
 *
 * Structure: the reference scatters edge terms into cells in edge order
 * (src/methods/fvm_tvd.cpp:253-292 gradients, 329-365 residual).  Here each
 * cell gathers its edges in ascending edge index, which adds the same terms
 * in the same order, so the result is bit-identical to the scatter and the
 * loops parallelise without atomics.  Primitive variables are cached per cell
 * per stage; convertConsToPar (fvm_tvd.cpp:803-813) is a pure function of the
 * cell's conserved state, so every use sees the value the reference
 * recomputes.
 *
 * Parity status per function:
 *   pinned by the reference's golden vectors (tests/golden, tests/test_oracle.py):
 *     mesh geometry, EOS/primitive conversion, Green-Gauss gradients,
 *     reconstruction, boundary ghosts, rim_orig, per-edge fluxes, residual,
 *     reference time step, Heun RK2, bounds flags and remediation, steady mode
 *   extensions pinned by invariants/closed forms (tests/test_oracle_ext.py):
 *     LSQ gradients (exact for linear fields), minmod limiter (bounds),
 *     LAX/HLLC/Roe fluxes (consistency, Rankine-Hugoniot, symmetry),
 *     viscous flux (Couette/shear closed form), SSP-RK3 (order on a linear
 *     ODE), periodic pairing (conservation, translation invariance),
 *     spectral time step (uniform-state closed form), pressure outlet.
 *   implicit-method matrix fill (or_implicit_assemble, or_roe_jacobians;
 *   the reference's fvm_tvd_implicit.cpp does not build here: it includes
 *   HYPRE headers from an empty submodule) pinned by tests/test_implicit_cpu.py:
 *     Roe property A(U_roe)(U_R - U_L) = F(U_R) - F(U_L), A+ + A- = A/2 at
 *     the Roe state against the analytic Euler Jacobian, spectra of A+ and
 *     A-, supersonic upwinding, zero column sums on a periodic mesh,
 *     RHS = the explicit residual (golden-pinned) on a periodic mesh,
 *     free-stream RHS = 0, ghost closed forms, steady tau closed form.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double GAS_R = 8.314472; /* src/global.cpp:6 */

struct or_mesh {
	int nn, nc, ne;
	double* xy;
	int* cnodes;   /* 3 per cell */
	int* cedges;   /* 3 per cell, ascending edge index */
	double *S, *cx, *cy;
	int *n1, *n2, *c1, *c2, *etype, *egroup;
	double *nx, *ny, *l;
	double* gp;    /* 6 per edge: midpoint, Gauss point 1, Gauss point 2 */
	double *shx, *shy;
};

typedef struct {
	double r, p, e, E, u, v, cz, T;
} prm;

struct or_solver {
	const or_mesh* m;
	or_cfg c;
	int* cell_mat;
	int* edge_bc;
	double *U, *U0, *U1, *R, *G, *F, *ctau, *lsq_w;
	prm* P;
	unsigned* flags;
	double TAU;
	long cap_hits;
};

/* ------------------------------------------------------------------ mesh */

typedef struct {
	int a, b, cell, slot;
} half_edge;

static int cmp_half(const void* x, const void* y)
{
	const half_edge* p = (const half_edge*)x;
	const half_edge* q = (const half_edge*)y;
	if (p->a != q->a) return p->a < q->a ? -1 : 1;
	if (p->b != q->b) return p->b < q->b ? -1 : 1;
	if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
	return p->slot - q->slot;
}

typedef struct {
	int a, b, idx;
} bkey;

static int cmp_bkey(const void* x, const void* y)
{
	const bkey* p = (const bkey*)x;
	const bkey* q = (const bkey*)y;
	if (p->a != q->a) return p->a < q->a ? -1 : 1;
	if (p->b != q->b) return p->b < q->b ? -1 : 1;
	return p->idx - q->idx;
}

/*
 * Topology and geometry as src/mesh/MeshReaderSalomeUnv.cpp:61-264 builds
 * them: centroids (89-90); neighbour across (node k, node k+1) = last cell in
 * ascending order sharing both nodes (100-115); edges created by the first
 * cell that meets them, in (cell, local edge) order (118-217), so each cell's
 * edge list is ascending; unit normal out of c1 (170-187); midpoint and the
 * two Gauss points mid -/+ (n2-n1)/(2 sqrt 3) (160-169); Heron area over the
 * cell's edge lengths in list order (250-255).  Boundary edges take the group
 * of the first listed boundary element with the same nodes (find_edge,
 * 14-22); none -> error 222 (TYPE_MESH_UNV_NOT_DEFINED_BND_EDGE).
 */
int or_mesh_build(int nn, const double* xy, int nc, const int* tri, int nb, const int* bedges,
                  const int* bgroup, or_mesh** out)
{
	or_mesh* m = (or_mesh*)calloc(1, sizeof(or_mesh));
	m->nn = nn;
	m->nc = nc;
	m->xy = (double*)malloc(sizeof(double) * 2 * (size_t)nn);
	memcpy(m->xy, xy, sizeof(double) * 2 * (size_t)nn);
	m->cnodes = (int*)malloc(sizeof(int) * 3 * (size_t)nc);
	memcpy(m->cnodes, tri, sizeof(int) * 3 * (size_t)nc);
	m->S = (double*)malloc(sizeof(double) * nc);
	m->cx = (double*)malloc(sizeof(double) * nc);
	m->cy = (double*)malloc(sizeof(double) * nc);
	for (int i = 0; i < nc; i++) {
		const int* t = tri + 3 * (size_t)i;
		m->cx[i] = (xy[2 * t[0]] + xy[2 * t[1]] + xy[2 * t[2]]) / 3.0;
		m->cy[i] = (xy[2 * t[0] + 1] + xy[2 * t[1] + 1] + xy[2 * t[2] + 1]) / 3.0;
	}
	/* neighbours through a sort of half edges */
	size_t nh = 3 * (size_t)nc;
	half_edge* h = (half_edge*)malloc(sizeof(half_edge) * nh);
	for (int i = 0; i < nc; i++)
		for (int k = 0; k < 3; k++) {
			int a = tri[3 * i + k], b = tri[3 * i + (k + 1) % 3];
			half_edge* e = &h[3 * (size_t)i + k];
			e->a = a < b ? a : b;
			e->b = a < b ? b : a;
			e->cell = i;
			e->slot = k;
		}
	qsort(h, nh, sizeof(half_edge), cmp_half);
	int* neigh = (int*)malloc(sizeof(int) * nh);
	for (size_t s = 0; s < nh;) {
		size_t t = s;
		while (t < nh && h[t].a == h[s].a && h[t].b == h[s].b) t++;
		for (size_t q = s; q < t; q++) {
			int self = h[q].cell, nb_cell = -2;
			for (size_t r = s; r < t; r++)
				if (h[r].cell != self) nb_cell = h[r].cell; /* last in ascending order */
			neigh[3 * (size_t)self + h[q].slot] = nb_cell;
		}
		s = t;
	}
	free(h);
	/* counting pass with the reference's erase-the-back-reference rule */
	int ne = 0;
	for (int i = 0; i < nc; i++)
		for (int j = 0; j < 3; j++) {
			int p = neigh[3 * i + j];
			if (p > -1) {
				for (int k = 0; k < 3; k++)
					if (neigh[3 * p + k] == i) neigh[3 * p + k] = -1;
				ne++;
			}
			if (p == -2) ne++;
		}
	m->ne = ne;
	m->n1 = (int*)malloc(sizeof(int) * ne);
	m->n2 = (int*)malloc(sizeof(int) * ne);
	m->c1 = (int*)malloc(sizeof(int) * ne);
	m->c2 = (int*)malloc(sizeof(int) * ne);
	m->etype = (int*)malloc(sizeof(int) * ne);
	m->egroup = (int*)malloc(sizeof(int) * ne);
	m->nx = (double*)malloc(sizeof(double) * ne);
	m->ny = (double*)malloc(sizeof(double) * ne);
	m->l = (double*)malloc(sizeof(double) * ne);
	m->gp = (double*)malloc(sizeof(double) * 6 * (size_t)ne);
	m->shx = (double*)calloc(ne, sizeof(double));
	m->shy = (double*)calloc(ne, sizeof(double));
	m->cedges = (int*)malloc(sizeof(int) * 3 * (size_t)nc);
	int* cfi = (int*)calloc(nc, sizeof(int));
	bkey* bk = (bkey*)malloc(sizeof(bkey) * (nb > 0 ? nb : 1));
	for (int k = 0; k < nb; k++) {
		int a = bedges[2 * k], b = bedges[2 * k + 1];
		bk[k].a = a < b ? a : b;
		bk[k].b = a < b ? b : a;
		bk[k].idx = k;
	}
	qsort(bk, nb, sizeof(bkey), cmp_bkey);
	const double sq3 = 1.0 / sqrt(3.0);
	int ie = 0, err = 0;
	for (int i = 0; i < nc; i++)
		for (int j = 0; j < 3; j++) {
			int p = neigh[3 * i + j];
			if (p == -1) continue;
			int a = tri[3 * i + j], b = tri[3 * i + (j + 1) % 3];
			double x1 = xy[2 * a], y1 = xy[2 * a + 1], x2 = xy[2 * b], y2 = xy[2 * b + 1];
			m->n1[ie] = a;
			m->n2[ie] = b;
			double* g = &m->gp[6 * (size_t)ie];
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
			double vx = m->cx[i] - g[0], vy = m->cy[i] - g[1];
			if (vx * nx + vy * ny > 0) {
				nx *= -1;
				ny *= -1;
			}
			m->nx[ie] = nx;
			m->ny[ie] = ny;
			m->l[ie] = l;
			m->c1[ie] = i;
			m->cedges[3 * (size_t)i + cfi[i]++] = ie;
			m->egroup[ie] = -1;
			if (p > -1) {
				m->c2[ie] = p;
				m->cedges[3 * (size_t)p + cfi[p]++] = ie;
				m->etype[ie] = 0;
			} else {
				m->c2[ie] = -1;
				m->etype[ie] = 0x100;
				bkey key = {a < b ? a : b, a < b ? b : a, -1};
				int lo = 0, hi = nb;
				while (lo < hi) {
					int mid = (lo + hi) / 2;
					if (cmp_bkey(&bk[mid], &key) < 0) lo = mid + 1; else hi = mid;
				}
				if (lo < nb && bk[lo].a == key.a && bk[lo].b == key.b)
					m->egroup[ie] = bgroup[bk[lo].idx];
				else
					err = 222;
			}
			ie++;
		}
	for (int i = 0; i < nc; i++) {
		double a = m->l[m->cedges[3 * i]], b = m->l[m->cedges[3 * i + 1]], c = m->l[m->cedges[3 * i + 2]];
		double p = (a + b + c) / 2.0;
		m->S[i] = sqrt(p * (p - a) * (p - b) * (p - c));
	}
	free(cfi);
	free(bk);
	free(neigh);
	*out = m;
	return err;
}

typedef struct {
	long long kx, ky;
	int e;
} mkey;

static int cmp_mkey(const void* x, const void* y)
{
	const mkey* p = (const mkey*)x;
	const mkey* q = (const mkey*)y;
	if (p->kx != q->kx) return p->kx < q->kx ? -1 : 1;
	if (p->ky != q->ky) return p->ky < q->ky ? -1 : 1;
	return p->e - q->e;
}

/*
 * Extension (no reference code): periodic pairing.  Each boundary edge of
 * group_a is joined with the edge of group_b whose midpoint plus (sx,sy)
 * coincides with its own.  The joined edge keeps group_a's index, c1 and
 * normal; c2 is group_b's cell, seen shifted by (sx,sy); group_b's edge is
 * removed and the remaining edges are renumbered in order.
 * Pinned by tests/test_oracle_ext.py::test_periodic_* (conservation,
 * translation invariance of a uniform flow).
 */
int or_mesh_periodic(or_mesh* m, int ga, int gb, double sx, double sy)
{
	int na = 0, nbb = 0;
	for (int e = 0; e < m->ne; e++) {
		if (m->c2[e] >= 0) continue;
		if (m->egroup[e] == ga) na++;
		if (m->egroup[e] == gb) nbb++;
	}
	if (na != nbb || na == 0) return 301;
	mkey* A = (mkey*)malloc(sizeof(mkey) * na);
	mkey* B = (mkey*)malloc(sizeof(mkey) * na);
	int ia = 0, ib = 0;
	for (int e = 0; e < m->ne; e++) {
		if (m->c2[e] >= 0) continue;
		if (m->egroup[e] == ga) {
			A[ia].kx = llround(m->gp[6 * (size_t)e] * 1e9);
			A[ia].ky = llround(m->gp[6 * (size_t)e + 1] * 1e9);
			A[ia++].e = e;
		} else if (m->egroup[e] == gb) {
			B[ib].kx = llround((m->gp[6 * (size_t)e] + sx) * 1e9);
			B[ib].ky = llround((m->gp[6 * (size_t)e + 1] + sy) * 1e9);
			B[ib++].e = e;
		}
	}
	qsort(A, na, sizeof(mkey), cmp_mkey);
	qsort(B, na, sizeof(mkey), cmp_mkey);
	int* remap = (int*)malloc(sizeof(int) * m->ne);
	for (int e = 0; e < m->ne; e++) remap[e] = e;
	for (int k = 0; k < na; k++) {
		if (A[k].kx != B[k].kx || A[k].ky != B[k].ky) {
			free(A); free(B); free(remap);
			return 302;
		}
		int ea = A[k].e, eb = B[k].e;
		m->c2[ea] = m->c1[eb];
		m->shx[ea] = sx;
		m->shy[ea] = sy;
		m->etype[ea] = 0;
		m->egroup[ea] = -1;
		remap[eb] = -1 - ea; /* removed; the cell refers to ea instead */
	}
	free(A);
	free(B);
	int* newid = (int*)malloc(sizeof(int) * m->ne);
	int ne2 = 0;
	for (int e = 0; e < m->ne; e++) newid[e] = remap[e] >= 0 ? ne2++ : -1;
	for (int e = 0; e < m->ne; e++)
		if (remap[e] < 0) newid[e] = newid[-1 - remap[e]];
	int w = 0;
	for (int e = 0; e < m->ne; e++) {
		if (remap[e] < 0) continue;
		m->n1[w] = m->n1[e]; m->n2[w] = m->n2[e];
		m->c1[w] = m->c1[e]; m->c2[w] = m->c2[e];
		m->etype[w] = m->etype[e]; m->egroup[w] = m->egroup[e];
		m->nx[w] = m->nx[e]; m->ny[w] = m->ny[e]; m->l[w] = m->l[e];
		memmove(&m->gp[6 * (size_t)w], &m->gp[6 * (size_t)e], 6 * sizeof(double));
		m->shx[w] = m->shx[e]; m->shy[w] = m->shy[e];
		w++;
	}
	for (int i = 0; i < m->nc; i++) {
		int* ce = &m->cedges[3 * (size_t)i];
		for (int k = 0; k < 3; k++) ce[k] = newid[ce[k]];
		for (int a = 0; a < 3; a++)
			for (int b = a + 1; b < 3; b++)
				if (ce[b] < ce[a]) { int t = ce[a]; ce[a] = ce[b]; ce[b] = t; }
	}
	m->ne = ne2;
	free(newid);
	free(remap);
	return 0;
}

void or_mesh_free(or_mesh* m)
{
	if (!m) return;
	free(m->xy); free(m->cnodes); free(m->cedges); free(m->S); free(m->cx); free(m->cy);
	free(m->n1); free(m->n2); free(m->c1); free(m->c2); free(m->etype); free(m->egroup);
	free(m->nx); free(m->ny); free(m->l); free(m->gp); free(m->shx); free(m->shy);
	free(m);
}

void or_mesh_counts(const or_mesh* m, int* nn, int* nc, int* ne)
{
	*nn = m->nn;
	*nc = m->nc;
	*ne = m->ne;
}

void or_mesh_export(const or_mesh* m, double* cell_geom, int* cell_nodes, int* cell_edges,
                    double* edge_geom, int* edge_int, int* edge_group, double* edge_shift)
{
	for (int i = 0; i < m->nc; i++) {
		if (cell_geom) {
			cell_geom[4 * i] = m->S[i];
			cell_geom[4 * i + 1] = m->cx[i];
			cell_geom[4 * i + 2] = m->cy[i];
			cell_geom[4 * i + 3] = 0.0;
		}
		for (int k = 0; k < 3; k++) {
			if (cell_nodes) cell_nodes[3 * i + k] = m->cnodes[3 * i + k];
			if (cell_edges) cell_edges[3 * i + k] = m->cedges[3 * i + k];
		}
	}
	for (int e = 0; e < m->ne; e++) {
		if (edge_geom) {
			double* g = &edge_geom[9 * (size_t)e];
			g[0] = m->nx[e];
			g[1] = m->ny[e];
			g[2] = m->l[e];
			memcpy(g + 3, &m->gp[6 * (size_t)e], 6 * sizeof(double));
		}
		if (edge_int) {
			int* t = &edge_int[5 * (size_t)e];
			t[0] = m->n1[e]; t[1] = m->n2[e]; t[2] = m->c1[e]; t[3] = m->c2[e]; t[4] = m->etype[e];
		}
		if (edge_group) edge_group[e] = m->egroup[e];
		if (edge_shift) {
			edge_shift[2 * e] = m->shx[e];
			edge_shift[2 * e + 1] = m->shy[e];
		}
	}
}

/* ------------------------------------------------------------ physics */

/* Material::URS, src/global.cpp:9-30 */
static void urs(const or_solver* s, int mat, prm* q, int flag)
{
	double M = s->c.M[mat], Cp = s->c.Cp[mat];
	double Cv = Cp - GAS_R / M;
	double gam = Cp / Cv;
	switch (flag) {
	case 0:
		q->p = q->r * q->e * (gam - 1);
		q->cz = sqrt(gam * q->p / q->r);
		break;
	case 1:
		q->e = q->p / (q->r * (gam - 1));
		q->T = q->e / Cv;
		break;
	case 2:
		q->r = q->p * M / (q->T * GAS_R);
		q->cz = sqrt(gam * q->p / q->r);
		break;
	}
}

/* FVM_TVD::convertConsToPar, src/methods/fvm_tvd.cpp:803-813 */
static void cons2prim(const or_solver* s, const double* U, int nc, int i, prm* q)
{
	q->r = U[i];
	q->u = U[nc + i] / U[i];
	q->v = U[2 * (size_t)nc + i] / U[i];
	q->E = U[3 * (size_t)nc + i] / U[i];
	q->e = q->E - 0.5 * (q->u * q->u + q->v * q->v);
	int mat = s->cell_mat[i];
	urs(s, mat, q, 0);
	urs(s, mat, q, 1);
}

static int cell_mat_of(const or_solver* s, int i) { return s->cell_mat[i]; }

/*
 * FVM_TVD::boundaryCond (src/methods/fvm_tvd.cpp:694-728) with the
 * CFDBoundary::run variants (src/bnd_cond.cpp:75-110): inlet sets u,v,T,p;
 * outlet copies; both wall types mirror the normal velocity (the factory
 * builds the slip class for BOUND_WALL_NO_SLIP, bnd_cond.cpp:50-51); then
 * URS(2) re-derives rho from p and T, URS(1) e and T, and E.
 * BOUND_PRESSURE_OUTLET is an extension: rho,u,v extrapolated, p set to the
 * given value while the normal velocity is subsonic.
 */
static void ghost(const or_solver* s, int e, const prm* pL, prm* pR)
{
	const or_mesh* m = s->m;
	int b = s->edge_bc[e];
	int mat = cell_mat_of(s, m->c1[e]);
	int type = s->c.bc_type[b];
	const double* par = s->c.bc_par[b];
	if (type == OR_BND_PRESSURE_OUTLET) {
		*pR = *pL;
		double un = pL->u * m->nx[e] + pL->v * m->ny[e];
		if (fabs(un) < pL->cz) pR->p = par[0];
		urs(s, mat, pR, 1);
		pR->E = pR->e + 0.5 * (pR->u * pR->u + pR->v * pR->v);
		return;
	}
	if (type == OR_BND_INLET) {
		pR->u = par[0];
		pR->v = par[1];
		pR->T = par[2];
		pR->p = par[3];
	} else if (type == OR_BND_OUTLET) {
		*pR = *pL;
	} else {
		*pR = *pL;
		double Un = pL->u * m->nx[e] + pL->v * m->ny[e];
		double Vx = m->nx[e] * Un * 2.0;
		double Vy = m->ny[e] * Un * 2.0;
		pR->u = pL->u - Vx;
		pR->v = pL->v - Vy;
	}
	urs(s, mat, pR, 2);
	urs(s, mat, pR, 1);
	pR->E = pR->e + 0.5 * (pR->u * pR->u + pR->v * pR->v);
}

/*
 * rim_orig, src/global.cpp:232-405: exact Riemann solver (Newton on p* from
 * the acoustic guess, P >= 1e-5 clamp, stop when |DU-F1-F2| <= 1e-5 after the
 * update, then the wave pattern sampled at x/t = 0).  Powers use
 * exp(log(x)*k) as the reference does.  max_iter > 0 caps the Newton loop
 * (the reference has no cap); *hit is set when the cap ends it.
 */
static void rim(double* RI, double* EI, double* PI, double* UI, double* VI, double* WI,
                double RB, double PB, double UB, double VB, double WB,
                double RE, double PE, double UE, double VE, double WE, double GAM,
                int max_iter, int* iters, int* hit)
{
	double AGAM = (GAM - 1.0);
	double DGAM = (2.0 / AGAM);
	double GGAM = (sqrt(GAM * AGAM));
	double HGAM = (AGAM / 2.0);
	double FGAM = (3.0 * GAM - 1.0);
	double OGAM = (AGAM / (2.0 * GAM));
	double QGAM = (GAM + 1.0);
	double PGAM = (QGAM / (2.0 * GAM));
	double RGAM = (4.0 * GAM);
	double SGAM = (GAM * AGAM);
	double TGAM = (QGAM / 2.0);

	double US = 0.0, UF = 0.0;
	double RF, RS, EF, ES, SBL, SFL, SSL, SEL, D;
	double P, F1, FS1, F2, FS2;
	const double eps = 1.0e-5;
	double CB = sqrt(GAM * PB / RB);
	double CE = sqrt(GAM * PE / RE);
	double EB = CB * CB / SGAM;
	double EE = CE * CE / SGAM;
	double RCB = RB * CB;
	double RCE = RE * CE;
	double DU = UB - UE;
	int it = 0;
	*hit = 0;
	if (DU < -2.0 * (CB + CE) / AGAM) {
		RF = 0.0; RS = 0.0; EF = 0.0; ES = 0.0;
		SBL = UB - CB;
		SFL = UB + 2.0 * CB / AGAM;
		SSL = UE - 2.0 * CE / AGAM;
		SEL = UE + CE;
	} else {
		P = (PB * RCE + PE * RCB + DU * RCB * RCE) / (RCB + RCE);
		for (;;) {
			if (P < eps) P = eps;
			double PPB = P / PB;
			if (PB > P) {
				double ZFB = CB * exp(log(PPB) * OGAM);
				F1 = DGAM * (ZFB - CB);
				FS1 = ZFB / (GAM * P);
			} else {
				double PKB = PGAM * PPB + OGAM;
				double ZNB = RCB * sqrt(PKB);
				F1 = (P - PB) / ZNB;
				FS1 = (QGAM * PPB + FGAM) / (RGAM * ZNB * PKB);
			}
			double PPE = P / PE;
			if (PE > P) {
				double ZFE = CE * exp(log(PPE) * OGAM);
				F2 = DGAM * (ZFE - CE);
				FS2 = ZFE / (GAM * P);
			} else {
				double PKE = PGAM * PPE + OGAM;
				double ZNE = RCE * sqrt(PKE);
				F2 = (P - PE) / ZNE;
				FS2 = (QGAM * PPE + FGAM) / (RGAM * ZNE * PKE);
			}
			double DP = (DU - F1 - F2) / (FS1 + FS2);
			P = P + DP;
			it++;
			if (!(fabs(DU - F1 - F2) > eps)) break;
			if (max_iter > 0 && it >= max_iter) { *hit = 1; break; }
		}
		double PPB = P / PB;
		double PPE = P / PE;
		double ZFB = CB * exp(log(PPB) * OGAM);
		double ZFE = CE * exp(log(PPE) * OGAM);
		if (PB > P) {
			EF = ZFB * ZFB / SGAM;
			UF = UB + DGAM * (CB - ZFB);
			RF = P / (AGAM * EF);
			SBL = UB - CB;
			SFL = UF - ZFB;
		} else {
			D = UB - sqrt((TGAM * P + HGAM * PB) / RB);
			double UBD = UB - D;
			double RUBD = RB * UBD;
			RF = RUBD * RUBD / (PB - P + RUBD * UBD);
			UF = D + RUBD / RF;
			EF = P / (AGAM * RF);
			SBL = D;
			SFL = D;
		}
		if (PE > P) {
			ES = ZFE * ZFE / SGAM;
			US = UE - DGAM * (CE - ZFE);
			RS = P / (AGAM * ES);
			SSL = US + ZFE;
			SEL = UE + CE;
		} else {
			D = UE + sqrt((TGAM * P + HGAM * PE) / RE);
			double UED = UE - D;
			double RUED = RE * UED;
			RS = RUED * RUED / (PE - P + RUED * UED);
			US = D + RUED / RS;
			ES = P / (AGAM * RS);
			SEL = D;
			SSL = D;
		}
	}
	if (iters) *iters = it;
	if (SEL <= 0.0) {
		*RI = RE; *EI = EE; *UI = UE; *VI = VE; *WI = WE;
	} else if (SBL >= 0.0) {
		*RI = RB; *EI = EB; *UI = UB; *VI = VB; *WI = WB;
	} else if ((SSL >= 0.0) && (SFL <= 0.0)) {
		if (US >= 0.0) {
			*RI = RF; *EI = EF; *UI = UF; *VI = VB; *WI = WB;
		} else {
			*RI = RS; *EI = ES; *UI = US; *VI = VE; *WI = WE;
		}
	} else if (SFL > 0.0) {
		*UI = (UB + DGAM * GGAM * sqrt(EB)) / (1 + DGAM);
		*VI = VB;
		*WI = WB;
		*EI = (*UI * *UI) / SGAM;
		*RI = RB * exp(log(*EI / EB) * (1 / AGAM));
	} else {
		*UI = (UE - DGAM * GGAM * sqrt(EE)) / (1 + DGAM);
		*VI = VE;
		*WI = WE;
		*EI = (*UI * *UI) / SGAM;
		*RI = RE * exp(log(*EI / EE) * (1 / AGAM));
	}
	*PI = AGAM * *EI * *RI;
}

/* FVM_TVD::calcFlux, Godunov branch (src/methods/fvm_tvd.cpp:602-622) */
static void flux_godunov(double* f, const prm* pL, const prm* pR, double nx, double ny, double GAM,
                         int max_iter, int* hit)
{
	double RI, EI, PI, UN, UT, WI;
	double unl = pL->u * nx + pL->v * ny;
	double unr = pR->u * nx + pR->v * ny;
	double utl = pL->u * ny - pL->v * nx;
	double utr = pR->u * ny - pR->v * nx;
	rim(&RI, &EI, &PI, &UN, &UT, &WI, pL->r, pL->p, unl, utl, 0, pR->r, pR->p, unr, utr, 0, GAM,
	    max_iter, NULL, hit);
	double UI = UN * nx + UT * ny;
	double VI = UN * ny - UT * nx;
	f[0] = RI * UN;
	f[1] = f[0] * UI + PI * nx;
	f[2] = f[0] * VI + PI * ny;
	f[3] = (RI * (EI + 0.5 * (UI * UI + VI * VI)) + PI) * UN;
}

/* FLUX=LAX as FVM_TVD_IMPLICIT::calcFlux defines it (src/methods/fvm_tvd_implicit.cpp:939-960);
 * pinned by tests/test_oracle_pins.py::test_lax_* (closed forms) */
static void flux_lax(double* f, const prm* pL, const prm* pR, double nx, double ny, double GAM)
{
	double unl = pL->u * nx + pL->v * ny;
	double unr = pR->u * nx + pR->v * ny;
	double al = fabs(unl) + sqrt(GAM * pL->p / pL->r);
	double ar = fabs(unr) + sqrt(GAM * pR->p / pR->r);
	double alpha = al > ar ? al : ar;
	double rol = pL->r, rul = pL->r * pL->u, rvl = pL->r * pL->v;
	double rel = pL->p / (GAM - 1.0) + 0.5 * pL->r * (pL->u * pL->u + pL->v * pL->v);
	double ror = pR->r, rur = pR->r * pR->u, rvr = pR->r * pR->v;
	double rer = pR->p / (GAM - 1.0) + 0.5 * pR->r * (pR->u * pR->u + pR->v * pR->v);
	double frl = rol * unl;
	double frr = ror * unr;
	f[0] = 0.5 * (frr + frl - alpha * (ror - rol));
	f[1] = 0.5 * (frr * pR->u + frl * pL->u + (pR->p + pL->p) * nx - alpha * (rur - rul));
	f[2] = 0.5 * (frr * pR->v + frl * pL->v + (pR->p + pL->p) * ny - alpha * (rvr - rvl));
	f[3] = 0.5 * ((rer + pR->p) * unr + (rel + pL->p) * unl - alpha * (rer - rel));
}

/*
 * Extension: HLLC (Toro, Riemann Solvers, 3rd ed., sec. 10.4-10.6) with
 * Davis wave-speed bounds S_L = min(un_L - c_L, un_R - c_R),
 * S_R = max(un_L + c_L, un_R + c_R), and the star states of Toro (10.39).
 */
static void flux_hllc(double* f, const prm* pL, const prm* pR, double nx, double ny, double GAM)
{
	double unl = pL->u * nx + pL->v * ny, unr = pR->u * nx + pR->v * ny;
	double cl = sqrt(GAM * pL->p / pL->r), cr = sqrt(GAM * pR->p / pR->r);
	double El = pL->p / (GAM - 1.0) + 0.5 * pL->r * (pL->u * pL->u + pL->v * pL->v);
	double Er = pR->p / (GAM - 1.0) + 0.5 * pR->r * (pR->u * pR->u + pR->v * pR->v);
	double sl = fmin(unl - cl, unr - cr);
	double sr = fmax(unl + cl, unr + cr);
	double fl[4] = {pL->r * unl, pL->r * pL->u * unl + pL->p * nx, pL->r * pL->v * unl + pL->p * ny,
	                (El + pL->p) * unl};
	double fr[4] = {pR->r * unr, pR->r * pR->u * unr + pR->p * nx, pR->r * pR->v * unr + pR->p * ny,
	                (Er + pR->p) * unr};
	if (sl >= 0.0) { memcpy(f, fl, sizeof(fl)); return; }
	if (sr <= 0.0) { memcpy(f, fr, sizeof(fr)); return; }
	double ml = pL->r * (sl - unl), mr = pR->r * (sr - unr);
	double ss = (pR->p - pL->p + pL->r * unl * (sl - unl) - pR->r * unr * (sr - unr)) / (ml - mr);
	if (ss >= 0.0) {
		double k = ml / (sl - ss);
		double us[4] = {k, k * (pL->u + (ss - unl) * nx), k * (pL->v + (ss - unl) * ny),
		                k * (El / pL->r + (ss - unl) * (ss + pL->p / (pL->r * (sl - unl))))};
		double ul[4] = {pL->r, pL->r * pL->u, pL->r * pL->v, El};
		for (int q = 0; q < 4; q++) f[q] = fl[q] + sl * (us[q] - ul[q]);
	} else {
		double k = mr / (sr - ss);
		double us[4] = {k, k * (pR->u + (ss - unr) * nx), k * (pR->v + (ss - unr) * ny),
		                k * (Er / pR->r + (ss - unr) * (ss + pR->p / (pR->r * (sr - unr))))};
		double ur[4] = {pR->r, pR->r * pR->u, pR->r * pR->v, Er};
		for (int q = 0; q < 4; q++) f[q] = fr[q] + sr * (us[q] - ur[q]);
	}
}

/*
 * Extension: Roe flux (Roe 1981; Toro sec. 11.2) on Roe averages as
 * roe_orig forms them (src/global.cpp:407-441), with Harten's entropy fix
 * |lambda| -> (lambda^2 + delta^2) / (2 delta) for |lambda| < delta,
 * delta = 0.1 * c_roe.
 */
static void flux_roe(double* f, const prm* pL, const prm* pR, double nx, double ny, double GAM)
{
	double unl = pL->u * nx + pL->v * ny, unr = pR->u * nx + pR->v * ny;
	double El = pL->p / (GAM - 1.0) + 0.5 * pL->r * (pL->u * pL->u + pL->v * pL->v);
	double Er = pR->p / (GAM - 1.0) + 0.5 * pR->r * (pR->u * pR->u + pR->v * pR->v);
	double Hl = (El + pL->p) / pL->r, Hr = (Er + pR->p) / pR->r;
	double sl = sqrt(pL->r), sr = sqrt(pR->r), w = 1.0 / (sl + sr);
	double u = (sl * pL->u + sr * pR->u) * w, v = (sl * pL->v + sr * pR->v) * w;
	double H = (sl * Hl + sr * Hr) * w;
	double q2 = u * u + v * v;
	double c = sqrt((GAM - 1.0) * (H - 0.5 * q2));
	double un = u * nx + v * ny, ut = -u * ny + v * nx;
	double dr = pR->r - pL->r, dp = pR->p - pL->p;
	double dun = unr - unl;
	double dut = (-pR->u * ny + pR->v * nx) - (-pL->u * ny + pL->v * nx);
	double rr = sl * sr;
	double a1 = (dp - rr * c * dun) / (2.0 * c * c);
	double a2 = dr - dp / (c * c);
	double a3 = rr * dut;
	double a4 = (dp + rr * c * dun) / (2.0 * c * c);
	double l1 = fabs(un - c), l2 = fabs(un), l4 = fabs(un + c);
	double dlt = 0.1 * c;
	if (l1 < dlt) l1 = (l1 * l1 + dlt * dlt) / (2.0 * dlt);
	if (l4 < dlt) l4 = (l4 * l4 + dlt * dlt) / (2.0 * dlt);
	double tx = -ny, ty = nx;
	double d[4];
	d[0] = l1 * a1 + l2 * a2 + l4 * a4;
	d[1] = l1 * a1 * (u - c * nx) + l2 * a2 * u + l2 * a3 * tx + l4 * a4 * (u + c * nx);
	d[2] = l1 * a1 * (v - c * ny) + l2 * a2 * v + l2 * a3 * ty + l4 * a4 * (v + c * ny);
	d[3] = l1 * a1 * (H - un * c) + l2 * a2 * 0.5 * q2 + l2 * a3 * ut + l4 * a4 * (H + un * c);
	double fl[4] = {pL->r * unl, pL->r * pL->u * unl + pL->p * nx, pL->r * pL->v * unl + pL->p * ny,
	                (El + pL->p) * unl};
	double fr[4] = {pR->r * unr, pR->r * pR->u * unr + pR->p * nx, pR->r * pR->v * unr + pR->p * ny,
	                (Er + pR->p) * unr};
	for (int q = 0; q < 4; q++) f[q] = 0.5 * (fl[q] + fr[q] - d[q]);
}

static void num_flux(const or_solver* s, double* f, const prm* pL, const prm* pR, double nx, double ny,
                     int* hit)
{
	double GAM = s->c.gam_flux;
	*hit = 0;
	switch (s->c.flux) {
	case OR_FLUX_LAX: flux_lax(f, pL, pR, nx, ny, GAM); break;
	case OR_FLUX_HLLC: flux_hllc(f, pL, pR, nx, ny, GAM); break;
	case OR_FLUX_ROE: flux_roe(f, pL, pR, nx, ny, GAM); break;
	default: flux_godunov(f, pL, pR, nx, ny, GAM, s->c.max_newton, hit); break;
	}
}

/* ------------------------------------------------------------ passes */

static void prim_pass(or_solver* s, const double* U)
{
	int nc = s->m->nc;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) cons2prim(s, U, nc, i, &s->P[i]);
}

/* the neighbour of cell i across edge e: cell index, or -1 with a ghost */
static int across(const or_solver* s, int e, int i, const prm* pi, prm* pj, double* dx, double* dy)
{
	const or_mesh* m = s->m;
	int c1 = m->c1[e], c2 = m->c2[e];
	if (c2 < 0) {
		ghost(s, e, pi, pj);
		/* ghost centroid: mirror image of the cell centroid in the edge line */
		double mx = m->gp[6 * (size_t)e], my = m->gp[6 * (size_t)e + 1];
		double dn = (mx - m->cx[i]) * m->nx[e] + (my - m->cy[i]) * m->ny[e];
		*dx = 2.0 * dn * m->nx[e];
		*dy = 2.0 * dn * m->ny[e];
		return -1;
	}
	if (i == c1) {
		*pj = s->P[c2];
		*dx = (m->cx[c2] + m->shx[e]) - m->cx[c1];
		*dy = (m->cy[c2] + m->shy[e]) - m->cy[c1];
		return c2;
	}
	*pj = s->P[c1];
	*dx = m->cx[c1] - (m->cx[c2] + m->shx[e]);
	*dy = m->cy[c1] - (m->cy[c2] + m->shy[e]);
	return c1;
}

/* Gauss-point offsets of cell i on edge e (reconstruct, fvm_tvd.cpp:661-664, 681-683) */
static void gp_offset(const or_mesh* m, int e, int i, int g, double* dx, double* dy)
{
	const double* P = &m->gp[6 * (size_t)e + 2 + 2 * g];
	if (i == m->c1[e]) {
		*dx = P[0] - m->cx[i];
		*dy = P[1] - m->cy[i];
	} else {
		*dx = P[0] - (m->cx[i] + m->shx[e]);
		*dy = P[1] - (m->cy[i] + m->shy[e]);
	}
}

/*
 * Green-Gauss gradients of (r, p, u, v), FVM_TVD::calcGrad
 * (src/methods/fvm_tvd.cpp:242-301): per edge the face value is the mean of
 * the two cell-centre primitives (ghost from boundaryCond on a boundary),
 * times n.x*l / n.y*l, added to c1 and subtracted from c2 in edge order,
 * then divided by S.
 */
static void grad_gg(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			int c1 = m->c1[e], c2 = m->c2[e];
			prm pL = s->P[c1], pR;
			if (c2 > -1) pR = s->P[c2]; else ghost(s, e, &pL, &pR);
			double nx = m->nx[e], ny = m->ny[e], l = m->l[e];
			double t[8];
			t[0] = (pL.r + pR.r) / 2 * nx * l;
			t[1] = (pL.r + pR.r) / 2 * ny * l;
			t[2] = (pL.p + pR.p) / 2 * nx * l;
			t[3] = (pL.p + pR.p) / 2 * ny * l;
			t[4] = (pL.u + pR.u) / 2 * nx * l;
			t[5] = (pL.u + pR.u) / 2 * ny * l;
			t[6] = (pL.v + pR.v) / 2 * nx * l;
			t[7] = (pL.v + pR.v) / 2 * ny * l;
			if (i == c1)
				for (int q = 0; q < 8; q++) g[q] += t[q];
			else
				for (int q = 0; q < 8; q++) g[q] -= t[q];
		}
		double si = m->S[i];
		for (int q = 0; q < 8; q++) s->G[(size_t)q * nc + i] = g[q] / si;
	}
}

/*
 * Extension: unweighted least-squares gradients.  For cell i with neighbour
 * offsets d_k (ghost at the mirror point on boundaries, periodic shifts
 * applied), A = sum d_k d_k^T and w_k = A^{-1} d_k; grad = sum_k w_k (phi_k -
 * phi_i) in ascending edge order.  Exact for linear fields
 * (tests/test_oracle_ext.py::test_lsq_exact_linear).
 */
static void lsq_weights(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	s->lsq_w = (double*)malloc(sizeof(double) * 6 * (size_t)nc);
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		double d[3][2];
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			int c1 = m->c1[e], c2 = m->c2[e];
			if (c2 < 0) {
				double mx = m->gp[6 * (size_t)e], my = m->gp[6 * (size_t)e + 1];
				double dn = (mx - m->cx[i]) * m->nx[e] + (my - m->cy[i]) * m->ny[e];
				d[k][0] = 2.0 * dn * m->nx[e];
				d[k][1] = 2.0 * dn * m->ny[e];
			} else if (i == c1) {
				d[k][0] = (m->cx[c2] + m->shx[e]) - m->cx[c1];
				d[k][1] = (m->cy[c2] + m->shy[e]) - m->cy[c1];
			} else {
				d[k][0] = m->cx[c1] - (m->cx[c2] + m->shx[e]);
				d[k][1] = m->cy[c1] - (m->cy[c2] + m->shy[e]);
			}
		}
		double a11 = 0, a12 = 0, a22 = 0;
		for (int k = 0; k < 3; k++) {
			a11 += d[k][0] * d[k][0];
			a12 += d[k][0] * d[k][1];
			a22 += d[k][1] * d[k][1];
		}
		double det = a11 * a22 - a12 * a12;
		for (int k = 0; k < 3; k++) {
			s->lsq_w[6 * (size_t)i + 2 * k] = (a22 * d[k][0] - a12 * d[k][1]) / det;
			s->lsq_w[6 * (size_t)i + 2 * k + 1] = (a11 * d[k][1] - a12 * d[k][0]) / det;
		}
	}
}

static void grad_lsq(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	if (!s->lsq_w) lsq_weights(s);
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		prm pi = s->P[i];
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			prm pj;
			double dx, dy;
			across(s, e, i, &pi, &pj, &dx, &dy);
			double wx = s->lsq_w[6 * (size_t)i + 2 * k], wy = s->lsq_w[6 * (size_t)i + 2 * k + 1];
			double dphi[4] = {pj.r - pi.r, pj.p - pi.p, pj.u - pi.u, pj.v - pi.v};
			for (int q = 0; q < 4; q++) {
				g[2 * q] += wx * dphi[q];
				g[2 * q + 1] += wy * dphi[q];
			}
		}
		for (int q = 0; q < 8; q++) s->G[(size_t)q * nc + i] = g[q];
	}
}

/*
 * Extension: minmod limiting in Barth-Jespersen form.  For each of
 * (r, p, u, v): phi_min/phi_max over the cell and its face neighbours
 * (ghosts on boundaries); at each of the cell's six Gauss points,
 * D = grad . d and psi_k = minmod(1, (phi_max - phi)/D) for D > 0,
 * minmod(1, (phi_min - phi)/D) for D < 0, 1 for D = 0; the gradient is scaled
 * by min_k psi_k, the largest scaling in [0, 1] that keeps every
 * reconstructed value inside the neighbour bounds
 * (tests/test_oracle_pins.py::test_minmod_is_the_largest_bounded_scaling,
 * tests/test_oracle_ext.py::test_minmod_bounds).
 */
static void limit_minmod(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		prm pi = s->P[i];
		double phi[4] = {pi.r, pi.p, pi.u, pi.v};
		double lo[4], hi[4];
		for (int q = 0; q < 4; q++) lo[q] = hi[q] = phi[q];
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			prm pj;
			double dx, dy;
			across(s, e, i, &pi, &pj, &dx, &dy);
			double nb[4] = {pj.r, pj.p, pj.u, pj.v};
			for (int q = 0; q < 4; q++) {
				if (nb[q] < lo[q]) lo[q] = nb[q];
				if (nb[q] > hi[q]) hi[q] = nb[q];
			}
		}
		for (int q = 0; q < 4; q++) {
			double gx = s->G[(size_t)(2 * q) * nc + i], gy = s->G[(size_t)(2 * q + 1) * nc + i];
			double psi = 1.0;
			for (int k = 0; k < 3; k++) {
				int e = m->cedges[3 * (size_t)i + k];
				for (int g = 0; g < 2; g++) {
					double dx, dy;
					gp_offset(m, e, i, g, &dx, &dy);
					double D = gx * dx + gy * dy;
					double r = 1.0;
					if (D > 0.0) r = (hi[q] - phi[q]) / D;
					else if (D < 0.0) r = (lo[q] - phi[q]) / D;
					if (r < psi) psi = r;
				}
			}
			if (psi < 0.0) psi = 0.0;
			s->G[(size_t)(2 * q) * nc + i] = psi * gx;
			s->G[(size_t)(2 * q + 1) * nc + i] = psi * gy;
		}
	}
}

static void gradients(or_solver* s)
{
	if (s->c.gradient == OR_GRAD_LSQ) grad_lsq(s); else grad_gg(s);
	if (s->c.limiter == OR_LIM_MINMOD) limit_minmod(s);
}

/* temperature gradient from those of p and rho: T = p M / (rho gR) */
static void grad_T(const or_solver* s, int i, double* tx, double* ty)
{
	int nc = s->m->nc;
	const prm* q = &s->P[i];
	double Rg = GAS_R / s->c.M[s->cell_mat[i]];
	double pr = q->p / q->r;
	double gx_r = s->G[i], gy_r = s->G[(size_t)nc + i];
	double gx_p = s->G[2 * (size_t)nc + i], gy_p = s->G[3 * (size_t)nc + i];
	*tx = (gx_p - pr * gx_r) / (q->r * Rg);
	*ty = (gy_p - pr * gy_r) / (q->r * Rg);
}

/*
 * Extension: Navier-Stokes viscous flux at a Gauss point.  Face gradients of
 * u, v and T are the means of the two cell gradients (the cell's own on a
 * boundary); the face velocity is the mean of the reconstructed left/right
 * velocities; mu = material ML, conductivity = material K, or mu*Cp/Pr when K
 * is zero.  Returns the viscous part of the normal flux, which is subtracted
 * from the inviscid flux.  Pinned by tests/test_oracle_ext.py::test_viscous_*
 * (stress) and tests/test_oracle_pins.py::test_heat_conduction_* (conduction).
 */
static void viscous_flux(const or_solver* s, int e, const prm* pL, const prm* pR, double* fv)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	int c1 = m->c1[e], c2 = m->c2[e];
	int mat = s->cell_mat[c1];
	double mu = s->c.ML[mat];
	double k = s->c.K[mat] > 0.0 ? s->c.K[mat] : mu * s->c.Cp[mat] / s->c.prandtl;
	double ux = s->G[4 * (size_t)nc + c1], uy = s->G[5 * (size_t)nc + c1];
	double vx = s->G[6 * (size_t)nc + c1], vy = s->G[7 * (size_t)nc + c1];
	double tx, ty;
	grad_T(s, c1, &tx, &ty);
	if (c2 >= 0) {
		double tx2, ty2;
		grad_T(s, c2, &tx2, &ty2);
		ux = 0.5 * (ux + s->G[4 * (size_t)nc + c2]);
		uy = 0.5 * (uy + s->G[5 * (size_t)nc + c2]);
		vx = 0.5 * (vx + s->G[6 * (size_t)nc + c2]);
		vy = 0.5 * (vy + s->G[7 * (size_t)nc + c2]);
		tx = 0.5 * (tx + tx2);
		ty = 0.5 * (ty + ty2);
	}
	double nx = m->nx[e], ny = m->ny[e];
	double div = ux + vy;
	double txx = mu * (2.0 * ux - (2.0 / 3.0) * div);
	double tyy = mu * (2.0 * vy - (2.0 / 3.0) * div);
	double txy = mu * (uy + vx);
	double uf = 0.5 * (pL->u + pR->u), vf = 0.5 * (pL->v + pR->v);
	double sx = txx * nx + txy * ny, sy = txy * nx + tyy * ny;
	fv[0] = 0.0;
	fv[1] = sx;
	fv[2] = sy;
	fv[3] = uf * sx + vf * sy + k * (tx * nx + ty * ny);
}

/*
 * Edge fluxes, FVM_TVD::run edge loop (src/methods/fvm_tvd.cpp:329-352) with
 * FVM_TVD::reconstruct (646-691): at each Gauss point the left state is the
 * c1 cell-centre primitive plus grad . (P - x_c1) for r, p, u, v (e, E, cz, T
 * stay cell-centre values); the right state likewise from c2, or the ghost of
 * the reconstructed left state on a boundary; the Gauss-point fluxes are
 * summed and multiplied by l/2.  Output: F[e*4+q].
 */
static void edge_pass(or_solver* s)
{
	const or_mesh* m = s->m;
	int ne = m->ne, nc = m->nc;
	const double* G = s->G;
	long hits = 0;
#pragma omp parallel for schedule(static) reduction(+ : hits)
	for (int e = 0; e < ne; e++) {
		int c1 = m->c1[e], c2 = m->c2[e];
		double nx = m->nx[e], ny = m->ny[e];
		double acc[4] = {0.0, 0.0, 0.0, 0.0};
		for (int g = 0; g < 2; g++) {
			prm pL = s->P[c1], pR;
			double dx, dy;
			gp_offset(m, e, c1, g, &dx, &dy);
			pL.r += G[c1] * dx + G[(size_t)nc + c1] * dy;
			pL.p += G[2 * (size_t)nc + c1] * dx + G[3 * (size_t)nc + c1] * dy;
			pL.u += G[4 * (size_t)nc + c1] * dx + G[5 * (size_t)nc + c1] * dy;
			pL.v += G[6 * (size_t)nc + c1] * dx + G[7 * (size_t)nc + c1] * dy;
			if (c2 > -1) {
				pR = s->P[c2];
				gp_offset(m, e, c2, g, &dx, &dy);
				pR.r += G[c2] * dx + G[(size_t)nc + c2] * dy;
				pR.p += G[2 * (size_t)nc + c2] * dx + G[3 * (size_t)nc + c2] * dy;
				pR.u += G[4 * (size_t)nc + c2] * dx + G[5 * (size_t)nc + c2] * dy;
				pR.v += G[6 * (size_t)nc + c2] * dx + G[7 * (size_t)nc + c2] * dy;
			} else {
				ghost(s, e, &pL, &pR);
			}
			double f[4];
			int hit;
			num_flux(s, f, &pL, &pR, nx, ny, &hit);
			hits += hit;
			if (s->c.viscous) {
				double fv[4];
				viscous_flux(s, e, &pL, &pR, fv);
				for (int q = 0; q < 4; q++) f[q] -= fv[q];
			}
			for (int q = 0; q < 4; q++) acc[q] += f[q];
		}
		double l = m->l[e] * 0.5;
		for (int q = 0; q < 4; q++) s->F[4 * (size_t)e + q] = acc[q] * l;
	}
	s->cap_hits += hits;
}

/* residual accumulation (fvm_tvd.cpp:353-363): -F into c1, +F into c2, edge order */
static void residual_pass(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		double r[4] = {0.0, 0.0, 0.0, 0.0};
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			const double* f = &s->F[4 * (size_t)e];
			if (m->c1[e] == i)
				for (int q = 0; q < 4; q++) r[q] -= f[q];
			else
				for (int q = 0; q < 4; q++) r[q] += f[q];
		}
		for (int q = 0; q < 4; q++) s->R[(size_t)q * nc + i] = r[q];
	}
}

static void stage_residual(or_solver* s, const double* U)
{
	prim_pass(s, U);
	gradients(s);
	edge_pass(s);
	residual_pass(s);
}

static inline int is_lim(const or_solver* s, int i) { return (s->flags[i] & 0x2u) > 0; }

/*
 * FVM_TVD::calcTimeStep (src/methods/fvm_tvd.cpp:216-240): CFL*S/max(|u|+c,
 * |v|+c) with the cell area S; unsteady: TAU = min(TAU, that) over cells,
 * every cell gets TAU; steady: local step per cell.
 * Extension OR_CFL_SPECTRAL: 2*CFL*S / sum_faces(lambda_f l_f) with
 * lambda_f the mean of (|V|+c) over the two sides (ghost on a boundary),
 * after the spectral-radius step of fvm_tvd_implicit.cpp:556-563, 644.
 */
void or_calc_tau(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	prim_pass(s, s->U);
	for (int i = 0; i < nc; i++) {
		prm p = s->P[i];
		double loc;
		if (s->c.cfl_mode == OR_CFL_SPECTRAL) {
			double sum = 0.0;
			double li = sqrt(p.u * p.u + p.v * p.v) + p.cz;
			for (int k = 0; k < 3; k++) {
				int e = m->cedges[3 * (size_t)i + k];
				prm pj;
				double dx, dy;
				across(s, e, i, &p, &pj, &dx, &dy);
				double lj = sqrt(pj.u * pj.u + pj.v * pj.v) + pj.cz;
				sum += 0.5 * (li + lj) * m->l[e];
			}
			loc = 2.0 * s->c.cfl * m->S[i] / sum;
		} else {
			double a = fabs(p.u) + p.cz, b = fabs(p.v) + p.cz;
			loc = s->c.cfl * m->S[i] / (a > b ? a : b);
		}
		if (s->c.steady) {
			s->ctau[i] = loc;
		} else if (s->TAU > loc) {
			s->TAU = loc;
		}
	}
	if (!s->c.steady)
		for (int i = 0; i < nc; i++) s->ctau[i] = s->TAU;
}

static void euler_update(or_solver* s, double* U)
{
	int nc = s->m->nc;
#pragma omp parallel for schedule(static)
	for (int i = 0; i < nc; i++) {
		if (is_lim(s, i)) continue;
		double cfl = s->ctau[i] / s->m->S[i];
		for (int q = 0; q < 4; q++) U[(size_t)q * nc + i] += cfl * s->R[(size_t)q * nc + i];
	}
}

/* bound flags (fvm_tvd.cpp:439-447); the clamped values are not written back */
static void bound_flags(or_solver* s, int i)
{
	prm par;
	cons2prim(s, s->U, s->m->nc, i, &par);
	int f = 0;
	if (par.r < s->c.ro_min) f = 1;
	if (par.r > s->c.ro_max) f = 1;
	if (par.p < s->c.p_min) f = 1;
	if (par.p > s->c.p_max) f = 1;
	if (fabs(par.u) > s->c.u_max) f = 1;
	if (fabs(par.v) > s->c.u_max) f = 1;
	if (f) s->flags[i] |= 0x2u;
}

/*
 * FVM_TVD::remediateLimCells (src/methods/fvm_tvd.cpp:464-499): in cell
 * order, a flagged cell takes the area-weighted mean of the c2 cells of its
 * edges (which may be the cell itself, line 479), in its edge-list order,
 * using values already remediated earlier in the sweep; then its counter
 * advances and after 32 steps the flag is cleared.
 */
static void remediate(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	double* U = s->U;
	for (int i = 0; i < nc; i++) {
		if (!is_lim(s, i)) continue;
		double sr[4] = {0.0, 0.0, 0.0, 0.0}, S = 0.0;
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			int j = m->c2[e];
			if (j >= 0) {
				double sj = m->S[j];
				S += sj;
				for (int q = 0; q < 4; q++) sr[q] += U[(size_t)q * nc + j] * sj;
			}
		}
		for (int q = 0; q < 4; q++) U[(size_t)q * nc + i] = sr[q] / S;
		s->flags[i] += 0x010000u;
		if (s->flags[i] & 0x200000u) s->flags[i] &= 0x001110u;
	}
}

/*
 * One time step, FVM_TVD::run body (src/methods/fvm_tvd.cpp:312-449):
 * Heun RK2 (two forward-Euler stages then U = (U_old + U)/2) when
 * rk_stages == 2.  Extension rk_stages == 3: SSP-RK3 of Shu & Osher,
 * U1 = U0 + dt L(U0); U2 = 3/4 U0 + 1/4 (U1 + dt L(U1));
 * U3 = 1/3 U0 + 2/3 (U2 + dt L(U2)).  LIM-flagged cells stay frozen in every
 * stage, as in the reference.
 */
static void step(or_solver* s)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	size_t n4 = 4 * (size_t)nc;
	if (s->c.steady) or_calc_tau(s);
	memcpy(s->U0, s->U, n4 * sizeof(double));
	if (s->c.rk_stages == 3) {
		stage_residual(s, s->U);
		euler_update(s, s->U);
		memcpy(s->U1, s->U, n4 * sizeof(double));
		stage_residual(s, s->U1);
		euler_update(s, s->U1);
		for (int i = 0; i < nc; i++) {
			if (is_lim(s, i)) continue;
			for (int q = 0; q < 4; q++) {
				size_t k = (size_t)q * nc + i;
				s->U[k] = 0.75 * s->U0[k] + 0.25 * s->U1[k];
			}
		}
		stage_residual(s, s->U);
		euler_update(s, s->U);
		for (int i = 0; i < nc; i++) {
			if (is_lim(s, i)) continue;
			for (int q = 0; q < 4; q++) {
				size_t k = (size_t)q * nc + i;
				s->U[k] = (1.0 / 3.0) * s->U0[k] + (2.0 / 3.0) * s->U[k];
			}
			bound_flags(s, i);
		}
	} else {
		stage_residual(s, s->U);
		euler_update(s, s->U);
		stage_residual(s, s->U);
		euler_update(s, s->U);
		for (int i = 0; i < nc; i++) {
			if (is_lim(s, i)) continue;
			for (int q = 0; q < 4; q++) {
				size_t k = (size_t)q * nc + i;
				s->U[k] = 0.5 * (s->U0[k] + s->U[k]);
			}
			bound_flags(s, i);
		}
	}
	remediate(s);
}

int or_run(or_solver* s, int nsteps)
{
	for (int k = 0; k < nsteps; k++) step(s);
	return nsteps;
}

long or_newton_cap_hits(const or_solver* s) { return s->cap_hits; }

/*
 * Pressure force on the edges selected by edge_mask, after
 * FVM_TVD_IMPLICIT::calcLiftForce (src/methods/fvm_tvd_implicit.cpp:1088-1110):
 * sum p_c1 * l * n over the edges, with the c1 cell-centre pressure.
 * Returns the y component and stores the x component in *fx.
 */
double or_wall_force(const or_solver* s, const int* edge_mask, double* fx)
{
	const or_mesh* m = s->m;
	double Fx = 0.0, Fy = 0.0;
	for (int e = 0; e < m->ne; e++) {
		if (!edge_mask[e]) continue;
		prm p;
		cons2prim(s, s->U, m->nc, m->c1[e], &p);
		Fx += p.p * m->l[e] * m->nx[e];
		Fy += p.p * m->l[e] * m->ny[e];
	}
	if (fx) *fx = Fx;
	return Fy;
}

/* ------------------------------------------------------------ implicit */

/*
 * FVM_TVD_IMPLICIT::boundaryCond (src/methods/fvm_tvd_implicit.cpp:963-1008):
 * inlet takes T, p, u, v from the boundary and rho from URS(2), then URS(1);
 * outlet copies the inside state; the wall copies it and mirrors the normal
 * velocity (no URS afterwards, unlike the explicit method's ghost).
 */
static void ghost_impl(const or_solver* s, int e, const prm* pL, prm* pR)
{
	const or_mesh* m = s->m;
	int b = s->edge_bc[e];
	int mat = s->cell_mat[m->c1[e]];
	const double* par = s->c.bc_par[b];
	if (s->c.bc_type[b] == OR_BND_INLET) {
		memset(pR, 0, sizeof(*pR));
		pR->T = par[2];
		pR->p = par[3];
		pR->u = par[0];
		pR->v = par[1];
		urs(s, mat, pR, 2);
		urs(s, mat, pR, 1);
	} else if (s->c.bc_type[b] == OR_BND_OUTLET) {
		*pR = *pL;
	} else {
		*pR = *pL;
		double Un = pL->u * m->nx[e] + pL->v * m->ny[e];
		double Vx = m->nx[e] * Un * 2.0;
		double Vy = m->ny[e] * Un * 2.0;
		pR->u = pL->u - Vx;
		pR->v = pL->v - Vy;
	}
}

/* FVM_TVD_IMPLICIT::reconstruct(iFace, pL, pR), fvm_tvd_implicit.cpp:431-442 */
static void impl_centre(const or_solver* s, int e, prm* pL, prm* pR)
{
	const or_mesh* m = s->m;
	*pL = s->P[m->c1[e]];
	if (m->c2[e] >= 0) *pR = s->P[m->c2[e]];
	else ghost_impl(s, e, pL, pR);
}

/* FVM_TVD_IMPLICIT::calcGrad, fvm_tvd_implicit.cpp:186-262 (edge scatter);
 * g[8][nc] as (rx ry px py ux uy vx vy) */
static void impl_grad(const or_solver* s, double* g)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	memset(g, 0, 8 * (size_t)nc * sizeof(double));
	for (int e = 0; e < m->ne; e++) {
		int c1 = m->c1[e], c2 = m->c2[e];
		prm pL, pR;
		impl_centre(s, e, &pL, &pR);
		double nx = m->nx[e], ny = m->ny[e], l = m->l[e];
		double t[8];
		t[0] = (pL.r + pR.r) / 2 * nx * l;
		t[1] = (pL.r + pR.r) / 2 * ny * l;
		t[2] = (pL.p + pR.p) / 2 * nx * l;
		t[3] = (pL.p + pR.p) / 2 * ny * l;
		t[4] = (pL.u + pR.u) / 2 * nx * l;
		t[5] = (pL.u + pR.u) / 2 * ny * l;
		t[6] = (pL.v + pR.v) / 2 * nx * l;
		t[7] = (pL.v + pR.v) / 2 * ny * l;
		for (int q = 0; q < 8; q++) g[(size_t)q * nc + c1] += t[q];
		if (c2 > -1)
			for (int q = 0; q < 8; q++) g[(size_t)q * nc + c2] -= t[q];
	}
	for (int i = 0; i < nc; i++) {
		double si = m->S[i];
		for (int q = 0; q < 8; q++) g[(size_t)q * nc + i] /= si;
	}
}

/* FVM_TVD_IMPLICIT::reconstruct(iFace, pL, pR, p), fvm_tvd_implicit.cpp:444-476:
 * linear from both centroids on an inner face, centre state and ghost on a
 * boundary face */
static void impl_recon(const or_solver* s, const double* g, int e, int gp, prm* pL, prm* pR)
{
	const or_mesh* m = s->m;
	int nc = m->nc;
	int c1 = m->c1[e], c2 = m->c2[e];
	*pL = s->P[c1];
	if (c2 < 0) {
		ghost_impl(s, e, pL, pR);
		return;
	}
	*pR = s->P[c2];
	double dx, dy;
	gp_offset(m, e, c1, gp, &dx, &dy);
	pL->r += g[c1] * dx + g[(size_t)nc + c1] * dy;
	pL->p += g[2 * (size_t)nc + c1] * dx + g[3 * (size_t)nc + c1] * dy;
	pL->u += g[4 * (size_t)nc + c1] * dx + g[5 * (size_t)nc + c1] * dy;
	pL->v += g[6 * (size_t)nc + c1] * dx + g[7 * (size_t)nc + c1] * dy;
	gp_offset(m, e, c2, gp, &dx, &dy);
	pR->r += g[c2] * dx + g[(size_t)nc + c2] * dy;
	pR->p += g[2 * (size_t)nc + c2] * dx + g[3 * (size_t)nc + c2] * dy;
	pR->u += g[4 * (size_t)nc + c2] * dx + g[5 * (size_t)nc + c2] * dy;
	pR->v += g[6 * (size_t)nc + c2] * dx + g[7 * (size_t)nc + c2] * dy;
}

/* roe_orig, src/global.cpp:407-441 */
static void roe_orig(double* RI, double* EI, double* PI, double* UI, double* VI, double* WI,
                     double RB, double PB, double UB, double VB, double WB,
                     double RE, double PE, double UE, double VE, double WE, double GAM)
{
	double AGAM = GAM - 1.0;
	double fSB = sqrt(RB), fSE = sqrt(RE);
	double fS_ = 1.0 / (fSB + fSE);
	*RI = fSB * fSE;
	*UI = (fSB * UB + fSE * UE) * fS_;
	*VI = (fSB * VB + fSE * VE) * fS_;
	*WI = (fSB * WB + fSE * WE) * fS_;
	double EB = PB / (RB * AGAM), EE = PE / (RE * AGAM);
	double HB = EB + (UB * UB + VB * VB + WB * WB) * 0.5 + PB / RB;
	double HE = EE + (UE * UE + VE * VE + WE * WE) * 0.5 + PE / RE;
	double HI = (fSB * HB + fSE * HE) * fS_;
	*PI = (HI - (*UI * *UI + *VI * *VI + *WI * *WI) * 0.5) * *RI * (GAM - 1.0) / GAM;
	*EI = *PI / (*RI * AGAM);
}

/* FVM_TVD_IMPLICIT::calcRoeAverage, fvm_tvd_implicit.cpp:408-429 */
static void roe_average(prm* a, const prm* pL, const prm* pR, double GAM)
{
	double WI;
	roe_orig(&a->r, &a->e, &a->p, &a->u, &a->v, &WI, pL->r, pL->p, pL->u, pL->v, 0.0, pR->r, pR->p, pR->u, pR->v,
	         0.0, GAM);
	a->cz = sqrt(GAM * a->p / a->r);
	a->E = a->e + 0.5 * (a->u * a->u + a->v * a->v);
}

/* eigenValues, rightEigenVector, leftEigenVector, fvm_tvd_implicit.cpp:335-369 */
static void eigen_system(double lam[4][4], double R[4][4], double L[4][4], const prm* a, double nx, double ny,
                         double H, double GAM)
{
	double c = a->cz, u = a->u, v = a->v;
	double qn = u * nx + v * ny, q2 = u * u + v * v;
	double lx = -ny, ly = nx, ql = u * lx + v * ly;
	double g1 = GAM - 1.0, c2 = c * c;
	memset(lam, 0, 16 * sizeof(double));
	lam[0][0] = qn - c;
	lam[1][1] = qn;
	lam[2][2] = qn + c;
	lam[3][3] = qn;
	R[0][0] = 1.0;        R[0][1] = 1.0;     R[0][2] = 1.0;        R[0][3] = 0.0;
	R[1][0] = u - c * nx; R[1][1] = u;       R[1][2] = u + c * nx; R[1][3] = lx;
	R[2][0] = v - c * ny; R[2][1] = v;       R[2][2] = v + c * ny; R[2][3] = ly;
	R[3][0] = H - qn * c; R[3][1] = q2 / 2;  R[3][2] = H + qn * c; R[3][3] = ql;
	L[0][0] = 0.5 * (0.5 * q2 * g1 / c2 + qn / c);
	L[0][1] = -0.5 * (g1 * u / c2 + nx / c);
	L[0][2] = -0.5 * (g1 * v / c2 + ny / c);
	L[0][3] = 0.5 * g1 / c2;
	L[1][0] = 1 - 0.5 * q2 * g1 / c2;
	L[1][1] = g1 * u / c2;
	L[1][2] = g1 * v / c2;
	L[1][3] = -g1 / c2;
	L[2][0] = 0.5 * (0.5 * q2 * g1 / c2 - qn / c);
	L[2][1] = -0.5 * (g1 * u / c2 - nx / c);
	L[2][2] = -0.5 * (g1 * v / c2 - ny / c);
	L[2][3] = 0.5 * g1 / c2;
	L[3][0] = -ql;
	L[3][1] = lx;
	L[3][2] = ly;
	L[3][3] = 0;
}

/* multMtx4, fvm_tvd_implicit.cpp:264-277 */
static void mult4(double d[4][4], double a[4][4], double b[4][4])
{
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++) {
			double sum = 0;
			for (int k = 0; k < 4; k++) sum += a[i][k] * b[k][j];
			d[i][j] = sum;
		}
}

/* calcAP / calcAM, fvm_tvd_implicit.cpp:371-406: 0.5 R diag(lambda+-) L */
static void split_part(double d[4][4], double R[4][4], double lam[4][4], double L[4][4], int plus)
{
	double t[4][4], D[4][4];
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			D[i][j] = plus ? (lam[i][j] > 0 ? lam[i][j] : 0) : (lam[i][j] < 0 ? lam[i][j] : 0);
	mult4(t, R, D);
	mult4(d, t, L);
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++) d[i][j] *= 0.5;
}

/* A+ l and A- l of one face for normal n at the Roe state a */
static void split_face(double Ap[4][4], double Am[4][4], const prm* a, double nx, double ny, double l, double GAM)
{
	double lam[4][4], R[4][4], L[4][4];
	double H = a->E + a->p / a->r;
	eigen_system(lam, R, L, a, nx, ny, H, GAM);
	split_part(Ap, R, lam, L, 1);
	split_part(Am, R, lam, L, 0);
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++) {
			Am[i][j] *= l;
			Ap[i][j] *= l;
		}
}

/* the block slot of edge e in cell i's row: 1 + its rank among the cell's edges */
static int slot_of(const or_mesh* m, int i, int e)
{
	for (int k = 0; k < 3; k++)
		if (m->cedges[3 * (size_t)i + k] == e) return k + 1;
	return -1;
}

static void add_block(double* blocks, int i, int slot, double A[4][4])
{
	double* b = &blocks[(size_t)i * 64 + 16 * (size_t)slot];
	for (int r = 0; r < 4; r++)
		for (int c = 0; c < 4; c++) b[4 * r + c] += A[r][c];
}

/*
 * The matrix fill of FVM_TVD_IMPLICIT::run (src/methods/fvm_tvd_implicit.cpp:
 * 534-663) for the current state: per edge in index order the Roe average of
 * the centre states, A+(n) l into (c1,c1), A-(n) l into (c1,c2), A+(-n) l into
 * (c2,c2), A-(-n) l into (c2,c1); the 2-Gauss-point flux times l/2 out of c1
 * and into c2; the steady spectral sum; then S/ctau on the diagonal.
 * Output as block rows: blocks[nc][4][16] (slot 0 diagonal, slot k+1 the
 * k-th face in ascending edge order), cols[nc][4], rhs[nc][4], ctau[nc].
 * GAM is the configured flux gamma (the reference's __GAM = 1.4).
 */
int or_implicit_assemble(or_solver* s, int steady, double cfl, double tau, double* blocks, int* cols, double* rhs,
                         double* ctau)
{
	const or_mesh* m = s->m;
	int nc = m->nc, ne = m->ne;
	double GAM = s->c.gam_flux;
	for (int b = 0; b < s->c.nbc; b++)
		if (s->c.bc_type[b] == OR_BND_PRESSURE_OUTLET) return 1;
	if (s->c.flux != OR_FLUX_GODUNOV && s->c.flux != OR_FLUX_LAX) return 2;
	prim_pass(s, s->U);
	double* g = (double*)malloc(8 * (size_t)nc * sizeof(double));
	double* lam_sum = (double*)calloc(nc, sizeof(double));
	memset(blocks, 0, 64 * (size_t)nc * sizeof(double));
	memset(rhs, 0, 4 * (size_t)nc * sizeof(double));
	for (int i = 0; i < nc; i++) {
		cols[4 * (size_t)i] = i;
		for (int k = 0; k < 3; k++) {
			int e = m->cedges[3 * (size_t)i + k];
			int j = m->c1[e] == i ? m->c2[e] : m->c1[e];
			cols[4 * (size_t)i + 1 + k] = j;
		}
	}
	impl_grad(s, g);
	long hits = 0;
	for (int e = 0; e < ne; e++) {
		int c1 = m->c1[e], c2 = m->c2[e];
		double l = m->l[e], nx = m->nx[e], ny = m->ny[e];
		prm cL, cR, avg;
		impl_centre(s, e, &cL, &cR);
		roe_average(&avg, &cL, &cR, GAM);
		if (steady) {
			double lambda = sqrt(avg.u * avg.u + avg.v * avg.v) + avg.cz;
			lambda *= l;
			lam_sum[c1] += lambda;
			if (c2 >= 0) lam_sum[c2] += lambda;
		}
		double Ap[4][4], Am[4][4];
		split_face(Ap, Am, &avg, nx, ny, l, GAM);
		add_block(blocks, c1, 0, Ap);
		if (c2 >= 0) {
			add_block(blocks, c1, slot_of(m, c1, e), Am);
			split_face(Ap, Am, &avg, -nx, -ny, l, GAM);
			add_block(blocks, c2, 0, Ap);
			add_block(blocks, c2, slot_of(m, c2, e), Am);
		}
		double fr[4] = {0.0, 0.0, 0.0, 0.0};
		for (int gp = 0; gp < 2; gp++) {
			prm pL, pR;
			impl_recon(s, g, e, gp, &pL, &pR);
			double f[4];
			int hit = 0;
			if (s->c.flux == OR_FLUX_LAX) flux_lax(f, &pL, &pR, nx, ny, GAM);
			else flux_godunov(f, &pL, &pR, nx, ny, GAM, s->c.max_newton, &hit);
			hits += hit;
			for (int q = 0; q < 4; q++) fr[q] += f[q];
		}
		for (int q = 0; q < 4; q++) rhs[4 * (size_t)c1 + q] -= 0.5 * l * fr[q];
		if (c2 >= 0)
			for (int q = 0; q < 4; q++) rhs[4 * (size_t)c2 + q] += 0.5 * l * fr[q];
	}
	for (int i = 0; i < nc; i++) {
		double ct = steady ? 2.0 * cfl * m->S[i] / (lam_sum[i] + 1.0e-100) : tau;
		if (ctau) ctau[i] = ct;
		for (int q = 0; q < 4; q++) blocks[(size_t)i * 64 + 5 * q] += m->S[i] / ct;
	}
	s->cap_hits += hits;
	free(g);
	free(lam_sum);
	return 0;
}

/* the face Jacobians alone: rows (rL,pL,uL,vL,rR,pR,uR,vR,nx,ny) ->
 * (Roe r,u,v,p,c,H, A+ [16], A- [16]) with l = 1 */
void or_roe_jacobians(int n, const double* in, double* out, double gam)
{
	for (int i = 0; i < n; i++) {
		const double* r = &in[10 * (size_t)i];
		double* o = &out[38 * (size_t)i];
		prm pL, pR, a;
		memset(&pL, 0, sizeof(pL));
		memset(&pR, 0, sizeof(pR));
		pL.r = r[0]; pL.p = r[1]; pL.u = r[2]; pL.v = r[3];
		pR.r = r[4]; pR.p = r[5]; pR.u = r[6]; pR.v = r[7];
		roe_average(&a, &pL, &pR, gam);
		double Ap[4][4], Am[4][4];
		split_face(Ap, Am, &a, r[8], r[9], 1.0, gam);
		o[0] = a.r; o[1] = a.u; o[2] = a.v; o[3] = a.p; o[4] = a.cz; o[5] = a.E + a.p / a.r;
		for (int k = 0; k < 16; k++) {
			o[6 + k] = Ap[k / 4][k % 4];
			o[22 + k] = Am[k / 4][k % 4];
		}
	}
}

/* ------------------------------------------------------------ API */

or_solver* or_solver_create(const or_mesh* m, const or_cfg* cfg, const int* cell_mat, const int* edge_bc)
{
	or_solver* s = (or_solver*)calloc(1, sizeof(or_solver));
	s->m = m;
	s->c = *cfg;
	int nc = m->nc, ne = m->ne;
	s->cell_mat = (int*)malloc(sizeof(int) * nc);
	memcpy(s->cell_mat, cell_mat, sizeof(int) * nc);
	s->edge_bc = (int*)malloc(sizeof(int) * ne);
	memcpy(s->edge_bc, edge_bc, sizeof(int) * ne);
	size_t n4 = 4 * (size_t)nc;
	s->U = (double*)calloc(n4, sizeof(double));
	s->U0 = (double*)calloc(n4, sizeof(double));
	s->U1 = (double*)calloc(n4, sizeof(double));
	s->R = (double*)calloc(n4, sizeof(double));
	s->G = (double*)calloc(2 * n4, sizeof(double));
	s->F = (double*)calloc(4 * (size_t)ne, sizeof(double));
	s->ctau = (double*)calloc(nc, sizeof(double));
	s->P = (prm*)calloc(nc, sizeof(prm));
	s->flags = (unsigned*)calloc(nc, sizeof(unsigned));
	s->TAU = cfg->tau;
	return s;
}

void or_solver_free(or_solver* s)
{
	if (!s) return;
	free(s->cell_mat); free(s->edge_bc); free(s->U); free(s->U0); free(s->U1); free(s->R);
	free(s->G); free(s->F); free(s->ctau); free(s->P); free(s->flags); free(s->lsq_w);
	free(s);
}

void or_set_state(or_solver* s, const double* u) { memcpy(s->U, u, 4 * (size_t)s->m->nc * sizeof(double)); }
void or_get_state(const or_solver* s, double* u) { memcpy(u, s->U, 4 * (size_t)s->m->nc * sizeof(double)); }
void or_set_flags(or_solver* s, const unsigned* f) { memcpy(s->flags, f, (size_t)s->m->nc * sizeof(unsigned)); }
void or_get_flags(const or_solver* s, unsigned* f) { memcpy(f, s->flags, (size_t)s->m->nc * sizeof(unsigned)); }

double or_get_tau(const or_solver* s, double* ctau)
{
	if (ctau) memcpy(ctau, s->ctau, (size_t)s->m->nc * sizeof(double));
	return s->TAU;
}

void or_set_tau(or_solver* s, double tau)
{
	s->TAU = tau;
	for (int i = 0; i < s->m->nc; i++) s->ctau[i] = tau;
}

void or_gradients(or_solver* s, double* g)
{
	prim_pass(s, s->U);
	gradients(s);
	if (g) memcpy(g, s->G, 8 * (size_t)s->m->nc * sizeof(double));
}

void or_edge_fluxes(or_solver* s, double* f)
{
	prim_pass(s, s->U);
	gradients(s);
	edge_pass(s);
	if (f) memcpy(f, s->F, 4 * (size_t)s->m->ne * sizeof(double));
}

void or_residual(or_solver* s, double* r)
{
	stage_residual(s, s->U);
	if (r) memcpy(r, s->R, 4 * (size_t)s->m->nc * sizeof(double));
}

void or_rim(int n, const double* in, double* out, int max_iter, int* iters)
{
	for (int i = 0; i < n; i++) {
		const double* r = &in[9 * (size_t)i];
		double* o = &out[6 * (size_t)i];
		int hit;
		rim(&o[0], &o[1], &o[2], &o[3], &o[4], &o[5], r[0], r[1], r[2], r[3], 0.0, r[4], r[5], r[6], r[7],
		    0.0, r[8], max_iter, iters ? &iters[i] : NULL, &hit);
	}
}

void or_flux(int kind, int n, const double* in, double* out, double gam)
{
	for (int i = 0; i < n; i++) {
		const double* r = &in[10 * (size_t)i];
		prm pL, pR;
		memset(&pL, 0, sizeof(pL));
		memset(&pR, 0, sizeof(pR));
		pL.r = r[0]; pL.p = r[1]; pL.u = r[2]; pL.v = r[3];
		pR.r = r[4]; pR.p = r[5]; pR.u = r[6]; pR.v = r[7];
		double* f = &out[4 * (size_t)i];
		int hit;
		switch (kind) {
		case OR_FLUX_LAX: flux_lax(f, &pL, &pR, r[8], r[9], gam); break;
		case OR_FLUX_HLLC: flux_hllc(f, &pL, &pR, r[8], r[9], gam); break;
		case OR_FLUX_ROE: flux_roe(f, &pL, &pR, r[8], r[9], gam); break;
		default: flux_godunov(f, &pL, &pR, r[8], r[9], gam, 0, &hit); break;
		}
	}
}

void or_set_threads(int n)
{
#ifdef _OPENMP
	omp_set_num_threads(n > 0 ? n : 1);
#else
	(void)n;
#endif
}
