// Assembly of the implicit method's linear system on the device (the
// reference's FVM_TVD_IMPLICIT::run matrix fill, src/methods/
// fvm_tvd_implicit.cpp:534-646); the solve stays on the host.  Included at
// the end of solver.hip.
//
// Per step the reference forms, for the increment dU,
//   [S/tau I + sum_faces A+(n_out) l] dU_i + sum_faces A-(n_out) l dU_j = RHS_i
// where A+- = 0.5 R diag(lambda+-) L at the Roe average of the two cells'
// centre states (ghost on a boundary), n_out the normal out of cell i, and
// RHS_i the explicit residual of the implicit method's own flux (2 Gauss
// points, linear reconstruction on interior faces, centre values on
// boundary faces).  Steady runs take tau_i = 2 CFL S_i / (sum l (|u|+c) of
// the Roe states + 1e-100).
//
// Layout of the result, one block row per cell (BSR with at most 4 blocks a
// row): blocks[nc][4][16] row-major 4x4 blocks, slot 0 the diagonal, slots
// 1..3 the faces in ascending edge order; cols[nc][4] the block columns
// (slot 0 = the cell, -1 for a boundary face); rhs[nc][4]; ctau[nc].
//
// Three passes: gradients (Green-Gauss with the implicit ghost), the edge
// flux (two lanes per edge), and one cell-centric pass gathering the RHS in
// ascending edge order and forming the cell's Jacobian blocks face by face
// (each cell recomputes the Roe average of its faces: no per-edge matrix
// buffer, no atomics).

namespace {

template <bool MM>
__device__ __forceinline__ Prim ghost_impl_of(const Bufs& b, const Phys& ph, int e, int c1, const Prim& inside)
{
	int bc = -edge_cells(b, e).y - 1;
	return ghost_implicit(b.bctype[bc], &b.bcpar[4 * bc], edge_geo(b, e)[0], edge_geo(b, e)[1], inside, mat_of<MM>(b, ph, c1));
}

// FVM_TVD_IMPLICIT::calcGrad (fvm_tvd_implicit.cpp:186-262): Green-Gauss of
// the cell-centre primitives, the implicit ghost on a boundary
template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_impl_gradient(Bufs b, Phys ph)
{
	__shared__ double lds[BLOCK * 9];
	int c0 = xcd_block() * BLOCK;
	int i = c0 + threadIdx.x;
	double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (i < b.nc) {
		Prim pi = load_prim(b, i);
		CellTop tp = load_top(b, i);
		for (int k = 0; k < 3; k++) {
			int j = tp.nb[k];
			Prim q = j >= 0 ? load_prim(b, j) : ghost_impl_of<MM>(b, ph, tp.f[k] >> 1, i, pi);
			// n l out of cell i (the reference adds for c1 and subtracts for c2)
			double fx = b.cfn[6 * (size_t)i + 2 * k], fy = b.cfn[6 * (size_t)i + 2 * k + 1];
			double s[4] = {pi.r + q.r, pi.p + q.p, pi.u + q.u, pi.v + q.v};
			for (int c = 0; c < 4; c++) {
				g[2 * c] += s[c] / 2 * fx;
				g[2 * c + 1] += s[c] / 2 * fy;
			}
		}
		double isi = 1.0 / b.S[i];
		for (int q = 0; q < 8; q++) g[q] *= isi;
	}
	store_records<8>(b.G, c0, min(BLOCK, b.nc - c0), g, lds);
}

// The RHS flux of each edge (fvm_tvd_implicit.cpp:612-627): sum over the two
// Gauss points, reconstructed on interior faces (reconstruct with a point,
// 444-476), centre state and ghost on boundary faces; two lanes per edge,
// F = (f1 + f2) l/2
template <bool MM, int FLUX>
__global__ __launch_bounds__(BLOCK) void k_impl_edge(Bufs b, Phys ph)
{
	int t = xcd_block() * BLOCK + threadIdx.x;
	int e = t >> 1, gp = t & 1;
	if (e >= b.ne) return;  // both lanes of a pair leave together
	int2 ce = edge_cells(b, e);
	int c1 = ce.x, c2 = ce.y;
	double nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
	Prim L = load_prim(b, c1), R;
	if (c2 >= 0) {
		recon(L, b, c1, b.edl[edl_index(e, 0, gp)], b.edl[edl_index(e, 0, gp) + 1]);
		R = load_prim(b, c2);
		recon(R, b, c2, b.edl[edl_index(e, 1, gp)], b.edl[edl_index(e, 1, gp) + 1]);
	} else {
		R = ghost_impl_of<MM>(b, ph, e, c1, L);
	}
	double f[4];
	int hit = 0;
	if (FLUX == CFD2D_FLUX_LAX) flux_lax(f, L, R, nx, ny, ph.gam_flux);
	else flux_godunov(f, L, R, nx, ny, ph.rk, ph.max_newton, hit);
	if (hit) atomicAdd(&b.err[0], 1ull);
	double o[4];
	for (int q = 0; q < 4; q++) o[q] = pair_swap(f[q]);
	store_edge_flux(b.F, e, gp, f, o, edge_geo(b, e)[2] * 0.5);
}

struct ImplOut {
	double* blocks;  // [nc][4][16]
	int* cols;       // [nc][4]
	double* rhs;     // [nc][4]
	double* ctau;    // [nc]
	double cfl, tau;
	int steady;
};

// As store_records, for records of W doubles that lie `stride` doubles apart
// (here one 4x4 block of each block row): eight lanes write one record's
// 128 bytes.  The trailing barrier lets the caller reuse lds at once.
template <int W>
__device__ __forceinline__ void store_strided(double* base, size_t stride, int c0, int nvalid, const double* v,
                                              double* lds)
{
	constexpr int PAD = W + 1;
	int t = threadIdx.x;
	for (int q = 0; q < W; q++) lds[t * PAD + q] = v[q];
	__syncthreads();
	int nd = nvalid * W;
	for (int d = 2 * t; d < nd; d += 2 * BLOCK) {
		int c = d / W, q = d % W;  // W even: d and d + 1 lie in one record
		double* o = static_cast<double*>(__builtin_assume_aligned(base + (size_t)(c0 + c) * stride + q, 16));
		__builtin_nontemporal_store(lds[c * PAD + q], o);  // written once, read by the host copy
		__builtin_nontemporal_store(lds[c * PAD + q + 1], o + 1);
	}
	__syncthreads();
}

// One block row per cell, one lane per cell: the faces' blocks in ascending
// edge order (the reference's order for the diagonal sum), each staged
// through LDS so that a store instruction writes whole 128-byte blocks; the
// RHS gathered in ascending edge order like the explicit residual.  Lanes
// past the last cell compute on the block's first cell and store nothing.
template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_impl_assemble(Bufs b, Phys ph, ImplOut o)
{
	__shared__ double lds[BLOCK * 17];
	int c0 = xcd_block() * BLOCK;
	int nvalid = min(BLOCK, b.nc - c0);
	int i = c0 + (threadIdx.x < nvalid ? threadIdx.x : 0);
	Prim pi = load_prim(b, i);
	RoeSide si = roe_side(pi.r);
	double diag[16];
	for (int q = 0; q < 16; q++) diag[q] = 0.0;
	double lam = 0.0;
	int cl[4];
	cl[0] = i;
	CellTop tp = load_top(b, i);
	for (int k = 0; k < 3; k++) {
		int f = tp.f[k];
		int e = f >> 1, side = f & 1;
		int j = tp.nb[k];
		// the edge's (c1, c2) centre states, as reconstruct(iEdge, L, R) gives them
		Prim q = j >= 0 ? load_prim(b, j) : ghost_impl_of<MM>(b, ph, e, i, pi);
		RoeSide sq = roe_side(q.r);
		RoeAvg a = side ? roe_average(q, sq, pi, si, ph.gam_flux) : roe_average(pi, si, q, sq, ph.gam_flux);
		double l = edge_geo(b, e)[2];
		if (o.steady) lam += (sqrt(a.u * a.u + a.v * a.v) + a.c) * l;
		double sn = side ? -1.0 : 1.0;
		double am[16];
		split_face(a, sn * edge_geo(b, e)[0], sn * edge_geo(b, e)[1], ph.gam_flux, l, diag, am);
		if (j < 0)
			for (int q2 = 0; q2 < 16; q2++) am[q2] = 0.0;
		cl[k + 1] = j;
		store_strided<16>(o.blocks + 16 * (k + 1), 64, c0, nvalid, am, lds);
	}
	double S = b.S[i];
	double ct = o.steady ? 2.0 * o.cfl * S / (lam + 1.0e-100) : o.tau;
	for (int q = 0; q < 4; q++) diag[5 * q] += S / ct;
	store_strided<16>(o.blocks, 64, c0, nvalid, diag, lds);
	if (threadIdx.x < nvalid) {
		double r[4];
		gather_residual(b, i, r);
		st4(o.rhs, i, r);
		int4* cp = reinterpret_cast<int4*>(o.cols) + i;
		*cp = make_int4(cl[0], cl[1], cl[2], cl[3]);
		o.ctau[i] = ct;
	}
}

}  // namespace

extern "C" int cfd2d_implicit_assemble(cfd2d_ctx* c, int steady, double cfl, double tau, double* blocks, int* cols,
                                       double* rhs, double* ctau)
{
	if (!c || (!steady && !(tau > 0.0)) || (steady && !(cfl > 0.0)))
		return fail(CFD2D_ERR_ARG, "cfd2d_implicit_assemble: bad arguments");
	if (c->p.flux != CFD2D_FLUX_GODUNOV && c->p.flux != CFD2D_FLUX_LAX)
		return fail(CFD2D_ERR_CONFIG, "implicit assembly: FLUX must be GODUNOV or LAX");
	std::vector<int> bt(c->nbc > 0 ? c->nbc : 1);
	if (c->nbc > 0) HIPCHK(hipMemcpy(bt.data(), c->b.bctype, c->nbc * sizeof(int), hipMemcpyDeviceToHost));
	for (int k = 0; k < c->nbc; k++)
		if (bt[k] == CFD2D_BND_PRESSURE_OUTLET)
			return fail(CFD2D_ERR_CONFIG, "implicit assembly: no pressure-outlet boundary in the implicit method");
	int nc = c->nc;
	size_t need = 64 * (size_t)nc * sizeof(double) + 4 * (size_t)nc * sizeof(int) + 5 * (size_t)nc * sizeof(double) + 1024;
	if (need > c->impl_ws_bytes) {
		if (c->impl_ws) (void)hipFree(c->impl_ws);
		c->impl_ws = nullptr;
		c->impl_ws_bytes = 0;
		HIPCHK(hipMalloc(&c->impl_ws, need));
		c->impl_ws_bytes = need;
	}
	char* base = (char*)c->impl_ws;
	ImplOut o;
	o.blocks = (double*)base;  // 512 B per cell, then rhs 32 B, cols 16 B, ctau 8 B: every array aligned
	o.rhs = o.blocks + 64 * (size_t)nc;
	o.cols = (int*)(o.rhs + 4 * (size_t)nc);
	o.ctau = (double*)(o.cols + 4 * (size_t)nc);
	o.cfl = cfl;
	o.tau = tau;
	o.steady = steady;
	rec_start(c, CFD2D_K_IMPL_GRADIENT);
	if (c->mm) hipLaunchKernelGGL((k_impl_gradient<true>), dim3(grid_of(nc)), dim3(BLOCK), 0, c->st, c->b, c->ph);
	else hipLaunchKernelGGL((k_impl_gradient<false>), dim3(grid_of(nc)), dim3(BLOCK), 0, c->st, c->b, c->ph);
	rec_stop(c);
	dim3 ge(grid_of(2 * c->ne));
	rec_start(c, CFD2D_K_IMPL_EDGE);
	if (c->p.flux == CFD2D_FLUX_LAX) {
		if (c->mm) hipLaunchKernelGGL((k_impl_edge<true, CFD2D_FLUX_LAX>), ge, dim3(BLOCK), 0, c->st, c->b, c->ph);
		else hipLaunchKernelGGL((k_impl_edge<false, CFD2D_FLUX_LAX>), ge, dim3(BLOCK), 0, c->st, c->b, c->ph);
	} else {
		if (c->mm) hipLaunchKernelGGL((k_impl_edge<true, CFD2D_FLUX_GODUNOV>), ge, dim3(BLOCK), 0, c->st, c->b, c->ph);
		else hipLaunchKernelGGL((k_impl_edge<false, CFD2D_FLUX_GODUNOV>), ge, dim3(BLOCK), 0, c->st, c->b, c->ph);
	}
	rec_stop(c);
	rec_start(c, CFD2D_K_IMPL_ASSEMBLE);
	if (c->mm) hipLaunchKernelGGL((k_impl_assemble<true>), dim3(grid_of(nc)), dim3(BLOCK), 0, c->st, c->b, c->ph, o);
	else hipLaunchKernelGGL((k_impl_assemble<false>), dim3(grid_of(nc)), dim3(BLOCK), 0, c->st, c->b, c->ph, o);
	rec_stop(c);
	int rc = check_launch("implicit assembly");
	if (rc) return rc;
	if (blocks) HIPCHK(hipMemcpyAsync(blocks, o.blocks, 64 * (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	if (cols) HIPCHK(hipMemcpyAsync(cols, o.cols, 4 * (size_t)nc * sizeof(int), hipMemcpyDeviceToHost, c->st));
	if (rhs) HIPCHK(hipMemcpyAsync(rhs, o.rhs, 4 * (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	if (ctau) HIPCHK(hipMemcpyAsync(ctau, o.ctau, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	return cfd2d_sync(c);  // waits, and reports a Newton cap reached in the flux
}

extern "C" int cfd2d_host_alloc(size_t bytes, void** p)
{
	if (!p) return fail(CFD2D_ERR_ARG, "cfd2d_host_alloc: null");
	*p = nullptr;
	HIPCHK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
	return CFD2D_OK;
}

extern "C" int cfd2d_host_free(void* p)
{
	if (p) HIPCHK(hipHostFree(p));
	return CFD2D_OK;
}
