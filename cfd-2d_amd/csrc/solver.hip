// Explicit finite-volume step on MI355X (gfx950): kernels, device context
// and the context half of the C ABI.
//
// Data layout (structure of arrays, FP64, cells in the reference's order):
//   cells  U[4][nc] conserved state, U0[4][nc] the RK stage buffer
//          (launch_update; structure of arrays: streamed), P[nc][4] primitives (r, p, u, v) of U (Po: where the
//          update writes them, = P except while the adjoint records stages)
//          and G[nc][8] gradients (rx ry px py ux uy vx vy) as one 32 B / 64 B
//          record per cell, because they are gathered by neighbours: one
//          record is two or four 16-byte loads instead of four or eight
//          8-byte gathers; dtS[nc] = tau_i / S_i,
//          S[nc], cface[nc][4] faces = edge*2 + side in ascending edge
//          order and cnb[nc][4] the neighbour across each face (-1 on a
//          boundary), each one 16 B record per cell (the fourth int unused;
//          the update pass reads only the faces),
//          lsqw[nc][6] least-squares weights (wx wy per face), cfn[nc][6]
//          the faces' outward n l (Green-Gauss weights), both one 48 B
//          record per cell (three 16-byte loads), cgp[12][nc] the cell's offsets
//          to its faces' Gauss points (limiter only), flags[nc]
//   edges  eg[ne][4] one 32 B record per edge: unit normal out of c1 and
//          length (nx, ny, l), then the two cells c1, c2 as two ints in the
//          last 8 bytes (c2 = -(bc+1) on a boundary),
//          edl[ne][2][4] Gauss-point offsets from the cell centroids, one
//          32 B record per lane of the edge's pair (edl_index): [0] c1 at
//          Gauss points 1, 2; [1] c2 at Gauss points 2, 1 (on a boundary
//          c1 at Gauss point 2, then zeros),
//          F[ne][4] = (flux at gp1 + flux at gp2) * l/2, one record per edge
//          written once
// Each RK stage is three passes: gradient (cell gather), edge flux (a lane
// pair per edge, result written once to F), update (cell gather of its
// three faces in ascending edge order, then the RK combination); without a
// limiter the edge pass forms the two cells' gradients itself and the
// gradient pass is not run.  No floating-point atomics: results do not
// depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/cfd2d.h"
#include "common.hpp"
#include "physics.hpp"

using namespace cfd2d;
using namespace cfd2d::dev;

namespace {

constexpr int BLOCK = 256;
constexpr int NXCD = 8;

struct Bufs {
	int nc, ne, nblocks, nmat;
	double *U, *U0, *P, *Po, *G, *dtS, *S, *lsqw, *cfn, *cgp, *loc, *R;
	double* Un;  // where the update pass writes the stage's state (launch_update)
	int *cface, *cnb, *cmat;
	unsigned* flags;
	int* blockflag;
	int* anyflag;  // [0] set by a final-stage update block holding a LIM cell,
	               // cleared by remediation; [1] the same, cleared only by the host
	double *eg, *edl, *F;  // eg [ne][4] (nx, ny, l, (c1, c2) as two ints)
	int* bctype;
	double* bcpar;
	double* mtab;  // 5 per material: M, Cv, gam, mu, kc
	unsigned long long* err;  // [0] Newton cap hits, [1] non-finite cells
	// tiled edge pass (k_edge_tile): per tile (int4) the first of its staged
	// cells in tcell, their count, its first edge and edge count; tcell the
	// staged cells of every tile, ascending; eslot per edge the LDS rows of
	// c1 (low 16 bits) and c2 (high 16 bits, 0xffff on a boundary)
	int* tiles;
	int* tcell;
	unsigned* eslot;
	int ntiles;  // 0: the edges are not in c1 order, no tiled pass
};

struct Phys {
	MatV m0;
	RimConst rk;
	double gam_flux, igm1, cfl;  // igm1 = 1 / (gam_flux - 1)
	double ro_min, ro_max, p_min, p_max, u_max;
	int max_newton, cfl_mode, steady;
	int store_g;  // the update pass also stores the gradient records of its input state (adjoint recording)
};

// Block index remapped so that consecutive index ranges stay on one XCD
// (workgroups are dispatched round-robin over the 8 XCDs): neighbouring
// cells/edges then share that XCD's L2.
__device__ __forceinline__ int xcd_block()
{
	int b = blockIdx.x, n = gridDim.x;
	int full = n / NXCD, rem = n % NXCD;
	int x = b % NXCD, loc = b / NXCD;
	int base = x < rem ? x * (full + 1) : rem * (full + 1) + (x - rem) * full;
	return base + loc;
}

template <bool MM>
__device__ __forceinline__ MatV mat_of(const Bufs& b, const Phys& ph, int i)
{
	if (!MM) return ph.m0;
	const double* t = &b.mtab[5 * b.cmat[i]];
	return MatV{t[0], t[1], t[2], t[3], t[4]};
}

// 32-byte records: four doubles per cell (buffers are 256-byte aligned, so
// each record is 32-byte aligned and loads/stores as 16-byte pieces)
__device__ __forceinline__ void ld4(const double* base, size_t i, double* v)
{
	const double* p = static_cast<const double*>(__builtin_assume_aligned(base + 4 * i, 32));
	v[0] = p[0];
	v[1] = p[1];
	v[2] = p[2];
	v[3] = p[3];
}

__device__ __forceinline__ void st4(double* base, size_t i, const double* v)
{
	double* p = static_cast<double*>(__builtin_assume_aligned(base + 4 * i, 32));
	p[0] = v[0];
	p[1] = v[1];
	p[2] = v[2];
	p[3] = v[3];
}

__device__ __forceinline__ void ld8(const double* base, size_t i, double* v)
{
	ld4(base, 2 * i, v);
	ld4(base, 2 * i + 1, v + 4);
}

// The value x of the other lane of this lane's pair (lanes 2k and 2k + 1):
// a DPP quad permutation [1, 0, 3, 2] of each 32-bit half, a VALU move,
// where __shfl_xor(x, 1) goes through the LDS crossbar and waits for it.
// The kernels keep both lanes of a pair active together.
__device__ __forceinline__ double pair_swap(double x)
{
	unsigned long long v = __builtin_bit_cast(unsigned long long, x);
	unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)v, 0xB1, 0xF, 0xF, false);
	unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(v >> 32), 0xB1, 0xF, 0xF, false);
	return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// A record that both lanes of a pair need (same index i in both): lane gp
// loads half gp and the pair swaps the halves -- half the loads per lane.
// N = 4 or 8 doubles.
template <int N>
__device__ __forceinline__ void pair_load(const double* base, size_t i, int gp, double* v)
{
	constexpr int H = N / 2;
	double m[H], o[H];
	const double* p = static_cast<const double*>(__builtin_assume_aligned(base + N * i + H * gp, 8 * H));
	for (int q = 0; q < H; q++) m[q] = p[q];
	for (int q = 0; q < H; q++) o[q] = pair_swap(m[q]);
	for (int q = 0; q < H; q++) {
		v[q] = gp ? o[q] : m[q];
		v[H + q] = gp ? m[q] : o[q];
	}
}

// An edge's record (32 bytes, two 16-byte loads): unit normal out of c1,
// length, and its two cells (c2 = -(bc+1) on a boundary) in the last 8 bytes.
__device__ __forceinline__ const double* edge_geo(const Bufs& b, size_t e)
{
	return static_cast<const double*>(__builtin_assume_aligned(b.eg + 4 * e, 32));
}

__device__ __forceinline__ int2 edge_cells(const Bufs& b, size_t e)
{
	return *reinterpret_cast<const int2*>(edge_geo(b, e) + 3);
}

// Where the offset (x, y) of Gauss point g (0, 1) from the centroid of the
// edge's side-th cell lies in edl: each lane of the edge's pair reads one
// 32-byte record holding its own cell's offsets at its own Gauss point
// first and at the partner's second.
__host__ __device__ __forceinline__ size_t edl_index(size_t e, int side, int g)
{
	return side ? 8 * e + 4 + 2 * (1 - g) : 8 * e + 2 * g;
}

// the cell's faces and neighbours (one 16-byte record each)
struct CellTop {
	int f[3], nb[3];
};

__device__ __forceinline__ CellTop load_top(const Bufs& b, size_t i)
{
	int4 a = reinterpret_cast<const int4*>(b.cface)[i], c = reinterpret_cast<const int4*>(b.cnb)[i];
	return CellTop{{a.x, a.y, a.z}, {c.x, c.y, c.z}};
}

__device__ __forceinline__ int4 load_faces(const Bufs& b, size_t i)
{
	return reinterpret_cast<const int4*>(b.cface)[i];
}

// Stores the W-double records of the block's cells [c0, c0 + nvalid) held in
// registers (lane t: cell c0 + t) through LDS, so that each store instruction
// covers contiguous memory instead of one 16-byte piece every W * 8 bytes.
// Every thread of the block calls it; lds holds BLOCK * (W + 1) doubles.
// NT: non-temporal stores, for records read back only much later.
template <int W, bool NT = false>
__device__ __forceinline__ void store_records(double* rec, int c0, int nvalid, const double* v, double* lds)
{
	constexpr int PAD = W + 1;  // odd row length spreads the lanes over the LDS banks
	int t = threadIdx.x;
	for (int q = 0; q < W; q++) lds[t * PAD + q] = v[q];
	__syncthreads();
	double* dst = rec + (size_t)c0 * W;
	int nd = nvalid * W;
	for (int d = 2 * t; d < nd; d += 2 * BLOCK) {
		int c = d / W, q = d % W;  // W even: d and d + 1 lie in one record
		double* o = static_cast<double*>(__builtin_assume_aligned(dst + d, 16));
		if (NT) {
			__builtin_nontemporal_store(lds[c * PAD + q], o);
			__builtin_nontemporal_store(lds[c * PAD + q + 1], o + 1);
		} else {
			o[0] = lds[c * PAD + q];
			o[1] = lds[c * PAD + q + 1];
		}
	}
}

// the same quantities as separate arrays ([4][n] / [8][n])
__device__ __forceinline__ void lda4(const double* base, size_t n, size_t i, double* v)
{
	for (int q = 0; q < 4; q++) v[q] = base[q * n + i];
}

__device__ __forceinline__ void sta4(double* base, size_t n, size_t i, const double* v)
{
	for (int q = 0; q < 4; q++) base[q * n + i] = v[q];
}

__device__ __forceinline__ void lda8(const double* base, size_t n, size_t i, double* v)
{
	for (int q = 0; q < 8; q++) v[q] = base[q * n + i];
}

__device__ __forceinline__ void sta8(double* base, size_t n, size_t i, const double* v)
{
	for (int q = 0; q < 8; q++) base[q * n + i] = v[q];
}

__device__ __forceinline__ Prim load_prim(const Bufs& b, int i)
{
	double v[4];
	ld4(b.P, i, v);
	return Prim{v[0], v[1], v[2], v[3]};
}

__device__ __forceinline__ void store_prim_to(double* P, size_t, int i, const Prim& q)
{
	double v[4] = {q.r, q.p, q.u, q.v};
	st4(P, i, v);
}

__device__ __forceinline__ void store_prim(const Bufs& b, int i, const Prim& q) { store_prim_to(b.P, b.nc, i, q); }

// ghost of cell i's (cell-centre or reconstructed) state across boundary edge e
template <bool MM>
__device__ __forceinline__ Prim ghost_of(const Bufs& b, const Phys& ph, int e, int c1, const Prim& inside,
                                         const Prim& centre)
{
	int bc = -edge_cells(b, e).y - 1;
	MatV m = mat_of<MM>(b, ph, c1);
	double Tc = temperature(centre, m);
	double czc = sqrt(m.gam * centre.p / centre.r);
	return ghost_state(b.bctype[bc], &b.bcpar[4 * bc], edge_geo(b, e)[0], edge_geo(b, e)[1], inside, Tc, czc, m);
}

// ---------------------------------------------------------------- kernels

template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_prim(Bufs b, Phys ph)
{
	__shared__ double lds[BLOCK * 5];
	int c0 = xcd_block() * BLOCK;
	int i = c0 + threadIdx.x;
	double v[4] = {0.0, 0.0, 0.0, 0.0};
	if (i < b.nc) {
		size_t n = b.nc;
		MatV m = mat_of<MM>(b, ph, i);
		Prim p = cons_to_prim(b.U[i], b.U[n + i], b.U[2 * n + i], b.U[3 * n + i], m);
		v[0] = p.r;
		v[1] = p.p;
		v[2] = p.u;
		v[3] = p.v;
	}
	store_records<4>(b.P, c0, min(BLOCK, b.nc - c0), v, lds);
}

// Barth-Jespersen scaling with minmod(1, r) at the cell's six Gauss points
// (extension; the oracle's limit_minmod)
__device__ __forceinline__ void minmod_limit(const Bufs& b, int i, const int* face, const Prim& pi,
                                             const Prim* nb, double* g)
{
	double phi[4] = {pi.r, pi.p, pi.u, pi.v};
	double lo[4], hi[4];
	for (int q = 0; q < 4; q++) lo[q] = hi[q] = phi[q];
	for (int k = 0; k < 3; k++) {
		double v[4] = {nb[k].r, nb[k].p, nb[k].u, nb[k].v};
		for (int q = 0; q < 4; q++) {
			if (v[q] < lo[q]) lo[q] = v[q];
			if (v[q] > hi[q]) hi[q] = v[q];
		}
	}
	double d[6][2];  // offsets to the Gauss points of the faces, per cell (coalesced)
	size_t n = b.nc;
	for (int k = 0; k < 6; k++) {
		d[k][0] = b.cgp[(size_t)(2 * k) * n + i];
		d[k][1] = b.cgp[(size_t)(2 * k + 1) * n + i];
	}
	// psi = min(1, min_k r_k), r_k = (hi - phi) / D or (lo - phi) / D by the sign
	// of D, always >= 0: the smallest ratio is picked by cross-multiplied
	// comparisons (a_k den < num b_k, all terms >= 0) and divided once
	for (int q = 0; q < 4; q++) {
		double gx = g[2 * q], gy = g[2 * q + 1];
		double num = 1.0, den = 1.0;
		for (int k = 0; k < 6; k++) {
			double D = gx * d[k][0] + gy * d[k][1];
			if (D == 0.0) continue;
			double a = D > 0.0 ? hi[q] - phi[q] : phi[q] - lo[q];
			double bb = fabs(D);
			if (a * den < num * bb) {
				num = a;
				den = bb;
			}
		}
		double psi = num / den;
		if (psi < 0.0) psi = 0.0;
		g[2 * q] = psi * gx;
		g[2 * q + 1] = psi * gy;
	}
}

// Green-Gauss gradients, FVM_TVD::calcGrad (src/methods/fvm_tvd.cpp:242-301)
// as a gather: the faces of cell i in ascending edge order, each adding
// (phi_c1 + phi_c2)/2 * n * l with + on c1's side and - on c2's, then / S.
// LSQ (extension): grad = sum_k w_k (phi_k - phi_i).  pi, face and nb (the
// neighbour states, ghosts on a boundary) are returned for the limiter.
template <bool MM, bool LSQ>
__device__ __forceinline__ void cell_gradient(const Bufs& b, const Phys& ph, int i, Prim& pi, int* face, Prim* nb,
                                              double* g)
{
	pi = load_prim(b, i);
	for (int q = 0; q < 8; q++) g[q] = 0.0;
	CellTop tp = load_top(b, i);
	for (int k = 0; k < 3; k++) {
		face[k] = tp.f[k];
		int j = tp.nb[k];
		if (j >= 0) nb[k] = load_prim(b, j);
		else nb[k] = ghost_of<MM>(b, ph, face[k] >> 1, i, pi, pi);
	}
	if (LSQ) {
		const double* w = static_cast<const double*>(__builtin_assume_aligned(b.lsqw + 6 * (size_t)i, 16));
		for (int k = 0; k < 3; k++) {
			double wx = w[2 * k], wy = w[2 * k + 1];
			double dphi[4] = {nb[k].r - pi.r, nb[k].p - pi.p, nb[k].u - pi.u, nb[k].v - pi.v};
			for (int q = 0; q < 4; q++) {
				g[2 * q] += wx * dphi[q];
				g[2 * q + 1] += wy * dphi[q];
			}
		}
	} else {
		const double* w = static_cast<const double*>(__builtin_assume_aligned(b.cfn + 6 * (size_t)i, 16));
		for (int k = 0; k < 3; k++) {
			// n l out of cell i (the reference adds for c1 and subtracts for c2)
			double fx = w[2 * k], fy = w[2 * k + 1];
			double s[4] = {pi.r + nb[k].r, pi.p + nb[k].p, pi.u + nb[k].u, pi.v + nb[k].v};
			for (int q = 0; q < 4; q++) {
				g[2 * q] += s[q] / 2 * fx;
				g[2 * q + 1] += s[q] / 2 * fy;
			}
		}
		double isi = 1.0 / b.S[i];
		for (int q = 0; q < 8; q++) g[q] *= isi;
	}
}

template <bool MM, bool LSQ, bool LIM>
__global__ __launch_bounds__(BLOCK) void k_gradient(Bufs b, Phys ph)
{
	__shared__ double lds[BLOCK * 9];
	int c0 = xcd_block() * BLOCK;
	int i = c0 + threadIdx.x;
	double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	if (i < b.nc) {
		Prim pi;
		int face[3];
		Prim nb[3];
		cell_gradient<MM, LSQ>(b, ph, i, pi, face, nb, g);
		if (LIM) minmod_limit(b, i, face, pi, nb, g);
	}
	store_records<8>(b.G, c0, min(BLOCK, b.nc - c0), g, lds);
}

// linear reconstruction at offset (dx, dy) with the gradient record g
__device__ __forceinline__ void recon_g(Prim& q, const double* g, double dx, double dy)
{
	q.r += g[0] * dx + g[1] * dy;
	q.p += g[2] * dx + g[3] * dy;
	q.u += g[4] * dx + g[5] * dy;
	q.v += g[6] * dx + g[7] * dy;
}

__device__ __forceinline__ void recon(Prim& q, const Bufs& b, int c, double dx, double dy)
{
	double g[8];
	ld8(b.G, c, g);
	recon_g(q, g, dx, dy);
}

// grad T from the gradient record's (rx ry px py) g
__device__ __forceinline__ void grad_T_g(const double* g, const Prim& q, const MatV& m, double& tx, double& ty)
{
	// T = p / (r Rg), Rg = gR / M: grad T = (grad p - (p/r) grad r) / (r Rg)
	double ir = 1.0 / q.r;
	double pr = q.p * ir;
	double s = ir * m.M * (1.0 / GAS_R);
	tx = (g[2] - pr * g[0]) * s;
	ty = (g[3] - pr * g[1]) * s;
}

// The lane pair's (f_gp1 + f_gp2) l/2 as the edge's 32-byte flux record: the
// even lane adds its partner's flux to its own (the reference's order), the
// odd lane the other way round (the same sums), and lane gp writes
// components 2gp and 2gp+1, so the 64 lanes of a wave store 1 KB
// contiguously
__device__ __forceinline__ void store_edge_flux(double* F, int e, int gp, const double* f, const double* o, double hl)
{
	double a0 = gp ? f[2] + o[2] : f[0] + o[0];
	double a1 = gp ? f[3] + o[3] : f[1] + o[1];
	double* p = static_cast<double*>(__builtin_assume_aligned(F + 4 * (size_t)e + 2 * gp, 16));
	p[0] = a0 * hl;
	p[1] = a1 * hl;
}

// Edge flux pass, FVM_TVD::run edge loop + reconstruct + calcFlux
// (src/methods/fvm_tvd.cpp:329-352, 646-691, 602-622).  Two adjacent lanes
// serve one edge, one Gauss point each: lane 0 works with c1, lane 1 with c2
// (c1 on a boundary); each reconstructs (r, p, u, v) of its cell linearly at
// both Gauss points and the pair swaps one of them, so each lane holds both
// sides at its Gauss point (the ghost of the reconstructed left state on a
// boundary); each evaluates the numerical flux and subtracts the viscous flux
// (extension); the even lane adds its partner's flux to its own (f1 + f2,
// the reference's order) and writes (f1 + f2) * l/2 once.  Halving the live
// state per lane doubles the waves a SIMD can hold.
//
// GR = 0 reads the gradients the gradient pass stored; GR = 1 (LSQ) or 2
// (Green-Gauss), 3 or 4 with the minmod limiter, has each lane form its own
// cell's gradient with the gradient pass's own code.  Every cell's gradient
// is then formed by each of its edges (three times) instead of once, but the
// gradient pass, its 64-byte record per cell written and read back, and its
// launch disappear; the extra arithmetic runs inside a pass that is bound by
// FP64 issue anyway, beside its memory traffic.  Results are the same bits as
// with GR = 0.
//
// SHAPE picks the launch shape the autotuner may try (EdgeShape): threads per
// workgroup and the waves per SIMD the register allocation must allow.
template <int SHAPE>
struct EdgeShape;
template <>
struct EdgeShape<0> {  // 256 threads, registers as the compiler allocates them
	static constexpr int NT = 256, WPE = 1;
};
template <>
struct EdgeShape<1> {  // 256 threads, at most 96 VGPRs: five waves per SIMD
	static constexpr int NT = 256, WPE = 5;
};
template <>
struct EdgeShape<2> {  // 128 threads
	static constexpr int NT = 128, WPE = 1;
};
template <>
struct EdgeShape<3> {  // 512 threads
	static constexpr int NT = 512, WPE = 1;
};
constexpr int EDGE_SHAPES = 4;

// One lane of an edge's pair from the primitives ps and gradients gs of the
// cell cs it reconstructs (c2 for the second lane of an inner edge, c1
// otherwise) onward: reconstruction, ghost, viscous terms, numerical flux and
// the pair's combined store.  Shared by the edge passes, so they give the
// same bits.
template <bool MM, int FLUX, bool VISC>
__device__ __forceinline__ void edge_lane(const Bufs& b, const Phys& ph, int e, int gp, int c1, bool inner, int cs,
                                          const Prim& ps, const double* gs, double nx, double ny)
{
	// Each lane extrapolates its own cell to its own Gauss point (kept) and to
	// the partner's (sent); on an inner edge the pair swaps the sent states,
	// so lane 0 holds (c1, c2) at Gauss point 1 and lane 1 at Gauss point 2 --
	// the same operations as each lane reconstructing both cells.  On a
	// boundary both lanes reconstruct c1 and build the ghost from it.
	double dl[4];  // this lane's record: offsets of cs at its own, then the partner's Gauss point
	ld4(b.edl, 2 * (size_t)e + gp, dl);
	Prim keep = ps, send = ps, recv;
	recon_g(keep, gs, dl[0], dl[1]);
	recon_g(send, gs, dl[2], dl[3]);
	recv.r = pair_swap(send.r);
	recv.p = pair_swap(send.p);
	recv.u = pair_swap(send.u);
	recv.v = pair_swap(send.v);
	Prim L, R;
	if (inner) {
		L = gp ? recv : keep;
		R = gp ? keep : recv;
	} else {
		L = keep;
		R = ghost_of<MM>(b, ph, e, c1, L, ps);
	}
	// Viscous terms first (only sx, sy and the energy term stay live through
	// the Riemann solve).  Face gradient = mean of the two cell gradients (the
	// cell's own on a boundary); face velocity = mean of the reconstructed
	// states.  Each lane forms grad T of its own cell and the pair swaps the
	// velocity and temperature gradients; a + b = b + a, so both lanes form
	// the same means.
	double sx = 0.0, sy = 0.0, ev = 0.0;
	if (VISC) {
		MatV m = mat_of<MM>(b, ph, c1);
		double own[6];
		for (int q = 0; q < 4; q++) own[q] = gs[4 + q];  // ux uy vx vy
		grad_T_g(gs, ps, mat_of<MM>(b, ph, cs), own[4], own[5]);
		double oth[6];
		for (int q = 0; q < 6; q++) oth[q] = pair_swap(own[q]);
		double ux = own[0], uy = own[1], vx = own[2], vy = own[3], tx = own[4], ty = own[5];
		if (inner) {
			ux = 0.5 * (ux + oth[0]);
			uy = 0.5 * (uy + oth[1]);
			vx = 0.5 * (vx + oth[2]);
			vy = 0.5 * (vy + oth[3]);
			tx = 0.5 * (tx + oth[4]);
			ty = 0.5 * (ty + oth[5]);
		}
		double div = ux + vy;
		double txx = m.mu * (2.0 * ux - (2.0 / 3.0) * div);
		double tyy = m.mu * (2.0 * vy - (2.0 / 3.0) * div);
		double txy = m.mu * (uy + vx);
		sx = txx * nx + txy * ny;
		sy = txy * nx + tyy * ny;
		double uf = 0.5 * (L.u + R.u), vf = 0.5 * (L.v + R.v);
		ev = uf * sx + vf * sy + m.kc * (tx * nx + ty * ny);
	}
	double f[4];
	int hit = 0;
	if (FLUX == CFD2D_FLUX_LAX) flux_lax(f, L, R, nx, ny, ph.gam_flux);
	else if (FLUX == CFD2D_FLUX_HLLC) flux_hllc(f, L, R, nx, ny, ph.gam_flux);
	else if (FLUX == CFD2D_FLUX_ROE) flux_roe(f, L, R, nx, ny, ph.gam_flux);
	else flux_godunov(f, L, R, nx, ny, ph.rk, ph.max_newton, hit);
	if (hit) atomicAdd(&b.err[0], 1ull);
	if (VISC) {
		f[1] -= sx;
		f[2] -= sy;
		f[3] -= ev;
	}
	double o[4];
	for (int q = 0; q < 4; q++) o[q] = pair_swap(f[q]);
	store_edge_flux(b.F, e, gp, f, o, edge_geo(b, e)[2] * 0.5);
}

template <bool MM, int FLUX, bool VISC, int GR, int SHAPE = 0>
__global__ __attribute__((amdgpu_flat_work_group_size(1, EdgeShape<SHAPE>::NT),
                          amdgpu_waves_per_eu(EdgeShape<SHAPE>::WPE))) void k_edge_flux(Bufs b, Phys ph)
{
	int t = xcd_block() * EdgeShape<SHAPE>::NT + threadIdx.x;
	int e = t >> 1, gp = t & 1;
	if (e >= b.ne) return;  // both lanes of a pair leave together
	int2 ce = edge_cells(b, e);
	int c1 = ce.x, c2 = ce.y;
	double nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
	bool inner = c2 >= 0;
	int cs = gp == 1 && inner ? c2 : c1;  // the cell this lane reconstructs
	double gs[8];                          // its gradients
	Prim ps;                               // and primitives
	if (GR) {
		int face[3];
		Prim nb[3];
		cell_gradient<MM, GR == 1 || GR == 3>(b, ph, cs, ps, face, nb, gs);
		if (GR >= 3) minmod_limit(b, cs, face, ps, nb, gs);
	} else {
		ps = load_prim(b, cs);
		ld8(b.G, cs, gs);
	}
	edge_lane<MM, FLUX, VISC>(b, ph, e, gp, c1, inner, cs, ps, gs, nx, ny);
}

// Tiled edge pass (fuse mode 2): one workgroup per tile of TILE_CELLS
// consecutive cells.  The tile owns the edges whose c1 it holds -- a
// contiguous edge range, since edges are numbered in c1 order -- and first
// forms, once each, the gradients of every cell those edges read (its own
// cells and the c2 cells beyond it) with the gradient pass's code, staging
// primitives and gradients in LDS; then the lane pairs of its edges run
// edge_lane from LDS.  A cell's gradient is formed by ~1.25 tiles instead of
// by each of its three edges, and the edge lanes gather nothing but their
// edge's own records.  The same bits as the other edge passes.
constexpr int TILE_CELLS = 256;  // cells per tile (smaller when the staged set would not fit)
constexpr int TILE_CAP = 384;    // staged cells per tile at most
constexpr int TILE_ROW = 12;     // LDS doubles per staged cell: r, p, u, v, then the 8 gradients

template <bool MM, int FLUX, bool VISC, int GR>
__global__ __launch_bounds__(BLOCK) void k_edge_tile(Bufs b, Phys ph)
{
	__shared__ double stage[TILE_CAP * TILE_ROW];
	int4 tl = reinterpret_cast<const int4*>(b.tiles)[xcd_block()];  // first staged cell, count, first edge, count
	for (int k = threadIdx.x; k < tl.y; k += BLOCK) {
		int c = b.tcell[tl.x + k];
		Prim pi;
		int face[3];
		Prim nb[3];
		double g[8];
		cell_gradient<MM, GR == 1 || GR == 3>(b, ph, c, pi, face, nb, g);
		if (GR >= 3) minmod_limit(b, c, face, pi, nb, g);
		double* r = stage + TILE_ROW * k;
		r[0] = pi.r;
		r[1] = pi.p;
		r[2] = pi.u;
		r[3] = pi.v;
		for (int q = 0; q < 8; q++) r[4 + q] = g[q];
	}
	__syncthreads();
	// lane pairs (t, t + 1) stay together: BLOCK is even
	for (int t = threadIdx.x; t < 2 * tl.w; t += BLOCK) {
		int e = tl.z + (t >> 1), gp = t & 1;
		int2 ce = edge_cells(b, e);
		bool inner = ce.y >= 0;
		bool second = gp == 1 && inner;
		unsigned sl = b.eslot[e];
		const double* r = stage + TILE_ROW * (int)(second ? sl >> 16 : sl & 0xffffu);
		Prim ps{r[0], r[1], r[2], r[3]};
		double gs[8];
		for (int q = 0; q < 8; q++) gs[q] = r[4 + q];
		edge_lane<MM, FLUX, VISC>(b, ph, e, gp, ce.x, inner, second ? ce.y : ce.x, ps, gs, edge_geo(b, e)[0],
		                          edge_geo(b, e)[1]);
	}
}

// residual of cell i: -F on c1's side, +F on c2's, ascending edge order
// (the scatter of fvm_tvd.cpp:353-363 seen from the cell); an edge's flux is
// one 32-byte record, two 16-byte loads
__device__ __forceinline__ void gather_residual(const Bufs& b, int i, double* r)
{
	r[0] = r[1] = r[2] = r[3] = 0.0;
	int4 fc = load_faces(b, i);
	int fs[3] = {fc.x, fc.y, fc.z};
	for (int k = 0; k < 3; k++) {
		int f = fs[k];
		double v[4];
		ld4(b.F, f >> 1, v);
		if (f & 1)
			for (int q = 0; q < 4; q++) r[q] += v[q];
		else
			for (int q = 0; q < 4; q++) r[q] -= v[q];
	}
}

__global__ __launch_bounds__(BLOCK) void k_residual(Bufs b)
{
	int i = xcd_block() * BLOCK + threadIdx.x;
	if (i >= b.nc) return;
	double r[4];
	gather_residual(b, i, r);
	for (int q = 0; q < 4; q++) b.R[(size_t)q * b.nc + i] = r[q];
}

enum { UPD_FIRST = 0, UPD_RK2_LAST = 2, UPD_RK3_MID = 3, UPD_RK3_LAST = 4 };

// The conserved state U and the stage-start copy U0 are touched once per
// update pass and not again until the next one.  When the context's arrays
// are well beyond the last-level (Infinity) cache, the update reads and writes them
// non-temporally (NT) so that they do not evict the primitives, fluxes and
// geometry that the next edge pass gathers; a case that fits the cache
// keeps them cached.  The values are the same either way.
template <bool NT>
__device__ __forceinline__ double ld_state(const double* p)
{
	if (NT) return __builtin_nontemporal_load(p);
	return *p;
}

template <bool NT>
__device__ __forceinline__ void st_state(double* p, double v)
{
	if (NT) __builtin_nontemporal_store(v, p);
	else *p = v;
}

// Update pass (fvm_tvd.cpp:366-374, 419-447): LIM-flagged cells stay frozen;
// the last stage forms the RK combination, refreshes the primitives and sets
// the bound flags (clamped values are not written back, 439-446).  The stage
// reads its input from b.U and the step-start state from b.U0 and writes
// b.Un; launch_update points them so that the step-start state is never
// copied (the reference's U_old copy, 318-321).  NT: see ld_state.  REC
// (adjoint recording behind an edge pass that forms its own gradients, where
// the stage's input primitives b.P and its output b.Po are separate slots):
// the pass first forms the gradient records of its input state, with the
// gradient pass's code, and stores them to b.G through LDS -- the edge pass
// formed the same values, but stored from the lane of each cell's first face
// they were single scattered records, slower in measurement.
template <bool MM, int MODE, bool NT, bool REC>
__global__ __launch_bounds__(BLOCK) void k_update(Bufs b, Phys ph)
{
	__shared__ double lds[BLOCK * (REC ? 9 : 5)];
	int lb = xcd_block();
	int i = lb * BLOCK + threadIdx.x;
	int flagged = 0;
	if (REC) {
		double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
		if (i < b.nc) {
			Prim pi;
			int face[3];
			Prim nb[3];
			if (b.lsqw) cell_gradient<MM, true>(b, ph, i, pi, face, nb, g);
			else cell_gradient<MM, false>(b, ph, i, pi, face, nb, g);
		}
		// read only by the reverse sweep, many stages later: kept out of the caches
		store_records<8, true>(b.G, lb * BLOCK, min(BLOCK, b.nc - lb * BLOCK), g, lds);
		__syncthreads();  // lds is reused below
	}
	double pv[4] = {0.0, 0.0, 0.0, 0.0};  // primitives written for this cell
	if (i < b.nc) {
		size_t n = b.nc;
		unsigned fl = b.flags[i];
		bool lim = (fl & 0x2u) != 0;
		double u[4], u0[4];
		for (int q = 0; q < 4; q++) u[q] = ld_state<NT>(&b.U[q * n + i]);
		if (MODE == UPD_FIRST && lim)  // a frozen cell's state still moves to the stage buffer
			for (int q = 0; q < 4; q++) st_state<NT>(&b.Un[q * n + i], u[q]);
		if (!lim) {
			double r[4];
			gather_residual(b, i, r);
			double cfl = b.dtS[i];
			if (MODE != UPD_FIRST)
				for (int q = 0; q < 4; q++) u0[q] = ld_state<NT>(&b.U0[q * n + i]);
			for (int q = 0; q < 4; q++) {
				double t = u[q] + cfl * r[q];
				if (MODE == UPD_FIRST) u[q] = t;
				else if (MODE == UPD_RK2_LAST) u[q] = 0.5 * (u0[q] + t);
				else if (MODE == UPD_RK3_MID) u[q] = 0.75 * u0[q] + 0.25 * t;
				else u[q] = (1.0 / 3.0) * u0[q] + (2.0 / 3.0) * t;
			}
			bool finite = true;
			for (int q = 0; q < 4; q++) {
				st_state<NT>(&b.Un[q * n + i], u[q]);
				finite = finite && isfinite(u[q]);
			}
			if (!finite) atomicAdd(&b.err[1], 1ull);
			MatV m = mat_of<MM>(b, ph, i);
			Prim p = cons_to_prim(u[0], u[1], u[2], u[3], m);
			pv[0] = p.r;
			pv[1] = p.p;
			pv[2] = p.u;
			pv[3] = p.v;
			if (MODE == UPD_RK2_LAST || MODE == UPD_RK3_LAST) {
				bool out = p.r < ph.ro_min || p.r > ph.ro_max || p.p < ph.p_min || p.p > ph.p_max ||
				           fabs(p.u) > ph.u_max || fabs(p.v) > ph.u_max;
				if (out) {
					fl |= 0x2u;
					b.flags[i] = fl;
				}
			}
		} else {
			ld4(b.P, i, pv);  // frozen cell: its primitives unchanged
		}
		flagged = (fl & 0x2u) != 0;
	}
	store_records<4>(b.Po, lb * BLOCK, min(BLOCK, b.nc - lb * BLOCK), pv, lds);
	if (MODE == UPD_RK2_LAST || MODE == UPD_RK3_LAST) {
		int any = __syncthreads_or(flagged);
		if (threadIdx.x == 0) {
			b.blockflag[lb] = any;
			if (any) {
				atomicOr(b.anyflag, 1);
				atomicOr(b.anyflag + 1, 1);
			}
		}
	}
}

// FVM_TVD::remediateLimCells (src/methods/fvm_tvd.cpp:464-499): each
// LIM-flagged cell, in ascending order, is replaced by the area-weighted mean
// of the c2 cells of its edges in edge-list order (which may be the cell
// itself) and its 32-step counter advances.  The sweep is sequential in the
// reference -- a later cell sees values remediated earlier -- and stays so
// here, but only where it must: one workgroup finds the 256-cell blocks that
// hold LIM cells (the update pass flags them); per such block all lanes
// gather their flagged cells' faces, c2 cells, areas and states at once into
// LDS (a block's earlier blocks are already stored), then one lane runs the
// sequential part over LDS alone, taking a c2 cell's new value from LDS when
// it is a flagged cell of the same block that comes earlier, and then all
// lanes store states, flags and primitives.  The same operations in the
// same order as the sequential sweep.
template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_remediate(Bufs b, Phys ph)
{
	__shared__ int pos[BLOCK];          // a lane's position among the block's flagged cells, -1 if none
	__shared__ int scan[BLOCK];
	__shared__ int bflag[BLOCK];        // the block flags of the current group of blocks
	__shared__ int nj[BLOCK][3];        // per position: c2 of each face (-1: boundary)
	__shared__ double ns[BLOCK][3];     // its area
	__shared__ double nu[BLOCK][3][4];  // its state when gathered
	__shared__ double nw[BLOCK][4];     // the remediated state
	size_t n = b.nc;
	int tid = threadIdx.x;
	if (*(volatile int*)b.anyflag == 0) return;  // no LIM cell anywhere this step
	for (int base = 0; base < b.nblocks; base += BLOCK) {
		int blk = base + tid;
		int mine = blk < b.nblocks ? b.blockflag[blk] : 0;
		bflag[tid] = mine;
		__syncthreads();
		if (!__syncthreads_or(mine)) continue;
		int lim = min(base + BLOCK, b.nblocks);
		for (int bb = base; bb < lim; bb++) {
			if (!bflag[bb - base]) continue;  // the same value in every lane
			int c0 = bb * BLOCK, i = c0 + tid;
			unsigned fl = i < b.nc ? b.flags[i] : 0u;
			int f = (fl & 0x2u) ? 1 : 0;
			// positions: inclusive prefix sum of the flags over the block
			scan[tid] = f;
			__syncthreads();
			for (int d = 1; d < BLOCK; d <<= 1) {
				int v = tid >= d ? scan[tid - d] : 0;
				__syncthreads();
				scan[tid] += v;
				__syncthreads();
			}
			int cnt = scan[BLOCK - 1], p = scan[tid] - 1;
			pos[tid] = f ? p : -1;
			if (f) {
				int4 fc = load_faces(b, i);
				int fs[3] = {fc.x, fc.y, fc.z};
				for (int k = 0; k < 3; k++) {
					int j = edge_cells(b, fs[k] >> 1).y;
					nj[p][k] = j;
					if (j >= 0) {
						ns[p][k] = b.S[j];
						for (int q = 0; q < 4; q++) nu[p][k][q] = b.U[q * n + j];
					}
				}
			}
			__syncthreads();
			if (tid == 0) {
				for (int r = 0; r < cnt; r++) {
					double sr[4] = {0.0, 0.0, 0.0, 0.0}, S = 0.0;
					for (int k = 0; k < 3; k++) {
						int j = nj[r][k];
						if (j < 0) continue;
						int jl = j - c0;  // a flagged cell of this block remediated before this one?
						int pj = jl >= 0 && jl < BLOCK ? pos[jl] : -1;
						const double* u = pj >= 0 && pj < r ? nw[pj] : nu[r][k];
						double s = ns[r][k];
						S += s;
						for (int q = 0; q < 4; q++) sr[q] += u[q] * s;
					}
					for (int q = 0; q < 4; q++) nw[r][q] = sr[q] / S;
				}
			}
			__syncthreads();
			if (f) {
				double u[4];
				for (int q = 0; q < 4; q++) {
					u[q] = nw[p][q];
					b.U[q * n + i] = u[q];
				}
				fl += 0x010000u;
				if (fl & 0x200000u) fl &= 0x001110u;
				b.flags[i] = fl;
				MatV m = mat_of<MM>(b, ph, i);
				store_prim(b, i, cons_to_prim(u[0], u[1], u[2], u[3], m));
			}
			__syncthreads();  // the next block gathers what this one stored
		}
	}
	if (tid == 0) *b.anyflag = 0;
}

// FVM_TVD::calcTimeStep (src/methods/fvm_tvd.cpp:216-240): local step
// CFL*S/max(|u|+c, |v|+c) with the cell area S (spectral variant:
// 2*CFL*S / sum_f lambda_f l_f); steady runs keep it per cell.
template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_tau(Bufs b, Phys ph)
{
	int i = xcd_block() * BLOCK + threadIdx.x;
	if (i >= b.nc) return;
	Prim p = load_prim(b, i);
	MatV m = mat_of<MM>(b, ph, i);
	double cz = sqrt(m.gam * p.p / p.r);
	double S = b.S[i];
	double loc;
	if (ph.cfl_mode == CFD2D_CFL_SPECTRAL) {
		double li = sqrt(p.u * p.u + p.v * p.v) + cz;
		double sum = 0.0;
		CellTop tp = load_top(b, i);
		for (int k = 0; k < 3; k++) {
			int f = tp.f[k];
			int e = f >> 1;
			int j = tp.nb[k];
			Prim q;
			double czj;
			if (j >= 0) {
				q = load_prim(b, j);
				MatV mj = mat_of<MM>(b, ph, j);
				czj = sqrt(mj.gam * q.p / q.r);
			} else {
				q = ghost_of<MM>(b, ph, e, i, p, p);
				czj = sqrt(m.gam * q.p / q.r);
			}
			double lj = sqrt(q.u * q.u + q.v * q.v) + czj;
			sum += 0.5 * (li + lj) * edge_geo(b, e)[2];
		}
		loc = 2.0 * ph.cfl * S / sum;
	} else {
		double a = fabs(p.u) + cz, c = fabs(p.v) + cz;
		loc = ph.cfl * S / (a > c ? a : c);
	}
	b.loc[i] = loc;
	if (ph.steady) b.dtS[i] = loc / S;
}

}  // namespace

// ======================================================================== context

struct cfd2d_ctx {
	int nc = 0, ne = 0, ncase = 1, nbc = 0;
	bool mm = false;
	cfd2d_params p{};
	Phys ph{};
	Bufs b{};
	hipStream_t st = nullptr;
	std::vector<double> hS;
	std::vector<int> hc1;  // edge c1, length, normal (host copies for the force integrals)
	std::vector<double> hel, hnx, hny;
	std::vector<int> case_off;
	std::vector<double> case_tau;
	// timing
	bool timing = false;
	unsigned timing_mask = ~0u;  // kernel ids (bit CFD2D_K_*) that get events
	int fuse_mode = -1;          // 1: the edge pass forms the gradients, 2: the tiled edge
	                             // pass does, 0: a separate gradient pass, -1: 1 without
	                             // a limiter, 0 with one
	bool prefer_graph = false;   // cfd2d_step replays a recorded step (chosen by cfd2d_autotune)
	int edge_shape = 0;          // launch shape of the edge pass (EdgeShape; chosen by cfd2d_autotune)
	int stream_mode = -1;        // update pass reads/writes the state non-temporally: 1 yes,
	                             // 0 no, -1 when the arrays exceed twice the Infinity Cache
	bool stream_auto = false;    // that size test, made once the arrays are allocated
	size_t dev_bytes = 0;        // bytes of the arrays allocated by dalloc
	bool ev_open = false;        // rec_start recorded a start event
	std::vector<hipEvent_t> ev0, ev1;
	std::vector<int> ev_kernel;
	int ev_used = 0;
	double tot_ms[CFD2D_K_COUNT] = {0};
	long long launches[CFD2D_K_COUNT] = {0};
	// recorded runs of steps (cfd2d_step_graph), by step count
	struct StepGraph {
		int steps;
		hipGraph_t g;
		hipGraphExec_t x;
	};
	std::vector<StepGraph> graphs;
	// the adjoint's recorded forward and reverse passes (adjoint.inl)
	struct AdjGraph {
		std::vector<long long> key;
		hipGraph_t g = nullptr;
		hipGraphExec_t x = nullptr;
	};
	AdjGraph adj_fwd, adj_rev;
	bool adj_use_graph = true;
	bool capturing = false;
	std::vector<void*> allocs;
	// adjoint topology, built on the first cfd2d_adjoint call
	int* adj_cslot = nullptr;
	int* adj_bedges = nullptr;  // grouped by boundary object, ascending within one
	int* adj_bcoff = nullptr;   // [nbc + 1] start of each object's run in adj_bedges
	int adj_nbedges = -1;
	// adjoint arena: checkpoints, stage records and work arrays, kept between
	// calls and grown on demand
	void* adj_ws = nullptr;
	size_t adj_ws_bytes = 0;
	std::vector<double> adj_Jcase;  // objective per case of the last adjoint call
	// the geometry adjoint's device copy of the mesh and work arrays, for the
	// mesh with id geo_id (0: none yet)
	void* geo_ws = nullptr;
	size_t geo_ws_bytes = 0;
	unsigned long long geo_id = 0;
	// implicit assembly output (blocks, rhs, ctau, cols), grown on demand
	void* impl_ws = nullptr;
	size_t impl_ws_bytes = 0;
};

namespace {

constexpr int EV_POOL = 8192;
constexpr size_t LLC_BYTES = (size_t)256 << 20;  // MI355X Infinity Cache (last-level, shared by the XCDs)

int hip_fail(hipError_t e, const char* what)
{
	return fail(CFD2D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(x)                                      \
	do {                                               \
		hipError_t _e = (x);                           \
		if (_e != hipSuccess) return hip_fail(_e, #x); \
	} while (0)

template <typename T>
int dalloc(cfd2d_ctx* c, T** p, size_t count)
{
	if (count == 0) count = 1;
	void* q = nullptr;
	hipError_t e = hipMalloc(&q, count * sizeof(T));
	if (e != hipSuccess) return hip_fail(e, "hipMalloc");
	c->allocs.push_back(q);
	c->dev_bytes += count * sizeof(T);
	*p = (T*)q;
	return 0;
}

template <typename T>
int upload(cfd2d_ctx* c, T** p, const std::vector<T>& v)
{
	int rc = dalloc(c, p, v.size());
	if (rc) return rc;
	if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
	return 0;
}

int grid_of(int n) { return (n + BLOCK - 1) / BLOCK; }

void rec_start(cfd2d_ctx* c, int k)
{
	c->ev_open = false;
	if (!c->timing || c->capturing || c->ev_used >= EV_POOL || !((c->timing_mask >> k) & 1u)) return;
	if ((int)c->ev0.size() <= c->ev_used) {
		hipEvent_t a, z;
		if (hipEventCreate(&a) != hipSuccess) return;
		if (hipEventCreate(&z) != hipSuccess) {
			(void)hipEventDestroy(a);
			return;
		}
		c->ev0.push_back(a);
		c->ev1.push_back(z);
		c->ev_kernel.push_back(k);
	}
	c->ev_kernel[c->ev_used] = k;
	(void)hipEventRecord(c->ev0[c->ev_used], c->st);
	c->ev_open = true;
}

void rec_stop(cfd2d_ctx* c)
{
	if (!c->ev_open) return;
	c->ev_open = false;
	if (!c->timing || c->capturing || c->ev_used >= EV_POOL || c->ev_used >= (int)c->ev0.size()) return;
	(void)hipEventRecord(c->ev1[c->ev_used], c->st);
	c->ev_used++;
}

void harvest(cfd2d_ctx* c)
{
	for (int k = 0; k < c->ev_used; k++) {
		float ms = 0.0f;
		if (hipEventElapsedTime(&ms, c->ev0[k], c->ev1[k]) == hipSuccess) {
			c->tot_ms[c->ev_kernel[k]] += ms;
			c->launches[c->ev_kernel[k]]++;
		}
	}
	c->ev_used = 0;
}

#define LAUNCH(c, kid, kern, n, ...)                                                        \
	do {                                                                                    \
		rec_start(c, kid);                                                                  \
		hipLaunchKernelGGL(kern, dim3(grid_of(n)), dim3(BLOCK), 0, (c)->st, __VA_ARGS__);   \
		rec_stop(c);                                                                        \
	} while (0)

template <bool MM>
void launch_prim(cfd2d_ctx* c)
{
	LAUNCH(c, CFD2D_K_PRIM, (k_prim<MM>), c->nc, c->b, c->ph);
}

template <bool MM>
void launch_gradient(cfd2d_ctx* c)
{
	bool lsq = c->p.gradient == CFD2D_GRAD_LSQ, lim = c->p.limiter == CFD2D_LIMITER_MINMOD;
	if (lsq && lim) LAUNCH(c, CFD2D_K_GRADIENT, (k_gradient<MM, true, true>), c->nc, c->b, c->ph);
	else if (lsq) LAUNCH(c, CFD2D_K_GRADIENT, (k_gradient<MM, true, false>), c->nc, c->b, c->ph);
	else if (lim) LAUNCH(c, CFD2D_K_GRADIENT, (k_gradient<MM, false, true>), c->nc, c->b, c->ph);
	else LAUNCH(c, CFD2D_K_GRADIENT, (k_gradient<MM, false, false>), c->nc, c->b, c->ph);
}

template <bool MM, int FLUX, bool VISC, int GR, int SHAPE>
void launch_edge_s(cfd2d_ctx* c)
{
	constexpr int NT = EdgeShape<SHAPE>::NT;
	int n = 2 * c->ne;
	rec_start(c, CFD2D_K_EDGE_FLUX);
	hipLaunchKernelGGL((k_edge_flux<MM, FLUX, VISC, GR, SHAPE>), dim3((n + NT - 1) / NT), dim3(NT), 0, c->st, c->b,
	                   c->ph);
	rec_stop(c);
}

// the launch shapes are instantiated for single-material contexts and the
// unlimited or stored-gradient edge passes (the variants the autotuner
// times); the others launch shape 0
template <bool MM, int FLUX, bool VISC, int GR>
void launch_edge_f(cfd2d_ctx* c)
{
	constexpr bool shapes = !MM && GR <= 2;
	switch (shapes ? c->edge_shape : 0) {
	case 1: launch_edge_s<MM, FLUX, VISC, GR, shapes ? 1 : 0>(c); break;
	case 2: launch_edge_s<MM, FLUX, VISC, GR, shapes ? 2 : 0>(c); break;
	case 3: launch_edge_s<MM, FLUX, VISC, GR, shapes ? 3 : 0>(c); break;
	default: launch_edge_s<MM, FLUX, VISC, GR, 0>(c); break;
	}
}

template <bool MM, bool VISC, int GR>
void launch_edge_v(cfd2d_ctx* c)
{
	switch (c->p.flux) {
	case CFD2D_FLUX_LAX: launch_edge_f<MM, CFD2D_FLUX_LAX, VISC, GR>(c); break;
	case CFD2D_FLUX_HLLC: launch_edge_f<MM, CFD2D_FLUX_HLLC, VISC, GR>(c); break;
	case CFD2D_FLUX_ROE: launch_edge_f<MM, CFD2D_FLUX_ROE, VISC, GR>(c); break;
	default: launch_edge_f<MM, CFD2D_FLUX_GODUNOV, VISC, GR>(c); break;
	}
}

// the edge pass reading the stored gradients
template <bool MM>
void launch_edge(cfd2d_ctx* c)
{
	if (c->p.viscous) launch_edge_v<MM, true, 0>(c);
	else launch_edge_v<MM, false, 0>(c);
}

// gradient + edge passes of one stage: by default the edge pass forms the
// gradients itself without a limiter (nothing reads the stored gradients
// afterwards) and a separate pass forms them with one (three limiter
// evaluations per cell cost more than the pass they save on large meshes);
// cfd2d_fuse_gradient and cfd2d_autotune change the choice
template <bool MM, bool VISC, int GR>
void launch_edge_tile(cfd2d_ctx* c)
{
	dim3 g(c->b.ntiles), blk(BLOCK);
	rec_start(c, CFD2D_K_EDGE_FLUX);
	switch (c->p.flux) {
	case CFD2D_FLUX_LAX: hipLaunchKernelGGL((k_edge_tile<MM, CFD2D_FLUX_LAX, VISC, GR>), g, blk, 0, c->st, c->b, c->ph); break;
	case CFD2D_FLUX_HLLC: hipLaunchKernelGGL((k_edge_tile<MM, CFD2D_FLUX_HLLC, VISC, GR>), g, blk, 0, c->st, c->b, c->ph); break;
	case CFD2D_FLUX_ROE: hipLaunchKernelGGL((k_edge_tile<MM, CFD2D_FLUX_ROE, VISC, GR>), g, blk, 0, c->st, c->b, c->ph); break;
	default: hipLaunchKernelGGL((k_edge_tile<MM, CFD2D_FLUX_GODUNOV, VISC, GR>), g, blk, 0, c->st, c->b, c->ph); break;
	}
	rec_stop(c);
}

template <bool MM, bool VISC>
void launch_edge_tiled(cfd2d_ctx* c, int gr)
{
	switch (gr) {
	case 1: launch_edge_tile<MM, VISC, 1>(c); break;
	case 2: launch_edge_tile<MM, VISC, 2>(c); break;
	case 3: launch_edge_tile<MM, VISC, 3>(c); break;
	default: launch_edge_tile<MM, VISC, 4>(c); break;
	}
}

template <bool MM>
void launch_gradient_edge(cfd2d_ctx* c)
{
	bool lim = c->p.limiter != CFD2D_LIMITER_NONE;
	bool fuse = c->fuse_mode < 0 ? !lim : c->fuse_mode != 0;
	if (!fuse) {
		launch_gradient<MM>(c);
		launch_edge<MM>(c);
		return;
	}
	bool lsq = c->p.gradient == CFD2D_GRAD_LSQ;
	int gr = (lsq ? 1 : 2) + (lim ? 2 : 0);
	if (c->fuse_mode == 2 && c->b.ntiles > 0) {
		if (c->p.viscous) launch_edge_tiled<MM, true>(c, gr);
		else launch_edge_tiled<MM, false>(c, gr);
		return;
	}
	if (c->p.viscous) {
		switch (gr) {
		case 1: launch_edge_v<MM, true, 1>(c); break;
		case 2: launch_edge_v<MM, true, 2>(c); break;
		case 3: launch_edge_v<MM, true, 3>(c); break;
		default: launch_edge_v<MM, true, 4>(c); break;
		}
	} else {
		switch (gr) {
		case 1: launch_edge_v<MM, false, 1>(c); break;
		case 2: launch_edge_v<MM, false, 2>(c); break;
		case 3: launch_edge_v<MM, false, 3>(c); break;
		default: launch_edge_v<MM, false, 4>(c); break;
		}
	}
}

template <bool MM, bool NT, bool REC>
void launch_update_nt(cfd2d_ctx* c, int mode, const Bufs& b)
{
	switch (mode) {
	case UPD_FIRST: LAUNCH(c, CFD2D_K_UPDATE, (k_update<MM, UPD_FIRST, NT, REC>), c->nc, b, c->ph); break;
	case UPD_RK2_LAST: LAUNCH(c, CFD2D_K_UPDATE, (k_update<MM, UPD_RK2_LAST, NT, REC>), c->nc, b, c->ph); break;
	case UPD_RK3_MID: LAUNCH(c, CFD2D_K_UPDATE, (k_update<MM, UPD_RK3_MID, NT, REC>), c->nc, b, c->ph); break;
	default: LAUNCH(c, CFD2D_K_UPDATE, (k_update<MM, UPD_RK3_LAST, NT, REC>), c->nc, b, c->ph); break;
	}
}

// the state streams past the caches (see ld_state) when the context's
// arrays are over twice the Infinity Cache, or as cfd2d_stream_state and
// cfd2d_autotune set it
bool stream_state(const cfd2d_ctx* c) { return c->stream_mode < 0 ? c->stream_auto : c->stream_mode != 0; }

// The state of a step lives in U; U0 is the stage buffer.  The first stage
// reads U and writes U0, so U keeps the step-start state; later stages read
// U0 (their input) and U (the step start) and write U0, the last one U:
// every stage reads and writes each cell's values in one lane, so the
// in-place stages are safe, and no step-start copy is made.
template <bool MM>
void launch_update(cfd2d_ctx* c, int mode)
{
	Bufs b = c->b;
	double *state = c->b.U, *stage = c->b.U0;
	bool last = mode == UPD_RK2_LAST || mode == UPD_RK3_LAST;
	b.U = mode == UPD_FIRST ? state : stage;
	b.U0 = state;
	b.Un = last ? state : stage;
	if (c->ph.store_g) launch_update_nt<MM, false, true>(c, mode, b);  // adjoint recording
	else if (stream_state(c)) launch_update_nt<MM, true, false>(c, mode, b);
	else launch_update_nt<MM, false, false>(c, mode, b);
}

template <bool MM>
void enqueue_step(cfd2d_ctx* c)
{
	if (c->p.steady) LAUNCH(c, CFD2D_K_TAU, (k_tau<MM>), c->nc, c->b, c->ph);
	if (c->p.rk_stages == 3) {
		const int modes[3] = {UPD_FIRST, UPD_RK3_MID, UPD_RK3_LAST};
		for (int s = 0; s < 3; s++) {
			launch_gradient_edge<MM>(c);
			launch_update<MM>(c, modes[s]);
		}
	} else {
		const int modes[2] = {UPD_FIRST, UPD_RK2_LAST};
		for (int s = 0; s < 2; s++) {
			launch_gradient_edge<MM>(c);
			launch_update<MM>(c, modes[s]);
		}
	}
	rec_start(c, CFD2D_K_REMEDIATE);
	hipLaunchKernelGGL((k_remediate<MM>), dim3(1), dim3(BLOCK), 0, c->st, c->b, c->ph);
	rec_stop(c);
}

void step_once(cfd2d_ctx* c)
{
	if (c->mm) enqueue_step<true>(c);
	else enqueue_step<false>(c);
}

void do_prim(cfd2d_ctx* c)
{
	if (c->mm) launch_prim<true>(c);
	else launch_prim<false>(c);
}

void do_gradient(cfd2d_ctx* c)
{
	if (c->mm) launch_gradient<true>(c);
	else launch_gradient<false>(c);
}

void do_edge(cfd2d_ctx* c)
{
	if (c->mm) launch_edge<true>(c);
	else launch_edge<false>(c);
}

int check_launch(const char* what)
{
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return hip_fail(e, what);
	return 0;
}

// least-squares weights w_k = A^{-1} d_k, A = sum d d^T (the oracle's lsq_weights)
// Tiles of the tiled edge pass (k_edge_tile): tiles of T consecutive cells,
// each owning the edges whose c1 it holds, staging c1 and (inner edges) c2
// of those edges.  False when the edges are not in c1 order or a tile would
// stage more than TILE_CAP cells.
bool build_tiles(const std::vector<int>& c1, const std::vector<int>& c2, int nc, int T, std::vector<int>& tiles,
                 std::vector<int>& tcell, std::vector<unsigned>& eslot)
{
	int ne = (int)c1.size();
	for (int e = 1; e < ne; e++)
		if (c1[e] < c1[e - 1]) return false;
	int nt = (nc + T - 1) / T;
	tiles.assign(4 * (size_t)nt, 0);
	tcell.clear();
	eslot.assign(ne, 0u);
	std::vector<int> need;
	int e = 0;
	for (int t = 0; t < nt; t++) {
		int e0 = e;
		while (e < ne && c1[e] < (t + 1) * T) e++;
		need.clear();
		for (int k = e0; k < e; k++) {
			need.push_back(c1[k]);
			if (c2[k] >= 0) need.push_back(c2[k]);
		}
		std::sort(need.begin(), need.end());
		need.erase(std::unique(need.begin(), need.end()), need.end());
		if ((int)need.size() > TILE_CAP) return false;
		tiles[4 * (size_t)t] = (int)tcell.size();
		tiles[4 * (size_t)t + 1] = (int)need.size();
		tiles[4 * (size_t)t + 2] = e0;
		tiles[4 * (size_t)t + 3] = e - e0;
		auto slot = [&](int cell) { return (unsigned)(std::lower_bound(need.begin(), need.end(), cell) - need.begin()); };
		for (int k = e0; k < e; k++) eslot[k] = slot(c1[k]) | ((c2[k] >= 0 ? slot(c2[k]) : 0xffffu) << 16);
		tcell.insert(tcell.end(), need.begin(), need.end());
	}
	if (tcell.empty()) tcell.push_back(0);
	return e == ne;
}

void lsq_weights(const Mesh& m, std::vector<double>& w)
{
	int nc = m.nc;
	w.assign(6 * (size_t)nc, 0.0);
	for (int i = 0; i < nc; i++) {
		double d[3][2];
		for (int k = 0; k < 3; k++) {
			int e = m.cedges[3 * (size_t)i + k];
			int c1 = m.c1[e], c2 = m.c2[e];
			if (c2 < 0) {
				double mx = m.gp[6 * (size_t)e], my = m.gp[6 * (size_t)e + 1];
				double dn = (mx - m.cx[i]) * m.nx[e] + (my - m.cy[i]) * m.ny[e];
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
		for (int k = 0; k < 3; k++) {
			w[6 * (size_t)i + 2 * k] = (a22 * d[k][0] - a12 * d[k][1]) / det;
			w[6 * (size_t)i + 2 * k + 1] = (a11 * d[k][1] - a12 * d[k][0]) / det;
		}
	}
}

void drop_adj_graph(cfd2d_ctx::AdjGraph& a)
{
	if (a.x) (void)hipGraphExecDestroy(a.x);
	if (a.g) (void)hipGraphDestroy(a.g);
	a.x = nullptr;
	a.g = nullptr;
	a.key.clear();
}

// the recorded runs hold the step variant and buffers of their recording
void drop_graph(cfd2d_ctx* c)
{
	for (auto& g : c->graphs) {
		(void)hipGraphExecDestroy(g.x);
		(void)hipGraphDestroy(g.g);
	}
	c->graphs.clear();
	drop_adj_graph(c->adj_fwd);
	drop_adj_graph(c->adj_rev);
}

void destroy_ctx(cfd2d_ctx* c)
{
	if (!c) return;
	if (c->st) (void)hipStreamSynchronize(c->st);
	drop_graph(c);
	for (auto e : c->ev0) (void)hipEventDestroy(e);
	for (auto e : c->ev1) (void)hipEventDestroy(e);
	for (void* p : c->allocs) (void)hipFree(p);
	if (c->adj_ws) (void)hipFree(c->adj_ws);
	if (c->impl_ws) (void)hipFree(c->impl_ws);
	if (c->geo_ws) (void)hipFree(c->geo_ws);
	if (c->st) (void)hipStreamDestroy(c->st);
	delete c;
}

}  // namespace

extern "C" {

int cfd2d_device_count(int* n)
{
	if (!n) return fail(CFD2D_ERR_ARG, "null");
	hipError_t e = hipGetDeviceCount(n);
	if (e != hipSuccess) {
		*n = 0;
		return hip_fail(e, "hipGetDeviceCount");
	}
	return CFD2D_OK;
}

int cfd2d_ctx_create(const cfd2d_mesh* mesh, const cfd2d_params* p, int n_bc, const int* bc_type,
                     const double* bc_par, const int* group_bc, const int* cell_mat, cfd2d_ctx** out)
{
	if (!mesh || !p || !out || n_bc < 0 || (n_bc > 0 && (!bc_type || !bc_par)))
		return fail(CFD2D_ERR_ARG, "cfd2d_ctx_create: bad arguments");
	const Mesh& m = mesh->m;
	if (p->nmat < 1 || p->nmat > CFD2D_MAXMAT) return fail(CFD2D_ERR_CONFIG, "materials count must be 1..8");
	if (p->rk_stages != 2 && p->rk_stages != 3) return fail(CFD2D_ERR_CONFIG, "rk_stages must be 2 or 3");
	if (p->flux < 0 || p->flux > 3 || p->gradient < 0 || p->gradient > 1 || p->limiter < 0 || p->limiter > 1)
		return fail(CFD2D_ERR_CONFIG, "unknown flux/gradient/limiter");
	for (int k = 0; k < n_bc; k++)
		if (bc_type[k] < CFD2D_BND_INLET || bc_type[k] > CFD2D_BND_PRESSURE_OUTLET)
			return fail(CFD2D_ERR_CONFIG, "unknown boundary type");
	int ngroups = (int)m.edge_groups.size();
	// every boundary edge needs a boundary object (fvm_tvd.cpp:132-174)
	std::vector<int> ec2(m.ne);
	for (int e = 0; e < m.ne; e++) {
		if (m.c2[e] >= 0) {
			ec2[e] = m.c2[e];
			continue;
		}
		int g = m.egroup[e];
		int bc = (group_bc && g >= 0 && (ngroups == 0 || g < ngroups)) ? group_bc[g] : -1;
		if (bc < 0 || bc >= n_bc) {
			char buf[96];
			snprintf(buf, sizeof(buf), "ERROR (boundary condition): unknown edge type of edge %d...", e);
			return fail(CFD2D_ERR_BOUNDARY, buf);
		}
		ec2[e] = -(bc + 1);
	}
	std::vector<int> cmat(m.nc, 0);
	bool mm = false;
	if (cell_mat) {
		for (int i = 0; i < m.nc; i++) {
			if (cell_mat[i] < 0 || cell_mat[i] >= p->nmat) return fail(CFD2D_ERR_CONFIG, "cell material out of range");
			cmat[i] = cell_mat[i];
			if (cmat[i] != 0) mm = true;
		}
	}
	auto* c = new (std::nothrow) cfd2d_ctx;
	if (!c) return fail(CFD2D_ERR_ARG, "out of host memory");
	c->nc = m.nc;
	c->ne = m.ne;
	c->nbc = n_bc;
	c->p = *p;
	c->mm = mm;
	c->ncase = (int)m.case_cell_offset.size() - 1;
	c->case_off = m.case_cell_offset;
	c->case_tau.assign(c->ncase, p->tau);
	c->hS = m.S;
	c->hc1 = m.c1;
	c->hel = m.l;
	c->hnx = m.nx;
	c->hny = m.ny;
	std::vector<double> mtab(5 * (size_t)p->nmat);
	for (int k = 0; k < p->nmat; k++) {
		double Cv = p->Cp[k] - GAS_R / p->M[k];
		double gam = p->Cp[k] / Cv;
		double kc = p->K[k] > 0.0 ? p->K[k] : p->ML[k] * p->Cp[k] / (p->prandtl > 0.0 ? p->prandtl : 0.72);
		mtab[5 * k] = p->M[k];
		mtab[5 * k + 1] = Cv;
		mtab[5 * k + 2] = gam;
		mtab[5 * k + 3] = p->ML[k];
		mtab[5 * k + 4] = kc;
	}
	c->ph.m0 = MatV{mtab[0], mtab[1], mtab[2], mtab[3], mtab[4]};
	c->ph.gam_flux = p->gam_flux;
	c->ph.igm1 = 1.0 / (p->gam_flux - 1.0);
	c->ph.rk = rim_const(p->gam_flux);
	c->ph.cfl = p->cfl;
	c->ph.ro_min = p->ro_min;
	c->ph.ro_max = p->ro_max;
	c->ph.p_min = p->p_min;
	c->ph.p_max = p->p_max;
	c->ph.u_max = p->u_max;
	// the device Newton loop always ends: no cap given means the library's
	c->ph.max_newton = p->max_newton > 0 ? p->max_newton : CFD2D_NEWTON_CAP;
	c->ph.cfl_mode = p->cfl_mode;
	c->ph.steady = p->steady;

	int nc = m.nc, ne = m.ne;
	std::vector<int> cface(4 * (size_t)nc, -1), cnb(4 * (size_t)nc, -1);
	std::vector<double> cfn(6 * (size_t)nc);
	for (int i = 0; i < nc; i++)
		for (int k = 0; k < 3; k++) {
			int e = m.cedges[3 * (size_t)i + k];
			int side = (m.c1[e] == i) ? 0 : 1;
			cface[4 * (size_t)i + k] = 2 * e + side;
			cnb[4 * (size_t)i + k] = side == 0 ? m.c2[e] : m.c1[e];
			double sg = side == 0 ? 1.0 : -1.0;
			cfn[6 * (size_t)i + 2 * k] = sg * (m.nx[e] * m.l[e]);
			cfn[6 * (size_t)i + 2 * k + 1] = sg * (m.ny[e] * m.l[e]);
		}
	std::vector<double> eg(4 * (size_t)ne, 0.0);
	for (int e = 0; e < ne; e++) {
		eg[4 * (size_t)e] = m.nx[e];
		eg[4 * (size_t)e + 1] = m.ny[e];
		eg[4 * (size_t)e + 2] = m.l[e];
		int cells[2] = {m.c1[e], ec2[e]};
		memcpy(&eg[4 * (size_t)e + 3], cells, sizeof(cells));
	}
	std::vector<double> edl(8 * (size_t)ne, 0.0);
	for (int e = 0; e < ne; e++) {
		int c1 = m.c1[e], c2 = m.c2[e];
		for (int g = 0; g < 2; g++) {
			double px = m.gp[6 * (size_t)e + 2 + 2 * g], py = m.gp[6 * (size_t)e + 3 + 2 * g];
			edl[edl_index(e, 0, g)] = px - m.cx[c1];
			edl[edl_index(e, 0, g) + 1] = py - m.cy[c1];
			if (c2 >= 0) {
				edl[edl_index(e, 1, g)] = px - (m.cx[c2] + m.shx[e]);
				edl[edl_index(e, 1, g) + 1] = py - (m.cy[c2] + m.shy[e]);
			} else if (g == 1) {  // the second lane of a boundary edge reconstructs c1
				edl[8 * (size_t)e + 4] = px - m.cx[c1];
				edl[8 * (size_t)e + 5] = py - m.cy[c1];
			}
		}
	}
	std::vector<double> bcp(4 * (size_t)(n_bc > 0 ? n_bc : 1), 0.0);
	std::vector<int> bct(n_bc > 0 ? n_bc : 1, 0);
	for (int k = 0; k < n_bc; k++) {
		bct[k] = bc_type[k];
		for (int q = 0; q < 4; q++) bcp[4 * k + q] = bc_par[4 * k + q];
	}
	Bufs& b = c->b;
	b.nc = nc;
	b.ne = ne;
	b.nblocks = grid_of(nc);
	b.nmat = p->nmat;
	int rc = 0;
	size_t n4 = 4 * (size_t)nc;
	if ((rc = dalloc(c, &b.U, n4)) || (rc = dalloc(c, &b.U0, n4)) || (rc = dalloc(c, &b.P, n4)) ||
	    (rc = dalloc(c, &b.G, 2 * n4)) || (rc = dalloc(c, &b.dtS, nc)) || (rc = upload(c, &b.S, m.S)) ||
	    (rc = dalloc(c, &b.loc, nc)) || (rc = dalloc(c, &b.R, n4)) || (rc = upload(c, &b.cface, cface)) ||
	    (rc = upload(c, &b.cnb, cnb)) || (rc = upload(c, &b.cfn, cfn)) || (rc = upload(c, &b.cmat, cmat)) || (rc = dalloc(c, &b.flags, nc)) ||
	    (rc = dalloc(c, &b.blockflag, b.nblocks)) || (rc = dalloc(c, &b.anyflag, 2)) || (rc = upload(c, &b.eg, eg)) || (rc = upload(c, &b.edl, edl)) || (rc = dalloc(c, &b.F, 4 * (size_t)ne)) ||
	    (rc = upload(c, &b.bctype, bct)) || (rc = upload(c, &b.bcpar, bcp)) || (rc = upload(c, &b.mtab, mtab)) ||
	    (rc = dalloc(c, &b.err, 2))) {
		destroy_ctx(c);
		return rc;
	}
	if (p->limiter != CFD2D_LIMITER_NONE) {
		// each cell's offsets to the Gauss points of its faces, faces in
		// ascending edge order (the limiter's order), from edl's own side
		std::vector<double> cgp(12 * (size_t)nc);
		for (int i = 0; i < nc; i++)
			for (int k = 0; k < 3; k++) {
				int f = cface[4 * (size_t)i + k];
				int e = f >> 1, side = f & 1;
				for (int g = 0; g < 2; g++) {
					cgp[(size_t)(2 * (2 * k + g)) * nc + i] = edl[edl_index(e, side, g)];
					cgp[(size_t)(2 * (2 * k + g) + 1) * nc + i] = edl[edl_index(e, side, g) + 1];
				}
			}
		if ((rc = upload(c, &b.cgp, cgp))) {
			destroy_ctx(c);
			return rc;
		}
	} else {
		b.cgp = nullptr;
	}
	if (p->gradient == CFD2D_GRAD_LSQ) {
		std::vector<double> w;
		lsq_weights(m, w);
		if ((rc = upload(c, &b.lsqw, w))) {
			destroy_ctx(c);
			return rc;
		}
	} else {
		b.lsqw = nullptr;
	}
	{
		// the tiled edge pass: the largest tile whose staged set fits
		std::vector<int> tiles, tcell;
		std::vector<unsigned> eslot;
		b.ntiles = 0;
		for (int T : {TILE_CELLS, 192, 128, 96, 64})
			if (build_tiles(m.c1, ec2, nc, T, tiles, tcell, eslot)) {
				if ((rc = upload(c, &b.tiles, tiles)) || (rc = upload(c, &b.tcell, tcell)) ||
				    (rc = upload(c, &b.eslot, eslot))) {
					destroy_ctx(c);
					return rc;
				}
				b.ntiles = (int)(tiles.size() / 4);
				break;
			}
	}
	hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
	if (e != hipSuccess) {
		destroy_ctx(c);
		return hip_fail(e, "hipStreamCreate");
	}
	b.Po = b.P;
	std::vector<unsigned> zf(nc, 0u);
	std::vector<int> zb(b.nblocks, 0);
	int zany[2] = {0, 0};
	unsigned long long zerr[2] = {0, 0};
	if (hipMemcpy(b.flags, zf.data(), nc * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b.blockflag, zb.data(), b.nblocks * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b.anyflag, zany, sizeof(zany), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(b.err, zerr, sizeof(zerr), hipMemcpyHostToDevice) != hipSuccess) {
		destroy_ctx(c);
		return fail(CFD2D_ERR_HIP, "initial upload failed");
	}
	// measured on the periodic NS case: streaming loses 5 % with 272 MB of
	// arrays, gains 1.4 % with 544 MB and 2.5 % with 1.1 GB
	c->stream_auto = c->dev_bytes > 2 * LLC_BYTES;
	*out = c;
	return CFD2D_OK;
}

void cfd2d_ctx_destroy(cfd2d_ctx* c) { destroy_ctx(c); }

int cfd2d_set_state(cfd2d_ctx* c, const double* u)
{
	if (!c || !u) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(c->b.U, u, 4 * (size_t)c->nc * sizeof(double), hipMemcpyHostToDevice, c->st));
	do_prim(c);
	HIPCHK(hipStreamSynchronize(c->st));
	return check_launch("k_prim");
}

int cfd2d_set_state_device(cfd2d_ctx* c, const double* du)
{
	if (!c || !du) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(c->b.U, du, 4 * (size_t)c->nc * sizeof(double), hipMemcpyDeviceToDevice, c->st));
	do_prim(c);
	HIPCHK(hipStreamSynchronize(c->st));
	return check_launch("k_prim");
}

int cfd2d_get_state(cfd2d_ctx* c, double* u)
{
	if (!c || !u) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(u, c->b.U, 4 * (size_t)c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_get_state_device(cfd2d_ctx* c, double* du)
{
	if (!c || !du) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(du, c->b.U, 4 * (size_t)c->nc * sizeof(double), hipMemcpyDeviceToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_set_flags(cfd2d_ctx* c, const unsigned* f)
{
	if (!c || !f) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(c->b.flags, f, (size_t)c->nc * sizeof(unsigned), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_get_flags(cfd2d_ctx* c, unsigned* f)
{
	if (!c || !f) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipMemcpyAsync(f, c->b.flags, (size_t)c->nc * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

// Unsteady: TAU_case = min(TAU, min over the case's cells of the local step),
// every cell of the case gets TAU_case (fvm_tvd.cpp:226-237).  Steady: local.
int cfd2d_calc_tau(cfd2d_ctx* c)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	if (c->mm) LAUNCH(c, CFD2D_K_TAU, (k_tau<true>), c->nc, c->b, c->ph);
	else LAUNCH(c, CFD2D_K_TAU, (k_tau<false>), c->nc, c->b, c->ph);
	int rc = check_launch("k_tau");
	if (rc) return rc;
	if (c->p.steady) {
		HIPCHK(hipStreamSynchronize(c->st));
		return CFD2D_OK;
	}
	std::vector<double> loc(c->nc);
	HIPCHK(hipMemcpyAsync(loc.data(), c->b.loc, c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	std::vector<double> dts(c->nc);
	for (int k = 0; k < c->ncase; k++) {
		double T = c->case_tau[k];
		for (int i = c->case_off[k]; i < c->case_off[k + 1]; i++)
			if (T > loc[i]) T = loc[i];
		c->case_tau[k] = T;
		for (int i = c->case_off[k]; i < c->case_off[k + 1]; i++) dts[i] = T / c->hS[i];
	}
	HIPCHK(hipMemcpyAsync(c->b.dtS, dts.data(), c->nc * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_get_tau(cfd2d_ctx* c, double* tau0, double* ctau)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	if (tau0) *tau0 = c->case_tau[0];
	if (ctau) {
		std::vector<double> dts(c->nc);
		HIPCHK(hipMemcpyAsync(dts.data(), c->b.dtS, c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
		HIPCHK(hipStreamSynchronize(c->st));
		if (c->p.steady) {
			std::vector<double> loc(c->nc);
			HIPCHK(hipMemcpy(loc.data(), c->b.loc, c->nc * sizeof(double), hipMemcpyDeviceToHost));
			for (int i = 0; i < c->nc; i++) ctau[i] = loc[i];
		} else {
			for (int k = 0; k < c->ncase; k++)
				for (int i = c->case_off[k]; i < c->case_off[k + 1]; i++) ctau[i] = c->case_tau[k];
		}
	}
	return CFD2D_OK;
}

int cfd2d_set_bc_par(cfd2d_ctx* c, const double* bc_par)
{
	if (!c || (c->nbc > 0 && !bc_par)) return fail(CFD2D_ERR_ARG, "cfd2d_set_bc_par: bad arguments");
	if (c->nbc == 0) return CFD2D_OK;
	HIPCHK(hipMemcpyAsync(c->b.bcpar, bc_par, 4 * (size_t)c->nbc * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_set_tau(cfd2d_ctx* c, double tau)
{
	if (!c || !(tau > 0.0)) return fail(CFD2D_ERR_ARG, "bad tau");
	std::vector<double> dts(c->nc);
	for (int k = 0; k < c->ncase; k++) c->case_tau[k] = tau;
	for (int i = 0; i < c->nc; i++) dts[i] = tau / c->hS[i];
	HIPCHK(hipMemcpyAsync(c->b.dtS, dts.data(), c->nc * sizeof(double), hipMemcpyHostToDevice, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_gradients(cfd2d_ctx* c, double* g)
{
	if (!c || !g) return fail(CFD2D_ERR_ARG, "null");
	do_gradient(c);
	int rc = check_launch("k_gradient");
	if (rc) return rc;
	std::vector<double> rec(8 * (size_t)c->nc);
	HIPCHK(hipMemcpyAsync(rec.data(), c->b.G, rec.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	for (int i = 0; i < c->nc; i++)
		for (int q = 0; q < 8; q++) g[(size_t)q * c->nc + i] = rec[8 * (size_t)i + q];
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_edge_fluxes(cfd2d_ctx* c, double* f)
{
	if (!c || !f) return fail(CFD2D_ERR_ARG, "null");
	do_gradient(c);
	do_edge(c);
	int rc = check_launch("k_edge_flux");
	if (rc) return rc;
	HIPCHK(hipMemcpyAsync(f, c->b.F, 4 * (size_t)c->ne * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_residual(cfd2d_ctx* c, double* r)
{
	if (!c || !r) return fail(CFD2D_ERR_ARG, "null");
	do_gradient(c);
	do_edge(c);
	hipLaunchKernelGGL(k_residual, dim3(grid_of(c->nc)), dim3(BLOCK), 0, c->st, c->b);
	int rc = check_launch("k_residual");
	if (rc) return rc;
	HIPCHK(hipMemcpyAsync(r, c->b.R, 4 * (size_t)c->nc * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	return CFD2D_OK;
}

int cfd2d_step(cfd2d_ctx* c, int nsteps)
{
	if (!c || nsteps < 0) return fail(CFD2D_ERR_ARG, "bad arguments");
	if (c->prefer_graph && !c->timing) return cfd2d_step_graph(c, nsteps);
	for (int s = 0; s < nsteps; s++) step_once(c);
	return check_launch("cfd2d_step");
}

// nsteps steps as recorded runs of up to GRAPH_STEPS steps each: a run of
// that many steps is recorded once (its kernels with their arguments) and
// replayed with one launch, so a pass of up to GRAPH_STEPS steps is one
// launch; at most GRAPH_KEEP recorded lengths are kept
constexpr int GRAPH_STEPS = 500, GRAPH_KEEP = 4;

int cfd2d_step_graph(cfd2d_ctx* c, int nsteps)
{
	if (!c || nsteps < 0) return fail(CFD2D_ERR_ARG, "bad arguments");
	while (nsteps > 0) {
		int k = std::min(nsteps, GRAPH_STEPS);
		hipGraphExec_t x = nullptr;
		for (auto& g : c->graphs)
			if (g.steps == k) x = g.x;
		if (!x) {
			if ((int)c->graphs.size() >= GRAPH_KEEP) {
				(void)hipGraphExecDestroy(c->graphs.front().x);
				(void)hipGraphDestroy(c->graphs.front().g);
				c->graphs.erase(c->graphs.begin());
			}
			hipGraph_t g = nullptr;
			c->capturing = true;
			HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
			for (int s = 0; s < k; s++) step_once(c);
			hipError_t e = hipStreamEndCapture(c->st, &g);
			c->capturing = false;
			if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
			e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
			if (e != hipSuccess) {
				(void)hipGraphDestroy(g);
				return hip_fail(e, "hipGraphInstantiate");
			}
			c->graphs.push_back(cfd2d_ctx::StepGraph{k, g, x});
		}
		HIPCHK(hipGraphLaunch(x, c->st));
		nsteps -= k;
	}
	return CFD2D_OK;
}

int cfd2d_sync(cfd2d_ctx* c)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipStreamSynchronize(c->st));
	harvest(c);
	unsigned long long err[2];
	HIPCHK(hipMemcpy(err, c->b.err, sizeof(err), hipMemcpyDeviceToHost));
	if (err[0]) return fail(CFD2D_ERR_NEWTON, "Riemann Newton iteration hit its cap " + std::to_string(err[0]) + " times");
	if (err[1]) return fail(CFD2D_ERR_NONFINITE, "non-finite state in " + std::to_string(err[1]) + " cell updates");
	return CFD2D_OK;
}

int cfd2d_status(cfd2d_ctx* c, long long* hits, int* nonfinite)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipStreamSynchronize(c->st));
	unsigned long long err[2];
	HIPCHK(hipMemcpy(err, c->b.err, sizeof(err), hipMemcpyDeviceToHost));
	if (hits) *hits = (long long)err[0];
	if (nonfinite) *nonfinite = (int)err[1];
	return CFD2D_OK;
}

int cfd2d_wall_force(cfd2d_ctx* c, const int* mask, double* fx, double* fy)
{
	if (!c || !mask) return fail(CFD2D_ERR_ARG, "null");
	std::vector<double> P(4 * (size_t)c->nc);
	HIPCHK(hipStreamSynchronize(c->st));
	HIPCHK(hipMemcpy(P.data(), c->b.P, P.size() * sizeof(double), hipMemcpyDeviceToHost));
	double X = 0.0, Y = 0.0;
	for (int e = 0; e < c->ne; e++) {
		if (!mask[e]) continue;
		double p = P[4 * (size_t)c->hc1[e] + 1];
		X += p * c->hel[e] * c->hnx[e];
		Y += p * c->hel[e] * c->hny[e];
	}
	if (fx) *fx = X;
	if (fy) *fy = Y;
	return CFD2D_OK;
}

int cfd2d_timing(cfd2d_ctx* c, int enable)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	HIPCHK(hipStreamSynchronize(c->st));
	harvest(c);
	c->timing = enable != 0;
	for (int k = 0; k < CFD2D_K_COUNT; k++) {
		c->tot_ms[k] = 0.0;
		c->launches[k] = 0;
	}
	return CFD2D_OK;
}

int cfd2d_stream_state(cfd2d_ctx* c, int mode)
{
	if (!c || mode < -1 || mode > 1) return fail(CFD2D_ERR_ARG, "cfd2d_stream_state: mode is -1, 0 or 1");
	c->stream_mode = mode;
	drop_graph(c);  // the recorded step holds the previous choice
	return CFD2D_OK;
}

int cfd2d_fuse_gradient(cfd2d_ctx* c, int mode)
{
	if (!c || mode < -1 || mode > 2) return fail(CFD2D_ERR_ARG, "cfd2d_fuse_gradient: mode is -1, 0, 1 or 2");
	if (mode == 2 && c->b.ntiles == 0)
		return fail(CFD2D_ERR_ARG, "cfd2d_fuse_gradient: no tiled edge pass for this mesh (edges not in c1 order)");
	c->fuse_mode = mode;
	drop_graph(c);  // the recorded step holds the previous choice
	return CFD2D_OK;
}

int cfd2d_autotune(cfd2d_ctx* c, int trial_steps, int* fuse_mode, int* use_graph, int* stream, double* ms_per_step)
{
	if (!c || trial_steps < 1) return fail(CFD2D_ERR_ARG, "cfd2d_autotune: trial_steps >= 1");
	size_t nc = c->nc;
	// everything a step reads and changes, restored after every trial
	struct Part {
		void* dev;
		size_t bytes;
	} parts[] = {{c->b.U, 4 * nc * sizeof(double)},      {c->b.P, 4 * nc * sizeof(double)},
	             {c->b.flags, nc * sizeof(unsigned)},     {c->b.dtS, nc * sizeof(double)},
	             {c->b.loc, nc * sizeof(double)},         {c->b.err, 2 * sizeof(unsigned long long)},
	             {c->b.anyflag, 2 * sizeof(int)},         {c->b.blockflag, (size_t)c->b.nblocks * sizeof(int)}};
	const int np = sizeof(parts) / sizeof(parts[0]);
	size_t total = 0, off[np];
	for (int k = 0; k < np; k++) {
		off[k] = total;
		total += (parts[k].bytes + 255) & ~(size_t)255;
	}
	HIPCHK(hipStreamSynchronize(c->st));
	char* save = nullptr;
	HIPCHK(hipMalloc(&save, total));
	hipEvent_t e0 = nullptr, e1 = nullptr;
	bool timing = c->timing;
	int fuse0 = c->fuse_mode, stream0 = c->stream_mode;
	bool graph0 = c->prefer_graph;
	auto copy_all = [&](bool to_save) -> hipError_t {
		for (int k = 0; k < np; k++) {
			void* d = to_save ? (void*)(save + off[k]) : parts[k].dev;
			const void* s2 = to_save ? parts[k].dev : (const void*)(save + off[k]);
			hipError_t r = hipMemcpyAsync(d, s2, parts[k].bytes, hipMemcpyDeviceToDevice, c->st);
			if (r != hipSuccess) return r;
		}
		return hipSuccess;
	};
	hipError_t err = copy_all(true);
	if (err == hipSuccess) err = hipEventCreate(&e0);
	if (err == hipSuccess) err = hipEventCreate(&e1);
	c->timing = false;
	float best = -1.0f;
	int shape0 = c->edge_shape;
	int best_fuse = 0, best_graph = 0, best_stream = 0, best_shape = 0;
	// the code-path variants with the default launch shape, then the launch
	// shapes of the per-edge pass with the fastest of them
	int nfuse = c->b.ntiles > 0 ? 3 : 2, nbase = 4 * nfuse;
	for (int v = 0; v < nbase + EDGE_SHAPES - 1 && err == hipSuccess; v++) {
		int fuse = v % nfuse, graph = (v / nfuse) & 1, strm = (v / nfuse / 2) & 1, shape = 0;
		if (v >= nbase) {
			if (best_fuse == 2) break;  // the tiled pass has one launch shape
			fuse = best_fuse;
			graph = best_graph;
			strm = best_stream;
			shape = v - nbase + 1;
		}
		c->fuse_mode = fuse;
		c->prefer_graph = graph != 0;
		c->stream_mode = strm;
		c->edge_shape = shape;
		drop_graph(c);
		if ((err = copy_all(false)) != hipSuccess) break;
		if (cfd2d_step(c, graph ? trial_steps : 1) != CFD2D_OK) {  // warm-up (records the timed graph)
			err = hipErrorUnknown;
			break;
		}
		if ((err = hipEventRecord(e0, c->st)) != hipSuccess) break;
		if (cfd2d_step(c, trial_steps) != CFD2D_OK) {
			err = hipErrorUnknown;
			break;
		}
		if ((err = hipEventRecord(e1, c->st)) != hipSuccess) break;
		if ((err = hipEventSynchronize(e1)) != hipSuccess) break;
		float ms = 0.0f;
		if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
		if (best < 0.0f || ms < best) {
			best = ms;
			best_fuse = fuse;
			best_graph = graph;
			best_stream = strm;
			best_shape = shape;
		}
	}
	if (err == hipSuccess) err = copy_all(false);
	if (err == hipSuccess) err = hipStreamSynchronize(c->st);
	c->timing = timing;
	drop_graph(c);
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	(void)hipFree(save);
	if (err != hipSuccess) {
		c->fuse_mode = fuse0;
		c->prefer_graph = graph0;
		c->stream_mode = stream0;
		c->edge_shape = shape0;
		return hip_fail(err, "cfd2d_autotune");
	}
	c->fuse_mode = best_fuse;
	c->prefer_graph = best_graph != 0;
	c->stream_mode = best_stream;
	c->edge_shape = best_shape;
	if (fuse_mode) *fuse_mode = best_fuse;
	if (use_graph) *use_graph = best_graph;
	if (stream) *stream = best_stream;
	if (ms_per_step) *ms_per_step = best / trial_steps;
	return CFD2D_OK;
}

int cfd2d_edge_shape(cfd2d_ctx* c, int shape)
{
	if (!c || shape < 0 || shape >= EDGE_SHAPES) return fail(CFD2D_ERR_ARG, "cfd2d_edge_shape: shape is 0..3");
	c->edge_shape = shape;
	drop_graph(c);  // the recorded step holds the previous choice
	return CFD2D_OK;
}

int cfd2d_tuning(cfd2d_ctx* c, int* fuse_mode, int* use_graph, int* stream_state, int* edge_shape)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	if (fuse_mode) *fuse_mode = c->fuse_mode;
	if (use_graph) *use_graph = c->prefer_graph ? 1 : 0;
	if (stream_state) *stream_state = c->stream_mode;
	if (edge_shape) *edge_shape = c->edge_shape;
	return CFD2D_OK;
}

int cfd2d_timing_kernels(cfd2d_ctx* c, unsigned mask)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	c->timing_mask = mask;
	return CFD2D_OK;
}

int cfd2d_timing_read(cfd2d_ctx* c, int kernel, double* total_ms, long long* launches)
{
	if (!c || kernel < 0 || kernel >= CFD2D_K_COUNT) return fail(CFD2D_ERR_ARG, "bad kernel id");
	HIPCHK(hipStreamSynchronize(c->st));
	harvest(c);
	if (total_ms) *total_ms = c->tot_ms[kernel];
	if (launches) *launches = c->launches[kernel];
	return CFD2D_OK;
}

int cfd2d_ctx_info(cfd2d_ctx* c, int* nc, int* ne, int* ncase)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	if (nc) *nc = c->nc;
	if (ne) *ne = c->ne;
	if (ncase) *ncase = c->ncase;
	return CFD2D_OK;
}

int cfd2d_ctx_stream(cfd2d_ctx* c, void** s)
{
	if (!c || !s) return fail(CFD2D_ERR_ARG, "null");
	*s = (void*)c->st;
	return CFD2D_OK;
}

}  // extern "C"

#include "adjoint.inl"
#include "geometry.inl"
#include "implicit.inl"
#include "pointwise.inl"
