// Reverse mode (discrete adjoint) of the explicit step.  Included at the end
// of solver.hip, whose kernels, buffers and context it reuses.
//
// Objective J = sum over the masked edges of p(c1) * l * n_y after nsteps
// steps (the pressure force of FVM_TVD_IMPLICIT::calcLiftForce,
// src/methods/fvm_tvd_implicit.cpp:1088-1110, y component).  The reverse
// sweep gives dJ/dU0 (every cell), dJ/d(boundary parameters) and, on
// request, the per-edge and per-cell geometry adjoints the host maps to node
// coordinates.  Scope: LAX, HLLC or the exact Godunov flux, Green-Gauss or
// LSQ gradients, Heun RK2 or SSP-RK3, all boundary types, periodic pairing;
// no limiter, no viscous terms; the time step is held at the value in effect
// (it is not differentiated).
//
// Per RK stage with input state U_s, U_{s+1} = a U0 + b (U_s + tau/S R(U_s)):
//   given W = dJ/dU_{s+1}:  acc0 += a W,  dJ/dU_s = b W + (dR/dU_s)^T (b tau/S W)
// The vector-Jacobian product of R runs over the primitives P and gradients
// G the forward pass recorded for that stage.  First the reverse of the flux,
// of the ghost on boundaries and of the reconstruction into (P, G) adjoints
// per cell: for HLLC and the exact Godunov flux (VJPs heavy in VALU
// work) once per edge, k_adj_edge (two lanes
// per edge, one Gauss point each, into a per-edge buffer for both sides) then
// k_adj_cell (each cell gathers its faces' entries in ascending edge order);
// for LAX k_adj_face (each cell evaluates its three faces from its own side,
// no edge buffer; with geometry terms k_adj_edge's geometry-only mode beside
// it).  Then k_adj_grad: reverse of the gradient operator (own
// and neighbours' gradient adjoints, ghost terms), the primitive->conserved
// Jacobian and the RK bookkeeping.  Boundary-parameter
// adjoints accumulate per boundary edge and are summed per boundary object
// once, in an order fixed by the launch shape.  No floating-point atomics:
// the result does not depend on scheduling.
//
// Storage: the forward kernels write each stage's P and G into per-stage
// slots (the update pass writes the next stage's P and, while recording,
// forms and stores the stage's gradient records), so recording costs no
// extra pass.  The state is checkpointed every ckpt steps; the last segment is
// recorded during the forward pass itself and every earlier segment is
// recomputed once from its checkpoint.  With ckpt >= nsteps nothing is
// recomputed.

namespace {

constexpr int EA = 16;  // per edge [side][Gauss point][r, p, u, v]; on a boundary the
                        // c2 half holds the adjoint of the ghost's cell-centre input

struct AdjBufs {
	double* Wn;     // [nc][4] dJ/dU_{s+1}
	double* Ws;     // [nc][4] dJ/dU_s
	double* Acc;    // [nc][4] sum of a * W over the stages of the current step
	double* Pb;     // [4][nc] adjoint of the primitives
	double* Gb;     // [8][nc] adjoint of the gradients
	                // (component arrays: held as per-cell records, the gfx950
	                // build of k_adj_grad returned wrong boundary-cell values
	                // at -O2/-O3, correct at -O1 or -fno-unroll-loops; both
	                // layouts compiled for the host run clean under ASan and
	                // UBSan with bounds checks and agree bitwise, so the source
	                // holds no out-of-bounds or uninitialised access that would
	                // explain it -- tests/test_hostemu.py)
	double* Ea;     // [16][ne]
	double* Ebc;    // [8][ne] boundary-parameter adjoints per edge, summed over stages
	                // (two slots of 4: one per Gauss point of the edge pass)
	double* bcbar;  // [nbc][4]
	const int* cslot;   // [3][nc] slot of this cell in its neighbour's face list
	const int* bedges;  // boundary edges grouped by boundary object, each run ascending
	const int* bcoff;   // [nbc + 1] start of each object's run in bedges
	int nbedges;
	// geometry adjoints, summed over every stage (GEO kernels only; each slot
	// is added to by one thread per kernel):
	double* Ge;  // [GE_N][ne] per edge: n-bar of the flux at Gauss points 1, 2
	             // (x, y each), the Gauss-point offset adjoints of side s at
	             // Gauss point g (x, y) at 4 + 4 s + 2 g, l-bar at GE_L
	double* WF;  // [2][ne] per stage (GEO): W of c1 and of c2 (0 on a boundary)
	             // dotted with the stage's flux F_e, for the area terms
	double* Gc;  // [GC_N][nc] per cell: S-bar, then per face k the adjoint of
	             // the Green-Gauss weight n l (x, y) or of the LSQ weight (x, y)
};

constexpr int GE_L = 12, GE_N = 13, GC_N = 7;

// d(ghost)/d(inside L, centre C, parameters) applied to gb (r, p, u, v
// order): the reverse of ghost_state.  Every output component is assigned on
// every path (the callers add them where they belong).
__device__ __forceinline__ void ghost_vjp(int type, const double* par, double nx, double ny, const Prim& L,
                                          const Prim& C, const MatV& m, const double* gb, double* dL, double* dC,
                                          double* dpar, double* dn = nullptr)
{
	if (dn) dn[0] = dn[1] = 0.0;  // only the walls' mirror depends on the normal
	if (type == BND_INLET) {
		// u, v, T, p from the parameters; r = P M / (T gR)
		double T = par[2], P = par[3];
		for (int q = 0; q < 4; q++) {
			dL[q] = 0.0;
			dC[q] = 0.0;
		}
		dpar[0] = gb[2];
		dpar[1] = gb[3];
		dpar[2] = -gb[0] * P * m.M / (T * T * GAS_R);
		dpar[3] = gb[1] + gb[0] * m.M / (T * GAS_R);
		return;
	}
	if (type == BND_PRESSURE_OUTLET) {
		// r, u, v copied; p fixed while |u.n| < c of the cell centre
		double un = L.u * nx + L.v * ny;
		double czc = sqrt(m.gam * C.p / C.r);
		bool sub = fabs(un) < czc;
		dL[0] = gb[0];
		dL[1] = sub ? 0.0 : gb[1];
		dL[2] = gb[2];
		dL[3] = gb[3];
		for (int q = 0; q < 4; q++) dC[q] = 0.0;
		dpar[0] = sub ? gb[1] : 0.0;
		dpar[1] = 0.0;
		dpar[2] = 0.0;
		dpar[3] = 0.0;
		return;
	}
	// outlet and walls: p copied, r = pL M / (Tc gR), Tc = pc / (rc (gamma-1) Cv);
	// walls mirror the normal velocity
	double Tc = temperature(C, m);
	double k = m.M / GAS_R;
	double Tb = -gb[0] * L.p * k / (Tc * Tc);
	double den = C.r * (m.gam - 1) * m.Cv;
	dL[0] = 0.0;
	dL[1] = gb[0] * k / Tc + gb[1];
	if (type == BND_OUTLET) {
		dL[2] = gb[2];
		dL[3] = gb[3];
	} else {
		double Unb = -2.0 * (nx * gb[2] + ny * gb[3]);
		dL[2] = gb[2] + Unb * nx;
		dL[3] = gb[3] + Unb * ny;
		if (dn) {
			// u' = u - 2 n (u.n): d/dn of (u', v') applied to gb
			double Un = L.u * nx + L.v * ny;
			dn[0] = -2.0 * (gb[2] * Un + (nx * gb[2] + ny * gb[3]) * L.u);
			dn[1] = -2.0 * (gb[3] * Un + (nx * gb[2] + ny * gb[3]) * L.v);
		}
	}
	dC[0] = -Tb * C.p / (C.r * den);
	dC[1] = Tb / den;
	dC[2] = 0.0;
	dC[3] = 0.0;
	for (int q = 0; q < 4; q++) dpar[q] = 0.0;
}

// reverse of flux_lax: Lb, Rb (r, p, u, v) = fb times the flux Jacobian.
// Written for few divisions: 1/r per side, and dc/dp = c/(2p),
// dc/dr = -c/(2r) on the side that sets the wave speed.
__device__ __forceinline__ void lax_vjp(const Prim& L, const Prim& R, double nx, double ny, double G, double igm1,
                                        const double* fb, double* Lb, double* Rb, double* nb = nullptr,
                                        double* fo = nullptr)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double irl = 1.0 / L.r, irr = 1.0 / R.r;
	double cl = sqrt(G * L.p * irl), cr = sqrt(G * R.p * irr);
	double al = fabs(unl) + cl, ar = fabs(unr) + cr;
	bool left = al > ar;
	double alpha = left ? al : ar;
	double kl = 0.5 * (L.u * L.u + L.v * L.v), kr = 0.5 * (R.u * R.u + R.v * R.v);
	double rel = L.p * igm1 + L.r * kl, rer = R.p * igm1 + R.r * kr;
	double frl = L.r * unl, frr = R.r * unr;
	double h0 = 0.5 * fb[0], h1 = 0.5 * fb[1], h2 = 0.5 * fb[2], h3 = 0.5 * fb[3];
	double alb = -(h0 * (R.r - L.r) + h1 * (R.r * R.u - L.r * L.u) + h2 * (R.r * R.v - L.r * L.v) + h3 * (rer - rel));
	double frrb = h0 + h1 * R.u + h2 * R.v, frlb = h0 + h1 * L.u + h2 * L.v;
	double unrb = h3 * (rer + R.p) + frrb * R.r, unlb = h3 * (rel + L.p) + frlb * L.r;
	// adjoints of the conserved variables (r, ru, rv, re) of each side
	double rorb = frrb * unr - h0 * alpha, rolb = frlb * unl + h0 * alpha;
	double rurb = -h1 * alpha, rulb = h1 * alpha;
	double rvrb = -h2 * alpha, rvlb = h2 * alpha;
	double rerb = h3 * (unr - alpha), relb = h3 * (unl + alpha);
	double Rub = h1 * frr, Lub = h1 * frl, Rvb = h2 * frr, Lvb = h2 * frl;
	double Rpb = h1 * nx + h2 * ny + h3 * unr, Lpb = h1 * nx + h2 * ny + h3 * unl;
	double Lrb = 0.0, Rrb = 0.0;
	// alpha = |u.n| + c of the side with the larger value
	double cs = left ? cl : cr, ps = left ? L.p : R.p, irs = left ? irl : irr;
	double ts = alb * 0.5 * cs;
	double psb = ts / ps, rsb = -ts * irs;
	if (left) {
		unlb += alb * (unl < 0 ? -1.0 : 1.0);
		Lpb += psb;
		Lrb += rsb;
	} else {
		unrb += alb * (unr < 0 ? -1.0 : 1.0);
		Rpb += psb;
		Rrb += rsb;
	}
	if (nb) {  // the normal enters through u.n of both sides and the pressure terms
		nb[0] = unlb * L.u + unrb * R.u + h1 * (R.p + L.p);
		nb[1] = unlb * L.v + unrb * R.v + h2 * (R.p + L.p);
	}
	Lrb += rolb + rulb * L.u + rvlb * L.v + relb * kl;
	Lub += rulb * L.r + relb * L.r * L.u + unlb * nx;
	Lvb += rvlb * L.r + relb * L.r * L.v + unlb * ny;
	Lpb += relb * igm1;
	Rrb += rorb + rurb * R.u + rvrb * R.v + rerb * kr;
	Rub += rurb * R.r + rerb * R.r * R.u + unrb * nx;
	Rvb += rvrb * R.r + rerb * R.r * R.v + unrb * ny;
	Rpb += rerb * igm1;
	Lb[0] = Lrb; Lb[1] = Lpb; Lb[2] = Lub; Lb[3] = Lvb;
	Rb[0] = Rrb; Rb[1] = Rpb; Rb[2] = Rub; Rb[3] = Rvb;
	if (fo) {  // the flux itself (flux_lax's expressions on these intermediates)
		fo[0] = 0.5 * (frr + frl - alpha * (R.r - L.r));
		fo[1] = 0.5 * (frr * R.u + frl * L.u + (R.p + L.p) * nx - alpha * (R.r * R.u - L.r * L.u));
		fo[2] = 0.5 * (frr * R.v + frl * L.v + (R.p + L.p) * ny - alpha * (R.r * R.v - L.r * L.v));
		fo[3] = 0.5 * ((rer + R.p) * unr + (rel + L.p) * unl - alpha * (rer - rel));
	}
}

// The adjoint of one side only: F(L, R; n) = -F(R, L; -n) holds exactly for
// the LAX flux (every term changes sign with n, alpha does not), so the
// adjoint of the cell's own reconstructed state O against the other side's
// X is lax_vjp's left part evaluated with (L, R, n, fb) = (O, X, s n, s fb),
// s = +1 for c1 and -1 for c2.  own_c2 keeps the forward pass's choice of
// the side whose wave speed is alpha when the two tie (the right one, c2).
// About half of lax_vjp's arithmetic: the other side's adjoint is formed by
// the other cell.
__device__ __forceinline__ void lax_vjp_own(const Prim& L, const Prim& R, double nx, double ny, double G,
                                            double igm1, const double* fb, bool own_c2, double* Lb)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double irl = 1.0 / L.r, irr = 1.0 / R.r;
	double cl = sqrt(G * L.p * irl), cr = sqrt(G * R.p * irr);
	double al = fabs(unl) + cl, ar = fabs(unr) + cr;
	bool own = own_c2 ? !(ar > al) : al > ar;  // the forward pass's al > ar, seen from this side
	double alpha = own ? al : ar;
	double kl = 0.5 * (L.u * L.u + L.v * L.v), kr = 0.5 * (R.u * R.u + R.v * R.v);
	double rel = L.p * igm1 + L.r * kl, rer = R.p * igm1 + R.r * kr;
	double frl = L.r * unl;
	double h0 = 0.5 * fb[0], h1 = 0.5 * fb[1], h2 = 0.5 * fb[2], h3 = 0.5 * fb[3];
	double alb = -(h0 * (R.r - L.r) + h1 * (R.r * R.u - L.r * L.u) + h2 * (R.r * R.v - L.r * L.v) + h3 * (rer - rel));
	double frlb = h0 + h1 * L.u + h2 * L.v;
	double unlb = h3 * (rel + L.p) + frlb * L.r;
	double rolb = frlb * unl + h0 * alpha, rulb = h1 * alpha, rvlb = h2 * alpha, relb = h3 * (unl + alpha);
	double Lub = h1 * frl, Lvb = h2 * frl;
	double Lpb = h1 * nx + h2 * ny + h3 * unl, Lrb = 0.0;
	if (own) {
		double ts = alb * 0.5 * cl;
		unlb += alb * (unl < 0 ? -1.0 : 1.0);
		Lpb += ts / L.p;
		Lrb += -ts * irl;
	}
	Lrb += rolb + rulb * L.u + rvlb * L.v + relb * kl;
	Lub += rulb * L.r + relb * L.r * L.u + unlb * nx;
	Lvb += rvlb * L.r + relb * L.r * L.v + unlb * ny;
	Lpb += relb * igm1;
	Lb[0] = Lrb;
	Lb[1] = Lpb;
	Lb[2] = Lub;
	Lb[3] = Lvb;
}

// one-direction forward-mode (dual) number: the exact Godunov flux's VJP
// past the reverse solver's iteration limit
struct D1 {
	double v, d;
};
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator/(D1 a, D1 b) { return D1{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ D1 operator*(double a, D1 b) { return D1{a * b.v, a * b.d}; }
__device__ __forceinline__ D1 operator+(D1 a, double b) { return D1{a.v + b, a.d}; }
__device__ __forceinline__ D1 operator/(D1 a, double b) { return D1{a.v / b, a.d / b}; }
__device__ __forceinline__ D1 dsqrt(D1 a)
{
	double s = sqrt(a.v);
	return D1{s, a.d / (2.0 * s)};
}

__device__ __forceinline__ D1 operator-(D1 a, double b) { return D1{a.v - b, a.d}; }
__device__ __forceinline__ D1 operator-(double a, D1 b) { return D1{a - b.v, -b.d}; }
__device__ __forceinline__ D1 operator+(double a, D1 b) { return D1{a + b.v, b.d}; }
__device__ __forceinline__ D1 operator*(D1 a, double b) { return D1{a.v * b, a.d * b}; }
__device__ __forceinline__ D1 operator/(double a, D1 b) { return D1{a / b.v, -a * b.d / (b.v * b.v)}; }
__device__ __forceinline__ D1 dsel(bool c, D1 a, D1 b) { return c ? a : b; }

// side_root and the fan power on duals: the value as the forward forms it,
// the derivative of t^(1/k) (k = 1/OGAM on a rarefaction, 2 on a shock)
__device__ __forceinline__ D1 dside_root(D1 t, bool rare, const RimConst& K)
{
	double y = side_root(t.v, rare, K);
	return D1{y, (rare ? K.OGAM : 0.5) * y / t.v * t.d};
}

__device__ __forceinline__ D1 dpow_iagam(D1 x, const RimConst& K)
{
	double y = pow_iagam(x.v, K);
	return D1{y, K.IAGAM * y / x.v * x.d};
}

// rim (physics.hpp) on one-direction duals: the same operations, branches
// and Newton iterations as the forward evaluation of the same inputs (every
// test reads the value parts), so the tangent is that of the discrete map
// the forward step evaluates, iteration count and stop rule included.
__device__ __forceinline__ void rim_d1(D1& RI, D1& EI, D1& PI, D1& UI, D1& VI, D1 RB, D1 PB, D1 UB, D1 VB, D1 RE,
                                       D1 PE, D1 UE, D1 VE, const RimConst& K, int max_iter)
{
	const double eps = 1.0e-5;
	D1 zero{0.0, 0.0};
	D1 US = zero, UF = zero, RF, RS, EF, ES, SBL, SFL, SSL, SEL;
	D1 rpB = RB * PB, rpE = RE * PE;
	D1 i4 = 1.0 / (rpB * rpE);
	D1 iB = rpE * i4, iE = rpB * i4;
	D1 IRB = PB * iB, IPB = RB * iB, IRE = PE * iE, IPE = RE * iE;
	D1 CB = dsqrt(K.GAM * PB * IRB), CE = dsqrt(K.GAM * PE * IRE);
	D1 RCB = RB * CB, RCE = RE * CE;
	D1 EB = CB * CB * K.ISGAM, EE = CE * CE * K.ISGAM;
	D1 DU = UB - UE;
	if (DU.v * K.AGAM < -2.0 * (CB.v + CE.v)) {
		RF = zero; RS = zero; EF = zero; ES = zero;
		SBL = UB - CB;
		SFL = UB + 2.0 * CB / K.AGAM;
		SSL = UE - 2.0 * CE / K.AGAM;
		SEL = UE + CE;
	} else {
		D1 P = (PB * RCE + PE * RCB + DU * RCB * RCE) / (RCB + RCE);
		int it = 0;
		for (;;) {
			if (P.v < eps) P = D1{eps, 0.0};
			D1 F[2], FS[2], x[2], t[2], y[2], den[2];
			bool rare[2];
			for (int sd = 0; sd < 2; sd++) {
				D1 PK = sd ? PE : PB, IPK = sd ? IPE : IPB, RC = sd ? RCE : RCB;
				x[sd] = P * IPK;
				rare[sd] = PK.v > P.v;
				t[sd] = dsel(rare[sd], x[sd], K.PGAM * x[sd] + K.OGAM);
				y[sd] = dside_root(t[sd], rare[sd], K);
				den[sd] = dsel(rare[sd], K.GAM * P, RC * y[sd] * t[sd]);
			}
			D1 ir = 1.0 / (den[0] * den[1]);
			for (int sd = 0; sd < 2; sd++) {
				D1 PK = sd ? PE : PB, C = sd ? CE : CB;
				D1 r = den[1 - sd] * ir;
				F[sd] = dsel(rare[sd], K.DGAM * (C * y[sd] - C), (P - PK) * t[sd] * r);
				FS[sd] = dsel(rare[sd], C * y[sd] * r, (K.QGAM * x[sd] + K.FGAM) * r * K.IRGAM);
			}
			D1 res = DU - F[0] - F[1];
			P = P + res / (FS[0] + FS[1]);
			++it;
			if (!(fabs(res.v) > eps)) break;
			if (max_iter > 0 && it >= max_iter) break;
		}
		D1 yy[2], den[2];
		bool rr[2];
		for (int sd = 0; sd < 2; sd++) {
			D1 PK = sd ? PE : PB, IPK = sd ? IPE : IPB, IRK = sd ? IRE : IRB, C = sd ? CE : CB;
			D1 RK = sd ? RE : RB;
			double sg = sd ? 1.0 : -1.0;
			rr[sd] = PK.v > P.v;
			D1 t = dsel(rr[sd], P * IPK, (K.TGAM * P + K.HGAM * PK) * IRK);
			yy[sd] = dside_root(t, rr[sd], K);
			D1 Z = C * yy[sd];
			D1 M = (-sg) * (RK * yy[sd]);
			D1 Q = PK - P + M * ((-sg) * yy[sd]);
			den[sd] = dsel(rr[sd], K.AGAM * (Z * Z * K.ISGAM), Q * M);
		}
		D1 is = 1.0 / (den[0] * den[1]);
		for (int sd = 0; sd < 2; sd++) {
			D1 PK = sd ? PE : PB, C = sd ? CE : CB;
			D1 RK = sd ? RE : RB, UK = sd ? UE : UB;
			double sg = sd ? 1.0 : -1.0;
			bool rare = rr[sd];
			D1 y = yy[sd];
			D1 Z = C * y;
			D1 Er = Z * Z * K.ISGAM;
			D1 D = UK + sg * y;
			D1 M = (-sg) * (RK * y);
			D1 Q = PK - P + M * ((-sg) * y);
			D1 r = den[1 - sd] * is;
			D1 Rs, Es, Us, head, tail;
			if (rare) {
				Es = Er;
				Us = UK + (sg * K.DGAM) * (Z - C);
				Rs = P * r;
				head = UK + sg * C;
				tail = Us + sg * Z;
			} else {
				D1 IM = Q * r;
				Rs = M * M * M * r;
				Us = D + Q * IM;
				Es = P * Q * IM * IM * K.IAGAM;
				head = D;
				tail = D;
			}
			if (sd) {
				RS = Rs; ES = Es; US = Us; SEL = head; SSL = tail;
			} else {
				RF = Rs; EF = Es; UF = Us; SBL = head; SFL = tail;
			}
		}
	}
	if (SEL.v <= 0.0) {
		RI = RE; EI = EE; UI = UE; VI = VE;
	} else if (SBL.v >= 0.0) {
		RI = RB; EI = EB; UI = UB; VI = VB;
	} else if ((SSL.v >= 0.0) && (SFL.v <= 0.0)) {
		if (US.v >= 0.0) {
			RI = RF; EI = EF; UI = UF; VI = VB;
		} else {
			RI = RS; EI = ES; UI = US; VI = VE;
		}
	} else if (SFL.v > 0.0) {
		UI = (UB + K.DGGAM * dsqrt(EB)) * K.IDGAM1;
		VI = VB;
		EI = (UI * UI) * K.ISGAM;
		RI = RB * dpow_iagam(EI / EB, K);
	} else {
		UI = (UE - K.DGGAM * dsqrt(EE)) * K.IDGAM1;
		VI = VE;
		EI = (UI * UI) * K.ISGAM;
		RI = RE * dpow_iagam(EI / EE, K);
	}
	PI = K.AGAM * EI * RI;
}

// Reverse mode of rim (physics.hpp): the VJP of the Riemann solver's five
// outputs with respect to its eight inputs, for the same evaluated map that
// rim_d1 differentiates forward -- its values, branches, Newton iterations
// and shared divisions, each operation reversed.  The iterates are not
// stored: the reverse of Newton iteration k runs the loop again from its
// start to k (n iterations cost n(n+1)/2 forward iterations).  With more than
// RIM_REV_MAX iterations it returns false and leaves the work to the caller's
// forward-mode path.
constexpr int RIM_REV_MAX = 8;

struct RimSetup {
	double rpB, rpE, i4, iB, iE, IRB, IPB, IRE, IPE, CB, CE, RCB, RCE, EB, EE, DU;
};

__device__ __forceinline__ RimSetup rim_setup(double RB, double PB, double UB, double RE, double PE, double UE,
                                              const RimConst& K)
{
	RimSetup f;
	f.rpB = RB * PB;
	f.rpE = RE * PE;
	f.i4 = 1.0 / (f.rpB * f.rpE);
	f.iB = f.rpE * f.i4;
	f.iE = f.rpB * f.i4;
	f.IRB = PB * f.iB;
	f.IPB = RB * f.iB;
	f.IRE = PE * f.iE;
	f.IPE = RE * f.iE;
	f.CB = sqrt(K.GAM * PB * f.IRB);
	f.CE = sqrt(K.GAM * PE * f.IRE);
	f.RCB = RB * f.CB;
	f.RCE = RE * f.CE;
	f.EB = f.CB * f.CB * K.ISGAM;
	f.EE = f.CE * f.CE * K.ISGAM;
	f.DU = UB - UE;
	return f;
}

// one Newton iteration of rim at P: the clamp, both sides' terms, the update
struct RimNewton {
	double Pc, x[2], t[2], y[2], den[2], ir, F[2], FS[2], res, S;
	bool clamped, rare[2];
};

__device__ __forceinline__ RimNewton rim_newton(const RimSetup& f, double P, double PB, double PE, const RimConst& K)
{
	RimNewton s;
	s.clamped = P < 1.0e-5;
	s.Pc = s.clamped ? 1.0e-5 : P;
	for (int sd = 0; sd < 2; sd++) {
		double PK = sd ? PE : PB, IPK = sd ? f.IPE : f.IPB, RC = sd ? f.RCE : f.RCB;
		s.x[sd] = s.Pc * IPK;
		s.rare[sd] = PK > s.Pc;
		s.t[sd] = s.rare[sd] ? s.x[sd] : K.PGAM * s.x[sd] + K.OGAM;
		s.y[sd] = side_root(s.t[sd], s.rare[sd], K);
		s.den[sd] = s.rare[sd] ? K.GAM * s.Pc : RC * s.y[sd] * s.t[sd];
	}
	s.ir = 1.0 / (s.den[0] * s.den[1]);
	for (int sd = 0; sd < 2; sd++) {
		double PK = sd ? PE : PB, C = sd ? f.CE : f.CB;
		double r = s.den[1 - sd] * s.ir;
		s.F[sd] = s.rare[sd] ? K.DGAM * (C * s.y[sd] - C) : (s.Pc - PK) * s.t[sd] * r;
		s.FS[sd] = s.rare[sd] ? C * s.y[sd] * r : (K.QGAM * s.x[sd] + K.FGAM) * r * K.IRGAM;
	}
	s.res = f.DU - s.F[0] - s.F[1];
	s.S = s.FS[0] + s.FS[1];
	return s;
}

// adjoints of the solver inputs and of the setup values
struct RimAdj {
	double RB, PB, UB, VB, RE, PE, UE, VE, IRB, IPB, IRE, IPE, CB, CE, RCB, RCE, EB, EE, DU;
};

// reverse of one Newton iteration: Pnb is the adjoint of the iterate it
// produced; returns the adjoint of the iterate it started from
__device__ __forceinline__ double rim_newton_rev(const RimSetup& f, const RimNewton& s, double PB, double PE,
                                                 const RimConst& K, double Pnb, RimAdj& a)
{
	double Pcb = Pnb;
	double resb = Pnb / s.S;
	double Sb = -Pnb * s.res / (s.S * s.S);
	a.DU += resb;
	double yb[2], tb[2], xb[2] = {0.0, 0.0}, denb[2] = {0.0, 0.0}, irb = 0.0;
	for (int sd = 0; sd < 2; sd++) {
		double PK = sd ? PE : PB, C = sd ? f.CE : f.CB;
		double r = s.den[1 - sd] * s.ir;
		double Fb = -resb, FSb = Sb, rb = 0.0, Cb = 0.0;
		yb[sd] = 0.0;
		tb[sd] = 0.0;
		if (s.rare[sd]) {
			Cb += Fb * K.DGAM * (s.y[sd] - 1.0);
			yb[sd] += Fb * K.DGAM * C;
			Cb += FSb * s.y[sd] * r;
			yb[sd] += FSb * C * r;
			rb += FSb * C * s.y[sd];
		} else {
			double dp = s.Pc - PK;
			Pcb += Fb * s.t[sd] * r;
			(sd ? a.PE : a.PB) -= Fb * s.t[sd] * r;
			tb[sd] += Fb * dp * r;
			rb += Fb * dp * s.t[sd];
			xb[sd] += FSb * K.QGAM * r * K.IRGAM;
			rb += FSb * (K.QGAM * s.x[sd] + K.FGAM) * K.IRGAM;
		}
		(sd ? a.CE : a.CB) += Cb;
		denb[1 - sd] += rb * s.ir;
		irb += rb * s.den[1 - sd];
	}
	double q = s.den[0] * s.den[1], qb = -irb / (q * q);
	denb[0] += qb * s.den[1];
	denb[1] += qb * s.den[0];
	for (int sd = 0; sd < 2; sd++) {
		double RC = sd ? f.RCE : f.RCB, IPK = sd ? f.IPE : f.IPB;
		if (s.rare[sd]) {
			Pcb += denb[sd] * K.GAM;
		} else {
			(sd ? a.RCE : a.RCB) += denb[sd] * s.y[sd] * s.t[sd];
			yb[sd] += denb[sd] * RC * s.t[sd];
			tb[sd] += denb[sd] * RC * s.y[sd];
		}
		tb[sd] += yb[sd] * ((s.rare[sd] ? K.OGAM : 0.5) * s.y[sd] / s.t[sd]);
		xb[sd] += s.rare[sd] ? tb[sd] : tb[sd] * K.PGAM;
		Pcb += xb[sd] * IPK;
		(sd ? a.IPE : a.IPB) += xb[sd] * s.Pc;
	}
	return s.clamped ? 0.0 : Pcb;
}

__device__ __forceinline__ bool rim_vjp(double RB, double PB, double UB, double VB, double RE, double PE, double UE,
                                        double VE, const RimConst& K, int max_iter, double RIb, double EIb, double PIb,
                                        double UIb, double VIb, double* xb)
{
	const double eps = 1.0e-5;
	RimAdj a{};
	RimSetup f = rim_setup(RB, PB, UB, RE, PE, UE, K);
	bool vac = f.DU * K.AGAM < -2.0 * (f.CB + f.CE);
	double P0 = 0.0, P = 0.0;
	int n = 0;
	// star states: y, Z, M, Q, den per side, their shared reciprocal, outputs
	double yy[2] = {0.0, 0.0}, ts[2] = {0.0, 0.0}, Zs[2] = {0.0, 0.0}, Ms[2] = {0.0, 0.0}, Qs[2] = {0.0, 0.0};
	double dens[2] = {1.0, 1.0}, is = 0.0;
	bool rr[2] = {false, false};
	double Rs[2] = {0.0, 0.0}, Es[2] = {0.0, 0.0}, Us[2] = {0.0, 0.0}, head[2], tail[2];
	if (vac) {
		head[0] = UB - f.CB;
		tail[0] = UB + 2.0 * f.CB / K.AGAM;
		tail[1] = UE - 2.0 * f.CE / K.AGAM;
		head[1] = UE + f.CE;
	} else {
		P0 = (PB * f.RCE + PE * f.RCB + f.DU * f.RCB * f.RCE) / (f.RCB + f.RCE);
		P = P0;
		for (;;) {
			RimNewton s = rim_newton(f, P, PB, PE, K);
			P = s.Pc + s.res / s.S;
			++n;
			if (!(fabs(s.res) > eps)) break;
			if (max_iter > 0 && n >= max_iter) break;
		}
		if (n > RIM_REV_MAX) return false;
		for (int sd = 0; sd < 2; sd++) {
			double PK = sd ? PE : PB, IPK = sd ? f.IPE : f.IPB, IRK = sd ? f.IRE : f.IRB, C = sd ? f.CE : f.CB;
			double RK = sd ? RE : RB, sg = sd ? 1.0 : -1.0;
			rr[sd] = PK > P;
			ts[sd] = rr[sd] ? P * IPK : (K.TGAM * P + K.HGAM * PK) * IRK;
			yy[sd] = side_root(ts[sd], rr[sd], K);
			Zs[sd] = C * yy[sd];
			Ms[sd] = (-sg) * (RK * yy[sd]);
			Qs[sd] = PK - P + Ms[sd] * ((-sg) * yy[sd]);
			dens[sd] = rr[sd] ? K.AGAM * (Zs[sd] * Zs[sd] * K.ISGAM) : Qs[sd] * Ms[sd];
		}
		is = 1.0 / (dens[0] * dens[1]);
		for (int sd = 0; sd < 2; sd++) {
			double C = sd ? f.CE : f.CB, UK = sd ? UE : UB, sg = sd ? 1.0 : -1.0;
			double r = dens[1 - sd] * is;
			if (rr[sd]) {
				Es[sd] = Zs[sd] * Zs[sd] * K.ISGAM;
				Us[sd] = UK + (sg * K.DGAM) * (Zs[sd] - C);
				Rs[sd] = P * r;
				head[sd] = UK + sg * C;
				tail[sd] = Us[sd] + sg * Zs[sd];
			} else {
				double IM = Qs[sd] * r, D = UK + sg * yy[sd];
				Rs[sd] = Ms[sd] * Ms[sd] * Ms[sd] * r;
				Us[sd] = D + Qs[sd] * IM;
				Es[sd] = P * Qs[sd] * IM * IM * K.IAGAM;
				head[sd] = D;
				tail[sd] = D;
			}
		}
	}
	// sampling (SBL, SFL = head[0], tail[0]; SEL, SSL = head[1], tail[1])
	double RI, EI;
	int br;
	if (head[1] <= 0.0) {
		br = 0;
		RI = RE;
		EI = f.EE;
	} else if (head[0] >= 0.0) {
		br = 1;
		RI = RB;
		EI = f.EB;
	} else if (tail[1] >= 0.0 && tail[0] <= 0.0) {
		br = Us[1] >= 0.0 ? 2 : 3;
		RI = br == 2 ? Rs[0] : Rs[1];
		EI = br == 2 ? Es[0] : Es[1];
	} else {
		br = tail[0] > 0.0 ? 4 : 5;
		RI = EI = 0.0;  // the fan's values are formed in its reverse below
	}
	// PI = AGAM EI RI; the fans' RI, EI depend on each other, handled there
	double RIt = RIb, EIt = EIb;
	if (br < 4) {
		RIt += PIb * K.AGAM * EI;
		EIt += PIb * K.AGAM * RI;
	}
	double Rsb[2] = {0.0, 0.0}, Esb[2] = {0.0, 0.0}, Usb[2] = {0.0, 0.0};
	switch (br) {
	case 0:
		a.RE += RIt;
		a.EE += EIt;
		a.UE += UIb;
		a.VE += VIb;
		break;
	case 1:
		a.RB += RIt;
		a.EB += EIt;
		a.UB += UIb;
		a.VB += VIb;
		break;
	case 2:
		Rsb[0] += RIt;
		Esb[0] += EIt;
		Usb[0] += UIb;
		a.VB += VIb;
		break;
	case 3:
		Rsb[1] += RIt;
		Esb[1] += EIt;
		Usb[1] += UIb;
		a.VE += VIb;
		break;
	default: {
		// UI = (UK +- DGGAM sqrt(EK)) IDGAM1, EI = UI^2 ISGAM, RI = RK (EI/EK)^(1/AGAM),
		// PI = AGAM EI RI (left fan: K = B and +; right: K = E and -)
		bool left = br == 4;
		double UK = left ? UB : UE, EK = left ? f.EB : f.EE, RK = left ? RB : RE, pm = left ? 1.0 : -1.0;
		double sq = sqrt(EK);
		double ui = (UK + pm * (K.DGGAM * sq)) * K.IDGAM1;
		double ei = ui * ui * K.ISGAM;
		double w = ei / EK, pw = pow_iagam(w, K);
		double ri = RK * pw;
		double rib = RIb + PIb * K.AGAM * ei, eib = EIb + PIb * K.AGAM * ri;
		double pwb = rib * RK;
		double RKb = rib * pw;
		double wb = pwb * (K.IAGAM * pw / w);
		eib += wb / EK;
		double EKb = -wb * ei / (EK * EK);
		double uib = UIb + eib * 2.0 * ui * K.ISGAM;
		double UKb = uib * K.IDGAM1;
		EKb += pm * uib * K.IDGAM1 * K.DGGAM / (2.0 * sq);
		if (left) {
			a.RB += RKb;
			a.EB += EKb;
			a.UB += UKb;
			a.VB += VIb;
		} else {
			a.RE += RKb;
			a.EE += EKb;
			a.UE += UKb;
			a.VE += VIb;
		}
	}
	}
	if (!vac) {
		// the star states
		double Pb = 0.0, isb = 0.0, denb[2] = {0.0, 0.0};
		double yb[2] = {0.0, 0.0}, Zb[2] = {0.0, 0.0}, Mb[2] = {0.0, 0.0}, Qb[2] = {0.0, 0.0};
		for (int sd = 0; sd < 2; sd++) {
			double C = sd ? f.CE : f.CB, sg = sd ? 1.0 : -1.0;
			double r = dens[1 - sd] * is, rb = 0.0;
			if (rr[sd]) {
				Zb[sd] += Esb[sd] * 2.0 * Zs[sd] * K.ISGAM;
				(sd ? a.UE : a.UB) += Usb[sd];
				Zb[sd] += Usb[sd] * (sg * K.DGAM);
				(sd ? a.CE : a.CB) -= Usb[sd] * (sg * K.DGAM);
				Pb += Rsb[sd] * r;
				rb += Rsb[sd] * P;
			} else {
				double IM = Qs[sd] * r, M = Ms[sd], Q = Qs[sd];
				Mb[sd] += Rsb[sd] * 3.0 * M * M * r;
				rb += Rsb[sd] * M * M * M;
				double IMb = Usb[sd] * Q;
				Qb[sd] += Usb[sd] * IM;
				(sd ? a.UE : a.UB) += Usb[sd];  // D = UK + sg y
				yb[sd] += Usb[sd] * sg;
				double e2 = Esb[sd] * K.IAGAM;
				Pb += e2 * Q * IM * IM;
				Qb[sd] += e2 * P * IM * IM;
				IMb += e2 * 2.0 * P * Q * IM;
				Qb[sd] += IMb * r;
				rb += IMb * Q;
			}
			denb[1 - sd] += rb * is;
			isb += rb * dens[1 - sd];
		}
		double q = dens[0] * dens[1], qb = -isb / (q * q);
		denb[0] += qb * dens[1];
		denb[1] += qb * dens[0];
		for (int sd = 0; sd < 2; sd++) {
			double C = sd ? f.CE : f.CB, RK = sd ? RE : RB, IPK = sd ? f.IPE : f.IPB, IRK = sd ? f.IRE : f.IRB;
			double PK = sd ? PE : PB, sg = sd ? 1.0 : -1.0;
			if (rr[sd]) {
				Zb[sd] += denb[sd] * K.AGAM * K.ISGAM * 2.0 * Zs[sd];
			} else {
				Qb[sd] += denb[sd] * Ms[sd];
				Mb[sd] += denb[sd] * Qs[sd];
			}
			(sd ? a.PE : a.PB) += Qb[sd];
			Pb -= Qb[sd];
			Mb[sd] += Qb[sd] * ((-sg) * yy[sd]);
			yb[sd] += Qb[sd] * Ms[sd] * (-sg);
			(sd ? a.RE : a.RB) += Mb[sd] * (-sg) * yy[sd];
			yb[sd] += Mb[sd] * (-sg) * RK;
			(sd ? a.CE : a.CB) += Zb[sd] * yy[sd];
			yb[sd] += Zb[sd] * C;
			double tb = yb[sd] * ((rr[sd] ? K.OGAM : 0.5) * yy[sd] / ts[sd]);
			if (rr[sd]) {
				Pb += tb * IPK;
				(sd ? a.IPE : a.IPB) += tb * P;
			} else {
				double u = K.TGAM * P + K.HGAM * PK;
				Pb += tb * IRK * K.TGAM;
				(sd ? a.PE : a.PB) += tb * IRK * K.HGAM;
				(sd ? a.IRE : a.IRB) += tb * u;
			}
		}
		// the Newton iterations, last to first, each from its recomputed start
		for (int k = n - 1; k >= 0; k--) {
			double Pk = P0;
			for (int m = 0; m < k; m++) {
				RimNewton s = rim_newton(f, Pk, PB, PE, K);
				Pk = s.Pc + s.res / s.S;
			}
			RimNewton s = rim_newton(f, Pk, PB, PE, K);
			Pb = rim_newton_rev(f, s, PB, PE, K, Pb, a);
		}
		// the acoustic initial guess
		double num = PB * f.RCE + PE * f.RCB + f.DU * f.RCB * f.RCE, dd = f.RCB + f.RCE;
		double numb = Pb / dd, ddb = -Pb * num / (dd * dd);
		a.PB += numb * f.RCE;
		a.RCE += numb * PB;
		a.PE += numb * f.RCB;
		a.RCB += numb * PE;
		a.DU += numb * f.RCB * f.RCE;
		a.RCB += numb * f.DU * f.RCE;
		a.RCE += numb * f.DU * f.RCB;
		a.RCB += ddb;
		a.RCE += ddb;
	}
	// the setup
	a.UB += a.DU;
	a.UE -= a.DU;
	a.CB += a.EB * 2.0 * f.CB * K.ISGAM;
	a.CE += a.EE * 2.0 * f.CE * K.ISGAM;
	a.RB += a.RCB * f.CB;
	a.CB += a.RCB * RB;
	a.RE += a.RCE * f.CE;
	a.CE += a.RCE * RE;
	double abB = a.CB / (2.0 * f.CB), abE = a.CE / (2.0 * f.CE);
	a.PB += abB * K.GAM * f.IRB;
	a.IRB += abB * K.GAM * PB;
	a.PE += abE * K.GAM * f.IRE;
	a.IRE += abE * K.GAM * PE;
	a.PB += a.IRB * f.iB;
	a.RB += a.IPB * f.iB;
	double iBb = a.IRB * PB + a.IPB * RB;
	a.PE += a.IRE * f.iE;
	a.RE += a.IPE * f.iE;
	double iEb = a.IRE * PE + a.IPE * RE;
	double rpEb = iBb * f.i4, rpBb = iEb * f.i4;
	double i4b = iBb * f.rpE + iEb * f.rpB;
	double q4 = f.rpB * f.rpE, q4b = -i4b / (q4 * q4);
	rpBb += q4b * f.rpE;
	rpEb += q4b * f.rpB;
	a.RB += rpBb * PB;
	a.PB += rpBb * RB;
	a.RE += rpEb * PE;
	a.PE += rpEb * RE;
	xb[0] = a.RB;
	xb[1] = a.PB;
	xb[2] = a.UB;
	xb[3] = a.VB;
	xb[4] = a.RE;
	xb[5] = a.PE;
	xb[6] = a.UE;
	xb[7] = a.VE;
	return true;
}

// The exact Godunov flux's vector-Jacobian product.  The solver's primal
// outputs first (rim), then the flux assembly and the rotation to the normal
// frame reversed by hand, then the solver itself by rim_vjp.  When the
// Newton loop ran too long for rim_vjp, forward mode instead: the solver's
// outputs depend on six inputs (density, pressure and normal velocity of
// each side; the tangential velocities only pass through, VI is VB or VE),
// so one dual pass (rim_d1) per one of those six, the first also seeding VB,
// whose tangent in VI tells which side was sampled.  The normal's adjoint (nb
// given) comes from the rotation and the assembly alone.
__device__ __forceinline__ void godunov_vjp(const Prim& L, const Prim& R, double nx, double ny, const RimConst& K,
                                            int max_iter, const double* fb, double* Lb, double* Rb,
                                            double* nb = nullptr, double* fo = nullptr)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double utl = L.u * ny - L.v * nx, utr = R.u * ny - R.v * nx;
	double RI, EI, PI, UN, UT;
	int hit = 0;
	rim(RI, EI, PI, UN, UT, L.r, L.p, unl, utl, R.r, R.p, unr, utr, K, max_iter, hit);
	// reverse of UI = UN nx + UT ny, VI = UN ny - UT nx, f0 = RI UN,
	// f1 = f0 UI + PI nx, f2 = f0 VI + PI ny, f3 = (RI (EI + (UI^2 + VI^2)/2) + PI) UN
	double ui = UN * nx + UT * ny, vi = UN * ny - UT * nx;
	double f0 = RI * UN, h = EI + 0.5 * (ui * ui + vi * vi), A = RI * h + PI;
	if (fo) {  // the flux itself, as flux_godunov forms it
		fo[0] = f0;
		fo[1] = f0 * ui + PI * nx;
		fo[2] = f0 * vi + PI * ny;
		fo[3] = A * UN;
	}
	double f0b = fb[0] + fb[1] * ui + fb[2] * vi;
	double uib = fb[1] * f0 + fb[3] * UN * RI * ui;
	double vib = fb[2] * f0 + fb[3] * UN * RI * vi;
	double PIb = fb[1] * nx + fb[2] * ny + fb[3] * UN;
	double RIb = f0b * UN + fb[3] * UN * h;
	double EIb = fb[3] * UN * RI;
	double UNb = f0b * RI + fb[3] * A + uib * nx + vib * ny;
	double UTb = uib * ny - vib * nx;
	double nxb = fb[1] * PI + uib * UN - vib * UT;
	double nyb = fb[2] * PI + uib * UT + vib * UN;
	double xb[8];  // adjoints of (RB, PB, UB, VB, RE, PE, UE, VE)
	if (!rim_vjp(L.r, L.p, unl, utl, R.r, R.p, unr, utr, K, max_iter, RIb, EIb, PIb, UNb, UTb, xb)) {
		const double in[6] = {L.r, L.p, unl, R.r, R.p, unr};
		double x6[6], sel = 0.0;
		for (int j = 0; j < 6; j++) {
			D1 x[6];
			for (int q = 0; q < 6; q++) x[q] = D1{in[q], q == j ? 1.0 : 0.0};
			D1 dRI, dEI, dPI, dUN, dUT;
			rim_d1(dRI, dEI, dPI, dUN, dUT, x[0], x[1], x[2], D1{utl, j == 0 ? 1.0 : 0.0}, x[3], x[4], x[5],
			       D1{utr, 0.0}, K, max_iter);
			if (j == 0) sel = dUT.d;
			x6[j] = RIb * dRI.d + EIb * dEI.d + PIb * dPI.d + UNb * dUN.d;
		}
		xb[0] = x6[0];
		xb[1] = x6[1];
		xb[2] = x6[2];
		xb[3] = UTb * sel;
		xb[4] = x6[3];
		xb[5] = x6[4];
		xb[6] = x6[5];
		xb[7] = UTb * (1.0 - sel);
	}
	Lb[0] = xb[0];
	Lb[1] = xb[1];
	Lb[2] = xb[2] * nx + xb[3] * ny;
	Lb[3] = xb[2] * ny - xb[3] * nx;
	Rb[0] = xb[4];
	Rb[1] = xb[5];
	Rb[2] = xb[6] * nx + xb[7] * ny;
	Rb[3] = xb[6] * ny - xb[7] * nx;
	if (nb) {  // u_n = u nx + v ny, u_t = u ny - v nx on each side
		nb[0] = nxb + xb[2] * L.u - xb[3] * L.v + xb[6] * R.u - xb[7] * R.v;
		nb[1] = nyb + xb[2] * L.v + xb[3] * L.u + xb[6] * R.v + xb[7] * R.u;
	}
}

// adjoint accumulators of one side's state (r, p, u, v), its normal
// velocity, energy r e + r |u|^2/2 and sound speed
struct SideAdj {
	double r = 0.0, p = 0.0, u = 0.0, v = 0.0, un = 0.0, E = 0.0, c = 0.0;
};

// reverse of the Euler flux (r un, r u un + p nx, r v un + p ny, (E + p) un)
// of one side for seeds g; nxb, nyb take the normal's direct terms
__device__ __forceinline__ void euler_flux_rev(const Prim& W, double un, double E, double nx, double ny,
                                               const double* g, SideAdj& a, double& nxb, double& nyb)
{
	a.r += g[0] * un + g[1] * W.u * un + g[2] * W.v * un;
	a.un += g[0] * W.r + g[1] * W.r * W.u + g[2] * W.r * W.v + g[3] * (E + W.p);
	a.u += g[1] * W.r * un;
	a.v += g[2] * W.r * un;
	a.p += g[1] * nx + g[2] * ny + g[3] * un;
	a.E += g[3] * un;
	nxb += g[1] * W.p;
	nyb += g[2] * W.p;
}

// the state's own terms: E = p/(G - 1) + r |u|^2 / 2, c = sqrt(G p / r),
// un = u nx + v ny
__device__ __forceinline__ void side_rev(const Prim& W, double c, double nx, double ny, double G, SideAdj& a,
                                         double* Wb, double& nxb, double& nyb)
{
	double r = a.r, p = a.p, u = a.u, v = a.v;
	p += a.E / (G - 1.0);
	r += a.E * 0.5 * (W.u * W.u + W.v * W.v);
	u += a.E * W.r * W.u;
	v += a.E * W.r * W.v;
	double ab = a.c / (2.0 * c);  // c = sqrt(a), a = G p / r
	p += ab * G / W.r;
	r -= ab * (G * W.p) / (W.r * W.r);
	u += a.un * nx;
	v += a.un * ny;
	nxb += a.un * W.u;
	nyb += a.un * W.v;
	Wb[0] = r;
	Wb[1] = p;
	Wb[2] = u;
	Wb[3] = v;
}

// The HLLC flux's vector-Jacobian product (flux_hllc reversed by hand: the
// outer states' Euler fluxes, or the contact speed, star state and wave
// speeds of the star region); nb (if given) takes the normal's adjoint
__device__ __forceinline__ void hllc_vjp(const Prim& L, const Prim& R, double nx, double ny, double G, const double* fb,
                                         double* Lb, double* Rb, double* nb = nullptr, double* fo = nullptr)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double cl = sqrt(G * L.p / L.r), cr = sqrt(G * R.p / R.r);
	double El = L.p / (G - 1.0) + 0.5 * L.r * (L.u * L.u + L.v * L.v);
	double Er = R.p / (G - 1.0) + 0.5 * R.r * (R.u * R.u + R.v * R.v);
	double a1 = unl - cl, a2 = unr - cr, b1 = unl + cl, b2 = unr + cr;
	bool sl1 = a1 < a2, sr1 = b1 > b2;
	double sl = sl1 ? a1 : a2, sr = sr1 ? b1 : b2;
	SideAdj al, ar;
	double nxb = 0.0, nyb = 0.0;
	if (sl >= 0.0 || sr <= 0.0) {
		const Prim& W = sl >= 0.0 ? L : R;
		double un = sl >= 0.0 ? unl : unr, E = sl >= 0.0 ? El : Er;
		euler_flux_rev(W, un, E, nx, ny, fb, sl >= 0.0 ? al : ar, nxb, nyb);
		if (fo) {  // the flux itself, as flux_hllc forms it
			fo[0] = W.r * un;
			fo[1] = W.r * W.u * un + W.p * nx;
			fo[2] = W.r * W.v * un + W.p * ny;
			fo[3] = (E + W.p) * un;
		}
	} else {
		double ml = L.r * (sl - unl), mr = R.r * (sr - unr);
		double num = R.p - L.p + L.r * unl * (sl - unl) - R.r * unr * (sr - unr), dens = ml - mr;
		double ss = num / dens;
		bool lft = ss >= 0.0;
		const Prim& K = lft ? L : R;
		SideAdj& ak = lft ? al : ar;
		double sk = lft ? sl : sr, unk = lft ? unl : unr, Ek = lft ? El : Er, mk = lft ? ml : mr;
		double g = sk - ss, k = mk / g;
		double d = ss - unk, e = sk - unk, he = K.r * e, h = K.p / he;
		double B3 = ss + h, inner = Ek / K.r + d * B3;
		double us1 = k * (K.u + d * nx), us2 = k * (K.v + d * ny), us3 = k * inner;
		if (fo) {
			fo[0] = K.r * unk + sk * (k - K.r);
			fo[1] = K.r * K.u * unk + K.p * nx + sk * (us1 - K.r * K.u);
			fo[2] = K.r * K.v * unk + K.p * ny + sk * (us2 - K.r * K.v);
			fo[3] = (Ek + K.p) * unk + sk * (us3 - Ek);
		}
		// f = F(K) + sk (us - U(K))
		double skb = fb[0] * (k - K.r) + fb[1] * (us1 - K.r * K.u) + fb[2] * (us2 - K.r * K.v) + fb[3] * (us3 - Ek);
		double us0b = fb[0] * sk, us1b = fb[1] * sk, us2b = fb[2] * sk, us3b = fb[3] * sk;
		ak.r -= fb[0] * sk + fb[1] * sk * K.u + fb[2] * sk * K.v;
		ak.u -= fb[1] * sk * K.r;
		ak.v -= fb[2] * sk * K.r;
		ak.E -= fb[3] * sk;
		euler_flux_rev(K, unk, Ek, nx, ny, fb, ak, nxb, nyb);
		// us3 = k (Ek/Kr + d (ss + Kp / (Kr e)))
		double kb = us3b * inner, innerb = us3b * k;
		ak.E += innerb / K.r;
		ak.r -= innerb * Ek / (K.r * K.r);
		double db = innerb * B3, B3b = innerb * d;
		double ssb = B3b, hb = B3b;
		ak.p += hb / he;
		double heb = -hb * K.p / (he * he);
		ak.r += heb * e;
		double eb = heb * K.r;
		skb += eb;
		ak.un -= eb;
		// us1, us2 = k (Ku + d nx), k (Kv + d ny); us0 = k
		kb += us1b * (K.u + d * nx) + us2b * (K.v + d * ny) + us0b;
		ak.u += us1b * k;
		ak.v += us2b * k;
		db += us1b * k * nx + us2b * k * ny;
		nxb += us1b * k * d;
		nyb += us2b * k * d;
		ssb += db;
		ak.un -= db;
		// k = mk / (sk - ss)
		double mkb = kb / g, gb = -kb * mk / (g * g);
		skb += gb;
		ssb -= gb;
		// ss = (Rp - Lp + Lr unl (sl - unl) - Rr unr (sr - unr)) / (ml - mr)
		double numb = ssb / dens, densb = -ssb * num / (dens * dens);
		double mlb = densb + (lft ? mkb : 0.0), mrb = -densb + (lft ? 0.0 : mkb);
		ar.p += numb;
		al.p -= numb;
		al.r += numb * unl * (sl - unl) + mlb * (sl - unl);
		al.un += numb * L.r * (sl - unl) - numb * L.r * unl - mlb * L.r;
		double slb = numb * L.r * unl + mlb * L.r + (lft ? skb : 0.0);
		ar.r += -numb * unr * (sr - unr) + mrb * (sr - unr);
		ar.un += -numb * R.r * (sr - unr) + numb * R.r * unr - mrb * R.r;
		double srb = -numb * R.r * unr + mrb * R.r + (lft ? 0.0 : skb);
		// sl = min(unl - cl, unr - cr), sr = max(unl + cl, unr + cr)
		SideAdj& as = sl1 ? al : ar;
		as.un += slb;
		as.c -= slb;
		SideAdj& at = sr1 ? al : ar;
		at.un += srb;
		at.c += srb;
	}
	side_rev(L, cl, nx, ny, G, al, Lb, nxb, nyb);
	side_rev(R, cr, nx, ny, G, ar, Rb, nxb, nyb);
	if (nb) {
		nb[0] = nxb;
		nb[1] = nyb;
	}
}

// Reverse of the edge pass for one stage.  Seed F_e-bar = (Rb_c2 - Rb_c1) l/2
// with Rb = scale * tau/S * W; each lane takes one Gauss point: the flux VJP,
// the ghost VJP on a boundary, then its (r, p, u, v) adjoints on both sides.
// Each lane adds its boundary-parameter adjoints into its own Ebc slot.
// GEO: the geometry terms too, with the stage's flux from the flux VJP's own
// evaluation.  GONLY (LAX with geometry terms, beside k_adj_face, which forms
// the (P, G) and parameter adjoints): only the geometry terms, which need
// both sides' adjoints of the face -- no Ea or Ebc.
template <bool MM, int FLUX, bool GEO, bool GONLY = false>
__global__ __launch_bounds__(BLOCK) void k_adj_edge(Bufs b, Phys ph, AdjBufs a, double scale)
{
	static_assert(!GONLY || (GEO && FLUX == CFD2D_FLUX_LAX), "geometry-only mode: LAX with geometry terms");
	int t = xcd_block() * BLOCK + threadIdx.x;
	int e = t >> 1, gp = t & 1;
	if (e >= b.ne) return;  // both lanes of a pair leave together
	size_t ne = b.ne;
	int2 ce = edge_cells(b, e);
	int c1 = ce.x, c2 = ce.y;
	double nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
	double hl = edge_geo(b, e)[2] * 0.5;
	double fb[4];
	double s1 = scale * b.dtS[c1], s2 = c2 >= 0 ? scale * b.dtS[c2] : 0.0;
	double w1[4], w2[4] = {0, 0, 0, 0};
	ld4(a.Wn, c1, w1);
	if (c2 >= 0) ld4(a.Wn, c2, w2);
	for (int q = 0; q < 4; q++) {
		double r1 = s1 * w1[q];
		double r2 = c2 >= 0 ? s2 * w2[q] : 0.0;
		fb[q] = (r2 - r1) * hl;
	}
	Prim p1 = load_prim(b, c1);
	Prim L = p1, R;
	recon(L, b, c1, b.edl[edl_index(e, 0, gp)], b.edl[edl_index(e, 0, gp) + 1]);
	if (c2 >= 0) {
		R = load_prim(b, c2);
		recon(R, b, c2, b.edl[edl_index(e, 1, gp)], b.edl[edl_index(e, 1, gp) + 1]);
	} else {
		R = ghost_of<MM>(b, ph, e, c1, L, p1);
	}
	double Lb[4], Rb[4], nb[2] = {0.0, 0.0};
	double Fe[4];  // GEO: the stage's flux through the edge, from the VJP's own evaluation
	if (FLUX == CFD2D_FLUX_HLLC) hllc_vjp(L, R, nx, ny, ph.gam_flux, fb, Lb, Rb, GEO ? nb : nullptr, GEO ? Fe : nullptr);
	else if (FLUX == CFD2D_FLUX_GODUNOV)
		godunov_vjp(L, R, nx, ny, ph.rk, ph.max_newton, fb, Lb, Rb, GEO ? nb : nullptr, GEO ? Fe : nullptr);
	else lax_vjp(L, R, nx, ny, ph.gam_flux, ph.igm1, fb, Lb, Rb, GEO ? nb : nullptr, GEO ? Fe : nullptr);
	if (GEO)  // F_e = (f_1 + f_2) l/2, as the edge pass forms it
		for (int q = 0; q < 4; q++) Fe[q] = (Fe[q] + pair_swap(Fe[q])) * hl;
	double Rsb[4];  // adjoint of c2's reconstructed state (inner edges)
	if (c2 < 0) {
		// the ghost's inputs: the reconstructed L, the cell centre, the parameters
		int bc = -c2 - 1;
		double dL[4], dC[4], dpar[4], dn[2];
		ghost_vjp(b.bctype[bc], &b.bcpar[4 * bc], nx, ny, L, p1, mat_of<MM>(b, ph, c1), Rb, dL, dC, dpar, dn);
		for (int q = 0; q < 4; q++) {
			Lb[q] += dL[q];
			Rsb[q] = 0.0;
			Rb[q] = dC[q];
			if (!GONLY) a.Ebc[(size_t)(4 * gp + q) * ne + e] += dpar[q];
		}
		nb[0] += dn[0];
		nb[1] += dn[1];
	} else {
		for (int q = 0; q < 4; q++) Rsb[q] = Rb[q];
	}
	if (!GONLY)
		for (int q = 0; q < 4; q++) {
			a.Ea[(size_t)(4 * gp + q) * ne + e] = Lb[q];
			a.Ea[(size_t)(8 + 4 * gp + q) * ne + e] = Rb[q];
		}
	if (GEO) {
		// this lane's Gauss point: the normal, and each side's offset (the
		// reconstruction phi + g . d gives d-bar = sum_q phi-bar_q g_q)
		double g1[8], g2[8];
		ld8(b.G, c1, g1);
		double d1x = 0.0, d1y = 0.0, d2x = 0.0, d2y = 0.0;
		for (int q = 0; q < 4; q++) {
			d1x += Lb[q] * g1[2 * q];
			d1y += Lb[q] * g1[2 * q + 1];
		}
		if (c2 >= 0) {
			ld8(b.G, c2, g2);
			for (int q = 0; q < 4; q++) {
				d2x += Rsb[q] * g2[2 * q];
				d2y += Rsb[q] * g2[2 * q + 1];
			}
		}
		a.Ge[(size_t)(2 * gp) * ne + e] += nb[0];
		a.Ge[(size_t)(2 * gp + 1) * ne + e] += nb[1];
		a.Ge[(size_t)(4 + 2 * gp) * ne + e] += d1x;
		a.Ge[(size_t)(5 + 2 * gp) * ne + e] += d1y;
		a.Ge[(size_t)(8 + 2 * gp) * ne + e] += d2x;
		a.Ge[(size_t)(9 + 2 * gp) * ne + e] += d2y;
		// the stage's flux enters R_c1 with -, R_c2 with + (F_e = (f_1 + f_2) l/2):
		// each lane gives its side's W . F_e for the area term, and the first
		// the length adjoint F_e . (Rb_c2 - Rb_c1) / l = F_e . F_e-bar / (l l/2)
		double wv[4] = {0.0, 0.0, 0.0, 0.0};  // this lane's side's W, loaded again (a cache hit)
		int cw = gp ? c2 : c1;
		if (cw >= 0) ld4(a.Wn, cw, wv);
		double wf = 0.0, lb = 0.0;
		for (int q = 0; q < 4; q++) {
			wf += wv[q] * Fe[q];
			lb += Fe[q] * fb[q];
		}
		a.WF[(size_t)gp * ne + e] = wf;
		if (gp == 0) a.Ge[(size_t)GE_L * ne + e] += lb / (hl * edge_geo(b, e)[2]);
	}
}

// Cell-centric reverse of the edge pass for fluxes whose VJP is cheap (LAX):
// each cell evaluates the flux VJP of its three faces and keeps its own side,
// so every interior face is evaluated by both of its cells.  Twice the
// arithmetic of k_adj_edge, but no per-edge buffer and no second gather:
// (P, G) adjoints go straight to Pb, Gb.  Two lanes per cell, one Gauss point
// each, summed by a lane shuffle.  A cell owns its boundary faces and adds
// their parameter adjoints into Ebc (one slot per Gauss point).  The
// geometry terms, which need both sides' adjoints of a face, are formed by
// edge (k_adj_edge's geometry-only mode).
template <bool MM, int FLUX>
__global__ __launch_bounds__(BLOCK) void k_adj_face(Bufs b, Phys ph, AdjBufs a, double scale)
{
	static_assert(FLUX == CFD2D_FLUX_LAX, "the one-sided face VJP holds for the LAX flux only");
	int t = xcd_block() * BLOCK + threadIdx.x;
	int i = t >> 1, gp = t & 1;
	if (i >= b.nc) return;  // both lanes of a pair leave together
	size_t ne = b.ne;
	// the records both lanes of the pair read are loaded half by each
	double pv[4];
	pair_load<4>(b.P, i, gp, pv);
	Prim pi{pv[0], pv[1], pv[2], pv[3]};
	double wi[4];
	double si = scale * b.dtS[i];
	pair_load<4>(a.Wn, i, gp, wi);
	for (int q = 0; q < 4; q++) wi[q] *= si;
	double pb[4] = {0, 0, 0, 0};
	double gb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	double gi[8];  // this cell's gradient record, used on every face
	pair_load<8>(b.G, i, gp, gi);
	CellTop tp = load_top(b, i);
	for (int k = 0; k < 3; k++) {
		int f = tp.f[k];
		int e = f >> 1, side = f & 1;
		int j = tp.nb[k];
		double nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
		double hl = edge_geo(b, e)[2] * 0.5;
		// F_e-bar = (Rb_c2 - Rb_c1) l/2; this cell is c1 (side 0) or c2 (side 1)
		double fb[4];
		Prim pj = pi;
		double gj[8];
		if (j >= 0) {
			double pv[4];
			pair_load<4>(b.P, j, gp, pv);
			pj = Prim{pv[0], pv[1], pv[2], pv[3]};
			pair_load<8>(b.G, j, gp, gj);
			double sj = scale * b.dtS[j];
			double wv[4];
			pair_load<4>(a.Wn, j, gp, wv);
			for (int q = 0; q < 4; q++) {
				double wj = sj * wv[q];
				fb[q] = (side ? wi[q] - wj : wj - wi[q]) * hl;
			}
		} else {
			for (int q = 0; q < 4; q++) fb[q] = -wi[q] * hl;
		}
		// this cell's and the neighbour's states at the Gauss point (the
		// reconstruction of each side, as the edge pass forms it)
		double dx = b.edl[edl_index(e, side, gp)], dy = b.edl[edl_index(e, side, gp) + 1];
		Prim own = pi, L, R;
		recon_g(own, gi, dx, dy);
		if (j >= 0) {
			Prim oth = pj;
			recon_g(oth, gj, b.edl[edl_index(e, 1 - side, gp)], b.edl[edl_index(e, 1 - side, gp) + 1]);
			L = side ? oth : own;
			R = side ? own : oth;
		} else {
			L = own;
			R = ghost_of<MM>(b, ph, e, i, L, pi);
		}
		double Ob[4];  // adjoint of this cell's reconstructed state
		if (j >= 0) {
			// an inner face: this side's adjoint only
			double sg = side ? -1.0 : 1.0, fs[4];
			for (int q = 0; q < 4; q++) fs[q] = sg * fb[q];
			lax_vjp_own(own, side ? L : R, sg * nx, sg * ny, ph.gam_flux, ph.igm1, fs, side != 0, Ob);
		} else {
			// a boundary face (always the cell's own side 0): both sides, the
			// ghost's adjoint taken back to the cell through the ghost map
			double Lb[4], Rb[4];
			lax_vjp(L, R, nx, ny, ph.gam_flux, ph.igm1, fb, Lb, Rb);
			int bc = -edge_cells(b, e).y - 1;
			double dL[4], dC[4], dpar[4];
			ghost_vjp(b.bctype[bc], &b.bcpar[4 * bc], nx, ny, L, pi, mat_of<MM>(b, ph, i), Rb, dL, dC, dpar, nullptr);
			for (int q = 0; q < 4; q++) {
				Ob[q] = Lb[q] + dL[q];
				pb[q] += dC[q];
				a.Ebc[(size_t)(4 * gp + q) * ne + e] += dpar[q];
			}
		}
		for (int q = 0; q < 4; q++) {
			double o = Ob[q];
			pb[q] += o;
			gb[2 * q] += o * dx;
			gb[2 * q + 1] += o * dy;
		}
	}
	for (int q = 0; q < 4; q++) pb[q] += pair_swap(pb[q]);
	for (int q = 0; q < 8; q++) gb[q] += pair_swap(gb[q]);
	if (gp == 0) {
		sta4(a.Pb, b.nc, i, pb);
		sta8(a.Gb, b.nc, i, gb);
	}
}

// each cell gathers its side of its three edges: P adjoint = sum of the
// Gauss-point adjoints, G adjoint = sum of adjoint x Gauss-point offset
__global__ __launch_bounds__(BLOCK) void k_adj_cell(Bufs b, AdjBufs a)
{
	int i = xcd_block() * BLOCK + threadIdx.x;
	if (i >= b.nc) return;
	size_t ne = b.ne;
	double pb[4] = {0, 0, 0, 0};
	double gb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	CellTop tp = load_top(b, i);
	for (int k = 0; k < 3; k++) {
		int f = tp.f[k];
		int e = f >> 1, side = f & 1;
		for (int gp = 0; gp < 2; gp++) {
			double dx = b.edl[edl_index(e, side, gp)];
			double dy = b.edl[edl_index(e, side, gp) + 1];
			for (int q = 0; q < 4; q++) {
				double v = a.Ea[(size_t)(8 * side + 4 * gp + q) * ne + e];
				pb[q] += v;
				gb[2 * q] += v * dx;
				gb[2 * q + 1] += v * dy;
			}
		}
		if (tp.nb[k] < 0)
			for (int gp = 0; gp < 2; gp++)
				for (int q = 0; q < 4; q++) pb[q] += a.Ea[(size_t)(8 + 4 * gp + q) * ne + e];
	}
	sta4(a.Pb, b.nc, i, pb);
	sta8(a.Gb, b.nc, i, gb);
}

// Reverse of the gradient operator and of the primitive conversion, then the
// RK bookkeeping: Ws = b W + (dP/dU)^T Pb, and acc0 per MODE
// (0: Acc = a W, 1: Acc += a W, 2: Ws += Acc, the step's first stage).
template <bool MM, bool LSQ, int MODE, bool GEO>
__global__ __launch_bounds__(BLOCK) void k_adj_grad(Bufs b, Phys ph, AdjBufs a, double wa, double wb)
{
	__shared__ double lds_w[BLOCK * 5], lds_a[BLOCK * 5];
	int c0 = xcd_block() * BLOCK;
	int i = c0 + threadIdx.x;
	double ws[4] = {0, 0, 0, 0}, acc[4] = {0, 0, 0, 0};
	if (i < b.nc) {
		size_t n = b.nc, ne = b.ne;
		Prim pc = load_prim(b, i);
		MatV m = mat_of<MM>(b, ph, i);
		double pb[4];
		lda4(a.Pb, b.nc, i, pb);
		double gb[8];
		lda8(a.Gb, b.nc, i, gb);
		double isi = LSQ ? 0.0 : 1.0 / b.S[i];
		double w[4];
		ld4(a.Wn, i, w);
		double Sb = 0.0;  // GEO: the cell's area adjoint
		if (GEO) {
			// U' = ... + wb (tau/S) R with the stage residual R_i = sum over the
			// faces of -F_e (as c1) or +F_e (as c2), so d/dS (tau/S) = -(tau/S)/S
			// gives S-bar -= wb (tau/S) (W . R) / S; W . F_e per side comes from
			// the edge reverse
			double si = wb * b.dtS[i], wr = 0.0;
			int4 fc = load_faces(b, i);
			int fs[3] = {fc.x, fc.y, fc.z};
			for (int k = 0; k < 3; k++) {
				double t = a.WF[(size_t)(fs[k] & 1) * ne + (fs[k] >> 1)];
				wr += (fs[k] & 1) ? t : -t;
			}
			Sb -= si * wr / b.S[i];
			if (!LSQ) {  // Green-Gauss: G = (sum over faces) / S
				double g[8];
				ld8(b.G, i, g);
				double gg = 0.0;
				for (int q = 0; q < 8; q++) gg += gb[q] * g[q];
				Sb -= gg * isi;
			}
		}
		CellTop tp = load_top(b, i);
		for (int k = 0; k < 3; k++) {
			int f = tp.f[k];
			int e = f >> 1, side = f & 1;
			int j = tp.nb[k];
			double nbb[4];  // adjoint of the neighbour value this face contributed
			if (LSQ) {
				double wx = b.lsqw[6 * (size_t)i + 2 * k], wy = b.lsqw[6 * (size_t)i + 2 * k + 1];
				for (int q = 0; q < 4; q++) {
					nbb[q] = wx * gb[2 * q] + wy * gb[2 * q + 1];
					pb[q] -= nbb[q];
				}
			} else {
				double l = edge_geo(b, e)[2], nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
				double sg = side ? -1.0 : 1.0;
				for (int q = 0; q < 4; q++) {
					nbb[q] = sg * 0.5 * l * (nx * gb[2 * q] + ny * gb[2 * q + 1]) * isi;
					pb[q] += nbb[q];
				}
			}
			Prim pk = pc;  // GEO: the value this face contributed (neighbour or ghost)
			if (j >= 0) {
				// the neighbour's gradient depends on this cell's value
				int s = a.cslot[k * n + i];
				double gj[8];
				lda8(a.Gb, b.nc, j, gj);
				if (LSQ) {
					double wx = b.lsqw[6 * (size_t)j + 2 * s], wy = b.lsqw[6 * (size_t)j + 2 * s + 1];
					for (int q = 0; q < 4; q++) pb[q] += wx * gj[2 * q] + wy * gj[2 * q + 1];
				} else {
					double l = edge_geo(b, e)[2], nx = edge_geo(b, e)[0], ny = edge_geo(b, e)[1];
					double sg = side ? 1.0 : -1.0;  // the neighbour sits on the other side
					double isj = 1.0 / b.S[j];
					for (int q = 0; q < 4; q++) pb[q] += sg * 0.5 * l * (nx * gj[2 * q] + ny * gj[2 * q + 1]) * isj;
				}
				if (GEO) pk = load_prim(b, j);
			} else {
				// ghost of the cell-centre state across a boundary edge
				int bc = -edge_cells(b, e).y - 1;
				double dL[4], dC[4], dpar[4], dn[2];
				ghost_vjp(b.bctype[bc], &b.bcpar[4 * bc], edge_geo(b, e)[0], edge_geo(b, e)[1], pc, pc, m, nbb, dL, dC,
				          dpar, GEO ? dn : nullptr);
				for (int q = 0; q < 4; q++) {
					pb[q] += dL[q] + dC[q];
					a.Ebc[(size_t)q * ne + e] += dpar[q];
				}
				if (GEO) {
					a.Ge[e] += dn[0];  // a boundary face is its cell's own (side 0)
					a.Ge[ne + e] += dn[1];
					pk = ghost_of<MM>(b, ph, e, i, pc, pc);
				}
			}
			if (GEO) {
				// the face's weight: Green-Gauss (sign n l) takes (phi_i + phi_k)/2 / S,
				// least squares w_k takes (phi_k - phi_i)
				double vi[4] = {pc.r, pc.p, pc.u, pc.v}, vk[4] = {pk.r, pk.p, pk.u, pk.v};
				double wbx = 0.0, wby = 0.0;
				for (int q = 0; q < 4; q++) {
					double t = LSQ ? vk[q] - vi[q] : 0.5 * (vi[q] + vk[q]) * isi;
					wbx += gb[2 * q] * t;
					wby += gb[2 * q + 1] * t;
				}
				a.Gc[(size_t)(1 + 2 * k) * n + i] += wbx;
				a.Gc[(size_t)(2 + 2 * k) * n + i] += wby;
			}
		}
		if (GEO) a.Gc[i] += Sb;
		// P = (r, p, u, v)(U): the transposed Jacobian
		double r = pc.r, u = pc.u, v = pc.v, g1 = m.gam - 1.0;
		double ub[4];
		ub[0] = pb[0] - pb[2] * u / r - pb[3] * v / r + pb[1] * g1 * 0.5 * (u * u + v * v);
		ub[1] = pb[2] / r - pb[1] * g1 * u;
		ub[2] = pb[3] / r - pb[1] * g1 * v;
		ub[3] = pb[1] * g1;
		for (int q = 0; q < 4; q++) ws[q] = wb * w[q] + ub[q];
		if (MODE == 0) {
			for (int q = 0; q < 4; q++) acc[q] = wa * w[q];
		} else if (MODE == 1) {
			ld4(a.Acc, i, acc);
			for (int q = 0; q < 4; q++) acc[q] += wa * w[q];
		} else {
			ld4(a.Acc, i, acc);
			for (int q = 0; q < 4; q++) ws[q] += acc[q];
		}
	}
	int nv = min(BLOCK, b.nc - c0);
	store_records<4>(a.Ws, c0, nv, ws, lds_w);
	if (MODE != 2) store_records<4>(a.Acc, c0, nv, acc, lds_a);
}

// boundary-parameter adjoints: one workgroup per boundary object sums the
// per-edge entries of that object's run of edges, in an order fixed by the
// launch shape (strided partial sums, then a tree over the workgroup)
__global__ __launch_bounds__(BLOCK) void k_bc_reduce(Bufs b, AdjBufs a)
{
	__shared__ double part[4][BLOCK];
	size_t ne = b.ne;
	int bc = blockIdx.x, t = threadIdx.x;
	double s[4] = {0, 0, 0, 0};
	for (int k = a.bcoff[bc] + t; k < a.bcoff[bc + 1]; k += BLOCK) {
		int e = a.bedges[k];
		for (int q = 0; q < 4; q++) s[q] += a.Ebc[(size_t)q * ne + e] + a.Ebc[(size_t)(4 + q) * ne + e];
	}
	for (int q = 0; q < 4; q++) part[q][t] = s[q];
	__syncthreads();
	for (int h = BLOCK / 2; h > 0; h >>= 1) {
		if (t < h)
			for (int q = 0; q < 4; q++) part[q][t] += part[q][t + h];
		__syncthreads();
	}
	if (t < 4) a.bcbar[4 * bc + t] = part[t][0];
}

// [n][4] records -> [4][n] arrays
__global__ __launch_bounds__(BLOCK) void k_records_to_arrays(const double* rec, double* arr, int n)
{
	int i = blockIdx.x * BLOCK + threadIdx.x;
	if (i >= n) return;
	double v[4];
	ld4(rec, i, v);
	for (int q = 0; q < 4; q++) arr[(size_t)q * n + i] = v[q];
}

__global__ __launch_bounds__(BLOCK) void k_fill(double* x, size_t count, double v)
{
	size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
	if (i < count) x[i] = v;
}

// objective seed at the final state: W(i) = (sum of l n_y over its masked
// edges) * dp/dU(i), and p(i) for the objective's value
template <bool MM>
__global__ __launch_bounds__(BLOCK) void k_seed(Bufs b, Phys ph, AdjBufs a, const int* cells, const double* wsum,
                                                int K, double* pout)
{
	int k = blockIdx.x * BLOCK + threadIdx.x;
	if (k >= K) return;
	int i = cells[k];
	Prim p = load_prim(b, i);
	MatV m = mat_of<MM>(b, ph, i);
	double g1 = m.gam - 1.0, w = wsum[k];
	double v[4] = {w * g1 * 0.5 * (p.u * p.u + p.v * p.v), -w * g1 * p.u, -w * g1 * p.v, w * g1};
	st4(a.Wn, i, v);
	pout[k] = p.p;
}

// device buffers of one adjoint call, carved from the context's adjoint
// arena (kept between calls, grown when a call needs more)
struct AdjWork {
	AdjBufs a{};
	std::vector<double*> ckpt;  // state at the start of every segment but the last
	double* Pst = nullptr;      // [nrec + 1][4][nc] stage primitives
	double* Gst = nullptr;      // [nrec][8][nc] stage gradients
	double* Ufin = nullptr;     // final state
	int* seed_cells = nullptr;
	double* seed_w = nullptr;
	double* seed_p = nullptr;
};

struct Carve {
	char* base;
	size_t off = 0;
	template <typename T>
	T* take(size_t count)
	{
		off = (off + 255) & ~(size_t)255;
		T* p = base ? (T*)(base + off) : nullptr;
		off += (count ? count : 1) * sizeof(T);
		return p;
	}
};

void carve_work(AdjWork& W, Carve& cv, int nck, size_t nrec, size_t nc, size_t ne, int nbc, bool edge_buffer,
                size_t K, bool geo)
{
	size_t n4 = 4 * nc, n8 = 8 * nc;
	W.ckpt.assign(nck - 1, nullptr);
	for (auto& p : W.ckpt) p = cv.take<double>(n4);
	W.Pst = cv.take<double>((nrec + 1) * n4);
	W.Gst = cv.take<double>(nrec * n8);
	W.Ufin = cv.take<double>(n4);
	W.a.Wn = cv.take<double>(n4);
	W.a.Ws = cv.take<double>(n4);
	W.a.Acc = cv.take<double>(n4);
	W.a.Pb = cv.take<double>(n4);
	W.a.Gb = cv.take<double>(n8);
	W.a.Ea = edge_buffer ? cv.take<double>((size_t)EA * ne) : nullptr;
	W.a.Ebc = cv.take<double>(8 * ne);
	W.a.bcbar = cv.take<double>(4 * (size_t)(nbc > 0 ? nbc : 1));
	W.a.Ge = geo ? cv.take<double>(GE_N * ne) : nullptr;
	W.a.Gc = geo ? cv.take<double>(GC_N * nc) : nullptr;
	W.a.WF = geo ? cv.take<double>(2 * ne) : nullptr;
	W.seed_cells = cv.take<int>(K);
	W.seed_w = cv.take<double>(K);
	W.seed_p = cv.take<double>(K);
}

int ensure_arena(cfd2d_ctx* c, size_t bytes)
{
	if (bytes <= c->adj_ws_bytes) return 0;
	if (c->adj_ws) {
		(void)hipFree(c->adj_ws);
		c->adj_ws = nullptr;
		c->adj_ws_bytes = 0;
	}
	HIPCHK(hipMalloc(&c->adj_ws, bytes));
	c->adj_ws_bytes = bytes;
	return 0;
}

// the slot of each cell in its neighbours' face lists and the boundary edges
// in ascending order, built once per context
int adj_topology(cfd2d_ctx* c)
{
	if (c->adj_nbedges >= 0) return 0;
	int nc = c->nc, ne = c->ne;
	std::vector<int> cface(4 * (size_t)nc), cnb(4 * (size_t)nc);
	std::vector<double> eg(4 * (size_t)ne);
	HIPCHK(hipMemcpy(cface.data(), c->b.cface, cface.size() * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(cnb.data(), c->b.cnb, cnb.size() * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(eg.data(), c->b.eg, eg.size() * sizeof(double), hipMemcpyDeviceToHost));
	std::vector<int> cslot(3 * (size_t)nc, 0), bed, bcoff((size_t)c->nbc + 1, 0);
	for (int i = 0; i < nc; i++)
		for (int k = 0; k < 3; k++) {
			int j = cnb[4 * (size_t)i + k];
			if (j < 0) continue;
			int e = cface[4 * (size_t)i + k] >> 1;
			int found = -1;
			for (int s = 0; s < 3; s++)
				if ((cface[4 * (size_t)j + s] >> 1) == e) found = s;
			if (found < 0) return fail(CFD2D_ERR_MESH, "adjoint: inconsistent face lists");
			cslot[(size_t)k * nc + i] = found;
		}
	// boundary edges grouped by object (a counting sort, stable in the edge
	// index); edges of an object index at or beyond nbc carry no parameter
	std::vector<int> ebc(ne, -1);
	for (int e = 0; e < ne; e++) {
		int cells[2];
		memcpy(cells, &eg[4 * (size_t)e + 3], sizeof(cells));
		int bc = -cells[1] - 1;
		if (cells[1] < 0 && bc < c->nbc) {
			ebc[e] = bc;
			bcoff[bc + 1]++;
		}
	}
	for (int k = 0; k < c->nbc; k++) bcoff[k + 1] += bcoff[k];
	bed.resize(bcoff[c->nbc]);
	{
		std::vector<int> fill(bcoff.begin(), bcoff.end() - 1);
		for (int e = 0; e < ne; e++)
			if (ebc[e] >= 0) bed[fill[ebc[e]]++] = e;
	}
	int rc;
	if ((rc = upload(c, &c->adj_cslot, cslot)) || (rc = upload(c, &c->adj_bedges, bed)) ||
	    (rc = upload(c, &c->adj_bcoff, bcoff)))
		return rc;
	c->adj_nbedges = bcoff[c->nbc];
	return 0;
}

void launch_fill(cfd2d_ctx* c, double* x, size_t count)
{
	hipLaunchKernelGGL(k_fill, dim3((unsigned)((count + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, c->st, x, count, 0.0);
}

// nsteps forward steps writing every stage's P and G into the stage slots
// (P slot 0 = primitives of the current state); U advances in place
template <bool MM>
void record_steps(cfd2d_ctx* c, AdjWork* A, int nsteps)
{
	size_t n4 = 4 * (size_t)c->nc, n8 = 8 * (size_t)c->nc;
	double *P0 = c->b.P, *Po0 = c->b.Po, *G0 = c->b.G;
	int S = c->p.rk_stages;
	const int modes2[2] = {UPD_FIRST, UPD_RK2_LAST}, modes3[3] = {UPD_FIRST, UPD_RK3_MID, UPD_RK3_LAST};
	c->b.P = A->Pst;
	c->b.Po = A->Pst;
	launch_prim<MM>(c);
	size_t t = 0;
	for (int s = 0; s < nsteps; s++) {
		for (int st = 0; st < S; st++, t++) {
			c->b.P = A->Pst + t * n4;
			c->b.Po = A->Pst + (t + 1) * n4;
			c->b.G = A->Gst + t * n8;
			launch_gradient_edge<MM>(c);
			// the stage's gradient records: the separate gradient pass (fuse mode
			// 0; no limiter here) has stored them already, otherwise the update
			// pass forms and stores them
			c->ph.store_g = c->fuse_mode != 0;
			launch_update<MM>(c, S == 3 ? modes3[st] : modes2[st]);
			c->ph.store_g = 0;
		}
		c->b.P = A->Pst + t * n4;
		c->b.Po = c->b.P;
		hipLaunchKernelGGL((k_remediate<MM>), dim3(1), dim3(BLOCK), 0, c->st, c->b, c->ph);
	}
	c->b.P = P0;
	c->b.Po = Po0;
	c->b.G = G0;
}

template <bool MM, bool LSQ, bool GEO>
void launch_adj_grad(cfd2d_ctx* c, AdjWork* A, double wa, double wb, int mode)
{
	dim3 g(grid_of(c->nc)), blk(BLOCK);
	if (mode == 0) hipLaunchKernelGGL((k_adj_grad<MM, LSQ, 0, GEO>), g, blk, 0, c->st, c->b, c->ph, A->a, wa, wb);
	else if (mode == 1) hipLaunchKernelGGL((k_adj_grad<MM, LSQ, 1, GEO>), g, blk, 0, c->st, c->b, c->ph, A->a, wa, wb);
	else hipLaunchKernelGGL((k_adj_grad<MM, LSQ, 2, GEO>), g, blk, 0, c->st, c->b, c->ph, A->a, wa, wb);
}

// GEO: geometry adjoints too (the length and area terms take the stage's
// fluxes from the edge reverse's own flux evaluation)
template <bool MM, bool GEO>
void launch_adj_stage(cfd2d_ctx* c, AdjWork* A, double wa, double wb, int mode)
{
	if (c->p.flux == CFD2D_FLUX_HLLC || c->p.flux == CFD2D_FLUX_GODUNOV) {
		// VJPs heavy in VALU work: once per edge, through the edge buffer
		if (c->p.flux == CFD2D_FLUX_HLLC)
			hipLaunchKernelGGL((k_adj_edge<MM, CFD2D_FLUX_HLLC, GEO>), dim3(grid_of(2 * c->ne)), dim3(BLOCK), 0,
			                   c->st, c->b, c->ph, A->a, wb);
		else
			hipLaunchKernelGGL((k_adj_edge<MM, CFD2D_FLUX_GODUNOV, GEO>), dim3(grid_of(2 * c->ne)), dim3(BLOCK), 0,
			                   c->st, c->b, c->ph, A->a, wb);
		hipLaunchKernelGGL(k_adj_cell, dim3(grid_of(c->nc)), dim3(BLOCK), 0, c->st, c->b, A->a);
	} else {
		// cheap VJP: evaluated from both sides, no edge buffer; the geometry
		// terms (both sides' adjoints per edge) and the stage's fluxes by edge
		hipLaunchKernelGGL((k_adj_face<MM, CFD2D_FLUX_LAX>), dim3(grid_of(2 * c->nc)), dim3(BLOCK), 0, c->st,
		                   c->b, c->ph, A->a, wb);
		if (GEO)
			hipLaunchKernelGGL((k_adj_edge<MM, CFD2D_FLUX_LAX, true, true>), dim3(grid_of(2 * c->ne)), dim3(BLOCK), 0,
			                   c->st, c->b, c->ph, A->a, wb);
	}
	if (c->p.gradient == CFD2D_GRAD_LSQ) launch_adj_grad<MM, true, GEO>(c, A, wa, wb, mode);
	else launch_adj_grad<MM, false, GEO>(c, A, wa, wb, mode);
}

// what a geometry-sensitivity call brings back from the reverse sweep
struct GeoOut {
	const double* Ge = nullptr;  // device [GE_N][ne], in the adjoint arena
	const double* Gc = nullptr;  // device [GC_N][nc], in the adjoint arena
	std::vector<double> lny;     // per edge: the objective's p(c1) where masked (adjoint of l n_y)
};

// steps per segment when the caller leaves it to the library: the largest
// interval whose stage records fit in half of the free device memory (the
// adjoint arena of earlier calls counted as free)
int auto_interval(cfd2d_ctx* c, int nsteps)
{
	size_t fr = 0, tot = 0;
	if (hipMemGetInfo(&fr, &tot) != hipSuccess) return std::min(nsteps, 20);
	fr += c->adj_ws_bytes;  // the arena of earlier calls is reused
	double per_step = (double)c->p.rk_stages * 12.0 * 8.0 * c->nc;
	double k = (0.5 * (double)fr - 64.0 * 8.0 * c->nc - 24.0 * 8.0 * c->ne) / per_step;
	if (k < 1.0) return 1;
	return k >= nsteps ? nsteps : (int)k;
}

__global__ __launch_bounds__(BLOCK) void k_copy(double* dst, const double* src, size_t count)
{
	size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
	if (i < count) dst[i] = src[i];
}

__global__ void k_zero_int(int* x) { *x = 0; }

// device-to-device copy as a kernel (recorded into the adjoint's graphs)
void launch_copy(cfd2d_ctx* c, double* dst, const double* src, size_t count)
{
	hipLaunchKernelGGL(k_copy, dim3((unsigned)((count + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, c->st, dst, src, count);
}

// Runs enq (work on the context stream only) recorded as a graph and
// replayed: recorded again only when key changes (sizes, checkpoint
// interval, objective size, the arena's address, the step variant), so a
// repeated call with the same shape is one graph launch per pass.  With
// graphs off the work is enqueued directly.
template <typename F>
int adj_pass(cfd2d_ctx* c, cfd2d_ctx::AdjGraph& G, const std::vector<long long>& key, F enq)
{
	if (!c->adj_use_graph) {
		enq();
		return 0;
	}
	if (!G.x || G.key != key) {
		drop_adj_graph(G);
		c->capturing = true;
		hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
		if (e != hipSuccess) {
			c->capturing = false;
			return hip_fail(e, "hipStreamBeginCapture");
		}
		enq();
		e = hipStreamEndCapture(c->st, &G.g);
		c->capturing = false;
		if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
		e = hipGraphInstantiate(&G.x, G.g, nullptr, nullptr, 0);
		if (e != hipSuccess) {
			drop_adj_graph(G);
			return hip_fail(e, "hipGraphInstantiate");
		}
		G.key = key;
	}
	HIPCHK(hipGraphLaunch(G.x, c->st));
	return 0;
}

template <bool MM>
int adjoint_run(cfd2d_ctx* c, int nsteps, int ck, const int* mask, double* J, double* dJdU0, double* dJdbc,
                GeoOut* geo = nullptr)
{
	int nc = c->nc, ne = c->ne, S = c->p.rk_stages;
	size_t n4 = 4 * (size_t)nc, n8 = 8 * (size_t)nc, bytes = n4 * sizeof(double);
	int rc = adj_topology(c);
	if (rc) return rc;
	if (ck <= 0) ck = auto_interval(c, nsteps);
	ck = std::min(ck, nsteps);
	int nck = (nsteps + ck - 1) / ck;
	size_t nrec = (size_t)ck * S;
	// objective edges grouped by c1, ascending edge order within a cell
	std::vector<std::pair<int, int>> me;
	for (int e = 0; e < ne; e++)
		if (mask[e]) me.emplace_back(c->hc1[e], e);
	std::stable_sort(me.begin(), me.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
		return x.first < y.first;
	});
	std::vector<int> cells;
	std::vector<double> wsum;
	for (size_t k = 0; k < me.size(); k++) {
		if (cells.empty() || cells.back() != me[k].first) {
			cells.push_back(me[k].first);
			wsum.push_back(0.0);
		}
		wsum.back() += c->hel[me[k].second] * c->hny[me[k].second];
	}
	int K = (int)cells.size();
	AdjWork work;
	AdjWork* A = &work;
	bool edge_buffer = c->p.flux == CFD2D_FLUX_HLLC || c->p.flux == CFD2D_FLUX_GODUNOV;
	Carve sizing{nullptr};
	carve_work(work, sizing, nck, nrec, nc, ne, c->nbc, edge_buffer, K, geo != nullptr);
	if ((rc = ensure_arena(c, sizing.off))) return rc;
	Carve cv{(char*)c->adj_ws};
	carve_work(work, cv, nck, nrec, nc, ne, c->nbc, edge_buffer, K, geo != nullptr);
	AdjBufs& a = A->a;
	a.cslot = c->adj_cslot;
	a.bedges = c->adj_bedges;
	a.bcoff = c->adj_bcoff;
	a.nbedges = c->adj_nbedges;
	if (K > 0) {
		HIPCHK(hipMemcpyAsync(A->seed_cells, cells.data(), K * sizeof(int), hipMemcpyHostToDevice, c->st));
		HIPCHK(hipMemcpyAsync(A->seed_w, wsum.data(), K * sizeof(double), hipMemcpyHostToDevice, c->st));
	}
	// what the recorded passes depend on besides the context itself
	std::vector<long long> key = {nsteps, ck, K, geo != nullptr, (long long)(size_t)c->adj_ws, c->fuse_mode,
	                              c->stream_mode, c->edge_shape, c->nbc, c->adj_nbedges};
	int last0 = (nck - 1) * ck;

	// forward pass: checkpoints, then the last segment recorded in place, and
	// the seed dJ/dU at the final state (its primitives are the last P slot)
	auto forward = [&]() {
		launch_fill(c, a.Ebc, 8 * (size_t)ne);
		if (geo) {
			launch_fill(c, a.Ge, GE_N * (size_t)ne);
			launch_fill(c, a.Gc, GC_N * (size_t)nc);
		}
		hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, c->st, c->b.anyflag + 1);
		for (int s = 0; s < last0; s++) {
			if (s % ck == 0) launch_copy(c, A->ckpt[s / ck], c->b.U, n4);
			step_once(c);
		}
		record_steps<MM>(c, A, nsteps - last0);
		launch_copy(c, A->Ufin, c->b.U, n4);
		launch_fill(c, a.Wn, n4);
		if (K > 0) {
			Bufs bf = c->b;
			bf.P = A->Pst + (size_t)(nsteps - last0) * S * n4;
			hipLaunchKernelGGL((k_seed<MM>), dim3(grid_of(K)), dim3(BLOCK), 0, c->st, bf, c->ph, a, A->seed_cells,
			                   A->seed_w, K, A->seed_p);
		}
	};
	if ((rc = adj_pass(c, c->adj_fwd, key, forward))) return rc;
	std::vector<double> pk(K > 0 ? K : 1);
	int lim = 0;
	if (K > 0) HIPCHK(hipMemcpyAsync(pk.data(), A->seed_p, K * sizeof(double), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipMemcpyAsync(&lim, c->b.anyflag + 1, sizeof(int), hipMemcpyDeviceToHost, c->st));
	HIPCHK(hipStreamSynchronize(c->st));
	if ((rc = check_launch("adjoint forward pass"))) return rc;
	if (lim) return fail(CFD2D_ERR_CONFIG, "adjoint: a cell left the bounds (LIM flag) during the steps");
	{
		// ascending edge order, as cfd2d_wall_force sums it; also per case of
		// a concatenated mesh (kept for cfd2d_adjoint_objectives)
		double Jv = 0.0;
		c->adj_Jcase.assign(c->ncase, 0.0);
		for (int e = 0; e < ne; e++) {
			if (!mask[e]) continue;
			int i = c->hc1[e];
			size_t k = std::lower_bound(cells.begin(), cells.end(), i) - cells.begin();
			double t = pk[k] * c->hel[e] * c->hny[e];
			Jv += t;
			int cs = (int)(std::upper_bound(c->case_off.begin(), c->case_off.end(), i) - c->case_off.begin()) - 1;
			c->adj_Jcase[cs] += t;
		}
		if (J) *J = Jv;
		if (geo) {
			geo->lny.assign(ne, 0.0);
			for (int e = 0; e < ne; e++)
				if (mask[e])
					geo->lny[e] = pk[std::lower_bound(cells.begin(), cells.end(), c->hc1[e]) - cells.begin()];
		}
	}
	// reverse sweep, segment by segment (the last one is already recorded),
	// the boundary-parameter sums, dJ/dU0 as [4][nc] (into Ws), and the
	// context left at the final state
	const double wa2[2] = {0.0, 0.5}, wb2[2] = {1.0, 0.5};
	const double wa3[3] = {0.0, 0.75, 1.0 / 3.0}, wb3[3] = {1.0, 0.25, 2.0 / 3.0};
	// every stage swaps the roles of Wn and Ws: after nsteps S stages dJ/dU0
	// (as records) is in finW, and the transposed copy goes to finT
	bool odd = ((long long)nsteps * S) % 2 == 1;
	double* finW = odd ? a.Ws : a.Wn;
	double* finT = odd ? a.Wn : a.Ws;
	auto reverse = [&]() {
		double *P0 = c->b.P, *G0 = c->b.G;
		double *Wn = a.Wn, *Ws = a.Ws;
		for (int g = nck - 1; g >= 0; g--) {
			int s0 = g * ck, len = std::min(nsteps, s0 + ck) - s0;
			if (g != nck - 1) {
				launch_copy(c, c->b.U, A->ckpt[g], n4);
				record_steps<MM>(c, A, len);
			}
			for (int t = len * S - 1; t >= 0; t--) {
				int st = t % S;
				c->b.P = A->Pst + (size_t)t * n4;
				c->b.G = A->Gst + (size_t)t * n8;
				double wa = S == 3 ? wa3[st] : wa2[st], wb = S == 3 ? wb3[st] : wb2[st];
				int mode = st == 0 ? 2 : (st == S - 1 ? 0 : 1);
				if (geo) launch_adj_stage<MM, true>(c, A, wa, wb, mode);
				else launch_adj_stage<MM, false>(c, A, wa, wb, mode);
				std::swap(a.Wn, a.Ws);
			}
			c->b.P = P0;
			c->b.G = G0;
		}
		if (c->nbc > 0 && a.nbedges > 0)
			hipLaunchKernelGGL(k_bc_reduce, dim3(c->nbc), dim3(BLOCK), 0, c->st, c->b, a);
		// per-cell records to the API's [4][nc], transposed on the device into
		// the other W buffer (free now)
		hipLaunchKernelGGL(k_records_to_arrays, dim3(grid_of(nc)), dim3(BLOCK), 0, c->st, (const double*)finW, finT,
		                   nc);
		launch_copy(c, c->b.U, A->Ufin, n4);
		launch_prim<MM>(c);
		a.Wn = Wn;  // the host's view back to the start (a replay swaps nothing on the host)
		a.Ws = Ws;
	};
	if ((rc = adj_pass(c, c->adj_rev, key, reverse))) return rc;
	if (dJdU0) HIPCHK(hipMemcpyAsync(dJdU0, finT, bytes, hipMemcpyDeviceToHost, c->st));
	if (dJdbc && c->nbc > 0) {
		if (a.nbedges > 0)
			HIPCHK(hipMemcpyAsync(dJdbc, a.bcbar, 4 * (size_t)c->nbc * sizeof(double), hipMemcpyDeviceToHost, c->st));
		else
			std::fill(dJdbc, dJdbc + 4 * (size_t)c->nbc, 0.0);
	}
	if (geo) {
		geo->Ge = a.Ge;
		geo->Gc = a.Gc;
	}
	HIPCHK(hipStreamSynchronize(c->st));
	return check_launch("adjoint reverse sweep");
}

}  // namespace

extern "C" int cfd2d_adjoint(cfd2d_ctx* c, int nsteps, int ckpt_every, const int* edge_mask, double* J,
                             double* dJdU0, double* dJdbc)
{
	if (!c || nsteps < 1 || !edge_mask) return fail(CFD2D_ERR_ARG, "cfd2d_adjoint: bad arguments");
	if (c->p.flux == CFD2D_FLUX_ROE || c->p.limiter != CFD2D_LIMITER_NONE || c->p.viscous || c->p.steady)
		return fail(CFD2D_ERR_CONFIG, "adjoint: needs FLUX GODUNOV, LAX or HLLC, no limiter, inviscid, unsteady");
	bool timing = c->timing;
	c->timing = false;  // the adjoint's launches are not part of the step timing
	int rc = c->mm ? adjoint_run<true>(c, nsteps, ckpt_every, edge_mask, J, dJdU0, dJdbc)
	               : adjoint_run<false>(c, nsteps, ckpt_every, edge_mask, J, dJdU0, dJdbc);
	c->timing = timing;
	return rc;
}

extern "C" int cfd2d_adjoint_graph(cfd2d_ctx* c, int enable)
{
	if (!c) return fail(CFD2D_ERR_ARG, "null");
	c->adj_use_graph = enable != 0;
	drop_adj_graph(c->adj_fwd);
	drop_adj_graph(c->adj_rev);
	return CFD2D_OK;
}

extern "C" int cfd2d_adjoint_objectives(cfd2d_ctx* c, double* J_case)
{
	if (!c || !J_case) return fail(CFD2D_ERR_ARG, "cfd2d_adjoint_objectives: bad arguments");
	if ((int)c->adj_Jcase.size() != c->ncase) return fail(CFD2D_ERR_ARG, "cfd2d_adjoint_objectives: no adjoint run yet");
	std::copy(c->adj_Jcase.begin(), c->adj_Jcase.end(), J_case);
	return CFD2D_OK;
}
