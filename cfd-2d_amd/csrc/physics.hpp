// Device physics of the explicit finite-volume step (gfx950, FP64).
//
// Each function restates one piece of the reference and says which; the
// extensions (LAX is the implicit method's, HLLC, Roe, viscous flux,
// pressure outlet) have no counterpart in FVM_TVD and are defined here.
#pragma once

#include <hip/hip_runtime.h>

namespace cfd2d {
namespace dev {

constexpr double GAS_R = 8.314472;  // src/global.cpp:6

struct MatV {
	double M, Cv, gam, mu, kc;  // molar mass, Cv, gamma, viscosity, conductivity
};

// Square root for the device passes: the hardware reciprocal square root
// refined by two coupled Newton-Raphson steps for g = sqrt(x) and
// h = 1/(2 sqrt(x)) (Markstein), the operations the compiler's sqrt performs
// for x >= 2^-767, without its rescaling of smaller arguments (no density,
// pressure or energy here comes near 1e-231); 0 and +inf return themselves,
// a negative x gives NaN.  The explicit passes give the same bits with
// either; the implicit assembly differs by 1 ulp in places, where the
// compiler arranged the surrounding arithmetic differently.
__device__ __forceinline__ double sqrt_dev(double x)
{
	double y = __builtin_amdgcn_rsq(x);
	double g = x * y, h = y * 0.5;
	double r = fma(-h, g, 0.5);
	g = fma(g, r, g);
	double d = fma(-g, g, x);
	h = fma(h, r, h);
	g = fma(d, h, g);
	d = fma(-g, g, x);
	g = fma(d, h, g);
	return (x == 0.0 || x == INFINITY) ? x : g;
}

struct Prim {
	double r, p, u, v;
};

// convertConsToPar + Material::URS(0) (src/methods/fvm_tvd.cpp:803-813,
// src/global.cpp:15-18): u = ru/r, v = rv/r, E = re/r, e = E - |u|^2/2,
// p = r e (gamma - 1)
__device__ __forceinline__ Prim cons_to_prim(double ro, double ru, double rv, double re, const MatV& m)
{
	Prim q;
	q.r = ro;
	q.u = ru / ro;
	q.v = rv / ro;
	double E = re / ro;
	double e = E - 0.5 * (q.u * q.u + q.v * q.v);
	q.p = q.r * e * (m.gam - 1);
	return q;
}

// Material::URS(1) temperature, e = p / (r (gamma - 1)), T = e / Cv
__device__ __forceinline__ double temperature(const Prim& q, const MatV& m)
{
	double e = q.p / (q.r * (m.gam - 1));
	return e / m.Cv;
}

enum { BND_INLET = 1, BND_OUTLET = 2, BND_WALL = 3, BND_PRESSURE_OUTLET = 4 };

// FVM_TVD::boundaryCond + CFDBoundary::run (src/methods/fvm_tvd.cpp:694-728,
// src/bnd_cond.cpp:75-110): inlet sets u, v, T, p; outlet copies; walls
// mirror the normal velocity; then URS(2) gives rho = p M / (T gR) with the
// cell-centre T (Tc) of the inside state.  Only (r, p, u, v) of the ghost
// reach the flux.  Pressure outlet (extension): rho, u, v copied, p fixed
// while |u.n| < c of the cell centre (czc).
__device__ __forceinline__ Prim ghost_state(int type, const double* par, double nx, double ny, const Prim& pL,
                                            double Tc, double czc, const MatV& m)
{
	Prim g = pL;
	double T = Tc;
	if (type == BND_INLET) {
		g.u = par[0];
		g.v = par[1];
		T = par[2];
		g.p = par[3];
	} else if (type == BND_PRESSURE_OUTLET) {
		double un = pL.u * nx + pL.v * ny;
		if (fabs(un) < czc) g.p = par[0];
		return g;
	} else if (type != BND_OUTLET) {
		double Un = pL.u * nx + pL.v * ny;
		double Vx = nx * Un * 2.0;
		double Vy = ny * Un * 2.0;
		g.u = pL.u - Vx;
		g.v = pL.v - Vy;
	}
	g.r = g.p * m.M / (T * GAS_R);
	return g;
}

// rim_orig (src/global.cpp:232-405): exact Riemann solver, Newton on p*
// from the acoustic guess with the P >= 1e-5 clamp, stopping when
// |DU - F1 - F2| <= 1e-5 is seen before the update; the rarefaction/shock
// branch of each side on PK > P as in the reference.  max_iter > 0 caps the
// loop and reports it in hit; iters (if given) receives the Newton
// iterations taken (0 on the vacuum branch).  Restructured for the device:
// each side's branch is one root and one reciprocal (side_newton_a/b,
// side_star_a/b), so lanes that take different branches do not run both;
// the two sides' reciprocals come from one division (1/(a b) times b and
// a), as do 1/rho and 1/p of both sides; divisions by constants are
// products with their reciprocals.  These change the rounding of single
// operations only (results agree with the reference to round-off, not bit
// for bit); the FP64 division is the costliest operation of this solver.
// Outputs the state sampled at x/t = 0: density RI, internal energy EI,
// pressure PI, normal velocity UI, tangential velocity VI.
// The gamma-dependent constants of rim_orig (src/global.cpp:236-250),
// formed once on the host with the same expressions instead of per call.
struct RimConst {
	double GAM, AGAM, DGAM, GGAM, HGAM, FGAM, OGAM, QGAM, PGAM, RGAM, SGAM, TGAM, DGAM1, IAGAM, DGGAM, ISGAM, IRGAM,
	    IDGAM1;
	int gamma14;  // OGAM = 1/7 and 1/AGAM = 5/2 to rounding (GAM = 1.4, the reference's value)
};

__host__ __device__ inline RimConst rim_const(double GAM)
{
	RimConst k;
	k.GAM = GAM;
	k.AGAM = (GAM - 1.0);
	k.DGAM = (2.0 / k.AGAM);
	k.GGAM = (sqrt(GAM * k.AGAM));
	k.HGAM = (k.AGAM / 2.0);
	k.FGAM = (3.0 * GAM - 1.0);
	k.OGAM = (k.AGAM / (2.0 * GAM));
	k.QGAM = (GAM + 1.0);
	k.PGAM = (k.QGAM / (2.0 * GAM));
	k.RGAM = (4.0 * GAM);
	k.SGAM = (GAM * k.AGAM);
	k.TGAM = (k.QGAM / 2.0);
	k.DGAM1 = (1 + k.DGAM);
	k.IAGAM = (1 / k.AGAM);
	k.DGGAM = k.DGAM * k.GGAM;
	k.ISGAM = 1.0 / k.SGAM;
	k.IRGAM = 1.0 / k.RGAM;
	k.IDGAM1 = 1.0 / k.DGAM1;
	k.gamma14 = fabs(7.0 * k.OGAM - 1.0) < 1e-14 && fabs(k.IAGAM - 2.5) < 1e-14;
	return k;
}

// x^(1/AGAM) of the fan density (rim_orig: exp(log(x) * (1/AGAM))); x^(5/2)
// = x^2 sqrt(x) for GAM = 1.4
__device__ __forceinline__ double pow_iagam(double x, const RimConst& K)
{
	if (K.gamma14 && x >= 0.0) return x * x * sqrt_dev(x);
	return exp(log(x) * K.IAGAM);
}

// The root rim_orig takes on one side: t^OGAM on a rarefaction (t = P/PK;
// x^(1/7) for GAM = 1.4) and sqrt(t) on a shock.  For GAM = 1.4 both are one
// computation, so the lanes of a wave that take different sides of the
// reference's branch still run the same instructions: a single-precision
// seed t^(1/k) (k = 7 or 2) refined by two Newton steps on y^k = t with
// refined hardware reciprocals of y^(k-1) (quadratic convergence from the
// ~1e-7 seed to an ulp or two; one refinement of the reciprocal in the first
// step, whose error the second removes, two in the second).  Other gammas and t
// outside [1e-30, 1e30] (negative t included) take exp(log(t) OGAM) and the
// square root.
__device__ __forceinline__ double side_root(double t, bool rare, const RimConst& K)
{
	if (K.gamma14 && t > 1e-30 && t < 1e30) {
		float tf = (float)t;
		double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(tf) * (rare ? (1.0f / 7.0f) : 0.5f));
		double ik = rare ? (1.0 / 7.0) : 0.5;
		double y2 = y * y;
		double pk = rare ? y2 * y2 * y2 : y;  // y^(k-1)
		double r = __builtin_amdgcn_rcp(pk);
		r = r * (2.0 - pk * r);
		y = y + (t * r - y) * ik;
		y2 = y * y;
		pk = rare ? y2 * y2 * y2 : y;
		r = __builtin_amdgcn_rcp(pk);
		r = r * (2.0 - pk * r);
		r = r * (2.0 - pk * r);
		return y + (t * r - y) * ik;
	}
	return rare ? exp(log(t) * K.OGAM) : sqrt_dev(t);
}

// one side of the Riemann problem: density, pressure, normal velocity, sound
// speed, rho c, 1/rho, 1/p
struct RimSide {
	double R, P, U, C, RC, IR, IP;
};

// rim_orig's Newton function of one side at P (F1, FS1 or F2, FS2 of
// src/global.cpp:278-303): rarefaction (PK > P) F = DGAM (Z - C),
// F' = Z / (GAM P) with Z = C (P/PK)^OGAM; shock F = (P - PK) / Z,
// F' = (QGAM P/PK + FGAM) / (RGAM Z PKK) with Z = R C sqrt(PKK),
// PKK = PGAM P/PK + OGAM.  The two sides of the branch share one root; the
// reciprocal each needs (of GAM P or Z PKK, "den") is formed for both sides
// of the problem from one division (side_newton_b takes it).
struct SideNewton {
	bool rare;
	double x, t, y, den;
};

__device__ __forceinline__ SideNewton side_newton_a(const RimSide& s, double P, const RimConst& K)
{
	SideNewton v;
	v.x = P * s.IP;
	v.rare = s.P > P;
	v.t = v.rare ? v.x : K.PGAM * v.x + K.OGAM;
	v.y = side_root(v.t, v.rare, K);
	v.den = v.rare ? K.GAM * P : s.RC * v.y * v.t;
	return v;
}

__device__ __forceinline__ void side_newton_b(const RimSide& s, const SideNewton& v, double P, double r,
                                              const RimConst& K, double& F, double& FS)
{
	F = v.rare ? K.DGAM * (s.C * v.y - s.C) : (P - s.P) * v.t * r;
	FS = v.rare ? s.C * v.y * r : (K.QGAM * v.x + K.FGAM) * r * K.IRGAM;
}

// The star state (Rs, Es, Us) of one side at the converged P and its wave's
// head and tail speeds (src/global.cpp:310-349); sg = -1 for the left side
// (B), +1 for the right (E).  Rarefaction: Z = C (P/PK)^OGAM, E = Z^2/SGAM,
// U = UK + sg DGAM (Z - C), R = P / (AGAM E), head UK + sg C, tail U + sg Z.
// Shock: w = sqrt((TGAM P + HGAM PK)/RK), D = UK + sg w, M = RK (UK - D),
// Q = PK - P + M (UK - D), R = M^2/Q, U = D + Q/M, E = P Q / (M^2 AGAM),
// head = tail = D.  One reciprocal per side, 1/(AGAM E) or 1/(Q M) ("den"),
// the two sides' from one division (side_star_b takes it).
struct SideStar {
	bool rare;
	double y, Z, Er, D, M, Q, den;
};

__device__ __forceinline__ SideStar side_star_a(const RimSide& s, double P, double sg, const RimConst& K)
{
	SideStar v;
	v.rare = s.P > P;
	double t = v.rare ? P * s.IP : (K.TGAM * P + K.HGAM * s.P) * s.IR;
	v.y = side_root(t, v.rare, K);
	v.Z = s.C * v.y;
	v.Er = v.Z * v.Z * K.ISGAM;
	v.D = s.U + sg * v.y;
	v.M = -sg * s.R * v.y;
	v.Q = s.P - P + v.M * (-sg * v.y);
	v.den = v.rare ? K.AGAM * v.Er : v.Q * v.M;
	return v;
}

__device__ __forceinline__ void side_star_b(const RimSide& s, const SideStar& v, double P, double sg, double r,
                                            const RimConst& K, double& Rs, double& Es, double& Us, double& head,
                                            double& tail)
{
	if (v.rare) {
		Es = v.Er;
		Us = s.U + sg * K.DGAM * (v.Z - s.C);
		Rs = P * r;
		head = s.U + sg * s.C;
		tail = Us + sg * v.Z;
	} else {
		double IM = v.Q * r;  // 1/M
		Rs = v.M * v.M * v.M * r;
		Us = v.D + v.Q * IM;
		Es = P * v.Q * IM * IM * K.IAGAM;
		head = v.D;
		tail = v.D;
	}
}

__device__ __forceinline__ void rim(double& RI, double& EI, double& PI, double& UI, double& VI, double RB,
                                    double PB, double UB, double VB, double RE, double PE, double UE, double VE,
                                    const RimConst& K, int max_iter, int& hit, int* iters = nullptr)
{
	const double AGAM = K.AGAM;
	const double eps = 1.0e-5;

	double US = 0.0, UF = 0.0;
	double RF, RS, EF, ES, SBL, SFL, SSL, SEL;
	// 1/rho and 1/p of both sides from one division
	const double rpB = RB * PB, rpE = RE * PE;
	const double i4 = 1.0 / (rpB * rpE);
	const double iB = rpE * i4, iE = rpB * i4;
	RimSide B, E;
	B.R = RB; B.P = PB; B.U = UB; B.IR = PB * iB; B.IP = RB * iB;
	E.R = RE; E.P = PE; E.U = UE; E.IR = PE * iE; E.IP = RE * iE;
	B.C = sqrt_dev(K.GAM * PB * B.IR);
	E.C = sqrt_dev(K.GAM * PE * E.IR);
	B.RC = RB * B.C;
	E.RC = RE * E.C;
	double CB = B.C, CE = E.C;
	double EB = CB * CB * K.ISGAM;
	double EE = CE * CE * K.ISGAM;
	double DU = UB - UE;
	hit = 0;
	if (iters) *iters = 0;
	// the vacuum test DU < -2 (CB + CE) / AGAM, multiplied through by AGAM > 0
	// (no division; the comparison can only differ within a rounding of the
	// threshold)
	if (DU * AGAM < -2.0 * (CB + CE)) {
		RF = 0.0; RS = 0.0; EF = 0.0; ES = 0.0;
		SBL = UB - CB;
		SFL = UB + 2.0 * CB / AGAM;
		SSL = UE - 2.0 * CE / AGAM;
		SEL = UE + CE;
	} else {
		double P = (PB * E.RC + PE * B.RC + DU * B.RC * E.RC) / (B.RC + E.RC);
		int it = 0;
		for (;;) {
			if (P < eps) P = eps;
			double F1, FS1, F2, FS2;
			SideNewton nb = side_newton_a(B, P, K), ne = side_newton_a(E, P, K);
			double ir = 1.0 / (nb.den * ne.den);
			side_newton_b(B, nb, P, ne.den * ir, K, F1, FS1);
			side_newton_b(E, ne, P, nb.den * ir, K, F2, FS2);
			double res = DU - F1 - F2;
			P = P + res / (FS1 + FS2);
			++it;
			if (!(fabs(res) > eps)) break;
			if (max_iter > 0 && it >= max_iter) {
				hit = 1;
				break;
			}
		}
		if (iters) *iters = it;
		SideStar sb = side_star_a(B, P, -1.0, K), se = side_star_a(E, P, 1.0, K);
		double is = 1.0 / (sb.den * se.den);
		side_star_b(B, sb, P, -1.0, se.den * is, K, RF, EF, UF, SBL, SFL);
		side_star_b(E, se, P, 1.0, sb.den * is, K, RS, ES, US, SEL, SSL);
	}
	if (SEL <= 0.0) {
		RI = RE; EI = EE; UI = UE; VI = VE;
	} else if (SBL >= 0.0) {
		RI = RB; EI = EB; UI = UB; VI = VB;
	} else if ((SSL >= 0.0) && (SFL <= 0.0)) {
		if (US >= 0.0) {
			RI = RF; EI = EF; UI = UF; VI = VB;
		} else {
			RI = RS; EI = ES; UI = US; VI = VE;
		}
	} else if (SFL > 0.0) {
		UI = (UB + K.DGGAM * sqrt_dev(EB)) * K.IDGAM1;
		VI = VB;
		EI = (UI * UI) * K.ISGAM;
		RI = RB * pow_iagam(EI / EB, K);
	} else {
		UI = (UE - K.DGGAM * sqrt_dev(EE)) * K.IDGAM1;
		VI = VE;
		EI = (UI * UI) * K.ISGAM;
		RI = RE * pow_iagam(EI / EE, K);
	}
	PI = AGAM * EI * RI;
}

// FVM_TVD::calcFlux, Godunov branch (src/methods/fvm_tvd.cpp:602-622)
__device__ __forceinline__ void flux_godunov(double* f, const Prim& L, const Prim& R, double nx, double ny,
                                             const RimConst& K, int max_iter, int& hit)
{
	double unl = L.u * nx + L.v * ny;
	double unr = R.u * nx + R.v * ny;
	double utl = L.u * ny - L.v * nx;
	double utr = R.u * ny - R.v * nx;
	double RI, EI, PI, UN, UT;
	rim(RI, EI, PI, UN, UT, L.r, L.p, unl, utl, R.r, R.p, unr, utr, K, max_iter, hit);
	double UI = UN * nx + UT * ny;
	double VI = UN * ny - UT * nx;
	f[0] = RI * UN;
	f[1] = f[0] * UI + PI * nx;
	f[2] = f[0] * VI + PI * ny;
	f[3] = (RI * (EI + 0.5 * (UI * UI + VI * VI)) + PI) * UN;
}

// FLUX=LAX as FVM_TVD_IMPLICIT::calcFlux defines it (src/methods/fvm_tvd_implicit.cpp:939-960)
__device__ __forceinline__ void flux_lax(double* f, const Prim& L, const Prim& R, double nx, double ny, double GAM)
{
	double unl = L.u * nx + L.v * ny;
	double unr = R.u * nx + R.v * ny;
	double al = fabs(unl) + sqrt_dev(GAM * L.p / L.r);
	double ar = fabs(unr) + sqrt_dev(GAM * R.p / R.r);
	double alpha = al > ar ? al : ar;
	double rol = L.r, rul = L.r * L.u, rvl = L.r * L.v;
	double rel = L.p / (GAM - 1.0) + 0.5 * L.r * (L.u * L.u + L.v * L.v);
	double ror = R.r, rur = R.r * R.u, rvr = R.r * R.v;
	double rer = R.p / (GAM - 1.0) + 0.5 * R.r * (R.u * R.u + R.v * R.v);
	double frl = rol * unl;
	double frr = ror * unr;
	f[0] = 0.5 * (frr + frl - alpha * (ror - rol));
	f[1] = 0.5 * (frr * R.u + frl * L.u + (R.p + L.p) * nx - alpha * (rur - rul));
	f[2] = 0.5 * (frr * R.v + frl * L.v + (R.p + L.p) * ny - alpha * (rvr - rvl));
	f[3] = 0.5 * ((rer + R.p) * unr + (rel + L.p) * unl - alpha * (rer - rel));
}

// The implicit method's ghost, FVM_TVD_IMPLICIT::boundaryCond
// (src/methods/fvm_tvd_implicit.cpp:963-1008): inlet takes T, p, u, v from
// the parameters (held here in the explicit order u, v, T, p) and rho =
// p M / (T gR) (URS(2)); outlet copies the inside state; walls copy it with
// the normal velocity mirrored.  Unlike the explicit ghost, rho is not
// recomputed from the cell-centre temperature.
__device__ __forceinline__ Prim ghost_implicit(int type, const double* par, double nx, double ny, const Prim& pL,
                                               const MatV& m)
{
	Prim g = pL;
	if (type == BND_INLET) {
		g.u = par[0];
		g.v = par[1];
		g.p = par[3];
		g.r = g.p * m.M / (par[2] * GAS_R);
	} else if (type != BND_OUTLET) {
		double Un = pL.u * nx + pL.v * ny;
		g.u = pL.u - nx * Un * 2.0;
		g.v = pL.v - ny * Un * 2.0;
	}
	return g;
}

// roe_orig (src/global.cpp:407-441) without the third velocity, with c and
// H as FVM_TVD_IMPLICIT::calcRoeAverage forms them (fvm_tvd_implicit.cpp:
// 408-429): Roe averages of u, v and the total enthalpy, p from H, c =
// sqrt(GAM p / r), H = e + q^2/2 + p/r.  Each side comes with sqrt(rho) and
// its reciprocal (a cell's are formed once and shared by its faces) and the
// divisions of the reference's expressions become products with
// reciprocals: the results differ from them by rounding only.
struct RoeSide {
	double s, is;  // sqrt_dev(rho), 1/sqrt_dev(rho)
};

__device__ __forceinline__ RoeSide roe_side(double r)
{
	RoeSide x;
	x.s = sqrt_dev(r);
	x.is = 1.0 / x.s;
	return x;
}

struct RoeAvg {
	double u, v, c, ic, H;  // ic = 1/c
};

__device__ __forceinline__ RoeAvg roe_average(const Prim& L, RoeSide sl, const Prim& R, RoeSide sr, double GAM)
{
	double AGAM = GAM - 1.0;
	double kH = GAM / AGAM;  // e + p/r = (p/r) GAM/(GAM-1)
	double s_ = 1.0 / (sl.s + sr.s);
	RoeAvg a;
	a.u = (sl.s * L.u + sr.s * R.u) * s_;
	a.v = (sl.s * L.v + sr.s * R.v) * s_;
	double HB = L.p * (sl.is * sl.is) * kH + (L.u * L.u + L.v * L.v) * 0.5;
	double HE = R.p * (sr.is * sr.is) * kH + (R.u * R.u + R.v * R.v) * 0.5;
	double HI = (sl.s * HB + sr.s * HE) * s_;
	double q2 = a.u * a.u + a.v * a.v;
	double r = sl.s * sr.s;
	double p = (HI - q2 * 0.5) * r * (AGAM / GAM);
	double por = p * (sl.is * sr.is);  // p / r
	a.c = sqrt_dev(GAM * por);
	a.ic = 1.0 / a.c;
	a.H = por * kH + 0.5 * q2;
	return a;
}

// A+ and A- of FVM_TVD_IMPLICIT (eigenValues, rightEigenVector,
// leftEigenVector, calcAP, calcAM: fvm_tvd_implicit.cpp:335-406):
// A+- = 0.5 R diag(max/min(lambda, 0)) L at the Roe state for normal n, the
// reference's factor 0.5 included, each times l.  A+ is added to diag, A- is
// written to am (both row-major 4x4); the sums over k of (R[i][k]
// lambda_k) L[k][j] run in the reference's order (multMtx4,
// fvm_tvd_implicit.cpp:264-277).
__device__ __forceinline__ void split_face(const RoeAvg& a, double nx, double ny, double GAM, double l, double* diag,
                                           double* am)
{
	double u = a.u, v = a.v, c = a.c, H = a.H;
	double qn = u * nx + v * ny, q2 = u * u + v * v;
	double lx = -ny, ly = nx, ql = u * lx + v * ly;
	double gc = (GAM - 1.0) * (a.ic * a.ic);  // g1 / c^2
	double h = 0.5 * q2 * gc, qc = qn * a.ic, gu = u * gc, xc = nx * a.ic, gv = v * gc, yc = ny * a.ic;
	double L[4][4] = {{0.5 * (h + qc), -0.5 * (gu + xc), -0.5 * (gv + yc), 0.5 * gc},
	                  {1 - h, gu, gv, -gc},
	                  {0.5 * (h - qc), -0.5 * (gu - xc), -0.5 * (gv - yc), 0.5 * gc},
	                  {-ql, lx, ly, 0}};
	double lam[4] = {qn - c, qn, qn + c, qn};
	double R[4][4] = {{1.0, 1.0, 1.0, 0.0},
	                  {u - c * nx, u, u + c * nx, lx},
	                  {v - c * ny, v, v + c * ny, ly},
	                  {H - qn * c, q2 / 2, H + qn * c, ql}};
	double hl = 0.5 * l;
#pragma unroll
	for (int i = 0; i < 4; i++) {
		double wp[4], wm[4];
		for (int k = 0; k < 4; k++) {
			wp[k] = R[i][k] * (lam[k] > 0 ? lam[k] : 0);
			wm[k] = R[i][k] * (lam[k] < 0 ? lam[k] : 0);
		}
		for (int j = 0; j < 4; j++) {
			double sp = 0.0, sm = 0.0;
			for (int k = 0; k < 4; k++) {
				sp += wp[k] * L[k][j];
				sm += wm[k] * L[k][j];
			}
			diag[4 * i + j] += sp * hl;
			am[4 * i + j] = sm * hl;
		}
	}
}

// Extension: HLLC (Toro, Riemann Solvers, 3rd ed., 10.4-10.6), Davis wave speeds
__device__ __forceinline__ void flux_hllc(double* f, const Prim& L, const Prim& R, double nx, double ny, double GAM)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double cl = sqrt_dev(GAM * L.p / L.r), cr = sqrt_dev(GAM * R.p / R.r);
	double El = L.p / (GAM - 1.0) + 0.5 * L.r * (L.u * L.u + L.v * L.v);
	double Er = R.p / (GAM - 1.0) + 0.5 * R.r * (R.u * R.u + R.v * R.v);
	double sl = fmin(unl - cl, unr - cr);
	double sr = fmax(unl + cl, unr + cr);
	if (sl >= 0.0) {
		f[0] = L.r * unl;
		f[1] = L.r * L.u * unl + L.p * nx;
		f[2] = L.r * L.v * unl + L.p * ny;
		f[3] = (El + L.p) * unl;
		return;
	}
	if (sr <= 0.0) {
		f[0] = R.r * unr;
		f[1] = R.r * R.u * unr + R.p * nx;
		f[2] = R.r * R.v * unr + R.p * ny;
		f[3] = (Er + R.p) * unr;
		return;
	}
	double ml = L.r * (sl - unl), mr = R.r * (sr - unr);
	double ss = (R.p - L.p + L.r * unl * (sl - unl) - R.r * unr * (sr - unr)) / (ml - mr);
	const Prim& K = ss >= 0.0 ? L : R;
	double sk = ss >= 0.0 ? sl : sr;
	double unk = ss >= 0.0 ? unl : unr;
	double Ek = ss >= 0.0 ? El : Er;
	double k = (ss >= 0.0 ? ml : mr) / (sk - ss);
	double us0 = k, us1 = k * (K.u + (ss - unk) * nx), us2 = k * (K.v + (ss - unk) * ny);
	double us3 = k * (Ek / K.r + (ss - unk) * (ss + K.p / (K.r * (sk - unk))));
	double fk0 = K.r * unk, fk1 = K.r * K.u * unk + K.p * nx, fk2 = K.r * K.v * unk + K.p * ny,
	       fk3 = (Ek + K.p) * unk;
	f[0] = fk0 + sk * (us0 - K.r);
	f[1] = fk1 + sk * (us1 - K.r * K.u);
	f[2] = fk2 + sk * (us2 - K.r * K.v);
	f[3] = fk3 + sk * (us3 - Ek);
}

// Extension: Roe flux on Roe averages (as roe_orig forms them, src/global.cpp:407-441)
// with Harten's entropy fix, delta = 0.1 c
__device__ __forceinline__ void flux_roe(double* f, const Prim& L, const Prim& R, double nx, double ny, double GAM)
{
	double unl = L.u * nx + L.v * ny, unr = R.u * nx + R.v * ny;
	double El = L.p / (GAM - 1.0) + 0.5 * L.r * (L.u * L.u + L.v * L.v);
	double Er = R.p / (GAM - 1.0) + 0.5 * R.r * (R.u * R.u + R.v * R.v);
	double Hl = (El + L.p) / L.r, Hr = (Er + R.p) / R.r;
	double sl = sqrt_dev(L.r), sr = sqrt_dev(R.r), w = 1.0 / (sl + sr);
	double u = (sl * L.u + sr * R.u) * w, v = (sl * L.v + sr * R.v) * w;
	double H = (sl * Hl + sr * Hr) * w;
	double q2 = u * u + v * v;
	double c = sqrt_dev((GAM - 1.0) * (H - 0.5 * q2));
	double un = u * nx + v * ny, ut = -u * ny + v * nx;
	double dr = R.r - L.r, dp = R.p - L.p;
	double dun = unr - unl;
	double dut = (-R.u * ny + R.v * nx) - (-L.u * ny + L.v * nx);
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
	double d0 = l1 * a1 + l2 * a2 + l4 * a4;
	double d1 = l1 * a1 * (u - c * nx) + l2 * a2 * u + l2 * a3 * tx + l4 * a4 * (u + c * nx);
	double d2 = l1 * a1 * (v - c * ny) + l2 * a2 * v + l2 * a3 * ty + l4 * a4 * (v + c * ny);
	double d3 = l1 * a1 * (H - un * c) + l2 * a2 * 0.5 * q2 + l2 * a3 * ut + l4 * a4 * (H + un * c);
	f[0] = 0.5 * (L.r * unl + R.r * unr - d0);
	f[1] = 0.5 * ((L.r * L.u * unl + L.p * nx) + (R.r * R.u * unr + R.p * nx) - d1);
	f[2] = 0.5 * ((L.r * L.v * unl + L.p * ny) + (R.r * R.v * unr + R.p * ny) - d2);
	f[3] = 0.5 * ((El + L.p) * unl + (Er + R.p) * unr - d3);
}

}  // namespace dev
}  // namespace cfd2d
