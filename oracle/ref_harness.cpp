// Golden-vector and timing harness around the reference FVM_TVD method.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
// reference's own sources (src/global.cpp, src/bnd_cond.cpp,
// src/methods/fvm_tvd.cpp, src/mesh/*.cpp, src/tinyxml/*.cpp) by
// oracle/Makefile into oracle/_ref/ref_harness; no reference source is
// copied here.  The reference members it reads are private/protected
// (src/methods/fvm_tvd.h:13-114, src/methods/method.h:129-134), so every
// translation unit of that build is compiled with -Dprivate=public
// -Dprotected=public (layout-neutral).
//
// Modes (all write raw little-endian arrays into <outdir>):
//   golden <task.xml> <outdir> <steps...>
//       geometry, time step, U0, Green-Gauss gradients of U0
//       (fvm_tvd.cpp:242-301), per-edge per-Gauss-point reconstruction and
//       Godunov flux (646-691, 602-622), the stage-1 residual (329-365) and
//       the state after each cumulative step count given (run(), 303-462).
//   rim <in.bin> <out.bin>
//       rim_orig (src/global.cpp:232-405) on rows of
//       (rb, pb, ub, vb, re, pe, ue, ve, gam) -> (RI, EI, PI, UI, VI, WI).
//   time <task.xml> <steps>
//       wall time of run() for <steps> steps; prints one JSON line.
#include "fvm_tvd.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

static void put(const std::string& dir, const char* name, const void* p, size_t bytes)
{
	std::string path = dir + "/" + name;
	FILE* f = fopen(path.c_str(), "wb");
	if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(3); }
	if (bytes) fwrite(p, 1, bytes, f);
	fclose(f);
}

static double now_s()
{
	timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

struct Harness : public FVM_TVD {
	void zero_flags()
	{
		// Cell::flag is left uninitialised by the UNV reader.
		for (int i = 0; i < grid.cCount; i++) grid.cells[i].flag = 0;
	}

	void set_steps(int n, double tmax)
	{
		STEP_MAX = n;
		TMAX = tmax;
		FILE_SAVE_STEP = 0x7fffffff;
		PRINT_STEP = 0x7fffffff;
	}

	void dump_state(const std::string& dir, const char* name)
	{
		int nc = grid.cCount;
		std::vector<double> u(4 * (size_t)nc);
		for (int i = 0; i < nc; i++) {
			u[i] = ro[i];
			u[nc + i] = ru[i];
			u[2 * nc + i] = rv[i];
			u[3 * nc + i] = re[i];
		}
		put(dir, name, u.data(), u.size() * 8);
	}

	void dump_flags(const std::string& dir, const char* name)
	{
		std::vector<unsigned int> f(grid.cCount);
		for (int i = 0; i < grid.cCount; i++) f[i] = grid.cells[i].flag;
		put(dir, name, f.data(), f.size() * 4);
	}

	void dump_geometry(const std::string& dir)
	{
		int nc = grid.cCount, ne = grid.eCount;
		std::vector<double> cell(4 * (size_t)nc);       // S, cx, cy, (unused 0)
		std::vector<int> cnodes(3 * (size_t)nc), cedges(3 * (size_t)nc);
		for (int i = 0; i < nc; i++) {
			Cell& c = grid.cells[i];
			cell[4 * i + 0] = c.S;
			cell[4 * i + 1] = c.c.x;
			cell[4 * i + 2] = c.c.y;
			cell[4 * i + 3] = 0.0;
			for (int k = 0; k < 3; k++) {
				cnodes[3 * i + k] = c.nodesInd[k];
				cedges[3 * i + k] = c.edgesInd[k];
			}
		}
		std::vector<double> edge(9 * (size_t)ne);      // nx, ny, l, c0x, c0y, c1x, c1y, c2x, c2y
		std::vector<int> eint(5 * (size_t)ne);         // n1, n2, c1, c2, type
		for (int e = 0; e < ne; e++) {
			Edge& g = grid.edges[e];
			edge[9 * e + 0] = g.n.x;
			edge[9 * e + 1] = g.n.y;
			edge[9 * e + 2] = g.l;
			for (int k = 0; k < 3; k++) {
				edge[9 * e + 3 + 2 * k] = (k < g.cCount) ? g.c[k].x : 0.0;
				edge[9 * e + 4 + 2 * k] = (k < g.cCount) ? g.c[k].y : 0.0;
			}
			eint[5 * e + 0] = g.n1;
			eint[5 * e + 1] = g.n2;
			eint[5 * e + 2] = g.c1;
			eint[5 * e + 3] = g.c2;
			eint[5 * e + 4] = g.type;
		}
		int counts[3] = {grid.nCount, nc, ne};
		put(dir, "counts.i32", counts, sizeof(counts));
		put(dir, "cell_geom.f64", cell.data(), cell.size() * 8);
		put(dir, "cell_nodes.i32", cnodes.data(), cnodes.size() * 4);
		put(dir, "cell_edges.i32", cedges.data(), cedges.size() * 4);
		put(dir, "edge_geom.f64", edge.data(), edge.size() * 8);
		put(dir, "edge_int.i32", eint.data(), eint.size() * 4);
	}

	void dump_gradients(const std::string& dir)
	{
		int nc = grid.cCount;
		calcGrad();
		std::vector<double> g(8 * (size_t)nc);     // rx ry px py ux uy vx vy (SoA)
		for (int i = 0; i < nc; i++) {
			g[0 * nc + i] = gradR[i].x; g[1 * nc + i] = gradR[i].y;
			g[2 * nc + i] = gradP[i].x; g[3 * nc + i] = gradP[i].y;
			g[4 * nc + i] = gradU[i].x; g[5 * nc + i] = gradU[i].y;
			g[6 * nc + i] = gradV[i].x; g[7 * nc + i] = gradV[i].y;
		}
		put(dir, "grad0.f64", g.data(), g.size() * 8);
	}

	// Per edge and Gauss point: reconstructed (r,p,u,v) left and right and
	// the Godunov flux; then the residual accumulated in edge order.
	void dump_edge_pass(const std::string& dir)
	{
		int nc = grid.cCount, ne = grid.eCount;
		calcGrad();
		std::vector<double> rec(2 * 8 * (size_t)ne), flx(2 * 4 * (size_t)ne), fe(4 * (size_t)ne);
		std::vector<double> res(4 * (size_t)nc, 0.0), absf(4 * (size_t)nc, 0.0);
		for (int e = 0; e < ne; e++) {
			Edge& g = grid.edges[e];
			double acc[4] = {0, 0, 0, 0};
			for (int ig = 1; ig < g.cCount && ig <= 2; ig++) {
				Param pL, pR;
				reconstruct(e, pL, pR, g.c[ig]);
				double f[4];
				calcFlux(f[0], f[1], f[2], f[3], pL, pR, g.n, 1.4);
				double* r = &rec[(size_t)(2 * e + ig - 1) * 8];
				r[0] = pL.r; r[1] = pL.p; r[2] = pL.u; r[3] = pL.v;
				r[4] = pR.r; r[5] = pR.p; r[6] = pR.u; r[7] = pR.v;
				for (int k = 0; k < 4; k++) {
					flx[(size_t)(2 * e + ig - 1) * 4 + k] = f[k];
					acc[k] += f[k];
				}
			}
			double hl = g.l * 0.5;
			for (int k = 0; k < 4; k++) {
				double fl = acc[k] * hl;
				fe[(size_t)e * 4 + k] = fl;
				res[(size_t)k * nc + g.c1] -= fl;
				absf[(size_t)k * nc + g.c1] += fabs(fl);
				if (g.c2 > -1) {
					res[(size_t)k * nc + g.c2] += fl;
					absf[(size_t)k * nc + g.c2] += fabs(fl);
				}
			}
		}
		put(dir, "recon0.f64", rec.data(), rec.size() * 8);
		put(dir, "gpflux0.f64", flx.data(), flx.size() * 8);
		put(dir, "edgeflux0.f64", fe.data(), fe.size() * 8);
		put(dir, "residual0.f64", res.data(), res.size() * 8);
		put(dir, "residual0_scale.f64", absf.data(), absf.size() * 8);
	}

	void dump_tau(const std::string& dir)
	{
		std::vector<double> t(grid.cCount + 1);
		t[0] = TAU;
		for (int i = 0; i < grid.cCount; i++) t[i + 1] = cTau[i];
		put(dir, "tau.f64", t.data(), t.size() * 8);
	}
};

static int mode_golden(int argc, char** argv)
{
	if (argc < 4) { fprintf(stderr, "golden <task.xml> <outdir> <steps...>\n"); return 2; }
	std::string dir = argv[3];
	Harness m;
	m.init(argv[2]);
	m.zero_flags();
	m.dump_geometry(dir);
	m.dump_tau(dir);
	m.dump_state(dir, "u0.f64");
	m.dump_gradients(dir);
	m.dump_edge_pass(dir);
	int done = 0;
	for (int a = 4; a < argc; a++) {
		int target = atoi(argv[a]);
		if (target > done) {
			m.set_steps(target - done, 1.0e300);
			m.run();
			done = target;
		}
		char name[64];
		snprintf(name, sizeof(name), "u_step%d.f64", target);
		m.dump_state(dir, name);
		snprintf(name, sizeof(name), "flags_step%d.u32", target);
		m.dump_flags(dir, name);
	}
	return 0;
}

static int mode_rim(int argc, char** argv)
{
	if (argc < 4) { fprintf(stderr, "rim <in.bin> <out.bin>\n"); return 2; }
	FILE* f = fopen(argv[2], "rb");
	if (!f) return 3;
	std::vector<double> in;
	double buf[9];
	while (fread(buf, 8, 9, f) == 9) in.insert(in.end(), buf, buf + 9);
	fclose(f);
	size_t n = in.size() / 9;
	std::vector<double> out(6 * n);
	for (size_t i = 0; i < n; i++) {
		const double* r = &in[9 * i];
		double* o = &out[6 * i];
		rim_orig(o[0], o[1], o[2], o[3], o[4], o[5], r[0], r[1], r[2], r[3], 0.0,
		         r[4], r[5], r[6], r[7], 0.0, r[8]);
	}
	FILE* g = fopen(argv[3], "wb");
	if (!g) return 3;
	fwrite(out.data(), 8, out.size(), g);
	fclose(g);
	return 0;
}

static int mode_time(int argc, char** argv)
{
	if (argc < 4) { fprintf(stderr, "time <task.xml> <steps>\n"); return 2; }
	Harness m;
	m.init(argv[2]);
	m.zero_flags();
	int steps = atoi(argv[3]);
	m.set_steps(steps, 1.0e300);
	double t0 = now_s();
	m.run();
	double t1 = now_s();
	double cu = (double)m.grid.cCount * steps / (t1 - t0);
	printf("{\"cells\": %d, \"steps\": %d, \"seconds\": %.6f, \"cell_updates_per_s\": %.6e}\n",
	       m.grid.cCount, steps, t1 - t0, cu);
	return 0;
}

int main(int argc, char** argv)
{
	hLog = fopen("task.log", "w");
	if (!hLog) hLog = fopen("/dev/null", "w");
	if (argc < 2) { fprintf(stderr, "usage: ref_harness golden|rim|time ...\n"); return 2; }
	if (!strcmp(argv[1], "golden")) return mode_golden(argc, argv);
	if (!strcmp(argv[1], "rim")) return mode_rim(argc, argv);
	if (!strcmp(argv[1], "time")) return mode_time(argc, argv);
	fprintf(stderr, "unknown mode %s\n", argv[1]);
	return 2;
}
