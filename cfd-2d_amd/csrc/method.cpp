#include "method.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>

#include "xml.hpp"

namespace cfd2d {

FILE* hLog = nullptr;

// log() of src/global.cpp:32-46: stdout and the task log
void log(const char* fmt, ...)
{
	va_list a, b;
	va_start(a, fmt);
	va_copy(b, a);
	vprintf(fmt, a);
	if (hLog) vfprintf(hLog, fmt, b);
	va_end(b);
	va_end(a);
	fflush(stdout);
	if (hLog) fflush(hLog);
}

namespace {

const double GAS_R = 8.314472;  // src/global.cpp:6

XmlNode* need(XmlNode* parent, const char* name)
{
	XmlNode* n = parent ? parent->first(name) : nullptr;
	if (!n) throw Exception(std::string("task file: missing <") + name + ">", 1);
	return n;
}

double attr_d(XmlNode* n, const char* key)
{
	double v;
	if (!n->attr(key, &v)) throw Exception("task file: <" + n->name + "> has no numeric '" + key + "'", 1);
	return v;
}

int attr_i(XmlNode* n, const char* key)
{
	int v;
	if (!n->attr(key, &v)) throw Exception("task file: <" + n->name + "> has no integer '" + key + "'", 1);
	return v;
}

double value_d(XmlNode* parent, const char* name) { return attr_d(need(parent, name), "value"); }

// optional <NAME value="..."/> in <control>
const char* opt_value(XmlNode* control, const char* name)
{
	XmlNode* n = control->first(name);
	return n ? n->attr("value") : nullptr;
}

int lookup(const char* v, const char* const* names, const int* codes, int n, const char* what)
{
	for (int i = 0; i < n; i++)
		if (!strcmp(v, names[i])) return codes[i];
	throw Exception(std::string("task file: unknown ") + what + " '" + v + "'", 1);
}

void check(int rc)
{
	if (rc != CFD2D_OK) throw Exception(cfd2d_last_error(), rc);
}

// Material::URS(2) then URS(1) then convertParToCons (fvm_tvd.cpp:91-92, 795-801)
void region_cons(const RegionSpec& r, const MaterialSpec& m, double* U)
{
	double Cv = m.Cp - GAS_R / m.M;
	double gam = m.Cp / Cv;
	double rho = r.p * m.M / (r.T * GAS_R);
	double e = r.p / (rho * (gam - 1));
	U[0] = rho;
	U[1] = rho * r.u;
	U[2] = rho * r.v;
	U[3] = rho * (e + 0.5 * (r.u * r.u + r.v * r.v));
}

}  // namespace

FVM_TVD::~FVM_TVD() { done(); }

void FVM_TVD::parse(const char* xmlFileName)
{
	XmlDoc doc;
	if (!doc.load_file(xmlFileName)) throw Exception("ERROR: " + doc.error(), 1);
	XmlNode* task = need(doc.root(), "task");
	XmlNode* control = need(task, "control");
	int steady = attr_i(need(control, "STEADY"), "value");
	STEADY = steady != 0;
	TAU = value_d(control, "TAU");
	TMAX = value_d(control, "TMAX");
	CFL = value_d(control, "CFL");
	STEP_MAX = attr_i(need(control, "STEP_MAX"), "value");
	FILE_SAVE_STEP = attr_i(need(control, "FILE_OUTPUT_STEP"), "value");
	PRINT_STEP = attr_i(need(control, "LOG_OUTPUT_STEP"), "value");

	memset(&prm, 0, sizeof(prm));
	prm.flux = CFD2D_FLUX_GODUNOV;
	prm.gradient = CFD2D_GRAD_GREEN_GAUSS;
	prm.limiter = CFD2D_LIMITER_NONE;
	prm.rk_stages = 2;
	prm.cfl_mode = CFD2D_CFL_REFERENCE;
	prm.max_newton = 500;  // the device loop is always bounded; hitting it is reported
	prm.prandtl = 0.72;
	if (const char* v = opt_value(control, "FLUX")) {
		static const char* n[] = {"GODUNOV", "LAX", "HLLC", "ROE"};
		static const int c[] = {CFD2D_FLUX_GODUNOV, CFD2D_FLUX_LAX, CFD2D_FLUX_HLLC, CFD2D_FLUX_ROE};
		prm.flux = lookup(v, n, c, 4, "FLUX");
	}
	if (const char* v = opt_value(control, "GRADIENT")) {
		static const char* n[] = {"GREEN_GAUSS", "GG", "LSQ", "LEAST_SQUARES"};
		static const int c[] = {0, 0, 1, 1};
		prm.gradient = lookup(v, n, c, 4, "GRADIENT");
	}
	if (const char* v = opt_value(control, "LIMITER")) {
		static const char* n[] = {"NONE", "MINMOD"};
		static const int c[] = {0, 1};
		prm.limiter = lookup(v, n, c, 2, "LIMITER");
	}
	if (const char* v = opt_value(control, "CFL_MODE")) {
		static const char* n[] = {"REFERENCE", "SPECTRAL"};
		static const int c[] = {0, 1};
		prm.cfl_mode = lookup(v, n, c, 2, "CFL_MODE");
	}
	if (const char* v = opt_value(control, "RK_STAGES")) prm.rk_stages = atoi(v);
	if (const char* v = opt_value(control, "VISCOUS")) prm.viscous = atoi(v);
	if (const char* v = opt_value(control, "PRANDTL")) prm.prandtl = atof(v);
	if (const char* v = opt_value(control, "MAX_NEWTON")) prm.max_newton = atoi(v);
	if (const char* v = opt_value(control, "RENUMBER")) {
		static const char* n[] = {"NONE", "HILBERT"};
		static const int c[] = {CFD2D_ORDER_NONE, CFD2D_ORDER_HILBERT};
		renumberOrder = lookup(v, n, c, 2, "RENUMBER");
	}

	XmlNode* lim = need(task, "limits");
	limitRmin = attr_d(need(lim, "ro"), "min");
	limitRmax = attr_d(need(lim, "ro"), "max");
	limitPmin = attr_d(need(lim, "p"), "min");
	limitPmax = attr_d(need(lim, "p"), "max");
	limitUmax = attr_d(need(lim, "u"), "max");

	XmlNode* mats = need(task, "materials");
	int matCount = attr_i(mats, "count");
	materials.clear();
	XmlNode* mn = mats->first("material");
	for (int i = 0; i < matCount; i++) {
		if (!mn) throw Exception("task file: fewer <material> than count", 1);
		MaterialSpec m;
		m.id = attr_i(mn, "id");
		m.name = need(mn, "name")->text;
		XmlNode* pp = need(mn, "parameters");
		m.M = value_d(pp, "M");
		m.Cp = value_d(pp, "Cp");
		m.K = value_d(pp, "K");
		m.ML = value_d(pp, "ML");
		materials.push_back(m);
		mn = mn->next("material");
	}
	XmlNode* regs = need(task, "regions");
	int regCount = attr_i(regs, "count");
	regions.clear();
	XmlNode* rn = regs->first("region");
	for (int i = 0; i < regCount; i++) {
		if (!rn) throw Exception("task file: fewer <region> than count", 1);
		RegionSpec r;
		r.id = attr_i(rn, "id");
		r.matId = attr_i(need(rn, "material"), "id");
		r.cellType = attr_i(need(rn, "cell"), "type");
		r.name = need(rn, "name")->text;
		XmlNode* pp = need(rn, "parameters");
		r.u = value_d(pp, "Vx");
		r.v = value_d(pp, "Vy");
		r.T = value_d(pp, "T");
		r.p = value_d(pp, "P");
		if (r.matId < 0 || r.matId >= matCount) throw Exception("task file: region material out of range", 1);
		regions.push_back(r);
		rn = rn->next("region");
	}
	// CFDBoundary::create (src/bnd_cond.cpp:10-72)
	XmlNode* bnds = need(task, "boundaries");
	boundaries.clear();
	for (XmlNode* bn = bnds->first("boundCond"); bn; bn = bn->next("boundCond")) {
		BoundarySpec b;
		bn->attr("edgeType", &b.edgeType);
		b.type = need(bn, "type")->text;
		b.name = need(bn, "name")->text;
		if (b.type == "BOUND_INLET") {
			XmlNode* pp = need(bn, "parameters");
			const char* keys[4] = {"Vx", "Vy", "T", "P"};
			for (int k = 0; k < 4; k++) {
				XmlNode* q = pp->first(keys[k]);
				if (!q)
					throw Exception(std::string("Parameter '") + keys[k] + "' isn't specified  for BOUND_INLET.",
					                Exception::TYPE_BOUND_NOPAR);
				b.par[k] = attr_d(q, "value");
			}
		} else if (b.type == "BOUND_PRESSURE_OUTLET") {
			b.par[0] = value_d(need(bn, "parameters"), "P");
		} else if (b.type == "BOUND_PERIODIC") {
			b.pair = need(bn, "pair")->text;
			XmlNode* s = need(bn, "shift");
			b.shift[0] = attr_d(s, "x");
			b.shift[1] = attr_d(s, "y");
		} else if (b.type != "BOUND_OUTLET" && b.type != "BOUND_WALL_SLIP" && b.type != "BOUND_WALL_NO_SLIP") {
			throw Exception("Unknown boundary type '" + b.type + "' specified.", Exception::TYPE_BOUND_UNKNOWN);
		}
		boundaries.push_back(b);
	}
	XmlNode* mesh = need(task, "mesh");
	const char* nm = need(mesh, "name")->attr("value");
	const char* ft = need(mesh, "filesType")->attr("value");
	if (!nm || !ft) throw Exception("task file: <mesh> needs name and filesType values", 1);
	meshName = nm;
	meshType = ft;
	if (XmlNode* ini = task->first("initial")) {
		const char* f = ini->attr("file");
		if (f) initialFile = f;
	}
}

void FVM_TVD::build()
{
	check(cfd2d_mesh_read(meshType.c_str(), meshName.c_str(), &msh));
	int ng = 0, ncg = 0, ne = 0;
	check(cfd2d_mesh_group_count(msh, 0, &ng));
	check(cfd2d_mesh_group_count(msh, 1, &ncg));
	std::vector<std::string> egn(ng), cgn(ncg);
	char buf[256];
	for (int i = 0; i < ng; i++) {
		check(cfd2d_mesh_group_name(msh, 0, i, buf, sizeof(buf)));
		egn[i] = buf;
	}
	for (int i = 0; i < ncg; i++) {
		check(cfd2d_mesh_group_name(msh, 1, i, buf, sizeof(buf)));
		cgn[i] = buf;
	}
	bool numeric = meshType == "berkeley_triangle";
	auto group_of = [&](const std::string& name) {
		for (int i = 0; i < ng; i++)
			if (egn[i] == name) return i;
		return -1;
	};
	for (auto& b : boundaries) {
		if (b.type != "BOUND_PERIODIC") continue;
		int ga = group_of(b.name), gb = group_of(b.pair);
		if (ga < 0 || gb < 0) throw Exception("periodic boundary: unknown group '" + b.name + "'/'" + b.pair + "'", 1);
		check(cfd2d_mesh_periodic(msh, ga, gb, b.shift[0], b.shift[1]));
	}
	check(cfd2d_mesh_counts(msh, nullptr, &nc, &ne));
	// the device steps the cells in Hilbert order (neighbours close in
	// memory); output and the initial-state file keep the file's order
	perm.clear();
	if (renumberOrder != CFD2D_ORDER_NONE) {
		cfd2d_mesh* r = nullptr;
		perm.assign(nc, 0);
		check(cfd2d_mesh_renumber(msh, renumberOrder, 0, perm.data(), &r));
		fileMsh = msh;
		msh = r;
	}
	// boundary objects, then group -> object (by name for UNV groups, by the
	// numeric edge type for Triangle markers; fvm_tvd.cpp:132-174)
	std::vector<int> bt;
	std::vector<double> bp;
	std::vector<const BoundarySpec*> objs;
	for (auto& b : boundaries) {
		if (b.type == "BOUND_PERIODIC") continue;
		int t = b.type == "BOUND_INLET" ? CFD2D_BND_INLET
		        : b.type == "BOUND_OUTLET" ? CFD2D_BND_OUTLET
		        : b.type == "BOUND_PRESSURE_OUTLET" ? CFD2D_BND_PRESSURE_OUTLET
		                                            : CFD2D_BND_WALL;
		bt.push_back(t);
		for (int q = 0; q < 4; q++) bp.push_back(b.par[q]);
		objs.push_back(&b);
	}
	std::vector<int> gbc(ng > 0 ? ng : 1, -1);
	for (int g = 0; g < ng; g++)
		for (size_t k = 0; k < objs.size(); k++) {
			bool hit = numeric ? (objs[k]->edgeType == atoi(egn[g].c_str())) : (objs[k]->name == egn[g]);
			if (hit) {
				gbc[g] = (int)k;
				break;
			}
		}
	// regions by cell group name (getRegionByName, fvm_tvd.cpp:773-781)
	std::vector<int> cg(nc);
	check(cfd2d_mesh_get_cell_groups(msh, cg.data()));
	std::vector<int> regOf(ncg, -1);
	for (int g = 0; g < ncg; g++)
		for (size_t r = 0; r < regions.size(); r++)
			if (regions[r].name == cgn[g]) {
				regOf[g] = (int)r;
				break;
			}
	std::vector<double> U(4 * (size_t)nc);
	std::vector<int> mat(nc, 0);  // device order
	for (int i = 0; i < nc; i++) {
		int r = cg[i] >= 0 ? regOf[cg[i]] : -1;
		if (r < 0) throw Exception("ERROR: unknown cell name '" + (cg[i] >= 0 ? cgn[cg[i]] : std::string("?")) + "'...", 1);
		const RegionSpec& R = regions[r];
		mat[i] = R.matId;
		double u[4];
		region_cons(R, materials[R.matId], u);
		for (int q = 0; q < 4; q++) U[(size_t)q * nc + i] = u[q];
	}
	if (!initialFile.empty()) {  // double[4][cells] in the file's cell order
		std::ifstream f(initialFile, std::ios::binary);
		if (!f) throw Exception("File '" + initialFile + "' opening error.", Exception::FILE_OPENING_ERROR);
		std::vector<double> Uf(U.size());
		f.read((char*)Uf.data(), Uf.size() * sizeof(double));
		if ((size_t)f.gcount() != Uf.size() * sizeof(double))
			throw Exception("initial state file is shorter than 4*cells doubles", 1);
		for (int q = 0; q < 4; q++)
			for (int k = 0; k < nc; k++)
				U[(size_t)q * nc + k] = Uf[(size_t)q * nc + (perm.empty() ? k : perm[k])];
	}
	cellMat.assign(nc, 0);
	for (int k = 0; k < nc; k++) cellMat[perm.empty() ? k : perm[k]] = mat[k];
	prm.steady = STEADY ? 1 : 0;
	prm.cfl = CFL;
	prm.tau = TAU;
	prm.gam_flux = 1.4;  // fvm_tvd.cpp:345, 398
	prm.ro_min = limitRmin;
	prm.ro_max = limitRmax;
	prm.p_min = limitPmin;
	prm.p_max = limitPmax;
	prm.u_max = limitUmax;
	if ((int)materials.size() > CFD2D_MAXMAT) throw Exception("at most 8 materials", 1);
	prm.nmat = (int)materials.size();
	for (int k = 0; k < prm.nmat; k++) {
		prm.M[k] = materials[k].M;
		prm.Cp[k] = materials[k].Cp;
		prm.K[k] = materials[k].K;
		prm.ML[k] = materials[k].ML;
	}
	check(cfd2d_ctx_create(msh, &prm, (int)bt.size(), bt.data(), bp.data(), gbc.data(), mat.data(), &ctx));
	check(cfd2d_set_state(ctx, U.data()));
	check(cfd2d_calc_tau(ctx));
	check(cfd2d_get_tau(ctx, &TAU, nullptr));
	if (!STEADY) log("time step: %25.16E\n", TAU);
}

void FVM_TVD::init(char* xmlFileName)
{
	parse(xmlFileName);
	build();
	save(0);
}

// FVM_TVD::run loop (src/methods/fvm_tvd.cpp:303-462): steps are enqueued in
// chunks up to the next output step; t advances by TAU per unsteady step.
void FVM_TVD::run()
{
	if (!ctx) throw Exception("run() before init()", 1);
	// long runs first let the autotuner pick the step variant (it restores
	// the state it trials on, and every variant gives the same results)
	if (STEP_MAX >= 200) check(cfd2d_autotune(ctx, 3, nullptr, nullptr, nullptr, nullptr));
	double t = 0.0;
	long step = 0;
	while (t < TMAX && step < STEP_MAX) {
		long next = STEP_MAX;
		if (FILE_SAVE_STEP > 0) next = std::min(next, (step / FILE_SAVE_STEP + 1) * (long)FILE_SAVE_STEP);
		if (PRINT_STEP > 0) next = std::min(next, (step / PRINT_STEP + 1) * (long)PRINT_STEP);
		long k = 0;
		double tt = t;
		while (step + k < next && tt < TMAX) {
			if (!STEADY) tt += TAU;
			k++;
		}
		check(cfd2d_step(ctx, (int)k));
		int rc = cfd2d_sync(ctx);
		if (rc != CFD2D_OK) throw Exception(cfd2d_last_error(), rc);
		step += k;
		t = tt;
		if (FILE_SAVE_STEP > 0 && step % FILE_SAVE_STEP == 0) saveAsync((int)step);
		if (PRINT_STEP > 0 && step % PRINT_STEP == 0) log("step: %ld\t\ttime step: %.16f\n", step, t);
	}
	joinWriter();  // every file is complete when run() returns
}

void FVM_TVD::done()
{
	if (writer.joinable()) writer.join();  // an error it held was reported by run()
	if (ctx) cfd2d_ctx_destroy(ctx);
	if (msh) cfd2d_mesh_destroy(msh);
	if (fileMsh) cfd2d_mesh_destroy(fileMsh);
	ctx = nullptr;
	msh = nullptr;
	fileMsh = nullptr;
}

// One output step copied to the host: the state and the per-cell time step.
struct FVM_TVD::Snapshot {
	int step = 0;
	std::vector<double> U, ctau;
};

std::shared_ptr<FVM_TVD::Snapshot> FVM_TVD::snapshot(int step)
{
	const cfd2d_mesh* om = fileMsh ? fileMsh : msh;  // files keep the mesh file's cell order
	if (outCells.empty()) {  // the mesh part of every file, read once
		int nn = 0, ne = 0;
		check(cfd2d_mesh_counts(om, &nn, nullptr, &ne));
		outCells.assign(3 * (size_t)nc, 0);
		check(cfd2d_mesh_export(om, nullptr, outCells.data(), nullptr, nullptr, nullptr, nullptr, nullptr));
		outXY.assign(2 * (size_t)nn, 0.0);
		check(cfd2d_mesh_nodes(om, outXY.data()));
		outNodes = nn;
	}
	auto s = std::make_shared<Snapshot>();
	s->step = step;
	s->U.resize(4 * (size_t)nc);
	s->ctau.resize(nc);
	check(cfd2d_get_state(ctx, s->U.data()));
	check(cfd2d_get_tau(ctx, nullptr, s->ctau.data()));
	if (!perm.empty()) {  // device order -> file order
		std::vector<double> U(s->U.size()), ct(nc);
		for (int k = 0; k < nc; k++) {
			for (int q = 0; q < 4; q++) U[(size_t)q * nc + perm[k]] = s->U[(size_t)q * nc + k];
			ct[perm[k]] = s->ctau[k];
		}
		s->U.swap(U);
		s->ctau.swap(ct);
	}
	return s;
}

void FVM_TVD::save(int step)
{
	joinWriter();
	writeVtk(*snapshot(step));
}

// The state is copied at once (one device-to-host copy); formatting and
// writing the text file, which takes far longer, runs on a writer thread
// while the device continues.  One file is in flight at a time.
void FVM_TVD::saveAsync(int step)
{
	joinWriter();
	std::shared_ptr<Snapshot> s = snapshot(step);
	writer = std::thread([this, s]() {
		try {
			writeVtk(*s);
		} catch (...) {
			writerError = std::current_exception();
		}
	});
}

void FVM_TVD::joinWriter()
{
	if (writer.joinable()) writer.join();
	if (writerError) {
		std::exception_ptr e = writerError;
		writerError = nullptr;
		std::rethrow_exception(e);
	}
}

// FVM_TVD::save (src/methods/fvm_tvd.cpp:501-600): legacy-ASCII VTK with the
// same fields, formats and line structure (including "MachNumber" =
// (u^2+v^2)/c and the single-line data blocks), so existing post-processing
// reads these files unchanged.
void FVM_TVD::writeVtk(const Snapshot& snap) const
{
	const int nn = outNodes;
	const std::vector<double>& U = snap.U;
	const std::vector<double>& ctau = snap.ctau;
	const std::vector<double>& xy = outXY;
	const std::vector<int>& cn = outCells;
	char fName[50];
	snprintf(fName, sizeof(fName), "res_%010d.vtk", snap.step);
	FILE* fp = fopen(fName, "w");
	if (!fp) throw Exception(std::string("File '") + fName + "' opening error.", Exception::FILE_OPENING_ERROR);
	fprintf(fp, "# vtk DataFile Version 2.0\n");
	fprintf(fp, "GASDIN data file\n");
	fprintf(fp, "ASCII\n");
	fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
	fprintf(fp, "POINTS %d float\n", nn);
	for (int i = 0; i < nn; i++) fprintf(fp, "%f %f %f  ", xy[2 * (size_t)i], xy[2 * (size_t)i + 1], 0.0);
	fprintf(fp, "\n");
	fprintf(fp, "CELLS %d %d\n", nc, 4 * nc);
	for (int i = 0; i < nc; i++)
		fprintf(fp, "3 %d %d %d\n", cn[3 * (size_t)i], cn[3 * (size_t)i + 1], cn[3 * (size_t)i + 2]);
	fprintf(fp, "\n");
	fprintf(fp, "CELL_TYPES %d\n", nc);
	for (int i = 0; i < nc; i++) fprintf(fp, "5\n");
	fprintf(fp, "\n");
	struct P {
		double r, p, u, v, T, cz, gam;
	};
	std::vector<P> pr(nc);
	for (int i = 0; i < nc; i++) {
		const MaterialSpec& m = materials[cellMat[i]];
		double Cv = m.Cp - GAS_R / m.M, gam = m.Cp / Cv;
		P q;
		q.r = U[i];
		q.u = U[(size_t)nc + i] / q.r;
		q.v = U[2 * (size_t)nc + i] / q.r;
		double E = U[3 * (size_t)nc + i] / q.r;
		double e = E - 0.5 * (q.u * q.u + q.v * q.v);
		q.p = q.r * e * (gam - 1);
		q.cz = sqrt(gam * q.p / q.r);
		e = q.p / (q.r * (gam - 1));
		q.T = e / Cv;
		q.gam = gam;
		pr[i] = q;
	}
	auto block = [&](const char* head, auto value) {
		fprintf(fp, "%s", head);
		for (int i = 0; i < nc; i++) {
			fprintf(fp, "%25.16f ", value(pr[i]));
			if (i + 1 == nc) fprintf(fp, "\n");
		}
	};
	fprintf(fp, "CELL_DATA %d\n", nc);
	block("SCALARS Density float 1\nLOOKUP_TABLE default\n", [](const P& q) { return q.r; });
	block("SCALARS Pressure float 1\nLOOKUP_TABLE default\n", [](const P& q) { return q.p; });
	block("SCALARS Temperature float 1\nLOOKUP_TABLE default\n", [](const P& q) { return q.T; });
	block("SCALARS MachNumber float 1\nLOOKUP_TABLE default\n", [](const P& q) { return (q.u * q.u + q.v * q.v) / q.cz; });
	fprintf(fp, "VECTORS Velosity float\n");
	for (int i = 0; i < nc; i++) {
		fprintf(fp, "%25.16f %25.16f %25.16f ", pr[i].u, pr[i].v, 0.0);
		if (i + 1 == nc) fprintf(fp, "\n");
	}
	fprintf(fp, "SCALARS Total_pressure float 1\nLOOKUP_TABLE default\n");
	for (int i = 0; i < nc; i++) {
		const P& q = pr[i];
		double agam = q.gam - 1.0;
		double M2 = (q.u * q.u + q.v * q.v) / (q.gam * q.p / q.r);
		fprintf(fp, "%f ", q.p * ::pow(1.0 + 0.5 * M2 * agam, q.gam / (q.gam - 1.0)));
		if ((i + 1) % 8 == 0 || i + 1 == nc) fprintf(fp, "\n");
	}
	fprintf(fp, "SCALARS TAU float 1\nLOOKUP_TABLE default\n");
	for (int i = 0; i < nc; i++) {
		fprintf(fp, "%25.16f ", ctau[i]);
		if (i + 1 == nc) fprintf(fp, "\n");
	}
	fclose(fp);
	printf("File '%s' saved...\n", fName);
}

// Solver::initMethod / runMethod / destroyMethod (src/methods/solver.cpp:13-72):
// <task method="FVM_TVD"> (or the explicit alias FVM_TVD_GPU) builds this
// method; a failure is logged and ends the process with its code, as the
// reference's log + exit/EXIT does.
namespace Solver {

Method* initMethod(char* fileName)
{
	XmlDoc doc;
	if (!doc.load_file(fileName)) {
		log("ERROR: %s\n", doc.error().c_str());
		exit(1);
	}
	XmlNode* task = doc.first("task");
	const char* methodName = task ? task->attr("method") : nullptr;
	if (!methodName) {
		log("ERROR: <task method=...> missing\n");
		exit(1);
	}
	Method* m = nullptr;
	if (!strcmp(methodName, "FVM_TVD") || !strcmp(methodName, "FVM_TVD_GPU")) {
		m = new FVM_TVD();
	} else {
		log("ERROR: method '%s' is not provided by this library (only FVM_TVD).\n", methodName);
		exit(1);
	}
	try {
		m->init(fileName);
	} catch (const Exception& e) {
		log("ERROR: %s\n", e.getMessage());
		exit(e.getType());
	}
	return m;
}

void runMethod(Method* m)
{
	try {
		m->run();
	} catch (const Exception& e) {
		log("ERROR: %s\n", e.getMessage());
		exit(e.getType());
	}
}

void destroyMethod(Method* m)
{
	m->done();
	delete m;
}

}  // namespace Solver

}  // namespace cfd2d
