// The drop-in C++ host for the explicit method: the reference's Method
// interface (src/methods/method.h:6-11) and string factory
// (src/methods/solver.cpp:13-72), with FVM_TVD running on the GPU through
// the C ABI of include/cfd2d.h.
#pragma once

#include <cstdio>
#include <exception>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cfd2d.h"

namespace cfd2d {

// Log sink of log(): stdout plus this file (the reference's hLog, opened by
// the driver before init, src/main.cpp:28).
extern FILE* hLog;
void log(const char* fmt, ...);

// Same shape as the reference's Exception (src/global.h:238-257).
class Exception {
public:
	static const int TYPE_BOUND_NOPAR = 101;
	static const int TYPE_BOUND_UNKNOWN = 102;
	static const int TYPE_MESH_WRONG_NAME = 201;
	static const int TYPE_MESH_UNV_UNKNOWN_ELEMENT = 221;
	static const int TYPE_MESH_UNV_NOT_DEFINED_BND_EDGE = 222;
	static const int FILE_OPENING_ERROR = 301;

	Exception(const std::string& msg, int t) : message(msg), type(t) {}
	const char* getMessage() const { return message.c_str(); }
	int getType() const { return type; }

private:
	std::string message;
	int type;
};

class Method {
public:
	virtual ~Method() {}
	virtual void init(char* xmlFileName) = 0;
	virtual void run() = 0;
	virtual void done() = 0;
};

struct BoundarySpec {
	int edgeType = 0;
	std::string type, name, pair;
	double par[4] = {0, 0, 0, 0};
	double shift[2] = {0, 0};
};

struct RegionSpec {
	int id = 0, matId = 0, cellType = 0;
	std::string name;
	double u = 0, v = 0, T = 0, p = 0;
};

struct MaterialSpec {
	int id = 0;
	std::string name;
	double M = 0, Cp = 0, K = 0, ML = 0;
};

// The explicit finite-volume method.  init() reads the task file with the
// reference's schema (src/methods/fvm_tvd.cpp:7-213) plus optional extension
// nodes, run() advances STEP_MAX steps or to TMAX (303-462), writing
// res_%010d.vtk and log lines at the reference's cadence, done() releases
// everything.  Errors throw cfd2d::Exception; Solver::initMethod turns them
// into log + exit(code) as the reference does.
class FVM_TVD : public Method {
public:
	~FVM_TVD() override;
	void init(char* xmlFileName) override;
	void run() override;
	void done() override;

	// for embedding hosts and tests
	cfd2d_ctx* context() const { return ctx; }
	cfd2d_mesh* mesh() const { return msh; }
	int cellCount() const { return nc; }
	double timeStep() const { return TAU; }
	void save(int step);

private:
	struct Snapshot;
	void parse(const char* xmlFileName);
	void build();
	std::shared_ptr<Snapshot> snapshot(int step);
	void writeVtk(const Snapshot& s) const;
	void saveAsync(int step);
	void joinWriter();

	double TMAX = 0, TAU = 0, CFL = 0;
	int STEP_MAX = 0, FILE_SAVE_STEP = 1, PRINT_STEP = 1;
	bool STEADY = true;
	double limitRmin = 0, limitRmax = 0, limitPmin = 0, limitPmax = 0, limitUmax = 0;
	cfd2d_params prm{};
	std::vector<MaterialSpec> materials;
	std::vector<RegionSpec> regions;
	std::vector<BoundarySpec> boundaries;
	std::string meshName, meshType, initialFile;
	cfd2d_mesh* msh = nullptr;      // the mesh the device steps (cells renumbered unless RENUMBER is NONE)
	cfd2d_mesh* fileMsh = nullptr;  // the mesh in the file's cell order (when renumbered)
	int renumberOrder = CFD2D_ORDER_HILBERT;
	std::vector<int> perm;          // device cell k is file cell perm[k] (empty: the same order)
	cfd2d_ctx* ctx = nullptr;
	int nc = 0;
	std::vector<int> cellMat;       // material of each cell in the file's order
	// output: mesh part of the VTK files (read once) and the writer thread
	std::vector<double> outXY;
	std::vector<int> outCells;
	int outNodes = 0;
	std::thread writer;
	std::exception_ptr writerError;
};

namespace Solver {
Method* initMethod(char* fileName);
void runMethod(Method* m);
void destroyMethod(Method* m);
}  // namespace Solver

}  // namespace cfd2d
