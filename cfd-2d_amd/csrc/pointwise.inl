// The Riemann solver and the numerical fluxes evaluated on given state
// pairs, outside any mesh: FVM_TVD::calcFlux (src/methods/fvm_tvd.cpp:
// 602-622) and rim_orig (src/global.cpp:232-405) by themselves, so that the
// device code can be checked row by row against the reference's outputs on
// a corpus of states (vacuum, strong shocks, Toro's tests) that no mesh case
// reaches.  Included at the end of solver.hip.
//
// Staging: the caller's rows are copied into device memory, one thread per
// row evaluates, and the outputs are copied back.  Layout of the staging
// block for n rows (offsets in doubles, each part 32-byte aligned):
//   rows  [n][RW]   inputs (RW = 8 for the Riemann solver, 10 for a flux)
//   out   [n][OW]   outputs (OW = 5 or 4)
//   its   [n]       Newton iterations per row (ints, Riemann solver only)
//   cnt   [1]       rows whose Newton loop hit the cap (unsigned long long)

namespace {

constexpr int PW_RIM = -1;  // kind of the Riemann-solver evaluation (fluxes use CFD2D_FLUX_*)

struct PwLayout {
	size_t rows, out, its, cnt, total;  // offsets and total size in doubles
};

PwLayout pw_layout(int kind, size_t n)
{
	auto up = [](size_t x) { return (x + 3) & ~(size_t)3; };
	size_t rw = kind == PW_RIM ? 8 : 10, ow = kind == PW_RIM ? 5 : 4;
	PwLayout L;
	L.rows = 0;
	L.out = up(n * rw);
	L.its = L.out + up(n * ow);
	L.cnt = L.its + up((n + 1) / 2);
	L.total = L.cnt + 4;
	return L;
}

__global__ void k_pw_clear(unsigned long long* cnt) { *cnt = 0ull; }

// one row per thread: kind PW_RIM evaluates rim (outputs RI, EI, PI, UI, VI
// and the iteration count), a CFD2D_FLUX_* kind the numerical flux
__global__ __launch_bounds__(BLOCK) void k_pointwise(int kind, int n, const double* rows, double* out, int* its,
                                                     unsigned long long* cnt, RimConst K, int max_newton)
{
	int i = blockIdx.x * BLOCK + threadIdx.x;
	if (i >= n) return;
	int hit = 0;
	if (kind == PW_RIM) {
		const double* r = rows + 8 * (size_t)i;
		double RI, EI, PI, UI, VI;
		int it = 0;
		rim(RI, EI, PI, UI, VI, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], K, max_newton, hit, &it);
		double* o = out + 5 * (size_t)i;
		o[0] = RI;
		o[1] = EI;
		o[2] = PI;
		o[3] = UI;
		o[4] = VI;
		its[i] = it;
	} else {
		const double* r = rows + 10 * (size_t)i;
		Prim L{r[0], r[1], r[2], r[3]}, R{r[4], r[5], r[6], r[7]};
		double f[4];
		if (kind == CFD2D_FLUX_LAX) flux_lax(f, L, R, r[8], r[9], K.GAM);
		else if (kind == CFD2D_FLUX_HLLC) flux_hllc(f, L, R, r[8], r[9], K.GAM);
		else if (kind == CFD2D_FLUX_ROE) flux_roe(f, L, R, r[8], r[9], K.GAM);
		else flux_godunov(f, L, R, r[8], r[9], K, max_newton, hit);
		double* o = out + 4 * (size_t)i;
		for (int q = 0; q < 4; q++) o[q] = f[q];
	}
	if (hit) atomicAdd(cnt, 1ull);
}

int pw_check(int kind, int n, const double* rows, double gam, const double* out)
{
	if (n < 0 || (n > 0 && (!rows || !out))) return fail(CFD2D_ERR_ARG, "pointwise evaluation: bad arguments");
	if (kind != PW_RIM && (kind < CFD2D_FLUX_GODUNOV || kind > CFD2D_FLUX_ROE))
		return fail(CFD2D_ERR_ARG, "pointwise evaluation: unknown flux kind");
	if (!(gam > 1.0)) return fail(CFD2D_ERR_ARG, "pointwise evaluation: gamma must exceed 1");
	return 0;
}

// copies the rows into the staging block st (device memory laid out by
// pw_layout), evaluates them on the null stream and copies the results back
int pw_run(int kind, int n, const double* rows, double gam, int max_newton, double* out, int* iters,
           long long* cap_hits, double* st)
{
	PwLayout L = pw_layout(kind, (size_t)n);
	size_t rw = kind == PW_RIM ? 8 : 10, ow = kind == PW_RIM ? 5 : 4;
	auto* cnt = reinterpret_cast<unsigned long long*>(st + L.cnt);
	int* its = reinterpret_cast<int*>(st + L.its);
	hipLaunchKernelGGL(k_pw_clear, dim3(1), dim3(1), 0, 0, cnt);
	HIPCHK(hipGetLastError());
	if (n > 0) {
		HIPCHK(hipMemcpy(st + L.rows, rows, (size_t)n * rw * sizeof(double), hipMemcpyHostToDevice));
		int mn = max_newton > 0 ? max_newton : CFD2D_NEWTON_CAP;
		hipLaunchKernelGGL(k_pointwise, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, 0, kind, n,
		                   (const double*)(st + L.rows), st + L.out, its, cnt, rim_const(gam), mn);
		HIPCHK(hipGetLastError());
		HIPCHK(hipMemcpy(out, st + L.out, (size_t)n * ow * sizeof(double), hipMemcpyDeviceToHost));
		if (iters) HIPCHK(hipMemcpy(iters, its, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
	}
	unsigned long long h = 0;
	HIPCHK(hipMemcpy(&h, cnt, sizeof(h), hipMemcpyDeviceToHost));
	if (cap_hits) *cap_hits = (long long)h;
	return CFD2D_OK;
}

int pw_alloc_run(int kind, int n, const double* rows, double gam, int max_newton, double* out, int* iters,
                 long long* cap_hits)
{
	int rc = pw_check(kind, n, rows, gam, out);
	if (rc) return rc;
	double* st = nullptr;
	HIPCHK(hipMalloc(&st, pw_layout(kind, (size_t)n).total * sizeof(double)));
	rc = pw_run(kind, n, rows, gam, max_newton, out, iters, cap_hits, st);
	hipError_t e = hipFree(st);
	if (rc) return rc;
	if (e != hipSuccess) return hip_fail(e, "hipFree");
	return CFD2D_OK;
}

// The staging block of cfd2d_flux_noalloc: a fixed array in device memory
// (each device holds its own copy for the life of the process), sized for
// CFD2D_NOALLOC_MAX_ROWS flux rows.  Host code must not pass the array's name
// to a kernel or a copy (on the host it names a stand-in, not device
// memory): noalloc_scratch() returns the calling device's address of it,
// looked up once per device with hipGetSymbolAddress.
// (pw_layout of a flux at the maximum: 10 + 4 doubles per row, half a
// double of iteration counts, the counter)
constexpr size_t NOALLOC_DOUBLES = 29 * (size_t)CFD2D_NOALLOC_MAX_ROWS / 2 + 64;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int noalloc_scratch(double** out)
{
	int dev = -1;
	HIPCHK(hipGetDevice(&dev));
	if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CFD2D_ERR_ARG, "cfd2d_flux_noalloc: device index too large");
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void* ptr = nullptr;
		HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr) return fail(CFD2D_ERR_HIP, "cfd2d_flux_noalloc: no device address for the scratch array");
		g_scratch_ptr[dev] = static_cast<double*>(ptr);
	}
	*out = g_scratch_ptr[dev];
	return 0;
}

}  // namespace

extern "C" int cfd2d_flux_noalloc(int kind, int n, const double* rows, double gam, int max_newton, double* out,
                                  long long* cap_hits)
{
	int rc = pw_check(kind, n, rows, gam, out);
	if (rc) return rc;
	if (kind == PW_RIM || n > CFD2D_NOALLOC_MAX_ROWS)
		return fail(CFD2D_ERR_ARG, "cfd2d_flux_noalloc: at most " + std::to_string(CFD2D_NOALLOC_MAX_ROWS) +
		                               " rows per call (got " + std::to_string(n) + ")");
	static_assert(CFD2D_NOALLOC_MAX_ROWS % 4 == 0 &&
	                  (10 + 4) * (size_t)CFD2D_NOALLOC_MAX_ROWS + CFD2D_NOALLOC_MAX_ROWS / 2 + 4 <= NOALLOC_DOUBLES,
	              "scratch smaller than the staging layout of the maximum row count");
	if (pw_layout(kind, (size_t)n).total > NOALLOC_DOUBLES)
		return fail(CFD2D_ERR_ARG, "cfd2d_flux_noalloc: staging layout exceeds the scratch array");
	double* st = nullptr;
	if ((rc = noalloc_scratch(&st))) return rc;
	return pw_run(kind, n, rows, gam, max_newton, out, nullptr, cap_hits, st);
}

extern "C" int cfd2d_riemann(int n, const double* rows, double gam, int max_newton, double* out, int* iters,
                             long long* cap_hits)
{
	return pw_alloc_run(PW_RIM, n, rows, gam, max_newton, out, iters, cap_hits);
}

extern "C" int cfd2d_flux(int kind, int n, const double* rows, double gam, int max_newton, double* out,
                          long long* cap_hits)
{
	return pw_alloc_run(kind, n, rows, gam, max_newton, out, nullptr, cap_hits);
}

namespace {

// one row per thread: the vector-Jacobian product of a flux the adjoint
// differentiates, (Lb, Rb, nb) = fb^T dF/d(L, R, n)
__global__ __launch_bounds__(BLOCK) void k_pw_vjp(int kind, int n, const double* rows, const double* fbar,
                                                  double* out, RimConst K, int max_newton)
{
	int i = blockIdx.x * BLOCK + threadIdx.x;
	if (i >= n) return;
	const double* r = rows + 10 * (size_t)i;
	Prim L{r[0], r[1], r[2], r[3]}, R{r[4], r[5], r[6], r[7]};
	double fb[4], Lb[4], Rb[4], nb[2];
	for (int q = 0; q < 4; q++) fb[q] = fbar[4 * (size_t)i + q];
	if (kind == CFD2D_FLUX_LAX) lax_vjp(L, R, r[8], r[9], K.GAM, 1.0 / (K.GAM - 1.0), fb, Lb, Rb, nb);
	else if (kind == CFD2D_FLUX_HLLC) hllc_vjp(L, R, r[8], r[9], K.GAM, fb, Lb, Rb, nb);
	else godunov_vjp(L, R, r[8], r[9], K, max_newton, fb, Lb, Rb, nb);
	double* o = out + 10 * (size_t)i;
	for (int q = 0; q < 4; q++) {
		o[q] = Lb[q];
		o[4 + q] = Rb[q];
	}
	o[8] = nb[0];
	o[9] = nb[1];
}

}  // namespace

extern "C" int cfd2d_flux_vjp(int kind, int n, const double* rows, const double* fbar, double gam, int max_newton,
                              double* out)
{
	if (n < 0 || (n > 0 && (!rows || !fbar || !out))) return fail(CFD2D_ERR_ARG, "cfd2d_flux_vjp: bad arguments");
	if (kind != CFD2D_FLUX_GODUNOV && kind != CFD2D_FLUX_LAX && kind != CFD2D_FLUX_HLLC)
		return fail(CFD2D_ERR_ARG, "cfd2d_flux_vjp: the adjoint's fluxes are GODUNOV, LAX and HLLC");
	if (!(gam > 1.0)) return fail(CFD2D_ERR_ARG, "cfd2d_flux_vjp: gamma must exceed 1");
	if (n == 0) return CFD2D_OK;
	double* st = nullptr;
	size_t nd = 24 * (size_t)n;  // rows [n][10], fbar [n][4], out [n][10]
	HIPCHK(hipMalloc(&st, nd * sizeof(double)));
	int rc = CFD2D_OK;
	hipError_t e = hipMemcpy(st, rows, 10 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess) e = hipMemcpy(st + 10 * (size_t)n, fbar, 4 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
	if (e == hipSuccess) {
		int mn = max_newton > 0 ? max_newton : CFD2D_NEWTON_CAP;
		hipLaunchKernelGGL(k_pw_vjp, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, 0, kind, n, (const double*)st,
		                   (const double*)(st + 10 * (size_t)n), st + 14 * (size_t)n, rim_const(gam), mn);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipMemcpy(out, st + 14 * (size_t)n, 10 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
	if (e != hipSuccess) rc = hip_fail(e, "cfd2d_flux_vjp");
	hipError_t e2 = hipFree(st);
	if (rc) return rc;
	if (e2 != hipSuccess) return hip_fail(e2, "hipFree");
	return CFD2D_OK;
}
