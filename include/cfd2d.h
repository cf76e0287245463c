/*
 * cfd2d -- MI355X (gfx950) explicit finite-volume step, C ABI.
 *
 * The drop-in boundary for the reference's explicit method FVM_TVD
 * (src/methods/fvm_tvd.h:7-115, src/methods/fvm_tvd.cpp): its Method
 * interface init/run/done (src/methods/method.h:6-11) is kept by the C++
 * host class in cfd-2d_amd/csrc/method.cpp, which calls the entry points
 * below; other hosts (ctypes, C/C++ tests) call them directly.
 *
 * Conventions: every function returns CFD2D_OK (0) or an error code;
 * cfd2d_last_error() gives the message of the last failure on the calling
 * thread.  No exception crosses the ABI.  Host buffers belong to the caller;
 * state arrays are structure-of-arrays double[4][n_cells] in the order
 * (rho, rho*u, rho*v, rho*E).  Each context owns one HIP stream and device
 * buffers; a context is not thread-safe, distinct contexts are.
 */
#ifndef CFD2D_H
#define CFD2D_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
	CFD2D_OK = 0,
	CFD2D_ERR_ARG = 1,          /* bad argument / size */
	CFD2D_ERR_HIP = 2,          /* HIP runtime failure */
	CFD2D_ERR_NEWTON = 3,       /* Riemann Newton iteration hit its cap */
	CFD2D_ERR_NONFINITE = 4,    /* non-finite state after an update */
	CFD2D_ERR_CONFIG = 5,       /* task / configuration error */
	CFD2D_ERR_BOUNDARY = 102,   /* edge without a boundary condition (Exception::TYPE_BOUND_UNKNOWN) */
	CFD2D_ERR_MESH_NAME = 201,  /* unknown mesh type (TYPE_MESH_WRONG_NAME) */
	CFD2D_ERR_UNV_ELEMENT = 221,/* TYPE_MESH_UNV_UNKNOWN_ELEMENT */
	CFD2D_ERR_UNV_BND_EDGE = 222,/* TYPE_MESH_UNV_NOT_DEFINED_BND_EDGE */
	CFD2D_ERR_MESH = 230,       /* malformed mesh input */
	CFD2D_ERR_FILE = 301,       /* FILE_OPENING_ERROR */
	CFD2D_ERR_PERIODIC = 302
};

enum { CFD2D_FLUX_GODUNOV = 0, CFD2D_FLUX_LAX = 1, CFD2D_FLUX_HLLC = 2, CFD2D_FLUX_ROE = 3 };
enum { CFD2D_GRAD_GREEN_GAUSS = 0, CFD2D_GRAD_LSQ = 1 };
enum { CFD2D_LIMITER_NONE = 0, CFD2D_LIMITER_MINMOD = 1 };
enum { CFD2D_CFL_REFERENCE = 0, CFD2D_CFL_SPECTRAL = 1 };
/* boundary types: CFDBndInlet / CFDBndOutlet / CFDBndWallSlip (src/bnd_cond.h:35-53);
 * BOUND_WALL_NO_SLIP maps to the slip wall as the reference's factory does
 * (src/bnd_cond.cpp:50-51); PRESSURE_OUTLET is an extension */
enum { CFD2D_BND_INLET = 1, CFD2D_BND_OUTLET = 2, CFD2D_BND_WALL = 3, CFD2D_BND_PRESSURE_OUTLET = 4 };
enum { CFD2D_MAXMAT = 8 };
enum { CFD2D_NEWTON_CAP = 500 };  /* Newton iterations when cfd2d_params.max_newton <= 0 */

/* kernels reported by cfd2d_timing_read */
enum {
	CFD2D_K_GRADIENT = 0, CFD2D_K_EDGE_FLUX = 1, CFD2D_K_UPDATE = 2, CFD2D_K_REMEDIATE = 3,
	CFD2D_K_TAU = 4, CFD2D_K_PRIM = 5, CFD2D_K_IMPL_GRADIENT = 6, CFD2D_K_IMPL_EDGE = 7,
	CFD2D_K_IMPL_ASSEMBLE = 8, CFD2D_K_COUNT = 9
};

/*
 * What FVM_TVD::init reads from <control>, <limits> and <materials>
 * (src/methods/fvm_tvd.cpp:27-68), plus extension switches whose zero
 * values are the reference behaviour.
 */
typedef struct cfd2d_params {
	int flux;          /* CFD2D_FLUX_*; the reference hard-wires GODUNOV (fvm_tvd.cpp:602-622) */
	int gradient;      /* CFD2D_GRAD_*; reference: Green-Gauss (fvm_tvd.cpp:242-301) */
	int limiter;       /* CFD2D_LIMITER_*; reference: none */
	int rk_stages;     /* 2 = the reference's Heun RK2 (fvm_tvd.cpp:323-438); 3 = SSP-RK3 */
	int viscous;       /* 1 = Navier-Stokes viscous flux (extension) */
	int steady;        /* <STEADY>: 1 = local time step recomputed every step */
	int cfl_mode;      /* CFD2D_CFL_* */
	int max_newton;    /* cap on rim_orig's Newton loop; <= 0 selects CFD2D_NEWTON_CAP (the reference's
	                      loop has no cap, but a device loop must end; hits raise CFD2D_ERR_NEWTON) */
	double cfl;        /* <CFL> */
	double tau;        /* <TAU>: upper bound of the unsteady time step */
	double gam_flux;   /* gamma used by the flux: 1.4 in the reference (fvm_tvd.cpp:345) */
	double ro_min, ro_max, p_min, p_max, u_max;   /* <limits> (fvm_tvd.cpp:43-48) */
	double prandtl;    /* conductivity = ML*Cp/Pr when K == 0 (extension) */
	int nmat;          /* <materials count>, at most CFD2D_MAXMAT */
	double M[CFD2D_MAXMAT], Cp[CFD2D_MAXMAT], K[CFD2D_MAXMAT], ML[CFD2D_MAXMAT];
} cfd2d_params;

typedef struct cfd2d_mesh cfd2d_mesh;
typedef struct cfd2d_ctx cfd2d_ctx;

const char* cfd2d_error_string(int code);
const char* cfd2d_last_error(void);
int cfd2d_device_count(int* n);

/* ---- mesh (host) ----
 * Replaces MeshReader::create/read (src/mesh/MeshReader.cpp:7-23) and the
 * Grid construction of MeshReaderSalomeUnv::read
 * (src/mesh/MeshReaderSalomeUnv.cpp:25-264). */
int  cfd2d_mesh_create(int n_nodes, const double* xy, int n_cells, const int* tri, int n_bedges,
                       const int* bedge_nodes, const int* bedge_group, cfd2d_mesh** out);
int  cfd2d_mesh_read(const char* files_type, const char* name, cfd2d_mesh** out);  /* "salome_unv" | "berkeley_triangle" */
int  cfd2d_mesh_set_groups(cfd2d_mesh* m, int kind /*0 edge, 1 cell*/, int n, const char* const* names);
int  cfd2d_mesh_set_cell_groups(cfd2d_mesh* m, const int* cell_group);
int  cfd2d_mesh_periodic(cfd2d_mesh* m, int group_a, int group_b, double sx, double sy);
int  cfd2d_mesh_concat(int n, const cfd2d_mesh* const* parts, cfd2d_mesh** out);
/* A new mesh with the cells of m in another order and the nodes numbered by
 * first use in it; groups, cell groups and periodic pairings carry over.
 * order CFD2D_ORDER_HILBERT sorts the cells along a Hilbert curve through
 * their centroids (locality: neighbouring cells and their edges close in
 * memory), CFD2D_ORDER_SHUFFLE by a seeded hash (to measure an unordered
 * mesh), CFD2D_ORDER_NONE keeps the order.  cell_perm int[n_cells] (may be
 * NULL): new cell k is old cell cell_perm[k].  One case only (renumber the
 * parts of a batch before cfd2d_mesh_concat).  The role of the reference's
 * METIS dual-graph partition (src/methods/decomp.cpp:86-104), here for
 * locality on one device. */
enum { CFD2D_ORDER_NONE = 0, CFD2D_ORDER_HILBERT = 1, CFD2D_ORDER_SHUFFLE = 2 };
int  cfd2d_mesh_renumber(const cfd2d_mesh* m, int order, int seed, int* cell_perm, cfd2d_mesh** out);
int  cfd2d_mesh_counts(const cfd2d_mesh* m, int* n_nodes, int* n_cells, int* n_edges);
int  cfd2d_mesh_nodes(const cfd2d_mesh* m, double* xy);   /* double[2][n_nodes] interleaved x,y */
int  cfd2d_mesh_group_count(const cfd2d_mesh* m, int kind, int* n);
int  cfd2d_mesh_group_name(const cfd2d_mesh* m, int kind, int i, char* buf, int len);
int  cfd2d_mesh_get_cell_groups(const cfd2d_mesh* m, int* cell_group);
int  cfd2d_mesh_case_offsets(const cfd2d_mesh* m, int* n_cases, int* cell_offsets, int* edge_offsets);
/* geometry in the reference's layout: cell (S, cx, cy, 0); edge (nx, ny, l, 3 points);
 * edge ints (n1, n2, c1, c2, type); any pointer may be NULL */
int  cfd2d_mesh_export(const cfd2d_mesh* m, double* cell_geom, int* cell_nodes, int* cell_edges,
                       double* edge_geom, int* edge_int, int* edge_group, double* edge_shift);
void cfd2d_mesh_destroy(cfd2d_mesh* m);
/* The reverse of the mesh geometry (extension): given adjoints of what the
 * explicit kernels read -- Ge double[13][n_edges] per edge: normal (x, y)
 * at Gauss points 1 and 2, side s's offset to Gauss point g (x, y) at
 * 4 + 4 s + 2 g, length at 12; Gc double[7][n_cells] per cell: area, then
 * per face (ascending edge order) its Green-Gauss weight sign * n l (x, y),
 * or with lsq != 0 its least-squares weight (x, y); lny double[n_edges]:
 * adjoint of l * n_y -- returns dxy double[n_nodes][2], the adjoint of every
 * node coordinate (inputs may be NULL: zero).  Host only;
 * cfd2d_geometry_vjp_device is the same on the GPU. */
int  cfd2d_mesh_geometry_vjp(const cfd2d_mesh* m, int lsq, const double* Ge, const double* Gc, const double* lny,
                             double* dxy);

/* ---- solver context (device) ----
 * Replaces the state arrays and passes of FVM_TVD (fvm_tvd.cpp:177-197,
 * 216-499).  bc_type/bc_par describe n_bc boundary objects (CFDBoundary,
 * src/bnd_cond.h:12-53: par = Vx, Vy, T, P for inlets, P for pressure
 * outlets); group_bc maps each edge group to a boundary object (-1: none);
 * cell_mat gives each cell's material (NULL: all 0).  Every boundary edge
 * must map to an object (else CFD2D_ERR_BOUNDARY, as fvm_tvd.cpp:148-152). */
int  cfd2d_ctx_create(const cfd2d_mesh* mesh, const cfd2d_params* p, int n_bc, const int* bc_type,
                      const double* bc_par, const int* group_bc, const int* cell_mat, cfd2d_ctx** out);
void cfd2d_ctx_destroy(cfd2d_ctx* c);
int  cfd2d_set_state(cfd2d_ctx* c, const double* u);          /* host double[4][nc] */
int  cfd2d_get_state(cfd2d_ctx* c, double* u);
int  cfd2d_set_state_device(cfd2d_ctx* c, const double* d_u); /* device double[4][nc] */
int  cfd2d_get_state_device(cfd2d_ctx* c, double* d_u);
int  cfd2d_set_flags(cfd2d_ctx* c, const unsigned* flags);   /* Cell::flag (src/mesh/grid.h:35) */
int  cfd2d_get_flags(cfd2d_ctx* c, unsigned* flags);
/* FVM_TVD::calcTimeStep (fvm_tvd.cpp:216-240) on the current state; per case for concatenated meshes */
int  cfd2d_calc_tau(cfd2d_ctx* c);
int  cfd2d_get_tau(cfd2d_ctx* c, double* tau_case0, double* ctau /* nc, may be NULL */);
int  cfd2d_set_tau(cfd2d_ctx* c, double tau);
/* Replace the parameters of every boundary object (double[n_bc][4], the
 * layout of cfd2d_ctx_create's bc_par).  Extension: the reference reads them
 * once from the task file (CFDBoundary::create, bnd_cond.cpp:10-41); the
 * adjoint's callers change inflow parameters between runs. */
int  cfd2d_set_bc_par(cfd2d_ctx* c, const double* bc_par);
/* stage diagnostics on the current state (parity checks) */
int  cfd2d_gradients(cfd2d_ctx* c, double* g);      /* double[8][nc]: rx ry px py ux uy vx vy */
int  cfd2d_edge_fluxes(cfd2d_ctx* c, double* f);    /* double[ne][4]: (sum of Gauss-point fluxes) * l/2 */
int  cfd2d_residual(cfd2d_ctx* c, double* r);       /* double[4][nc]: the *_int arrays after one stage */
/* nsteps time steps of FVM_TVD::run (fvm_tvd.cpp:310-449), enqueued on the context stream */
int  cfd2d_step(cfd2d_ctx* c, int nsteps);
/* the same steps, recorded once as a HIP graph and replayed (use_graph != 0) */
int  cfd2d_step_graph(cfd2d_ctx* c, int nsteps);
int  cfd2d_sync(cfd2d_ctx* c);                      /* waits; returns a device error if one was raised */
int  cfd2d_status(cfd2d_ctx* c, long long* newton_cap_hits, int* nonfinite);
/* pressure force over the edges with mask != 0 (calcLiftForce, fvm_tvd_implicit.cpp:1088-1110) */
int  cfd2d_wall_force(cfd2d_ctx* c, const int* edge_mask, double* fx, double* fy);
/* per-kernel HIP-event timing on the context stream */
int  cfd2d_timing(cfd2d_ctx* c, int enable);
/* Where the cell gradients are formed: mode 1 in the edge pass itself (each
 * edge's lane pair forms its two cells'), 2 in the tiled edge pass (a
 * workgroup per tile of consecutive cells forms the gradients its edges
 * read once each, staged in LDS; CFD2D_ERR_ARG when the mesh's edges are
 * not in c1 order), 0 in a separate gradient pass whose records the edge
 * pass reads, -1 (default) 1 without a limiter and 0 with one.  All modes
 * give the same bits.  No reference counterpart. */
int  cfd2d_fuse_gradient(cfd2d_ctx* c, int mode);
/* Whether the update pass reads and writes the conserved state with
 * non-temporal loads and stores (kept out of the caches, leaving them to the
 * data the edge pass gathers): 1 yes, 0 no, -1 (default) when the context's
 * device arrays are over twice the 256 MB Infinity Cache.  The same bits
 * either way.  No reference counterpart. */
int  cfd2d_stream_state(cfd2d_ctx* c, int mode);
/* Launch shape of the edge pass (no reference counterpart): 0 = 256 threads
 * per workgroup with the compiler's register allocation, 1 = 256 threads
 * held to five waves per SIMD (at most 96 VGPRs), 2 = 128 threads, 3 = 512
 * threads.  The same results in every shape.  Shapes apply to
 * single-material contexts' unlimited and stored-gradient edge passes. */
int  cfd2d_edge_shape(cfd2d_ctx* c, int shape);
/* Autotuner (no reference counterpart): times trial_steps steps of each
 * variant -- gradients in their own pass, in the edge pass or in the tiled
 * edge pass (when the mesh has tiles), eager launches
 * or a replayed recorded step, cached or streamed state in the update pass
 * -- on the context's current state, then the edge pass's launch shapes
 * (cfd2d_edge_shape) with the fastest of them, restores that state exactly,
 * and keeps the fastest for cfd2d_step (same results either way).  Reports
 * the choice and its ms per step (any pointer may be NULL; cfd2d_tuning
 * reads the launch shape).  Uses a temporary device copy of the state
 * (~100 B per cell). */
int  cfd2d_autotune(cfd2d_ctx* c, int trial_steps, int* fuse_mode, int* use_graph, int* stream_state,
                    double* ms_per_step);
/* the step variant in effect (as cfd2d_autotune or the setters left it) */
int  cfd2d_tuning(cfd2d_ctx* c, int* fuse_mode, int* use_graph, int* stream_state, int* edge_shape);
/* which kernels get events while timing is on: bit k = CFD2D_K_* id
 * (default all); fewer events disturb the stream less */
int  cfd2d_timing_kernels(cfd2d_ctx* c, unsigned mask);
int  cfd2d_timing_read(cfd2d_ctx* c, int kernel, double* total_ms, long long* launches);
int  cfd2d_ctx_info(cfd2d_ctx* c, int* n_cells, int* n_edges, int* n_cases);
int  cfd2d_ctx_stream(cfd2d_ctx* c, void** hip_stream);
/* Discrete adjoint of nsteps steps (extension; no reference counterpart: the
 * reference's steady implicit method FVM_TVD_IMPLICIT only evaluates this
 * objective, calcLiftForce at fvm_tvd_implicit.cpp:1088-1110).  Objective
 * J = sum over edges with edge_mask != 0 of p(c1) * l * n_y after the steps;
 * on return *J, dJ_dU0 = dJ/d(state at the call) as double[4][nc], and
 * dJ_dbc = dJ/d(parameters of each boundary object) as double[n_bc][4] (NULL
 * skips either).  Forward states are kept every ckpt_every steps and each
 * segment but the last is recomputed once in the reverse sweep; ckpt_every
 * <= 0 picks the largest interval whose stage records (96 B per cell per RK
 * stage) fit in half of the free device memory.  Supports FLUX LAX, HLLC
 * and GODUNOV (the exact solver differentiated through its Newton
 * iterations as evaluated), no limiter, inviscid, unsteady; the time step is
 * held fixed.  The context is left at the final state, as after
 * cfd2d_step(c, nsteps). */
int  cfd2d_adjoint(cfd2d_ctx* c, int nsteps, int ckpt_every, const int* edge_mask, double* J, double* dJ_dU0,
                   double* dJ_dbc);
/* cfd2d_adjoint plus dJ_dxy = dJ/d(node coordinates) as double[n_nodes][2]
 * (x, y) of mesh m, the mesh the context was created from: the reverse sweep
 * also accumulates the adjoints of the geometry the kernels read (edge
 * normals and lengths, Gauss-point offsets, cell areas, Green-Gauss or
 * least-squares weights; the wall-force objective's l n_y), and the host
 * carries them back through the mesh builder (centroids, normals, Gauss
 * points, Heron areas, MeshReaderSalomeUnv.cpp:89-255, and the weights) to
 * the nodes.  The time step is held at its value (as in cfd2d_adjoint) and
 * so are the normals' orientation and periodic shifts.  Each stage also
 * forms its forward fluxes again.  The geometry adjoints stay on the device:
 * the reverse through the mesh builder runs there (cfd2d_geometry_vjp_device)
 * on a device copy of m kept in the context while the same mesh is passed
 * (a new handle or a periodic pairing counts as a new mesh).  Extension, no
 * reference counterpart. */
int  cfd2d_adjoint_geometry(cfd2d_ctx* c, const cfd2d_mesh* m, int nsteps, int ckpt_every, const int* edge_mask,
                            double* J, double* dJ_dU0, double* dJ_dbc, double* dJ_dxy);
/* cfd2d_mesh_geometry_vjp on the context's GPU, for host arrays of the same
 * layouts (Ge and Gc required, lny may be NULL: zero): the same arithmetic
 * with every sum in the host function's order, so the same bits.  m must be
 * the context's mesh (same cell and edge counts).  Allocates device copies
 * of Ge and Gc for the call (cfd2d_adjoint_geometry reads its own from the
 * adjoint arena).  Extension, no reference counterpart. */
int  cfd2d_geometry_vjp_device(cfd2d_ctx* c, const cfd2d_mesh* m, int lsq, const double* Ge, const double* Gc,
                               const double* lny, double* dxy);
/* Whether cfd2d_adjoint records its forward pass (checkpoints, recorded
 * segment, seed) and its reverse sweep (recomputed segments, every stage's
 * reverse kernels, the boundary sums) each as a HIP graph and replays it:
 * enable != 0 (the default) records on the first call of a shape (step
 * count, interval, objective size) and replays on later calls, one launch
 * per pass; 0 launches every kernel directly.  The same results either way.
 * No reference counterpart. */
int  cfd2d_adjoint_graph(cfd2d_ctx* c, int enable);
/* The objective of the last cfd2d_adjoint call split by case of a
 * concatenated mesh (double[n_cases]; each case's own masked edges, summed
 * in ascending edge order).  Extension, no reference counterpart. */
int  cfd2d_adjoint_objectives(cfd2d_ctx* c, double* J_case);

/* The implicit method's linear system for the context's current state,
 * the matrix fill of FVM_TVD_IMPLICIT::run (src/methods/fvm_tvd_implicit.cpp:
 * 534-663: calcGrad 186-262, reconstruct 431-476, calcRoeAverage 408-429,
 * eigenValues/right/leftEigenVector/calcAP/calcAM 335-406, calcFlux
 * 917-961, boundaryCond 963-1008), assembled on the device; the solve
 * (MatrixSolver, HYPRE GMRES) stays with the caller.
 * One block row per cell, at most 4 blocks (BSR):
 *   blocks double[n_cells][4][16]  row-major 4x4 blocks; slot 0 = diagonal
 *                                  (sum of A+ l over the faces + S/ctau I),
 *                                  slots 1..3 = A- l of the faces in
 *                                  ascending edge order (zero on a boundary)
 *   cols   int[n_cells][4]         block column of each slot (-1: boundary)
 *   rhs    double[n_cells][4]      right4: -sum over faces of the 2-point
 *                                  flux times l/2 (outward from the cell)
 *   ctau   double[n_cells]         steady: 2 cfl S / (sum l (|u|+c) of the
 *                                  faces' Roe states + 1e-100); else tau
 * steady != 0 uses cfl, otherwise tau.  FLUX must be GODUNOV or LAX (the
 * implicit method's calcFlux); GAM = 1.4 in the flux and the Jacobians as in
 * the reference.  Any output pointer may be NULL.  Boundary objects of type
 * PRESSURE_OUTLET are rejected (the implicit method has none).  Uses the
 * context's gradient and edge-flux buffers as scratch. */
int  cfd2d_implicit_assemble(cfd2d_ctx* c, int steady, double cfl, double tau, double* blocks, int* cols,
                             double* rhs, double* ctau);
/* ---- the Riemann solver and the numerical fluxes on given state pairs ----
 * Parity checks of the device physics row by row, outside any mesh (states
 * no mesh case reaches: vacuum, strong shocks, Toro's tests).  Host arrays;
 * the rows are staged through device memory allocated for the call.
 * max_newton as cfd2d_params.max_newton (<= 0: CFD2D_NEWTON_CAP);
 * *cap_hits (may be NULL) = rows whose Newton loop hit the cap.
 *
 * rim_orig (src/global.cpp:232-405, as FVM_TVD::calcFlux calls it,
 * fvm_tvd.cpp:602-622): rows double[n][8] = (rho, p, u_n, u_t) of the left
 * then the right side; out double[n][5] = (rho, e, p, u_n, u_t) sampled at
 * x/t = 0 (RI, EI, PI, UI, VI); iters int[n] (may be NULL) = Newton
 * iterations of each row (0 on the vacuum branch); gam = GAM. */
int  cfd2d_riemann(int n, const double* rows, double gam, int max_newton, double* out, int* iters,
                   long long* cap_hits);
/* The numerical flux kind (CFD2D_FLUX_*: GODUNOV = FVM_TVD::calcFlux,
 * fvm_tvd.cpp:602-622; LAX = FVM_TVD_IMPLICIT::calcFlux,
 * fvm_tvd_implicit.cpp:939-960; HLLC and ROE extensions) of n state pairs:
 * rows double[n][10] = (rL, pL, uL, vL, rR, pR, uR, vR, nx, ny), out
 * double[n][4] = (mass, x-momentum, y-momentum, energy) flux through the
 * unit normal. */
int  cfd2d_flux(int kind, int n, const double* rows, double gam, int max_newton, double* out, long long* cap_hits);
/* cfd2d_flux without any device allocation, for callers that cannot
 * allocate (tight serving loops, code that must not call hipMalloc).  No
 * counterpart in the reference.  The rows are staged through a fixed device
 * array of the library (7.6 MB per device, held until the process
 * ends) instead of memory allocated for the call: n <= CFD2D_NOALLOC_MAX_ROWS
 * (larger n returns CFD2D_ERR_ARG).  Every call on one device shares that
 * array, so calls that overlap on one device (from several threads) race;
 * the caller serialises them.  Same arguments and results as cfd2d_flux. */
enum { CFD2D_NOALLOC_MAX_ROWS = 65536 };
int  cfd2d_flux_noalloc(int kind, int n, const double* rows, double gam, int max_newton, double* out,
                        long long* cap_hits);
/* The vector-Jacobian product of a flux the adjoint differentiates (kind
 * CFD2D_FLUX_GODUNOV, LAX or HLLC) on n state pairs: rows double[n][10] as
 * cfd2d_flux, fbar double[n][4] the seeds; out double[n][10] = fbar^T dF/dx
 * for x = (rL, pL, uL, vL, rR, pR, uR, vR, nx, ny), the normal's components
 * taken as independent.  The same device functions as cfd2d_adjoint's
 * reverse sweep (the exact Godunov flux: the Newton iterations as evaluated,
 * reversed).  Host arrays, staged through device memory allocated for the
 * call.  No counterpart in the reference. */
int  cfd2d_flux_vjp(int kind, int n, const double* rows, const double* fbar, double gam, int max_newton,
                    double* out);

/* Page-locked host memory (hipHostMalloc): output buffers of
 * cfd2d_implicit_assemble placed here are copied at the link's full rate.
 * No reference counterpart (the reference's matrix lives in host memory). */
int  cfd2d_host_alloc(size_t bytes, void** p);
int  cfd2d_host_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
