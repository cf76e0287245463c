/*
 * CPU restatement of the reference's explicit finite-volume step (FVM_TVD).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline.  Nothing in the product library links or calls it.
 *
 * Reference-exact functions follow src/methods/fvm_tvd.cpp,
 * src/global.cpp, src/bnd_cond.cpp and src/mesh/MeshReaderSalomeUnv.cpp
 * (citations at each function in fvm_oracle.c).  Functions marked
 * "extension" have no reference code; each is pinned by invariants and
 * closed forms in tests/ (see the test names at the function).
 */
#ifndef CFD2D_ORACLE_H
#define CFD2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_FLUX_GODUNOV = 0, OR_FLUX_LAX = 1, OR_FLUX_HLLC = 2, OR_FLUX_ROE = 3 };
enum { OR_GRAD_GG = 0, OR_GRAD_LSQ = 1 };
enum { OR_LIM_NONE = 0, OR_LIM_MINMOD = 1 };
enum { OR_BND_INLET = 1, OR_BND_OUTLET = 2, OR_BND_WALL = 3, OR_BND_PRESSURE_OUTLET = 4 };
enum { OR_CFL_REFERENCE = 0, OR_CFL_SPECTRAL = 1 };

#define OR_MAXMAT 8
#define OR_MAXBC 64

typedef struct or_mesh or_mesh;
typedef struct or_solver or_solver;

typedef struct or_cfg {
	int flux, gradient, limiter, rk_stages, viscous, steady, cfl_mode, max_newton;
	double cfl, tau, gam_flux;
	double ro_min, ro_max, p_min, p_max, u_max;
	double prandtl;
	int nmat;
	double M[OR_MAXMAT], Cp[OR_MAXMAT], K[OR_MAXMAT], ML[OR_MAXMAT];
	int nbc;
	int bc_type[OR_MAXBC];
	double bc_par[OR_MAXBC][4];
} or_cfg;

/* mesh */
int  or_mesh_build(int nn, const double* xy, int nc, const int* tri, int nb, const int* bedges,
                   const int* bgroup, or_mesh** out);
int  or_mesh_periodic(or_mesh* m, int group_a, int group_b, double sx, double sy);
void or_mesh_free(or_mesh* m);
void or_mesh_counts(const or_mesh* m, int* nn, int* nc, int* ne);
void or_mesh_export(const or_mesh* m, double* cell_geom, int* cell_nodes, int* cell_edges,
                    double* edge_geom, int* edge_int, int* edge_group, double* edge_shift);

/* solver */
or_solver* or_solver_create(const or_mesh* m, const or_cfg* cfg, const int* cell_mat, const int* edge_bc);
void   or_solver_free(or_solver* s);
void   or_set_state(or_solver* s, const double* u);
void   or_get_state(const or_solver* s, double* u);
void   or_set_flags(or_solver* s, const unsigned* f);
void   or_get_flags(const or_solver* s, unsigned* f);
void   or_calc_tau(or_solver* s);
double or_get_tau(const or_solver* s, double* ctau);
void   or_set_tau(or_solver* s, double tau);
void   or_gradients(or_solver* s, double* g);
void   or_edge_fluxes(or_solver* s, double* f);
void   or_residual(or_solver* s, double* r);
int    or_run(or_solver* s, int nsteps);
long   or_newton_cap_hits(const or_solver* s);
double or_wall_force(const or_solver* s, const int* edge_mask, double* fx);

/* the Riemann solver alone: rows (rb,pb,ub,vb,re,pe,ue,ve,gam) -> (RI,EI,PI,UI,VI,WI) */
void or_rim(int n, const double* in, double* out, int max_iter, int* iters);
/* a numerical flux alone: rows (rL,pL,uL,vL,rR,pR,uR,vR,nx,ny) -> (fr,fu,fv,fe) */
void or_flux(int kind, int n, const double* in, double* out, double gam);

/* the implicit method's block rows for the current state (0, or 1 = pressure
 * outlet present, 2 = flux not GODUNOV/LAX) */
int  or_implicit_assemble(or_solver* s, int steady, double cfl, double tau, double* blocks, int* cols,
                          double* rhs, double* ctau);
/* face Jacobians: rows (rL,pL,uL,vL,rR,pR,uR,vR,nx,ny) -> (Roe r,u,v,p,c,H, A+[16], A-[16]) */
void or_roe_jacobians(int n, const double* in, double* out, double gam);

void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
