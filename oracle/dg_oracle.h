/*
 * dg_oracle.h -- CPU oracle for the DG residual + RK path of IhmeGroup/DG-MPI.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (dense per-element contractions in the reference's own loop
 * order, the reference's quadrature rules, face tables and flux formulas).
 * It is the checker for the HIP product path: only tests/, the smoke check
 * and the CPU-baseline leg of the benchmark may load it.  Nothing in the
 * product links against it.
 *
 * Every function cites the reference file:line it restates.  The
 * reference itself cannot be built here (its Kokkos/KokkosKernels/toml11
 * submodules are empty), so each restated function is pinned by the tests
 * under tests/ against the reference's own golden vectors, closed forms and
 * invariants.  Parts that no such check pins are marked "parity unpinned".
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NS 5          /* rho, rho u, rho v, rho w, rho E (3D Euler, 1 species) */

enum orc_basis { ORC_LAGRANGE_HEX = 0, ORC_LEGENDRE_HEX = 1 };
enum orc_flux  { ORC_FLUX_LLF = 0, ORC_FLUX_HLLC = 1, ORC_FLUX_ROE = 2 };
enum orc_ic {
    ORC_IC_UNIFORM = 0,          /* data = 5 conservative values            */
    ORC_IC_SPHERE = 1,           /* reference 3D tanh sphere (ignores data) */
    ORC_IC_VORTEX3D = 2,         /* reference 3D vortex (ignores data)      */
    ORC_IC_VORTEX_EXTRUDED = 3,  /* 2D Shu vortex extruded in z: [rho,u,v,p,beta] */
    ORC_IC_GAUSSIAN_BUMP = 4,    /* [A, W, Vx, Vy, Vz, p0, x0, y0, z0]       */
    ORC_IC_TAYLOR_GREEN = 5      /* [Mach]                                   */
};
enum orc_stepper { ORC_FE = 0, ORC_RK4 = 1 };

typedef struct {
    int order;            /* p                                          */
    int basis;            /* enum orc_basis                             */
    int flux;             /* enum orc_flux                              */
    int nelem[3];         /* periodic box: elements per direction       */
    double lo[3], hi[3];  /* box corners                                */
    double gamma, R;      /* ideal gas                                  */
    int ic;               /* enum orc_ic                                */
    int n_ic_data;
    double ic_data[10];
    double warp;          /* 0: affine box; else periodic node displacement
                             (orc_box_node, a trilinear non-affine mesh) */
    int exact_normals;    /* 0: normals where the reference evaluates them
                             (basis.cpp:332, the face-0 points' first two
                             coordinates, i.e. each face's (s,t) swapped);
                             1: at the flux points themselves (new).  Same
                             on affine faces. */
} orc_params;

typedef struct orc_solver orc_solver;

/* ---- reference-element building blocks (exported for the pin tests) ---- */
int  orc_nq1d_from_qorder(int qorder);
int  orc_hex_qorder(int order);
int  orc_quad_face_qorder(int order);
void orc_gl_rule(int nq1d, double* pts, double* wts);
void orc_equidistant_nodes(double start, double stop, int nnodes, double* x);
void orc_basis_1d(int basis, int p, double x, double* val, double* grad);
void orc_basis_3d(int basis, int p, int npts, const double* pts,
                  double* val, double* grad);
void orc_face_pts_order(int orient, int npts, int* order);
void orc_points_on_face(int face_id, int orient, int np, const double* face_pts,
                        double* elem_pts);
void orc_local_nodes_on_face(int face_id, int gorder, int* lfnodes);

/* ---- pointwise physics ---- */
double orc_pressure(double gamma, const double* U);
double orc_maxwavespeed(double gamma, const double* U);
void   orc_flux_interior(double gamma, const double* U, double* Fdir);
void   orc_flux_numerical(int flux, double gamma, const double* UL,
                          const double* UR, const double* N, double* F);
/* the same two functions over n points (U, UL, UR [n][5], N [n][3],
 * F [n][5]; Fdir [n][3][5]) -- batch entry points for the 2D oracle, which
 * embeds (rho, rho u, rho v, rho E) as (rho, rho u, rho v, 0, rho E) */
void   orc_flux_interior_n(double gamma, int n, const double* U, double* Fdir);
void   orc_flux_numerical_n(int flux, double gamma, int n, const double* UL,
                            const double* UR, const double* N, double* F);
void   orc_ic_eval(const orc_params* prm, const double* x, double t, double* U);

/* ---- solver ---- */
/* Node (i, j, k) of the periodic box, i in [0, nelem[0]] etc.:
 *   x = lo_x + i h_x + warp h_x sin(2 pi eta) sin(2 pi zeta)
 *   y = lo_y + j h_y + warp h_y sin(2 pi zeta) sin(2 pi xi)
 *   z = lo_z + k h_z + warp h_z sin(2 pi xi) sin(2 pi eta)
 * with xi = (i mod nx)/nx, eta = (j mod ny)/ny, zeta = (k mod nz)/nz.  Each
 * displacement component depends on the other two node indices only, so
 * opposite box faces stay translates of each other (periodic pairing). */
void orc_box_node(const orc_params* prm, long i, long j, long k, double* x);

orc_solver* orc_create(const orc_params* prm);
orc_solver* orc_create_mesh(const orc_params* prm, long nnodes, const double* coords, long E,
                            const long* e2n, long F, const int* iface6);
void   orc_destroy(orc_solver* s);
long   orc_num_elems(const orc_solver* s);
int    orc_num_basis(const orc_solver* s);
long   orc_num_ifaces(const orc_solver* s);
void   orc_get_face_table(const orc_solver* s, unsigned* faces /* [F][8] */);
void   orc_get_normals(const orc_solver* s, double* normals /* [F][nqf][3] */);
void   orc_get_imm(const orc_solver* s, long elem, double* imm /* [nb][nb] */);
void   orc_init_state(orc_solver* s);
void   orc_set_state(orc_solver* s, const double* U);
void   orc_get_state(const orc_solver* s, double* U);
void   orc_residual(orc_solver* s, const double* U, double* R);
void   orc_inv_mass(const orc_solver* s, double dt, const double* R, double* dU);
void   orc_step(orc_solver* s, int stepper, double dt, int nsteps);
double orc_time(const orc_solver* s);
void   orc_l2_error(const orc_solver* s, double t, double* err /* [5] */);
double orc_max_wavespeed_state(const orc_solver* s);
double orc_total_volume(const orc_solver* s);

#ifdef __cplusplus
}
#endif
#endif
