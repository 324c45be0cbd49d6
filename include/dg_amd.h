/*
 * dg_amd.h -- C ABI of the MI355X-native DG residual + RK path.
 *
 * Drop-in boundary for the hot path of IhmeGroup/DG-MPI: the 3D Euler
 * discontinuous-Galerkin residual on periodic hexahedral boxes and the
 * FE/RK4 stepper that drives it.  The reference has no C ABI; its surface is
 * the templated C++ Solver<3> over Kokkos views.  Each entry point below
 * names the reference interface it replaces (paths relative to the
 * reference repository root).
 *
 * Conventions
 *   - status: every call returns DG_OK (0) or a dg_status code; the message
 *     is available from dg_last_error(handle) (dg_last_error(NULL) for a
 *     failed dg_create).  No C++ exception crosses this boundary.
 *   - ownership: the handle owns its device buffers (solution, RK
 *     registers); callers own host buffers and the device buffers they pass
 *     to dg_residual / dg_apply_inv_mass.
 *   - layout: solution and residual coefficients are FP64, [E][nb][5]
 *     row-major (element, basis k*(p+1)^2 + j*(p+1) + i, variable rho, rho u,
 *     rho v, rho w, rho E) -- the reference's LayoutRight Uc
 *     (src/solver/base.h:47).  DG_LAYOUT_LEFT selects the column-major
 *     variant the reference uses on GPUs (src/solver/base.cpp:248-275).
 *   - elements of the periodic box are numbered e = ix + nx*(iy + ny*iz).
 *   - streams: calls taking a `stream` argument are stream-ordered on that
 *     HIP stream (NULL = the handle's own stream); the others synchronise
 *     before returning.  One host thread per handle.
 */
#ifndef DG_AMD_H
#define DG_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dg_solver dg_solver;

enum dg_status {
    DG_OK = 0,
    DG_ERR_INPUT = 1,        /* bad case input (reference: InputException)   */
    DG_ERR_DEVICE = 2,       /* HIP runtime error or no usable GPU           */
    DG_ERR_ARGUMENT = 3,     /* bad pointer / size / layout argument         */
    DG_ERR_UNSUPPORTED = 4,  /* valid reference input outside this path      */
    DG_ERR_NUMERICS = 5      /* non-finite state (the reference asserts)     */
};

enum dg_layout { DG_LAYOUT_RIGHT = 0, DG_LAYOUT_LEFT = 1 };

typedef struct {
    long num_elems;          /* E                                            */
    int order;               /* p                                            */
    int num_basis;           /* (p+1)^3                                      */
    int num_state_vars;      /* 5                                            */
    int nelem[3];
    double lower[3], upper[3];
    double gamma, gas_constant;
    double timestep;         /* Stepper.timestep, or the CFL-derived dt      */
    double end_time;         /* Stepper.end                                  */
    double time;             /* solver time                                  */
    int num_iterations;      /* (int)(end / timestep), as Solver::solve      */
    int stepper;             /* 0 FE, 1 RK4                                  */
    int flux;                /* 0 LaxFriedrichs, 1 HLLC, 2 Roe               */
    int basis;               /* 0 LagrangeHex, 1 LegendreHex                 */
    int ic;                  /* 0 Uniform, 1 Sphere, 2 IsentropicVortex,
                                3 ExtrudedIsentropicVortex, 4 GaussianBump,
                                5 TaylorGreen, 6 Gaussian (2D),
                                7 VortexPulseMix                             */
    int dim;                 /* 3: hexahedra, 5 state variables; 2:
                                quadrilaterals, 4 state variables (rho,
                                rho u, rho v, rho E), nelem[2] = 1           */
    int precision;           /* 64, or 32: Numerics.precision = "FP32" (face
                                and volume work of the stage in FP32; RK
                                registers, M^-1 and reductions in FP64), or
                                16: "BF16" / "BF16x3" (the FP32 form with
                                the x-y projection on bf16 matrix cores)     */
} dg_info;

/* Parse a TOML case (the keys of src/exec/main.cpp:38-83,
 * src/solver/base.cpp:20-67, src/physics/base/base.cpp:24-32,
 * src/numerics/numerics_data.cpp:229-300; in place of Mesh.file the mesh
 * keys are type = "PeriodicBox", nelem, lower, upper, and optionally
 * warp (periodic node displacement, element widths) and normals =
 * "reference" | "exact"),
 * build the periodic mesh and the reference-element operators
 * (Solver<dim>::Solver + precompute_matrix_helpers,
 * src/solver/base.cpp:20-111) and allocate device state on the current
 * HIP device.  The state is left at zero; call dg_init_state. */
int dg_create(const char* toml_text, dg_solver** out);

/* The same for a hex mesh given as the arrays of the reference's mesh file
 * (Mesh<dim>::read_mesh_file, src/mesh/mesh.cpp:79-175): node_coords
 * [num_nodes][3], elem_nodes [num_elems][8] (node 4k + 2j + i of the
 * element's own reference frame, Q1 hexes) and iface_data [num_ifaces][6]
 * = (elemL, faceL, orientL, elemR, faceR, orientR) with the face ids and
 * orientations of src/numerics/basis/shape.cpp:296-428.  Every element
 * face must appear in exactly one interior face (periodic or closed
 * meshes; boundary face groups are not supported).  The TOML case holds
 * no [Mesh] keys other than normals = "reference" | "exact"; the element
 * metric is computed at every quadrature point (non-affine elements are
 * allowed) and the per-element inverse mass matrices are formed on the
 * device (orders 1-4).  With Physics.dim = 2 the mesh holds bilinear
 * quadrilaterals: node_coords [num_nodes][2], elem_nodes [num_elems][4]
 * (node 2j + i), face ids 0..3 and orientations 0..1 of
 * src/numerics/basis/shape.cpp:69-122 (orders 1-5). */
int dg_create_mesh(const char* toml_text, long num_nodes, const double* node_coords, long num_elems,
                   const long* elem_nodes, long num_ifaces, const int* iface_data, dg_solver** out);

/* Release device buffers and the handle.  Accepts NULL. */
int dg_destroy(dg_solver* s);

/* Last error message of the handle (NULL: of the last failed dg_create on
 * this thread).  Never returns NULL. */
const char* dg_last_error(const dg_solver* s);

/* Sizes and parameters of the case. */
int dg_get_info(const dg_solver* s, dg_info* info);

/* The mesh of the case: node coordinates [num_nodes][3] and element nodes
 * [num_elems][8] (node 4k + 2j + i of each element's frame), as given to
 * dg_create_mesh, or the (nx+1)(ny+1)(nz+1) grid of the box (warp included,
 * node i + (nx+1)(j + (ny+1)k), element ix + nx(iy + ny iz)) -- the arrays
 * the reference's Writer stores (src/io/writer.cpp:68-98).  Call with NULL
 * arrays to get num_nodes; then node_coords holds num_nodes*3 doubles and
 * elem_nodes num_elems*8 longs.  In 2D: [num_nodes][2] and [num_elems][4]
 * (node 2j + i, element ix + nx iy). */
int dg_get_mesh(const dg_solver* s, long* num_nodes, double* node_coords, long* elem_nodes);

/* L2 projection of the initial condition on (2p+2)^3 Gauss points, time
 * set to 0 (Solver::init_state_from_fcn, src/solver/base.cpp:113-223;
 * SolverTools::L2_projection, src/solver/tools.cpp:66-122).  With
 * Numerics.subtract_mean (reduced-precision forms) it also sets the
 * reference state of the FP32 stage: the domain mean of the new state. */
int dg_init_state(dg_solver* s);

/* Host <-> device copies of the solution coefficients (the h_Uc mirror,
 * src/solver/base.cpp:226-230, and the two layouts of
 * read_in_coefficients, src/solver/base.cpp:233-279).  host holds
 * E*(p+1)^3*5 doubles.  dg_set_state sets the FP32 stage's reference
 * state as dg_init_state does; a state written through
 * dg_state_device_ptr keeps the previous one (any reference gives the same
 * residual; the closer to the state's mean, the smaller its FP32
 * rounding). */
int dg_set_state(dg_solver* s, const double* host, int layout);
int dg_get_state(const dg_solver* s, double* host, int layout);

/* Device pointer to the solution coefficients ([E][nb][5], owned by the
 * handle; valid until dg_destroy). */
int dg_state_device_ptr(dg_solver* s, double** dev);

/* Residual R(U) = volume term - face term, not mass-inverted
 * (Solver::get_residual, src/solver/base.cpp:401-561).  U_dev and R_dev
 * are caller-owned device buffers of E*nb*5 doubles and must not alias. */
int dg_residual(dg_solver* s, const double* U_dev, double* R_dev, void* stream);

/* dU = dt * M^-1 R per element (SolverTools::mult_inv_mass_matrix,
 * src/solver/tools.cpp:124-139).  In-place (R_dev == dU_dev) is allowed. */
int dg_apply_inv_mass(dg_solver* s, double dt, const double* R_dev, double* dU_dev,
                      void* stream);

/* nsteps time steps of the configured stepper on the handle's state
 * (StepperBase::take_time_step: FE src/numerics/timestepping/stepper.cpp:23-55,
 * RK4 :71-142).  dt <= 0 uses the case time step. */
int dg_step(dg_solver* s, double dt, int nsteps);

/* Solver::solve (src/solver/base.cpp:377-399): (int)(end/dt) steps. */
int dg_solve(dg_solver* s);

/* Per-variable L2 error against the exact solution at time t, normalised
 * by the domain volume, on (2p+2)^3 points (get_error with ord = 2,
 * src/post/main.cpp:30-201).  2D: (2p+2)^2 points, err[0..3]. */
int dg_l2_error(dg_solver* s, double t, double err[5]);

/* Domain integrals of rho, rho u, rho v, rho w, rho E and of the kinetic
 * energy 1/2 rho |u|^2 on (2p+2)^3 points, compensated FP64 sums (no
 * reference counterpart; conservation and kinetic-energy diagnostics).
 * 2D: rho, rho u, rho v, rho E, kinetic energy, 0. */
int dg_integrals(dg_solver* s, double out[6]);

/* Largest |u| + c over the volume quadrature points of the state (the
 * wave speed of src/physics/euler/functions.cpp:25-40); the CFL rule
 * dt = CFL * min(h) / max(|u|+c) (Stepper.cfl) uses it. */
int dg_max_wavespeed(dg_solver* s, double* amax);

/* Set the solver time (the reference hard-codes 0, src/solver/base.cpp:25). */
int dg_set_time(dg_solver* s, double t);

/* Make the handle launch on a caller-owned HIP stream (NULL: back to the
 * handle's own stream).  No reference counterpart (Kokkos uses the default
 * execution space instance); lets callers order and time the path's
 * kernels on their own stream. */
int dg_set_stream(dg_solver* s, void* stream);

/* Batches of independent cases (no reference counterpart: the reference
 * runs one case per process).  A batch advances many periodic-box cases of
 * uneven size, order, flux and precision together: per RK stage one launch
 * per (order, flux, precision) group covers the elements of all its cases,
 * and one RK4 step of every group is captured once as a HIP graph and
 * replayed.  Each case keeps its own handle, registers, time step
 * (Stepper.timestep or the CFL-derived one) and time; after
 * dg_batch_step(b, n) every case holds what dg_step(case, 0, n) would have
 * produced (the same stage arithmetic; the compiler may fuse multiply-adds
 * differently in the batched kernel, which moves round-off only).  Cases: distinct handles on one device, 3D affine
 * boxes (no mesh, no warp), RK4.  The batch borrows the handles: destroy it
 * before them.  dg_batch_step waits for each case's stream before it starts
 * and returns when the steps are done. */
typedef struct dg_batch dg_batch;
int dg_batch_create(int num_cases, dg_solver* const* cases, dg_batch** out);
int dg_batch_step(dg_batch* b, int nsteps);
int dg_batch_destroy(dg_batch* b);
/* Last error of the batch (NULL: of the last failed dg_batch_create on this
 * thread).  Never returns NULL. */
const char* dg_batch_last_error(const dg_batch* b);
/* Each (order, basis, flux, precision) group's launch has tuning
 * candidates: two team sizes x two element-brick caps (the cases of the
 * group are traversed in bricks of at most 8^3 or 4^3 elements).  The
 * first dg_batch_step call with nsteps >= 6 runs its first six steps as
 * tuning steps (one with each team size, not counted: code loading; then
 * one with each candidate), keeps each group's fastest candidate and
 * captures the step again.  This reports, per group, the order, the team
 * size in use and the two team sizes' timings at the chosen cap (ms per
 * RK4 step of the group run alone; 0 before tuning).  Arrays of
 * *num_groups entries; NULL arrays: only the count. */
int dg_batch_tuning(const dg_batch* b, int* num_groups, int* order, int* team, double* ms_default,
                    double* ms_alternate);

/* The tile side of the tuning (no reference counterpart): per group the
 * brick cap in use (log2 of the largest brick edge), the number of cases and
 * elements, and the timing of every candidate (ms per RK4 step of the group
 * run alone; ms_candidates[group * num_candidates + team + 2 cap_index],
 * cap_index 0 = 8^3 bricks, 1 = 4^3; 0 before tuning).  NULL arrays are
 * skipped. */
int dg_batch_tuning_tiles(const dg_batch* b, int* num_groups, int* num_candidates, int* brick_cap, int* cases,
                          long* elements, double* ms_candidates);

/* Learned closure for coarse runs (no reference counterpart).  A per-element
 * MLP -- 70 features -> 128 -> 128 -> 40, GELU, BF16 weights with FP32
 * accumulation, FP32 master weights trained with Adam by this library's
 * own forward / backward kernels -- corrects the tendency of a p = 1
 * LegendreHex periodic box: dU/dt = M^-1 R(U) + C(U).  Features of an
 * element: its 40 coefficients, then the jumps of the face-mean state
 * (neighbour minus own) across its x-, x+, y-, y+, z-, z+ faces (5 each).
 * Targets: with a finer (p >= 2, Legendre) solver on the same mesh, Uc =
 * the p = 1 truncation of the fine state (its L2 projection) and
 * Y = trunc(M_f^-1 R_f(U_f)) - M_c^-1 R_c(Uc).  The network sees features
 * normalised as (x - x_mean) * x_scale and predicts Y / y_scale.  Device
 * arrays are caller-owned: X [rows][70], Y [rows][40] float, states double.
 * Parameters (dg_mlp_num_params floats): W1 [70][128], b1, W2 [128][128],
 * b2, W3 [128][40], b3.  Calls on one dg_mlp must not overlap (from
 * several host threads): it owns training buffers and the inference
 * kernel's weight image. */
typedef struct dg_mlp dg_mlp;
int dg_mlp_create(unsigned long long seed, dg_mlp** out);
int dg_mlp_destroy(dg_mlp* m);
const char* dg_mlp_last_error(const dg_mlp* m);
int dg_mlp_num_params(long* n);
int dg_mlp_get_params(const dg_mlp* m, float* host);
int dg_mlp_set_params(dg_mlp* m, const float* host);
int dg_mlp_set_scaling(dg_mlp* m, const float* x_mean, const float* x_scale, const float* y_scale);
/* Y_dev = C(X_dev), de-normalised: the closure step's inference kernel
 * (matrix cores: bf16 weights and layer inputs, FP32 accumulation) */
int dg_mlp_forward(dg_mlp* m, const float* X_dev, float* Y_dev, long rows);
/* gradient (host, num_params floats) and loss of the mean squared error
 * in normalised target units over all rows */
int dg_mlp_gradient(dg_mlp* m, const float* X_dev, const float* Y_dev, long rows, float* grad_host, double* loss);
/* Adam steps (lr; beta 0.9 / 0.999) over consecutive mini-batches of
 * `batch` rows named by idx_dev[0 .. nidx), the features jittered by
 * Gaussian noise of standard deviation `noise` (normalised units; 0: none);
 * mean loss of those steps */
int dg_mlp_train(dg_mlp* m, const float* X_dev, const float* Y_dev, long nrows, const int* idx_dev, long nidx,
                 int batch, float lr, float noise, double* mean_loss);
/* features of the p = 1 state U_dev of solver s */
int dg_closure_features(dg_solver* s, const double* U_dev, float* X_dev);
/* coarse state Uc_dev and target Y_dev of the fine state Uf_dev (uses the
 * handles' stage registers as scratch) */
int dg_closure_target(dg_solver* fine, const double* Uf_dev, dg_solver* coarse, double* Uc_dev, float* Y_dev);
/* target for a coarse state of its own (a coarse run beside the fine one):
 * Y = trunc(M_f^-1 R_f(U_f)) - M_c^-1 R_c(Uc) + (trunc(U_f) - Uc) / tau,
 * the tendency mismatch at Uc plus a relaxation toward the fine state */
int dg_closure_target_at(dg_solver* fine, const double* Uf_dev, dg_solver* coarse, const double* Uc_dev, double tau,
                         float* Y_dev);
/* nsteps RK4 steps of size dt (<= 0: the case's time step) of the p = 1
 * solver s with the corrected tendency; elements whose correction has a
 * 2-norm above `bound` get none (the plain residual), counted over all
 * stages in *n_fallback */
int dg_closure_step(dg_solver* s, dg_mlp* m, double dt, double bound, int nsteps, long* n_fallback);
/* dg_closure_step without any memory allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The feature rows, corrections and fallback counter live in
 * a static device array of the library (14.4 MB, held by each GPU the
 * library runs on until the process ends) instead of buffers on the
 * handle; it takes at most 32768 elements (32^3) and returns
 * DG_ERR_ARGUMENT above that.  Every call on one GPU shares that GPU's
 * array, which keeps the last call's contents: calls that overlap on one
 * GPU (several host threads, several streams) race.  Results equal
 * dg_closure_step's. */
int dg_closure_step_noalloc(dg_solver* s, dg_mlp* m, double dt, double bound, int nsteps, long* n_fallback);

#ifdef __cplusplus
}
#endif
#endif /* DG_AMD_H */
