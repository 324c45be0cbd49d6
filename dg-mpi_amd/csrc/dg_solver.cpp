// C ABI of the DG residual + RK path (include/dg_amd.h).
#include <cstdlib>
#include <algorithm>
#include "dg_amd.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "dg_config.hpp"
#include "dg_kernels.hpp"
#include "dg_tables.hpp"

#include "dg_solver_impl.hpp"


namespace {

thread_local std::string g_create_error;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::ostringstream o;
        o << what << ": " << hipGetErrorString(e);
        throw DeviceError(o.str());
    }
}

// run f, map its exception to a status and leave the message in err
template <class F>
int guarded_to(std::string& err, F&& f) {
    try {
        f();
        err.clear();
        return DG_OK;
    } catch (const dg::InputError& e) {
        err = e.what();
        return DG_ERR_INPUT;
    } catch (const dg::UnsupportedError& e) {
        err = e.what();
        return DG_ERR_UNSUPPORTED;
    } catch (const DeviceError& e) {
        err = e.what();
        return DG_ERR_DEVICE;
    } catch (const std::invalid_argument& e) {
        err = e.what();
        return DG_ERR_ARGUMENT;
    } catch (const std::domain_error& e) {
        err = e.what();
        return DG_ERR_NUMERICS;
    } catch (const std::bad_alloc&) {
        err = "host out of memory";
        return DG_ERR_DEVICE;
    } catch (const std::exception& e) {
        err = e.what();
        return DG_ERR_INPUT;
    }
}

template <class F>
int guarded(dg_solver* s, F&& f) {
    return guarded_to(s ? s->err : g_create_error, f);
}

void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string(what) + " is NULL");
}

hipStream_t pick_stream(dg_solver* s, void* stream) {
    return stream ? (hipStream_t)stream : s->stream;
}

void set_device(const dg_solver* s) { hip_check(hipSetDevice(s->device), "hipSetDevice"); }

// Neumaier-compensated sum of x[i*stride + off]
double csum(const std::vector<double>& x, size_t n, size_t stride, size_t off) {
    double sum = 0.0, c = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const double v = x[i * stride + off];
        const double t = sum + v;
        if (std::fabs(sum) >= std::fabs(v)) c += (sum - t) + v;
        else c += (v - t) + sum;
        sum = t;
    }
    return sum + c;
}

double max_wavespeed(dg_solver* s, const double* U) {
    hip_check(dg::launch_wavespeed(s->P, U, s->work, s->dtab, s->geo, s->stream), "wavespeed kernel");
    std::vector<double> h(s->geo.E);
    hip_check(hipMemcpyAsync(h.data(), s->work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
              "hipMemcpy");
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    double m = 0.0;
    for (double v : h) {
        if (!std::isfinite(v)) throw std::domain_error("non-finite state (wave speed)");
        if (v > m) m = v;
    }
    return m;
}

// one time step of the configured stepper on U0 (stream-ordered)
void step_once(dg_solver* s, double dt) {
    using namespace dg;
    StageArgs a{};
    a.dt = dt;
    a.cmp = s->cmp;
    hipStream_t st = s->stream;
    const int flux = (int)s->cfg.flux;
    if (s->cfg.stepper == Stepper::FE) {
        // the stage reads neighbours of U0 while other elements update, so
        // the new state goes to Ua and is copied back (the state pointer
        // handed out by dg_state_device_ptr stays valid)
        a.Ust = s->U0; a.U0 = s->U0; a.Uout = s->Ua; a.mode = MODE_FE;
        hip_check(launch_stage(s->P, flux, a, s->dtab, s->geo, st), "stage kernel");
        hip_check(hipMemcpyAsync(s->U0, s->Ua, s->ncoef * sizeof(double), hipMemcpyDeviceToDevice, st),
                  "hipMemcpyAsync");
        s->time += dt;
        return;
    }
    // RK4 (src/numerics/timestepping/stepper.cpp:71-142), fused stages
    a.U0 = s->U0; a.acc = s->acc;
    a.Ust = s->U0; a.Uout = s->Ua; a.alpha = 0.5; a.mode = MODE_RK_FIRST;
    hip_check(launch_stage(s->P, flux, a, s->dtab, s->geo, st), "stage kernel");
    s->time += dt / 2.0;
    a.Ust = s->Ua; a.Uout = s->Ub; a.alpha = 0.5; a.mode = MODE_RK_MID;
    hip_check(launch_stage(s->P, flux, a, s->dtab, s->geo, st), "stage kernel");
    a.Ust = s->Ub; a.Uout = s->Ua; a.alpha = 1.0; a.mode = MODE_RK_MID;
    hip_check(launch_stage(s->P, flux, a, s->dtab, s->geo, st), "stage kernel");
    s->time += dt / 2.0;
    a.Ust = s->Ua; a.Uout = s->U0; a.mode = MODE_RK_LAST;
    hip_check(launch_stage(s->P, flux, a, s->dtab, s->geo, st), "stage kernel");
}

// a new state starts without carried rounding error
void clear_cmp(dg_solver* s) {
    if (s->cmp) hip_check(dg::launch_zero(s->cmp, s->ncoef * sizeof(float), s->stream), "zero kernel (cmp)");
}

// Numerics.subtract_mean (FP32 forms): the stage kernels subtract a
// reference state from the coefficients before rounding them to FP32; it is
// the domain mean of the current state (conserved by the scheme), rounded
// so that the kernels' deviation fluxes are consistent: rho_r and v_r are
// FP32 values, m_r = rho_r v_r exactly in FP64, p_r = (gamma - 1)(E_r -
// m_r.v_r / 2).  No usable mean (a non-positive density or pressure): no
// reference.
void set_reference_state(dg_solver* s) {
    std::memset(s->geo.uref, 0, sizeof(s->geo.uref));
    std::memset(s->geo.vref, 0, sizeof(s->geo.vref));
    s->geo.pref = 0.0;
    // the stage subtracts the reference's coefficients for bases whose
    // constant function is their first function or the sum of all of them
    // (Legendre, Lagrange)
    if (!s->cfg.subtract_mean || s->tab.one_kind == 2) return;
    hip_check(dg::launch_eval(s->P, 1, s->time, s->U0, s->work, s->dtab, s->geo, s->ic, s->stream), "eval kernel");
    std::vector<double> h((size_t)s->geo.E * 6);
    hip_check(hipMemcpyAsync(h.data(), s->work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
              "hipMemcpy");
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    double I[5];
    for (int v = 0; v < 5; ++v) I[v] = csum(h, s->geo.E, 6, v);
    const double vol = s->geo.period[0] * s->geo.period[1] * s->geo.period[2];
    const double rho = (double)(float)(I[0] / vol);
    if (!(rho > 0.0) || !std::isfinite(rho)) return;
    double u[5], v[3], mv = 0.0;
    u[0] = rho;
    for (int d = 0; d < 3; ++d) {
        v[d] = (double)(float)(I[1 + d] / I[0]);
        u[1 + d] = rho * v[d];            // exact: a product of two FP32 values
        mv += u[1 + d] * v[d];
    }
    u[4] = I[4] / vol;
    const double p = (s->cfg.gamma - 1.0) * (u[4] - 0.5 * mv);
    if (!(p > 0.0) || !std::isfinite(p) || !std::isfinite(mv)) return;
    for (int k = 0; k < 5; ++k) s->geo.uref[k] = u[k];
    for (int d = 0; d < 3; ++d) s->geo.vref[d] = v[d];
    s->geo.pref = p;
}

void free_all(dg_solver* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (double* p : {s->U0, s->Ua, s->Ub, s->acc, s->work, s->X, s->imm, s->nev})
        if (p) (void)hipFree(p);
    if (s->cmp) (void)hipFree(s->cmp);
    s->cmp = nullptr;
    if (s->frec) (void)hipFree(s->frec);
    s->frec = nullptr;
    s->nev = nullptr;
    if (s->dtab) (void)hipFree(s->dtab);
    s->dtab = nullptr;
    if (s->cl_X) (void)hipFree(s->cl_X);
    if (s->cl_C) (void)hipFree(s->cl_C);
    if (s->cl_cnt) (void)hipFree(s->cl_cnt);
    s->cl_X = s->cl_C = nullptr;
    s->cl_cnt = nullptr;
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    s->U0 = s->Ua = s->Ub = s->acc = s->work = s->X = s->imm = nullptr;
    s->stream = s->own_stream = nullptr;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

struct MeshInput {
    long nnodes = 0, E = 0, F = 0;
    const double* coords = nullptr;
    const long* e2n = nullptr;
    const int* iface6 = nullptr;
};

int create(const char* toml_text, const MeshInput* mesh, dg_solver** out) {
    if (out) *out = nullptr;
    dg_solver* s = nullptr;
    int rc = guarded(nullptr, [&] {
        need(toml_text, "toml_text");
        need(out, "out");
        dg::CaseConfig cfg = dg::parse_case(toml_text, mesh != nullptr);
        // a given mesh: hexahedra of 8 nodes in 3D, quadrilaterals of 4 in 2D
        const int dim = cfg.dim, nv = dim == 2 ? 4 : 8;
        if (mesh) {
            if (mesh->E < 1 || mesh->nnodes < 1 || mesh->F < 1 || mesh->E > 0x7fffffffL)
                throw dg::InputError("mesh: empty or too large");
            need(mesh->coords, "coords");
            need(mesh->e2n, "elem_nodes");
            need(mesh->iface6, "iface_data");
            for (long i = 0; i < mesh->E * nv; ++i)
                if (mesh->e2n[i] < 0 || mesh->e2n[i] >= mesh->nnodes)
                    throw dg::InputError("mesh: element node index out of range");
            for (long i = 0; i < mesh->nnodes * dim; ++i)
                if (!std::isfinite(mesh->coords[i])) throw dg::InputError("mesh: non-finite node coordinate");
            cfg.nelem[0] = (int)mesh->E;
            cfg.nelem[1] = cfg.nelem[2] = 1;
        }
        // host tables of the general path (warped box or a given mesh):
        // element nodes, face records and normal points; mesh errors are
        // reported here, before any device work
        std::vector<double> X;
        dg::FaceTables ft;
        if (mesh) {
            X.resize((size_t)mesh->E * nv * dim);
            for (long e = 0; e < mesh->E; ++e)
                for (int n = 0; n < nv; ++n)
                    for (int d = 0; d < dim; ++d)
                        X[((size_t)e * nv + n) * dim + d] = mesh->coords[mesh->e2n[e * nv + n] * dim + d];
            ft = dim == 2 ? dg::mesh_face_tables_2d(cfg, mesh->E, mesh->F, mesh->iface6)
                          : dg::mesh_face_tables(cfg, mesh->E, mesh->F, mesh->iface6);
        } else if (cfg.warp > 0.0) {
            X = dg::element_nodes(cfg);
            ft = dg::box_face_tables(cfg);
        }
        s = new dg_solver();
        s->cfg = cfg;
        if (mesh) {
            s->mesh_xyz.assign(mesh->coords, mesh->coords + mesh->nnodes * dim);
            s->mesh_e2n.assign(mesh->e2n, mesh->e2n + mesh->E * nv);
        }
        s->P = cfg.order;
        s->tab = dg::make_tables(cfg);
        s->geo = dg::make_geometry(cfg);
        s->ic = dg::make_ic(cfg);
        s->dt = cfg.timestep;
        const size_t nb = (size_t)(s->P + 1) * (s->P + 1) * (cfg.dim == 3 ? (s->P + 1) : 1);
        s->ns = cfg.dim + 2;
        s->ncoef = (size_t)s->geo.E * nb * s->ns;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
            throw DeviceError("no HIP device available");
        hip_check(hipGetDevice(&s->device), "hipGetDevice");
        hip_check(hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking), "hipStreamCreate");
        s->stream = s->own_stream;
        const size_t bytes = s->ncoef * sizeof(double);
        hip_check(hipMalloc(&s->U0, bytes), "hipMalloc(U0)");
        hip_check(hipMalloc(&s->Ua, bytes), "hipMalloc(Ua)");
        hip_check(hipMalloc(&s->Ub, bytes), "hipMalloc(Ub)");
        hip_check(hipMalloc(&s->acc, bytes), "hipMalloc(acc)");
        hip_check(hipMalloc(&s->work, (size_t)s->geo.E * 6 * sizeof(double)), "hipMalloc(work)");
        hip_check(hipMalloc(&s->dtab, sizeof(dg::KTables)), "hipMalloc(tables)");
        hip_check(hipMemcpy(s->dtab, &s->tab, sizeof(dg::KTables), hipMemcpyHostToDevice), "hipMemcpy(tables)");
        hip_check(dg::launch_zero(s->U0, bytes, s->stream), "zero kernel (U0)");
        if (cfg.compensated) {
            hip_check(hipMalloc(&s->cmp, s->ncoef * sizeof(float)), "hipMalloc(cmp)");
            hip_check(dg::launch_zero(s->cmp, s->ncoef * sizeof(float), s->stream), "zero kernel (cmp)");
        }
        if (s->geo.mapped) {
            // upload the general-path tables, then form the per-element
            // inverse mass matrices (precompute_matrix_helpers,
            // src/solver/helpers.cpp:53-100)
            hip_check(hipMalloc(&s->X, X.size() * sizeof(double)), "hipMalloc(X)");
            hip_check(hipMemcpy(s->X, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy(X)");
            hip_check(hipMalloc(&s->frec, ft.frec.size() * sizeof(unsigned)), "hipMalloc(frec)");
            hip_check(hipMemcpy(s->frec, ft.frec.data(), ft.frec.size() * sizeof(unsigned), hipMemcpyHostToDevice),
                      "hipMemcpy(frec)");
            if (!ft.nev.empty()) {   // 3D only: 2D faces are straight
                hip_check(hipMalloc(&s->nev, ft.nev.size() * sizeof(double)), "hipMalloc(nev)");
                hip_check(hipMemcpy(s->nev, ft.nev.data(), ft.nev.size() * sizeof(double), hipMemcpyHostToDevice),
                          "hipMemcpy(nev)");
            }
            hip_check(hipMalloc(&s->imm, (size_t)s->geo.E * nb * nb * sizeof(double)), "hipMalloc(imm)");
            s->geo.X = s->X;
            s->geo.imm = s->imm;
            s->geo.frec = s->frec;
            s->geo.nev = s->nev;
            hip_check(dg::launch_mass(s->P, s->imm, s->work, s->dtab, s->geo, s->stream), "mass kernel");
            std::vector<double> jm(s->geo.E);
            hip_check(hipMemcpyAsync(jm.data(), s->work, jm.size() * sizeof(double), hipMemcpyDeviceToHost,
                                     s->stream), "hipMemcpy");
            hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
            double m = jm[0];
            for (double v : jm) m = std::fmin(m, v);
            if (!(m > 0.0)) throw dg::InputError("mesh: an element folds (|J| <= 0 at a quadrature point)");
        }
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    });
    if (rc != DG_OK) {
        free_all(s);
        delete s;
        return rc;
    }
    *out = s;
    return DG_OK;
}

}  // namespace

extern "C" {

int dg_create(const char* toml_text, dg_solver** out) { return create(toml_text, nullptr, out); }

int dg_create_mesh(const char* toml_text, long num_nodes, const double* node_coords, long num_elems,
                   const long* elem_nodes, long num_ifaces, const int* iface_data, dg_solver** out) {
    MeshInput m;
    m.nnodes = num_nodes;
    m.coords = node_coords;
    m.E = num_elems;
    m.e2n = elem_nodes;
    m.F = num_ifaces;
    m.iface6 = iface_data;
    return create(toml_text, &m, out);
}

int dg_destroy(dg_solver* s) {
    if (!s) return DG_OK;
    free_all(s);
    delete s;
    return DG_OK;
}

const char* dg_last_error(const dg_solver* s) {
    return s ? s->err.c_str() : g_create_error.c_str();
}

int dg_get_info(const dg_solver* s, dg_info* info) {
    return guarded(const_cast<dg_solver*>(s), [&] {
        need(s, "solver");
        need(info, "info");
        std::memset(info, 0, sizeof(*info));
        const auto& c = s->cfg;
        info->num_elems = s->geo.E;
        info->order = c.order;
        info->num_basis = (c.order + 1) * (c.order + 1) * (c.dim == 3 ? c.order + 1 : 1);
        info->num_state_vars = s->ns;
        info->dim = c.dim;
        info->precision = c.proj ? 16 : (c.fp32 ? 32 : 64);
        for (int d = 0; d < 3; ++d) {
            info->nelem[d] = c.nelem[d];
            info->lower[d] = c.lower[d];
            info->upper[d] = c.upper[d];
        }
        info->gamma = c.gamma;
        info->gas_constant = c.gas_constant;
        info->timestep = s->dt;
        info->end_time = c.end_time;
        info->time = s->time;
        info->num_iterations = s->dt > 0 ? (int)(c.end_time / s->dt) : 0;
        info->stepper = (int)c.stepper;
        info->flux = (int)c.flux;
        info->basis = (int)c.basis;
        info->ic = (int)c.ic;
    });
}

int dg_get_mesh(const dg_solver* s, long* num_nodes, double* node_coords, long* elem_nodes) {
    return guarded(const_cast<dg_solver*>(s), [&] {
        need(s, "solver");
        need(num_nodes, "num_nodes");
        const auto& c = s->cfg;
        if (!s->mesh_e2n.empty()) {
            *num_nodes = (long)(s->mesh_xyz.size() / s->cfg.dim);
            if (node_coords) std::memcpy(node_coords, s->mesh_xyz.data(), s->mesh_xyz.size() * sizeof(double));
            if (elem_nodes) std::memcpy(elem_nodes, s->mesh_e2n.data(), s->mesh_e2n.size() * sizeof(long));
            return;
        }
        if (c.dim == 2) {
            // quadrilaterals: (nx+1)(ny+1) grid nodes i + (nx+1) j, element
            // e = ix + nx iy with node 2j + i of its frame
            const long nx = c.nelem[0], ny = c.nelem[1], Nx = nx + 1;
            *num_nodes = Nx * (ny + 1);
            if (node_coords)
                for (long j = 0; j <= ny; ++j)
                    for (long i = 0; i <= nx; ++i) {
                        double* x = node_coords + 2 * (i + Nx * j);
                        x[0] = c.lower[0] + (c.upper[0] - c.lower[0]) * (double)i / (double)nx;
                        x[1] = c.lower[1] + (c.upper[1] - c.lower[1]) * (double)j / (double)ny;
                    }
            if (elem_nodes)
                for (long j = 0; j < ny; ++j)
                    for (long i = 0; i < nx; ++i)
                        for (int n = 0; n < 4; ++n)
                            elem_nodes[4 * (i + nx * j) + n] = (i + (n & 1)) + Nx * (j + (n >> 1));
            return;
        }
        // the box: (n+1)^3 grid nodes i + (nx+1)(j + (ny+1) k), the warp
        // displacement included; element e = ix + nx (iy + ny iz)
        const long nx = c.nelem[0], ny = c.nelem[1], nz = c.nelem[2];
        const long Nx = nx + 1, Ny = ny + 1;
        *num_nodes = Nx * Ny * (nz + 1);
        if (node_coords)
            for (long k = 0; k <= nz; ++k)
                for (long j = 0; j <= ny; ++j)
                    for (long i = 0; i <= nx; ++i) dg::box_node(c, i, j, k, node_coords + 3 * (i + Nx * (j + Ny * k)));
        if (elem_nodes)
            for (long k = 0; k < nz; ++k)
                for (long j = 0; j < ny; ++j)
                    for (long i = 0; i < nx; ++i) {
                        long* en = elem_nodes + 8 * (i + nx * (j + ny * k));
                        for (int n = 0; n < 8; ++n)
                            en[n] = (i + (n & 1)) + Nx * ((j + ((n >> 1) & 1)) + Ny * (k + (n >> 2)));
                    }
    });
}

int dg_init_state(dg_solver* s) {
    return guarded(s, [&] {
        need(s, "solver");
        set_device(s);
        s->time = 0.0;
        hip_check(dg::launch_init_state(s->P, s->time, s->U0, s->dtab, s->geo, s->ic, s->stream), "init kernel");
        clear_cmp(s);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        set_reference_state(s);
        if (s->cfg.timestep <= 0.0) {
            // CFL rule (new): dt = cfl * min(h) / max(|u| + c) of the initial state
            const double amax = max_wavespeed(s, s->U0);
            if (!(amax > 0.0)) throw std::domain_error("zero wave speed: cannot derive dt from Stepper.cfl");
            const double hmin = s->cfg.dim == 2 ? std::fmin(s->geo.h[0], s->geo.h[1])
                                                : std::fmin(s->geo.h[0], std::fmin(s->geo.h[1], s->geo.h[2]));
            s->dt = s->cfg.cfl * hmin / amax;
        }
    });
}

int dg_set_state(dg_solver* s, const double* host, int layout) {
    return guarded(s, [&] {
        need(s, "solver");
        need(host, "host");
        set_device(s);
        // stream-ordered work queued on the handle's stream (a residual of
        // the state buffer, a step) finishes before U0 is overwritten
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        if (layout == DG_LAYOUT_RIGHT) {
            hip_check(hipMemcpy(s->U0, host, s->ncoef * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        } else if (layout == DG_LAYOUT_LEFT) {
            const size_t E = s->geo.E, ns = s->ns, nb = s->ncoef / (E * ns);
            std::vector<double> t(s->ncoef);
            for (size_t v = 0; v < ns; ++v)
                for (size_t b = 0; b < nb; ++b)
                    for (size_t e = 0; e < E; ++e) t[(e * nb + b) * ns + v] = host[(v * nb + b) * E + e];
            hip_check(hipMemcpy(s->U0, t.data(), s->ncoef * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        } else {
            throw std::invalid_argument("unknown layout");
        }
        clear_cmp(s);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        set_reference_state(s);
    });
}

int dg_get_state(const dg_solver* cs, double* host, int layout) {
    dg_solver* s = const_cast<dg_solver*>(cs);
    return guarded(s, [&] {
        need(s, "solver");
        need(host, "host");
        set_device(s);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        if (layout == DG_LAYOUT_RIGHT) {
            hip_check(hipMemcpy(host, s->U0, s->ncoef * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        } else if (layout == DG_LAYOUT_LEFT) {
            const size_t E = s->geo.E, ns = s->ns, nb = s->ncoef / (E * ns);
            std::vector<double> t(s->ncoef);
            hip_check(hipMemcpy(t.data(), s->U0, s->ncoef * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
            for (size_t v = 0; v < ns; ++v)
                for (size_t b = 0; b < nb; ++b)
                    for (size_t e = 0; e < E; ++e) host[(v * nb + b) * E + e] = t[(e * nb + b) * ns + v];
        } else {
            throw std::invalid_argument("unknown layout");
        }
    });
}

int dg_state_device_ptr(dg_solver* s, double** dev) {
    return guarded(s, [&] {
        need(s, "solver");
        need(dev, "dev");
        *dev = s->U0;
    });
}

int dg_residual(dg_solver* s, const double* U_dev, double* R_dev, void* stream) {
    return guarded(s, [&] {
        need(s, "solver");
        need(U_dev, "U_dev");
        need(R_dev, "R_dev");
        if ((const void*)U_dev == (const void*)R_dev) throw std::invalid_argument("U_dev and R_dev alias");
        set_device(s);
        dg::StageArgs a{};
        a.Ust = U_dev;
        a.R = R_dev;
        a.mode = dg::MODE_RESIDUAL;
        hip_check(dg::launch_stage(s->P, (int)s->cfg.flux, a, s->dtab, s->geo, pick_stream(s, stream)),
                  "stage kernel");
    });
}

int dg_apply_inv_mass(dg_solver* s, double dt, const double* R_dev, double* dU_dev, void* stream) {
    return guarded(s, [&] {
        need(s, "solver");
        need(R_dev, "R_dev");
        need(dU_dev, "dU_dev");
        set_device(s);
        hip_check(dg::launch_inv_mass(s->P, dt, R_dev, dU_dev, s->dtab, s->geo, pick_stream(s, stream)),
                  "inverse mass kernel");
    });
}

int dg_step(dg_solver* s, double dt, int nsteps) {
    return guarded(s, [&] {
        need(s, "solver");
        if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
        set_device(s);
        if (dt <= 0.0) dt = s->dt;
        if (!(dt > 0.0)) throw std::invalid_argument("no time step: give dt, Stepper.timestep or call dg_init_state with Stepper.cfl");
        for (int i = 0; i < nsteps; ++i) step_once(s, dt);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    });
}

int dg_solve(dg_solver* s) {
    return guarded(s, [&] {
        need(s, "solver");
        if (!(s->dt > 0.0)) throw std::invalid_argument("no time step: call dg_init_state first with Stepper.cfl");
        set_device(s);
        const int n = (int)(s->cfg.end_time / s->dt);     // src/solver/base.cpp:383
        for (int i = 0; i < n; ++i) step_once(s, s->dt);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    });
}

int dg_l2_error(dg_solver* s, double t, double err[5]) {
    return guarded(s, [&] {
        need(s, "solver");
        need(err, "err");
        set_device(s);
        hip_check(dg::launch_eval(s->P, 0, t, s->U0, s->work, s->dtab, s->geo, s->ic, s->stream), "eval kernel");
        std::vector<double> h((size_t)s->geo.E * 6);
        hip_check(hipMemcpyAsync(h.data(), s->work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
                  "hipMemcpy");
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        // total volume: sum of element volumes on the (p+2)^3 rule
        // (src/mesh/tools.cpp:27-47, src/post/main.cpp:49-54); on the mapped
        // box the sum of w |J| that the eval kernel returns per element
        // (the (2p+2)^3 rule, exact for trilinear |J| as the (p+2)^3 one is)
        double vol;
        if (s->geo.mapped) {
            vol = csum(h, s->geo.E, 6, 5);
        } else {
            double wsum = 0.0;
            for (int a = 0; a < s->P + 2; ++a)
                for (int b = 0; b < s->P + 2; ++b) {
                    if (s->cfg.dim == 2) {
                        wsum += s->tab.w[a] * s->tab.w[b];
                        continue;
                    }
                    for (int c = 0; c < s->P + 2; ++c) wsum += s->tab.w[a] * s->tab.w[b] * s->tab.w[c];
                }
            vol = (double)s->geo.E * (wsum * s->geo.jdet);
        }
        for (int v = 0; v < s->ns; ++v) err[v] = std::pow(csum(h, s->geo.E, 6, v) / vol, 1. / 2);
    });
}

int dg_integrals(dg_solver* s, double out[6]) {
    return guarded(s, [&] {
        need(s, "solver");
        need(out, "out");
        set_device(s);
        hip_check(dg::launch_eval(s->P, 1, s->time, s->U0, s->work, s->dtab, s->geo, s->ic, s->stream), "eval kernel");
        std::vector<double> h((size_t)s->geo.E * 6);
        hip_check(hipMemcpyAsync(h.data(), s->work, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s->stream),
                  "hipMemcpy");
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        for (int v = 0; v < 6; ++v) out[v] = csum(h, s->geo.E, 6, v);
    });
}

int dg_max_wavespeed(dg_solver* s, double* amax) {
    return guarded(s, [&] {
        need(s, "solver");
        need(amax, "amax");
        set_device(s);
        *amax = max_wavespeed(s, s->U0);
    });
}

int dg_set_time(dg_solver* s, double t) {
    return guarded(s, [&] {
        need(s, "solver");
        s->time = t;
    });
}

int dg_set_stream(dg_solver* s, void* stream) {
    return guarded(s, [&] {
        need(s, "solver");
        set_device(s);
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        s->stream = stream ? (hipStream_t)stream : s->own_stream;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of independent cases (dg_batch_*)
// ---------------------------------------------------------------------------

// Tuning candidates of a batch group: team-size variant (dg::batch_team) x
// cap on the log2 edge of the element bricks each case is traversed in
// (8^3 or 4^3 bricks, never larger than the case's own default)
constexpr int NCAPS = 2;
constexpr int BRICK_CAPS[NCAPS] = {3, 2};
constexpr int NCAND = 2 * NCAPS;

struct dg_batch {
    // cases of one (order, basis, flux, precision): one launch per RK stage
    struct Group {
        int P = 0, basis = 0, flux = 0, fp32 = 0, proj = 0;
        unsigned nblocks = 0;
        std::vector<int> members;             // indices into cases
        dg::BatchCase* dcases = nullptr;      // device descriptors
        unsigned* blk_case = nullptr;         // device: case (within the group) of each block
        int variant = 0;                      // team-size variant of the launch (dg::batch_team)
        int cap = BRICK_CAPS[0];              // log2 edge cap of the element bricks of its cases
        double ms[2] = {0.0, 0.0};            // one RK4 step with each team variant at the chosen cap
        double ms_cand[NCAND] = {};           // one RK4 step per tuning candidate (team variant, cap)
        hipStream_t stream = nullptr;         // the group's branch of the captured step
        hipEvent_t done = nullptr;            // end of that branch
    };
    std::vector<dg_solver*> cases;
    std::vector<Group> groups;
    std::vector<double> dts;                  // time steps the descriptors were built with
    int device = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;               // one RK4 step of every group
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;                // start of a step (captured)
    bool tuned = false;                       // variants chosen by timing
    std::string err;
};

namespace {

thread_local std::string g_batch_error;

// descriptors of a group's cases: the case's registers (as byte offsets
// from the descriptor array, the base the launch receives), dt, geometry
void upload_group(dg_batch* b, dg_batch::Group& g) {
    using namespace dg;
    std::vector<BatchCase> h(g.members.size());
    const char* base = reinterpret_cast<const char*>(g.dcases);
    auto off = [&](const void* p) { return (long long)(reinterpret_cast<const char*>(p) - base); };
    unsigned first = 0;
    for (size_t k = 0; k < g.members.size(); ++k) {
        const dg_solver* s = b->cases[g.members[k]];
        BatchCase& c = h[k];
        std::memset(&c, 0, sizeof(c));
        c.off_U0 = off(s->U0);
        c.off_Ua = off(s->Ua);
        c.off_Ub = off(s->Ub);
        c.off_acc = off(s->acc);
        c.off_cmp = s->cmp ? off(s->cmp) : -1;
        c.dt = s->dt;
        c.g = s->geo;
        // the traversal bricks of the group's cap (the elements and their
        // arithmetic are the same in any traversal order)
        for (int d = 0; d < 3; ++d) c.g.tile[d] = std::min(s->geo.tile[d], g.cap);
        c.first_block = first;
        first += s->geo.E;
    }
    // ordered on the batch stream (the group branches of a step fork from it)
    hip_check(hipMemcpyAsync(g.dcases, h.data(), h.size() * sizeof(BatchCase), hipMemcpyHostToDevice, b->stream),
              "hipMemcpy(batch cases)");
    hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

// one RK4 step of every group, to be captured: the groups are independent
// (disjoint cases), so each runs its four stages on its own branch (a
// stream forked from the batch stream and joined back at the end), and a
// group's stages overlap the other groups' launches
// (DG_BATCH_SERIAL=1 in the environment launches every group on the batch
// stream instead, one after the other: the same work, for comparisons)
hipError_t launch_step(dg_batch* b) {
    static const bool serial = std::getenv("DG_BATCH_SERIAL") != nullptr;
    hipError_t e = hipEventRecord(b->fork, b->stream);
    for (size_t gi = 0; gi < b->groups.size() && e == hipSuccess; ++gi) {
        auto& g = b->groups[gi];
        const bool branch = gi > 0 && !serial;
        hipStream_t st = branch ? g.stream : b->stream;
        if (branch) e = hipStreamWaitEvent(st, b->fork, 0);
        for (int k = 0; k < 4 && e == hipSuccess; ++k)
            e = dg::launch_stage_batch(g.P, g.flux, g.fp32, g.proj, g.variant, g.dcases, g.blk_case, g.nblocks, k,
                                       b->cases[g.members[0]]->dtab, st);
        if (e == hipSuccess && branch) e = hipEventRecord(g.done, st);
    }
    for (size_t gi = 1; gi < b->groups.size() && e == hipSuccess && !serial; ++gi)
        e = hipStreamWaitEvent(b->stream, b->groups[gi].done, 0);
    return e;
}

void capture_step(dg_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const hipError_t le = launch_step(b);
    const hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
    hip_check(le, "batched stage launch (capture)");
    hip_check(ce, "hipStreamEndCapture");
    hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

// Autotuning: TUNE_STEPS eager RK4 steps, each group's launches bracketed
// by events: first one step with each team variant at the default cap (the
// first launch of a kernel pays for code loading; not counted), then one
// step with each candidate (team variant, brick cap); each group keeps its
// fastest candidate.  Each output value is computed by one thread with the
// same operations whatever the team size or the traversal order (each
// variant being compiled on its own, multiply-add fusion may differ:
// round-off), so these are ordinary steps of the cases.
constexpr int TUNE_STEPS = 2 + NCAND;

void tune(dg_batch* b) {
    const size_t ng = b->groups.size();
    std::vector<hipEvent_t> ev(ng * 8);
    for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    for (int k = 0; k < TUNE_STEPS; ++k) {
        const int cand = k < 2 ? k : k - 2;            // candidate = variant + 2 cap index
        const int v = cand & 1, ci = cand >> 1;
        for (auto& g : b->groups) {
            g.variant = v;
            if (g.cap != BRICK_CAPS[ci]) {
                g.cap = BRICK_CAPS[ci];
                upload_group(b, g);
            }
        }
        for (int st = 0; st < 4; ++st)
            for (size_t gi = 0; gi < ng; ++gi) {
                auto& g = b->groups[gi];
                hip_check(hipEventRecord(ev[gi * 8 + st * 2], b->stream), "hipEventRecord");
                hip_check(dg::launch_stage_batch(g.P, g.flux, g.fp32, g.proj, v, g.dcases, g.blk_case, g.nblocks, st,
                                                 b->cases[g.members[0]]->dtab, b->stream),
                          "batched stage launch");
                hip_check(hipEventRecord(ev[gi * 8 + st * 2 + 1], b->stream), "hipEventRecord");
            }
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        if (k < 2) continue;
        for (size_t gi = 0; gi < ng; ++gi) {
            double ms = 0.0;
            for (int st = 0; st < 4; ++st) {
                float t = 0.f;
                hip_check(hipEventElapsedTime(&t, ev[gi * 8 + st * 2], ev[gi * 8 + st * 2 + 1]),
                          "hipEventElapsedTime");
                ms += t;
            }
            b->groups[gi].ms_cand[cand] = ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& g : b->groups) {
        int best = 0;
        for (int c = 1; c < NCAND; ++c)
            if (g.ms_cand[c] < g.ms_cand[best]) best = c;
        g.variant = best & 1;
        g.cap = BRICK_CAPS[best >> 1];
        g.ms[0] = g.ms_cand[(best >> 1) * 2];
        g.ms[1] = g.ms_cand[(best >> 1) * 2 + 1];
        upload_group(b, g);
    }
    b->tuned = true;
}

void free_batch(dg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& g : b->groups) {
        if (g.dcases) (void)hipFree(g.dcases);
        if (g.blk_case) (void)hipFree(g.blk_case);
        if (g.stream) (void)hipStreamDestroy(g.stream);
        if (g.done) (void)hipEventDestroy(g.done);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

}  // namespace

extern "C" {

int dg_batch_create(int num_cases, dg_solver* const* cases, dg_batch** out) {
    if (out) *out = nullptr;
    dg_batch* b = nullptr;
    const int rc = guarded_to(g_batch_error, [&] {
        need(cases, "cases");
        need(out, "out");
        if (num_cases < 1) throw std::invalid_argument("dg_batch_create: num_cases < 1");
        b = new dg_batch();
        for (int i = 0; i < num_cases; ++i) {
            dg_solver* s = cases[i];
            need(s, "case handle");
            for (dg_solver* o : b->cases)
                if (o == s) throw std::invalid_argument("dg_batch_create: a case appears twice");
            if (i > 0 && s->device != b->cases[0]->device)
                throw std::invalid_argument("dg_batch_create: the cases live on different devices");
            if (s->cfg.dim != 3 || s->geo.mapped)
                throw dg::UnsupportedError("dg_batch_create: case " + std::to_string(i) +
                                           " is not a 3D affine periodic box (the batched stage covers those)");
            if (s->cfg.stepper != dg::Stepper::RK4)
                throw dg::UnsupportedError("dg_batch_create: case " + std::to_string(i) + " does not use RK4");
            b->cases.push_back(s);
        }
        b->device = b->cases[0]->device;
        hip_check(hipSetDevice(b->device), "hipSetDevice");
        // group by (order, basis, flux, precision), in order of first
        // appearance; the cases of a group share their reference-element
        // tables (they depend on the order and the basis only)
        for (int i = 0; i < num_cases; ++i) {
            const dg_solver* s = b->cases[i];
            const int flux = (int)s->cfg.flux, fp32 = s->geo.fp32, basis = (int)s->cfg.basis, proj = s->geo.proj;
            dg_batch::Group* g = nullptr;
            for (auto& q : b->groups)
                if (q.P == s->P && q.basis == basis && q.flux == flux && q.fp32 == fp32 && q.proj == proj) g = &q;
            if (!g) {
                b->groups.emplace_back();
                g = &b->groups.back();
                g->P = s->P;
                g->basis = basis;
                g->flux = flux;
                g->fp32 = fp32;
                g->proj = proj;
            }
            if ((unsigned long long)g->nblocks + s->geo.E > 0x7ffffff0ULL)
                throw dg::UnsupportedError("dg_batch_create: more than 2^31 - 1 elements of one order");
            g->members.push_back(i);
            g->nblocks += s->geo.E;
        }
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
        for (size_t gi = 1; gi < b->groups.size(); ++gi) {
            // default priority, as every stream of this library
            hip_check(hipStreamCreateWithFlags(&b->groups[gi].stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&b->groups[gi].done, hipEventDisableTiming), "hipEventCreate");
        }
        for (auto& g : b->groups) {
            hip_check(hipMalloc(&g.dcases, g.members.size() * sizeof(dg::BatchCase)), "hipMalloc(batch cases)");
            hip_check(hipMalloc(&g.blk_case, (size_t)g.nblocks * sizeof(unsigned)), "hipMalloc(batch blocks)");
            std::vector<unsigned> bc(g.nblocks);
            size_t k = 0;
            for (size_t m = 0; m < g.members.size(); ++m)
                for (unsigned e = 0; e < b->cases[g.members[m]]->geo.E; ++e) bc[k++] = (unsigned)m;
            hip_check(hipMemcpy(g.blk_case, bc.data(), bc.size() * sizeof(unsigned), hipMemcpyHostToDevice),
                      "hipMemcpy(batch blocks)");
            upload_group(b, g);
        }
        b->dts.resize(b->cases.size());
        for (size_t i = 0; i < b->cases.size(); ++i) b->dts[i] = b->cases[i]->dt;
        // one RK4 step of every group (variant 0 until tuned), captured once
        // and replayed per step
        capture_step(b);
    });
    if (rc != DG_OK) {
        free_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return DG_OK;
}

int dg_batch_step(dg_batch* b, int nsteps) {
    return guarded_to(b ? b->err : g_batch_error, [&] {
        need(b, "batch");
        if (nsteps < 0) throw std::invalid_argument("nsteps < 0");
        hip_check(hipSetDevice(b->device), "hipSetDevice");
        bool changed = false;
        for (size_t i = 0; i < b->cases.size(); ++i) {
            const dg_solver* s = b->cases[i];
            if (!(s->dt > 0.0))
                throw std::invalid_argument("dg_batch_step: case " + std::to_string(i) +
                                            " has no time step (Stepper.timestep, or dg_init_state with Stepper.cfl)");
            if (s->dt != b->dts[i]) changed = true;
            // earlier work on the case's own stream comes first
            hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
        }
        if (changed) {
            for (auto& g : b->groups) upload_group(b, g);
            for (size_t i = 0; i < b->cases.size(); ++i) b->dts[i] = b->cases[i]->dt;
        }
        int n0 = 0;
        if (!b->tuned && nsteps >= TUNE_STEPS) {
            // the first steps of a long enough call pick each group's team
            // size and bricks, then the step is captured again with them
            tune(b);
            capture_step(b);
            n0 = TUNE_STEPS;
        }
        for (int n = n0; n < nsteps; ++n) hip_check(hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
        hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
        for (dg_solver* s : b->cases)
            for (int n = 0; n < nsteps; ++n) {   // as step_once advances it
                s->time += s->dt / 2.0;
                s->time += s->dt / 2.0;
            }
    });
}

int dg_batch_destroy(dg_batch* b) {
    if (!b) return DG_OK;
    free_batch(b);
    delete b;
    return DG_OK;
}

const char* dg_batch_last_error(const dg_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

int dg_batch_tuning(const dg_batch* b, int* num_groups, int* order, int* team, double* ms_default,
                    double* ms_alternate) {
    return guarded_to(b ? const_cast<dg_batch*>(b)->err : g_batch_error, [&] {
        need(b, "batch");
        need(num_groups, "num_groups");
        *num_groups = (int)b->groups.size();
        for (size_t i = 0; i < b->groups.size(); ++i) {
            const auto& g = b->groups[i];
            if (order) order[i] = g.P;
            if (team) team[i] = dg::batch_team(g.P, g.variant);
            if (ms_default) ms_default[i] = b->tuned ? g.ms[0] : 0.0;
            if (ms_alternate) ms_alternate[i] = b->tuned ? g.ms[1] : 0.0;
        }
    });
}

int dg_batch_tuning_tiles(const dg_batch* b, int* num_groups, int* num_candidates, int* brick_cap, int* cases,
                          long* elements, double* ms_candidates) {
    return guarded_to(b ? const_cast<dg_batch*>(b)->err : g_batch_error, [&] {
        need(b, "batch");
        need(num_groups, "num_groups");
        *num_groups = (int)b->groups.size();
        if (num_candidates) *num_candidates = NCAND;
        for (size_t i = 0; i < b->groups.size(); ++i) {
            const auto& g = b->groups[i];
            if (brick_cap) brick_cap[i] = g.cap;
            if (cases) cases[i] = (int)g.members.size();
            if (elements) elements[i] = (long)g.nblocks;
            if (ms_candidates)
                for (int c = 0; c < NCAND; ++c) ms_candidates[i * NCAND + c] = b->tuned ? g.ms_cand[c] : 0.0;
        }
    });
}

}  // extern "C"
