// data.h5 solution files (the reference's Writer / read_in_coefficients
// layout) and hex mesh files (the layout the reference's mesh reader takes)
// for the DG path.  Host code over the HDF5 C library.
//
// Mesh files (src/mesh/mesh.cpp:17-26,79-200): scalar datasets Dimension,
// QOrder, nElem, nNode, nIFace, nNodePerElem, nNodePerFace, nBFaceGroup
// (64-bit integers) and NodeCoords [nNode][dim] (double), Elem2Nodes
// [nElem][nNodePerElem], IFaceData [nIFace][6] (64-bit integers).
//
// Layout written (one rank), after src/io/writer.cpp:10-65 and the
// parallel-dataset convention of src/io/HDF5Wrapper.h:780-877:
//   attributes   "Number of Ranks", "Number of Dimensions", "Number of
//                Elements", "Number of Nodes", "Number of Nodes Per Element",
//                "Number of Basis Functions", "Number of State Variables"
//                (unsigned int), "Solver Final Time" (double), "Stored
//                Layout" (int, 1 = row major)
//   dataset      "Number of Elements per Partition" [ranks] unsigned int
//   per array X  X [total] (flat), X_offsets [ranks] hsize_t with attributes
//                nDims (unsigned long), nRanks (int), totalSize (hsize_t),
//                X_dims [nDims*ranks] hsize_t
//   arrays       "Node Coordinates" [nodes][3] double, "Local to Global
//                Node ID" [nodes], "Element Global Node IDs" [E][8], "Local
//                to Global Element IDs" [E] (unsigned int), "Solution
//                Coefficients" [E][nb][ns] double
#include "dg_amd_h5.h"

#include <hdf5.h>

#include "dg_config.hpp"

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct H5Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// RAII for HDF5 identifiers
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) {
        if (id < 0) throw H5Error("HDF5: cannot " + what);
    }
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

void check(herr_t e, const std::string& what) {
    if (e < 0) throw H5Error("HDF5: cannot " + what);
}

void quiet_hdf5() {
    // our own messages replace HDF5's error-stack printout
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_err.clear();
        return DG_OK;
    } catch (const ArgError& e) {
        g_err = e.what();
        return DG_ERR_ARGUMENT;
    } catch (const dg::UnsupportedError& e) {
        g_err = e.what();
        return DG_ERR_UNSUPPORTED;
    } catch (const H5Error& e) {
        g_err = e.what();
        return DG_ERR_INPUT;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return DG_ERR_DEVICE;
    } catch (const std::exception& e) {
        g_err = e.what();
        return DG_ERR_INPUT;
    }
}

template <class T>
void write_attr(hid_t loc, const char* name, hid_t type, const T& v) {
    Hid space(H5Screate(H5S_SCALAR), H5Sclose, std::string("create a scalar space for ") + name);
    Hid att(H5Acreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
            std::string("create attribute ") + name);
    check(H5Awrite(att, type, &v), std::string("write attribute ") + name);
}

template <class T>
T read_attr(hid_t loc, const char* name, hid_t type) {
    if (H5Aexists(loc, name) <= 0) throw H5Error(std::string("attribute '") + name + "' is missing");
    Hid att(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose, std::string("open attribute ") + name);
    T v{};
    check(H5Aread(att, type, &v), std::string("read attribute ") + name);
    return v;
}

void write_1d(hid_t file, const std::string& name, hid_t type, hsize_t n, const void* data) {
    Hid space(H5Screate_simple(1, &n, &n), H5Sclose, "create space for " + name);
    Hid ds(H5Dcreate2(file, name.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
           H5Dclose, "create dataset " + name);
    check(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write dataset " + name);
}

// one-rank "parallel dataset": flat data + _offsets (+ attributes) + _dims
void write_parallel(hid_t file, const std::string& name, hid_t type, const std::vector<hsize_t>& dims,
                    const void* data) {
    hsize_t total = 1;
    for (hsize_t d : dims) total *= d;
    write_1d(file, name, type, total, data);
    const hsize_t offset = 0;
    const std::string on = name + "_offsets";
    write_1d(file, on, H5T_NATIVE_ULLONG, 1, &offset);
    Hid od(H5Dopen2(file, on.c_str(), H5P_DEFAULT), H5Dclose, "open " + on);
    const unsigned long ndims = dims.empty() ? 1 : dims.size();
    const int nranks = 1;
    write_attr(od, "nDims", H5T_NATIVE_ULONG, ndims);
    write_attr(od, "nRanks", H5T_NATIVE_INT, nranks);
    write_attr(od, "totalSize", H5T_NATIVE_ULLONG, (unsigned long long)total);
    write_1d(file, name + "_dims", H5T_NATIVE_ULLONG, dims.size(), dims.data());
}

std::vector<hsize_t> read_1d_ull(hid_t file, const std::string& name) {
    if (H5Lexists(file, name.c_str(), H5P_DEFAULT) <= 0) throw H5Error("dataset '" + name + "' is missing");
    Hid ds(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose, "open " + name);
    Hid sp(H5Dget_space(ds), H5Sclose, "get the space of " + name);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    if (n < 0) throw H5Error("bad extent of " + name);
    std::vector<hsize_t> v((size_t)n);
    if (n) check(H5Dread(ds, H5T_NATIVE_ULLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), "read " + name);
    return v;
}

struct ParallelInfo {
    std::vector<hsize_t> dims;
    hsize_t total = 0, offset = 0;
};

// the reader side of the convention (HDF5Wrapper.h:920-1000), one rank
ParallelInfo parallel_info(hid_t file, const std::string& name) {
    const std::string on = name + "_offsets";
    if (H5Lexists(file, on.c_str(), H5P_DEFAULT) <= 0) throw H5Error("dataset '" + on + "' is missing");
    Hid od(H5Dopen2(file, on.c_str(), H5P_DEFAULT), H5Dclose, "open " + on);
    const unsigned long long ndims = read_attr<unsigned long long>(od, "nDims", H5T_NATIVE_ULLONG);
    const int nranks = read_attr<int>(od, "nRanks", H5T_NATIVE_INT);
    if (nranks != 1)
        throw H5Error("'" + name + "' was written by " + std::to_string(nranks) +
                      " ranks; this reader takes single-rank files");
    ParallelInfo pi;
    pi.total = read_attr<unsigned long long>(od, "totalSize", H5T_NATIVE_ULLONG);
    pi.dims = read_1d_ull(file, name + "_dims");
    if (pi.dims.size() != ndims) throw H5Error("'" + name + "_dims' does not hold nDims entries");
    const std::vector<hsize_t> off = read_1d_ull(file, on);
    if (off.size() != 1) throw H5Error("'" + on + "' does not hold one offset");
    pi.offset = off[0];
    return pi;
}

std::string errstr_path(const char* path) { return path ? std::string(path) : std::string("(null)"); }

// data.h5 of a hex (dim 3) or quadrilateral (dim 2) mesh (Writer::write,
// src/io/writer.cpp:10-65, one rank): xyz [nn][dim], e2n [E][2^dim]
void write_solution_file(const char* path, int dim, long nn, const double* xyz, long E, const long* e2n,
                         int num_basis, int num_state_vars, double time, const double* Uc, int layout) {
    if (layout != DG_LAYOUT_RIGHT && layout != DG_LAYOUT_LEFT)
        throw ArgError("layout must be DG_LAYOUT_RIGHT or DG_LAYOUT_LEFT");
    if (num_basis < 1 || num_state_vars < 1 || E < 1 || nn < 1) throw ArgError("bad sizes");
    if (E > 0xffffffffL || nn > 0xffffffffL) throw ArgError("mesh too large for 32-bit IDs");
    quiet_hdf5();
    const unsigned npe = dim == 2 ? 4u : 8u;
    std::vector<unsigned> nid((size_t)nn), eid((size_t)E), en((size_t)E * npe);
    for (long n = 0; n < nn; ++n) nid[(size_t)n] = (unsigned)n;
    for (long e = 0; e < E; ++e) eid[(size_t)e] = (unsigned)e;
    for (size_t i = 0; i < en.size(); ++i) {
        if (e2n[i] < 0 || e2n[i] >= nn) throw ArgError("element node index out of range");
        en[i] = (unsigned)e2n[i];
    }
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "create " + errstr_path(path));
    write_attr(file, "Number of Ranks", H5T_NATIVE_UINT, 1u);
    write_attr(file, "Number of Dimensions", H5T_NATIVE_UINT, (unsigned)dim);
    write_attr(file, "Number of Elements", H5T_NATIVE_UINT, (unsigned)E);
    write_attr(file, "Number of Nodes", H5T_NATIVE_UINT, (unsigned)nn);
    write_attr(file, "Number of Nodes Per Element", H5T_NATIVE_UINT, npe);
    write_attr(file, "Number of Basis Functions", H5T_NATIVE_UINT, (unsigned)num_basis);
    write_attr(file, "Number of State Variables", H5T_NATIVE_UINT, (unsigned)num_state_vars);
    write_attr(file, "Solver Final Time", H5T_NATIVE_DOUBLE, time);
    const unsigned epart = (unsigned)E;
    write_1d(file, "Number of Elements per Partition", H5T_NATIVE_UINT, 1, &epart);
    write_attr(file, "Stored Layout", H5T_NATIVE_INT, layout == DG_LAYOUT_RIGHT ? 1 : 0);
    write_parallel(file, "Node Coordinates", H5T_NATIVE_DOUBLE, {(hsize_t)nn, (hsize_t)dim}, xyz);
    write_parallel(file, "Local to Global Node ID", H5T_NATIVE_UINT, {(hsize_t)nn}, nid.data());
    write_parallel(file, "Element Global Node IDs", H5T_NATIVE_UINT, {(hsize_t)E, (hsize_t)npe}, en.data());
    write_parallel(file, "Local to Global Element IDs", H5T_NATIVE_UINT, {(hsize_t)E}, eid.data());
    write_parallel(file, "Solution Coefficients", H5T_NATIVE_DOUBLE,
                   {(hsize_t)E, (hsize_t)num_basis, (hsize_t)num_state_vars}, Uc);
}

// ---- mesh files (the reference's mesh reader, src/mesh/mesh.cpp:17-26,79-200) ----

bool has_link(hid_t file, const char* name) { return H5Lexists(file, name, H5P_DEFAULT) > 0; }

long read_scalar(hid_t file, const char* name, bool required, long dflt) {
    if (!has_link(file, name)) {
        if (required) throw H5Error(std::string("mesh dataset '") + name + "' is missing");
        return dflt;
    }
    Hid ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose, std::string("open ") + name);
    Hid sp(H5Dget_space(ds), H5Sclose, std::string("get the space of ") + name);
    if (H5Sget_simple_extent_npoints(sp) != 1) throw H5Error(std::string("'") + name + "' is not a scalar");
    long v = 0;
    check(H5Dread(ds, H5T_NATIVE_LONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), std::string("read ") + name);
    return v;
}

// whole array `name` of `count` values as mem_type
void read_array(hid_t file, const char* name, hid_t mem_type, long count, void* out) {
    if (!has_link(file, name)) throw H5Error(std::string("mesh dataset '") + name + "' is missing");
    Hid ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose, std::string("open ") + name);
    Hid sp(H5Dget_space(ds), H5Sclose, std::string("get the space of ") + name);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    if (n != count)
        throw H5Error(std::string("'") + name + "' holds " + std::to_string((long long)n) + " values, expected " +
                      std::to_string(count));
    if (count) check(H5Dread(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out), std::string("read ") + name);
}

void write_scalar(hid_t file, const char* name, long v) {
    Hid space(H5Screate(H5S_SCALAR), H5Sclose, std::string("create a scalar space for ") + name);
    Hid ds(H5Dcreate2(file, name, H5T_STD_I64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           std::string("create dataset ") + name);
    check(H5Dwrite(ds, H5T_NATIVE_LONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), std::string("write ") + name);
}

void write_2d(hid_t file, const char* name, hid_t file_type, hid_t mem_type, hsize_t rows, hsize_t cols,
              const void* data) {
    const hsize_t dims[2] = {rows, cols};
    Hid space(H5Screate_simple(2, dims, dims), H5Sclose, std::string("create space for ") + name);
    Hid ds(H5Dcreate2(file, name, file_type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
           std::string("create dataset ") + name);
    if (rows > 0 && cols > 0)
        check(H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), std::string("write ") + name);
}

dg_h5_mesh_info mesh_info(hid_t file) {
    dg_h5_mesh_info m{};
    m.dim = read_scalar(file, "Dimension", true, 0);
    m.qorder = read_scalar(file, "QOrder", true, 0);
    m.num_elems = read_scalar(file, "nElem", true, 0);
    m.num_nodes = read_scalar(file, "nNode", true, 0);
    m.num_ifaces = read_scalar(file, "nIFace", true, 0);
    m.num_nodes_per_elem = read_scalar(file, "nNodePerElem", true, 0);
    // absent in old files; the reference then takes 1 (mesh.cpp:98-107)
    m.num_nodes_per_face = read_scalar(file, "nNodePerFace", false, 1);
    m.num_bface_groups = read_scalar(file, "nBFaceGroup", false, 0);
    if (m.dim < 1 || m.num_elems < 0 || m.num_nodes < 0 || m.num_ifaces < 0 || m.num_nodes_per_elem < 1)
        throw H5Error("mesh file sizes are negative or zero");
    return m;
}

}  // namespace

extern "C" {

const char* dg_h5_last_error(void) { return g_err.c_str(); }

int dg_h5_write_box(const char* path, const int nelem[3], const double lower[3], const double upper[3],
                    int num_basis, int num_state_vars, double time, const double* Uc, int layout) {
    return guarded([&] {
        if (!path || !nelem || !lower || !upper || !Uc) throw ArgError("dg_h5_write_box: NULL argument");
        for (int d = 0; d < 3; ++d)
            if (nelem[d] < 1 || !(upper[d] > lower[d])) throw ArgError("dg_h5_write_box: bad box");
        const long nx = nelem[0], ny = nelem[1], nz = nelem[2];
        const long E = nx * ny * nz;
        const long Nx = nx + 1, Ny = ny + 1, Nz = nz + 1, nn = Nx * Ny * Nz;
        // mesh arrays of the structured box (reference hex node n = 4k + 2j + i)
        std::vector<double> xyz((size_t)nn * 3);
        std::vector<long> e2n((size_t)E * 8);
        for (long k = 0; k < Nz; ++k)
            for (long j = 0; j < Ny; ++j)
                for (long i = 0; i < Nx; ++i) {
                    const long n = i + Nx * (j + Ny * k);
                    const long ijk[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d)
                        xyz[(size_t)n * 3 + d] =
                            lower[d] + (upper[d] - lower[d]) * (double)ijk[d] / (double)nelem[d];
                }
        for (long k = 0; k < nz; ++k)
            for (long j = 0; j < ny; ++j)
                for (long i = 0; i < nx; ++i) {
                    const long e = i + nx * (j + ny * k);
                    for (int n = 0; n < 8; ++n)
                        e2n[(size_t)e * 8 + n] = (i + (n & 1)) + Nx * ((j + ((n >> 1) & 1)) + Ny * (k + (n >> 2)));
                }
        write_solution_file(path, 3, nn, xyz.data(), E, e2n.data(), num_basis, num_state_vars, time, Uc, layout);
    });
}

int dg_h5_read_header(const char* path, dg_h5_header* h) {
    return guarded([&] {
        if (!path || !h) throw ArgError("dg_h5_read_header: NULL argument");
        quiet_hdf5();
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "open " + errstr_path(path));
        dg_h5_header r{};
        r.num_ranks = (int)read_attr<unsigned>(file, "Number of Ranks", H5T_NATIVE_UINT);
        r.dim = (int)read_attr<unsigned>(file, "Number of Dimensions", H5T_NATIVE_UINT);
        r.num_elems = (long)read_attr<unsigned>(file, "Number of Elements", H5T_NATIVE_UINT);
        r.num_nodes = (long)read_attr<unsigned>(file, "Number of Nodes", H5T_NATIVE_UINT);
        r.num_nodes_per_elem = (int)read_attr<unsigned>(file, "Number of Nodes Per Element", H5T_NATIVE_UINT);
        r.num_basis = read_attr<int>(file, "Number of Basis Functions", H5T_NATIVE_INT);
        r.num_state_vars = read_attr<int>(file, "Number of State Variables", H5T_NATIVE_INT);
        r.time = read_attr<double>(file, "Solver Final Time", H5T_NATIVE_DOUBLE);
        r.stored_layout = read_attr<int>(file, "Stored Layout", H5T_NATIVE_INT) != 0 ? 1 : 0;
        *h = r;
    });
}

int dg_h5_read_coefficients(const char* path, double* Uc, long capacity, int layout) {
    return guarded([&] {
        if (!path || !Uc) throw ArgError("dg_h5_read_coefficients: NULL argument");
        if (layout != DG_LAYOUT_RIGHT && layout != DG_LAYOUT_LEFT)
            throw ArgError("dg_h5_read_coefficients: layout must be DG_LAYOUT_RIGHT or DG_LAYOUT_LEFT");
        quiet_hdf5();
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "open " + errstr_path(path));
        const int nb = read_attr<int>(file, "Number of Basis Functions", H5T_NATIVE_INT);
        const int ns = read_attr<int>(file, "Number of State Variables", H5T_NATIVE_INT);
        const bool row = read_attr<int>(file, "Stored Layout", H5T_NATIVE_INT) != 0;
        // elements of this (single) partition, as the reference reads them
        const std::vector<hsize_t> part = read_1d_ull(file, "Number of Elements per Partition");
        if (part.size() != 1) throw H5Error("file has more than one partition");
        const long E = (long)part[0];
        const ParallelInfo pi = parallel_info(file, "Solution Coefficients");
        hsize_t local = 1;
        for (hsize_t d : pi.dims) local *= d;
        if (local != (hsize_t)E * nb * ns || pi.offset + local > pi.total)
            throw H5Error("'Solution Coefficients' does not hold E*nb*ns values");
        if ((long)local > capacity) throw ArgError("dg_h5_read_coefficients: buffer too small");
        std::vector<double> buf((size_t)local);
        {
            Hid ds(H5Dopen2(file, "Solution Coefficients", H5P_DEFAULT), H5Dclose, "open Solution Coefficients");
            Hid fs(H5Dget_space(ds), H5Sclose, "get the space of Solution Coefficients");
            if ((hsize_t)H5Sget_simple_extent_npoints(fs) != pi.total)
                throw H5Error("'Solution Coefficients' size differs from totalSize");
            const hsize_t one = 1, off = pi.offset, cnt = local;
            check(H5Sselect_hyperslab(fs, H5S_SELECT_SET, &off, nullptr, &one, &cnt), "select hyperslab");
            Hid ms(H5Screate_simple(1, &cnt, &cnt), H5Sclose, "create memory space");
            check(H5Dread(ds, H5T_NATIVE_DOUBLE, ms, fs, H5P_DEFAULT, buf.data()), "read Solution Coefficients");
        }
        // stored row major: (e, b, s) at (e*nb + b)*ns + s; column major: (s*nb + b)*E + e
        const bool want_row = layout == DG_LAYOUT_RIGHT;
        if (row == want_row) {
            std::memcpy(Uc, buf.data(), buf.size() * sizeof(double));
        } else {
            for (long e = 0; e < E; ++e)
                for (int b = 0; b < nb; ++b)
                    for (int s = 0; s < ns; ++s) {
                        const size_t r = ((size_t)e * nb + b) * ns + s, c = ((size_t)s * nb + b) * E + e;
                        if (row) Uc[c] = buf[r];
                        else Uc[r] = buf[c];
                    }
        }
    });
}

int dg_h5_read_mesh_info(const char* path, dg_h5_mesh_info* info) {
    return guarded([&] {
        if (!path || !info) throw ArgError("dg_h5_read_mesh_info: NULL argument");
        quiet_hdf5();
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "open " + errstr_path(path));
        *info = mesh_info(file);
    });
}

int dg_h5_read_mesh(const char* path, double* node_coords, long coords_capacity, long* elem_nodes,
                    long elem_nodes_capacity, int* iface_data, long iface_capacity) {
    return guarded([&] {
        if (!path || !node_coords || !elem_nodes || !iface_data) throw ArgError("dg_h5_read_mesh: NULL argument");
        quiet_hdf5();
        Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "open " + errstr_path(path));
        const dg_h5_mesh_info m = mesh_info(file);
        const long nc = m.num_nodes * m.dim, ne = m.num_elems * m.num_nodes_per_elem, nf = m.num_ifaces * 6;
        if (nc > coords_capacity || ne > elem_nodes_capacity || nf > iface_capacity)
            throw ArgError("dg_h5_read_mesh: buffer too small");
        read_array(file, "NodeCoords", H5T_NATIVE_DOUBLE, nc, node_coords);
        read_array(file, "Elem2Nodes", H5T_NATIVE_LONG, ne, elem_nodes);
        read_array(file, "IFaceData", H5T_NATIVE_INT, nf, iface_data);
    });
}

int dg_h5_write_mesh(const char* path, int dim, long num_nodes, const double* node_coords, long num_elems,
                     const long* elem_nodes, long num_ifaces, const int* iface_data) {
    return guarded([&] {
        if (!path || !node_coords || !elem_nodes || (num_ifaces > 0 && !iface_data))
            throw ArgError("dg_h5_write_mesh: NULL argument");
        if (dim != 2 && dim != 3) throw ArgError("dg_h5_write_mesh: dim must be 2 or 3");
        if (num_nodes < 1 || num_elems < 1 || num_ifaces < 0) throw ArgError("dg_h5_write_mesh: bad sizes");
        const int nv = dim == 2 ? 4 : 8;
        for (long i = 0; i < num_elems * nv; ++i)
            if (elem_nodes[i] < 0 || elem_nodes[i] >= num_nodes)
                throw ArgError("dg_h5_write_mesh: element node index out of range");
        quiet_hdf5();
        std::vector<long> ifc((size_t)num_ifaces * 6);
        for (size_t i = 0; i < ifc.size(); ++i) ifc[i] = iface_data[i];
        Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "create " + errstr_path(path));
        write_scalar(file, "Dimension", dim);
        write_scalar(file, "QOrder", 1);
        write_scalar(file, "nElem", num_elems);
        write_scalar(file, "nNode", num_nodes);
        write_scalar(file, "nIFace", num_ifaces);
        write_scalar(file, "nNodePerElem", nv);
        write_scalar(file, "nNodePerFace", nv / 2);
        write_scalar(file, "nBFaceGroup", 0);
        write_2d(file, "NodeCoords", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, num_nodes, dim, node_coords);
        write_2d(file, "Elem2Nodes", H5T_STD_I64LE, H5T_NATIVE_LONG, num_elems, nv, elem_nodes);
        write_2d(file, "IFaceData", H5T_STD_I64LE, H5T_NATIVE_LONG, num_ifaces, 6, ifc.data());
    });
}

int dg_create_from_mesh_file(const char* toml_text, const char* mesh_file, dg_solver** out) {
    if (out) *out = nullptr;
    std::string file;
    dg_h5_mesh_info m{};
    std::vector<double> xyz;
    std::vector<long> e2n;
    std::vector<int> ifc;
    int rc = guarded([&] {
        if (!toml_text || !out) throw ArgError("dg_create_from_mesh_file: NULL argument");
        file = mesh_file ? std::string(mesh_file) : dg::case_mesh_file(toml_text);
        if (file.empty()) return;
        quiet_hdf5();
        Hid f(H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose, "open mesh file " + file);
        m = mesh_info(f);
        // the file's dimension is the case's (Mesh<dim>, mesh.cpp:28-60)
        const int dim = dg::parse_case(toml_text, true).dim, nv = dim == 2 ? 4 : 8;
        if (m.dim != dim)
            throw dg::InputError(file + ": Dimension = " + std::to_string(m.dim) +
                                 " does not match Physics.dim = " + std::to_string(dim));
        if (m.qorder != 1 || m.num_nodes_per_elem != nv)
            throw dg::UnsupportedError(file + ": QOrder = " + std::to_string(m.qorder) + " with " +
                                       std::to_string(m.num_nodes_per_elem) + " nodes per element; this path takes " +
                                       (dim == 2 ? "bilinear (QOrder 1, 4-node) quadrilaterals"
                                                 : "trilinear (QOrder 1, 8-node) hexahedra"));
        if (m.num_bface_groups > 0)
            throw dg::UnsupportedError(file + ": boundary face groups are not on this path (periodic meshes only)");
        xyz.resize((size_t)m.num_nodes * dim);
        e2n.resize((size_t)m.num_elems * nv);
        ifc.resize((size_t)m.num_ifaces * 6);
        read_array(f, "NodeCoords", H5T_NATIVE_DOUBLE, m.num_nodes * dim, xyz.data());
        read_array(f, "Elem2Nodes", H5T_NATIVE_LONG, m.num_elems * nv, e2n.data());
        read_array(f, "IFaceData", H5T_NATIVE_INT, m.num_ifaces * 6, ifc.data());
    });
    if (rc != DG_OK) return rc;
    rc = file.empty() ? dg_create(toml_text, out)
                      : dg_create_mesh(toml_text, m.num_nodes, xyz.data(), m.num_elems, e2n.data(), m.num_ifaces,
                                       ifc.data(), out);
    if (rc != DG_OK) g_err = dg_last_error(nullptr);
    return rc;
}

int dg_write_solution(const dg_solver* s, const char* path) {
    dg_info info;
    int rc = dg_get_info(s, &info);
    if (rc != DG_OK) {
        g_err = dg_last_error(s);
        return rc;
    }
    std::vector<double> U;
    try {
        U.resize((size_t)info.num_elems * info.num_basis * info.num_state_vars);
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return DG_ERR_DEVICE;
    }
    rc = dg_get_state(s, U.data(), DG_LAYOUT_RIGHT);
    long nn = 0;
    if (rc == DG_OK) rc = dg_get_mesh(s, &nn, nullptr, nullptr);
    if (rc != DG_OK) {
        g_err = dg_last_error(s);
        return rc;
    }
    return guarded([&] {
        if (!path) throw ArgError("dg_write_solution: NULL path");
        std::vector<double> xyz((size_t)nn * info.dim);
        std::vector<long> e2n((size_t)info.num_elems * (info.dim == 2 ? 4 : 8));
        if (dg_get_mesh(s, &nn, xyz.data(), e2n.data()) != DG_OK) throw std::runtime_error(dg_last_error(s));
        write_solution_file(path, info.dim, nn, xyz.data(), info.num_elems, e2n.data(), info.num_basis,
                            info.num_state_vars, info.time, U.data(), DG_LAYOUT_RIGHT);
    });
}

int dg_read_solution(dg_solver* s, const char* path, int set_time) {
    dg_info info;
    int rc = dg_get_info(s, &info);
    if (rc != DG_OK) {
        g_err = dg_last_error(s);
        return rc;
    }
    dg_h5_header h;
    rc = dg_h5_read_header(path, &h);
    if (rc != DG_OK) return rc;
    if (h.num_elems != info.num_elems || h.num_basis != info.num_basis || h.num_state_vars != info.num_state_vars) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "%s holds E=%ld nb=%d ns=%d; the case has E=%ld nb=%d ns=%d", path, h.num_elems,
                      h.num_basis, h.num_state_vars, info.num_elems, info.num_basis, info.num_state_vars);
        g_err = msg;
        return DG_ERR_INPUT;
    }
    std::vector<double> U;
    try {
        U.resize((size_t)info.num_elems * info.num_basis * info.num_state_vars);
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return DG_ERR_DEVICE;
    }
    rc = dg_h5_read_coefficients(path, U.data(), (long)U.size(), DG_LAYOUT_RIGHT);
    if (rc != DG_OK) return rc;
    rc = dg_set_state(s, U.data(), DG_LAYOUT_RIGHT);
    if (rc == DG_OK && set_time) rc = dg_set_time(s, h.time);
    if (rc != DG_OK) g_err = dg_last_error(s);
    return rc;
}

}  // extern "C"
