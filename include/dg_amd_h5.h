/* Solution files in the reference's data.h5 layout and hex mesh files in
 * the layout of the reference's mesh reader, for the DG path of libdg_amd.
 *
 * Implemented by libdg_amd_h5.so (host code over the HDF5 C library; no GPU
 * code).  The file layout is the one Writer::write produces
 * (src/io/writer.cpp:10-65) with one writing rank, using the "parallel
 * dataset" convention of the reference's HDF5 wrapper
 * (src/io/HDF5Wrapper.h:780-877): every array is a flat 1D dataset `name`
 * plus `name_offsets` (per-rank offsets; attributes nDims, nRanks,
 * totalSize) and `name_dims` (per-rank dimensions).  Files written here are
 * read by the reference's Solver::read_in_coefficients
 * (src/solver/base.cpp:233-279) and the reverse.
 *
 * Status codes are those of dg_amd.h; dg_h5_last_error() gives the message
 * of the last failed call on this thread.
 */
#ifndef DG_AMD_H5_H
#define DG_AMD_H5_H

#include "dg_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Header attributes of a data.h5 file (src/io/writer.cpp:42-54). */
typedef struct dg_h5_header {
    long num_elems;            /* "Number of Elements" */
    long num_nodes;            /* "Number of Nodes" */
    int num_ranks;             /* "Number of Ranks" */
    int dim;                   /* "Number of Dimensions" */
    int num_nodes_per_elem;    /* "Number of Nodes Per Element" */
    int num_basis;             /* "Number of Basis Functions" */
    int num_state_vars;        /* "Number of State Variables" */
    int stored_layout;         /* "Stored Layout": 1 row major (DG_LAYOUT_RIGHT), 0 column major */
    double time;               /* "Solver Final Time" */
} dg_h5_header;

/* Write a data.h5 file for a structured periodic box of nelem[0] x nelem[1]
 * x nelem[2] hexahedra on [lower, upper]: the header attributes, the mesh
 * datasets (node coordinates of the (n+1)^3 grid, trilinear element nodes
 * in the reference's hex node order 4k+2j+i, identity local-to-global
 * maps) and "Solution Coefficients" [E][nb][ns] in the given layout
 * (Writer::write, src/io/writer.cpp:10-65; Writer::Writer,
 * src/io/writer.cpp:68-98).  Uc holds E*nb*ns doubles in `layout`.
 * Element e = ix + nx*(iy + ny*iz). */
int dg_h5_write_box(const char* path, const int nelem[3], const double lower[3],
                    const double upper[3], int num_basis, int num_state_vars, double time,
                    const double* Uc, int layout);

/* Read the header attributes (src/solver/base.cpp:242-251). */
int dg_h5_read_header(const char* path, dg_h5_header* header);

/* Read "Solution Coefficients" into Uc (num_elems*nb*ns doubles), converted
 * from the stored layout to `layout` (src/solver/base.cpp:253-275).
 * `capacity` is the number of doubles Uc holds. */
int dg_h5_read_coefficients(const char* path, double* Uc, long capacity, int layout);

/* Sizes of a mesh file (Mesh::read_mesh, src/mesh/mesh.cpp:79-107). */
typedef struct dg_h5_mesh_info {
    long dim;                  /* "Dimension" */
    long qorder;               /* "QOrder" (geometric order) */
    long num_elems;            /* "nElem" */
    long num_nodes;            /* "nNode" */
    long num_ifaces;           /* "nIFace" */
    long num_nodes_per_elem;   /* "nNodePerElem" */
    long num_nodes_per_face;   /* "nNodePerFace" (1 when absent, as the reference takes it) */
    long num_bface_groups;     /* "nBFaceGroup" (0 when absent) */
} dg_h5_mesh_info;

int dg_h5_read_mesh_info(const char* path, dg_h5_mesh_info* info);

/* NodeCoords [num_nodes][dim], Elem2Nodes [num_elems][num_nodes_per_elem]
 * and IFaceData [num_ifaces][6] of a mesh file (src/mesh/mesh.cpp:120-160),
 * any dimension; the capacities are the lengths of the three buffers. */
int dg_h5_read_mesh(const char* path, double* node_coords, long coords_capacity, long* elem_nodes,
                    long elem_nodes_capacity, int* iface_data, long iface_capacity);

/* Write a mesh file of QOrder 1 elements without boundary face groups:
 * dim 3, trilinear hexahedra (8 nodes per element, 4 per face); dim 2,
 * bilinear quadrilaterals (4 per element, 2 per face).  The datasets the
 * reference's mesh reader opens, with the arrays of dg_create_mesh. */
int dg_h5_write_mesh(const char* path, int dim, long num_nodes, const double* node_coords, long num_elems,
                     const long* elem_nodes, long num_ifaces, const int* iface_data);

/* dg_create_mesh with the arrays of a mesh file (the Mesh constructor,
 * src/mesh/mesh.cpp:28-60, and Mesh::read_mesh, 79-200).  mesh_file NULL:
 * the case's Mesh.file, a path relative to the working directory as in the
 * reference; a case without Mesh.file (a PeriodicBox case) goes to
 * dg_create.  The file must hold a 3D, QOrder 1 hex mesh without boundary
 * face groups.  On failure dg_h5_last_error() holds the message. */
int dg_create_from_mesh_file(const char* toml_text, const char* mesh_file, dg_solver** out);

/* Write the solver's state and time to `path` (src/exec/main.cpp:106-112),
 * with the solver's mesh (dg_get_mesh). */
int dg_write_solution(const dg_solver* s, const char* path);

/* Load the coefficients of `path` into the solver's state
 * (Solver::read_in_coefficients, src/solver/base.cpp:233-279).  The file
 * must have the solver's E, nb and ns.  With set_time != 0 the solver time
 * is set to the file's "Solver Final Time"; the reference's post reads that
 * attribute but leaves its time at 0 (set_time = 0 reproduces that). */
int dg_read_solution(dg_solver* s, const char* path, int set_time);

/* Message of the last failed dg_h5_* / dg_*_solution call on this thread. */
const char* dg_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DG_AMD_H5_H */
