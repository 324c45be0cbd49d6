// Case configuration: a TOML-subset reader and the reference's input keys.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace dg {

// Input error, the counterpart of the reference's InputException
// (src/common/my_exceptions.h; thrown by enum_from_string,
// src/common/enums.h:33-54).
struct InputError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// Valid reference input that this path does not implement.
struct UnsupportedError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// One TOML value: string, number, bool or a flat array of those.
struct TomlValue {
    enum Kind { STRING, NUMBER, BOOL, ARRAY } kind = STRING;
    std::string str;
    double num = 0.0;
    bool is_int = false;
    bool b = false;
    std::vector<TomlValue> arr;
};

// table name -> key -> value ("" is the root table).
using TomlDoc = std::map<std::string, std::map<std::string, TomlValue>>;

// Parses tables, key = value pairs, strings, integers, floats, booleans,
// (multi-line) arrays and comments -- the subset the reference's input
// decks use (examples/*/input.toml).
TomlDoc parse_toml(const std::string& text);

enum class Basis { LagrangeHex = 0, LegendreHex = 1 };
enum class Flux { LaxFriedrichs = 0, HLLC = 1, Roe = 2 };
enum class Stepper { FE = 0, RK4 = 1 };
enum class IC { Uniform = 0, Sphere = 1, IsentropicVortex = 2, ExtrudedIsentropicVortex = 3,
                GaussianBump = 4, TaylorGreen = 5, Gaussian = 6, VortexPulseMix = 7 };

struct CaseConfig {
    // [Physics] dim: 3 (hexahedra, 5 state variables) or 2 (quadrilaterals,
    // 4 state variables; the box has nelem[2] = 1)
    int dim = 3;
    // [Numerics] (src/numerics/numerics_data.cpp:229-300)
    int order = 0;
    Basis basis = Basis::LagrangeHex;
    // precision = "FP64" (default) | "FP32" (a new key, no
    // reference counterpart): FP32 face and volume work in the stage
    // kernel, FP64 RK registers, M^-1 and reductions; 3D periodic box only
    bool fp32 = false;
    // "BF16" (1) / "BF16x3" (3): the FP32 form with the x-y projection on
    // bf16 matrix cores (one term / a hi-lo split with three products)
    int proj = 0;
    // subtract_mean = true (new key; the default of the reduced-precision
    // forms): the FP32 stage subtracts the domain-mean state from the
    // coefficients before rounding them and forms deviation fluxes
    bool subtract_mean = false;
    // [Mesh]: structured periodic box, type = "PeriodicBox", nelem, lower,
    // upper; warp > 0 displaces the nodes periodically (trilinear,
    // non-affine elements); normals = "reference" (the reference's
    // evaluation points) or "exact" (at the flux points)
    int nelem[3] = {0, 0, 0};
    double lower[3] = {0, 0, 0}, upper[3] = {1, 1, 1};
    double warp = 0.0;
    bool exact_normals = false;
    bool mesh_input = false;   // the mesh came with dg_create_mesh
    // [Stepper] (src/solver/base.cpp:46-57)
    Stepper stepper = Stepper::FE;
    double timestep = 0.0;     // 0 => derived from cfl
    double cfl = 0.0;          // new key, used when timestep is absent
    bool compensated = false;  // new key: compensated (TwoSum) solution update
    double end_time = 0.0;
    // [Physics] (src/solver/base.cpp:34-44; src/physics/base/base.cpp:24-32)
    double gamma = 1.4;
    double gas_constant = 287.05;
    Flux flux = Flux::LaxFriedrichs;
    // [InitialCondition] (src/exec/main.cpp:73-83)
    IC ic = IC::Uniform;
    std::vector<double> ic_data;
};

// mesh_input: the mesh is given with the call (no box keys required)
CaseConfig parse_case(const std::string& toml_text, bool mesh_input = false);

// Mesh.file of a case ("" when the case has none: a PeriodicBox case).
std::string case_mesh_file(const std::string& toml_text);

}  // namespace dg
