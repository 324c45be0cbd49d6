// TOML-subset reader and case-key interpretation.
#include "dg_config.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <sstream>

namespace dg {

namespace {

struct Cursor {
    const std::string& s;
    size_t i = 0;
    int line = 1;
    explicit Cursor(const std::string& t) : s(t) {}
    bool eof() const { return i >= s.size(); }
    char peek() const { return eof() ? '\0' : s[i]; }
    char get() {
        char c = s[i++];
        if (c == '\n') ++line;
        return c;
    }
    [[noreturn]] void fail(const std::string& msg) const {
        std::ostringstream o;
        o << "TOML parse error at line " << line << ": " << msg;
        throw InputError(o.str());
    }
    void skip_ws_inline() {
        while (!eof() && (peek() == ' ' || peek() == '\t' || peek() == '\r')) get();
    }
    void skip_comment() {
        if (peek() == '#')
            while (!eof() && peek() != '\n') get();
    }
    // whitespace, newlines and comments (inside arrays)
    void skip_all() {
        for (;;) {
            skip_ws_inline();
            if (peek() == '#') { skip_comment(); continue; }
            if (peek() == '\n') { get(); continue; }
            break;
        }
    }
};

std::string parse_key(Cursor& c) {
    c.skip_ws_inline();
    std::string k;
    if (c.peek() == '"') {
        c.get();
        while (!c.eof() && c.peek() != '"') k += c.get();
        if (c.eof()) c.fail("unterminated quoted key");
        c.get();
        return k;
    }
    while (!c.eof() && (std::isalnum((unsigned char)c.peek()) || c.peek() == '_' || c.peek() == '-'))
        k += c.get();
    if (k.empty()) c.fail("expected a key");
    return k;
}

TomlValue parse_value(Cursor& c) {
    c.skip_ws_inline();
    TomlValue v;
    char ch = c.peek();
    if (ch == '"' || ch == '\'') {
        char q = c.get();
        v.kind = TomlValue::STRING;
        while (!c.eof() && c.peek() != q) {
            char x = c.get();
            if (x == '\n') c.fail("newline in string");
            if (x == '\\' && q == '"' && !c.eof()) {
                char e = c.get();
                switch (e) {
                case 'n': x = '\n'; break;
                case 't': x = '\t'; break;
                default: x = e;
                }
            }
            v.str += x;
        }
        if (c.eof()) c.fail("unterminated string");
        c.get();
        return v;
    }
    if (ch == '[') {
        c.get();
        v.kind = TomlValue::ARRAY;
        c.skip_all();
        if (c.peek() == ']') { c.get(); return v; }
        for (;;) {
            c.skip_all();
            v.arr.push_back(parse_value(c));
            c.skip_all();
            if (c.peek() == ',') { c.get(); c.skip_all(); if (c.peek() == ']') { c.get(); break; } continue; }
            if (c.peek() == ']') { c.get(); break; }
            c.fail("expected ',' or ']' in array");
        }
        return v;
    }
    std::string tok;
    while (!c.eof() && c.peek() != ',' && c.peek() != ']' && c.peek() != '\n' && c.peek() != '#' &&
           c.peek() != ' ' && c.peek() != '\t' && c.peek() != '\r')
        tok += c.get();
    if (tok == "true" || tok == "false") {
        v.kind = TomlValue::BOOL;
        v.b = tok == "true";
        return v;
    }
    std::string t2;
    for (char x : tok) if (x != '_') t2 += x;
    if (t2.empty()) c.fail("expected a value");
    char* end = nullptr;
    v.kind = TomlValue::NUMBER;
    v.num = std::strtod(t2.c_str(), &end);
    if (!end || *end != '\0') c.fail("bad value '" + tok + "'");
    v.is_int = t2.find_first_of(".eE") == std::string::npos && t2 != "inf" && t2 != "nan";
    return v;
}

}  // namespace

TomlDoc parse_toml(const std::string& text) {
    TomlDoc doc;
    Cursor c(text);
    std::string table;
    doc[table];
    while (!c.eof()) {
        c.skip_ws_inline();
        if (c.peek() == '#') { c.skip_comment(); continue; }
        if (c.peek() == '\n') { c.get(); continue; }
        if (c.eof()) break;
        if (c.peek() == '[') {
            c.get();
            if (c.peek() == '[') c.fail("arrays of tables are not supported");
            table = parse_key(c);
            while (c.peek() == '.') { c.get(); table += "." + parse_key(c); }
            c.skip_ws_inline();
            if (c.peek() != ']') c.fail("expected ']' after table name");
            c.get();
            doc[table];
        } else {
            std::string key = parse_key(c);
            c.skip_ws_inline();
            if (c.peek() != '=') c.fail("expected '=' after key '" + key + "'");
            c.get();
            TomlValue v = parse_value(c);
            if (doc[table].count(key)) c.fail("duplicate key '" + key + "'");
            doc[table][key] = v;
        }
        c.skip_ws_inline();
        c.skip_comment();
        if (!c.eof() && c.peek() != '\n') c.fail("trailing characters after value");
    }
    return doc;
}

namespace {

const TomlValue* find(const TomlDoc& d, const std::string& table, const std::string& key) {
    auto t = d.find(table);
    if (t == d.end()) return nullptr;
    auto k = t->second.find(key);
    return k == t->second.end() ? nullptr : &k->second;
}

[[noreturn]] void missing(const std::string& table, const std::string& key) {
    throw InputError("missing required key " + table + "." + key);
}

double get_num(const TomlDoc& d, const std::string& t, const std::string& k, bool required, double dflt) {
    const TomlValue* v = find(d, t, k);
    if (!v) { if (required) missing(t, k); return dflt; }
    if (v->kind != TomlValue::NUMBER) throw InputError(t + "." + k + " must be a number");
    return v->num;
}

int get_int(const TomlDoc& d, const std::string& t, const std::string& k, bool required, int dflt) {
    const TomlValue* v = find(d, t, k);
    if (!v) { if (required) missing(t, k); return dflt; }
    if (v->kind != TomlValue::NUMBER || !v->is_int) throw InputError(t + "." + k + " must be an integer");
    return (int)v->num;
}

std::string get_str(const TomlDoc& d, const std::string& t, const std::string& k, bool required,
                    const std::string& dflt) {
    const TomlValue* v = find(d, t, k);
    if (!v) { if (required) missing(t, k); return dflt; }
    if (v->kind != TomlValue::STRING) throw InputError(t + "." + k + " must be a string");
    return v->str;
}

bool get_bool(const TomlDoc& d, const std::string& t, const std::string& k, bool dflt) {
    const TomlValue* v = find(d, t, k);
    if (!v) return dflt;
    if (v->kind != TomlValue::BOOL) throw InputError(t + "." + k + " must be true or false");
    return v->b;
}

std::vector<double> get_num_array(const TomlDoc& d, const std::string& t, const std::string& k,
                                  bool required) {
    const TomlValue* v = find(d, t, k);
    if (!v) { if (required) missing(t, k); return {}; }
    if (v->kind != TomlValue::ARRAY) throw InputError(t + "." + k + " must be an array");
    std::vector<double> out;
    for (const auto& e : v->arr) {
        if (e.kind != TomlValue::NUMBER) throw InputError(t + "." + k + " must hold numbers");
        out.push_back(e.num);
    }
    return out;
}

// enum_from_string (src/common/enums.h:33-54): same message shape.
[[noreturn]] void bad_choice(const std::string& str, const std::vector<std::string>& valid) {
    std::ostringstream m;
    m << "Oh no! " << str << " is not a valid choice! Instead, try one of these:";
    for (auto& v : valid) m << " " << v;
    throw InputError(m.str());
}

int pick(const std::string& str, const std::vector<std::string>& valid) {
    for (size_t i = 0; i < valid.size(); ++i)
        if (valid[i] == str) return (int)i;
    bad_choice(str, valid);
}

}  // namespace

std::string case_mesh_file(const std::string& toml_text) {
    TomlDoc d = parse_toml(toml_text);
    const TomlValue* v = find(d, "Mesh", "file");
    if (!v) return "";
    if (v->kind != TomlValue::STRING) throw InputError("Mesh.file must be a string");
    return v->str;
}

CaseConfig parse_case(const std::string& toml_text, bool mesh_input) {
    TomlDoc d = parse_toml(toml_text);
    CaseConfig c;
    c.mesh_input = mesh_input;

    // [Physics]
    std::string phys = get_str(d, "Physics", "name", true, "");
    int pt = pick(phys, {"Euler", "NavierStokes"});          // enums.h:56-66
    if (pt != 0) throw UnsupportedError("Physics.name = NavierStokes is not implemented by this path");
    int dim = get_int(d, "Physics", "dim", true, 3);
    if (dim != 2 && dim != 3) throw InputError("Physics.dim must be 2 or 3");
    c.dim = dim;
    c.gamma = get_num(d, "Physics", "gamma", false, 1.4);
    c.gas_constant = get_num(d, "Physics", "GasConstant", false, 287.05);
    std::string nf = get_str(d, "Physics", "convective_flux_fcn", false, "LaxFriedrichs");
    c.flux = (Flux)pick(nf, {"LaxFriedrichs", "HLLC", "Roe"});  // enums.h:81-91 + Roe (new)
    if (!(c.gamma > 1.0)) throw InputError("Physics.gamma must be > 1");

    // [Numerics]
    c.order = get_int(d, "Numerics", "order", true, 0);
    std::string bname = get_str(d, "Numerics", "basis", true, "");
    static const std::vector<std::string> bases = {
        "LagrangeSeg", "LagrangeQuad", "LagrangeHex", "LegendreSeg", "LegendreQuad", "LegendreHex",
        "LagrangeGLL1D", "LagrangeGLL2D", "LagrangeGLL3D", "LagrangeGL2D", "LagrangeGL3D",
        "TriangleLagrange"};                                     // enums.h:109-139
    int bi = pick(bname, bases);
    // the tensor bases share their 1D functions; the shape follows Physics.dim
    if (bi == 2 || bi == 1) c.basis = Basis::LagrangeHex;
    else if (bi == 5 || bi == 4) c.basis = Basis::LegendreHex;
    else throw UnsupportedError("Numerics.basis = " + bname + " is not implemented by this path "
                                "(LagrangeHex, LegendreHex; LagrangeQuad, LegendreQuad with dim = 2)");
    if ((bi == 1 || bi == 4) != (dim == 2))
        throw InputError("Numerics.basis = " + bname + " does not match Physics.dim = " + std::to_string(dim));
    std::string quad = get_str(d, "Numerics", "quad", false, "GaussLegendre");
    int qi = pick(quad, {"GaussLegendre", "GaussLobatto", "TriangleDunavant", "TriangleGaussLegendre"});
    if (qi != 0) throw UnsupportedError("Numerics.quad = " + quad + " is not implemented");
    if (c.order < 1) throw UnsupportedError("Numerics.order must be >= 1 on this path");
    // FP32: the FP32 form; BF16 / BF16x3: the FP32 form with the x-y part of
    // the weak-gradient projection on the bf16 matrix cores (FP32
    // accumulation; one bf16 term, or the value and its rounding error as
    // two bf16 terms and three products)
    const int prec = pick(get_str(d, "Numerics", "precision", false, "FP64"), {"FP64", "FP32", "BF16", "BF16x3"});
    c.fp32 = prec >= 1;
    c.proj = prec == 2 ? 1 : (prec == 3 ? 3 : 0);
    if (c.fp32 && dim != 3) throw UnsupportedError("Numerics.precision = FP32 with Physics.dim = 2 (3D box only)");
    c.subtract_mean = get_bool(d, "Numerics", "subtract_mean", c.fp32);
    if (c.subtract_mean && !c.fp32)
        throw UnsupportedError("Numerics.subtract_mean applies to Numerics.precision = FP32, BF16 or BF16x3");
    if (c.basis == Basis::LegendreHex && c.order > 7)
        throw InputError("Legendre polynomial P = " + std::to_string(c.order) + " not implemented.");

    // [Stepper]
    std::string st = get_str(d, "Stepper", "type", false, "FE");
    c.stepper = (Stepper)pick(st, {"FE", "RK4"});            // enums.h:68-78
    c.end_time = get_num(d, "Stepper", "end", true, 0.0);
    c.timestep = get_num(d, "Stepper", "timestep", false, 0.0);
    c.cfl = get_num(d, "Stepper", "cfl", false, 0.0);
    // compensated = true (new key; the default of the reduced-precision
    // forms): the solution update U0 + increment is carried as a pair of
    // FP64 registers (the sum and its rounding error, TwoSum)
    c.compensated = get_bool(d, "Stepper", "compensated", c.fp32);
    if (c.compensated && dim != 3) throw UnsupportedError("Stepper.compensated with Physics.dim = 2 (3D only)");
    if (c.timestep <= 0.0 && c.cfl <= 0.0)
        throw InputError("missing required key Stepper.timestep (or Stepper.cfl)");

    // [InitialCondition]
    std::string icn = get_str(d, "InitialCondition", "name", true, "");
    // enums.h:93-107 plus three new names
    {
        int idx = pick(icn, {"Uniform", "Gaussian", "Sphere", "IsentropicVortex", "ExtrudedIsentropicVortex",
                             "GaussianBump", "TaylorGreen", "VortexPulseMix"});
        static const IC map[] = {IC::Uniform, IC::Gaussian, IC::Sphere, IC::IsentropicVortex,
                                 IC::ExtrudedIsentropicVortex, IC::GaussianBump, IC::TaylorGreen,
                                 IC::VortexPulseMix};
        c.ic = map[idx];
        // Physics<2>::call_IC and Physics<3>::call_IC (physics/base/base.cpp:77-103)
        if (dim == 2 && !(c.ic == IC::Uniform || c.ic == IC::Gaussian || c.ic == IC::IsentropicVortex))
            throw UnsupportedError("InitialCondition.name = " + icn + " is a 3D initial condition");
        if (dim == 3 && c.ic == IC::Gaussian)
            throw UnsupportedError("InitialCondition.name = Gaussian is a 2D initial condition");
    }
    c.ic_data = get_num_array(d, "InitialCondition", "data", false);
    if (c.ic_data.size() > 24) throw InputError("InitialCondition.data holds at most 24 values");
    if (c.ic != IC::VortexPulseMix && c.ic_data.size() > 10)
        throw InputError("InitialCondition.data holds at most 10 values");
    size_t need = 0;
    switch (c.ic) {
    case IC::Uniform: need = dim == 2 ? 4 : 5; break;
    case IC::Gaussian: need = 4; break;
    case IC::IsentropicVortex: need = dim == 2 ? 5 : 0; break;
    case IC::ExtrudedIsentropicVortex: need = 5; break;
    case IC::GaussianBump: need = 9; break;
    case IC::TaylorGreen: need = 1; break;
    case IC::VortexPulseMix: need = 19; break;
    default: need = 0;
    }
    if (c.ic_data.size() < need)
        throw InputError("InitialCondition.data needs " + std::to_string(need) + " values for " + icn);
    if (((dim == 2 && c.ic == IC::IsentropicVortex) || c.ic == IC::ExtrudedIsentropicVortex ||
         c.ic == IC::VortexPulseMix) && c.gas_constant != 1.0)
        throw InputError("ExtrudedIsentropicVortex requires Physics.GasConstant = 1 "
                         "(src/physics/euler/functions.cpp:107)");

    // [Mesh]
    if (mesh_input) {
        // the mesh comes with the call (dg_create_mesh); lower / upper, if
        // given, are the periodic cell of the advected exact solutions
        std::vector<double> lo = get_num_array(d, "Mesh", "lower", false);
        std::vector<double> hi = get_num_array(d, "Mesh", "upper", false);
        if (lo.size() == (size_t)dim && hi.size() == (size_t)dim)
            for (int i = 0; i < dim; ++i) {
                c.lower[i] = lo[i];
                c.upper[i] = hi[i];
            }
        const std::string nrm = get_str(d, "Mesh", "normals", false, "reference");
        if (nrm == "exact") c.exact_normals = true;
        else if (nrm != "reference") throw InputError("Mesh.normals must be \"reference\" or \"exact\"");
        if (get_num(d, "Mesh", "warp", false, 0.0) != 0.0)
            throw InputError("Mesh.warp applies to the periodic box only");
        if (c.fp32) throw UnsupportedError("Numerics.precision = FP32 on a mesh (the periodic box only)");
        if (c.timestep <= 0.0) throw InputError("Stepper.timestep is required with a mesh (Stepper.cfl needs the box)");
        if (dim == 3 && c.order > 4)
            throw UnsupportedError("meshes with order > 4: the per-element mass matrix inverse of the "
                                   "non-affine path is formed in LDS and holds at most (4+1)^3 basis functions");
        return c;
    }
    std::string mtype = get_str(d, "Mesh", "type", false, "");
    if (mtype.empty()) {
        if (find(d, "Mesh", "file"))
            throw UnsupportedError("Mesh.file: HDF5 mesh files are read by dg_create_from_mesh_file "
                                   "(libdg_amd_h5); dg_create takes Mesh.type = \"PeriodicBox\"");
        missing("Mesh", "type");
    }
    if (mtype != "PeriodicBox") throw InputError("Mesh.type must be \"PeriodicBox\"");
    // dim entries each (2D: the box is nelem[0] x nelem[1], nelem[2] = 1)
    const size_t nd = (size_t)dim;
    const std::string ndw = dim == 2 ? "2" : "3";
    std::vector<double> ne = get_num_array(d, "Mesh", "nelem", true);
    if (ne.size() != nd) throw InputError("Mesh.nelem must hold " + ndw + " integers");
    for (size_t i = 0; i < nd; ++i) {
        if (ne[i] < 1 || ne[i] != std::floor(ne[i])) throw InputError("Mesh.nelem entries must be positive integers");
        c.nelem[i] = (int)ne[i];
    }
    std::vector<double> lo = get_num_array(d, "Mesh", "lower", false);
    std::vector<double> hi = get_num_array(d, "Mesh", "upper", false);
    if (!lo.empty()) {
        if (lo.size() != nd) throw InputError("Mesh.lower must hold " + ndw + " numbers");
        for (size_t i = 0; i < nd; ++i) c.lower[i] = lo[i];
    }
    if (!hi.empty()) {
        if (hi.size() != nd) throw InputError("Mesh.upper must hold " + ndw + " numbers");
        for (size_t i = 0; i < nd; ++i) c.upper[i] = hi[i];
    }
    if (dim == 2) {
        c.nelem[2] = 1;
        c.lower[2] = 0.0;
        c.upper[2] = 1.0;
    }
    for (int i = 0; i < 3; ++i)
        if (!(c.upper[i] > c.lower[i])) throw InputError("Mesh.upper must exceed Mesh.lower");
    if ((double)c.nelem[0] * c.nelem[1] * c.nelem[2] > 2.0e9)
        throw InputError("Mesh.nelem: too many elements");
    c.warp = get_num(d, "Mesh", "warp", false, 0.0);
    if (!(c.warp >= 0.0 && c.warp <= 0.25))
        throw InputError("Mesh.warp must lie in [0, 0.25] (node displacement in element widths)");
    if (dim == 2 && c.warp > 0.0) throw UnsupportedError("Mesh.warp with Physics.dim = 2 (affine quadrilaterals only)");
    const std::string nrm = get_str(d, "Mesh", "normals", false, "reference");
    if (nrm == "exact") c.exact_normals = true;
    else if (nrm != "reference") throw InputError("Mesh.normals must be \"reference\" or \"exact\"");
    if (c.warp > 0.0 && c.order > 4)
        throw UnsupportedError("Mesh.warp with order > 4: the per-element mass matrix inverse of the "
                               "non-affine path is formed in LDS and holds at most (4+1)^3 basis functions");
    if (c.fp32 && c.warp > 0.0) throw UnsupportedError("Numerics.precision = FP32 with Mesh.warp (affine box only)");
    return c;
}

}  // namespace dg
