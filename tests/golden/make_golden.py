"""Transcribe the reference's own golden data into JSON fixtures.

Reads the reference source tree as text (nothing is compiled, imported or
run) and writes numeric data only:

  gl_tables.json          Gauss-Legendre nodes and weights, one entry per
                          point count, from src/numerics/nodes.cpp:97-500 and
                          src/numerics/quadrature/segment.cpp:7-411 (the
                          32-digit literals are kept as strings).
  legendre_p4.json        expected values/gradients of the LegendreQuad and
                          LegendreHex p=4 bases at two points, from
                          test/numerics/basis/test_basis.cpp:363-686.
  vortex3d_regression.json
                          L2 errors of the 3D isentropic vortex case, from
                          examples/euler/3D/isentropic_vortex/regression_data.txt
                          with its settings from input.toml in that folder.
  vortex2d_convergence.json
                          2D vortex L2 density errors per p and mesh from
                          examples/isentropic_vortex/convergence.py:22-86.

Usage:  python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _cases_to_tables(text, var, defects=None):
    """Parse `case a: case b: var(i) = (literal); ... break;` blocks.

    Blocks with an index that is never assigned are reported in `defects`
    (the reference's 19-point node table assigns pts(16) twice and never
    pts(15), src/numerics/nodes.cpp case 36/37)."""
    out = {}
    blocks = re.split(r"\bbreak\s*;", text)
    for blk in blocks:
        cases = re.findall(r"case\s+(\d+)\s*:", blk)
        vals = re.findall(var + r"\((\d+)\)\s*=\s*\(?\s*([-+0-9.eE]+)\s*\)?\s*;", blk)
        if not cases or not vals:
            continue
        n = max(int(i) for i, _ in vals) + 1
        arr = [None] * n
        for i, v in vals:
            arr[int(i)] = v
        if any(a is None for a in arr):
            if defects is not None:
                missing = [i for i, a in enumerate(arr) if a is None]
                dup = sorted({i for i, _ in vals if [j for j, _ in vals].count(i) > 1}, key=int)
                for c in cases:
                    defects[str(n)] = {"qorder_case": int(c), "unassigned": missing,
                                       "assigned_twice": [int(d) for d in dup]}
            continue
        for c in cases:
            out[int(c)] = arr
    return out


def gl_tables(ref):
    nodes_src = open(os.path.join(ref, "src/numerics/nodes.cpp")).read()
    # only the live get_gauss_legendre_segment_nodes body (commented copies follow)
    start = nodes_src.index("void get_gauss_legendre_segment_nodes")
    body = nodes_src[start:]
    # stop at the first commented-out duplicate block
    cut = body.find("//     case 0")
    if cut > 0:
        body = body[:cut]
    defects = {}
    pts = _cases_to_tables(body, "pts", defects)
    seg_src = open(os.path.join(ref, "src/numerics/quadrature/segment.cpp")).read()
    start = seg_src.index("void get_segment_weights_gl")
    wts = _cases_to_tables(seg_src[start:], "wts")
    table = {}
    for qorder in sorted(pts):
        if qorder in wts:
            n = len(pts[qorder])
            assert len(wts[qorder]) == n, qorder
            table[str(n)] = {"qorders": sorted(set([qorder]) | set(table.get(str(n), {}).get("qorders", []))),
                             "pts": pts[qorder], "wts": wts[qorder]}
    table["defects"] = defects
    return table


def _vector(text, name):
    m = re.search(r"std::vector<rtype>\s+" + name + r"\s*\{(.*?)\};", text, re.S)
    assert m, name
    return [float(x) for x in re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", m.group(1))]


def legendre_p4(ref):
    src = open(os.path.join(ref, "test/numerics/basis/test_basis.cpp")).read()
    return {
        "source": "test/numerics/basis/test_basis.cpp:363-686",
        "tolerance": 1e-5,
        "order": 4,
        "points_2d": [[0.3, -0.2], [0.7, -0.3]],
        "values_2d": _vector(src, "exact_vals_2d"),
        "grads_2d": _vector(src, "exact_grads_2d"),
        "points_3d": [[0.3, -0.2, 0.5], [0.7, -0.3, 0.1]],
        "values_3d": _vector(src, "exact_vals_3d"),
        "grads_3d": _vector(src, "exact_grads_3d"),
        "layout": "values[iq*nb + ib]; grads[idim*2*nb + iq*nb + ib]",
    }


def vortex3d(ref):
    d = os.path.join(ref, "examples/euler/3D/isentropic_vortex")
    txt = open(os.path.join(d, "regression_data.txt")).read()
    errs = {}
    for i, v in re.findall(r"norm_total_error\((\d)\)=([-+0-9.eE]+)", txt):
        errs[int(i)] = v
    return {
        "source": "examples/euler/3D/isentropic_vortex/regression_data.txt:11-17",
        "settings_source": "examples/euler/3D/isentropic_vortex/input.toml:1-23",
        "mesh": "cube_4.h5 (periodic 4x4x4 hexahedra, unit cube)",
        "order": 3, "basis": "LagrangeHex", "flux": "HLLC", "stepper": "RK4",
        "timestep": 1e-3, "end": 1.0, "gamma": 1.4, "GasConstant": 1.0,
        "ic": "IsentropicVortex",
        "l2_errors": [errs[i] for i in range(5)],
    }


def vortex2d(ref):
    src = open(os.path.join(ref, "examples/isentropic_vortex/convergence.py")).read()
    out = {"source": "examples/isentropic_vortex/convergence.py:22-86", "series": {}}
    for name, vals in re.findall(r"(\w+)\s*=\s*np\.array\(\[(.*?)\]\)", src, re.S):
        nums = re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", vals)
        out["series"][name] = [float(x) for x in nums]
    # the active series of the study (the last assignment of each name, in
    # uncommented lines), as written, without the entries marked "HOLD FOR
    # NOW" (repeats of the entry before them)
    out["settings_source"] = ("examples/isentropic_vortex/input.toml:1-23 (LagrangeQuad, HLLC, RK4, dt 0.001, "
                              "end 1.0, IsentropicVortex [1,1,1,1,5], GasConstant 1)")
    out["mesh"] = ("periodic N x N quadrilaterals on [-5,5]^2 (the reference's box_NxN.h5 mesh package is not "
                   "shipped; the domain is inferred from these values, which it reproduces to ~1e-7)")
    live = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
    for name, vals in re.findall(r"(\w+)\s*=\s*\(?np\.array\(\[(.*?)\]\)", live, re.S):
        if name == "Nelem":
            out["nelem_1d"] = [int(float(x)) for x in re.findall(r"[0-9]+\.?[0-9]*", vals)]
        elif re.fullmatch(r"err_p\d", name):
            keep = [l for l in vals.splitlines() if "HOLD" not in l]
            out[name] = re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", "\n".join(keep))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    files = {
        "gl_tables.json": gl_tables(ref),
        "legendre_p4.json": legendre_p4(ref),
        "vortex3d_regression.json": vortex3d(ref),
        "vortex2d_convergence.json": vortex2d(ref),
    }
    for name, data in files.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
