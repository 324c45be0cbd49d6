"""data.h5 solution files with the GPU solver: write (src/exec/main.cpp:
106-112), read back (Solver::read_in_coefficients, src/solver/base.cpp:
233-279) and restart.  Bit-exact: the file stores FP64 coefficients and the
stage kernels are deterministic."""
import os

import numpy as np
import pytest

import dg_amd
from dg_amd import hip

pytestmark = pytest.mark.gpu

h5 = pytest.importorskip("dg_amd.h5")
if not os.path.exists(h5.H5_LIB_PATH):
    pytest.skip("libdg_amd_h5.so not built (no HDF5 at build time)", allow_module_level=True)


def _case(ne=(3, 4, 2), order=2):
    return dg_amd.case_toml(order=order, basis="LegendreHex", flux="Roe", nelem=ne, lower=(-5, -5, -5),
                            upper=(5, 5, 5), stepper="RK4", timestep=0.02, end=1.0,
                            ic="ExtrudedIsentropicVortex", ic_data=(1, 1, 1, 1, 5))


@pytest.fixture(scope="module")
def gpu():
    assert hip.device_count() > 0, "GPU test needs a GPU"


def test_write_then_read_solution(gpu, tmp_path):
    f = str(tmp_path / "data.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case()) as b, dg_amd.Solver(_case()) as c:
        a.init_state_from_fcn()
        a.take_time_step(0.02, 3)
        a.write(f)
        hdr = h5.read_header(f)
        assert (hdr["num_elems"], hdr["num_basis"], hdr["num_state_vars"]) == a.shape
        assert hdr["time"] == a.time and hdr["stored_layout"] == 1
        assert np.array_equal(h5.read_coefficients(f), a.get_state())
        b.read_in_coefficients(f)
        assert np.array_equal(b.get_state(), a.get_state())
        assert b.time == a.time
        assert np.array_equal(b.get_error(), a.get_error())
        # the reference's post keeps its time at 0
        c.read_in_coefficients(f, set_time=False)
        assert c.time == 0.0
        assert np.array_equal(c.get_state(), a.get_state())


def test_restart_is_bit_exact(gpu, tmp_path):
    f = str(tmp_path / "restart.h5")
    with dg_amd.Solver(_case(order=3)) as full, dg_amd.Solver(_case(order=3)) as first, \
            dg_amd.Solver(_case(order=3)) as second:
        full.init_state_from_fcn()
        full.take_time_step(0.02, 6)
        first.init_state_from_fcn()
        first.take_time_step(0.02, 3)
        first.write(f)
        second.read_in_coefficients(f)
        second.take_time_step(0.02, 3)
        assert np.array_equal(second.get_state(), full.get_state())
        assert abs(second.time - full.time) < 1e-15


def test_read_solution_rejects_other_cases(gpu, tmp_path):
    f = str(tmp_path / "data.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case(ne=(3, 4, 3))) as b, \
            dg_amd.Solver(_case(order=3)) as c:
        a.init_state_from_fcn()
        a.write(f)
        for s in (b, c):
            with pytest.raises(dg_amd.DGError) as ei:
                s.read_in_coefficients(f)
            assert ei.value.status == 1 and "the case has" in str(ei.value)
        with pytest.raises(dg_amd.DGError):
            b.read_in_coefficients(str(tmp_path / "missing.h5"))


def test_file_written_in_column_major_reads_back(gpu, tmp_path):
    """A file with "Stored Layout" = 0 (the reference's LayoutLeft GPU
    builds, src/solver/base.cpp:478-489) loads into the same state."""
    f = str(tmp_path / "left.h5")
    with dg_amd.Solver(_case()) as a, dg_amd.Solver(_case()) as b:
        a.init_state_from_fcn()
        U = a.get_state()
        i = a.info
        h5.write_box(f, tuple(i.nelem), tuple(i.lower), tuple(i.upper),
                     np.ascontiguousarray(U.transpose(2, 1, 0)), time=0.0, layout=dg_amd.DG_LAYOUT_LEFT)
        b.read_in_coefficients(f)
        assert np.array_equal(b.get_state(), U)
