"""The device output path: ps.nc and t.nc written from a handle while the
steps go on, read back with scipy and compared with the fields."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.io import netcdf_file

import dynamical_model_amd as dm
from dynamical_model_amd.output import Output
from parity import small_config2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [dm.MLC_FP64, dm.MLC_FP32])
def test_output_records_match_fields(tmp_path, precision):
    case = small_config2(40, 36, 12)
    hs = dm.build_case(case)
    m = dm.Moloch()
    m.allocate_moloch(hs.jx, hs.iy, hs.kz, dm.case_params(case, precision=precision))
    m.load_state(hs)
    m.init_moloch()
    print(f"case 40x36x12 precision={precision}: output at steps 1, 3, 5", flush=True)
    out = Output(m, str(tmp_path))
    expect = []
    for n in range(1, 6):
        if n in (1, 3, 5):  # output(istep) runs before moloch(istep)
            out.write()
            m.sync()
            expect.append((dm.interior(m.get_field("ps")), dm.interior(m.get_field("t"))))
        m.moloch(n)
    out.close()
    with netcdf_file(str(tmp_path / "ps.nc"), "r", mmap=False) as nc:
        ps = nc.variables["ps"].data.copy()
    with netcdf_file(str(tmp_path / "t.nc"), "r", mmap=False) as nc:
        t = nc.variables["t"].data.copy()
    assert ps.shape == (3, 36, 40) and t.shape == (3, 12, 36, 40)
    for r, (eps, et) in enumerate(expect):
        np.testing.assert_array_equal(ps[r], eps)
        np.testing.assert_array_equal(t[r], et)
    # the records differ: the state moved between them
    assert not np.array_equal(t[0], t[2])
    m.close()
