"""The device output path: ps.nc and t.nc written from a handle while the
steps go on, read back with scipy and compared with the fields and, in FP64,
with the CPU oracle's ps and t at the same steps."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.io import netcdf_file

import dynamical_model_amd as dm
import oracle_lib
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
    if precision == dm.MLC_FP64:
        # the file's content against the CPU oracle's ps and t after 0, 2 and
        # 4 steps (t bit for bit; ps through the device exp, to the last ulp)
        o = oracle_lib.oracle_from_host(hs, dm.case_params(case))
        done = 0
        for r, nsteps in enumerate((0, 2, 4)):
            while done < nsteps:
                done += 1
                o.moloch(done, False)
            np.testing.assert_array_equal(t[r], dm.interior(o.field("t")))
            np.testing.assert_allclose(ps[r], dm.interior(o.field("ps")), rtol=1e-14, atol=0)
