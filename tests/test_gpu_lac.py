"""The learned acoustic corrector on the GPU.

The corrector replaces the first nsound-1 acoustic substeps of a phase by one
substep of (nsound-1)*dts plus a learned increment.  Checked here:

* its training pairs against the CPU oracle: the inputs are the increments
  of the long substep (and theta, qv) on the 3 x 3 stencil, the targets the
  difference between nsound-1 ordinary substeps and the long substep, both
  computed by the oracle's own
  acoustic loop (nompi/mod_moloch.F90:483-851) with nsound and dts set to
  the two schedules;
* that collecting pairs leaves the trajectory bit-identical to ordinary steps;
* that the trained increment beats the uncorrected long substep (the gain);
* that a failing check falls back to the ordinary loop, bit-identically.
"""
from __future__ import annotations

import numpy as np
import pytest

import dynamical_model_amd as dm
import oracle_lib
from dynamical_model_amd import _lib
from dynamical_model_amd.lac import Corrector, feature_index, nin_for, nout_for
from parity import PROGNOSTIC, small_config2

pytestmark = pytest.mark.gpu


def _model(case, **pover):
    hs = dm.build_case(case)
    params = dm.case_params(case, **pover)
    m = dm.Moloch()
    m.allocate_moloch(hs.jx, hs.iy, hs.kz, params)
    m.load_state(hs)
    m.init_moloch()
    return m, hs, params


def _fields(m, names=PROGNOSTIC):
    m.sync()
    return {n: m.get_field(n) for n in names}


def _same_interior(fa, fb):
    """Bitwise equality over the points a step defines (halo corners beyond
    the 1-wide ring are never read and may hold stale copies)."""
    for n in PROGNOSTIC:
        assert np.array_equal(dm.interior(fa[n]), dm.interior(fb[n])), n


def _relerr(a, b):
    a, b = dm.interior(a), dm.interior(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_training_pairs_match_oracle():
    case = small_config2(40, 36, 16)
    m, hs, params = _model(case)
    spp = 300
    corr = Corrector(m, nhid=64, samples_per_phase=spp, capacity=4 * spp, chunk_rows=1024, check_every=0, seed=7)
    print(f"case 40x36x16 fp64, collect 1 step, {spp} samples per phase", flush=True)
    corr.collect(1)
    assert corr.stats()["samples"] == 3 * spp
    X, T, cols = corr.dataset(0, spp)  # the first acoustic phase
    m.close()

    kz, jx, iy = case.kz, case.jx, case.iy
    dts = params.dt / params.nadv / params.nsound
    nb = params.nsound - 1

    def phase_start():
        o = oracle_lib.oracle_from_host(hs, params)
        o.call("ora_reset_tendencies")
        o.call("ora_prologue")
        return o

    start, long_, plain = phase_start(), phase_start(), phase_start()
    long_.S.nsound = 1
    long_.call("ora_sound", nb * dts)
    plain.S.nsound = nb
    plain.call("ora_sound", dts)

    j = cols % jx + 1
    i = cols // jx + 1
    # inputs: increments of the long substep, then theta and qv after it
    src = [long_.field(n) - start.field(n) for n in ("u", "v", "w", "pai")]
    src[2] = src[2][1:]
    src += [long_.field("tetav"), long_.field("qv")]
    fl, p = feature_index(kz)
    dj, di = p % 3 - 1, p // 3 - 1
    f, lev = fl // kz, fl % kz
    Xo = np.zeros((spp, nin_for(kz)), np.float32)
    for r in range(spp):
        jj = (j[r] + dj - 1) % jx + 1
        ii = (i[r] + di - 1) % iy + 1
        val = np.zeros(fl.size)
        ok = f < 6
        for ff in range(6):
            sel = ok & (f == ff)
            val[sel] = src[ff][lev[sel], ii[sel] + 1, jj[sel] + 1]
        Xo[r] = val.astype(np.float32)
    assert np.array_equal(X, Xo)
    # targets
    To = np.zeros((spp, nout_for(kz)))
    for ff, name in enumerate(["u", "v", "w", "pai"]):
        a, b = plain.field(name), long_.field(name)
        if name == "w":
            a, b = a[1:], b[1:]
        To[:, ff * kz:(ff + 1) * kz] = (a - b)[:, i + 1, j + 1].T
    for ff, name in enumerate(["u", "v", "w", "pai"]):
        t, to = T[:, ff * kz:(ff + 1) * kz], To[:, ff * kz:(ff + 1) * kz]
        err = np.abs(t - to).max() / np.abs(to).max()
        print(f"target {name}: max |T| {np.abs(to).max():.3e} rel err {err:.2e}")
        assert np.abs(to).max() > 0, name
        assert err < 1e-6, name


@pytest.mark.parametrize("precision", [dm.MLC_FP64, dm.MLC_FP32])
def test_collect_is_bitwise_ordinary(precision):
    case = small_config2(48, 40, 16)
    a, _, _ = _model(case, precision=precision)
    b, _, _ = _model(case, precision=precision)
    corr = Corrector(a, nhid=64, samples_per_phase=256, capacity=4096, chunk_rows=1024, check_every=0)
    print(f"case 48x40x16 precision={precision}: collect 3 steps vs 3 ordinary steps", flush=True)
    corr.collect(3)
    for n in range(1, 4):
        b.moloch(n)
    _same_interior(_fields(a), _fields(b))
    st = corr.stats()
    assert st["samples"] == 3 * 3 * 256 and st["plain_phases"] == 9
    a.close()
    b.close()


def _train(corr, iters):
    losses = [corr.train(iters // 2, 256, 1e-3), corr.train(iters // 4, 256, 3e-4),
              corr.train(iters // 4, 256, 1e-4)]
    print("losses", " ".join(f"{x:.4f}" for x in losses), flush=True)
    return losses


@pytest.mark.parametrize("precision,kz,nhid", [(dm.MLC_FP32, 16, 256), (dm.MLC_FP64, 16, 256),
                                               (dm.MLC_FP32, 40, 512)],
                         ids=["fp32-kz16", "fp64-kz16", "fp32-kz40-corrector-shape"])
def test_learned_increment_beats_long_substep(precision, kz, nhid):
    """The last case is the corrector of the 512 x 512 x 40 training set:
    2160 inputs, two hidden layers of 512, 160 outputs."""
    case = small_config2(64, 64, kz)
    learned, _, _ = _model(case, precision=precision)
    ordinary, _, _ = _model(case, precision=precision)
    reduced, _, _ = _model(case, precision=precision)
    cfg = dict(nhid=nhid, samples_per_phase=2048, capacity=65536, chunk_rows=4096, check_every=0, seed=3)
    corr = Corrector(learned, **cfg)
    rcorr = Corrector(reduced, **cfg)
    if kz == 40:
        assert corr.network().nin == 2160 and corr.network().nout == 160
    nwarm = 8
    print(f"case 64x64x{kz} precision={precision} nhid={nhid}: collect {nwarm} steps", flush=True)
    corr.collect(nwarm)
    for n in range(1, nwarm + 1):
        ordinary.moloch(n)
        reduced.moloch(n)
    losses = _train(corr, 2000)
    assert losses[-1] < 0.5
    corr.mode("on")
    rcorr.mode("reduced")
    for n in range(nwarm + 1, nwarm + 5):
        learned.moloch(n)
        ordinary.moloch(n)
        reduced.moloch(n)
    fl, fo, fr = _fields(learned), _fields(ordinary), _fields(reduced)
    st = corr.stats()
    assert st["learned_phases"] == 12 and not st["fallback"]
    gains = {}
    for n in ["u", "v", "w", "pai"]:
        el, er = _relerr(fl[n], fo[n]), _relerr(fr[n], fo[n])
        gains[n] = er / el
        print(f"{n:4s} learned {el:.3e}  long substep {er:.3e}  gain {er / el:.2f}")
        assert np.isfinite(fl[n]).all()
    for n in ["u", "v", "w", "pai"]:
        assert gains[n] > 1.5, (n, gains)
    for m in (learned, ordinary, reduced):
        m.close()


@pytest.mark.parametrize("shape,nhid", [((64, 64, 16), 128), ((70, 45, 20), 128), ((64, 64, 16), 512),
                                        ((64, 64, 40), 512), ((64, 48, 60), 512), ((64, 64, 9), 512)])
def test_fused_engine_matches_generic(shape, nhid):
    """The fused kernels (inputs gathered from the fields inside the MFMA
    kernel, packed weights; at 512 hidden units layers 2 and 3 in one kernel,
    with 64, 240 and 36 outputs, the last filling one 32-wide tile and part of
    the next) against materialised features + the generic GEMM: same trained
    network, one learned step each."""
    jx, iy, kz = shape
    case = small_config2(jx, iy, kz)
    runs = {}
    for engine in (0, 1):
        m, _, _ = _model(case, precision=dm.MLC_FP32)
        corr = Corrector(m, nhid=nhid, samples_per_phase=1024, capacity=8192, chunk_rows=1024, check_every=0,
                         seed=5, engine=engine)
        print(f"case {jx}x{iy}x{kz} nhid={nhid} fp32 engine={engine}", flush=True)
        corr.collect(2)
        corr.train(200, 256, 1e-3)
        w = [corr.network().get_weights(l) for l in range(3)]
        corr.mode("on")
        m.moloch(3)
        runs[engine] = (w, _fields(m), corr.stats())
        m.close()
    for l in range(3):
        assert np.array_equal(runs[0][0][l][0], runs[1][0][l][0])
    assert runs[1][2]["learned_phases"] == 3
    for n in PROGNOSTIC:
        a, b = dm.interior(runs[1][1][n]), dm.interior(runs[0][1][n])
        rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(f"{n:6s} fused vs generic rel {rel:.2e}")
        # both engines issue the same MFMA sequence in the same k order
        assert np.array_equal(a, b), n


def test_failed_check_falls_back_bitwise():
    case = small_config2(48, 40, 16)
    a, _, _ = _model(case, precision=dm.MLC_FP32)
    b, _, _ = _model(case, precision=dm.MLC_FP32)
    corr = Corrector(a, nhid=64, samples_per_phase=512, capacity=4096, chunk_rows=1024, check_every=1,
                     check_tol=0.0, retrain_steps=0)
    print("case 48x40x16 fp32: check tolerance 0 forces the fallback", flush=True)
    corr.collect(2)
    for n in (1, 2):
        b.moloch(n)
    corr.train(20, 128, 1e-3)
    corr.mode("on")
    for n in range(3, 6):
        a.moloch(n)
        b.moloch(n)
    _same_interior(_fields(a), _fields(b))
    st = corr.stats()
    print(st)
    assert st["fallback"] and st["checks"] == 3 and st["learned_phases"] == 0
    assert all(r > 0 for r in st["ratio"])
    # training with no iterations leaves the fallback; retraining clears it
    corr.train(0, 128, 1e-4)
    assert corr.stats()["fallback"]
    corr.train(1, 128, 1e-4)
    assert not corr.stats()["fallback"]
    a.close()
    b.close()


def test_nan_network_falls_back_bitwise():
    """A network that puts out NaN (its last bias NaN) fails its first check
    although NaN compares false with any tolerance; the run goes on with the
    ordinary substeps, bit-identical to a run without the corrector."""
    case = small_config2(48, 40, 16)
    a, _, _ = _model(case, precision=dm.MLC_FP32)
    b, _, _ = _model(case, precision=dm.MLC_FP32)
    corr = Corrector(a, nhid=64, samples_per_phase=512, capacity=4096, chunk_rows=1024, check_every=1,
                     check_tol=1.0, retrain_steps=0)
    print("case 48x40x16 fp32: NaN network output", flush=True)
    corr.collect(2)
    for n in (1, 2):
        b.moloch(n)
    corr.train(20, 128, 1e-3)
    net = corr.network()
    w, bias = net.get_weights(2)
    bias[:] = np.nan
    net.set_weights(2, w, bias)
    corr.mode("on")
    for n in range(3, 6):
        a.moloch(n)
        b.moloch(n)
    fa = _fields(a)
    _same_interior(fa, _fields(b))
    st = corr.stats()
    print(st)
    assert st["fallback"] and st["checks"] == 3 and st["learned_phases"] == 0
    assert all(np.isfinite(dm.interior(fa[n])).all() for n in fa)
    a.close()
    b.close()


def test_checks_pass_with_loose_tolerance():
    case = small_config2(48, 40, 16)
    a, _, _ = _model(case, precision=dm.MLC_FP32)
    corr = Corrector(a, nhid=64, samples_per_phase=512, capacity=4096, chunk_rows=1024, check_every=2,
                     check_tol=1e30, rel_tol=0.0)
    print("case 48x40x16 fp32: checks every 2 steps, loose tolerance", flush=True)
    corr.collect(2)
    corr.train(50, 128, 1e-3)
    corr.mode("on")
    for n in range(3, 8):
        a.moloch(n)
    st = corr.stats()
    print(st)
    # learned steps 0, 2, 4 are checked (3 phases each), the others use the increment
    assert st["checks"] == 9 and st["learned_phases"] == 6 and not st["fallback"]
    assert all(np.isfinite(st["rel"]))
    net = corr.network()
    w, b = net.get_weights(0)
    assert w.shape == (nin_for(case.kz), 64) and np.isfinite(w).all()
    a.close()


def test_corrector_errors():
    case = small_config2(32, 32, 8)
    a, _, _ = _model(case, nsound=1)
    with pytest.raises(dm.MolochError) as e:
        Corrector(a, nhid=64, samples_per_phase=16, capacity=64, chunk_rows=256)
    assert e.value.code == _lib.MLC_EINVAL
    a.close()
    b, _, _ = _model(case)
    corr = Corrector(b, nhid=64, samples_per_phase=16, capacity=64, chunk_rows=256)
    with pytest.raises(dm.MolochError) as e:
        corr.mode("on")  # not trained
    assert e.value.code == _lib.MLC_ESTATE
    with pytest.raises(dm.MolochError) as e:
        corr.train(1, 8, 1e-3)  # nothing collected
    assert e.value.code == _lib.MLC_ESTATE
    with pytest.raises(dm.MolochError):
        Corrector(b, nhid=64, samples_per_phase=16, capacity=64, chunk_rows=256)  # second corrector
    b.close()


def test_relative_error_budget_gates_the_check():
    """A trained network beats the uncorrected long substep (error ratios
    below 1, which check_tol = 1 accepts) but not a tight budget on the
    relative error: with rel_tol below its error the check fails and the run
    goes on with the ordinary loop; with a loose rel_tol it passes."""
    case = small_config2(64, 64, 16)
    out = {}
    for tag, rel_tol in (("tight", 1e-9), ("loose", 1e3)):
        m, _, _ = _model(case, precision=dm.MLC_FP32)
        corr = Corrector(m, nhid=256, samples_per_phase=2048, capacity=65536, chunk_rows=4096, check_every=1,
                         check_tol=1.0, rel_tol=rel_tol, retrain_steps=0, seed=3)
        print(f"case 64x64x16 fp32: rel_tol {rel_tol}", flush=True)
        corr.collect(8)
        _train(corr, 1000)
        corr.mode("on")
        m.moloch(9)
        out[tag] = corr.stats()
        print(tag, out[tag])
        m.close()
    for tag in out:
        assert out[tag]["worst_ratio"] < 1.0, tag  # better than no increment
        assert 0.0 < out[tag]["worst_rel"] < 1.0
    assert out["tight"]["fallback"] and out["tight"]["failed_checks"] == 1
    assert not out["loose"]["fallback"] and out["loose"]["failed_checks"] == 0


def test_failed_check_recollects_retrains_and_resumes():
    """After a failed check the handle collects fresh pairs for
    retrain_steps ordinary steps, retrains and resumes with a check; after
    max_retrains retrainings without a passed check it stays off.  With a
    budget no network meets, every step is a failed check (which continues
    from the ordinary substeps), a collection step or an ordinary step, so
    the trajectory stays bit-identical to the ordinary loop."""
    case = small_config2(48, 40, 16)
    a, _, _ = _model(case, precision=dm.MLC_FP32)
    b, _, _ = _model(case, precision=dm.MLC_FP32)
    corr = Corrector(a, nhid=64, samples_per_phase=512, capacity=4096, chunk_rows=1024, check_every=50,
                     check_tol=1e30, rel_tol=1e-12, retrain_steps=2, retrain_iters=20, retrain_batch=128,
                     retrain_lr=1e-3, max_retrains=2)
    print("case 48x40x16 fp32: unreachable budget, 2 retrainings", flush=True)
    corr.collect(2)
    for n in (1, 2):
        b.moloch(n)
    corr.train(20, 128, 1e-3)
    corr.mode("on")
    for n in range(3, 11):
        a.moloch(n)
        b.moloch(n)
    _same_interior(_fields(a), _fields(b))
    st = corr.stats()
    print(st)
    # steps 3, 6, 9 checked (the first learned step after each training),
    # 4-5 and 7-8 collect, 10 ordinary
    assert st["failed_checks"] == 3 and st["retrains"] == 2 and st["fallback"]
    assert st["collect_phases"] == 4 * case_nadv(a) and st["learned_phases"] == 0
    assert st["checks"] == 3 * case_nadv(a)
    assert st["samples"] == 2 * case_nadv(a) * 512  # the set restarted with the last retraining
    a.close()
    b.close()


def case_nadv(m):
    return m.params.nadv
