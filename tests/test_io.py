"""Files of the reference's command-line driver: the initial-condition reader
(generic/read_initial_conditions.c:30-78, CONP records [t, y_0, param,
y_1 .. y_{N-1}]), the state log (generic/solver_main.c:47-66: t, then N values
per condition) and the binary network file.  Host code only: no GPU needed."""
import os

import numpy as np
import pytest

import accelerint_amd as acc
from accelerint_amd import mechanism as M


def _write_ics(path, y, var, t=0.0):
    n, num = y.shape
    rec = np.empty((num, n + 2))
    rec[:, 0] = t
    rec[:, 1] = y[0]
    rec[:, 2] = var
    rec[:, 3:] = y[1:].T
    rec.tofile(path)


@pytest.fixture
def fresh():
    acc.lib()
    yield acc
    acc.select_mechanism(acc.MECH_VAN_DER_POL)


def test_read_initial_conditions_vdp(fresh, tmp_path):
    acc.select_mechanism(acc.MECH_VAN_DER_POL)
    y0, mu = M.van_der_pol_conditions(37, seed=3)
    p = str(tmp_path / "ign_data.bin")
    _write_ics(p, y0, mu, t=0.5)
    y, var = acc.read_initial_conditions(p, 37)
    assert np.array_equal(y, y0) and np.array_equal(var, mu)
    # fewer conditions than the file holds: the first NUM records
    y, var = acc.read_initial_conditions(p, 5)
    assert np.array_equal(y, y0[:, :5]) and np.array_equal(var, mu[:5])


def test_read_initial_conditions_kinetics(fresh, tmp_path):
    mech = M.network(54)
    acc.set_kinetics(mech)
    y0, P = M.initial_conditions(mech, 19, seed=4, T_range=(800.0, 2200.0))
    p = str(tmp_path / "ign_data.bin")
    _write_ics(p, y0, P)
    y, var = acc.read_initial_conditions(p, 19)
    assert y.shape == (54, 19)
    assert np.array_equal(y, y0) and np.array_equal(var, P)


def test_short_or_missing_file_is_an_error(fresh, tmp_path):
    acc.select_mechanism(acc.MECH_VAN_DER_POL)
    y0, mu = M.van_der_pol_conditions(4, seed=5)
    p = str(tmp_path / "short.bin")
    _write_ics(p, y0, mu)
    with pytest.raises(acc.AccelerIntError, match="incorrectly formatted"):
        acc.read_initial_conditions(p, 5)
    with pytest.raises(acc.AccelerIntError, match="Could not open"):
        acc.read_initial_conditions(str(tmp_path / "missing.bin"), 1)


def test_write_log_records(fresh, tmp_path):
    acc.select_mechanism(acc.MECH_VAN_DER_POL)
    rng = np.random.default_rng(6)
    p = str(tmp_path / "log.bin")
    ys = [rng.standard_normal((2, 7)) for _ in range(3)]
    for k, y in enumerate(ys):
        acc.write_log(p, 1e-3 * k, y, append=k > 0)
    raw = np.fromfile(p).reshape(3, 1 + 7 * 2)
    for k, y in enumerate(ys):
        assert raw[k, 0] == 1e-3 * k
        assert np.array_equal(raw[k, 1:].reshape(7, 2), y.T)


def test_network_file_round_trip(fresh, tmp_path):
    mech = M.network(14)
    p = str(tmp_path / "net.bin")
    mech.write_binary(p)
    acc.load_kinetics(p)
    assert acc.num_states() == 14
    raw = np.fromfile(p, dtype=np.int32, count=2)
    assert tuple(raw) == (mech.ns, mech.nr)
    assert os.path.getsize(p) == 8 + 8 * (mech.ns * 6 + mech.nr * 3) + 4 * 6 * mech.nr
    bad = str(tmp_path / "bad.bin")
    with open(bad, "wb") as f:
        f.write(b"\x05\x00")
    with pytest.raises(acc.AccelerIntError, match="not a kinetics network file"):
        acc.load_kinetics(bad)
