"""The command-line driver (accelerint_amd/accelerint_main, the counterpart of
generic/solver_main.cu) end to end: ign_data.bin in, log/<solver>-log.bin out,
final states against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from accelerint_amd import mechanism as M

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "accelerint_amd", "accelerint_main")


def _ics(path, y, var):
    n, num = y.shape
    rec = np.zeros((num, n + 2))
    rec[:, 1] = y[0]
    rec[:, 2] = var
    rec[:, 3:] = y[1:].T
    rec.tofile(path)


def _run(cwd, *args):
    assert os.path.exists(CLI), "build the CLI first (python -m accelerint_amd.build)"
    r = subprocess.run([CLI, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    return r.stdout


def test_cli_van_der_pol_logged_steps(tmp_path, oracle_lib):
    num, steps, dt = 256, 10, 1e-3
    y0, mu = M.van_der_pol_conditions(num, seed=1234)
    _ics(tmp_path / "ign_data.bin", y0, mu)
    (tmp_path / "log").mkdir()
    out = _run(tmp_path, num, 0, "--end", steps * dt, "--step", dt, "--log")
    assert "TFinal:" in out and "Time per step:" in out
    log = np.fromfile(tmp_path / "log" / "radau2a-int-gpu-log.bin").reshape(steps + 1, 1 + num * 2)
    assert np.allclose(log[:, 0], dt * np.arange(steps + 1))
    assert np.array_equal(log[0, 1:].reshape(num, 2).T, y0)
    yo, co, _ = oracle_lib.integrate_steps(oracle_lib.Mechanism.van_der_pol(), y0, mu, dt, steps, nthreads=8)
    y = log[-1, 1:].reshape(num, 2).T
    we = np.abs(y - yo) / (1e-10 + 1e-6 * np.abs(yo))
    print("CLI vdP final state: max weighted err vs oracle %.3g" % we.max())
    assert (co == 0).all() and we.max() <= 10.0


def test_cli_kinetics_network_file(tmp_path, oracle_lib):
    mech = M.network(14)
    num, steps, dt = 512, 5, 1e-6
    y0, P = M.initial_conditions(mech, num, seed=77)
    mech.write_binary(str(tmp_path / "net.bin"))
    _ics(tmp_path / "ign_data.bin", y0, P)
    (tmp_path / "log").mkdir()
    _run(tmp_path, num, 0, "--end", steps * dt, "--step", dt, "--mech", "net.bin", "--log-end-only")
    log = np.fromfile(tmp_path / "log" / "radau2a-int-gpu-log.bin").reshape(2, 1 + num * 14)
    assert log[1, 0] == steps * dt
    y = log[1, 1:].reshape(num, 14).T
    om = oracle_lib.Mechanism.kinetics(mech)
    yo, co, _ = oracle_lib.integrate_steps(om, y0, P, dt, steps, nthreads=8)
    we = np.abs(y - yo) / (1e-10 + 1e-6 * np.abs(yo))
    print("CLI 14-state final state: max weighted err vs oracle %.3g" % we.max())
    assert (co == 0).all() and we.max() <= 10.0
