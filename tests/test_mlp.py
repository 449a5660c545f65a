"""Learned Newton start: the network's kernels against an FP64 CPU restatement
(tests/mlp_ref.py), and the trained network inside the integrator.

* CPU: the restatement's backward pass against central differences (FP64, no
  BF16 rounding), which pins the gradient formulas it shares with mlp.hip.
* GPU: loss, gradient and batched forward pass of the BF16-MFMA kernels
  against the restatement with the same BF16 operand rounding (norm-wise
  relative error <= 2e-3: FP32 versus FP64 accumulation); the integrator's
  per-wavefront inference path against the batched path; training reduces the
  loss; and an A/B run on fresh conditions shows what the trained start changes
  (Newton iterations, accepted steps) with both runs held to the oracle by
  tests/arbiter.py.
"""
import numpy as np
import pytest

import mlp_ref
from arbiter import arbitrate

N54 = 54


def _rand_params(net, rng, scale=0.3):
    parts = []
    for s in net.shapes:
        a = rng.standard_normal(s) * (scale if len(s) == 2 else 0.1)
        parts.append(a)
    W1, b1, W2, b2, W3, b3 = parts
    W1[net.inn:, :] = 0.0
    W3[:, net.out:] = 0.0
    b3[net.out:] = 0.0
    return net.pack([W1, b1, W2, b2, W3, b3]).astype(np.float32)


def test_reference_gradient_matches_central_differences():
    rng = np.random.default_rng(5)
    net = mlp_ref.Net(3)
    p = _rand_params(net, rng).astype(np.float64)
    x = rng.standard_normal((16, net.inn))
    t = rng.standard_normal((16, net.out))
    L, g = net.loss_grad(p, x, t, bf=False)
    idx = rng.choice(np.flatnonzero(np.abs(g) > 0), 60, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (net.loss(pp, x, t, bf=False) - net.loss(pm, x, t, bf=False)) / (2 * h)
        # truncation O(h^2) plus the rounding of the two loss values, eps |L| / h
        assert abs(fd - g[i]) <= 1e-5 * abs(g[i]) + 16 * np.finfo(float).eps * abs(L) / h, (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == 1.0 + 2 * 2.0 ** -7      # tie -> even (up)
    assert r[3] == -2.5
    assert np.isfinite(r[4])


@pytest.fixture
def accmod():
    import accelerint_amd as acc
    acc.lib()
    yield acc
    acc.lib().accelerInt_mlp_release()


@pytest.mark.gpu
def test_mlp_loss_gradient_against_cpu_restatement(accmod):
    acc = accmod
    rng = np.random.default_rng(11)
    net = mlp_ref.Net(N54)
    acc.mlp_init(N54, 3)
    p = _rand_params(net, rng)
    acc.mlp_set_params(p)
    assert np.array_equal(acc.mlp_params(), p)
    B = 64
    x = rng.standard_normal((B, net.inn)).astype(np.float32)
    t = rng.standard_normal((B, net.out)).astype(np.float32)
    loss_g, g_g = acc.mlp_loss_grad(x, t)
    loss_c, g_c = net.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss_g - loss_c) <= 1e-4 * loss_c
    o = 0
    for s, name in zip(net.shapes, ["W1", "b1", "W2", "b2", "W3", "b3"]):
        k = int(np.prod(s))
        a, b = g_g[o:o + k].astype(np.float64), g_c[o:o + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print("grad %s: relative error %.2e" % (name, rel))
        assert rel <= 2e-3, name
        o += k
    y_g = acc.mlp_predict(x, net.out)
    y_c = net.predict(p.astype(np.float64), x.astype(np.float64), bf=True)
    assert np.abs(y_g - y_c).max() <= 2e-3 * np.abs(y_c).max()


@pytest.mark.gpu
def test_mlp_in_kernel_inference_matches_batched(accmod):
    acc = accmod
    from accelerint_amd import mechanism as M
    rng = np.random.default_rng(12)
    net = mlp_ref.Net(N54)
    acc.mlp_init(N54, 4)
    p = _rand_params(net, rng)
    acc.mlp_set_params(p)
    mech = M.network(54)
    y0, P0 = M.initial_conditions(mech, 64, seed=3, T_range=(800.0, 2200.0))
    h = 5e-7
    phi = np.concatenate([1000.0 / y0[:1].T, np.log(np.maximum(y0[1:].T, 1e-30))], axis=1)
    feats = np.concatenate([phi, np.full((64, 1), np.log(h)), np.log(P0)[:, None]], axis=1)
    mu = feats.mean(0)
    sd = np.where(feats.std(0) > 0, feats.std(0), 1.0)
    zs = rng.uniform(0.5, 2.0, net.out)
    acc.mlp_set_normalization(mu, sd, zs)
    z_k = acc.mlp_predict_states(np.ascontiguousarray(y0.T), h, P0)
    x = ((feats - mu) / sd).astype(np.float32)
    z_b = acc.mlp_predict(x, net.out).astype(np.float64) * zs
    err = np.abs(z_k - z_b).max() / np.abs(z_b).max()
    print("in-kernel vs batched inference: max relative difference %.2e" % err)
    assert err <= 2e-2


@pytest.mark.gpu
def test_learned_start_ab(accmod, oracle_lib):
    acc = accmod
    from accelerint_amd import learned, mechanism as M
    mech = M.network(54)
    dt = 1e-6
    acc.set_kinetics(mech)
    acc.set_tolerances(1e-10, 1e-6)
    acc.set_const_time_step(False)
    acc.set_abort_on_error(False)
    acc.initialize(8192, 0)
    try:
        y_tr, P_tr = M.initial_conditions(mech, 8192, seed=777, T_range=(800.0, 2200.0))
        X, Z = learned.collect(y_tr, P_tr, 8, dt)
        print("collected %d samples" % X.shape[0])
        assert X.shape[0] >= 8192 * 7
        hist = learned.train(X, Z, epochs=30, batch=512, lr=2e-3, seed=1)
        print("training loss per epoch: first %.4g last %.4g" % (hist[0], hist[-1]))
        assert hist[-1] < 0.5 * hist[0]
        # A/B on conditions not seen in training
        num, steps = 2048, 4
        y0, P = M.initial_conditions(mech, num, seed=778, T_range=(800.0, 2200.0))
        res = {}
        for on in (False, True):
            y, codes, cnt = learned.integrate_ab(y0, P, steps, dt, acc.PRECISION_MIXED, on)
            res[on] = (y, codes, cnt)
        om = oracle_lib.Mechanism.kinetics(mech)
        yo, co, cnt_o = oracle_lib.integrate_steps(om, y0, P, dt, steps, nthreads=8)
        for on in (False, True):
            y, codes, cnt = res[on]
            assert (codes == co).all()
            print("learned=%s: newton %d, steps %d, f %d, lu %d"
                  % (on, cnt[:, 6].sum(), cnt[:, 0].sum(), cnt[:, 7].sum(), cnt[:, 4].sum()))
            arbitrate("learned=%s" % on, y, yo, oracle_lib, om, y0, P, dt, steps)
        gain = 1.0 - res[True][2][:, 6].sum() / res[False][2][:, 6].sum()
        print("Newton iterations saved by the learned start: %.2f %%" % (100 * gain))
    finally:
        acc.cleanup()


def _garbage_net_run(acc, oracle_lib, params_fn, label):
    from accelerint_amd import learned, mechanism as M
    mech = M.network(54)
    dt, num, steps = 1e-6, 512, 2
    acc.set_kinetics(mech)
    acc.set_tolerances(1e-10, 1e-6)
    acc.set_const_time_step(False)
    acc.set_abort_on_error(False)
    acc.initialize(num, 0)
    try:
        net = mlp_ref.Net(N54)
        acc.mlp_init(N54, 5)
        acc.mlp_set_params(params_fn(net))
        y0, P = M.initial_conditions(mech, num, seed=779, T_range=(800.0, 2200.0))
        acc.mlp_set_normalization(np.zeros(net.inn), np.ones(net.inn), np.full(net.out, 1e-2))
        out = {}
        for on in (False, True):
            out[on] = learned.integrate_ab(y0, P, steps, dt, acc.PRECISION_FP64, on)
        om = oracle_lib.Mechanism.kinetics(mech)
        yo, co, _ = oracle_lib.integrate_steps(om, y0, P, dt, steps, nthreads=8)
        y, codes, cnt = out[True]
        assert np.isfinite(y).all()
        assert (codes == co).all()
        arbitrate(label, y, yo, oracle_lib, om, y0, P, dt, steps)
        return out
    finally:
        acc.cleanup()


@pytest.mark.gpu
def test_learned_start_nan_network_falls_back(accmod, oracle_lib):
    """A network whose prediction is not finite is not used: the run is the
    zero-start run, bit for bit."""
    out = _garbage_net_run(accmod, oracle_lib, lambda net: np.full(net.count, np.nan, np.float32), "nan network")
    assert np.array_equal(out[True][0], out[False][0])
    assert np.array_equal(out[True][2], out[False][2])


@pytest.mark.gpu
def test_learned_start_garbage_network_falls_back(accmod, oracle_lib):
    """A network predicting large finite garbage: its first Newton correction
    (non-finite residuals included) is not smaller than the corrected stages,
    so the step restarts from the zero start; the results stay finite and
    held to the oracle, at the price of the discarded iterations."""
    rng = np.random.default_rng(13)
    out = _garbage_net_run(accmod, oracle_lib, lambda net: (_rand_params(net, rng, scale=3.0) * 50).astype(np.float32),
                           "garbage network")
    print("Newton iterations: learned %d, zero start %d" % (out[True][2][:, 6].sum(), out[False][2][:, 6].sum()))
    assert out[True][2][:, 6].sum() > out[False][2][:, 6].sum()


@pytest.mark.gpu
def test_sample_log_into_caller_buffers(accmod):
    """accelerInt_log_first_steps: one sample per condition and global step in
    caller-owned device buffers, identical to what the library-owned log
    (accelerInt_start_log / accelerInt_read_log) records for the same
    integration; the first global step's samples carry the features of the
    initial states (1000/T, log Y_k, log p) and an h no larger than the
    initial 5e-7 (rejections before the first acceptance only shrink it)."""
    from hipmem import DeviceArray
    from accelerint_amd import mechanism as M
    acc = accmod
    mech = M.network(54)
    num, steps, dt = 256, 2, 1e-6
    y0, P = M.initial_conditions(mech, num, seed=95, T_range=(800.0, 2200.0))
    acc.set_kinetics(mech)
    acc.set_tolerances(1e-10, 1e-6)
    acc.set_const_time_step(False)
    acc.set_abort_on_error(False)
    acc.set_precision(acc.PRECISION_FP64)
    acc.initialize(num, 0)

    def rows_sorted(a):
        return a[np.lexsort(a.T[::-1])]

    try:
        cap = num * steps + 16
        xd = DeviceArray(np.zeros((cap, 56), np.float32))
        zd = DeviceArray(np.zeros((cap, 162), np.float32))
        cd = DeviceArray(np.zeros(1, np.int32))
        acc.log_first_steps(xd.addr, zd.addr, cap, cd.addr)
        y = y0.copy()
        acc.integrate(num, 0.0, dt, dt, y, P)  # first global step: one sample per condition
        n1 = int(cd.get()[0])
        X1 = xd.get()[:n1]
        acc.integrate(num, dt, steps * dt, dt, y, P)
        n = int(cd.get()[0])
        acc.log_first_steps(0, 0, 0, 0)
        X, Z = xd.get()[:n], zd.get()[:n]
        assert n1 == num and n == num * steps
        assert np.isfinite(X).all() and np.isfinite(Z).all()
        ref = np.concatenate([1000.0 / y0[:1].T, np.log(np.maximum(y0[1:].T, 1e-30)),
                              np.log(P)[:, None]], axis=1).astype(np.float32)
        got = X1[:, list(range(54)) + [55]]
        assert np.allclose(rows_sorted(got), rows_sorted(ref), rtol=1e-5, atol=1e-5)
        assert (X1[:, 54] <= np.float32(np.log(5e-7)) + 1e-5).all()

        acc.start_log(cap)
        y2 = y0.copy()
        acc.integrate(num, 0.0, steps * dt, dt, y2, P)
        X2, Z2 = acc.read_log(cap, 54)  # also ends the library-owned log
        assert X2.shape[0] == n
        XZ, XZ2 = np.concatenate([X, Z], 1), np.concatenate([X2, Z2], 1)
        assert np.array_equal(rows_sorted(XZ), rows_sorted(XZ2))
    finally:
        acc.cleanup()


@pytest.mark.gpu
def test_device_resident_training_matches_host_entry(accmod):
    """accelerInt_mlp_train on device buffers and accelerInt_mlp_train_host on
    the same samples from the same initial network give the same parameters."""
    from hipmem import DeviceArray
    acc = accmod
    rng = np.random.default_rng(17)
    net = mlp_ref.Net(N54)
    n = 2048
    x = rng.standard_normal((n, net.inn)).astype(np.float32)
    t = (0.3 * rng.standard_normal((n, net.out))).astype(np.float32)
    acc.mlp_init(N54, 9)
    h_host = acc.mlp_train_host(x, t, 2, 256, 1e-3, 5)
    p_host = acc.mlp_params()
    acc.mlp_init(N54, 9)
    xd, td = DeviceArray(x), DeviceArray(t)
    h_dev = acc.mlp_train(n, xd.addr, td.addr, 2, 256, 1e-3, 5)
    p_dev = acc.mlp_params()
    print("loss per epoch: host %s device %s" % (h_host, h_dev))
    assert h_dev[-1] < h_dev[0]
    assert np.array_equal(p_host, p_dev) and np.array_equal(h_host, h_dev)
