"""Host logic of the surrogate and the fit loop (no GPU): parameter encoding,
curve normalisation, the NRMSE and objective definitions, differential evolution."""
import numpy as np
import pytest

from mcluminescence_amd import fit as F
from mcluminescence_amd import surrogate as S
from mcluminescence_amd.workloads import rho_prime_for


def test_encode_decode_round_trip_and_box():
    rng = np.random.default_rng(0)
    th = S.sample_params(1000, rng)
    assert np.all(th >= S.LOW * (1 - 1e-12)) and np.all(th <= S.HIGH * (1 + 1e-12))
    u = S.encode(th)
    assert np.all(np.abs(u) <= 1 + 1e-6)
    np.testing.assert_allclose(S.decode(u), th, rtol=1e-6)
    # log axes: the box centre of E is the geometric mean
    assert abs(S.decode(np.zeros((1, 6)))[0, 0] - np.sqrt(1.0 * 1.8)) < 1e-12


@pytest.mark.parametrize("c0", [1e-3, 1e-4])
def test_normalise_round_trip(c0):
    c = np.array([0.0, 1e-6, 1e-4, 0.01, 0.3, 1.0])
    y = S.normalise(c, c0)
    assert y[0] == 0 and abs(y[-1] - 1) < 1e-6 and np.all(np.diff(y) > 0)
    np.testing.assert_allclose(S.denormalise(y, c0), c, rtol=1e-5, atol=1e-9)


def test_nrmse_definition():
    ref = np.array([[0.0, 1.0, 4.0, 1.0]])
    assert S.nrmse(ref, ref)[0] == 0
    pred = ref + np.array([[0.0, 0.0, 0.4, 0.0]])
    assert abs(S.nrmse(pred, ref)[0] - np.sqrt(0.16 / 4) / 4.0) < 1e-15


def test_objective_is_squared_nrmse_against_the_measured_curve():
    c = np.array([0.1, 0.5, 0.2])
    y = c * 1000.0
    assert F.objective(c, y, 1000.0)[0] == 0
    off = F.objective(c * 1.1, y, 1000.0)[0]
    assert abs(np.sqrt(off) - S.nrmse(c * 1.1 * 1000, y)[0]) < 1e-12


def test_full_theta_fixes_the_heating_rate():
    th = F.full_theta(np.zeros(5), 2.5)
    assert th.shape == (1, 6) and th[0, 5] == 2.5


def test_differential_evolution_finds_a_quadratic_minimum():
    rng = np.random.default_rng(1)
    de = F.DifferentialEvolution(5, 48, rng)
    target = np.array([0.3, -0.2, 0.5, 0.1, -0.6])
    f = lambda x: ((x - target) ** 2).sum(axis=1)
    de.seed_fitness(f(de.x))
    for _ in range(150):
        trial = de.propose()
        assert np.all(np.abs(trial) <= 1)
        de.accept(trial, f(trial))
    x, fx = de.best
    assert fx < 1e-10
    np.testing.assert_allclose(x, target, atol=1e-5)


def test_recovery_uses_rho_prime():
    th = np.array([1.3, 1e12, 4e9, 1e23, 0.5, 1.0])
    # same rho' with another alpha: no rho' error
    th2 = th.copy()
    th2[2] = 5e9
    th2[3] = 1e23 * (5e9 / 4e9) ** 3
    r = F.FitResult(th2, 0.0, 0.0, 0.0, 0, 0).recovery(th)
    assert abs(rho_prime_for(th2[3], th2[2]) / rho_prime_for(th[3], th[2]) - 1) < 1e-12
    assert r["log10_rho_prime"] < 1e-12 and r["E"] == 0


def test_run_table_is_byte_identical_to_the_per_set_runs():
    """surrogate.tl_table (column-wise) gives exactly the mcl_run_params of tl_run
    per parameter set, and run tables pass through run_array unchanged."""
    from mcluminescence_amd import surrogate as S
    from mcluminescence_amd._abi import RUN_DTYPE, run_array
    theta = S.sample_params(257, np.random.default_rng(5))
    per_set = bytes(run_array([S.tl_run(th, 400, 8) for th in theta]))
    table = S.tl_table(theta, 400, 8)
    assert table.dtype == RUN_DTYPE
    assert bytes(run_array(table)) == per_set
    arr = run_array(table)
    assert run_array(arr) is arr


def test_iso_edges_are_log_time_bins():
    e = S.iso_edges()
    assert e[0] == 0 and e[-1] == S.ISO_STEPS and e.size == S.N_BINS + 1
    d = np.diff(e)
    assert np.all(d >= 1) and np.all(d[:S.ISO_LINEAR_BINS] == 1)
    # geometric beyond the single steps: within half a step of L r^k, r = (4000 / L)^(1/209)
    L = S.ISO_LINEAR_BINS
    r = (S.ISO_STEPS / L) ** (1.0 / (S.N_BINS - L))
    k = np.arange(S.N_BINS - L + 1)
    assert np.all(np.abs(e[L:] - L * r ** k) <= 0.5 + 1e-9)
    np.testing.assert_array_equal(np.asarray(S.ISO.edges), e)
    t = S.ISO.axis()
    assert t[0] == S.ISO_DT and t[-1] == 2000.0


def test_iso_kind_box_and_run_table():
    from mcluminescence_amd._abi import run_array
    rng = np.random.default_rng(2)
    th = S.ISO.sample_params(300, rng)
    assert np.all(th[:, 5] >= 400) and np.all(th[:, 5] <= 600)
    np.testing.assert_allclose(S.ISO.decode(S.ISO.encode(th)), th, rtol=1e-6)
    per_set = bytes(run_array([S.ISO.run(t, 400, 8) for t in th]))
    assert bytes(run_array(S.ISO.table(th, 400, 8))) == per_set
    r = S.ISO.run(th[0], 400, 8)
    assert r.T_rate == 0 and r.dt == S.ISO_DT and r.n_steps == S.ISO_STEPS
    assert abs(r.T_start + 273.15 - th[0, 5]) < 1e-9


def test_iso_recovery_is_in_the_rate_and_fill():
    th = np.array([1.4, 1e13, 4e9, 1e23, 0.5, 500.0])
    # another (E, s) with the same k at T_iso: no rate error
    th2 = th.copy()
    th2[0] = 1.5
    th2[1] = th[1] * np.exp((th2[0] - th[0]) / (8.617343e-5 * 500.0))
    r = F.FitResult(th2, 0.0, 0.0, 0.0, 0, 0, kind="iso").recovery(th)
    assert set(r) == {"log10_k", "fill"} and r["log10_k"] < 1e-12
    assert F.iso_rate(th)[0] == pytest.approx(0.07738, rel=1e-3)


def test_iso_truths_resolve_their_decay():
    from mcluminescence_amd import fit_bench
    rng = np.random.default_rng(4)
    for _ in range(20):
        th = fit_bench.draw_truth_iso(rng)
        tau = 1.0 / F.iso_rate(th)[0]
        assert fit_bench.ISO_TIME_CONSTANT[0] <= tau <= fit_bench.ISO_TIME_CONSTANT[1]
        assert 420 <= th[5] <= 580 and 0.3 <= th[4] <= 1.0
