"""Tunnelling-only isothermal decay against Huntley's closed form, the reference's
analy_TL_iso (src/est_params/functions.py:264-268): n(t) = n0 exp(-rho' ln(1.8 s t)^3).

Three layers, each with its own stated tolerance:

1. The closed form against the exact model it approximates (oracle.analytic.
   nn_survival: nearest-hole tunnelling at rate s e^{-alpha r} in an infinite
   medium of Poisson holes), itself checked against an independent sampling of
   the nearest-hole distance.  Huntley's step approximation (the 1.8 = e^gamma
   shift) is good to 0.001 in n / n0 at rho' = 1e-5, 0.004 at 1e-4 and 0.02 at
   1e-3 over s t in [1, 1e10].
2. The replay harness (fixed-dt and event mode) on a tunnelling-only run with
   few electrons against many holes (5 % fill: 100 electrons, 3455 holes)
   against the expectation over its own geometry, mean of exp(-s t w_e) with the
   electrons' initial nearest holes: equal within 4 standard errors plus the
   fraction of electrons that share a nearest hole with another (only those can
   lose it, and a lost hole is replaced by a farther one).  The fixed-dt step
   needs no tolerance of its own: with a constant rate, survival over n steps of
   p = 1 - e^{-k dt} is e^{-k n dt} exactly.
3. The replay against the closed form itself, within the sum of the named
   errors: the approximation (1), hole sharing (2), and the corner placement of
   the electron cube in the hole box (electrons lie in [0, d)^3, holes in
   [0, 1.2 d)^3, so electrons near the three faces at 0 have fewer holes around
   them than in an infinite medium and decay more slowly: up to 0.025 in n / n0
   at rho' = 1e-4 by s t = 4e9, 0.002 at 1e-5).

The GPU tests run the same workloads through the fixed-dt and event kernels,
bit-exact against the replay and so inside the same bands.
"""
import math

import numpy as np
import pytest

from mcluminescence_amd import workloads as W
from oracle import analytic as A
from oracle import orc

DTS = (1.0, 1e3, 1e6)                  # three runs cover s t from 1 to 4e9 (s_eff = 1/s)
STEPS_CHECKED = (1, 3, 10, 30, 100, 300, 1000, 4000)
# (rho', closed form vs exact model, corner effect, MC vs closed form)
BANDS = {1e-5: (0.001, 0.003, 0.006), 1e-4: (0.004, 0.025, 0.045)}
SIMS = 128


def test_huntley_closed_form_against_nearest_neighbour_integral():
    st = np.logspace(0, 10, 21)
    for rp, tol in ((1e-5, 0.001), (1e-4, 0.004), (1e-3, 0.02)):
        h = A.analy_tl_iso(1.0, rp, 1.0, st)
        e = A.nn_survival(rp, 1.0, st)
        assert np.max(np.abs(h - e)) <= tol, (rp, np.max(np.abs(h - e)))
    # the shift is what makes it work: without the 1.8 the error at 1e-4 is 8x larger
    bad = np.exp(-1e-4 * np.log(st) ** 3)
    assert np.max(np.abs(bad - A.nn_survival(1e-4, 1.0, st))) > 0.025
    # special values: 1.8 s t = e^k gives n / n0 = exp(-rho' k^3); n0 scales linearly
    for k in (5.0, 10.0, 20.0):
        t = math.exp(k) / 1.8
        assert A.analy_tl_iso(250.0, 3e-4, 1.0, t) == pytest.approx(250.0 * math.exp(-3e-4 * k ** 3),
                                                                     rel=1e-12)


def test_nearest_neighbour_integral_against_sampling():
    # nearest-hole distance by inversion: P(u* > u) = exp(-rho' u^3)
    rng = np.random.default_rng(5)
    for rp in (1e-4, 1e-3):
        u = (-np.log(rng.uniform(size=400_000)) / rp) ** (1.0 / 3.0)
        for st in (1.0, 1e2, 1e4, 1e6, 1e8):
            s_mc = np.exp(-st * np.exp(-u))
            se = s_mc.std() / math.sqrt(u.size)
            assert abs(s_mc.mean() - A.nn_survival(rp, 1.0, st)[0]) < 5 * se + 1e-9, (rp, st)
    assert A.nn_survival(1e-4, 1.0, 0.0)[0] == pytest.approx(1.0, abs=1e-9)


def _runs(rp):
    return [W.tunnelling_iso(rp, dt=dt, sims=SIMS)[0] for dt in DTS]


def _geometry(runs):
    """Per run: initial w of every electron of every sim, and the fraction of
    electrons whose initial nearest hole is also another electron's."""
    out = []
    for r in runs:
        ws, shared, n = [], 0, 0
        for sim in range(r.sims):
            st = orc.init_state(r, 0, sim)
            ws.append(st["w"].astype(np.float64))
            _, inv, cnt = np.unique(st["nn"], return_inverse=True, return_counts=True)
            shared += int(np.sum(cnt[inv] > 1))
            n += st["nn"].size
        out.append((np.concatenate(ws), shared / n))
    return out


def _fixed_dt_survival(counts, run):
    n0 = orc.geometry(run).n_e0 * run.sims
    gone = np.cumsum(counts.sum(axis=0))
    return {n: 1.0 - gone[n - 1] / n0 for n in STEPS_CHECKED}


def _event_survival(lum, t, run):
    n0 = orc.geometry(run).n_e0 * run.sims
    return {n: 1.0 - lum[t <= n * run.dt + 1e-9 * run.dt].sum() / n0 for n in STEPS_CHECKED}


def _check_bands(rp, runs, geo, survivals, s_eff=1.0):
    """survivals: per run {step: surviving fraction}.  Every band adds 4 standard
    errors of the sample it is computed from (the Monte Carlo's electrons, or the
    geometry's, for the expectation over them)."""
    approx, corner, total = BANDS[rp]
    assert total >= approx + corner      # the closed-form band is the sum of named errors
    for r, (w, shared), surv in zip(runs, geo, survivals):
        n0 = orc.geometry(r).n_e0 * r.sims
        for n, s in surv.items():
            t = n * r.dt
            f = np.exp(-s_eff * t * w)
            g = float(f.mean())
            se_g = float(f.std()) / math.sqrt(f.size)
            se = math.sqrt(max(s * (1 - s), 1e-4) / n0)
            assert abs(s - g) <= 4 * se + shared, (rp, r.dt, n, s, g, se, shared)
            # the geometry's nearest-hole distances against the infinite medium
            inf = A.nn_survival(rp, s_eff, t)[0]
            assert abs(g - inf) <= corner + 4 * se_g, (rp, r.dt, n, g, inf, se_g)
            h = A.analy_tl_iso(1.0, rp, s_eff, t)
            assert abs(s - h) <= total + 4 * se, (rp, r.dt, n, s, h, se)


@pytest.mark.parametrize("rp", [1e-5, 1e-4])
def test_replay_fixed_dt_tunnelling_decay_follows_huntley(rp):
    runs = _runs(rp)
    counts, _ = orc.run_fixed_dt(runs, threads=8)
    sims = [counts[k * SIMS:(k + 1) * SIMS] for k in range(len(runs))]
    _check_bands(rp, runs, _geometry(runs), [_fixed_dt_survival(c, r) for c, r in zip(sims, runs)])


@pytest.mark.parametrize("rp", [1e-5, 1e-4])
def test_replay_event_mode_tunnelling_decay_follows_huntley(rp):
    runs = _runs(rp)
    ev = [_event_run(r) for r in runs]
    lum, t, _, steps = orc.run_event(ev, [r.dt * r.n_steps for r in runs], 128, threads=8)
    assert np.all(steps > 0)
    surv = [_event_survival(lum[k * SIMS:(k + 1) * SIMS], t[k * SIMS:(k + 1) * SIMS], r)
            for k, r in enumerate(runs)]
    _check_bands(rp, runs, _geometry(runs), surv)


def _event_run(r):
    """The same run for the event kernel: n_steps unused, no step cap at T_rate = 0."""
    from mcluminescence_amd._abi import RunParams
    return RunParams.make(**{**r.as_dict(), "n_steps": 1})


@pytest.mark.gpu
@pytest.mark.parametrize("rp", [1e-5, 1e-4])
def test_gpu_tunnelling_decay_bit_exact_and_follows_huntley(gpu_ctx, rp):
    runs = _runs(rp)
    print(f"tunnelling-only decay rho'={rp}: fixed-dt, {len(runs)} runs x {SIMS} clusters",
          flush=True)
    g, gl, _ = gpu_ctx.run_fixed_dt(runs)
    o, ol = orc.run_fixed_dt(runs, threads=8)
    np.testing.assert_array_equal(g, o)
    np.testing.assert_array_equal(gl, ol)
    geo = _geometry(runs)
    sims = [g[k * SIMS:(k + 1) * SIMS] for k in range(len(runs))]
    _check_bands(rp, runs, geo, [_fixed_dt_survival(c, r) for c, r in zip(sims, runs)])
    print(f"tunnelling-only decay rho'={rp}: event mode", flush=True)
    ev = [_event_run(r) for r in runs]
    dur = [r.dt * r.n_steps for r in runs]
    lum, t, er, steps, _ = gpu_ctx.run_event(ev, dur, 128)
    rl, rt, rer, rsteps = orc.run_event(ev, dur, 128, threads=8)
    np.testing.assert_array_equal(steps, rsteps)
    np.testing.assert_array_equal(lum, rl)
    np.testing.assert_array_equal(t, rt)
    np.testing.assert_array_equal(er, rer)
    surv = [_event_survival(lum[k * SIMS:(k + 1) * SIMS], t[k * SIMS:(k + 1) * SIMS], r)
            for k, r in enumerate(runs)]
    _check_bands(rp, runs, geo, surv)
