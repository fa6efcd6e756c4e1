"""The allocation-free entry point (mcl_run_fixed_dt_noalloc) against
mcl_run_fixed_dt with device outputs: identical counts, live electron-steps and
work totals, on two calls in a row (the second smaller, over the scratch the
first one left), and its rejection of inputs beyond the static scratch."""
import numpy as np
import pytest

from mcluminescence_amd import workloads as W
import ctypes as C

from mcluminescence_amd._abi import F_DEVICE_PTRS, MCL_ERANGE, MclError, Stats, run_array

pytestmark = pytest.mark.gpu


def _reference(ctx, runs, n_bins):
    import torch
    n = sum(r.sims for r in runs)
    row = n_bins if n_bins > 0 else max(r.n_steps for r in runs if r.sims > 0)
    counts = torch.full((n, row), -1, dtype=torch.int32, device="cuda")
    live = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    st = Stats()
    rc = ctx.lib.mcl_run_fixed_dt(ctx.handle, run_array(runs), len(runs),
                                  C.c_uint64(0x4D434C554D000001), int(n_bins), F_DEVICE_PTRS,
                                  C.c_void_p(counts.data_ptr()), C.c_void_p(live.data_ptr()),
                                  C.byref(st))
    ctx.check(rc, "mcl_run_fixed_dt")
    return counts.cpu().numpy(), live.cpu().numpy(), st


def _noalloc(ctx, runs, n_bins):
    import torch
    n = sum(r.sims for r in runs)
    row = n_bins if n_bins > 0 else max(r.n_steps for r in runs if r.sims > 0)
    counts = torch.full((n, row), -7, dtype=torch.int32, device="cuda")
    live = torch.full((n,), -7, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    st = ctx.run_fixed_dt_noalloc(runs, counts.data_ptr(), live.data_ptr(), n_bins=n_bins)
    return counts.cpu().numpy(), live.cpu().numpy(), st


def test_noalloc_matches_run_fixed_dt_on_two_calls(gpu_ctx):
    big = W.uneven_sweep(n_sets=600, n_steps=256, seed=5)     # uneven sizes, binned
    small = W.tl_glow_16x500(sims=4)                          # one run, per-step counts
    for name, runs, nb in (("600-set uneven sweep, 64 bins", big, 64),
                           ("16x500 TL glow, 4 sims, per step", small, 0)):
        print(f"no-alloc entry point: {name}", flush=True)
        c, l, st = _noalloc(gpu_ctx, runs, nb)
        rc, rl, rst = _reference(gpu_ctx, runs, nb)
        np.testing.assert_array_equal(c, rc)
        np.testing.assert_array_equal(l, rl)
        assert st.live_electron_steps == rst.live_electron_steps
        assert st.recombinations == rst.recombinations == int(rc.sum())
        assert st.cluster_steps == rst.cluster_steps
        assert st.clusters == rst.clusters
        assert st.kernel_ms > 0


def test_noalloc_rejects_inputs_beyond_its_scratch(gpu_ctx):
    import torch
    runs = W.iso_decay_4096x400(sims=110_000)      # 40 B per cluster: more than 4 MiB
    counts = torch.zeros((1, 1), dtype=torch.int32, device="cuda")
    with pytest.raises(MclError) as e:
        gpu_ctx.run_fixed_dt_noalloc(runs, counts.data_ptr(), None, n_bins=1)
    assert e.value.code == MCL_ERANGE
    assert "scratch" in str(e.value)
