"""Batch plans (sfbhip_batch_plan_*): the batched mixed decomposition planned
once, enqueued as kernels only, captured in a hipGraph and replayed.  Every
run is checked against the FP64 batch (sfbhip_direct_batched, itself checked
against the oracle in test_batched_gpu.py); graph replays are checked
bit for bit against an eager enqueue of the same plan, with eager torch work
and synchronisations between the replays."""
import math

import numpy as np
import pytest
import torch

from catana_amd import capi, workloads

pytestmark = pytest.mark.gpu

SPECS = [  # (N, lmax, nmax, rmax, k_max, window volume, generator)
    (3000, 12, 8, 300.0, 0.0, 0.0, "uniform"),
    (20000, 40, 30, 800.0, 0.12, 0.0, "clustered"),
    (517, 9, 5, 150.0, 0.0, 2.5e7, "uniform"),
    (9000, 33, 20, 1200.0, 0.05, 0.0, "uniform"),
    (1, 6, 4, 50.0, 0.0, 0.0, "uniform"),
    (40000, 70, 40, 1500.0, 0.08, 0.0, "uniform"),
]


def _catalogs(specs=SPECS):
    out = []
    for i, (n, lmax, nmax, rmax, kmax, vol, gen) in enumerate(specs):
        if gen == "uniform":
            pts = workloads.uniform_ball(n, rmax, seed=140 + i)
        else:
            pts = workloads.clustered_ball(n // 40, rmax, mean_children=40, sigma=6.0, seed=140 + i)
        out.append(dict(np=pts, points=torch.from_numpy(pts).cuda(), lmax=lmax, nmax=nmax, rmax=rmax, k_max=kmax,
                        window_volume=vol))
    return out


def _flmn_err(a, b):
    return max(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300) for x, y in zip(a, b))


def _check_vs_fp64(cats, outs, handle, tol=1e-6):
    r64 = capi.sfb_direct_batched_torch(handle, cats, store_flmn=True, precision=capi.PREC_FP64)
    torch.cuda.synchronize()
    for c, (k64, c64, f64), (k, cl, f) in zip(cats, r64, outs):
        np.testing.assert_array_equal(k.cpu().numpy(), k64.cpu().numpy())
        a, b = cl.cpu().numpy(), c64.cpu().numpy()
        assert np.all((a == 0) == (b == 0))
        m = b != 0
        assert np.max(np.abs(a[m] / b[m] - 1)) < tol
        if f is not None:
            L, N = c["lmax"], c["nmax"]
            assert _flmn_err(capi.split_flmn(f.cpu().numpy(), L, N), capi.split_flmn(f64.cpu().numpy(), L, N)) < tol


def _snapshot(outs):
    return [tuple(None if t is None else t.clone() for t in o) for o in outs]


def test_plan_eager_vs_fp64_batch(handle):
    cats = _catalogs()
    print("case: batch plan, eager enqueue, 6 catalogues", flush=True)
    plan = capi.BatchPlan(handle, cats, store_flmn=True)
    plan.enqueue()
    ms = plan.sync()
    torch.cuda.synchronize()
    assert ms[1] > 0 and ms[3] >= ms[1]
    _check_vs_fp64(cats, plan.outputs, handle)
    # a second enqueue of the same plan gives identical results (fixed-order reduction)
    first = _snapshot(plan.outputs)
    plan.enqueue()
    plan.sync()
    torch.cuda.synchronize()
    for a, b in zip(first, plan.outputs):
        for x, y in zip(a, b):
            if x is not None:
                assert torch.equal(x, y)
    plan.close()


def test_plan_graph_replays_match_eager(handle):
    cats = _catalogs()
    print("case: batch plan, graph capture and 3 replays with new points", flush=True)
    plan = capi.BatchPlan(handle, cats, store_flmn=True)
    plan.capture()
    for rep in range(3):
        # new points in the bound tensors (eager torch work on the default stream)
        for c in cats:
            p = c["points"]
            p[:, 2] = torch.remainder(p[:, 2] + 0.37 * (rep + 1), 2 * math.pi)
            p[:, 0] *= 0.97
        torch.cuda.synchronize()
        plan.enqueue()
        plan.sync()
        torch.cuda.synchronize()
        eager = _snapshot(plan.outputs)
        for o in plan.outputs:  # the replay must rewrite every output
            for t in o:
                if t is not None:
                    t.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.launch()
        ms = plan.sync()
        torch.cuda.synchronize()
        assert ms[3] > 0
        for a, b in zip(eager, plan.outputs):
            for x, y in zip(a, b):
                if x is not None:
                    assert torch.equal(x, y), f"replay {rep} differs from the eager enqueue"
        if rep == 0:
            _check_vs_fp64(cats, plan.outputs, handle)
    plan.close()


def test_plan_radius_bound_and_bad_points(handle):
    cats = _catalogs(SPECS[:2])
    far = cats[0]["points"].clone()
    far[7, 0] = 1.1 * cats[0]["rmax"]
    cats_far = [dict(cats[0], points=far), cats[1]]
    print("case: batch plan, a point beyond rmax", flush=True)
    plan = capi.BatchPlan(handle, cats_far, store_flmn=True)
    plan.enqueue()
    with pytest.raises(ValueError, match="radius bound"):
        plan.sync()
    plan.close()
    # with a bound that covers it the result matches the FP64 batch
    plan = capi.BatchPlan(handle, cats_far, store_flmn=True, r_bound=[1.2 * cats[0]["rmax"], 0.0])
    plan.enqueue()
    plan.sync()
    torch.cuda.synchronize()
    _check_vs_fp64(cats_far, plan.outputs, handle)
    plan.close()
    bad = cats[0]["points"].clone()
    bad[3, 1] = float("nan")
    plan = capi.BatchPlan(handle, [dict(cats[0], points=bad)], store_flmn=False)
    plan.enqueue()
    with pytest.raises(ValueError, match="r < 0 or a non-finite"):
        plan.sync()
    plan.close()


def test_plan_refusals(handle):
    cats = _catalogs(SPECS[:1])
    with pytest.raises(capi.SfbError, match="SPLIT_FP16"):
        capi.BatchPlan(handle, cats, mixed_operands=capi.OPERANDS_FP32)
    with pytest.raises(capi.SfbError, match="MIXED_KERNEL_512 is not one"):
        capi.BatchPlan(handle, cats, mixed_kernel=capi.MIXED_KERNEL_512)
    with pytest.raises(ValueError, match="r_bound"):
        capi.BatchPlan(handle, cats, r_bound=[float("inf")])
    t = capi.get_tile_table(handle)
    bad = dict(t, chunks=[3] * capi.TILE_BANDS)
    with pytest.raises(ValueError, match="chunks"):
        capi.set_tile_table(handle, bad)
    assert capi.get_tile_table(handle) == t


def test_autotune_and_tuned_plan(handle):
    cats = _catalogs([SPECS[1], SPECS[3], SPECS[5]])
    print("case: autotuner on 3 catalogues", flush=True)
    capi.set_tile_table(handle, None)
    t = capi.batch_autotune(handle, cats, reps=1)
    try:
        assert all(c in (1, 2, 4) for c in t["chunks"]) and all(k in (1, 2, 4) for k in t["point_split"])
        assert 1 <= t["items_per_cu"] <= 4096
        assert capi.get_tile_table(handle) == t
        plan = capi.BatchPlan(handle, cats, store_flmn=True)
        plan.enqueue()
        plan.sync()
        torch.cuda.synchronize()
        _check_vs_fp64(cats, plan.outputs, handle)
        plan.close()
        # every (chunks, point split) choice in the populated bands gives the same result to the mixed tolerance
        for ch, ks in [(1, 1), (2, 4), (4, 2)]:
            capi.set_tile_table(handle, dict(chunks=[ch] * capi.TILE_BANDS, point_split=[ks] * capi.TILE_BANDS,
                                             items_per_cu=16))
            plan = capi.BatchPlan(handle, cats, store_flmn=True)
            plan.enqueue()
            plan.sync()
            torch.cuda.synchronize()
            _check_vs_fp64(cats, plan.outputs, handle)
            plan.close()
    finally:
        capi.set_tile_table(handle, None)
