"""The persistent schedule's tuning and its hipGraph replay, on a reduced batch of uneven
jobs (the 256-job workload's first 12 jobs at a quarter of their size):

- results do not depend on the tile shape (4, 8, 16 or 32 pixels wide per wave64 tile) or the
  refill period (march steps between refills): bitwise equal outputs;
- the captured passes replayed several times, with other eager work in between, equal the
  pass-by-pass launches bitwise every time, and still report per-pass times;
- the autotuner's choice gives the same outputs and is no slower than the default.
"""
import time

import numpy as np
import pytest

import numericdream_amd as nd
from numericdream_amd import abi

pytestmark = pytest.mark.gpu

FIELDS = ("depth", "normal", "position", "trap", "hit", "shadow", "evals")


def small_jobs(n=12, div=4):
    jobs = []
    for i in range(n):
        u = nd.config_uniforms(4, i)
        s = max(16, int(u.resolution[0]) // div)
        nd.resize(u, s, s)
        jobs.append((u, nd.GBuffer(s, s)))
    return jobs


def outputs(jobs):
    return [g.numpy() for _, g in jobs]


def same(a, b):
    for x, y in zip(a, b):
        for k in FIELDS:
            assert np.array_equal(x[k], y[k]), k


@pytest.fixture
def tuned(renderer):
    def apply(**kw):
        renderer.set_tuning(**kw)
        return renderer
    yield apply
    renderer.set_tuning()


@pytest.mark.parametrize("kernel", [abi.KERNEL_PERSISTENT, abi.KERNEL_PLAIN])
def test_tile_shape_and_refill_period_do_not_change_results(tuned, kernel):
    jobs = small_jobs()
    r = tuned()
    r.render_batch(jobs, shadows=True, kernel=kernel)
    ref = outputs(jobs)
    for tw in (4, 16, 32):
        for period in (1, 3):
            r = tuned(tile_width=tw, refill_period=period)
            jobs2 = [(u, nd.GBuffer(g.w, g.h)) for u, g in jobs]
            st = r.render_batch(jobs2, shadows=True, kernel=kernel)
            assert st.rays == sum(g.w * g.h for _, g in jobs)
            same(ref, outputs(jobs2))


def test_graph_replays_equal_pass_by_pass_launches(tuned):
    import torch
    jobs = small_jobs()
    other = small_jobs(6, 8)
    r = tuned()  # pass by pass (the default)
    st0 = r.render_batch(jobs, shadows=True)
    assert st0.graph == 0
    ref = outputs(jobs)
    r = tuned(gbuffer_graphs=True)
    replays = 0
    for rep in range(4):
        for _, g in jobs:
            for k in FIELDS:
                getattr(g, k)[...] = 0
        st = r.render_batch(jobs, shadows=True)
        replays += st.graph
        same(ref, outputs(jobs))
        assert st.de_evals == st0.de_evals and st.hits == st0.hits
        assert st.ms_march > 0 and st.ms_surface >= 0 and st.ms_total >= st.ms_march
        # eager work between replays: another batch pass by pass, and device work of torch
        r.set_tuning()
        r.render_batch(other, shadows=False)
        r.set_tuning(gbuffer_graphs=True)
        x = torch.ones(1 << 20, device="cuda:0")
        torch.cuda.synchronize()
        assert float(x.sum()) == float(1 << 20)
    print("\n[graph] %d of 4 calls replayed a captured graph" % replays, flush=True)
    assert replays >= 2  # the first call of a shape runs pass by pass, the second captures, later ones replay


def test_autotune_same_outputs_and_not_slower(tuned):
    jobs = small_jobs(16, 2)
    r = tuned()
    r.render_batch(jobs, shadows=True)
    ref = outputs(jobs)

    def median_ms(reps=7):
        ms = []
        for _ in range(reps):
            ms.append(r.render_batch(jobs, shadows=True).ms_total)
        return float(np.median(ms))

    default_ms = median_ms()
    t = r.autotune(jobs, shadows=True)
    print("\n[tune] waves/CU %d refill %d chunk %d tile %d period %d" % (t.waves_per_cu, t.refill_min_idle,
                                                                           t.chunk_rays, t.tile_width,
                                                                           t.refill_period), flush=True)
    r.render_batch(jobs, shadows=True)
    same(ref, outputs(jobs))
    tuned_ms = median_ms()
    print("[tune] default %.3f ms, tuned %.3f ms" % (default_ms, tuned_ms), flush=True)
    assert tuned_ms <= default_ms * 1.10
