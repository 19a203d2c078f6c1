"""Batched form and autotuner: every image of a batch, solved inside one graph on several
streams, is bit-identical to the oracle (f64 norm mode), on every replay; a tuned plan
too, whatever strip height and projection kernel the tuner picked."""
import numpy as np
import pytest

import artefact_amd
import oracle
from artefact_amd import synth

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_batch_matches_oracle_on_every_replay():
    imgs = synth.batch_workload(7, seed=2, min_side=32, max_side=160, iterations=6)
    refs = [oracle.compute(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                           norm_mode="f64") for w in imgs]
    print("case batch of 7, 3 streams", flush=True)
    with artefact_amd.Batch([artefact_amd.BatchImage(w["coefs"], w["weight"], w["pweight"], w["iterations"],
                                                     w["max_w"], w["max_h"], w["mode"]) for w in imgs],
                            streams=3, tune=True) as batch:
        for rep in range(3):
            batch.run()
            for i, w in enumerate(imgs):
                out = batch.download(i)
                for c, (a, b) in enumerate(zip(out, refs[i])):
                    assert bits_equal(a, b), f"replay {rep} image {w['name']} channel {c}"


def test_autotuned_stream_count_matches_oracle():
    """streams = 0 with tuning: the graph is timed with 1 to 4 streams and the fastest kept;
    the images still come out bit-exact."""
    imgs = synth.batch_workload(6, seed=6, min_side=32, max_side=128, iterations=4)
    print("case batch of 6, autotuned streams", flush=True)
    with artefact_amd.Batch([artefact_amd.BatchImage(w["coefs"], w["weight"], w["pweight"], w["iterations"],
                                                     w["max_w"], w["max_h"], w["mode"]) for w in imgs],
                            streams=0, tune=True) as batch:
        assert 1 <= batch.streams <= 4
        for rep in range(2):
            batch.run()
            for i, w in enumerate(imgs):
                ref = oracle.compute(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                                     norm_mode="f64")
                for a, b in zip(batch.download(i), ref):
                    assert bits_equal(a, b), f"replay {rep} {w['name']}"


def test_single_stream_batch_matches_oracle():
    imgs = synth.batch_workload(3, seed=4, min_side=48, max_side=96, iterations=3)
    print("case batch of 3, 1 stream", flush=True)
    with artefact_amd.Batch([artefact_amd.BatchImage(w["coefs"], w["weight"], w["pweight"], w["iterations"],
                                                     w["max_w"], w["max_h"], w["mode"]) for w in imgs],
                            streams=1, tune=False) as batch:
        batch.run()
        for i, w in enumerate(imgs):
            ref = oracle.compute(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                                 norm_mode="f64")
            for a, b in zip(batch.download(i), ref):
                assert bits_equal(a, b), w["name"]


@pytest.mark.parametrize("name", ["tgv_420_208x144", "tgv_444_72x40", "tv_luma_64x64"])
def test_tuned_plan_bit_exact(name):
    wl = synth.workload(name, seed=9)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case tune {name}", flush=True)
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], 7, W, H, mode=wl["mode"]) as plan:
        t = plan.tune()
        assert t["kg_rows"] >= 2 and t["ms_per_iteration"] > 0
        plan.init()
        plan.iterate()
        out = plan.download()
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], 7, W, H, norm_mode="f64")
    for a, b in zip(out, ref):
        assert bits_equal(a, b)
