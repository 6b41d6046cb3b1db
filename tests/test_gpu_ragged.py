"""Config 4: ragged mixed-width modexp, persistent + hipGraph vs plain, bit-exact."""

import random

import pytest

from oracle import gmp_engine as G

pytestmark = pytest.mark.gpu


def _powm():
    return G.powm if G.available() else pow


@pytest.fixture(scope="module")
def workload(ctx):
    from cuhepy_amd.ragged import DeviceWorkload, make_subbatches
    subs = make_subbatches(n_subs=24, max_size=300, seed=4040)
    subs += make_subbatches(n_subs=2, max_size=1, seed=4041)  # single-item sub-batches
    return DeviceWorkload(ctx, subs)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_ragged_persistent_bit_exact(ctx, workload, variant):
    from cuhepy_amd.engine import RaggedPlan
    plan = RaggedPlan(ctx, workload.desc, variant=variant, window=0, waves_per_cu=2)
    workload.clear_outputs()
    plan.run(graph=False)
    ctx.sync()
    assert workload.spot_check(6, random.Random(variant), _powm()) == []
    plan.close()


def test_ragged_graph_replays_match_eager(ctx, workload):
    """Replay the captured graph several times, with eager work and
    synchronisation between replays; every replay must equal the eager result."""
    import numpy as np
    from cuhepy_amd.engine import RaggedPlan
    plan = RaggedPlan(ctx, workload.desc, variant=1, window=4, waves_per_cu=2)
    workload.clear_outputs()
    plan.run(graph=False)
    ctx.sync()
    eager = [workload.result(i).copy() for i in range(len(workload.subs))]
    assert workload.spot_check(4, random.Random(1), _powm()) == []
    for rep in range(3):
        workload.clear_outputs()       # eager work between replays
        ctx.sync()
        plan.run(graph=True)
        ctx.sync()
        for i in range(len(workload.subs)):
            assert np.array_equal(workload.result(i), eager[i]), f"replay {rep} sub {i}"
    plan.close()


def test_ragged_plain_matches(ctx, workload):
    import numpy as np
    from cuhepy_amd.engine import RaggedPlan
    plan = RaggedPlan(ctx, workload.desc, variant=0, window=5, waves_per_cu=4)
    workload.clear_outputs()
    plan.run(graph=True)
    ctx.sync()
    first = [workload.result(i).copy() for i in range(len(workload.subs))]
    workload.clear_outputs()
    plan.run_plain()
    ctx.sync()
    for i in range(len(workload.subs)):
        assert np.array_equal(workload.result(i), first[i])
    plan.close()


def test_autotune_picks_a_valid_shape(ctx, workload):
    from cuhepy_amd.autotune import tune_ragged
    best, log = tune_ragged(ctx, workload, "test-device", variants=(0, 2), windows=(4, 5), waves=(2,), reps=1,
                            force=True)
    assert best["variant"] in (0, 2) and best["window"] in (4, 5)
    assert len(log) >= 3
