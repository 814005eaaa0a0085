"""Batched ensembles: each member steps bit-identically alone and in the
batch (same kernel variant), through graph replays interleaved with direct
steps and host reads."""
import numpy as np
import pytest

import coolbubble_amd as cb

pytestmark = pytest.mark.gpu


def members(precision):
    base = [
        dict(D=2, nGrids=64, L=8.0, c=4.0, readxns=1, nBubbles=5),
        dict(D=3, nGrids=32, L=4.0, c=2.0, readxns=-2),
        dict(D=3, nGrids=24, L=3.0, c=2.0, readxns=-3, Trad=0.3, expansion=1, Mpl=10.0),
        dict(D=1, nGrids=256, L=16.0, c=4.0),
        dict(D=2, nGrids=45, L=5.625, c=4.0, readxns=1, nBubbles=3),
        dict(D=3, nGrids=32, L=4.0, c=4.0, readxns=1, nBubbles=4, rand_seed=7),
        dict(D=2, nGrids=128, L=16.0, c=6.0, readxns=-1, DeltaR=0.5),
    ]
    # no three-step tile of 256 threads fits a 256-wide FP64 row: in FP64
    # this member takes two steps per launch beside three-step ones
    base.append(dict(D=3, nGrids=256, L=32.0, c=4.0, readxns=-2))
    for m in base:
        m["dt"] = 0.2 * m["L"] / m["nGrids"]
        m["precision"] = precision
    return base


@pytest.mark.parametrize("two_step", [0, 1])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_members_bit_identical_alone_and_batched(precision, two_step):
    """two_step: static members take two or three leapfrog steps per batched
    launch (the steps left over on their one-step companion), as a lone run
    with the same variant does; the chunks leave 0, 1 and 2 steps over."""
    ms = members(precision)
    print("batch", precision, len(ms), "two_step", two_step, flush=True)
    b = cb.Batch(ms)
    b.set_shape((2, 1, 1, 32, two_step))
    if two_step:
        fused = [v.variant_info(v.kernel_variants()[1])["fuse"] for v in b.members if v.kernel_variants()[1] >= 0]
        assert 3 in fused
        if precision == "fp64":
            assert 2 in fused
    alone = []
    for m, view in zip(ms, b.members):
        s = cb.Bubble(**m)
        s.select_kernel(view.kernel_variants()[1])
        alone.append(s)
    # graph segments, direct steps, host reads in between
    for chunk in (40, 3, 40, 40, 17, 1):
        b.step(chunk)
        for s in alone:
            s.step(chunk)
        for s, view in zip(alone, b.members):
            np.testing.assert_array_equal(view.energy_row(), s.energy_row())
    for s, view in zip(alone, b.members):
        pa, qa = s.fields()
        pb, qb = view.fields()
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(qa, qb)
        assert view.state().step == s.state().step == 141
    b.close()


def test_three_step_members_only():
    """A batch whose multi-step members all take three steps per launch runs
    the kernel that holds only the three-step tile: bit-identical to the
    lone runs, through a graph replay and chunks leaving 1 and 2 steps."""
    ms = [dict(D=3, nGrids=192, L=24.0, c=4.0, readxns=-2, rand_seed=3),
          dict(D=3, nGrids=256, L=32.0, c=2.0, readxns=1, nBubbles=3, rand_seed=5),
          dict(D=2, nGrids=2048, L=128.0, c=4.0, readxns=-1)]
    for m in ms:
        m["dt"] = 0.2 * m["L"] / m["nGrids"]
        m["precision"] = "fp32"
    print("batch three-step only", len(ms), flush=True)
    b = cb.Batch(ms)
    b.set_shape((2, 1, 1, 32, 1))
    assert [v.variant_info(v.kernel_variants()[1])["fuse"] for v in b.members] == [3, 3, 3]
    alone = []
    for m, view in zip(ms, b.members):
        s = cb.Bubble(**m)
        s.select_kernel(view.kernel_variants()[1])
        alone.append(s)
    for chunk in (40, 40, 7, 5, 1):
        b.step(chunk)
        for s in alone:
            s.step(chunk)
        for s, view in zip(alone, b.members):
            np.testing.assert_array_equal(view.energy_row(), s.energy_row())
    for s, view in zip(alone, b.members):
        pa, qa = s.fields()
        pb, qb = view.fields()
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(qa, qb)
        s.close()
    b.close()


def test_batch_member_cannot_be_destroyed_alone():
    b = cb.Batch(members("fp64")[:2])
    assert cb.lib().cb_destroy(b.members[0].handle) == 1
    b.close()
