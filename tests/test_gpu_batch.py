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
    for m in base:
        m["dt"] = 0.2 * m["L"] / m["nGrids"]
        m["precision"] = precision
    return base


@pytest.mark.parametrize("two_step", [0, 1])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_members_bit_identical_alone_and_batched(precision, two_step):
    """two_step: static members take two leapfrog steps per batched launch
    (odd remainders on their one-step companion), as a lone run with the
    same two-step variant does."""
    ms = members(precision)
    print("batch", precision, len(ms), "two_step", two_step, flush=True)
    b = cb.Batch(ms)
    b.set_shape((2, 1, 1, 32, two_step))
    if two_step:
        fused = [v.variant_info(v.kernel_variants()[1])["fuse"] for v in b.members if v.kernel_variants()[1] >= 0]
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


def test_batch_member_cannot_be_destroyed_alone():
    b = cb.Batch(members("fp64")[:2])
    assert cb.lib().cb_destroy(b.members[0].handle) == 1
    b.close()
