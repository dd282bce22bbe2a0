"""The batched sweep (one launch per step over all cases, device cadence,
HIP-graph replay) against the same cases run one after another."""
import numpy as np
import pytest

from tests.helpers import channel128
from water_flow_simulation_amd import reach3d as R
from water_flow_simulation_amd.batch import BatchRun
from water_flow_simulation_amd.sweep import ensemble

pytestmark = pytest.mark.gpu


def _cases():
    small, _ = ensemble(3, seed=1, scale=0.2)
    from tests.test_builder import CASES
    nowall = R.build_reach(*CASES["nowall_longtail"]())        # no wall model, no uref
    return small + [channel128(), nowall]


def _state(run):
    sums, navg = run.lbm.sums()
    return sums, navg, run.lbm.macro(), run.lbm.populations()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_sequential_bitwise(use_graph):
    reaches = _cases()
    pattern = (100, 35, 40)              # graph chunks, an odd remainder (parity flip), a re-captured graph
    seq = []
    for r in reaches:
        run = R.Run3D(r)
        for m in pattern:
            run.advance(m)
        seq.append(run)
    b = BatchRun(reaches)
    for m in pattern:
        b.advance(m, use_graph=use_graph)
        # eager readout between graph replays
        _ = b.runs[0].lbm.macro()
    for k, (rs, rb) in enumerate(zip(seq, b.runs)):
        s_sums, s_navg, s_mac, s_pop = _state(rs)
        b_sums, b_navg, b_mac, b_pop = _state(rb)
        assert s_navg == b_navg and rs.t == rb.t, k
        fl = ~reaches[k].solid.ravel()
        for name in s_sums:
            assert np.array_equal(s_sums[name][fl], b_sums[name][fl]), (k, name)
        for a, c in zip(s_mac, b_mac):
            assert np.array_equal(a[fl], c[fl]), k
        m19 = np.tile(fl, 19)
        assert np.array_equal(s_pop[m19], b_pop[m19]), k
        assert rb.lbm.nstep == sum(pattern)
    for run in seq:
        run.close()
    b.close()
