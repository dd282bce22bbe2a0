"""The two device measurements the benchmark reports beside its roofline:
the plain copy bandwidth and the streaming ceiling of the step's
19-population pull pattern. Both must run, report a plausible rate and
refuse arguments they cannot honour."""
import pytest

from water_flow_simulation_amd import _native

pytestmark = pytest.mark.gpu


def test_copy_bandwidth_runs():
    print("case: copy probe, 64 MiB x 5", flush=True)
    gbs = _native.bandwidth_probe(64 << 20, 5)
    assert 50.0 < gbs < 20000.0


@pytest.mark.parametrize("elem", [4, 2])
@pytest.mark.parametrize("stagger", [0, 4096])
def test_stream_ceiling_runs(elem, stagger):
    print(f"case: 19-stream probe 128x64x32, {elem}-byte populations, stagger {stagger} B", flush=True)
    gbs = _native.stream_probe(128, 64, 32, elem, stagger, 5)
    assert 50.0 < gbs < 20000.0


@pytest.mark.parametrize("args", [
    (32, 64, 32, 4, 0),        # rows narrower than one wavefront
    (128, 2, 32, 4, 0),        # no interior in y
    (128, 64, 32, 8, 0),       # element size not 2 or 4
    (128, 64, 32, 4, 6),       # stagger not a whole element
    (128, 64, 32, 4, -4),      # negative stagger
])
def test_stream_ceiling_refuses_bad_arguments(args):
    with pytest.raises(ValueError):
        _native.stream_probe(*args, iters=1)
