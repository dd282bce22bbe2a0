"""The autotuner's choice and cache (CPU: the timing functions are replaced
by fixed tables, so nothing runs on a GPU): the fastest candidate is picked,
cached per case shape, storage format and closure, kept in the JSON file,
and reused without timing again until asked to retune."""
import json
from types import SimpleNamespace

import pytest

from water_flow_simulation_amd import autotune


@pytest.fixture
def tuner(tmp_path, monkeypatch):
    monkeypatch.setenv("WFS_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", None)
    calls = []

    def widths(reach, storage="f32", closure=None):
        calls.append(("w", reach.nx, storage, closure is not None))
        base = {64: 0.50, 128: 0.47, 192: 0.45, 256: 0.46}
        if closure is not None:
            base = {64: 0.44, 128: 0.45, 192: 0.46, 256: 0.46}
        return dict(base)

    def staggers(reach, storage="f32"):
        calls.append(("s", reach.nx, storage))
        return {0: 0.345, 512: 0.341, 1024: 0.347, 1536: 0.339}

    monkeypatch.setattr(autotune, "time_widths", widths)
    monkeypatch.setattr(autotune, "time_staggers", staggers)
    return SimpleNamespace(calls=calls, path=tmp_path / "tune.json")


def _reach(nx=1500, ny=522, nz=30):
    return SimpleNamespace(nx=nx, ny=ny, nz=nz)


def test_fastest_width_per_closure_is_cached(tuner):
    r = _reach()
    assert autotune.block_x_for(r, "f16dev") == 192
    assert autotune.block_x_for(r, "f16dev", closure=[0.0]) == 64
    assert autotune.block_x_for(r, "f16dev") == 192            # cached: no new timing
    assert len(tuner.calls) == 2
    saved = json.loads(tuner.path.read_text())
    assert saved["f16dev/smag/1500x522x30"]["block_x"] == 192
    assert saved["f16dev/mlp/1500x522x30"]["block_x"] == 64
    autotune.block_x_for(r, "f16dev", retune=True)
    assert len(tuner.calls) == 3


def test_fastest_stagger_per_format_and_shape(tuner, monkeypatch):
    r = _reach()
    assert autotune.pop_stagger_for(r, "f16dev") == 1536
    assert autotune.pop_stagger_for(r, "f16dev") == 1536
    assert autotune.pop_stagger_for(r, "f32") == 1536
    assert autotune.pop_stagger_for(_reach(nx=1000), "f16dev") == 1536
    assert [c[0] for c in tuner.calls] == ["s", "s", "s"]    # one timing per (format, shape)
    saved = json.loads(tuner.path.read_text())
    assert {k for k in saved if k.startswith("stagger/")} == {
        "stagger/f16dev/smag/1500x522x30", "stagger/f32/smag/1500x522x30", "stagger/f16dev/smag/1000x522x30"}
    # a fresh process reads the file instead of timing
    monkeypatch.setattr(autotune, "_cache", None)
    assert autotune.pop_stagger_for(r, "f16dev") == 1536
    assert len(tuner.calls) == 3


def test_unwritable_cache_still_serves_the_process(tmp_path, monkeypatch):
    (tmp_path / "file").write_text("")                       # a file where the cache folder would go
    monkeypatch.setenv("WFS_AUTOTUNE_CACHE", str(tmp_path / "file" / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", None)
    monkeypatch.setattr(autotune, "time_staggers", lambda reach, storage="f32": {0: 1.0, 512: 0.5})
    assert autotune.pop_stagger_for(_reach(), "f32") == 512
    monkeypatch.setattr(autotune, "time_staggers", lambda reach, storage="f32": pytest.fail("timed again"))
    assert autotune.pop_stagger_for(_reach(), "f32") == 512
