"""Host-side checks of the frequency sweep: argument validation and the autotuner cache (no GPU)."""

import json

import numpy as np
import pytest


@pytest.fixture(scope="module")
def spaces():
    from bempp_amd.api import function_space
    from bempp_amd.api.grid.shapes import icosphere

    g = icosphere(1)
    return function_space(g, "DP", 0), function_space(g, "P", 1)


def test_rejects_bad_wavenumbers_and_spaces(spaces):
    from bempp_amd.core.frequency_sweep import hip_kbatch_assemble

    dp0, p1 = spaces
    with pytest.raises(ValueError):
        hip_kbatch_assemble(dp0, dp0, np.linspace(1, 2, 65))
    with pytest.raises(ValueError):
        hip_kbatch_assemble(dp0, dp0, [])
    with pytest.raises(ValueError):
        hip_kbatch_assemble(dp0, dp0, [1.0 + 0.5j])
    with pytest.raises(ValueError):
        hip_kbatch_assemble(dp0, dp0, [0.0, 1.0])
    with pytest.raises(ValueError):
        hip_kbatch_assemble(p1, p1, [1.0])
    with pytest.raises(ValueError):
        hip_kbatch_assemble(dp0, dp0, [1.0], precision="half")


def test_api_rejects_other_device_interfaces(spaces):
    from bempp_amd.api.operators.boundary.helmholtz import single_layer_frequency_sweep

    dp0, _ = spaces
    with pytest.raises(ValueError):
        single_layer_frequency_sweep(dp0, dp0, dp0, [1.0], device_interface="opencl")
    with pytest.raises(ValueError):
        single_layer_frequency_sweep(dp0, dp0, dp0, [0.0])


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    from bempp_amd.core import autotune

    cache = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("BEMPP_AMD_TUNE_CACHE", str(cache))
    autotune.clear_memory()
    key = autotune.make_key("helmholtz_single_layer", "p0", "p0", 5120, 5120, 32, "single", "dev")
    assert autotune.lookup(key) is None
    autotune._store(key, {"config": [32, 2], "ms": 1.0, "all": {}})
    assert json.loads(cache.read_text())[key]["config"] == [32, 2]
    assert autotune.lookup(key) == (32, 2)
    cache.write_text("not json")
    autotune.clear_memory()
    assert autotune.lookup(key) is None
    autotune.clear_memory()


def test_autotune_candidates_fit_the_abi():
    from bempp_amd.core import autotune

    assert set(autotune.CANDIDATE_TILES) <= {8, 16, 32}
    assert all(1 <= w <= 16 for w in autotune.CANDIDATE_WG_PER_CU)
    assert autotune.DEFAULT_CONFIG[0] in autotune.CANDIDATE_TILES


@pytest.mark.parametrize("entry", [
    {"config": [12, 4]},          # tile outside the candidates
    {"config": [16, 3]},          # workgroups per CU outside the candidates
    {"config": [16]},             # wrong length
    {"config": "16,4"},           # wrong type
    {"config": [16.5, 4]},        # not an integer
    {"config": [True, 4]},        # bool is not a tile size
    {"ms": 1.0},                  # no config
    [16, 4],                      # not a mapping
])
def test_autotune_treats_invalid_entries_as_a_miss(tmp_path, monkeypatch, entry):
    from bempp_amd.core import autotune

    cache = tmp_path / "tune.json"
    monkeypatch.setenv("BEMPP_AMD_TUNE_CACHE", str(cache))
    autotune.clear_memory()
    key = autotune.make_key("helmholtz_single_layer", "p0", "p0", 320, 320, 4, "single", "dev")
    cache.write_text(json.dumps({key: entry}))
    assert autotune.lookup(key) is None
    cache.write_text(json.dumps({key: {"config": [8, 2]}}))
    assert autotune.lookup(key) == (8, 2)
    autotune.clear_memory()
