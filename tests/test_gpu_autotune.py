"""Autotuner: searches the launch knobs of a plan and leaves it set to a candidate whose
transforms still match the oracle (every knob selects among parity-tested kernels)."""
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "bf16x6"])
def test_autotune_picks_a_valid_setting(gpu, prec, tmp_path, monkeypatch):
    import torch
    import dsht
    import oracle
    from dsht import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tuning.json"))
    N, I, J = 42, 144, 72
    p = dsht.Plan(N, I, J, radius=6.371e6)
    p.set_precision({v: k for k, v in dsht.PREC_NAMES.items()}[prec])
    entry = autotune.tune(p, 96, reps=1)
    cands = autotune.candidates(p, 96)
    assert {k: entry[k] for k in ("chunk", "cb", "leg")} in cands
    assert len(entry["candidates"]) == len(cands)
    for name, key in autotune.KNOBS:
        assert p.get_tuning(key) == entry[name]
    # the cached entry is reused and re-applied
    p.set_tuning(dsht.TUNE_LEG_CB, 0)
    again = autotune.tune(p, 96)
    assert again["ms"] == entry["ms"] and p.get_tuning(dsht.TUNE_LEG_CB) == entry["cb"]
    o = oracle.Oracle(N, I, J, 6.371e6)
    w = red_spectrum(N, 5, 2)
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert rel_l2(p._inv(w, dsht.GRADLAT), o.inv(w, oracle.GRADLAT)) < tol
