"""Host logic of the autotuner (aok.autotune): the JSON table, signatures and applying a
record to a context.  The timing search itself runs on the GPU (test_gpu_autotune.py)."""
import json

import torch

from aok import autotune


class _Ctx:
    def __init__(self):
        self.knobs = {}

    def tune(self, k, v):
        self.knobs[k] = v


def test_table_roundtrip_and_apply(tmp_path):
    path = str(tmp_path / "sub" / "t.json")
    t = autotune.TuneTable(path)
    assert t.get("x") is None and not t.apply(_Ctx(), "x")
    t.put("x", {"ovl_cap": 56, "ovl_split": 2}, 1.25, 10)
    t.save()
    with open(path) as f:
        raw = json.load(f)
    assert raw == {"x": {"knobs": {"ovl_cap": 56, "ovl_split": 2}, "ms": 1.25, "candidates": 10}}
    t2 = autotune.TuneTable(path)
    c = _Ctx()
    assert t2.apply(c, "x") and c.knobs == {"ovl_cap": 56, "ovl_split": 2}


def test_signatures_bucket_batch_sizes(monkeypatch):
    monkeypatch.setattr(autotune, "_device_name", lambda d: "gfx950")
    s1 = autotune.overlap_signature([0, 1, 2], 9, 3, 65536, torch.float32, None)
    s2 = autotune.overlap_signature([0, 1, 2], 9, 3, 60000, torch.float32, None)
    s3 = autotune.overlap_signature([0, 1, 2], 9, 3, 65536, torch.float64, None)
    assert s1 == s2 == "ovl2c|float32|ls=0,1,2|K=9|NZ=3|B~2^16|gfx950"
    assert s3 != s1
    a = autotune.ao_signature(6, 4, 3, 30, 2, 2 ** 24, torch.float32, False, None)
    assert a == "ao|float32|nRad=6|K=4|NZ=3|nOrb=30|L=2|grad=0|nX~2^24|gfx950"


def test_env_path(monkeypatch, tmp_path):
    monkeypatch.setenv("AOK_AUTOTUNE_PATH", str(tmp_path / "a.json"))
    assert autotune.default_path() == str(tmp_path / "a.json")
