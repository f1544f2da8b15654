"""Autotuner table logic (no GPU)."""

from forte2_amd import autotune


def test_lookup_picks_nearest_size_class():
    table = {
        "fp64:24:139": {"target_blocks": 256, "xchunk_bytes": 1 << 30, "ms": 0.1},
        "fp64:384:2224": {"target_blocks": 1024, "xchunk_bytes": 1 << 31, "ms": 3.0},
        "fp64:1536:8896": {"target_blocks": 2048, "xchunk_bytes": 1 << 32, "ms": 466.0},
        "bf16x3:1536:8896": {"target_blocks": 4096, "xchunk_bytes": 1 << 32, "ms": 110.0},
    }
    assert autotune.lookup(48, 278, "fp64", table)["target_blocks"] == 256
    assert autotune.lookup(400, 2300, "fp64", table)["target_blocks"] == 1024
    assert autotune.lookup(1500, 9000, "bf16x3", table)["target_blocks"] == 4096
    assert autotune.lookup(1500, 9000, "fp32", table) is None


def test_save_load_roundtrip(tmp_path):
    t = {"fp64:24:139": {"target_blocks": 512, "xchunk_bytes": 1 << 29, "ms": 0.05}}
    p = tmp_path / "t.json"
    autotune.save(t, str(p))
    assert autotune.load(str(p)) == t
