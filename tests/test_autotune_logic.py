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


def test_grouped_entries_are_separate_and_applied():
    table = {
        "fp64:576:3336": {"target_blocks": 2048, "xchunk_bytes": 1 << 32, "ms": 9.0},
        "group:fp64:576:3336": {"target_blocks": 512, "ms": 30.0, "group": 4},
        "group:bf16x3:576:3336": {"target_blocks": 1024, "bf16_nbw_max": 6, "ms": 12.0, "group": 4},
    }
    assert autotune.lookup(600, 3300, "fp64", table)["target_blocks"] == 2048
    assert autotune.lookup(600, 3300, "fp64", table, grouped=True)["target_blocks"] == 512

    class FB:
        nbf, naux = 600, 3300

        def __init__(self):
            self.set = {}

        def set_tuning(self, k, v):
            self.set[k] = v

    fb = FB()
    autotune.apply(fb, "bf16x3", table, grouped=True)
    assert fb.set == {"target_blocks": 1024, "bf16_nbw_max": 6}
    fb = FB()
    assert autotune.apply(fb, "fp32", table, grouped=True) is None and fb.set == {}


def test_candidate_sets():
    single64 = list(autotune._candidates("fp64", grouped=False))
    assert len(single64) == 28 and all("bf16_nbw_max" not in c for c in single64)
    grouped_bf = list(autotune._candidates("bf16x3", grouped=True))
    assert len(grouped_bf) == 35 and all("xchunk_bytes" not in c and "kx_waves" not in c for c in grouped_bf)
    single_bf = list(autotune._candidates("bf16x3", grouped=False))
    # 7 block targets x 2 chunk sizes x 5 widths x 2 contraction tiles x
    # (4 or 8 waves of the 128-row half-transform, or the 256-row one)
    assert len(single_bf) == 420 and {c["kx_waves"] for c in single_bf} == {4, 8}
    assert {c["ks_tile"] for c in single_bf} == {128, 256}
    assert {c["kx_rows"] for c in single_bf} == {128, 256}
    assert not any(c["kx_rows"] == 256 and c["kx_waves"] == 4 for c in single_bf)
    N, Q, o = autotune.alkane_shapes([20])[0]
    assert (N, Q, o) == (1188, 2840, [81, 80])
