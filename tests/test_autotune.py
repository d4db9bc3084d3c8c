"""Host-side behaviour of the autotune table (no GPU): bucketing, lookup,
hand-set entries, JSON save / load round trip, malformed files, and the
SGEMM_AMD_TUNE_TABLE environment hook.  Timing candidates needs a GPU and is
covered in tests/test_grouped_gpu.py."""
import os
import subprocess
import sys
import tempfile

import pytest

import sgemm_amd as sa


@pytest.fixture(autouse=True)
def clean_table():
    if not os.path.exists(sa.LIB_PATH):
        sa.build()
    sa.autotune_clear()
    yield
    sa.autotune_clear()


def test_default_heuristic_is_deterministic_and_valid():
    for form in (sa.FORM_F32, sa.FORM_BF16X3):
        n_cfg = len(sa.configs(form))
        for shape in [(1, 1, 1), (32, 2048, 77), (1024, 1024, 1024), (8192, 8192, 8192)]:
            a = sa.autotune_lookup(form, *shape)
            assert a == sa.autotune_lookup(form, *shape)
            assert 0 <= a[0] < n_cfg and a[1] is False
            g = sa.autotune_lookup(form, *shape, grouped=True)
            assert g[1] is False and g[0] in ((1, 2, 5) if form == sa.FORM_F32 else (1, 8))


def test_set_applies_to_whole_bucket_only():
    sa.autotune_set(sa.FORM_F32, 1000, 1000, 1000, 1)
    assert sa.autotune_lookup(sa.FORM_F32, 513, 600, 1024) == (1, True)     # same bucket
    assert sa.autotune_lookup(sa.FORM_F32, 1025, 600, 1024)[1] is False      # next m bucket
    assert sa.autotune_lookup(sa.FORM_F32, 1000, 1000, 1000, grouped=True)[1] is False
    assert sa.autotune_lookup(sa.FORM_BF16X3, 1000, 1000, 1000)[1] is False


def test_invalid_entries_rejected():
    with pytest.raises(sa.SgemmError):
        sa.autotune_set(sa.FORM_F32, 10, 10, 10, 99)
    with pytest.raises(sa.SgemmError):
        sa.autotune_set(sa.FORM_F32, 10, 10, 10, 3, grouped=True)  # not in the grouped kernel
    with pytest.raises(sa.SgemmError):
        sa.autotune_set(7, 10, 10, 10, 0)


def test_save_load_roundtrip_and_bad_file():
    sa.autotune_set(sa.FORM_F32, 64, 64, 64, 1)
    sa.autotune_set(sa.FORM_BF16X3, 4096, 4096, 4096, 8, grouped=True)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "t.json")
        sa.autotune_save(path)
        import json
        data = json.load(open(path))
        assert len(data["entries"]) == 2
        sa.autotune_clear()
        assert sa.autotune_size() == 0
        sa.autotune_load(path)
        assert sa.autotune_size() == 2
        assert sa.autotune_lookup(sa.FORM_F32, 60, 50, 40) == (1, True)
        assert sa.autotune_lookup(sa.FORM_BF16X3, 4000, 3000, 2100, grouped=True) == (8, True)
        bad = os.path.join(d, "bad.json")
        with open(bad, "w") as f:
            f.write('{"entries": [\n  {"form": 9, "grouped": 0}\n]}\n')
        with pytest.raises(sa.SgemmError) as e:
            sa.autotune_load(bad)
        assert e.value.status == 8
        assert sa.autotune_size() == 2  # a failed load keeps the old table
        with pytest.raises(sa.SgemmError):
            sa.autotune_load(os.path.join(d, "missing.json"))


def test_environment_table_loaded_on_first_use():
    sa.autotune_set(sa.FORM_F32, 300, 300, 300, 2)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "env.json")
        sa.autotune_save(path)
        code = ("import sys; sys.path.insert(0, %r); import sgemm_amd as sa; "
                "print(sa.autotune_lookup(0, 300, 300, 300))" % os.path.dirname(sa.ROOT + "/x"))
        env = dict(os.environ, SGEMM_AMD_TUNE_TABLE=path)
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                             text=True, check=True).stdout.strip()
        assert out == "(2, True)"


def test_launch_config_follows_operand_alignment():
    """sgemm_launch_config reports what the default entry point launches
    (pointers are not read).  FP32 with 16-byte aligned columns short of a
    round of 1024 direct-loaded 128x128 tiles picks by the fill model: at
    3584^3 (784 tiles, 3.06 per CU, 12.25 64x64 tiles per CU) the 64x64
    tile (config 7); at 3968^3 (961 tiles, 3.75 per CU) config 5.  With odd
    leading dimensions the register-staged tile (768 resident) takes the
    tail split (config 9) from 700 tiles.  Whole rounds of 256x256 tiles with
    aligned columns and k a multiple of 32 take config 15, partial rounds
    config 16 where the time model says its stream-K tail pays.  k == 0 and leading dimensions
    past 2^21 run no tile config (-1)."""
    F = sa.FORM_F32
    base = 1 << 20                                # 16-byte aligned fake addresses
    m = n = 3584                                  # 784 tiles of 128x128
    k = 3584
    aligned = sa.launch_config(F, m, n, k, base, m, base, k)
    assert aligned == sa.autotune_lookup(F, m, n, k)[0] == 7
    assert sa.launch_config(F, 3968, 3968, 3968, base, 3968, base, 3968) == 5
    assert sa.launch_config(F, m, n, k, base, m + 1, base, k) == 9      # odd lda
    assert sa.launch_config(F, m, n, k, base + 4, m, base, k) == 9      # A 4 bytes off
    # whole rounds of 256x256 tiles with k % 32 == 0: config 15; odd
    # columns: config 9; a partial last round (6144^3: 2.25 rounds) that
    # stream-K spreads cheaply: config 16
    assert sa.launch_config(F, 8192, 8192, 8192, base, 8192, base, 8192) == 15
    assert sa.launch_config(F, 8192, 8192, 8192, base, 8193, base, 8192) == 9
    assert sa.launch_config(F, 6144, 6144, 6144, base, 6144, base, 6144) == 16
    # a partial round that stream-K spreads cheaply: config 16 (13312^3: 144
    # tiles past ten rounds); a few tiles cut into many pieces: config 9
    assert sa.launch_config(F, 13312, 13312, 13312, base, 13312, base, 13312) == 16
    assert sa.launch_config(F, 7168, 7168, 7168, base, 7168, base, 7168) == 9
    # short of a round of 128x128 tiles with long k: stream-K over 256x256
    # tiles where the model puts it ahead of the 64x64 / 128x128 fill model
    assert sa.launch_config(F, 3072, 3072, 8192, base, 3072, base, 8192) == 16
    assert sa.launch_config(F, 8192, 8192, 127, base, 8193, base, 127) == 14
    assert sa.launch_config(F, m, n, 0, base, m, base, 1) == -1
    assert sa.launch_config(F, m, n, k, base, (1 << 21) + 4, base, k) == -1
    for form in (sa.FORM_F32, sa.FORM_BF16X3):                          # table entries win
        sa.autotune_set(form, m, n, k, 2)
        assert sa.launch_config(form, m, n, k, base, m + 1, base, k) == 2
    with pytest.raises(sa.SgemmError):
        sa.launch_config(F, -1, n, k, base, m, base, k)
