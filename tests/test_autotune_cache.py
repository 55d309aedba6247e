"""The autotuner's cache file (host logic, no GPU): lines of key and shape
are read back when a path is set; malformed lines are skipped."""


def test_cache_file_is_read(tmp_path):
    from bml_amd import hip
    p = tmp_path / "tune.txt"
    p.write_text("cu256:p2:f1:e0:n44:a28:b28:t40 16 512 4 0 1 0 0 1 0.123000\n"
                 "cu256:p1:f1:e2:n60:a47:b47:t55 64 3000 2 0 8 2 10 4 86.000000\n"
                 "not a shape\n")
    hip.autotune_clear()
    hip.set_autotune(False, str(p))
    try:
        assert not hip.get_autotune()
        assert hip.autotune_entries() == 2
        hip.autotune_clear()
        assert hip.autotune_entries() == 0
    finally:
        hip.autotune_clear()
