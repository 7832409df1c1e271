"""The autotuner's command line on the CPU: its defaults are the full grid
(so `python -m gallatin_amd.autotune` sweeps every point), and the
one-point form used by the kernel-trace commands parses to exactly that
point."""
import pytest

from gallatin_amd import autotune


def test_defaults_are_the_grid():
    a = autotune.parse_args([])
    assert tuple(m << 20 for m in a.segments_mib) == autotune.SEGMENTS
    assert tuple(a.smallest) == autotune.SMALLEST
    assert tuple(a.threads) == autotune.THREADS
    assert tuple(a.forms) == tuple(autotune.FORMS)
    assert a.pool_gib == 64 and a.nbatches == 256 and a.flags == 0


def test_one_point():
    a = autotune.parse_args("--segments-mib 64 --smallest 256 --threads 1024 "
                            "--forms workgroup --flags 2 --reps 10".split())
    assert a.segments_mib == [64] and a.smallest == [256] and a.threads == [1024]
    assert a.forms == ["workgroup"] and a.flags == 2 and a.reps == 10


def test_unknown_form_is_refused():
    with pytest.raises(SystemExit):
        autotune.parse_args(["--forms", "warp"])


def test_biggest_slice_rule():
    # min(segment / 4096, 4096): a segment holds one block of the biggest class
    assert autotune.biggest_for(4 << 20) == 1024
    assert autotune.biggest_for(16 << 20) == 4096
    assert autotune.biggest_for(64 << 20) == 4096
