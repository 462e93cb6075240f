"""bench.py's contract pieces that need no GPU: the algorithmic-bytes figure
the roofline is priced on, the committed rocprofv3 traffic summary it reads,
and the one-GPU rule."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_alg_bytes_of_the_4m_instance():
    # 4,194,304 boxes; at the reference rule's cell size 256: 65,536 cells and
    # 4,751,038 memberships; 2,230,679 pairs -> 32N + 12C + 24M + 8P.
    b = bench.alg_bytes(4194304, 65536, 4751038, 2230679)
    assert b["total"] == 266874504
    assert b["count"] + b["scan"] + b["scatter"] + b["test"] == b["total"]
    assert b["test"] == 20 * 4751038 + 8 * 2230679


def test_committed_traffic_profile_is_read():
    # bench.py's defaults run the single-pass launch (place, test)
    traffic, src = bench.read_traffic("onepass", 32.0, 4096)
    assert src == os.path.join("profiles", "traffic_plain_uniform_4m_onepass_cs32_bt4096.txt")
    for k in ("bin_place", "bin_test"):
        assert traffic[k] > 0, k
    # the place kernel reads every box and writes a 20-byte bucket entry per
    # coarse membership; the test kernel reads at least the bucket
    assert traffic["bin_place"] > 16 * 4194304 + 20 * 4194304
    assert traffic["bin_test"] > 20 * 4194304


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
