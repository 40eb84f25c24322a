"""bench.py's reading of the committed traffic profile: a record is used only
for its own launch shape and the tick-kernel source it measured."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import bench  # noqa: E402

SHAPE = {"hub": 2048, "heavy": 512, "apl": 16, "grid": 1024}


def _write(tmp_path, sha):
    p = tmp_path / "traffic.txt"
    p.write_text("# header\n"
                 "shape: hub=4096 heavy=512 apl=16 grid=1024\n"
                 f"kernel_source_sha256: {sha}\n"
                 "traffic_bytes_per_launch: 111.0\n"
                 "traffic_bytes_range: 55.0 111.0\n"
                 "shape: hub=2048 heavy=512 apl=16 grid=1024\n"
                 f"kernel_source_sha256: {sha}\n"
                 "traffic_bytes_per_launch: 222.0\n"
                 "traffic_bytes_range: 66.0 222.0\n")
    return str(p)


def test_record_of_this_shape_and_kernel(tmp_path):
    b, rng, src = bench.read_traffic(_write(tmp_path, bench.kernel_source_sha()), SHAPE)
    assert b == 222.0 and rng == (66.0, 222.0) and "this launch shape" in src


def test_stale_kernel_or_other_shape_gives_none(tmp_path):
    b, rng, why = bench.read_traffic(_write(tmp_path, "0" * 64), SHAPE)
    assert b is None and rng is None and "another version" in why
    b, rng, why = bench.read_traffic(_write(tmp_path, bench.kernel_source_sha()), dict(SHAPE, hub=1024))
    assert b is None and "no traffic record" in why
    b, rng, why = bench.read_traffic(str(tmp_path / "missing.txt"), SHAPE)
    assert b is None and "no traffic profile" in why
