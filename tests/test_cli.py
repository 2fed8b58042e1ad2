"""mdmc_cli, the command-line front end: the .bin reader/writer (UniformGrid.inl:512-571,
:639-695) round trip and argument errors, with no GPU needed."""
import os
import struct
import subprocess

import numpy as np
import pytest

import fixtures
import oracle_api as oa

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "manifolddmc_amd", "mdmc_cli")

pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="mdmc_cli is not built (python -m manifolddmc_amd.build)")


@pytest.mark.parametrize("problem", ["1-Edge", "Iterative-Fix", "12-Merge-iso-241.525"])
def test_grid_round_trip_is_byte_exact(tmp_path, problem):
    src = os.path.join(fixtures.GOLDEN, "Base-Problems", problem + ".bin")
    out = str(tmp_path / "copy")  # ".bin" is appended, as writeSelectionToFile does
    r = subprocess.run([CLI, src, "--write-grid", out], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert open(out + ".bin", "rb").read() == open(src, "rb").read()


def test_written_grid_reads_back_with_the_reference_geometry(tmp_path):
    # a synthetic grid through the writer: header fields and x-fastest values survive
    dims = (5, 4, 3)
    vals = (np.arange(np.prod(dims)) * 7 % 1000).astype("<u2")
    src = tmp_path / "g.bin"
    src.write_bytes(struct.pack("<3i", *dims) + struct.pack("<3f", 0.5, 0.25, 2.0) + vals.tobytes())
    r = subprocess.run([CLI, str(src), "--write-grid", str(tmp_path / "h.bin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    v, origin, deltas = oa.read_bin(str(tmp_path / "h.bin"))
    assert v.shape == (3, 4, 5) and np.array_equal(v.reshape(-1), vals.astype(np.float32))
    assert np.array_equal(deltas, np.float32([0.5, 0.25, 2.0]))
    assert np.array_equal(origin, np.float32([-1.0, -0.375, -2.0]))  # -0.5 * (dims - 1) * deltas


def test_argument_errors(tmp_path):
    src = os.path.join(fixtures.GOLDEN, "Base-Problems", "1-Edge.bin")
    assert subprocess.run([CLI], capture_output=True).returncode == 2
    assert subprocess.run([CLI, src, "--bogus"], capture_output=True).returncode == 2
    assert subprocess.run([CLI, src, "--iso", "900"], capture_output=True).returncode == 2  # no --out
    assert subprocess.run([CLI, src, "--iso"], capture_output=True).returncode == 2
    bad = tmp_path / "short.bin"
    bad.write_bytes(struct.pack("<3i", 4, 4, 4) + struct.pack("<3f", 1, 1, 1) + b"\0" * 10)
    r = subprocess.run([CLI, str(bad), "--write-grid", str(tmp_path / "o.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "short read" in r.stderr
    assert subprocess.run([CLI, str(tmp_path / "missing.bin")], capture_output=True).returncode == 1
