"""Host logic of the serving loops under AddressSanitizer, on the CPU.

tests/standin builds the library's host code (host-only compilation, host
ASAN) against a stand-in of the HIP runtime (tests/standin/mockhip.cpp: device
memory is host memory, kernels are no-ops except the list builds' count
export, graph capture is recorded and replayed, destroyed handles are caught)
and runs serve_replay.cpp: 40 water-box jobs through the stream server's
create / chunked propagate / read back / destroy sequence, six more one after
another, 20 through fused batches re-formed at every membership change, and 16
in the many-system benchmark's order (all created, each warmed and advanced
alone through the launch trial, then one autotuned fused batch).  With
APO_STANDIN_GROW=1 every list build reports more entries than the last, which
drives list_export_check -> list_reserve -> new generation -> graph
recapture.  Every call must succeed, ASAN must stay silent, and no device
allocation, stream, event, graph or pinned buffer may outlive its job."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
STANDIN = os.path.join(HERE, "standin")
BUILD = os.path.join(STANDIN, "_build")

pytestmark = pytest.mark.skipif(
    not os.path.isdir(STANDIN) or shutil.which("g++") is None
    or not os.path.exists("/opt/rocm/bin/hipcc"), reason="needs g++ and hipcc")


@pytest.fixture(scope="module")
def replay():
    r = subprocess.run(["bash", os.path.join(STANDIN, "build.sh")], capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return os.path.join(BUILD, "serve_replay")


def _run(exe, grow):
    env = dict(os.environ, APO_STANDIN_GROW="1" if grow else "0",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:detect_stack_use_after_return=1")
    r = subprocess.run([exe, "40"], capture_output=True, text=True, timeout=900, env=env)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "LeakSanitizer" not in out
    m = re.search(r"graphs instantiated (\d+)", out)
    a = re.search(r"device allocations (\d+),", out)
    return int(m.group(1)), int(a.group(1)), out


def test_serving_sequences_clean_under_asan(replay):
    inst, allocs, out = _run(replay, grow=False)
    assert "stream server: 40 jobs" in out and "plain run: 6 jobs" in out
    assert "batch server: 20 jobs" in out
    assert "many-system bench order: 16 systems" in out


def test_list_growth_path_clean_under_asan(replay):
    """Growth must really happen (more allocations and graph recaptures than
    without it) and stay clean."""
    inst0, allocs0, _ = _run(replay, grow=False)
    inst1, allocs1, out = _run(replay, grow=True)
    assert allocs1 > allocs0 and inst1 > inst0, (allocs0, allocs1, inst0, inst1)
