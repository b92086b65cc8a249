"""The factorization's round schedule on generated problems (host only, no GPU).

Builds tests/sched_check.cpp with the symbolic analysis (csrc/problem.cpp) and the problem
generator (csrc/generator.cpp) and runs it: every node's panel steps run once, in order;
a parent steps only after every child's last step; children of one parent end in distinct
rounds (their extend-adds are plain stores); exactly one child per parent, the last to end,
factors the parent's first diagonal block; each step's tiles are the lower triangle of its
trailing region; split rounds list every panel-row chunk; the tiles that factor a diagonal
block come first in their round; a child's update rows keep their order in the parent.  A second build runs the same checks
under the host sanitizers.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "sophus-rs_amd" / "csrc"


# The sanitized build runs the same checks with the host code under AddressSanitizer (leaks
# included) and UndefinedBehaviorSanitizer, undefined behaviour fatal.
SANITIZE = ["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=undefined"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("sanitize", [False, True], ids=["plain", "asan_ubsan"])
def test_round_schedule_invariants(tmp_path, sanitize):
    exe = tmp_path / "sched_check"
    cmd = ["g++", "-O1", *(SANITIZE if sanitize else []), "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include",
           f"-I{CSRC}", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "sched_check.cpp"),
           str(CSRC / "problem.cpp"), str(CSRC / "generator.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(run.stdout)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-2000:]
    lines = [l for l in run.stdout.splitlines() if l.startswith("ok ")]
    assert len(lines) == 6
    # the benchmark's camera graph runs in well under one round per panel of the level
    # schedule (76 panel steps there)
    rounds_2000 = int(lines[3].split()[1])
    assert rounds_2000 < 70
