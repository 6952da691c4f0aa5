"""CPU sanitizer builds (tests/sanitize/): the host library and the oracle
under AddressSanitizer + UndefinedBehaviorSanitizer (leaks included), and
the host library's threaded diagram builders under ThreadSanitizer.  The
driver builds every fixture on one and on eight host threads, with and
without state-dominance partners, checks that both give the same arrays,
and (ASan build) runs the oracle's three methods and the frontier sort on
them.  Any sanitizer report fails the run."""
import os
import subprocess

import pytest

from conftest import INSTANCES, problem_type_of
from pmdd_amd import instances as gen

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def _args(extra):
    out = []
    for name in sorted(os.listdir(INSTANCES)):
        if "ncities15" in name:
            continue  # the large TSP fixtures take minutes under the sanitizers
        out.append(f"{os.path.join(INSTANCES, name)}:{problem_type_of(name)}")
    return out + extra


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_sanitized_host_and_oracle(tmp_path, kind):
    subprocess.run(["make", "-s", "-C", HERE, kind], check=True)
    extra = []
    if kind == "asan":
        extra.append(gen.write(gen.knapsack(7, 30, 4), str(tmp_path / "kp.dat")) + ":1")
        extra.append(gen.write(gen.tsp(9, 9, 3), str(tmp_path / "tsp.dat")) + ":3")
    else:
        # wide enough (layers past 65,536 states) for the threaded builder paths
        extra.append(gen.write(gen.setpacking(5, 130, 2, m=26), str(tmp_path / "sp.dat")) + ":2")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(HERE, "_build", f"driver_{kind}")] + _args(extra),
                       capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout[-2000:], r.stderr[-4000:])
