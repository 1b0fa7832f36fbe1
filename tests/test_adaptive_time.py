"""AdaptiveTime state machine (AdaptiveTime.hh, restating the reference's
AdaptiveTime.hh:33-86) driven by a scripted sequence of inspectVelocity
results through a mock particle set: growth by sqrt(2) after an all-slow
step with one step of waiting, rollback over two calls with dt / 4 and
three steps of waiting, and the time bookkeeping.  Host only (g++)."""
import math
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def trace(tmp_path_factory):
    exe = tmp_path_factory.mktemp("at") / "adaptive_time_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpp", "adaptive_time_test.cpp")], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = []
    for line in out.strip().splitlines():
        j, t, dt, rb, calls = line.split()
        rows.append((int(j), float(t), float(dt), int(rb), calls))
    return rows


def test_state_machine(trace):
    s2 = math.sqrt(2.0)
    f = lambda x: pytest.approx(x, rel=1e-6)
    # script: 0, -1, -1, -1, 1, 0, -1, -1, -1, -1, -1, 0
    # step 1 (res 0): t=1, dt=1
    j, t, dt, rb, calls = trace[0]
    assert (j, t, dt, rb, calls) == (1, f(1.0), f(1.0), 0, "-")
    # step 2 (slow, no wait): calcVinit((3 - sqrt2) dt), dt *= sqrt2, wait 1 -> 0 at once
    j, t, dt, rb, calls = trace[1]
    assert j == 2 and t == f(2.0) and dt == f(s2) and calls.startswith("V(")
    assert float(calls[2:-1]) == f(3.0 - s2)
    # step 3 (slow again, wait was decremented to 0 in the same call): grows again
    j, t, dt, rb, calls = trace[2]
    assert j == 3 and t == f(2.0 + s2) and dt == f(2.0) and calls.startswith("V(")
    # step 4 (slow): grows to 2 sqrt2
    j, t, dt, rb, calls = trace[3]
    assert dt == f(2.0 * s2)
    t4 = t
    # step 5 (too fast): time advances, j advances, rollback(dt) requested
    j, t, dt, rb, calls = trace[4]
    assert j == 5 and t == f(t4 + 2.0 * s2) and rb == 1 and calls.startswith("R(")
    # step 6 (second stage, whatever the result): rollback2(dt), t -= dt, dt /= 4,
    # j unchanged, then three steps of waiting
    j, t, dt, rb, calls = trace[5]
    assert j == 5 and t == f(t4) and dt == f(2.0 * s2 / 4.0) and rb == 0
    assert calls.startswith("R2(")
    for k in (6, 7, 8):       # slow, but waiting (wait 3 -> 2 -> 1 -> 0)
        assert trace[k][4] == "-" and trace[k][2] == f(2.0 * s2 / 4.0)
    j, t, dt, rb, calls = trace[9]   # the wait is over: grows
    assert calls.startswith("V(") and dt == f(2.0 * s2 / 4.0 * s2)
    assert trace[11][4] == "-"
