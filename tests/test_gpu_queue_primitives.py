"""Costs of the primitives a persistent device-side request queue would use
(tests/host/queue_primitives.hip): work claims by agent-scope fetch-add on
one and on eight counters, dependent polling loads of pinned host memory,
and release fences after 4 KiB writes. The counters must end at their exact
counts; the timings are printed, not asserted (they are the figures behind
the server's host-side admission)."""
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host", "libtrove_devtests.so")


def test_queue_primitives(cuda):
    assert os.path.exists(LIB), f"{LIB} not built (run __graft_entry__.build())"
    L = ctypes.CDLL(LIB)
    L.devtest_queue_primitives.argtypes = [ctypes.c_void_p]
    L.devtest_queue_primitives.restype = ctypes.c_int
    out = (ctypes.c_double * 8)()
    assert L.devtest_queue_primitives(ctypes.cast(out, ctypes.c_void_p)) == 0
    claims1, claims8, counts_ok, poll_us, pub_none, pub_agent, pub_system, done_ok = list(out)
    assert counts_ok == 1.0, "a claim counter lost or gained an update"
    assert done_ok == 1.0, "the done counter lost or gained an update"
    assert min(claims1, claims8, poll_us, pub_none, pub_agent, pub_system) > 0
    print(f"\nqueue primitives: claims {claims1:.1f} M/s on one counter, {claims8:.1f} M/s over eight; "
          f"host-memory poll {poll_us:.3f} us per dependent load; 4 KiB publish per wavefront "
          f"{pub_none:.2f} us (no fence), {pub_agent:.2f} us (agent release), {pub_system:.2f} us (system release)")
