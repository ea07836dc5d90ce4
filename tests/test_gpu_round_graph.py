"""The server's captured batched rounds (nuif_fusion_round_prepare + a graph replayed while
the round's launches stay the same) against rounds enqueued launch by launch: two servers
play the same frames, one replaying captured rounds and one not; a session joins after
the third round (another launch signature: the round is captured anew) and one leaves
after the fifth.  Every session's poses of every frame and its maps equal bit for bit."""
import sys

import numpy as np
import pytest

from natureuserinterfacestudio_amd import nuif
from natureuserinterfacestudio_amd import workloads as W

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def test_captured_rounds_equal_enqueued_rounds(gpu):
    wl = W.SMALL
    rounds, first = 7, 4
    announce(f"server rounds replayed from graphs vs enqueued: {first}+1 sessions 320x240 F16, {rounds} rounds")
    frames = [nuif.DeviceArray.from_host(wl.render(f)[0]) for f in range(rounds + first + 2)]
    desc = wl.kinfu_desc(storage=nuif.TSDF_F16, use_graph=True)
    results = []
    for graph in (True, False):
        srv = nuif.Server((first + 2) * nuif.Server.session_bytes(desc) + (1 << 30))
        srv.set_round_graph(graph)
        ids = [srv.open(desc) for _ in range(first)]
        late = None
        played = {}
        for r in range(rounds):
            if r == 3:
                late = srv.open(desc)
                ids.append(late)
            if r == 5:
                srv.close_session(ids.pop(0))
            srv.submit(ids, [frames[r + k] for k in range(len(ids))])
            srv.wait()
            for sid in ids:
                played[sid] = played.get(sid, 0) + 1
        out = {}
        for sid in ids:
            s = srv.sessions[sid]
            out[sid] = (s.poses(played[sid]).copy(), *[m.copy() for m in s.maps()])
        results.append(out)
        srv.close()
    a, b = results
    assert sorted(a) == sorted(b) and len(a) == first
    for sid in a:
        for x, y in zip(a[sid], b[sid]):
            np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32), err_msg=str(sid))
    announce("captured rounds equal enqueued rounds")
