"""Serving loop (pe_mid_server_*): every answer equals the plain per-board solve, alone and
in a mix of concurrent requests, under a workspace pool that forces requests to wait."""
import threading
import time

import numpy as np
import pytest

import pentago_amd as P
from pentago_amd.workloads import board_spots, seeded_boards

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    P.init()


def test_same_alone_and_in_the_mix():
    boards = seeded_boards(48, 31, 22, 34)
    expect = [P.midsolve_internal(b) for b in boards]
    server = P.MidServer()
    alone = []
    for b in boards:
        alone.append(server.wait(server.submit(b))[0])
    # the mix: submitted from another thread with small gaps while earlier ones run
    tickets = []

    def feed():
        rng = np.random.default_rng(1)
        for b in boards:
            tickets.append(server.submit(b))
            time.sleep(float(rng.exponential(0.0005)))

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    mix = [server.wait(t) for t in tickets]
    stats = server.stats()
    server.close()
    print("stats", stats, flush=True)
    assert alone == expect
    assert [m[0] for m in mix] == expect
    assert all(m[1] > 0 for m in mix)
    assert stats["completed"] == 2 * len(boards)


def test_pool_budget_makes_requests_wait():
    boards = seeded_boards(40, 32, 20, 30)
    big = max(board_spots(b) for b in boards)
    from tests.test_mid_schedule import table_bytes, work_bytes
    budget = int(1.5 * (table_bytes(big) + work_bytes(big))) + (4 << 20)
    server = P.MidServer(workspace_budget=budget, streams=2, max_batch=8)
    tickets = [server.submit(b) for b in boards]
    got = [server.wait(t)[0] for t in tickets]
    stats = server.stats()
    server.close()
    print("stats", stats, flush=True)
    assert stats["peak_workspace"] <= budget
    assert stats["batches"] >= 5
    assert got == [P.midsolve_internal(b) for b in boards]


def test_poll_and_rejections():
    server = P.MidServer(workspace_budget=64 << 20)
    with pytest.raises(P.PentagoError):           # 17 empty cells need far more than 64 MB
        server.submit(seeded_boards(1, 3, 19, 19)[0])
    with pytest.raises(P.PentagoError):           # not a midsolve board (too few stones)
        server.submit(P.high_board(1, 2, 4))
    b = seeded_boards(1, 4, 30, 30)[0]
    t = server.submit(b)
    out = None
    for _ in range(20000):
        out = server.poll(t)
        if out is not None:
            break
        time.sleep(0.0005)
    server.close()
    assert out is not None and out[0] == P.midsolve_internal(b)


def test_http_front_end_known_answer():
    """GET /<board> through the HTTP front end: the 163 known values of
    pentago/mid/benchmark_test.cc:39-58 as a JSON object of board names."""
    import json
    import urllib.request
    from pentago_amd import http_server as H
    from tests.common import golden
    kat = golden("mid_kat.json")
    httpd, solver = H.serve(port=0)
    try:
        port = httpd.server_address[1]
        with urllib.request.urlopen("http://127.0.0.1:%d/%d" % (port, kat["board"]), timeout=60) as r:
            got = json.loads(r.read().decode())
    finally:
        httpd.shutdown()
        solver.close()
    assert got == {"%d%s" % (b, "m" if m else ""): v for b, m, v in kat["values"]}
