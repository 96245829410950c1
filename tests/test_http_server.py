"""HTTP front end (pentago_amd/http_server.py): request parsing and replies in the shape of
web/server/index.js:16-58.  The solver is replaced by a stand-in here (no GPU); the GPU test
in test_gpu_mid_server.py answers a real request."""
import json
import threading
import urllib.error
import urllib.request
from http.server import ThreadingHTTPServer

import pytest

from pentago_amd import http_server as H
from tests.common import golden


def test_parse_known_answer_board():
    kat = golden("mid_kat.json")
    b = H.parse_board("/%d" % kat["board"])
    assert (b.ply + 1) >> 1 == bin(b.side[0] | b.side[1]).count("1")
    bm = H.parse_board("/%dm" % kat["board"])
    assert bm.ply == b.ply - 1


@pytest.mark.parametrize("path", ["/", "/abc", "/12x", "/%d" % (1 << 64), "/%d" % (3 ** 9),
                                  "/%d" % (1 + 3 + 9)])
def test_parse_rejects(path):
    with pytest.raises(ValueError):
        H.parse_board(path)


def test_name_of():
    assert H.name_of(5) == "5"
    assert H.name_of(5 | 1 << 63) == "5m"


class FakeSolver:
    def values(self, board):
        return {"1": 1, "2m": -1}, 0.1


def get(port, path):
    try:
        with urllib.request.urlopen("http://127.0.0.1:%d%s" % (port, path), timeout=10) as r:
            return r.status, r.headers, r.read().decode()
    except urllib.error.HTTPError as e:
        return e.code, e.headers, e.read().decode()


def test_replies():
    httpd = ThreadingHTTPServer(("127.0.0.1", 0), H.make_handler(FakeSolver()))
    th = threading.Thread(target=httpd.serve_forever, daemon=True)
    th.start()
    port = httpd.server_address[1]
    try:
        status, headers, body = get(port, "/bad")
        assert status == 404 and "bad url" in body
        status, _, body = get(port, "/1")            # one stone: too few for a solve
        assert status == 400
        kat = golden("mid_kat.json")
        status, headers, body = get(port, "/%d" % kat["board"])
        assert status == 200
        assert headers["content-type"].startswith("application/json")
        assert headers["access-control-allow-origin"] == "*"
        assert json.loads(body) == {"1": 1, "2m": -1}
    finally:
        httpd.shutdown()
