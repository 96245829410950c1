"""HTTP front end for position solves, in the request shape of the reference's server
(web/server/index.js:16-58): GET /<board> where <board> is the decimal board number
(four base-3 quadrants of 16 bits, high_board_t's raw encoding) with a trailing "m" when the
player to move still has to rotate.  The reply is a JSON object mapping the names of the
board, its children and (after a rotation) grandchildren to their values (1 win, 0 tie,
-1 loss for the player to move), as midsolve_traverse produces them
(pentago/mid/midengine.cc:127-158).

Requests go through one serving loop (pe_mid_server_*): concurrent requests are batched on
the device as they arrive.  Boards with fewer than 18 stones are answered by the
reference's precomputed database, not by a solve; they get status 400 here.

    python -m pentago_amd.http_server --port 8000
"""
import argparse
import ctypes as C
import json
import re
import threading
from http.server import BaseHTTPRequestHandler, ThreadingHTTPServer

from . import HalfSupers, MidServer, MidValue, PentagoError, _check, high_from_board, lib

USAGE = ("usage: GET /<board>[m], <board> the decimal board number (quadrant q in bits 16q..16q+15, "
         "base 3, 0 empty / 1 black / 2 white), m if the player to move must still rotate; "
         "at least 18 stones\n")
NAME = re.compile(r"^/(\d+)(m?)$")


def parse_board(path):
    """The high_board_t for a request path, or raise ValueError."""
    m = NAME.match(path)
    if not m:
        raise ValueError("expected /(\\d+)m?")
    board, middle = int(m.group(1)), m.group(2) == "m"
    if board >> 64:
        raise ValueError("board number out of range")
    black = white = 0
    for q in range(4):
        quad = (board >> (16 * q)) & 0xffff
        if quad >= 3 ** 9:
            raise ValueError("quadrant %d out of range" % q)
        for _ in range(9):
            d = quad % 3
            quad //= 3
            black += d == 1
            white += d == 2
    if black - white not in (0, 1):
        raise ValueError("stone counts are not those of a game")
    return high_from_board(board, middle)


def name_of(raw):
    return "%d%s" % (raw & ((1 << 63) - 1), "m" if raw >> 63 else "")


class Solver:
    """One serving loop shared by all request threads."""

    def __init__(self, **server_options):
        self.server = MidServer(**server_options)

    def values(self, board):
        ticket = self.server.submit(board)
        res = (HalfSupers * 19)()
        lat = C.c_double()
        self.server.spots.pop(ticket, None)
        _check(lib().pe_mid_server_wait(self.server.ptr, ticket, res, C.byref(lat)), "pe_mid_server_wait")
        vals = (MidValue * 163)()
        n = C.c_int()
        _check(lib().pe_midsolve_traverse(C.byref(board), res, vals, C.byref(n)), "pe_midsolve_traverse")
        return {name_of(int(vals[i].raw)): int(vals[i].value) for i in range(n.value)}, lat.value

    def close(self):
        self.server.close()


def make_handler(solver):
    class Handler(BaseHTTPRequestHandler):
        def _reply(self, status, body, ctype):
            data = body.encode()
            self.send_response(status)
            self.send_header("content-type", ctype)
            self.send_header("x-content-type-options", "nosniff")
            if status == 200:
                self.send_header("cache-control", "public")
                self.send_header("access-control-allow-origin", "*")
            self.send_header("content-length", str(len(data)))
            self.end_headers()
            self.wfile.write(data)

        def do_GET(self):
            try:
                board = parse_board(self.path)
            except ValueError as e:
                self._reply(404, "bad url %s: %s\n\n%s" % (self.path, e, USAGE), "text/plain; charset=utf-8")
                return
            stones = (board.ply + 1) >> 1
            if stones < 18:
                self._reply(400, "lookup failed: boards with fewer than 18 stones are not solved here\n",
                            "text/plain; charset=utf-8")
                return
            try:
                values, _ = solver.values(board)
            except PentagoError as e:
                self._reply(500, "lookup failed: %s\n" % e, "text/plain; charset=utf-8")
                return
            self._reply(200, json.dumps(values), "application/json; charset=utf-8")

        def log_message(self, *args):
            pass

    return Handler


def serve(host="127.0.0.1", port=8000, **server_options):
    """Start the HTTP server in a background thread; returns (httpd, solver)."""
    solver = Solver(**server_options)
    httpd = ThreadingHTTPServer((host, port), make_handler(solver))
    threading.Thread(target=httpd.serve_forever, daemon=True).start()
    return httpd, solver


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=8000)
    args = ap.parse_args()
    from . import init
    init()
    solver = Solver()
    httpd = ThreadingHTTPServer((args.host, args.port), make_handler(solver))
    try:
        httpd.serve_forever()
    finally:
        solver.close()


if __name__ == "__main__":
    main()
