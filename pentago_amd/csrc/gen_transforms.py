"""Generates transforms.h: the line transforms of unreflected lines as straight-line 32-bit code.

A super (256 bits, index r0 + 4 r1 + 16 r2 + 64 r3) is held as eight 32-bit dwords; bit
position = index, so index bits 0-2 select the bit inside a byte, bits 3-4 the byte inside a
dword and bits 5-7 the dword.  A line whose child section standardizes through a rotation
by r (no reflection) maps child supers to parent supers by
    Y(x) = X(commute[g][x] + y),   (g, y) = (-r, r (1,1,1,1))    (pentago/end/compute.cc:69)
and for these four symmetries commute[g] is a plain permutation of the coordinates, so
    Y(x) = X(j),  j_q = x_{pi(q)} + s   with s = r    (every coordinate shifted by s).
That is an affine permutation of the 8 index bits.  Each plan below realizes it with:
  * exchanges of an in-byte index bit with a dword index bit (a word-pair swap),
  * exchanges of an in-byte bit with a byte index bit, or of two in-byte bits (delta swaps),
  * rotations of the coordinate held in index bits 0-1 or 2-3 (masked shifts),
  * renames of dwords (free) and byte gathers (v_perm_b32, one per output dword per pair of
    source dwords) for everything that lives in index bits 3-7.
The generator simulates every step on the 256 bit positions and only emits a plan whose
result equals the target map; the tests compare the emitted functions with rearrange and
shift_local bit by bit.

Usage: python3 gen_transforms.py > transforms.h
"""
import sys

# j_q = x_{pi[q]} + s: (pi, s) per inverse global rotation g (inverse of the standardizing
# rotation r = (4 - g) & 3; s = r)
CASES = {
    1: ((2, 0, 3, 1), 3),
    2: ((3, 2, 1, 0), 2),
    3: ((1, 3, 0, 2), 1),
}


def coords(i):
    return [(i >> (2 * q)) & 3 for q in range(4)]


def index(c):
    return sum((c[q] & 3) << (2 * q) for q in range(4))


def target(pi, s):
    return [index([coords(x)[pi[q]] + s for q in range(4)]) for x in range(256)]


class Gen:
    def __init__(self):
        self.st = list(range(256))   # st[pos] = input index held at bit position pos
        self.lines = []
        self.cost = 0                # 32-bit VALU operations per super
        self.names = ["d%d" % i for i in range(8)]
        self.tmp = 0

    def fresh(self):
        self.tmp += 1
        return "t%d" % self.tmp

    # exchange in-byte index bit a (0-2) with dword index bit b (5-7)
    def xchg_dw(self, a, b):
        k, s = b - 5, 1 << a
        m = {0: 0x55555555, 1: 0x33333333, 2: 0x0F0F0F0F}[a]
        new = self.st[:]
        for pos in range(256):
            if (pos >> a) & 1 and not (pos >> b) & 1:
                q = pos ^ (1 << a) ^ (1 << b)
                new[pos], new[q] = self.st[q], self.st[pos]
        self.st = new
        for D in range(8):
            if D >> k & 1:
                continue
            lo, hi = self.names[D], self.names[D | 1 << k]
            t = self.fresh()
            self.lines.append("const uint32_t %s = lop3_32<kAndXor>(%s >> %d, %s, 0x%08Xu); %s ^= %s; %s ^= %s << %d;"
                              % (t, lo, s, hi, m, hi, t, lo, t, s))
            self.cost += 5

    # exchange index bits a < b inside each dword (a in-byte; b in-byte or a byte bit)
    def delta(self, a, b):
        s = (1 << b) - (1 << a)
        m = 0
        for p in range(32):
            if (p >> a) & 1 and not (p >> b) & 1:
                m |= 1 << p
        new = self.st[:]
        for pos in range(256):
            if (pos >> a) & 1 and not (pos >> b) & 1:
                q = pos ^ (1 << a) ^ (1 << b)
                new[pos], new[q] = self.st[q], self.st[pos]
        self.st = new
        for D in range(8):
            x = self.names[D]
            t = self.fresh()
            self.lines.append("const uint32_t %s = lop3_32<kAndXor>(%s, %s >> %d, 0x%08Xu); %s = lop3_32<kXor3>(%s, %s, %s << %d);"
                              % (t, x, x, s, m, x, x, t, t, s))
            self.cost += 4

    # Y(z) = X(z + k e) for the coordinate held in index bits (lo, lo + 1), lo = 0 or 2
    def rot(self, lo, k):
        k &= 3
        if not k:
            return
        new = self.st[:]
        for pos in range(256):
            c = (pos >> lo) & 3
            src = (pos & ~(3 << lo)) | (((c + k) & 3) << lo)
            new[pos] = self.st[src]
        self.st = new
        if lo == 0:
            sh, m = k, 0x11111111 * (0xF >> k)
            width = 4
        else:
            sh, m = 4 * k, 0x00010001 * (0xFFFF >> (4 * k))
            width = 16
        for D in range(8):
            x = self.names[D]
            self.lines.append("%s = lop3_32<kSel>(%s >> %d, %s << %d, 0x%08Xu);"
                              % (x, x, sh, x, width - sh, m))
            self.cost += 3

    # any map new[pos] = st[f(pos)] with f moving only index bits 3-7: dword renames and
    # byte gathers
    def bytemap(self, f):
        want = [self.st[f(pos)] for pos in range(256)]
        self.gather(want)

    def gather(self, want):
        where = {}
        for D in range(8):
            for k in range(4):
                key = tuple(self.st[D * 32 + k * 8 + b] for b in range(8))
                where[key] = (D, k)
        out = []
        srcs_all = []
        for D2 in range(8):
            src = []
            for k2 in range(4):
                key = tuple(want[D2 * 32 + k2 * 8 + b] for b in range(8))
                if key not in where:
                    raise ValueError("byte %d of dword %d holds no whole source byte" % (k2, D2))
                src.append(where[key])
            srcs_all.append(src)
        new_names = []
        for D2 in range(8):
            src = srcs_all[D2]
            dws = sorted(set(d for d, _ in src))
            if len(dws) == 1 and all(k == k2 for k2, (_, k) in enumerate(src)):
                new_names.append(self.names[dws[0]])   # rename
                continue
            v = self.fresh()
            if len(dws) <= 2:
                hi = dws[0]
                lo = dws[1] if len(dws) == 2 else dws[0]
                sel = 0
                for k2, (d, k) in enumerate(src):
                    sel |= (k if d == lo else 4 + k) << (8 * k2)
                out.append("const uint32_t %s = pe_perm(%s, %s, 0x%08Xu);" % (v, self.names[hi], self.names[lo], sel))
                self.cost += 1
            else:
                a_hi, a_lo = dws[0], dws[1]
                b_hi, b_lo = dws[2], dws[3] if len(dws) == 4 else dws[2]
                ta, tb = self.fresh(), self.fresh()
                sa = sb = sf = 0
                for k2, (d, k) in enumerate(src):
                    if d in (a_hi, a_lo):
                        sa |= (k if d == a_lo else 4 + k) << (8 * k2)
                        sb |= 0x0C << (8 * k2)
                        sf |= (4 + k2) << (8 * k2)
                    else:
                        sb |= (k if d == b_lo else 4 + k) << (8 * k2)
                        sa |= 0x0C << (8 * k2)
                        sf |= k2 << (8 * k2)
                out.append("const uint32_t %s = pe_perm(%s, %s, 0x%08Xu);" % (ta, self.names[a_hi], self.names[a_lo], sa))
                out.append("const uint32_t %s = pe_perm(%s, %s, 0x%08Xu);" % (tb, self.names[b_hi], self.names[b_lo], sb))
                out.append("const uint32_t %s = pe_perm(%s, %s, 0x%08Xu);" % (v, ta, tb, sf))
                self.cost += 3
            new_names.append(v)
        self.lines.extend(out)
        # new values become the working dwords (mutable copies)
        fixed = []
        for D2 in range(8):
            w = self.fresh()
            fixed.append(w)
            self.lines.append("uint32_t %s = %s;" % (w, new_names[D2]))
        self.names = fixed
        self.st = want

    def finish(self, tgt):
        self.gather(tgt)
        assert self.st == tgt


def swap_bits(a, b):
    def f(pos):
        x, y = (pos >> a) & 1, (pos >> b) & 1
        return (pos & ~((1 << a) | (1 << b))) | (x << b) | (y << a)
    return f


def add_at(lo, hi, k):
    """position map adding k to the 2-bit number held in position bits (lo, hi): used in a
    bytemap step it shifts the coordinate whose bits sit there, Y(z) = X(z + k e)"""
    def f(pos):
        c = ((pos >> lo) & 1) | (((pos >> hi) & 1) << 1)
        c = (c + k) & 3
        return (pos & ~((1 << lo) | (1 << hi))) | ((c & 1) << lo) | ((c >> 1) << hi)
    return f


def cycle3(a, b, c):
    """position map for moving the content of bit a to bit b, b to c and c to a"""
    def f(pos):
        x = [(pos >> i) & 1 for i in range(8)]
        y = x[:]
        y[a], y[b], y[c] = x[b], x[c], x[a]
        return sum(v << i for i, v in enumerate(y))
    return f


def compose(*fs):
    """bytemap of the steps in order: Y1(p) = X(f1(p)), Y2(p) = Y1(f2(p)), ..."""
    def f(pos):
        for g in reversed(fs):
            pos = g(pos)
        return pos
    return f


def plan(g):
    pi, s = CASES[g]
    tgt = target(pi, s)
    G = Gen()
    # Every coordinate is shifted by s while its two bits sit in index bits 3-7 (a dword
    # rename or part of a byte gather), except where noted.
    if g == 1:
        # j = (x2, x0, x3, x1) + 3: input bit b moves to 0->4 1->5 2->0 3->1 4->6 5->7 6->2 7->3
        G.bytemap(add_at(6, 7, 3))                        # coordinate 3 (rename)
        G.xchg_dw(2, 6)                                   # bit 6 -> 2, bit 2 -> 6
        G.bytemap(compose(swap_bits(3, 7), add_at(6, 7, 3)))   # bit 3 -> 7; coordinate 1 now in 6-7
        G.xchg_dw(0, 6)
        G.xchg_dw(1, 7)
    elif g == 2:
        # j = (x3, x2, x1, x0) + 2: bits 0<->6, 1<->7, 2<->4, 3<->5; +2 flips every high bit
        G.bytemap(add_at(6, 7, 2))                        # coordinate 3 (rename)
        G.xchg_dw(0, 6)
        G.xchg_dw(1, 7)
        G.bytemap(swap_bits(4, 5))
        G.xchg_dw(2, 5)
    elif g == 3:
        # j = (x1, x3, x0, x2) + 1: input bit b moves to 0->2 1->3 2->6 3->7 4->0 5->1 6->4 7->5
        G.bytemap(compose(add_at(6, 7, 1), add_at(4, 5, 1), cycle3(4, 5, 6)))   # coordinates 3, 2; 4 -> 5 -> 6 -> 4
        G.xchg_dw(0, 5)
        G.xchg_dw(1, 6)
        G.bytemap(add_at(5, 6, 1))                        # coordinate 0, now in 5-6 (rename)
        G.xchg_dw(2, 5)
    G.finish(tgt)
    return G


def emit():
    out = []
    out.append("// Generated by gen_transforms.py; do not edit.")
    out.append("// Line transforms of unreflected lines, Y(x) = X(commute[g][x] + y) for")
    out.append("// (g, y) = (1, 0xff), (2, 0xaa), (3, 0x55), as straight-line 32-bit code.")
    out.append("#pragma once")
    out.append('#include "super.h"')
    out.append("")
    out.append("namespace pe {")
    out.append("")
    for g in (1, 2, 3):
        G = plan(g)
        pi, s = CASES[g]
        out.append("// g = %d: j = (%s) + %d; %d 32-bit operations per super" %
                   (g, ", ".join("x%d" % p for p in pi), s, G.cost))
        out.append("PE_HD S4 line_transform%d(S4 a) {" % g)
        out.append("  uint32_t %s;" % ", ".join("d%d = uint32_t(a.w[%d]%s)" % (D, D // 2, " >> 32" if D & 1 else "")
                                               for D in range(8)))
        for l in G.lines:
            out.append("  " + l)
        n = G.names
        out.append("  S4 r;")
        for w in range(4):
            out.append("  r.w[%d] = uint64_t(%s) | uint64_t(%s) << 32;" % (w, n[2 * w], n[2 * w + 1]))
        out.append("  return r;")
        out.append("}")
        out.append("")
    out.append("}  // namespace pe")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    sys.stdout.write(emit())
