"""numpy restatement of the seeded child-data generator (pe_fill_seeded), used to build
the same inputs on the host that the device generator writes."""
import numpy as np

M1 = np.uint64(0xbf58476d1ce4e5b9)
M2 = np.uint64(0x94d049bb133111eb)


def mix64(z):
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def seeded_nodes(nodes, seed, first=0):
    """uint64 array of 8*nodes words: node = (win[4], known[4])."""
    key = mix64(np.array([seed], dtype=np.uint64))[0]
    idx = np.arange(nodes, dtype=np.uint64) + np.uint64(first)
    out = np.zeros((nodes, 2, 4), dtype=np.uint64)
    digits = 0xE666
    for w in range(4):
        ctr = (idx * np.uint64(4) + np.uint64(w)) * np.uint64(32)
        known = np.zeros(nodes, dtype=np.uint64)
        for k in range(16):
            r = mix64(key ^ (ctr + np.uint64(k)))
            known = (known | r) if (digits >> k & 1) else (known & r)
        out[:, 1, w] = known
        out[:, 0, w] = mix64(key ^ (ctr + np.uint64(16))) & known
    return out.reshape(-1)
