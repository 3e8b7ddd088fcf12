"""Convert the reference's safeprimes32.txt into the binary table read at run time.

Usage: python make_safeprimes.py <path to data/safeprimes32.txt> [out]

Each line of the text file is ``a  a*2^32-1  (a*2^32-2)/2``; the first column
is the multiply-with-carry multiplier of stream i (line i, file order).  The
output is the first column as little-endian uint32, in file order.
"""
import sys

import numpy as np


def main(src, out):
    vals = []
    with open(src) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3:
                a, p, q = int(parts[0]), int(parts[1]), int(parts[2])
                assert p == a * 2**32 - 1 and q == (p - 1) // 2, line
                vals.append(a)
    np.asarray(vals, dtype="<u4").tofile(out)
    print("wrote %d multipliers to %s" % (len(vals), out))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2] if len(sys.argv) > 2 else "safeprimes32.u32")
