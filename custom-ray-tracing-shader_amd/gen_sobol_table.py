"""Generate csrc/sobol_directions.inc: 260 dimensions x 32 Sobol direction numbers.

The numbers follow Joe & Kuo's published new-joe-kuo-6.21201 primitive polynomials and
initial direction numbers (S. Joe and F. Y. Kuo, "Constructing Sobol sequences with better
two-dimensional projections", SIAM J. Sci. Comput. 30, 2008), with dimension 0 the van der
Corput sequence.  That is the table the reference obtains from cuRAND's
CURAND_DIRECTION_VECTORS_32_JOEKUO6 (Native/sampling.cuh:6-15); the same numbers ship in
rocRAND as ROCRAND_DIRECTION_VECTORS_32_JOEKUO6, which the tests compare against.

The polynomial/initial-number data is read from SciPy's copy of the Joe-Kuo file.
Run: python gen_sobol_table.py  (writes the .inc next to the C++ sources).
"""
import os
from pathlib import Path

import numpy as np
import scipy

DIMENSIONS = 4 + 8 * 32  # camera dims + 8 per bounce for 32 bounces (Native/sampling.cuh:6)


def direction_numbers(ndim=DIMENSIONS):
    data = np.load(os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz"))
    poly, vinit = data["poly"], data["vinit"]
    table = np.zeros((ndim, 32), dtype=np.uint64)
    table[0] = [1 << (31 - k) for k in range(32)]
    for dim in range(1, ndim):
        p = int(poly[dim])
        degree = p.bit_length() - 1
        v = [int(vinit[dim, k]) << (31 - k) for k in range(degree)] + [0] * (32 - degree)
        for k in range(degree, 32):
            x = v[k - degree] ^ (v[k - degree] >> degree)
            for i in range(1, degree):
                if (p >> (degree - i)) & 1:
                    x ^= v[k - i]
            v[k] = x & 0xFFFFFFFF
        table[dim] = v
    return table.astype(np.uint32)


def main():
    table = direction_numbers()
    out = Path(__file__).resolve().parent / "csrc" / "sobol_directions.inc"
    lines = ["// Generated by gen_sobol_table.py: Joe-Kuo 6.21201 direction numbers, 260 x 32."]
    flat = table.reshape(-1)
    for i in range(0, flat.size, 8):
        lines.append("    " + ", ".join(f"0x{int(x):08x}u" for x in flat[i:i + 8]) + ",")
    out.write_text("\n".join(lines) + "\n")
    print(f"wrote {out} ({flat.size} words)")


if __name__ == "__main__":
    main()
