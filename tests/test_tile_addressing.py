"""The tile body's operand addressing (blas-for-multi-core-dsp_amd/csrc/
tile_body.h, Stager::load with buffer loads, and Stager::store's unshift),
restated on the CPU and checked for every thread of every staging shape the
kernels use: each element that reaches the LDS stage is the matrix element of
its (row, column) or zero outside the matrix, and every address a live load
reads lies inside its own row's columns.

The restatement: thread tid's vector i is tile row tid / CPR + i RSTEP,
columns (tid % CPR) V + [0, V); a vector whose row is past R or whose first
column is past Cn reads nothing (offset past the resource); one that ends
past Cn is loaded sh = first + V - Cn elements early (sh < V since Cn >= V)
from a resource based V elements before the row origin, and unshifted at the
store (element e = loaded[e + sh], zero past the end).  No GPU needed.
"""
import random

import pytest

# (elements per 16-byte vector, K tile) for FP32 and FP64; tile edge and waves
CONFIGS = [(4, 32), (2, 16)]
TILES = [(64, 4), (128, 8), (128, 4)]


def check_tile(V, BK, BM, NW, KC, R, Cn, ld, r0, c0):
    threads = 64 * NW
    TR, TC = (BM, BK) if KC else (BK, BM)
    CPR = TC // V
    NCH = TR * CPR // threads
    RSTEP = threads // CPR
    for tid in range(threads):
        row, col = tid // CPR, (tid % CPR) * V
        gc = c0 + col
        col_in = gc < Cn
        over = gc + V - Cn
        sh = over if (col_in and over > 0) else 0
        assert 0 <= sh < V
        for i in range(NCH):
            gr = r0 + row + i * RSTEP
            live = col_in and gr < R
            base = (r0 + i * RSTEP) * ld + c0 - V   # resource base, in elements
            off = row * ld + col + V - sh           # lane offset, in elements
            assert off >= 0                          # an unsigned 32-bit offset must not wrap
            loaded = []
            for e in range(V):
                if live:
                    a = base + off + e
                    assert gr * ld <= a < gr * ld + Cn, (tid, i, e)
                    loaded.append((a // ld, a % ld))
                else:
                    loaded.append(None)
            stored = [loaded[e + sh] if e + sh < V else None for e in range(V)]
            for e in range(V):
                want = (gr, gc + e) if (gr < R and gc + e < Cn) else None
                assert stored[e] == want, (tid, i, e, stored[e], want)


@pytest.mark.parametrize("V,BK", CONFIGS)
@pytest.mark.parametrize("BM,NW", TILES)
@pytest.mark.parametrize("KC", [True, False])
def test_stager_addressing(V, BK, BM, NW, KC):
    TR, TC = (BM, BK) if KC else (BK, BM)
    rng = random.Random(V * 1000 + BM * 10 + NW + (1 if KC else 0))
    cases = [(TR, TC, TC), (1, V, V), (TR + 1, TC + 1, TC + 3), (3 * TR - 1, 2 * TC + V + 1, 2 * TC + V + 1)]
    for _ in range(40):
        R = rng.randint(1, 3 * TR)
        Cn = rng.randint(V, 3 * TC)
        cases.append((R, Cn, Cn + rng.randint(0, 5)))
    for R, Cn, ld in cases:
        for r0 in range(0, R, TR):
            for c0 in range(0, Cn, TC):
                check_tile(V, BK, BM, NW, KC, R, Cn, ld, r0, c0)
