"""gp_check_partitions (av1mi_check_partitions, host code of the library) against a feasibility recursion written
here from the AV1 partition semantics (test_oracle_pins.partition_blocks): for random shape masks, partition masks and
maxPart, the library refuses exactly the options under which some superblock -- whatever part of it the frame covers --
has no partition into evaluated shapes that the DP may choose (inside the frame the options' partitions, past its
bottom / right edge SPLIT or HORZ / VERT, past both SPLIT)."""
import functools

import numpy as np
import pytest

import av1gpu_amd as A
from test_oracle_pins import BS, HORZ, SPLIT, VERT, partition_blocks


def feasible(shape_mask, part8, partN, max_part):
    def evaluated(w, h):
        bs = BS.get((w, h))
        return bs is not None and (shape_mask >> bs) & 1

    @functools.lru_cache(maxsize=None)
    def node(n, x, y, FW, FH):
        if x >= FW or y >= FH:
            return True
        h = n // 2
        rows, cols = y + h < FH, x + h < FW
        if rows and cols:
            pm = part8 if n == 8 else partN
            cand = [SPLIT] + ([p for p in range(10) if (pm >> p) & 1 and (n > 8 or p <= SPLIT) and p != SPLIT] if n <= max_part else [])
        elif cols:
            cand = [SPLIT] + ([HORZ] if n <= max_part else [])
        elif rows:
            cand = [SPLIT] + ([VERT] if n <= max_part else [])
        else:
            cand = [SPLIT]
        for p in cand:
            ok = True
            for b in partition_blocks(p, n, x, y):
                if b[0] == "node":
                    ok = node(b[1], b[2], b[3], FW, FH)
                elif b[2] < FW and b[3] < FH:
                    ok = evaluated(b[0], b[1])
                if not ok:
                    break
            if ok:
                return True
        return False

    return all(node(64, 0, 0, fw, fh) for fw in range(8, 65, 8) for fh in range(8, 65, 8))


def test_partition_check_equals_recursion():
    rng = np.random.default_rng(11)
    shapes = [s for s in range(22) if not 13 <= s <= 15]
    seen = {True: 0, False: 0}
    for trial in range(400):
        mask = sum(1 << s for s in shapes if rng.random() < rng.choice([0.3, 0.6, 0.9]))
        part8 = int(rng.integers(0, 16))
        partN = int(rng.integers(0, 1024))
        max_part = int(rng.choice([16, 32, 64]))
        want = feasible(mask, part8, partN, max_part)
        P = A.GpParams(shapeMask=mask, partMask8=part8, partMaskN=partN, maxPart=max_part)
        try:
            A.check_partitions(P)
            got = True
        except ValueError:
            got = False
        assert got == want, (hex(mask), hex(part8), hex(partN), max_part)
        seen[want] += 1
    assert seen[True] > 20 and seen[False] > 20
