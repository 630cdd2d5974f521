"""The join's distribution passes on the stride lattice (k_jp_dist_lattice in
velodb_amd/csrc/join_pairs.hip) compute what the per-output recursion
computes: levels j_top, j_top/2, .., u of the KKS move rule
(obliviator/join_kks_INT_INT/Include/sort.h:103-121), where each output p
reads the slots p - c u, c < 2^S.  A model of the kernel's block schedule
(256 residues, K rows plus the 2^S - 1 halo rows, levels applied top row
down) is run against the recursion on random targets; it also checks that
every access stays in [0, m_pad) and every slot is written exactly once.
CPU only: this pins the schedule the kernel implements, not the kernel."""
import random

import pytest


def _step(cur, prv, p, j):
    # an element moves from p - j to p when its target is >= p; the element at
    # p leaves when its target is >= p + j (None = empty slot)
    if prv is not None and prv >= p:
        return prv
    if cur is not None and cur >= p + j:
        return None
    return cur


def _recursive(a, s, p, j_top):
    if s == 0:
        return a[p] if p >= 0 else None
    j = j_top >> (s - 1)
    return _step(_recursive(a, s - 1, p, j_top), _recursive(a, s - 1, p - j, j_top), p, j)


def _lattice(a, m_pad, s, j_top, k_rows):
    u = j_top >> (s - 1)
    h = (1 << s) - 1
    w = k_rows + h
    out = [False] * m_pad
    res = [None] * m_pad
    res_blocks = u // 256
    for b in range(m_pad // (256 * k_rows)):
        for t in range(256):
            q = (b % res_blocks) * 256 + t
            k0 = (b // res_blocks) * k_rows - h
            v = []
            for r in range(w):
                k = k0 + r
                if k >= 0:
                    pp = q + k * u
                    assert 0 <= pp < m_pad
                    v.append(a[pp])
                else:
                    v.append(None)
            for lvl in range(s):
                st = 1 << (s - 1 - lvl)
                for r in range(w - 1, st - 1, -1):
                    v[r] = _step(v[r], v[r - st], q + (k0 + r) * u, st * u)
            for r in range(h, w):
                pp = q + (k0 + r) * u
                assert 0 <= pp < m_pad and not out[pp]
                out[pp] = True
                res[pp] = v[r]
    assert all(out)
    return res


@pytest.mark.parametrize("k_cap", [16, 32])
def test_lattice_schedule_equals_recursion(k_cap):
    rnd = random.Random(k_cap)
    for m_pad in (512, 1024, 2048):
        for s in (1, 2, 3):
            j_top = m_pad // 2
            while (j_top >> (s - 1)) >= 256:
                a = [None if rnd.random() < 0.3 else rnd.randrange(0, m_pad + 600) for _ in range(m_pad)]
                rows = m_pad // (j_top >> (s - 1))
                k_rows = min(k_cap, rows)
                want = [_recursive(a, s, p, j_top) for p in range(m_pad)]
                assert _lattice(a, m_pad, s, j_top, k_rows) == want, (m_pad, s, j_top)
                j_top //= 2
