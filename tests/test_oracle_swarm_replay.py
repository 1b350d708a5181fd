"""Pins the oracle's batched swarm (the stream the GPU swarms are checked
against) by replaying it in plain Python on small cases (CPU only):

  * xoshiro128** against values worked by hand from the published algorithm
    (Blackman and Vigna, "Scrambled linear pseudorandom number generators");
  * a Python swarm written from the algorithm statement -- Philox initial
    swarm, per-particle xoshiro128** velocity stream seeded by the Philox block
    (band, particle, 0xFFFFFFFE, 0), inertia 0.9 -> 0.4, c1 = c2 = 1.5,
    velocity clamp to half the box, wall absorption, strict '<' bests in
    particle order, the reference's stall rule -- using the oracle's Philox
    and cost (both pinned elsewhere), must reproduce the oracle's parameters,
    cost and iteration count exactly for two bands of one batch.
"""
import math

import numpy as np
import pytest

MASK = 0xFFFFFFFF


class Xoshiro:
    def __init__(self, r):
        self.s = [int(x) & MASK for x in r]
        if not any(self.s):
            self.s[0] = 1

    @staticmethod
    def rotl(x, k):
        return ((x << k) | (x >> (32 - k))) & MASK

    def next(self):
        s = self.s
        out = (self.rotl((s[1] * 5) & MASK, 7) * 9) & MASK
        t = (s[1] << 9) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self.rotl(s[3], 11)
        return out


def test_xoshiro_worked_values():
    # state {1, 2, 3, 4}: out = rotl(2*5, 7)*9 = 11520; the update gives
    # {7, 0, 1026, 12288} (out 0), then {12295, 1029, 1029, 25165824}
    # (out rotl(1029*5, 7)*9 = 5927040)
    g = Xoshiro([1, 2, 3, 4])
    assert g.next() == 11520
    assert g.s == [7, 0, 1026, 12288]
    assert g.next() == 0
    assert g.s == [12295, 1029, 1029, 25165824]
    assert g.next() == 5927040
    z = Xoshiro([0, 0, 0, 0])
    assert z.s == [1, 0, 0, 0]


def u24(x):
    return float(int(x) >> 8) * (1.0 / 16777216.0)


def replay(orc, m, src, band, P, max_iters, stall_iters, seed):
    np_ = orc.n_params(m)
    lo, hi = orc.bounds(m)
    lo, hi = [float(x) for x in lo], [float(x) for x in hi]
    vmax = [0.5 * (hi[d] - lo[d]) for d in range(np_)]
    key = [seed & MASK, (seed >> 32) & MASK]
    pos = [[0.0] * np_ for _ in range(P)]
    vel = [[0.0] * np_ for _ in range(P)]
    for p in range(P):
        for k in range((np_ + 1) // 2):
            r = orc.philox([band, p, 0xFFFFFFFF, k], key)
            for h in range(2):
                d = 2 * k + h
                if d >= np_:
                    break
                pos[p][d] = lo[d] + u24(r[2 * h]) * (hi[d] - lo[d])
                vel[p][d] = vmax[d] * 0.2 * (2.0 * u24(r[2 * h + 1]) - 1.0)
    rng = [Xoshiro(orc.philox([band, p, 0xFFFFFFFE, 0], key)) for p in range(P)]
    pb = [[0.0] * np_ for _ in range(P)]
    pc = [math.inf] * P
    gb, gc = [0.0] * np_, math.inf
    prev, stall, done = math.inf, 0, 0
    inv_max = 1.0 / float(max_iters)
    obs = [src[k] for k in ("times", "flux", "obs_var", "is_upper", "upper_flux")]
    for it in range(max_iters):
        w = 0.9 - (0.9 - 0.4) * float(it) * inv_max
        cost = [orc.pso_cost(m, np.array(pos[p]), *obs) for p in range(P)]
        done = it + 1
        for p in range(P):
            if cost[p] < pc[p]:
                pc[p] = cost[p]
                pb[p] = list(pos[p])
                if cost[p] < gc:
                    gc = cost[p]
                    gb = list(pos[p])
        for p in range(P):
            for d in range(np_):
                r1, r2 = u24(rng[p].next()), u24(rng[p].next())
                x = pos[p][d]
                v = w * vel[p][d] + 1.5 * r1 * (pb[p][d] - x) + 1.5 * r2 * (gb[d] - x)
                v = min(max(v, -vmax[d]), vmax[d])
                nx = x + v
                if nx <= lo[d]:
                    pos[p][d], vel[p][d] = lo[d], 0.0
                elif nx >= hi[d]:
                    pos[p][d], vel[p][d] = hi[d], 0.0
                else:
                    pos[p][d], vel[p][d] = nx, v
        if prev - gc > 0.01 * max(abs(prev), 1e-10):
            stall, prev = 0, gc
        else:
            stall += 1
            if stall >= stall_iters:
                break
    return np.array(gb), gc, done


def small_batch():
    rng = np.random.default_rng(3)
    srcs = []
    for n in (14, 9):
        t = np.sort(rng.uniform(0.0, 60.0, n))
        f = np.exp(-(t - 20.0) / 15.0) / (1.0 + np.exp(-(t - 20.0) / 3.0)) + rng.normal(0.0, 0.03, n)
        up = np.zeros(n, np.int32)
        up[[0, 3]] = 1
        srcs.append({"times": t, "flux": f, "obs_var": np.full(n, 9e-4), "is_upper": up,
                     "upper_flux": np.full(n, 0.09)})
    off = np.array([0, srcs[0]["times"].size, srcs[0]["times"].size + srcs[1]["times"].size], np.int64)
    csr = {"offsets": off}
    for k in srcs[0]:
        csr[k] = np.concatenate([s[k] for s in srcs])
    return csr, srcs


@pytest.mark.parametrize("m", [0, 2])
def test_python_swarm_reproduces_oracle(orc, m):
    csr, srcs = small_batch()
    P, iters, stall, seed = 6, 12, 5, 0x1234_5678_9ABC
    prm, cost, it = orc.batch_pso(m, csr, P, iters, stall, seed)
    for b in range(2):
        gb, gc, done = replay(orc, m, srcs[b], b, P, iters, stall, seed)
        assert done == it[b], (b, done, it[b])
        assert gc == cost[b], (b, gc, cost[b])
        assert np.array_equal(gb, prm[b]), (b, gb, prm[b])
