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


# ---------------------------------------------------------------------------
# The CPU path's asynchronous swarm (oracle/cpu_pso.cpp), replayed from the
# reference's pso_minimize / pso_minimize_seeded (src/parametric.rs:1160-1266,
# 1646-1757): particles initialised in order (particle 0 at the seed when
# seeded; every particle draws its velocity), then per iteration inertia
# 0.9 - 0.5 it / max_iters, per particle and dimension r1, r2, velocity
# clamp to half the box, wall absorption, the cost, the personal best and --
# immediately, before the next particle moves -- the global best (strict
# '<'); the stall rule with the initial best as baseline.  Random numbers:
# the job's Philox stream (counter (stream, model | kind << 8 | fixed_t0 <<
# 16, slot, particle x 32 + d / 2), slot 0 for the initial swarm and it + 1
# for iteration it; odd dimensions take the block's upper pair).  Costs come
# from the oracle's PsoCost (pinned against a sequential brute force in
# test_oracle_cpu_pso.py), evaluated one point at a time.
# ---------------------------------------------------------------------------
def _cpso_draw(orc, seed, stream, tag, slot, p, d):
    r = orc.philox([stream, tag, slot, p * 32 + (d >> 1)], [seed & MASK, (seed >> 32) & MASK])
    return (u24(r[2]), u24(r[3])) if d & 1 else (u24(r[0]), u24(r[1]))


def replay_async(orc, cost, lo, hi, P, max_iters, stall_iters, seed, tag, x0=None, stream=0):
    D = len(lo)
    vmax = [0.5 * (hi[d] - lo[d]) for d in range(D)]
    pos, vel, pb, pc = [], [], [], []
    gb, gc = [0.0] * D, math.inf
    for p in range(P):
        x, v = [0.0] * D, [0.0] * D
        for d in range(D):
            u1, u2 = _cpso_draw(orc, seed, stream, tag, 0, p, d)
            x[d] = x0[d] if (x0 is not None and p == 0) else lo[d] + u1 * (hi[d] - lo[d])
            v[d] = vmax[d] * 0.2 * (2.0 * u2 - 1.0)
        c = cost(x)
        pos.append(x), vel.append(v), pb.append(list(x)), pc.append(c)
        if c < gc:
            gc, gb = c, list(x)
    prev, stall, it = gc, 0, 0
    while it < max_iters:
        w = 0.9 - (0.9 - 0.4) * float(it) * (1.0 / float(max_iters))
        for p in range(P):
            for d in range(D):
                u1, u2 = _cpso_draw(orc, seed, stream, tag, it + 1, p, d)
                x = pos[p][d]
                v = w * vel[p][d] + 1.5 * u1 * (pb[p][d] - x) + 1.5 * u2 * (gb[d] - x)
                v = min(max(v, -vmax[d]), vmax[d])
                nx = x + v
                if nx <= lo[d]:
                    pos[p][d], vel[p][d] = lo[d], 0.0
                elif nx >= hi[d]:
                    pos[p][d], vel[p][d] = hi[d], 0.0
                else:
                    pos[p][d], vel[p][d] = nx, v
            c = cost(pos[p])
            if c < pc[p]:
                pc[p], pb[p] = c, list(pos[p])
                if c < gc:  # asynchronous: the next particle already sees it
                    gc, gb = c, list(pos[p])
        it += 1
        if prev - gc > 0.01 * max(abs(prev), 1e-10):
            stall, prev = 0, gc
        else:
            stall += 1
            if stall >= stall_iters:
                break
    return np.array(gb), gc, it


@pytest.mark.parametrize("m,seeded", [(0, False), (0, True), (3, False), (1, True)])
def test_python_async_swarm_reproduces_oracle(orc, m, seeded):
    from test_oracle_cpu_pso import evaluate, job

    csr, srcs = small_batch()
    band = {"offsets": np.array([0, srcs[0]["times"].size], np.int64), **srcs[0]}
    lo, hi = orc.bounds(m, cpu=True)
    lo, hi = [float(x) for x in lo], [float(x) for x in hi]
    x0 = [0.5 * (a + b) for a, b in zip(lo, hi)] if seeded else None
    P, iters, stall, seed = 7, 9, 4, 0xABCDEF012345
    j = job(orc, kind=orc.CPSO_SWARM_BAND, model=m, band=0, dim=len(lo), n_particles=P, max_iters=iters,
            stall_iters=stall, use_prior=1, seeded=int(seeded), lo=lo, hi=hi, x0=x0 if seeded else lo, seed=seed)
    o = orc.cpu_pso(band, np.array([j]))
    cost = lambda x: float(evaluate(orc, band, m, np.array(x), prior=1))
    gb, gc, it = replay_async(orc, cost, lo, hi, P, iters, stall, seed, tag=m, x0=x0)
    assert it == o["iters"][0]
    assert gc == o["cost"][0]
    assert np.array_equal(gb, o["x"][0, :len(lo)])


def test_synchronous_update_would_differ(orc):
    """The replay is sensitive to the update order: a synchronous global
    best (updated only after all particles moved, as the GPU path does)
    takes another trajectory on the same stream."""
    from test_oracle_cpu_pso import evaluate, job

    csr, srcs = small_batch()
    band = {"offsets": np.array([0, srcs[0]["times"].size], np.int64), **srcs[0]}
    m = 0
    lo, hi = orc.bounds(m, cpu=True)
    lo, hi = [float(x) for x in lo], [float(x) for x in hi]
    P, iters, stall, seed = 7, 9, 9, 99
    j = job(orc, kind=orc.CPSO_SWARM_BAND, model=m, band=0, dim=len(lo), n_particles=P, max_iters=iters,
            stall_iters=stall, use_prior=1, lo=lo, hi=hi, seed=seed)
    o = orc.cpu_pso(band, np.array([j]))
    cost = lambda x: float(evaluate(orc, band, m, np.array(x), prior=1))

    def replay_sync():
        D = len(lo)
        vmax = [0.5 * (hi[d] - lo[d]) for d in range(D)]
        pos, vel, pb, pc = [], [], [], []
        gb, gc = [0.0] * D, math.inf
        for p in range(P):
            x, v = [0.0] * D, [0.0] * D
            for d in range(D):
                u1, u2 = _cpso_draw(orc, seed, 0, m, 0, p, d)
                x[d], v[d] = lo[d] + u1 * (hi[d] - lo[d]), vmax[d] * 0.2 * (2.0 * u2 - 1.0)
            c = cost(x)
            pos.append(x), vel.append(v), pb.append(list(x)), pc.append(c)
            if c < gc:
                gc, gb = c, list(x)
        for it in range(iters):
            w = 0.9 - 0.5 * it / iters
            snap = list(gb)
            for p in range(P):
                for d in range(D):
                    u1, u2 = _cpso_draw(orc, seed, 0, m, it + 1, p, d)
                    x = pos[p][d]
                    v = min(max(w * vel[p][d] + 1.5 * u1 * (pb[p][d] - x) + 1.5 * u2 * (snap[d] - x), -vmax[d]),
                            vmax[d])
                    nx = x + v
                    pos[p][d], vel[p][d] = (lo[d], 0.0) if nx <= lo[d] else ((hi[d], 0.0) if nx >= hi[d] else (nx, v))
                c = cost(pos[p])
                if c < pc[p]:
                    pc[p], pb[p] = c, list(pos[p])
            k = int(np.argmin(pc))
            if pc[k] < gc:
                gc, gb = pc[k], list(pb[k])
        return gc

    assert replay_sync() != o["cost"][0]
