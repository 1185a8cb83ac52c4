"""Synthetic workloads of the batched and served configurations.

Every draw comes from one splitmix64 stream (state += 0x9E3779B97F4A7C15, then
the Stafford-13 finalizer), so a workload is fixed by its start state.
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, state):
        self.state = state & MASK64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def signed(self):
        v = self.next()
        return v - (1 << 64) if v >> 63 else v

    def uniform(self):
        """Float in [0, 1) from the top 53 bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


BATCH_STATE = 0x5EED000000000003
SERVED_STATE = 0x5EED000000000005
FULL_WORLD_RADIUS = 1_875_000


def batched_searches(n=256, state=BATCH_STATE, r_lo=256, r_hi=200_000, center=1_800_000,
                     thr_lo=50, thr_hi=62):
    """n searches: (seed, rd_max, min_size, max_size, rd_min, cx, cz) with the
    radius log-uniform in [r_lo, r_hi], the center uniform in +-center and the
    threshold uniform in [thr_lo, thr_hi]."""
    g = SplitMix64(state)
    out = []
    for _ in range(n):
        seed = g.signed()
        rd = int(round(math.exp(math.log(r_lo) + g.uniform() * (math.log(r_hi) - math.log(r_lo)))))
        cx = g.randint(-center, center)
        cz = g.randint(-center, center)
        thr = g.randint(thr_lo, thr_hi)
        out.append((seed, rd, thr, 221, 0, cx, cz))
    return out


def served_requests(n, rate_per_s, state=SERVED_STATE, full_world=True, small_only=False):
    """n requests of the served mix with Poisson arrivals: dicts with the
    search, its arrival time (s) and whether it pauses (then resumes) or
    cancels. Mix: 70% radius uniform in [256, 4096], 25% radius 50,000,
    5% full world (radius 1,875,000 around 0,0; with full_world False the 5%
    are radius 200,000; with small_only every request is of the first kind);
    thresholds uniform in [50, 65]; 10% pause/resume, 5% cancel."""
    g = SplitMix64(state)
    t = 0.0
    out = []
    for i in range(n):
        t += -math.log(1.0 - g.uniform()) / rate_per_s
        u = g.uniform()
        if u < 0.70 or small_only:
            rd = g.randint(256, 4096)
            cx, cz = g.randint(-1_800_000, 1_800_000), g.randint(-1_800_000, 1_800_000)
        elif u < 0.95:
            rd = 50_000
            cx, cz = g.randint(-1_800_000, 1_800_000), g.randint(-1_800_000, 1_800_000)
        else:
            rd = FULL_WORLD_RADIUS if full_world else 200_000
            cx = cz = 0
        thr = g.randint(50, 65)
        a = g.uniform()
        action = "pause" if a < 0.10 else ("cancel" if a < 0.15 else None)
        out.append({"search": (g.signed(), rd, thr, 221, 0, cx, cz), "arrival": t,
                    "action": action, "action_at": g.uniform()})
    return out
