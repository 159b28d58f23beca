"""MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998) and libstdc++'s
uniform_real_distribution<double> on it, to regenerate the inputs of the reference's seeded tests
(tests/ssdf_brute_force_test.cpp:11-31 uses std::mt19937 rng(1) and U(-1, 1))."""


class MT19937:
    def __init__(self, seed: int = 5489):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.idx = 624

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            v = mt[(i + 397) % 624] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= 624:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def generate_canonical_double(rng: MT19937) -> float:
    """std::generate_canonical<double, 53> over a 32-bit engine: two draws, sum r1 + r2 2^32 rounded
    to double, divided by 2^64 (values >= 1 become the largest double below 1)."""
    s = float(rng()) + float(rng()) * 4294967296.0
    r = s / 18446744073709551616.0
    if r >= 1.0:
        r = 0.9999999999999999
    return r


def uniform(rng: MT19937, a: float, b: float) -> float:
    return generate_canonical_double(rng) * (b - a) + a
