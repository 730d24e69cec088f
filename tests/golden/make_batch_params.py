"""Writes batch_taylor_params.csv: the 512-bar parameter table of the batched
Taylor workload (seed 20261203), drawn with an independent pure-Python
MT19937-64 (Matsumoto & Nishimura's 64-bit Mersenne Twister, the
std::mt19937_64 parameter set) in the documented order: nr ~ U{5..20},
v0 ~ U(150, 300) m/s, sigma_y0 ~ U(300, 500) MPa, H ~ U(50, 300) MPa.

Run: python tests/golden/make_batch_params.py
"""
import os

MASK = (1 << 64) - 1


class MT64:
    """MT19937-64: w=64, n=312, m=156, r=31, a=0xB5026F5AA96619E9,
    (u,d)=(29,0x5555555555555555), (s,b)=(17,0x71D67FFFEDA60000),
    (t,c)=(37,0xFFF7EEE000000000), l=43, f=6364136223846793005."""

    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK
        self.idx = 312

    def _twist(self):
        mt = self.mt
        up, lo = 0xFFFFFFFF80000000, 0x7FFFFFFF
        for i in range(312):
            x = (mt[i] & up) | (mt[(i + 1) % 312] & lo)
            xa = x >> 1
            if x & 1:
                xa ^= 0xB5026F5AA96619E9
            mt[i] = mt[(i + 156) % 312] ^ xa
        self.idx = 0

    def __call__(self):
        if self.idx >= 312:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK


def draw(n, seed=20261203):
    g = MT64(seed)

    def u(a, b):
        return a + (b - a) * ((g() >> 11) * (1.0 / 9007199254740992.0))
    rows = []
    for _ in range(n):
        nr = 5 + g() % 16
        v0 = u(150.0, 300.0)
        sy = u(300e6, 500e6)
        h = u(50e6, 300e6)
        rows.append((nr, v0, sy, h))
    return rows


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_taylor_params.csv")
    with open(out, "w") as f:
        f.write("bar,nr,nz,v0,sigma_y0,hardening_modulus\n")
        for i, (nr, v0, sy, h) in enumerate(draw(512)):
            f.write("%d,%d,%d,%.17g,%.17g,%.17g\n" % (i, nr, 5 * nr, v0, sy, h))
    print(out)
