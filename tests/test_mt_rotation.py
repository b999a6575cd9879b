"""The generator facts the device's mt19937_64 handling rests on, checked on the CPU with a plain
restatement of the published algorithm (Matsumoto-Nishimura, 64-bit parameters; the C++ standard's
std::mt19937_64, [rand.eng.mers] / [rand.predef]):

- the lazy twist (one element regenerated as it is output, from the elements 1 and 156 places after
  it, mod 312) gives the batch twist's stream: the known answer of [rand.predef], the 10,000th
  output of a default-seeded engine, is 9981545732273789042;
- a state rotated together with its next index is the same generator (the realignment of the
  walkers' states before a launch of all walkers relies on it), at every next index."""
import random

N, M = 312, 156
MASK = (1 << 64) - 1


def seed_state(s):
    x = [s & MASK]
    for i in range(1, N):
        x.append((6364136223846793005 * (x[-1] ^ (x[-1] >> 62)) + i) & MASK)
    return x


def twist(xk, xk1, xkm):
    y = (xk & 0xFFFFFFFF80000000) | (xk1 & 0x7FFFFFFF)
    return xkm ^ (y >> 1) ^ (0xB5026F5AA96619E9 if y & 1 else 0)


def temper(z):
    z ^= (z >> 29) & 0x5555555555555555
    z ^= (z << 17) & 0x71D67FFFEDA60000
    z ^= (z << 37) & 0xFFF7EEE000000000
    z ^= z >> 43
    return z & MASK


class Lazy:
    """The device's form: element k regenerated when output (mci_device.hpp mt_next)."""

    def __init__(self, x, p=0):
        self.x, self.p = list(x), p

    def __call__(self):
        k = self.p
        nv = twist(self.x[k], self.x[(k + 1) % N], self.x[(k + M) % N])
        self.x[k] = nv
        self.p = (k + 1) % N
        return temper(nv)

    def rotated(self):
        return Lazy(self.x[self.p:] + self.x[:self.p], 0)


class Batch:
    """The textbook form: all 312 elements twisted in order when the index reaches 312."""

    def __init__(self, x):
        self.x, self.i = list(x), N

    def __call__(self):
        if self.i == N:
            for k in range(N):
                self.x[k] = twist(self.x[k], self.x[(k + 1) % N], self.x[(k + M) % N])
            self.i = 0
        z = self.x[self.i]
        self.i += 1
        return temper(z)


def test_lazy_twist_gives_the_standard_known_answer():
    g = Lazy(seed_state(5489))
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_lazy_twist_equals_batch_twist():
    for s in (1, 1337, 2**63 + 5):
        a, b = Lazy(seed_state(s)), Batch(seed_state(s))
        assert [a() for _ in range(1000)] == [b() for _ in range(1000)]


def test_rotated_state_is_the_same_generator_at_every_index():
    rng = random.Random(7)
    base = Lazy(seed_state(20240601))
    for _ in range(3 * N):
        base()
    for p in range(N):  # every next index, including the first and second halves of a round
        g = Lazy(base.x, p)
        h = g.rotated()
        assert h.p == 0
        k = rng.randrange(1, 2 * N)
        assert [g() for _ in range(k)] == [h() for _ in range(k)], p
