"""Per-sample parameter generation (alsuq/samples, generator, distribution).

Two generators:
  * STLMersenne (the reference's `stlmersenne`/`auto`, STLMersenne.cpp:25-85):
    one default-seeded std::mt19937_64 shared by every parameter; the value of
    (sample s, component c) of a parameter of dimension D is drawn after
    discarding forward to draw index s*D + c, through
    std::uniform_real_distribution<double>(0, 1) (STLMersenne.hpp:37).
  * seeded: sample i draws from its own std::mt19937_64(seed + i).
Uniform scaling is x*(upper-lower)+lower (Uniform.cpp:27-34).

libstdc++'s uniform_real_distribution<double> over a 64-bit engine takes one
draw x and returns double(x) / 2^64 (round-to-nearest conversion), clamped to
the largest double below 1 (bits/random.tcc, generate_canonical).
"""
import numpy as np

_MASK = (1 << 64) - 1
_NN, _MM = 312, 156
_MATRIX_A = 0xB5026F5AA96619E9
_UM, _LM = 0xFFFFFFFF80000000, 0x7FFFFFFF


class MT19937_64:
    """std::mt19937_64 (C++ [rand.eng.mers] with the mt19937_64 parameters)."""

    def __init__(self, seed=5489):
        mt = np.zeros(_NN, dtype=np.uint64)
        x = seed & _MASK
        mt[0] = x
        for i in range(1, _NN):
            x = (6364136223846793005 * (x ^ (x >> 62)) + i) & _MASK
            mt[i] = x
        self.mt = mt
        self.idx = _NN

    def _twist(self):
        mt = self.mt
        um, lm, a = np.uint64(_UM), np.uint64(_LM), np.uint64(_MATRIX_A)
        one = np.uint64(1)
        # Blocks that only read untouched words can be vectorized.
        y = (mt[: _NN - _MM] & um) | (mt[1: _NN - _MM + 1] & lm)
        mt[: _NN - _MM] = mt[_MM:_NN] ^ (y >> one) ^ np.where((y & one) != 0, a, np.uint64(0))
        for i in range(_NN - _MM, _NN - 1, _MM):
            j = min(i + _MM, _NN - 1)
            y = (mt[i:j] & um) | (mt[i + 1:j + 1] & lm)
            mt[i:j] = mt[i + _MM - _NN:j + _MM - _NN] ^ (y >> one) ^ np.where((y & one) != 0, a, np.uint64(0))
        y = (mt[_NN - 1] & um) | (mt[0] & lm)
        mt[_NN - 1] = mt[_MM - 1] ^ (y >> one) ^ (a if int(y) & 1 else np.uint64(0))
        self.idx = 0

    def raw(self, n):
        """Next n outputs as uint64."""
        out = np.empty(n, dtype=np.uint64)
        k = 0
        while k < n:
            if self.idx >= _NN:
                self._twist()
            take = min(n - k, _NN - self.idx)
            y = self.mt[self.idx:self.idx + take].copy()
            y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
            y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
            y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
            y ^= y >> np.uint64(43)
            out[k:k + take] = y
            self.idx += take
            k += take
        return out

    def discard(self, n):
        while n > 0:
            if self.idx >= _NN:
                self._twist()
            take = min(n, _NN - self.idx)
            self.idx += take
            n -= take


_BELOW_ONE = np.nextafter(1.0, 0.0)


def canonical(raw):
    """uniform_real_distribution<double>(0,1) values from raw 64-bit draws."""
    # Python int -> float is correctly rounded; numpy's uint64 -> float64 too.
    u = raw.astype(np.float64) / 18446744073709551616.0
    return np.where(u >= 1.0, _BELOW_ONE, u)


def stl_mersenne_uniform(nsamples, dim, lower=0.0, upper=1.0, sample_start=0):
    """(nsamples, dim) values as the reference Runner draws them for one
    parameter (samples sample_start, sample_start+1, ... in order)."""
    eng = MT19937_64()
    count = 0
    out = np.empty((nsamples, dim))
    for i in range(nsamples):
        s = sample_start + i
        if count // dim < s:
            eng.discard(s * dim - count)
            count = s * dim
        x = canonical(eng.raw(dim))
        count += dim
        out[i] = x * (upper - lower) + lower
    return out


def seeded_uniform(seed, nsamples, dim, lower=0.0, upper=1.0):
    """Sample i drawn from std::mt19937_64(seed + i)."""
    out = np.empty((nsamples, dim))
    for i in range(nsamples):
        x = canonical(MT19937_64(seed + i).raw(dim))
        out[i] = x * (upper - lower) + lower
    return out
