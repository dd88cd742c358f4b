"""Independent skip-ahead of GSL's ranlxs2 stream, for checking windows far
into a draw without generating everything before them serially.

Test infrastructure.  The 48-bit subtract-with-borrow recurrence
x_n = x_{n-5} - x_{n-12} - c_{n-1} (mod b = 2^48), which ranlxs2 runs 397
times per block of 24 outputs, is a linear congruential generator in
disguise (Marsaglia & Zaman 1991; Tezuka, L'Ecuyer & Couture 1993): with the
state (x_{n-12}, ..., x_{n-1}; c_{n-1}) mapped to

    Y_n = P - Q + c_{n-1}  (mod m),   m = b^12 - b^5 + 1,
    P = sum_{j<12} x_{n-12+j} b^j,    Q = sum_{j>=7} x_{n-12+j} b^(j-7),

every step multiplies Y by b^-1 modulo m.  (Derivation: with
Z = b Y = sum_j (b^(j+1) - [j>=7] b^(j-6)) x_{n-12+j} + b c, the difference
b Z_{n+1} - Z_n reduces, using b^12 = b^5 - 1 mod m, to
b (x_{n-5} - x_{n-12} - c_{n-1} - x_n + b c_n) = 0.)

The arithmetic here is Python's arbitrary-precision integers (pow with a
negative exponent for the inverse), so it shares nothing with the HIP
library's 576-bit limb code.  A decoded state can differ from the true one
in its carry representation; the difference is gone after 12 steps, and the
serial generator continued from a block boundary runs 397 steps before it
emits anything.  tests/test_cpu_oracle.py pins all of this against the
oracle's serial stream.
"""
import numpy as np

B = 1 << 48
M = B ** 12 - B ** 5 + 1
LUX = 397
MASK48 = B - 1


def seed_words(seed):
    """The 12 initial 48-bit words of ranlxs seeded with `seed` (Luescher's
    31-bit shift register, clocked 48 times per word, from the seed's low 32
    bits; seed 0 means 1), as GSL's ranlxs_set_lux fills them."""
    if seed == 0:
        seed = 1
    i = seed & 0xFFFFFFFF
    # seeds at or above 2^31 turn negative in the reference's 32-bit int and
    # take C's signed remainders; the tests use smaller seeds
    assert i < 1 << 31, "seed_words: seeds below 2^31 only"
    bits = []
    for _ in range(31):
        bits.append(i % 2)
        i //= 2
    ib, jb = 0, 18
    words = []
    for _ in range(12):
        xv = 0
        for _ in range(48):
            xv += xv + bits[ib]
            bits[ib] = (bits[ib] + bits[jb]) % 2
            ib = (ib + 1) % 31
            jb = (jb + 1) % 31
        words.append(xv & MASK48)
    return words


def encode(words, carry):
    """Y of the state (words oldest first, carry)."""
    P = 0
    for j, w in enumerate(words):
        P += int(w) << (48 * j)
    Q = P >> (48 * 7)
    return (P - Q + int(carry)) % M


def decode(Y):
    """A state (words, carry = 0) whose Y is the given residue: P with
    P - floor(P / b^7) = Y, found by fixed-point iteration."""
    P = Y
    for _ in range(8):
        Pn = Y + (P >> (48 * 7))
        if Pn == P:
            break
        P = Pn
    assert P - (P >> (48 * 7)) == Y and P < B ** 12
    return [(P >> (48 * j)) & MASK48 for j in range(12)], 0


def state_at_block(seed, k):
    """(words, carry) at the block boundary before block k (k*397 steps from
    the seeded state): the serial generator continued from it emits blocks
    k, k+1, ... of the stream, 24 numbers each."""
    y0 = encode(seed_words(seed), 0)
    binv = pow(B, -1, M)
    yk = pow(binv, LUX * k, M) * y0 % M
    return decode(yk)


def gaussians_near(po, seed, g_index, mean_raw_per_gaussian, count, margin=20000):
    """Gaussians of the serial stream starting a little before the one of
    index ~g_index: the oracle's parse from a block boundary `margin`
    Gaussians early, with its first values dropped (they may belong to a
    parse that has not yet merged with the stream's).  Returns the values;
    the caller locates them in the stream under test."""
    raw0 = max(0, int((g_index - margin) * mean_raw_per_gaussian))
    k = raw0 // 24
    words, carry = state_at_block(seed, k)
    vals = po.gaussians_from_state(words, carry, 0, count + 2000)
    return vals[2000:]


def locate(window, probe, tol=1e-13, run=16):
    """Index in `window` where the first `run` values of `probe` appear
    (within tol), or -1."""
    head = probe[:run]
    cand = np.flatnonzero(np.abs(window[:len(window) - run + 1] - head[0]) <= tol)
    for i in cand:
        if np.all(np.abs(window[i:i + run] - head) <= tol):
            return int(i)
    return -1
