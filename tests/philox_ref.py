"""Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) in plain Python, used
by the tests to generate the uniform stream the device swap consumes."""

M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> 32, p0 & MASK
        hi1, lo1 = p1 >> 32, p1 & MASK
        c = [(hi1 ^ c[1] ^ k0) & MASK, lo1, (hi0 ^ c[3] ^ k1) & MASK, lo0]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c


def uniforms(seed, counter, count):
    """The stream PhiloxStream(seed, counter).uniform() yields on the device."""
    k = (seed & MASK, (seed >> 32) & MASK)
    words = []
    ctr = counter
    while len(words) < 2 * count:
        words += philox4x32_10([ctr & MASK, (ctr >> 32) & MASK, 0x5EED0001, 0], k)
        ctr += 1
    out = []
    for i in range(count):
        x = (words[2 * i] << 32) | words[2 * i + 1]
        out.append((x >> 11) * (1.0 / 9007199254740992.0))
    return out
