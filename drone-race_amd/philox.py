"""Philox4x32-10 (Salmon et al., SC'11) in numpy, for host-side scenario
generation (gate jitter, sizes, seeds).  The same generator runs in the HIP
kernels (csrc/dr_rng.h); uniforms use the top 24 bits of a word."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)
STREAM_CONTROL, STREAM_TRACK, STREAM_BRANCH = 0, 1, 2


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32, key: (2,) uint32 -> (..., 4) uint32."""
    c = np.array(ctr, dtype=np.uint32)
    k0, k1 = np.uint32(key[0]), np.uint32(key[1])
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = np.uint32(k0 + W0)
                k1 = np.uint32(k1 + W1)
            p0 = M0 * c[..., 0].astype(np.uint64)
            p1 = M1 * c[..., 2].astype(np.uint64)
            n = np.empty_like(c)
            n[..., 0] = (p1 >> np.uint64(32)).astype(np.uint32) ^ c[..., 1] ^ k0
            n[..., 1] = (p1 & MASK).astype(np.uint32)
            n[..., 2] = (p0 >> np.uint64(32)).astype(np.uint32) ^ c[..., 3] ^ k1
            n[..., 3] = (p0 & MASK).astype(np.uint32)
            c = n
    return c


def uniforms(seed, n, stream=STREAM_TRACK, salt=0):
    """n uniforms in [0, 1) from counters (i // 4, salt, 0, stream)."""
    seed = int(seed)
    key = (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)
    m = (n + 3) // 4
    ctr = np.zeros((m, 4), dtype=np.uint32)
    ctr[:, 0] = np.arange(m, dtype=np.uint32)
    ctr[:, 1] = np.uint32(salt)
    ctr[:, 3] = np.uint32(stream)
    words = philox4x32_10(ctr, key).reshape(-1)[:n]
    return (words >> np.uint32(8)).astype(np.float64) * 2.0 ** -24


def normals(seed, n, stream=STREAM_BRANCH, salt=0):
    """n standard normals (Box-Muller in float64) for host-side draws."""
    m = (n + 1) // 2
    u = uniforms(seed, 2 * m, stream, salt)
    u1 = 1.0 - u[0::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u[1::2]), r * np.sin(2 * np.pi * u[1::2])])
    return z[:n]
