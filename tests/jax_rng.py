"""NumPy restatement of the JAX 0.8.0 PRNG used by the reference's tests.

The reference's tests draw their points with ``jax.random.normal(jax.random.key(137), shape)``
under ``jax_enable_x64`` (tests/test_tree.py:3,9,28; tests/test_graph.py:3,11,19;
tests/test_refine.py:13,21,48).  JAX is not installed here, so this module restates the
published algorithm of the pinned dependency ``jax==0.8.0`` (uv.lock:651-652):

* key = threefry seed: ``[seed >> 32, seed & 0xffffffff]`` (``jax._src.prng._threefry_seed``);
* random bits (``jax_threefry_partitionable=True``, the default since JAX 0.5): one
  Threefry-2x32 (20 rounds, Salmon et al. 2011) evaluation per element on the 64-bit
  row-major iota split into (hi, lo) words; 64-bit bits = ``hi_out << 32 | lo_out``,
  32-bit bits = ``hi_out ^ lo_out``;
* uniform: mantissa fill ``((bits >> (nbits - nmant)) | bits(1.0)) - 1``, scaled to
  ``[nextafter(-1, 0), 1)`` and clamped below;
* normal: ``sqrt(2) * erfinv(u)``.

It is test infrastructure only: the canary ``points[0, 0] == -0.61777326``
(tests/test_tree.py:29) pins it.
"""
import numpy as np

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))


def _rotl(x, r):
    return ((x << np.uint32(r)) | (x >> np.uint32(32 - r))).astype(np.uint32)


def threefry2x32(k1, k2, x0, x1):
    k1 = np.uint32(k1)
    k2 = np.uint32(k2)
    ks = (k1, k2, np.uint32(k1 ^ k2 ^ np.uint32(0x1BD11BDA)))
    x0 = (np.asarray(x0, dtype=np.uint32) + ks[0]).astype(np.uint32)
    x1 = (np.asarray(x1, dtype=np.uint32) + ks[1]).astype(np.uint32)
    with np.errstate(over="ignore"):
        for block in range(5):
            for r in _ROT[block % 2]:
                x0 = (x0 + x1).astype(np.uint32)
                x1 = _rotl(x1, r)
                x1 = (x0 ^ x1).astype(np.uint32)
            x0 = (x0 + ks[(block + 1) % 3]).astype(np.uint32)
            x1 = (x1 + ks[(block + 2) % 3] + np.uint32(block + 1)).astype(np.uint32)
    return x0, x1


def key(seed):
    seed = int(seed)
    return (np.uint32((seed >> 32) & 0xFFFFFFFF), np.uint32(seed & 0xFFFFFFFF))


def random_bits(k, bit_width, shape):
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.uint64)
    hi = (idx >> np.uint64(32)).astype(np.uint32)
    lo = (idx & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    b1, b2 = threefry2x32(k[0], k[1], hi, lo)
    if bit_width == 64:
        bits = (b1.astype(np.uint64) << np.uint64(32)) | b2.astype(np.uint64)
    elif bit_width == 32:
        bits = (b1 ^ b2).astype(np.uint32)
    else:
        raise ValueError(bit_width)
    return bits.reshape(shape)


def uniform(k, shape, dtype=np.float64, minval=0.0, maxval=1.0):
    dtype = np.dtype(dtype)
    finfo = np.finfo(dtype)
    nbits, nmant = finfo.bits, finfo.nmant
    uint = {32: np.uint32, 64: np.uint64}[nbits]
    bits = random_bits(k, nbits, shape).astype(uint)
    one_bits = np.array(1.0, dtype=dtype).view(uint)
    float_bits = (bits >> uint(nbits - nmant)) | one_bits
    floats = float_bits.view(dtype) - dtype.type(1.0)
    minval = dtype.type(minval)
    maxval = dtype.type(maxval)
    return np.maximum(minval, floats * (maxval - minval) + minval).astype(dtype)


def normal(k, shape, dtype=np.float64):
    from scipy.special import erfinv

    dtype = np.dtype(dtype)
    lo = np.nextafter(np.array(-1.0, dtype), np.array(0.0, dtype))
    u = uniform(k, shape, dtype, lo, 1.0)
    return (dtype.type(np.sqrt(2)) * erfinv(u)).astype(dtype)
