"""Synthetic inputs: the host twin of gt355_fill_uniform.

Counter-based SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): element n of a field is
mix(base + (n+1) * 0x9E3779B97F4A7C15) with base = seed ^ (field_id * 0xD1B54A32D192ED03), mapped to
U(a, b) as a + (b - a) * ((x >> 11) * 2**-53) in FP64.  n runs over the logical box, i fastest, then j, then k,
halo included -- the reference initializer fills every index of the allocation (storage/builder.hpp:67-84).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
FIELD_MUL = np.uint64(0xD1B54A32D192ED03)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

# field ids of the benchmark workloads
FIELD_IDS = {"in": 1, "coeff": 2, "u_stage": 3, "u_pos": 4, "utens": 5, "utens_stage": 6, "wcon": 7}


def uniform(lengths, seed, field_id, a, b, dtype=np.float64):
    """Returns an array indexed [i, j, k] of the given lengths."""
    ni, nj, nk = lengths
    n = np.arange(1, ni * nj * nk + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = np.uint64(seed) ^ (np.uint64(field_id) * FIELD_MUL)
        z = base + n * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    v = a + (b - a) * u
    return v.reshape(nk, nj, ni).transpose(2, 1, 0).astype(dtype)
