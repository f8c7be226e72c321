"""Helpers for the GPU parity tests."""
import numpy as np

INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]
FP_TYPES = [np.float32, np.float64]
ALL_OPS = list(range(9))
FP_OPS = [0, 1, 2, 3]


def dev(torch, a: np.ndarray):
    """Copy a numpy array to the GPU (contiguous)."""
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def empty_like_dev(torch, a: np.ndarray, n=None):
    return torch.empty(a.size if n is None else n, dtype=torch.from_numpy(a[:0]).dtype,
                       device="cuda")


def rand(dtype, n, seed=0, lo=None, hi=None):
    rng = np.random.default_rng(seed)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        lo = info.min if lo is None else lo
        hi = info.max if hi is None else hi
        return rng.integers(lo, hi, size=n, dtype=dtype, endpoint=True)
    return rng.standard_normal(n).astype(dtype)


def op_input(dtype, op, n, seed):
    """Inputs that keep each operator's result informative."""
    x = rand(dtype, n, seed)
    if op == 1 and np.issubdtype(dtype, np.floating):  # products: values near 1
        x = (1.0 + (rand(np.float64, n, seed) * 1e-3)).astype(dtype)
    if op in (7, 8):  # logical: a mix of zeros and non-zeros
        x[np.random.default_rng(seed + 1).random(n) < (0.999 if op == 8 else 0.001)] = 0
    if op == 4:  # bit_and: mostly all-ones so the result is not trivially zero
        x = np.bitwise_or(x, ~np.array(1 << (np.random.default_rng(seed).integers(0, 7)),
                                       dtype=dtype))
    return x


def zipf_keys(n, num_bins, s=1.1, seed=0):
    """Bounded Zipf(s) keys on {0..num_bins-1} by inverse CDF
    (p_k proportional to (k+1)^-s)."""
    rng = np.random.default_rng(seed)
    p = np.arange(1, num_bins + 1, dtype=np.float64) ** -s
    cdf = np.cumsum(p)
    cdf /= cdf[-1]
    u = rng.random(n)
    return np.minimum(np.searchsorted(cdf, u, side="right"), num_bins - 1).astype(np.uint32)


def lognormal_offsets(num_segments, median=512, sigma=2.0, max_len=1 << 20, seed=0, min_len=1):
    """Segment offsets with log-normal lengths clip(round(exp(N(ln median, sigma))))."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.round(np.exp(rng.normal(np.log(median), sigma, num_segments))),
                   min_len, max_len).astype(np.uint64)
    return np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
