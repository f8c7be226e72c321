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


# ---------------------------------------------------------------------------
# Floating-point error bounds.  A sum computed by any evaluation tree whose
# longest leaf-to-root path holds d rounded additions satisfies
# |computed - exact| <= gamma_d * sum|x|, gamma_d = d u / (1 - d u)
# (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 4.2).
# The depths below are the kernels' own, for the library's default launch
# geometry.
# ---------------------------------------------------------------------------
def gamma(d, eps):
    return d * eps / (1.0 - d * eps)


def reduce_depth(n, cus):
    """Longest rounding chain of the device-wide reduction of n elements.

    The default geometry runs one work-group per CU with at least 2048
    accumulators each (512 work-items x 4 loads, 1024 x 8 for dot products,
    1024 x 2 for compensated ones), so at most ceil(n / (2048 cus)) vectors'
    worth of elements reach one accumulator in the main loop; leftover
    vectors (<= 3 x 8 elements) and one scalar element can join it.  Then
    come the accumulator combine (<= 7), the DPP wave reduce (6), the
    cross-wave step (4), the last group's fold of the partials (<= 1 + 6 +
    4), init (1) and, for dot products, the product's own rounding (1)."""
    return -(-n // (2048 * cus)) + 8 * 4 + 1 + 31


def scan_depth(i, tile_elems):
    """Longest rounding chain behind element i of a decoupled look-back
    scan whose tiles hold at least tile_elems elements: each tile's
    inclusive prefix is op(exclusive prefix, aggregate), so the chain
    through the prefixes of earlier tiles grows by at most one per tile,
    plus the in-tile depth (vector 3, rows 8, wave 6, (row, wave) totals
    7, aggregate combine 2, final op 1) and the look-back's combine of up
    to 64 aggregates (6 when paired, taken as 64 here)."""
    return np.asarray(i) // tile_elems + 1 + 27 + 64


def fold_bound(abs_sum, depth, eps):
    return gamma(depth, eps) * abs_sum


def compensated_bound(exact, abs_sum, n, cus):
    """Compensated fp32 sum / Dot2: the (sum, error) pairs keep the running
    error of every addition exactly (TwoSum, TwoProduct by FMA); what is
    lost is the rounding of the fp32 error-term accumulation, at most
    gamma_d^2 sum|x| (Ogita, Rump, Oishi 2005, Prop. 4.5), the FP64
    cross-group fold and the final rounding to fp32."""
    d = reduce_depth(n, cus)
    return (2.0 ** -24 * abs(exact) + gamma(d, 2.0 ** -24) ** 2 * abs_sum
            + gamma(d, 2.0 ** -53) * abs_sum)


def exact_ints(torch, n, seed, lo=1, hi=8):
    """Non-zero integers in +-[lo, hi] on the GPU (int32): float32 and bf16
    hold them exactly, so sums of them are exact as long as every partial
    sum stays below 2^24 in magnitude; with random signs a partial sum over
    m elements is of order sqrt(m) * hi, far below that at 2^28 elements.
    A lost element shifts the sum by at least lo."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    mag = torch.randint(lo, hi + 1, (n,), generator=g, device="cuda", dtype=torch.int32)
    sign = torch.randint(0, 2, (n,), generator=g, device="cuda", dtype=torch.int32) * 2 - 1
    return mag * sign
