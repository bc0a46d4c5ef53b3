"""The error bound behind the RDF pass's fixed-point classification
(lennard-jones-cuda_amd/csrc/lj_kernels.hip, k_rdf_fx), checked on the CPU.

k_rdf_fx bins a pair from the fixed-point minimum image (q = rint(x 2^32 / L)
mod 2^32, the force pass's coordinates) whenever the interval
[r2q - delta, r2q + delta] lies inside one bin, and runs the reference
arithmetic only for the other pairs.  That is bit-exact only if delta bounds
|r2_ref S^2 - r2q| for every pair.  Here the reference arithmetic is emulated
in float32/float64 numpy, following the two rules:
  * CPU rule (MDSystem.cpp:269-279): float difference, image
    (float)(r - L fast_round((float)(r / L))) in double, r^2 summed in float
    without contraction;
  * GPU rule (MDSystem.cu:26-35): float difference, image d - Lf lroundf(d / Lf)
    in float, r^2 as fmaf(rz, rz, fmaf(ry, ry, rx rx)).
The pairs stress the bound's terms: separations near bin edges, near half a
box on an axis (the image choice), coordinates outside the box, near-coincident
particles, and several box sizes.  The bound's constants are the kernel's; the
second half checks the classification itself (r-bin fractions, table
intervals, the far-pair threshold) against the reference bins.
"""
import numpy as np
import pytest

U = 2.0 ** -24
F32 = np.float32


def fma32(a, b, c):
    # a*b is exact in float64 (24 + 24 bits); one rounding of the sum to
    # float64 then float32 (a double rounding can differ from fmaf by one ulp,
    # far inside the bound's slack)
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def fast_round(x):  # MDSystem.cpp:732-739, on float32
    return np.where(x > 0, np.trunc(x + F32(0.5)), np.trunc(x - F32(0.5))).astype(np.int64)


def ref_r2(pi, pj, L, rule):
    d = (pi - pj).astype(F32)  # float differences
    if rule == "cpu":
        q = (d.astype(np.float64) / L).astype(F32)
        k = fast_round(q)
        d = (d.astype(np.float64) - L * k).astype(F32)
        return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    Lf = F32(L)
    q = d / Lf
    k = np.where(q >= 0, np.floor(q + F32(0.5)), np.ceil(q - F32(0.5))).astype(F32)  # lroundf
    d = (d - Lf * k).astype(F32)
    return fma32(d[:, 2], d[:, 2], fma32(d[:, 1], d[:, 1], d[:, 0] * d[:, 0]))


def fixq(x, qs):
    return np.mod(np.rint(x.astype(np.float64) * qs), 2.0 ** 32).astype(np.uint64).astype(np.uint32)


def fixed_r2(pi, pj, qs):
    d = (fixq(pi, qs) - fixq(pj, qs)).view(np.int32).astype(F32)  # wraps modulo 2^32
    return fma32(d[:, 2], d[:, 2], fma32(d[:, 1], d[:, 1], d[:, 0] * d[:, 0]))


def bound_consts(A, L, qs):
    """k_rdf_fx's Kr, K2, K0 (fixed-point units) for the coordinate span A."""
    Kr = 1.01 * 3.4641016151377546 * (2.0 * A + 3.0 * L) * U * qs
    K2 = 1.01 * 16.0 * U
    K0 = 1.01 * (3.0 * U * U * (2.0 * A + 5.0 * L) ** 2 * qs * qs + Kr * Kr)
    return Kr, K2, K0


def make_pairs(rng, L, n, edges_r=None):
    """Pairs of float32 positions: uniform in the box, plus separations at
    given radii (bin edges), near half a box on an axis, outside the box and
    near-coincident."""
    parts = []
    pi = rng.uniform(0, L, (n, 3))
    parts.append((pi, rng.uniform(0, L, (n, 3))))
    if edges_r is not None:
        r = rng.choice(edges_r, n) * (1.0 + rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-8, -3, n))
        v = rng.normal(size=(n, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        parts.append((pi, np.mod(pi + r[:, None] * v, L)))
    pj = rng.uniform(0, L, (n, 3))
    ax = rng.integers(0, 3, n)
    pj[np.arange(n), ax] = np.mod(pi[np.arange(n), ax] + L / 2 + rng.normal(0, 1e-6 * L, n), L)
    parts.append((pi, pj))
    parts.append((rng.uniform(-0.2 * L, 1.2 * L, (n, 3)), rng.uniform(-0.2 * L, 1.2 * L, (n, 3))))
    parts.append((pi, pi + rng.normal(0, 1e-3, (n, 3))))
    a = np.concatenate([p[0] for p in parts]).astype(F32)
    b = np.concatenate([p[1] for p in parts]).astype(F32)
    return a, b


@pytest.mark.parametrize("L", [8.0, 26.842, 100.5])
@pytest.mark.parametrize("rule", ["cpu", "gpu"])
def test_fixed_point_r2_within_bound(L, rule):
    rng = np.random.default_rng(7)
    qs = 2.0 ** 32 / L
    pi, pj = make_pairs(rng, L, 40000, edges_r=np.linspace(0.05, 0.5, 64) * L)
    r2ref = ref_r2(pi, pj, L, rule).astype(np.float64) * qs * qs
    r2q = fixed_r2(pi, pj, qs).astype(np.float64)
    both = np.concatenate([pi, pj], axis=1).astype(np.float64)
    A = both.max(axis=1) - both.min(axis=1)  # the pair's coordinate span
    Kr, K2, K0 = bound_consts(A, L, qs)
    delta = Kr * np.sqrt(r2q) + K2 * r2q + K0
    err = np.abs(r2ref - r2q)
    # the image choice matters: pairs near half a box must be in the sample
    assert np.any(err > 0)
    worst = (err / delta).max()
    print(f"\n  L={L} rule={rule}: max |r2_ref - r2q| / delta = {worst:.3f}")
    assert worst <= 1.0
    # and the sample reaches the bound's scale (0.1-0.2 measured): delta is not
    # vacuous by orders of magnitude
    assert worst > 0.02


def ref_bins(r2, rule, nbins, dr2, rmax):
    """The reference bin of each r2 (-1: not counted; the GPU rule clamps)."""
    if rule == "rbins":
        dr = rmax / nbins
        e2 = ((np.arange(nbins + 1) * dr) ** 2).astype(F32)
        b = np.searchsorted(e2, r2, side="right") - 1
        return np.where((r2 < e2[-1]) & (b >= 0), b, -1)
    if rule == "cpu":
        b = np.floor(r2.astype(np.float64) / np.float64(F32(dr2))).astype(np.int64)
        return np.where(b < nbins, b, -1)
    q = (r2 / F32(dr2)).astype(F32)
    b = np.floor(q).astype(np.int64)
    return np.where(r2 > F32(1e-6), np.minimum(b, nbins - 1), -1)


@pytest.mark.parametrize("rule", ["rbins", "cpu", "gpu"])
def test_classification_matches_reference_bins(rule):
    """Every pair the kernel's classification decides lands in the reference
    bin; the undecided remainder is small."""
    rng = np.random.default_rng(3)
    L = 26.842
    qs = 2.0 ** 32 / L
    if rule == "rbins":
        nbins, rmax, dr2 = 500, L / 2, 0.0
        edges_r = np.arange(1, nbins + 1) * (rmax / nbins)
    else:
        nbins, rmax, dr2 = 256, 0.0, 0.1
        edges_r = np.sqrt(np.arange(1, nbins + 1) * dr2)
    pi, pj = make_pairs(rng, L, 40000, edges_r=edges_r)
    ok_box = (pi >= 0).all(1) & (pi < L).all(1) & (pj >= 0).all(1) & (pj < L).all(1)
    pi, pj = pi[ok_box], pj[ok_box]
    r2ref = ref_r2(pi, pj, L, "gpu" if rule == "gpu" else "cpu")
    want = ref_bins(r2ref, rule, nbins, dr2, rmax)
    r2q = fixed_r2(pi, pj, qs)
    A = L  # coordinates in [0, L)
    Kr, K2, K0 = bound_consts(A, L, qs)
    if rule == "rbins":
        es = F32(nbins / rmax / qs)
        M0 = F32(1.01 * (float(es) * Kr + float(es) ** 2 * K0))
        M1 = F32(1.01 * (K2 + 6 * U))
        x = (np.sqrt(r2q) * es).astype(F32)
        m = fma32(np.full_like(x, M1), x, np.full_like(x, M0))
        f = (x - np.floor(x)).astype(F32)
        k = x.astype(np.int64)
        dec = (x >= 1) & (f >= m) & (f <= F32(1) - m) & (k < nbins)
        gone = (x - m) >= F32(nbins)
        got_gone = want[gone]
        assert np.all(got_gone == -1)
    else:
        # the table path: T[k] = the least float r2 whose reference bin is >= k
        # (within a few ulps of k dr2: a short walk on the float grid)
        d32 = F32(dr2)
        if rule == "cpu":
            reached = lambda x, kk: np.floor(float(x) / float(d32)) >= kk  # noqa: E731
        else:
            reached = lambda x, kk: np.floor(x / d32) >= kk  # noqa: E731
        T = np.zeros(nbins + 1, dtype=F32)
        for kk in range(1, nbins + 1):
            x = F32(kk * float(d32))
            while reached(x, kk):
                x = np.nextafter(x, F32(0))
            while not reached(x, kk):
                x = np.nextafter(x, F32(np.inf))
            T[kk] = x
        if rule == "gpu":
            T[nbins] = np.inf
        Tend = float(T[nbins - 1] if rule == "gpu" else T[nbins])
        S2 = qs * qs
        T2 = (T.astype(np.float64) * S2).astype(F32)
        r = (np.frombuffer(((r2q.view(np.uint32) >> 1) + np.uint32(0x1FC00000)).tobytes(), dtype=F32))
        assert np.all(r.astype(np.float64) >= np.sqrt(r2q.astype(np.float64)) * (1 - 1e-7))
        dl = fma32(np.full_like(r2q, F32(Kr)), r, fma32(np.full_like(r2q, F32(K2)), r2q,
                                                        np.full_like(r2q, F32(K0))))
        lo = (r2q - dl).astype(F32)
        hi = (r2q + dl).astype(F32)
        es = F32(1.0 / dr2 / S2)
        k = np.clip((r2q * es).astype(F32), 0, nbins - 1).astype(np.int64)
        gone = lo >= F32(Tend * S2)
        dec = (T2[k] <= lo) & (hi < T2[k + 1])
        if rule == "gpu":
            dec &= (lo > F32(1e-6 * S2)) & (k < nbins - 1)
            assert np.all(want[gone] == nbins - 1)
        else:
            assert np.all(want[gone] == -1)
    assert np.all(want[dec] == k[dec])
    und = ~dec & ~gone
    print(f"\n  {rule}: {dec.sum()} decided, {gone.sum()} beyond, {und.sum()} to the exact path "
          f"({und.mean():.2e} of {und.size})")
    # the near-edge half of the sample is built to sit on bin edges; the
    # uniform pairs are decided almost always
    n_uni = 40000
    assert und[:n_uni].mean() < 0.01
