"""Inputs and output layout of the reference known-answer driver (oracle/ref_kat.f90).

The inputs are a deterministic SplitMix64 stream, so only the reference's
outputs need to be stored as a fixture.  The compiled reference uses its own
compile-time parameters (program/parameters.f90: K0=4, MM=NN=32, Lx=12, Lz=6,
dt=0.01, Re=200, drag=1e-2, dragall), recorded in the output header.
"""
import numpy as np

K0, MM, NN = 4, 32, 32
K, KK, M = 2 * K0 - 1, 3 * K0 - 1, 2 * MM - 1
NV = 64
SEED = 0x4D5746


def splitmix_uniform(seed, count):
    """count doubles uniform in [-1, 1) from SplitMix64(seed)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -52) - 1.0


def inputs():
    """A (mpt), S (spec), V, VX, VZ in reference layout (C-order, reversed axes)."""
    nm, ns = K * M * NN, KK * M * NN
    r = splitmix_uniform(SEED, 2 * nm + 2 * ns + 3 * KK * NV)
    o = 0
    def take(n, shape):
        nonlocal o
        a = r[o:o + n].reshape(shape)
        o += n
        return a
    Are, Aim = take(nm, (NN, M, K)), take(nm, (NN, M, K))
    Sre, Sim = take(ns, (NN, M, KK)), take(ns, (NN, M, KK))
    V, VX, VZ = take(KK * NV, (NV, KK)), take(KK * NV, (NV, KK)), take(KK * NV, (NV, KK))
    return Are, Aim, Sre, Sim, V, VX, VZ


def write_inputs(path):
    with open(path, "wb") as f:
        for a in inputs():
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def read_outputs(path):
    d = np.fromfile(path, dtype="<f8")
    nm, ns = K * M * NN, KK * M * NN
    out, o = {}, 0
    def take(name, n, shape):
        nonlocal o
        out[name] = d[o:o + n].reshape(shape)
        o += n
    take("hdr", 12, (12,))
    take("lhs", nm, (NN, M, K)); take("rhs", nm, (NN, M, K))
    take("m2s_re", ns, (NN, M, KK)); take("m2s_im", ns, (NN, M, KK))
    take("ux_re", ns, (NN, M, KK)); take("ux_im", ns, (NN, M, KK))
    take("uz_re", ns, (NN, M, KK)); take("uz_im", ns, (NN, M, KK))
    take("s2m_re", nm, (NN, M, K)); take("s2m_im", nm, (NN, M, K))
    per = [("udotgradu", KK), ("udotgradu2", KK), ("nluw", K0), ("nlvv", K0),
           ("nluv", K0 - 1), ("nluyv", K0), ("nlvvy", K0 - 1)]
    stride = sum(n for _, n in per)
    blk = d[o:o + NV * stride].reshape(NV, stride)
    c = 0
    for name, n in per:
        out[name] = blk[:, c:c + n].copy()
        c += n
    o += NV * stride
    assert o == d.size, (o, d.size)
    return out


# Fixture keeps n < NSUB of the per-mode arrays (every (m, k), and the (0,0)
# special cases) plus everything per vector.
NSUB = 8
