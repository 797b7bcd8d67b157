"""Known-answer tests of the numeric building blocks shared by the oracle and
the GPU path: the RNG (native/common/barney-common.h:21-108) against a
Python-integer restatement, fp16 rounding against numpy.float16 (IEEE
round-to-nearest-even), and bnx_numeric.h sin/cos against libm."""
import ctypes as C

import numpy as np

M64 = (1 << 64) - 1


class PyRNG:
    """RNGSeed / Random2 re-derived from barney-common.h:21-67 with Python ints."""

    def __init__(self, state=0xCAFEF00DD15EA5E5):
        self.state = state

    def seed(self, a, b):
        self.state ^= (2 * a + 1) & 0xFFFFFFFF
        self.next()
        self.state ^= (2 * b) & 0xFFFFFFFF
        self.next()

    def next(self):
        x = self.state
        count = x >> 61
        self.state = (x * 6364136223846793005) & M64
        x ^= x >> 22
        return (x >> (22 + count)) & 0xFFFFFFFF

    def next_h(self, h):
        self.state ^= (h << 1) & M64
        return self.next()


def test_rng_sequences_match_closed_form(orc):
    L = orc.lib()
    for a, b in [(0, 0), (1, 2), (511, 0), (0xFFFFFFFF, 7), (123456, 654321), (0x80000000, 0x7FFFFFFF)]:
        out = (C.c_uint32 * 16)()
        L.orc_rng_seq(a, b, 16, out)
        r = PyRNG()
        r.seed(a, b)
        assert list(out) == [r.next() for _ in range(16)], (a, b)


def test_random2_child_stream(orc):
    L = orc.lib()
    for a, b, h in [(3, 4, 0), (3, 4, 0x3F800000), (100, 200, 0xDEADBEEF)]:
        out = (C.c_float * 8)()
        st = C.c_uint64()
        L.orc_random2_seq(a, b, h, 8, out, C.byref(st))
        p = PyRNG()
        p.seed(a, b)
        child = PyRNG(p.state)
        p.next()
        child.next_h(h if h else 290374)
        exp = [np.float32((child.next() & 0xFFFFFF) * (1.0 / 16777216.0)) for _ in range(8)]
        assert np.array_equal(np.array(list(out), np.float32), np.array(exp, np.float32))
        assert st.value == p.state


def test_fp16_all_halves_widen_exactly(orc):
    L = orc.lib()
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float32)
    got = np.array([L.orc_f16_to_f32(int(b)) for b in bits], dtype=np.float32)
    finite = np.isfinite(ref)
    assert np.array_equal(got[finite].view(np.uint32), ref[finite].view(np.uint32))
    assert np.all(np.isnan(got[~finite]) == np.isnan(ref[~finite]))


def test_fp16_rounding_is_rne(orc):
    L = orc.lib()
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(3.0),
        (rng.random(20000) * 70000).astype(np.float32),
        (rng.random(20000) * 1e-4).astype(np.float32),               # subnormal halves
        # exact ties between neighbouring halves
        ((np.arange(1, 4000, dtype=np.float64) + 0.5) * 2.0 ** -10).astype(np.float32),
        ((np.arange(1, 1000, dtype=np.float64) + 0.5) * 2.0 ** -24).astype(np.float32),
        np.array([65504.0, 65519.99, 65520.0, 1e9, -1e9, 0.0, -0.0, 6.0e-8, 2.9e-8, 3.0e-8], np.float32),
    ])
    ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_f32_to_f16(float(v)) for v in vals], dtype=np.uint16)
    bad = np.nonzero(got != ref)[0]
    assert len(bad) == 0, [(float(vals[i]), hex(got[i]), hex(ref[i])) for i in bad[:5]]


def _ulp_err(got, ref64):
    ref32 = ref64.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    ulp = np.maximum(ulp, np.float64(np.spacing(np.float32(1e-30))))
    return np.abs(got.astype(np.float64) - ref64) / ulp


def test_sincos_within_two_ulp_of_libm(orc):
    L = orc.lib()
    xs = np.concatenate([np.linspace(0.0, 2 * np.pi, 40001, dtype=np.float32),
                         np.linspace(-30.0, 30.0, 30001, dtype=np.float32)])
    s = np.zeros(len(xs), np.float32)
    c = np.zeros(len(xs), np.float32)
    fs, fc = C.c_float(), C.c_float()
    for i, x in enumerate(xs):
        L.orc_sincos(float(x), C.byref(fs), C.byref(fc))
        s[i], c[i] = fs.value, fc.value
    x64 = xs.astype(np.float64)
    es, ec = _ulp_err(s, np.sin(x64)), _ulp_err(c, np.cos(x64))
    # away from zeros of the function the error is a few ulp; near zeros the
    # absolute error is below 1e-7
    big_s, big_c = np.abs(np.sin(x64)) > 1e-3, np.abs(np.cos(x64)) > 1e-3
    assert es[big_s].max() <= 2.0 and ec[big_c].max() <= 2.0, (es[big_s].max(), ec[big_c].max())
    assert np.abs(s - np.sin(x64)).max() < 2e-7 and np.abs(c - np.cos(x64)).max() < 2e-7
