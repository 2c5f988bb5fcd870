"""CPU checks of the cube input path: the shared decimal -> binary64
conversion (critic2_amd/csrc/number.hpp, the code the GPU parser runs) built
for the host and compared bit for bit with Python's float(); and the oracle's
cube readers against the reference's own cube files."""
import glob
import os
import random
import shutil
import struct
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

from tests.cubetext import bincube, expected_grid, fortran_e, fortran_float
from tests.helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CUBES = sorted(glob.glob(os.path.join(GOLDEN, "cube", "*.cube")))


@pytest.fixture(scope="module")
def number_check(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("num") / "number_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-x", "c++",
                    os.path.join(ROOT, "tests", "number_check.cpp"), "-o", exe], check=True)

    def run(tokens, mode=None):
        args = [exe] + ([mode] if mode else [])
        out = subprocess.run(args, input="\n".join(tokens) + "\n", capture_output=True, text=True, check=True)
        if mode == "fmt":
            return [l[1:-1] for l in out.stdout.splitlines()]
        return out.stdout.split()
    return run


def _bits(x: float) -> str:
    return struct.pack(">d", x).hex()


EDGE = [
    "0", "-0.0", "+0.0E+000", "1", "1.0", "0.1", "1e23", "8.98846567431158e307", "1.7976931348623157e308",
    "1.7976931348623158e308", "1.797693134862315807e308", "1.7976931348623159e308", "2e308", "1e99999",
    "1e-99999", "0e99999", "4.9406564584124654e-324", "2.4703282292062327e-324", "2.4703282292062328e-324",
    "2.2250738585072011e-308", "2.2250738585072012e-308", "2.2250738585072014e-308", "2.2250738585072009e-308",
    "9007199254740993", "9007199254740995", "9007199254740992.5", "9007199254740993.0000000001",
    "123456789012345678901234567890", "0.000000000000000000000000000000000000001", "1.0D-05", "1.0d+05",
    "-1.234E-005", "0.21032651859020E+001", "0.38636354665094E-016", "1e-400", "1e400", "1.5-3", "2.5+2",
    "3.0q2", "5e-324", "3e-324", "2.5e-324", "2.4e-324", "1e-323", "1.e5", ".5", "5.", "00012.500E-0001",
    "7.2057594037927933e16", "0.0000000000000000000000000000000000000000000000000000000001234",
    "1.00000000000000011102230246251565404",  # 37 digits just above the 1 + 2^-53 midpoint
]
REJECT = {"1.00000000000000011102230246251565404236316680908203125": "E2", "3*1.0": "E2", "1.0,2.0": "E2",
          "1.0/": "E2", "abc": "E1", "1.0x": "E1", "1.0E": "E1", ".": "E1", "-": "E1"}


def _random_tokens(seed, count):
    rng = random.Random(seed)
    getcontext().prec = 60
    out = []
    while len(out) < count:
        kind = rng.randrange(7)
        bits = rng.getrandbits(63)
        x = struct.unpack("<d", struct.pack("<Q", bits))[0]
        if not np.isfinite(x):
            continue
        if kind == 0:
            out.append(repr(x))
        elif kind == 1:
            out.append(f"{x:.{rng.randrange(25)}e}")
        elif kind == 2:  # cube writer styles
            y = rng.uniform(-1, 1) * 10.0 ** rng.randrange(-30, 5)
            out.append(rng.choice([f"{y:.5E}", f"{y:22.14E}".strip(), f"{y:.5E}".replace("E", "D")]))
        elif kind == 3:  # random digit strings, 1..40 significant digits, any exponent
            nd = rng.randrange(1, 41)
            s = "".join(rng.choice("0123456789") for _ in range(nd))
            dp = rng.randrange(nd + 1)
            out.append(s[:dp] + "." + s[dp:] + f"E{rng.randrange(-350, 330)}")
        elif kind == 4:  # near the midpoint of two adjacent doubles
            if x == 0.0:
                continue
            y = np.nextafter(x, np.inf)
            m = (Decimal(x) + Decimal(float(y))) / 2
            d = rng.randrange(17, 40)
            t = f"{m:.{d}e}"
            mant, ex = t.split("e")
            last = int(mant[-1]) + rng.choice([-1, 0, 1])
            if 0 <= last <= 9:
                mant = mant[:-1] + str(last)
            out.append(mant + "e" + ex)
        elif kind == 5:  # subnormal range
            y = rng.getrandbits(rng.randrange(1, 53)) * 2.0 ** -1074
            out.append(f"{y:.{rng.randrange(1, 20)}e}")
        else:  # exact ties of 54-bit integers scaled by powers of two and ten
            v = (rng.getrandbits(53) | (1 << 53)) | 1
            out.append(str(v) + ("" if rng.random() < 0.5 else "e" + str(rng.randrange(-5, 6))))
    return out


def test_conversion_edge_cases_bit_exact(number_check):
    toks = EDGE + list(REJECT)
    got = number_check(toks)
    assert len(got) == len(toks)
    for t, g in zip(toks, got):
        if t in REJECT:
            assert g == REJECT[t], t
        else:
            assert g == _bits(fortran_float(t)), t


@pytest.mark.parametrize("seed", [1, 2])
def test_conversion_random_bit_exact(number_check, seed):
    toks = _random_tokens(seed, 40000)
    got = number_check(toks)
    bad = [(t, g, _bits(fortran_float(t))) for t, g in zip(toks, got) if g != _bits(fortran_float(t))]
    assert not bad, bad[:5]


def test_conversion_rejects_planted_error(number_check):
    """A mis-rounded expectation (one ulp off) must not match."""
    t = "0.1"
    wrong = _bits(np.nextafter(0.1, 1.0))
    assert number_check([t])[0] != wrong


@pytest.mark.parametrize("path", [p for p in CUBES if "header" not in p], ids=os.path.basename)
def test_oracle_read_cube_reference_files(path):
    from oracle import oracle
    hdr, f = oracle.read_cube(path)
    text = open(path, "rb").read()
    # data offset: after 6 + nat (+1) records
    lines = text.split(b"\n")
    nrec = 6 + hdr["nat"] + (1 if hdr["ismo"] else 0)
    off = sum(len(l) + 1 for l in lines[:nrec])
    want = expected_grid(text, hdr["n"], off)
    assert np.array_equal(f.view(np.uint64), want.view(np.uint64))
    assert hdr["nat"] == 16


def test_oracle_header_only_reference_file():
    from oracle import oracle
    path = [p for p in CUBES if "header" in p][0]
    with pytest.raises(ValueError):
        oracle.read_cube(path)  # the header-only file holds no grid values


@pytest.mark.parametrize("split", [None, 3])
def test_oracle_read_bincube_records(tmp_path, split):
    from oracle import oracle
    n = (5, 4, 3)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(60)
    step = np.array([[0.1, 0.0, 0.0], [0.02, 0.2, 0.0], [0.0, 0.03, 0.3]]).T  # columns = axes
    p = tmp_path / "g.bincube"
    p.write_bytes(bincube(f, n, step, x0=(1.0, 2.0, 3.0), atoms=[(6, (0.1, 0.2, 0.3))], split=split))
    hdr, g = oracle.read_bincube(str(p))
    assert hdr["n"] == n and hdr["nat"] == 1
    assert np.array_equal(g, f)
    assert np.array_equal(hdr["step"], step)


def _fmt_values(seed, count):
    rng = random.Random(seed)
    vals = [0.0, -0.0, 1.0, -1.5, 0.1, 1e308, 1.7976931348623157e308, 5e-324, 2.2250738585072014e-308,
            999999.5, 9999995.0, 1234565.0, 1234575.0, 0.5, 12345650000000.0, 99999950000000.0, 1e-5, 1e22,
            float("nan"), float("inf"), float("-inf"), 9.5367431640625e-07]
    while len(vals) < count:
        k = rng.randrange(4)
        if k == 0:
            x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
            if x != x:
                continue
        elif k == 1:
            x = rng.getrandbits(53) * 2.0 ** rng.randrange(-80, 40) * rng.choice([-1, 1])
        elif k == 2:  # decimal ties at the 7th or 15th digit
            x = float((rng.getrandbits(40) * 10 + 5) * 10 ** rng.randrange(0, 4))
        else:
            x = rng.getrandbits(rng.randrange(1, 60)) * 2.0 ** -1074
        vals.append(x)
    return vals


@pytest.mark.parametrize("seed", [1, 2])
def test_formatting_matches_python(number_check, seed):
    vals = _fmt_values(seed, 20000)
    lines = [f"{st} {struct.unpack('<Q', struct.pack('<d', v))[0]:x}" for v in vals for st in (0, 1, 2, 3)]
    got = number_check(lines, mode="fmt")
    want = [fortran_e(v, st) for v in vals for st in (0, 1, 2, 3)]
    bad = [(l, g, w) for l, g, w in zip(lines, got, want) if g != w]
    assert not bad, bad[:5]


def test_formatting_planted_error_is_caught(number_check):
    """A field one unit off in the last digit must not match."""
    got = number_check(["0 " + struct.pack(">d", 0.1).hex()], mode="fmt")[0]
    assert got == fortran_e(0.1, 0)
    assert got != fortran_e(0.1 + 1e-14, 0)  # 0.10000000000001E+000
