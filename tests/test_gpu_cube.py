"""GPU cube readers (c2gpu_cube_read / c2gpu_cube_parse / c2gpu_bincube_read)
against the oracle's restatement of grid3%read_cube / read_bincube and
against Python's float() of every token: bit-exact, header included."""
import glob
import os
import random
import struct

import numpy as np
import pytest

from critic2_amd import cube
from critic2_amd._lib import C2GpuError
from tests.cubetext import (atoms_of, bincube, body, expected_body, expected_grid, fortran_e, fortran_float,
                            header)
from tests.helpers import GOLDEN

pytestmark = pytest.mark.gpu

CUBES = sorted(glob.glob(os.path.join(GOLDEN, "cube", "*.cube")))


def _same(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


GRID_CUBES = [p for p in CUBES if "nciplot" not in p]
NCI_CUBES = [p for p in CUBES if "nciplot" in p]


@pytest.mark.parametrize("path", [p for p in CUBES if "header" not in p], ids=os.path.basename)
def test_reference_cube_files(ctx, path):
    from oracle import oracle
    print("case", os.path.basename(path), flush=True)
    hdr, f = cube.read_cube(ctx, path)
    ohdr, of = oracle.read_cube(path)
    assert hdr["n"] == ohdr["n"] and hdr["nat"] == ohdr["nat"] and hdr["ismo"] == ohdr["ismo"]
    assert _same(hdr["x0_cart"], ohdr["x0_cart"]) and _same(hdr["step"], ohdr["step"])
    assert _same(f, of)
    text = open(path, "rb").read()
    assert _same(f, expected_grid(text, hdr["n"], hdr["data_offset"]))
    assert np.allclose(hdr["x2cl"], hdr["step"] * np.array(hdr["n"])[None, :], rtol=0, atol=0)


def test_reference_header_only_file(ctx):
    path = [p for p in CUBES if "header" in p][0]
    hdr = cube.read_header(ctx, path)
    assert hdr["n"] == (1, 1, 1) and hdr["nat"] == 16
    with pytest.raises(C2GpuError) as e:
        cube.read_cube(ctx, path)
    assert e.value.code == 9


def _tokens(rng, count):
    out = []
    for _ in range(count):
        kind = rng.randrange(9)
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(63)))[0]
        if not np.isfinite(x):
            x = 1.0
        if kind == 0:
            out.append(f"{rng.uniform(-1, 1) * 10.0 ** rng.randrange(-12, 3):.5E}")
        elif kind == 1:
            out.append(f"{rng.uniform(-1, 1) * 10.0 ** rng.randrange(-12, 3):.14E}")
        elif kind == 2:
            out.append(repr(x))
        elif kind == 3:
            nd = rng.randrange(1, 41)
            s = "".join(rng.choice("0123456789") for _ in range(nd))
            dp = rng.randrange(nd + 1)
            out.append(rng.choice("+-") * rng.randrange(2) + s[:dp] + "." + s[dp:] +
                       rng.choice("EeDd") + str(rng.randrange(-340, 320)))
        elif kind == 4:
            out.append(f"{rng.getrandbits(40) * 2.0 ** -1074:.{rng.randrange(1, 18)}e}")
        elif kind == 5:
            out.append(rng.choice(["0", "-0.0", "1", "-1", "1e308", "2e308", "1e-330", "1.5-3", "2.5D+2"]))
        elif kind == 6:  # long tokens past the staged LDS tail: leading zeros
            out.append("0." + "0" * rng.randrange(250, 600) + "1234" + "E+" + str(rng.randrange(250, 600)))
        elif kind == 7:
            out.append(str(rng.getrandbits(60)))
        else:
            out.append(f"{x:.{rng.randrange(20)}e}")
    return out


@pytest.mark.parametrize("seed,n,per_line,sep,eol", [
    (1, (13, 7, 29), 6, " ", "\n"),
    (2, (5, 11, 3), 1, "\t", "\r\n"),
    (3, (17, 3, 64), 64, "   ", "\n"),
    (4, (2, 2, 2), 3, " ", "\n"),
])
def test_synthetic_cubes_bit_exact(ctx, tmp_path, seed, n, per_line, sep, eol):
    print("case seed", seed, "n", n, flush=True)
    rng = random.Random(seed)
    npts = n[0] * n[1] * n[2]
    toks = _tokens(rng, npts)
    step = [[0.1 * (i + 1) if a == i else 0.01 * (a - i) for a in range(3)] for i in range(3)]
    atoms = [(6, (0.1, 0.2, 0.3)), (1, (1.0, 1.5, -0.5))]
    ismo = seed == 2
    text = header(n, step, x0=(-1.25, 0.5, 2.0), atoms=atoms, ismo=ismo, precise=seed % 2 == 1)
    text += body(toks, per_line=per_line, n3=n[2], sep=sep, eol=eol)
    text += " 9.99E+00 trailing values are ignored\n"
    p = tmp_path / "g.cube"
    p.write_text(text)
    raw = text.encode()
    hdr, f = cube.read_cube(ctx, str(p))
    assert hdr["n"] == n and hdr["nat"] == 2 and hdr["ismo"] == ismo
    want = np.array([fortran_float(t) for t in toks]).reshape(n).transpose(2, 1, 0).reshape(-1)
    assert _same(f, want)
    assert _same(f, expected_grid(raw, n, hdr["data_offset"]))
    # same bytes through the in-memory entry, host and device (unaligned)
    h2, f2 = cube.parse_cube(ctx, raw)
    assert _same(f2, want)
    for shift in (0, 3, 13):
        d = ctx.empty((len(raw) + shift,), np.uint8)
        d.copy_from(np.frombuffer(b"#" * shift + raw, dtype=np.uint8))

        class View:  # the text starts `shift` bytes into the allocation
            def data_ptr(self, _d=d, _s=shift):
                return _d.data_ptr() + _s
        h3, f3 = cube.parse_cube(ctx, (View(), len(raw)))
        assert _same(f3.numpy(), want), shift
        assert h3["data_offset"] == hdr["data_offset"]


def test_multi_workgroup_cube_matches_oracle(ctx, tmp_path):
    n = (40, 36, 52)  # ~75k values, ~1 MB of text: a few hundred workgroups
    rng = np.random.default_rng(7)
    v = rng.standard_normal(n[0] * n[1] * n[2]) * np.exp(rng.uniform(-30, 5, n[0] * n[1] * n[2]))
    toks = [f"{x:.5E}" for x in v]
    text = header(n, np.eye(3) * 0.2) + body(toks, per_line=6, n3=n[2])
    p = tmp_path / "big.cube"
    p.write_text(text)
    from oracle import oracle
    _, of = oracle.read_cube(str(p))
    hdr, f = cube.read_cube(ctx, str(p))
    assert _same(f, of)
    hdr, fd = cube.read_cube(ctx, str(p), device=True)
    assert _same(fd.numpy(), of)


@pytest.mark.parametrize("bad,why", [
    ("", "fewer values"),
    (" abc", "non-numeric"),
    (" 3*1.0", "repeat count"),
    (" 1.0,2.0", "comma"),
    (" 1." + "1" * 45, "more than 40 significant digits"),
])
def test_malformed_values_are_errors(ctx, tmp_path, bad, why):
    n = (2, 2, 2)
    toks = [f"{0.5 * i:.5E}" for i in range(7)]
    text = header(n, np.eye(3)) + body(toks, n3=2) + bad + "\n"
    p = tmp_path / "bad.cube"
    p.write_text(text)
    with pytest.raises(C2GpuError) as e:
        cube.read_cube(ctx, str(p))
    assert e.value.code == 9, why


def test_truncated_header_is_error(ctx, tmp_path):
    p = tmp_path / "h.cube"
    p.write_text("title\ntitle\n  2  0.0 0.0 0.0\n  4 0.1 0.0 0.0\n")
    with pytest.raises(C2GpuError) as e:
        cube.read_header(ctx, str(p))
    assert e.value.code == 9


@pytest.mark.parametrize("split,ismo,device", [(None, False, False), (4, True, False), (3, False, True)])
def test_bincube_matches_oracle(ctx, tmp_path, split, ismo, device):
    from oracle import oracle
    n = (9, 5, 7)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(n[0] * n[1] * n[2])
    step = np.array([[0.3, 0.0, 0.0], [0.05, 0.25, 0.0], [0.0, 0.02, 0.2]]).T
    p = tmp_path / "g.bincube"
    p.write_bytes(bincube(f, n, step, x0=(0.5, -0.5, 1.0), atoms=[(8, (0.0, 0.1, 0.2))] * 3, ismo=ismo,
                          split=split))
    hdr, g = cube.read_bincube(ctx, str(p), device=device)
    if device:
        g = g.numpy()
    ohdr, og = oracle.read_bincube(str(p))
    assert _same(g, f) and _same(g, og)
    assert hdr["n"] == n and hdr["nat"] == 3 and hdr["ismo"] == ismo
    assert _same(hdr["step"], step)


def test_bincube_short_record_is_error(ctx, tmp_path):
    n = (4, 4, 4)
    data = bincube(np.zeros(60), n, np.eye(3))  # 60 of 64 values
    p = tmp_path / "s.bincube"
    p.write_bytes(data)
    with pytest.raises(C2GpuError) as e:
        cube.read_bincube(ctx, str(p))
    assert e.value.code == 9


# ---- writer

@pytest.mark.parametrize("path", GRID_CUBES, ids=os.path.basename)
def test_writer_reproduces_reference_grid_files(ctx, tmp_path, path):
    """writegrid_cube output of the reference, rewritten from the values read
    back, byte for byte (header, atoms, records, trailing blanks)."""
    print("case", os.path.basename(path), flush=True)
    text = open(path, "rb").read()
    precise = b"E+" in text.split(b"\n")[2]
    style = 0 if precise else 1
    if "header" in path:
        hdr, f = cube.read_header(ctx, path), None
    else:
        hdr, f = cube.read_cube(ctx, path)
    z, x = atoms_of(text, hdr["nat"])
    out = tmp_path / "w.cube"
    cube.write_cube(ctx, str(out), hdr["n"], hdr["x0_cart"], hdr["step"], z, x, f, style=style)
    assert out.read_bytes() == text


@pytest.mark.parametrize("path", NCI_CUBES, ids=os.path.basename)
def test_nci_body_reproduces_reference(ctx, path):
    """nciplot's write_cube_body records (1p,e13.5e3) of the reference."""
    text = open(path, "rb").read()
    hdr, f = cube.read_cube(ctx, path)
    assert cube.format_body(ctx, f, hdr["n"], style=2) == text[hdr["data_offset"]:]


@pytest.mark.parametrize("style", [0, 1, 2])
@pytest.mark.parametrize("n,zfast,device", [((7, 5, 13), False, False), ((3, 4, 12), True, False),
                                            ((9, 2, 31), False, True), ((1, 1, 1), True, True)])
def test_format_matches_python(ctx, style, n, zfast, device):
    rng = random.Random(hash((style, n)) & 0xffff)
    npts = n[0] * n[1] * n[2]
    vals = []
    specials = [0.0, -0.0, 5e-324, 1.7976931348623157e308, float("inf"), float("-inf"), float("nan"), 1234565.0]
    for t in range(npts):
        k = rng.randrange(5)
        if k == 0 and t < len(specials) * 3:
            vals.append(specials[t % len(specials)])
        elif k == 1:
            vals.append(struct.unpack("<d", struct.pack("<Q", rng.getrandbits(63) | (rng.getrandbits(1) << 63)))[0])
        else:
            vals.append(rng.uniform(-1, 1) * 10.0 ** rng.randrange(-20, 5))
    f = np.array(vals)
    # expected from the x-fastest view
    fx = f if not zfast else f.reshape(n[0], n[1], n[2]).transpose(2, 1, 0).reshape(-1)
    want = expected_body(fx, n, style).encode()
    src = ctx.to_device(f) if device else f
    got = cube.format_body(ctx, src, n, style=style, zfast=zfast)
    assert got == want


def test_write_read_round_trip_large(ctx, tmp_path):
    """Precise cube of a 96x80x64 grid: reading back gives each value
    rounded to 14 digits; writing that again gives the same bytes."""
    n = (96, 80, 64)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(n[0] * n[1] * n[2]) * np.exp(rng.uniform(-40, 40, n[0] * n[1] * n[2]))
    p1, p2 = tmp_path / "a.cube", tmp_path / "b.cube"
    step = np.diag([0.1, 0.11, 0.12])
    cube.write_cube(ctx, str(p1), n, (0.0, 0.0, 0.0), step, f=f, style=0)
    hdr, g = cube.read_cube(ctx, str(p1))
    want = np.array([float(f"{v:.13e}") for v in f[:5000]])
    assert _same(g[:5000], want)
    cube.write_cube(ctx, str(p2), n, (0.0, 0.0, 0.0), step, f=ctx.to_device(g), style=0)
    assert p1.read_bytes() == p2.read_bytes()
    from oracle import oracle
    _, og = oracle.read_cube(str(p1))
    assert _same(g, og)


def test_standard_style_negative_values_overflow_field(ctx):
    """1p,E12.5E3 has no room for a sign: negative values print asterisks."""
    got = cube.format_body(ctx, np.array([-1.0, 2.0]), (1, 1, 2), style=1)
    assert got == b" ************ 2.00000E+000 \n"
    assert fortran_e(-1.0, 1) == "*" * 12
