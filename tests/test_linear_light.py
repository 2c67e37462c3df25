"""The oracle's MMFX linear-light operator (oracle/compose.c: LinearPixel, crates/mmfx/src/render.rs
:1780-1880) pinned by the reference's own known answer and by the mathematics:
* composites_group_opacity_in_linear_light (render.rs:1996-2006): a #f00 rectangle in a group of
  opacity 0.5 over a #000 background reads back as [187..188, 0, 0, 255];
* srgb_to_linear_u16 and linear_u16_to_srgb for every input equal the sRGB transfer functions
  evaluated in 40-digit decimal arithmetic and rounded half away from zero;
* multiply_channel is the nearest integer to l * r / 65535 (65535 is odd, so there are no ties);
  over() is the identity at opacity 0, replaces with an opaque source at opacity 65535 and stays
  within one step of the exact Porter-Duff value;
* to_rgba8(from_rgba(p)) returns every opaque sRGBA8 pixel unchanged, and its alpha byte is the
  nearest integer to alpha / 257.
Planted mistakes in the oracle's sources (multiply_channel truncating, the sRGB exponent 2.2, the
alpha byte divided by 256) each fail one of these."""
import ctypes as C
import os
import shutil
import subprocess
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

from mmrecode_amd import abi
from oracle import binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _px(*v):
    return (C.c_uint16 * 4)(*v)


def group_opacity_scene(L):
    """The reference test's scene on the oracle: background #000 painted over a transparent 1x1
    surface, the group's surface holding #f00, blended at opacity round(0.5 * 65535) = 32768."""
    background = (C.c_uint16 * 4)()
    L.oracle_linear_from_rgba((C.c_uint8 * 4)(0, 0, 0, 255), background)
    target = _px(0, 0, 0, 0)
    L.oracle_linear_over(background, target, 65535)
    red = (C.c_uint16 * 4)()
    L.oracle_linear_from_rgba((C.c_uint8 * 4)(255, 0, 0, 255), red)
    group = _px(0, 0, 0, 0)
    L.oracle_linear_over(red, group, 65535)
    job = abi.new(abi.LinearCompositeJob, target=C.addressof(target), width=1, height=1, stride=1, layer_count=1)
    job.layers[0] = abi.LinearLayer(pixels=C.addressof(group), format=abi.LINEAR_SOURCE_SURFACE, stride=1, x=0, y=0, width=1,
                                    height=1, clip_left=0, clip_top=0, clip_right=1, clip_bottom=1, opacity=32768)
    assert L.oracle_linear_composite(C.byref(job)) == 0
    out = (C.c_uint8 * 4)()
    L.oracle_linear_to_rgba8(target, out)
    return list(target), list(out)


def exact_round(d):
    """Round half away from zero of a positive Decimal."""
    return int((d + Decimal("0.5")).to_integral_value(rounding="ROUND_FLOOR"))


def srgb_to_linear_exact(v):
    e = Decimal(v) / Decimal(255)
    lin = e / Decimal("12.92") if e <= Decimal("0.04045") else ((e + Decimal("0.055")) / Decimal("1.055")) ** Decimal("2.4")
    return lin * 65535


def linear_to_srgb_exact(v):
    lin = Decimal(v) / Decimal(65535)
    if lin <= Decimal("0.0031308"):
        enc = lin * Decimal("12.92")
    else:
        enc = Decimal("1.055") * lin ** (Decimal(1) / Decimal("2.4")) - Decimal("0.055")
    return enc * 255


def check_group_opacity(L):
    _, out = group_opacity_scene(L)
    return 187 <= out[0] <= 188 and out[1:] == [0, 0, 255]


def check_srgb_to_linear(L):
    getcontext().prec = 40
    bad = 0
    for v in range(256):
        q = srgb_to_linear_exact(v)
        if abs(q - int(q) - Decimal("0.5")) < Decimal("1e-9"):
            continue
        bad += L.oracle_srgb_to_linear_u16(v) != min(exact_round(q), 65535)
    return bad


def check_linear_to_srgb(L, values):
    getcontext().prec = 40
    bad = 0
    for v in values:
        q = linear_to_srgb_exact(v)
        if abs(q - int(q) - Decimal("0.5")) < Decimal("1e-9"):
            continue
        bad += L.oracle_linear_u16_to_srgb(v) != min(max(exact_round(q), 0), 255)
    return bad


def check_multiply(L, n=20000):
    rng = np.random.default_rng(3)
    pairs = [(a, b) for a in (0, 1, 2, 32767, 32768, 65534, 65535) for b in (0, 1, 32768, 65535)]
    pairs += [tuple(int(x) for x in rng.integers(0, 65536, 2)) for _ in range(n)]
    bad = 0
    for a, b in pairs:
        want = round(Fraction(a * b, 65535))   # exact: 65535 is odd, so a * b / 65535 is never k + 1/2
        bad += L.oracle_linear_multiply_channel(a, b) != want
    return bad


def check_round_trip(L):
    bad = 0
    for v in range(256):
        for c in range(3):
            p = [0, 0, 0, 255]
            p[c] = v
            lin = (C.c_uint16 * 4)()
            L.oracle_linear_from_rgba((C.c_uint8 * 4)(*p), lin)
            out = (C.c_uint8 * 4)()
            L.oracle_linear_to_rgba8(lin, out)
            bad += list(out) != p
    return bad


def check_alpha_byte(L):
    """to_rgba8's alpha byte is the nearest integer to alpha / 257 (257 is odd: no ties)."""
    bad = 0
    out = (C.c_uint8 * 4)()
    for a in range(0, 65536, 7):
        L.oracle_linear_to_rgba8(_px(0, 0, 0, a), out)
        bad += out[3] != round(Fraction(a, 257))
    return bad


def test_reference_group_opacity_known_answer():
    target, out = group_opacity_scene(ob.lib())
    assert target == [32768, 0, 0, 65535]
    assert 187 <= out[0] <= 188 and out[1:] == [0, 0, 255], out


def test_srgb_to_linear_is_the_transfer_function():
    assert check_srgb_to_linear(ob.lib()) == 0


def test_linear_to_srgb_is_the_inverse_transfer_function():
    assert check_linear_to_srgb(ob.lib(), range(65536)) == 0


def test_multiply_channel_is_the_nearest_integer():
    assert check_multiply(ob.lib()) == 0


def test_over_identities_and_porter_duff():
    L = ob.lib()
    rng = np.random.default_rng(11)
    for _ in range(3000):
        a = int(rng.integers(0, 65536))
        src = [min(int(rng.integers(0, 65536)), a) for _ in range(3)] + [a]   # premultiplied: channels <= alpha
        b = int(rng.integers(0, 65536))
        dst = [min(int(rng.integers(0, 65536)), b) for _ in range(3)] + [b]
        op = int(rng.choice([0, 65535, int(rng.integers(0, 65536))]))
        d = _px(*dst)
        L.oracle_linear_over(_px(*src), d, op)
        got = list(d)
        if op == 0:
            assert got == dst
        if op == 65535 and a == 65535:
            assert got == src
        sa = Fraction(a * op, 65535 * 65535)
        for c in range(4):
            exact = Fraction(src[c] * op, 65535) + dst[c] * (1 - sa)
            assert abs(got[c] - min(exact, 65535)) <= 1.5, (src, dst, op, got)


def test_opaque_round_trip_is_lossless():
    assert check_round_trip(ob.lib()) == 0


def test_alpha_byte_is_the_nearest_integer_to_alpha_over_257():
    assert check_alpha_byte(ob.lib()) == 0


PLANTS = {
    "multiply_truncates": ("const uint32_t q = ((uint32_t)left * right + LINEAR_MAX / 2) / LINEAR_MAX;",
                           "const uint32_t q = ((uint32_t)left * right) / LINEAR_MAX;"),
    "gamma_2_2": ("pow((encoded + 0.055) / 1.055, 2.4)", "pow((encoded + 0.055) / 1.055, 2.2)"),
    "alpha_by_256": ("const uint32_t a8 = (alpha + 128) / 257;", "const uint32_t a8 = alpha / 256;"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    base = tmp_path_factory.mktemp("linear_planted")
    libs = {}
    for name, (old, new) in PLANTS.items():
        d = base / name
        (d / "oracle").mkdir(parents=True)
        (d / "include").mkdir()
        shutil.copy(os.path.join(ROOT, "include", "mmr_recon.h"), d / "include")
        srcs = []
        for f in os.listdir(os.path.join(ROOT, "oracle")):
            if f.endswith((".c", ".h")):
                text = open(os.path.join(ROOT, "oracle", f)).read()
                if f == "compose.c":
                    assert old in text, name
                    text = text.replace(old, new)
                (d / "oracle" / f).write_text(text)
                if f.endswith(".c"):
                    srcs.append(str(d / "oracle" / f))
        so = d / "liboracle_planted.so"
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", "-D_GNU_SOURCE", "-shared",
                        "-o", str(so), *srcs, "-lm"], check=True)
        L = C.CDLL(str(so))
        L.oracle_linear_multiply_channel.restype = C.c_uint16
        L.oracle_linear_multiply_channel.argtypes = [C.c_uint16, C.c_uint16]
        L.oracle_srgb_to_linear_u16.restype = C.c_uint16
        L.oracle_srgb_to_linear_u16.argtypes = [C.c_uint8]
        L.oracle_linear_u16_to_srgb.restype = C.c_uint8
        L.oracle_linear_u16_to_srgb.argtypes = [C.c_uint16]
        L.oracle_linear_over.argtypes = [C.c_void_p, C.c_void_p, C.c_uint16]
        libs[name] = L
    return libs


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistakes_fail(planted, plant):
    L = planted[plant]
    failed = (not check_group_opacity(L) or check_srgb_to_linear(L) > 0 or check_multiply(L, 2000) > 0
              or check_round_trip(L) > 0 or check_alpha_byte(L) > 0)
    assert failed, plant
