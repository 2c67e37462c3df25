"""The device's fast IDCT (fused sums, device_common.h idct8x8_*_fast) of the MPEG-2 and MJPEG
kernels is exact where it is used: its window bound, checked arithmetically, and an empirical check
of the same algorithm (tests/fastpath/idct_fast_check.c) against the oracle's reference-order IDCTs
on 2 M blocks each (dense, the largest coefficient sums, sparse small blocks where ties live; MJPEG
also past its row limit, which must fall back) and on the tie / FMA fixture blocks, every one of
which the window must catch. A planted checker without the window (it never falls back) lets the
fixtures' fused results through and must then disagree with the oracle."""
import json
import math
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

from idct_reference import basis  # noqa: E402

PINS = json.load(open(os.path.join(HERE, "golden", "idct_pins.json")))
WINDOW = 2.0 ** -30   # kIdctTieWindow


def test_window_covers_the_error_bound_of_every_mpeg2_block():
    u = 2.0 ** -53
    g8 = 8 * u / (1 - 8 * u)
    m = max(abs(v) for row in basis() for v in row)
    l1 = 63 * 2048 + 32768 * 8          # 63 saturated coefficients, an int16 DC shifted by 3
    bound = 2 * (2 * g8 + g8 * g8) * m * m * l1
    assert bound * 2 < WINDOW, (bound, WINDOW)


def _checker(tmp_path, window=None):
    exe = tmp_path / ("idct_fast_check" + ("_planted" if window else ""))
    src = os.path.join(HERE, "fastpath", "idct_fast_check.c")
    text = open(src).read()
    if window:
        text = text.replace("static const double kWindow = 0x1p-30;", f"static const double kWindow = {window};")
        assert window in text
    c = tmp_path / (exe.name + ".c")
    c.write_text(text.replace('#include "../../oracle/tables.h"', f'#include "{ROOT}/oracle/tables.h"'))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe), str(c), "-ldl", "-lm"], check=True)
    return str(exe)


def test_mjpeg_row_limit_keeps_blocks_inside_the_bound():
    """MJPEG coefficients (level x quantiser) are not saturated: the kernel takes the fast form only
    when every row's sum of |c| is at most 391168 / 8, so the block's sum stays within the bound
    above; the +128 adds one rounding of at most 2^-45 to both evaluations."""
    u = 2.0 ** -53
    g8 = 8 * u / (1 - 8 * u)
    m = max(abs(v) for row in basis() for v in row)
    bound = 2 * (2 * g8 + g8 * g8) * m * m * (391168 // 8) * 8 + 2 * 2.0 ** -45
    assert bound < WINDOW


def _run(exe, n, seed, blocks=(), mode="mpeg2"):
    from oracle import binding

    binding.lib()
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    stdin = "\n".join(" ".join(str(v) for v in b) for b in blocks)
    r = subprocess.run([exe, so, str(n), str(seed), mode], input=stdin, capture_output=True, text=True, check=True,
                       timeout=600)
    count, flagged, bad = map(int, r.stdout.split())
    return count, flagged, bad


def _fixtures(kinds=("mpeg2", "mpeg2_odd")):
    out = []
    for kind in kinds:
        for cat in ("tie", "fused"):
            for d in PINS[kind][cat]:
                c = [0] * 64
                for k, v in d.items():
                    c[int(k)] = v
                out.append(c)
    return out


@pytest.mark.timeout(600)
def test_fast_path_rounds_like_the_reference_outside_the_window(tmp_path):
    exe = _checker(tmp_path)
    count, flagged, bad = _run(exe, 2_000_000, 1)
    assert count == 2_000_000 and bad == 0
    assert 0 < flagged < count // 100     # ties do occur (and fall back), rarely
    fx = _fixtures()
    count, flagged, bad = _run(exe, 0, 1, fx)
    assert count == len(fx) and flagged == len(fx) and bad == 0


@pytest.mark.timeout(600)
def test_mjpeg_fast_path_rounds_like_the_reference_outside_the_window(tmp_path):
    exe = _checker(tmp_path)
    count, flagged, bad = _run(exe, 2_000_000, 3, mode="mjpeg")
    assert count == 2_000_000 and bad == 0
    assert count // 8 < flagged < count // 4   # every block past the row limit, and the ties
    fx = _fixtures(("mjpeg",))
    count, flagged, bad = _run(exe, 0, 1, fx, mode="mjpeg")
    assert count == len(fx) and flagged == len(fx) and bad == 0


@pytest.mark.parametrize("mode, kinds", [("mpeg2", ("mpeg2", "mpeg2_odd")), ("mjpeg", ("mjpeg",))])
def test_planted_missing_window_fails_on_the_fixtures(tmp_path, mode, kinds):
    exe = _checker(tmp_path, "-1.0")
    count, flagged, bad = _run(exe, 0, 1, _fixtures(kinds), mode=mode)
    assert bad > 0


# ---------------------------------------------------------------- DV (dv.hip: fused row / column passes)
DV_WINDOW = 2.0 ** -25     # kDvTieWindow
DV_BLOCKS = json.load(open(os.path.join(HERE, "golden", "dv_fallback_blocks.json")))


def test_dv_window_covers_the_error_bound():
    """Reference order: 64 sequential sums of three-factor products (gamma_66); fused form: two
    8-term fused chains and tables rounded once (2 gamma_8 + gamma_8^2 + 3u); every factor of a term
    is at most 1 in magnitude, and |c| <= 132092 (AC: int32 >> 14; DC: int16 * 4 + 1024)."""
    u = 2.0 ** -53
    g = lambda n: n * u / (1 - n * u)
    a = 64 * (32767 * 4 + 1024)
    assert (1 << 31) >> 14 <= 32767 * 4 + 1024
    bound = (2 * g(8) + g(8) ** 2 + 3 * u + g(66)) * a / 4
    assert bound < DV_WINDOW, (bound, DV_WINDOW)


def _dv_checker(tmp_path):
    exe = tmp_path / "dv_fast_check"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", str(exe),
                    os.path.join(HERE, "fastpath", "dv_fast_check.c"), "-ldl", "-lm"], check=True)
    return str(exe)


def _dv_run(exe, n, seed, mode, blocks=(), window=None):
    from oracle import binding

    binding.lib()
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    stdin = "\n".join(" ".join(str(v) for v in b) for b in blocks)
    args = [exe, so, str(n), str(seed), mode] + ([window] if window else [])
    r = subprocess.run(args, input=stdin, capture_output=True, text=True, check=True, timeout=600)
    return tuple(map(int, r.stdout.split()))


def _dv_pins(kind):
    out = []
    for cat in ("tie", "fused"):
        for d in PINS[kind][cat]:
            c = [0] * 64
            for k, v in d.items():
                c[int(k)] = v
            out.append(c)
    return out


def _dv_fallback_coefficients(oracle, kind):
    import ctypes as C

    out = []
    for b in DV_BLOCKS[kind]:
        lv = (C.c_int16 * 64)()
        for p, v in b["levels"].items():
            lv[int(p)] = v
        c = (C.c_int32 * 64)()
        oracle.oracle_dv_dequantize(lv, C.c_uint8(b["info"]), c)
        assert (b["info"] & 1) == (kind == "dv248")
        out.append(list(c))
    return out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["88", "248"])
def test_dv_fast_path_rounds_like_the_reference_outside_the_window(tmp_path, mode):
    exe = _dv_checker(tmp_path)
    count, flagged, bad = _dv_run(exe, 1_000_000, 7, mode)
    assert count == 1_000_000 and bad == 0
    assert 0 < flagged < count // 100    # ties do occur (and fall back), rarely
    fx = _dv_pins("dv" + mode)
    count, flagged, bad = _dv_run(exe, 0, 1, mode, fx)
    assert count == len(fx) and flagged == len(fx) and bad == 0


@pytest.mark.parametrize("mode", ["88", "248"])
def test_dv_planted_missing_window_fails(tmp_path, oracle, mode):
    """Without the window the fused form rounds samples of the tie / FMA pins and of the committed
    level blocks differently from the reference; with it, every such block is flagged."""
    exe = _dv_checker(tmp_path)
    for blocks in (_dv_pins("dv" + mode), _dv_fallback_coefficients(oracle, "dv" + mode)):
        assert _dv_run(exe, 0, 1, mode, blocks, window="-1.0")[2] > 0
        count, flagged, bad = _dv_run(exe, 0, 1, mode, blocks)
        assert flagged == count == len(blocks) and bad == 0
    for c in _dv_fallback_coefficients(oracle, "dv" + mode):   # each committed block on its own
        assert _dv_run(exe, 0, 1, mode, [c], window="-1.0")[2] > 0
