"""The traversal's triangle screen (triangle_screen_rejects, core/geometry.h)
never rejects a triangle that the exact Moller-Trumbore test accepts within
the current closest distance, for reciprocals up to 2 ulps off the correctly
rounded 1/a (the hardware v_rcp_f32 is within 1 ulp). Compiled for the host
from the same header the GPU kernels include."""
import ctypes as C
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "native" / "screen_check.cpp"


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    so = tmp_path_factory.mktemp("screen") / "screen_check.so"
    subprocess.run(["g++", "-O2", "-std=c++20", "-ffp-contract=off", "-shared", "-fPIC",
                    str(SRC), "-o", str(so)], check=True)
    lib = C.CDLL(str(so))
    f = lib.screen_violations
    f.restype = C.c_int64
    f.argtypes = [C.c_uint64, C.c_int64, C.c_float, C.c_int, C.POINTER(C.c_int64),
                  C.POINTER(C.c_int64)]

    def run(seed, n, margin, ulps):
        rej, acc = C.c_int64(), C.c_int64()
        bad = f(seed, n, margin, ulps, C.byref(rej), C.byref(acc))
        assert bad >= 0, "intersect_triangle_exact differs from intersect_triangle"
        return bad, rej.value, acc.value
    return run


def test_screen_is_conservative(check):
    bad, rejected, accepted = check(1, 400_000, 2.0 ** -16, 2)
    assert accepted > 100_000          # many exact hits near the boundaries
    assert rejected > 0.3 * 5 * 400_000  # the screen decides many candidates alone
    assert bad == 0


def test_check_detects_a_margin_below_the_reciprocal_error(check):
    # with no margin, a reciprocal 1-2 ulps high pushes boundary hits out
    bad, _, _ = check(2, 400_000, 0.0, 2)
    assert bad > 0
