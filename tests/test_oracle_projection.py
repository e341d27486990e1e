"""The rotation that carries a direction sampled in the local frame (z = the
frame axis) into world space (find_relative_vec, restating
include/lib/projection.h:9-32), through the shared per-vertex code's diffuse
sample: the reference's accuracy test (tests/src/projection.cpp:5-34, 1e-6
in float) restated on this rotation, plus the properties that fix it (polar
angle, azimuth, unit length), in float and in double."""
import numpy as np
import pytest

from oracle import oracle as O
from path_amd.renderer import material

# the input vectors of tests/src/projection.cpp:17-33 and their negations
REF_VECS = [(0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 1), (0.382, 0.38, 1), (0.7, 0.8, 1),
            (0.7, 0.8, 0.1), (0.7, 0.8, -0.1)]
AXES = [np.array(v, np.float64) / np.linalg.norm(v) for v in REF_VECS]
AXES += [-a for a in AXES]
TOL = {32: 1e-6, 64: 1e-12}


def _sample(axis, phi, v, precision):
    """Diffuse sample at a surface with normal `axis` hit from its front:
    polar cosine sqrt(1 - v), azimuth phi turns."""
    m = material("Diffuse", diffuse=(1.0, 1.0, 1.0))
    d, _, disc, used = O.bsdf_sample(m, -axis, axis, [phi, v], precision=precision)
    assert not disc and used == 2
    return d


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("k", range(len(AXES)))
def test_local_z_maps_to_frame_axis(k, precision):
    axis = AXES[k]
    for phi in (0.0, 0.3, 0.77):
        d = _sample(axis, phi, 0.0, precision)
        assert np.abs(d - axis).max() < TOL[precision], (axis, phi, d)


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("k", range(len(AXES)))
def test_rotation_keeps_polar_angle_and_azimuth(k, precision):
    axis = AXES[k]
    tol = TOL[precision] * 4
    for v in (0.1, 0.5, 0.9):
        ds = [_sample(axis, phi, v, precision) for phi in (0.1, 0.35, 0.6)]
        for d in ds:
            assert abs(np.linalg.norm(d) - 1) < tol
            assert abs(d @ axis - np.sqrt(1 - v)) < tol
        # a quarter turn apart: tangential parts orthogonal; half a turn:
        # opposite
        t = [d - (d @ axis) * axis for d in ds]
        assert abs(t[0] @ t[1]) < tol
        assert abs(t[0] @ t[2] + np.sqrt(v) ** 2) < tol


def test_rotation_is_not_a_reflection():
    """Away from the antiparallel axis the map is a proper rotation:
    azimuth increases counter-clockwise about the axis for every frame
    (a reflected frame would reverse it)."""
    for axis in AXES:
        if axis[2] < -0.999:
            continue  # -z: the reference's own map is -I there
        a = _sample(axis, 0.0, 0.5, 64)
        b = _sample(axis, 0.25, 0.5, 64)
        ta, tb = a - (a @ axis) * axis, b - (b @ axis) * axis
        assert np.cross(ta, tb) @ axis > 0, axis
