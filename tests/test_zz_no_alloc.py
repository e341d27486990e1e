"""pt_render_noalloc: the mega-kernel render that allocates nothing (scratch
in a file-scope device array) against pt_render's device output, through the
C ABI, on two calls in a row (the second smaller) and with partial sums in the
scratch (several sample blocks per pixel)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(x, y, ftw):
    from path_amd import _lib as L
    spec = L.PtSampleSpec()
    spec.type = L.SAMPLE_SPEC["SquareImage"]
    spec.x_dim, spec.y_dim = x, y
    spec.film_to_world[:] = [float(v) for v in np.asarray(ftw).ravel()]
    return spec


def _noalloc(renderer, scene, ftw, x, y, spp, settings, out):
    from path_amd import _lib as L
    secs = C.c_double()
    return L.lib().pt_render_noalloc(renderer.handle, C.byref(_spec(x, y, ftw)),
                                     C.c_void_p(out.ptr), scene.handle, spp,
                                     C.byref(settings.raw), C.byref(secs))


def _reference(renderer, scene, ftw, x, y, spp, settings):
    """pt_render into device memory (the existing entry point)."""
    from path_amd.renderer import DeviceBuffer
    buf = DeviceBuffer(renderer, x * y * 12)
    renderer.render(scene, ftw, x, y, spp, settings, output="DeviceFloatRGB", device_out=buf)
    return buf.numpy((y, x, 3)).copy()


def _settings(block_size=256, per_thread=0):
    from path_amd.renderer import Settings
    s = Settings()
    s.term_prob = "NIters"
    s.n_iters = 4
    if per_thread:
        wd = s.render_work_division
        wd.block_size = block_size
        wd.target_x_block_size = 32
        wd.force_target_samples = 1
        wd.forced_target_samples_per_thread = per_thread
    return s


def test_noalloc_matches_pt_render_two_calls(renderer, cornell_sphere):
    from path_amd import _lib as L
    from path_amd.renderer import DeviceBuffer, work_division
    scene, ftw = cornell_sphere
    # call 1: two sample blocks per pixel (partial sums in the scratch array)
    s1 = _settings(512, 1)
    assert work_division(s1.render_work_division, 1024, 24, 24)["num_sample_blocks"] == 2
    ref1 = _reference(renderer, scene, ftw, 24, 24, 1024, s1)
    out = DeviceBuffer(renderer, 24 * 24 * 12)
    assert _noalloc(renderer, scene, ftw, 24, 24, 1024, s1, out) == L.PT_OK, L.lib().pt_last_error()
    got1 = out.numpy((24, 24, 3)).copy()
    assert np.array_equal(got1.view(np.uint32), ref1.view(np.uint32))
    assert renderer.last_stats().samples == 24 * 24 * 1024
    # call 2: a smaller image, one sample block (written straight to the output)
    s2 = _settings()
    ref2 = _reference(renderer, scene, ftw, 16, 12, 16, s2)
    out2 = DeviceBuffer(renderer, 16 * 12 * 12)
    assert _noalloc(renderer, scene, ftw, 16, 12, 16, s2, out2) == L.PT_OK, L.lib().pt_last_error()
    got2 = out2.numpy((12, 16, 3)).copy()
    assert np.array_equal(got2.view(np.uint32), ref2.view(np.uint32))
    # call 3: several sample blocks again, smaller than call 1 (the array
    # keeps call 1's partials; they must not leak into this image)
    ref3 = _reference(renderer, scene, ftw, 8, 8, 1024, s1)
    out3 = DeviceBuffer(renderer, 8 * 8 * 12)
    assert _noalloc(renderer, scene, ftw, 8, 8, 1024, s1, out3) == L.PT_OK, L.lib().pt_last_error()
    assert np.array_equal(out3.numpy((8, 8, 3)).view(np.uint32), ref3.view(np.uint32))


def test_noalloc_rejections(renderer, cornell, cornell_sphere):
    from path_amd import _lib as L
    from path_amd.renderer import DeviceBuffer
    scene, ftw = cornell_sphere
    other, oftw = cornell
    s = _settings()
    _reference(renderer, scene, ftw, 16, 16, 16, s)
    out = DeviceBuffer(renderer, 16 * 16 * 12)
    # another scene than the resident one
    assert _noalloc(renderer, other, oftw, 16, 16, 16, s, out) == L.PT_ERR_INVALID
    # a Sobol table for another sample count than the resident one
    assert _noalloc(renderer, scene, ftw, 16, 16, 32, s, out) == L.PT_ERR_INVALID
    # Streaming is not served
    s_w = _settings()
    s_w.kernel_approach = "Streaming"
    assert _noalloc(renderer, scene, ftw, 16, 16, 16, s_w, out) == L.PT_ERR_INVALID
    # more partial sums than the scratch array holds: 1024^2 pixels x 2 blocks
    big = _settings(512, 1)
    _reference(renderer, scene, ftw, 4, 4, 1024, big)  # resident table for 1024 spp
    assert _noalloc(renderer, scene, ftw, 1024, 1024, 1024, big, out) == L.PT_ERR_INVALID
    assert b"no-allocation limit" in L.lib().pt_last_error()
    # host memory as the output
    host = np.zeros(16 * 16 * 3, np.float32)
    secs = C.c_double()
    rc = L.lib().pt_render_noalloc(renderer.handle, C.byref(_spec(16, 16, ftw)),
                                   C.c_void_p(host.ctypes.data), scene.handle, 16,
                                   C.byref(s.raw), C.byref(secs))
    assert rc == L.PT_ERR_INVALID
