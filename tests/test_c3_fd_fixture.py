"""The committed finite-difference reference of the 512 x 512 adjoint workload
(tests/golden/c3_fd_gradient.npz, written by tests/golden/make_c3_fd.py from
the FP64 oracle with its FP64 colour path).  The GPU test compares the
adjoint with it; these checks pin the fixture itself on the CPU:

* its weights are the workload's dL/dI = 2 (I - T) / n at p0: re-rendered by
  the oracle on seeded pixels, (I - T) agrees with the stored float16 values;
* its central differences have converged: the four steps h = 1e-6 ... 1.25e-7
  agree to well within the adjoint's bar for every parameter;
* the pixels it leaves out are few (class changes within +-h, and pixels
  whose own differences had not converged), while the differences over every
  pixel have not converged (which is why they are left out).
"""
import os

import numpy as np
import pytest

import oracle as O
from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "c3_fd_gradient.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIXTURE)


def test_weights_are_the_loss_gradient_at_p0(fx):
    size = int(fx["size"])
    s16 = fx["s16"].astype(np.float64)
    assert s16.shape == (size, size, 3) and np.isfinite(s16).all()
    excluded = np.concatenate([fx["excluded_class"], fx["excluded_fd"]])
    assert np.all(s16.reshape(-1, 3)[excluded] == 0.0)
    S = O.Scene(A.PRR_KERR_BL, 1.0, 0.7, CF.tracer_kerr(tol=1e-12))
    S.set_texture(0, CF.sky_texture(), A.PRR_FILTER_BILINEAR)
    S.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    tmpl = CF.frame_c3(size, size)
    kept = np.setdiff1d(np.arange(size * size), excluded)
    ids = np.sort(np.random.default_rng(3).choice(kept, 48, replace=False)).astype(np.int64)
    I = S.render_rays(api.frame_from_orbit(tmpl, CF.c3_orbit()), ids, want_color64=True)["color64"]
    T = S.render_rays(api.frame_from_orbit(tmpl, CF.c3_orbit(target=True)), ids, want_color64=True)["color64"]
    d = np.clip(I, 0, 1) - np.clip(T, 0, 1)
    stored = s16.reshape(-1, 3)[ids]
    # float16 keeps 11 significant bits
    assert np.all(np.abs(stored - d) <= 2.0 ** -11 * np.abs(d) + 1e-7), np.abs(stored - d).max()


def test_differences_have_converged(fx):
    fd = fx["fd"]
    assert fd.shape == (4, 4) and np.all(np.isfinite(fd))
    spread = np.abs(fd - fd[-1]).max(axis=0) / np.abs(fd[-1])
    assert np.all(spread < 2e-6), spread


def test_few_pixels_left_out_and_the_whole_image_does_not_converge(fx):
    size = int(fx["size"])
    left_out = fx["excluded_class"].size + fx["excluded_fd"].size
    assert left_out < 0.002 * size * size
    fd_all = fx["fd_all"]
    spread_all = np.abs(fd_all - fd_all[-1]).max(axis=0) / np.abs(fd_all[-1])
    assert spread_all.max() > 1e-2
