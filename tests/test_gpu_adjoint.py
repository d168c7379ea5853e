"""The Kerr adjoint kernel (prr_adjoint) on the GPU.

The gradient of L = mean((I - T)^2) w.r.t. (spin, r_cam, incl, sphere_phi):

* on the 512 x 512 workload, against the committed central differences of
  the FP64 oracle with its FP64 colour path at tol 1e-12
  (tests/golden/c3_fd_gradient.npz), per component;
* per pixel on a 96 x 96 crop of it, against central differences of the
  GPU's FP64 renders at h = 1e-6 (escape and sphere pixels);
* the FP32-bulk gradient against the FP64 reference gradient;
* independent of the checkpoint interval, the replay chunking and where rays
  are binned, and through the torch.autograd wrapper.

Pixels whose rays change class between p + h and p - h carry a
discontinuity that finite differences see and a gradient cannot; they are
left out of both sides.
"""
import os

import numpy as np
import pytest

from parrelray_amd import _abi as A
from parrelray_amd import api
from parrelray_amd import configs as CF

pytestmark = pytest.mark.gpu


def context(precision, tol):
    ctx = api.Context(0)
    ctx.set_spacetime(A.PRR_KERR_BL, 1.0, 0.7)
    ctx.set_tracer(CF.tracer_kerr(precision=precision, tol=tol))
    ctx.set_texture(0, CF.sky_texture(), A.PRR_FILTER_BILINEAR)
    ctx.set_texture(1, CF.sphere_texture(), A.PRR_FILTER_BILINEAR)
    return ctx


def render(ctx, tmpl, orbit):
    f = api.frame_from_orbit(tmpl, orbit)
    out = ctx.render(f, rays=True)[0]
    cls = out["cls"].reshape(f.height, f.width, f.samples)
    return out["rgbf"].astype(np.float64), cls


def with_params(orbit, p):
    o = A.Orbit.from_buffer_copy(orbit)
    o.spin, o.r_cam, o.incl, o.sphere_phi = p
    return o


SIZE = 96


@pytest.fixture(scope="module")
def setup():
    tmpl = CF.frame_c3(SIZE, SIZE)
    o0, ot = CF.c3_orbit(), CF.c3_orbit(target=True)
    ref = context(A.PRR_FP64, 1e-12)
    T, _ = render(ref, tmpl, ot)
    return tmpl, o0, T, ref


def test_adjoint_image_is_the_forward_render(setup):
    tmpl, o0, T, _ = setup
    ctx = context(A.PRR_FP32_BULK, 1e-8)
    I, _ = render(ctx, tmpl, o0)
    img = np.empty((SIZE, SIZE, 3), np.float32)
    g = ctx.adjoint(tmpl, o0, np.zeros((SIZE, SIZE, 3), np.float32), img_out=img)
    # the adjoint's own forward pass is the same algorithm in another kernel;
    # in the FP32-bulk form it chooses the stage precision per ray, the render
    # per wave (FP64 stages for the whole wave when one ray needs them), so the
    # rays whose wave-mates needed FP64 stages differ by the FP32 stage rounding
    d = np.abs(img - I.astype(np.float32))
    print("adjoint image vs render: max", float(d.max()), "within 1e-6", float((d <= 1e-6).mean()),
          "identical", float((d == 0).mean()))
    assert d.max() <= 1e-3 and (d <= 1e-6).mean() >= 0.995, (float(d.max()), float((d <= 1e-6).mean()))
    assert np.all(g == 0.0)
    st = ctx.stats()
    assert st["adjoint_ms"] > 0 and st["adjoint_steps"] > 0


def test_fp32_bulk_gradient_matches_the_fp64_reference_gradient(setup):
    """FP32-bulk adjoint (tol 1e-8) against the plain FP64 adjoint at tol 1e-12."""
    tmpl, o0, T, ref = setup
    I, _ = render(ref, tmpl, o0)
    dLdI = (2.0 * (I - T) / I.size).astype(np.float32)
    g64 = ref.adjoint(tmpl, o0, dLdI)
    g32 = context(A.PRR_FP32_BULK, 1e-8).adjoint(tmpl, o0, dLdI)
    rel = np.abs(g32 - g64).max() / np.abs(g64).max()
    print("fp64 reference", g64, "fp32-bulk", g32, "relative", rel)
    assert rel < 1e-3


def test_adjoint_single_pixels_match_central_differences(setup):
    """dI/dp of single pixels (dL/dI one-hot on one pixel) against central
    differences of that pixel (h = 1e-6), for escape and sphere pixels whose
    colour changes with the parameters and is differentiable on [p - h, p + h]."""
    tmpl, o0, _, ref = setup
    p0 = np.array([o0.spin, o0.r_cam, o0.incl, o0.sphere_phi])
    I0, c0 = render(ref, tmpl, o0)
    h = 1e-6
    plus, minus = [], []
    for k in range(4):
        e = np.zeros(4)
        e[k] = h
        plus.append(render(ref, tmpl, with_params(o0, p0 + e)))
        minus.append(render(ref, tmpl, with_params(o0, p0 - e)))
    rng = np.random.default_rng(5)
    ctx = context(A.PRR_FP64, 1e-10)
    checked = {A.PRR_ESCAPE: 0, A.PRR_SPHERE: 0}
    bad = []
    wch = np.array([0.3, 0.5, 0.2])
    for pix in rng.permutation(SIZE * SIZE):
        j, i = divmod(int(pix), SIZE)
        cls = c0[j, i, 0]
        if cls not in checked or checked[cls] >= 5:
            continue
        if any(plus[k][1][j, i, 0] != cls or minus[k][1][j, i, 0] != cls for k in range(4)):
            continue
        fwd = np.array([(plus[k][0][j, i] - I0[j, i]) / h for k in range(4)]) @ wch
        bwd = np.array([(I0[j, i] - minus[k][0][j, i]) / h for k in range(4)]) @ wch
        fdw = (fwd + bwd) / 2
        if np.abs(fdw).max() < 1.0:  # inside a checker cell: no signal above the float32 rounding
            continue
        # A texel edge (a kink of the bilinear lookup) inside [p - h, p + h]: the
        # one-sided differences disagree by more than the float32 rounding of the
        # image (an ulp of 0.5 over h is 0.06), and the pixel has no derivative to
        # compare against there.
        if np.abs(fwd - bwd).max() > 0.15 + 0.05 * np.abs(fdw).max():
            continue
        dLdI = np.zeros((SIZE, SIZE, 3), np.float32)
        dLdI[j, i] = wch
        g = ctx.adjoint(tmpl, o0, dLdI)
        checked[cls] += 1
        if not np.all(np.abs(g - fdw) <= 1e-2 * np.abs(fdw).max() + 0.05):
            bad.append((int(cls), j, i, g.tolist(), fdw.tolist()))
        if min(checked.values()) >= 5:
            break
    print("checked", checked, "bad", bad[:4])
    assert checked[A.PRR_ESCAPE] >= 3 and checked[A.PRR_SPHERE] >= 3
    assert not bad


def test_autograd_wrapper_backward_calls_the_adjoint(setup):
    torch = pytest.importorskip("torch")
    from parrelray_amd.autograd import KerrOrbitRender
    tmpl, o0, T, _ = setup
    ctx = context(A.PRR_FP32_BULK, 1e-8)
    mod = KerrOrbitRender(ctx, tmpl, o0)
    params = torch.tensor([o0.spin, o0.r_cam, o0.incl, o0.sphere_phi], dtype=torch.float64, requires_grad=True)
    img = mod(params)
    target = torch.from_numpy(T.astype(np.float32)).cuda()
    loss = ((img - target) ** 2).mean()
    loss.backward()
    dLdI = (2.0 * (img.detach() - target) / img.numel()).cpu().numpy()
    g = ctx.adjoint(tmpl, o0, dLdI)
    assert np.allclose(params.grad.numpy(), g, rtol=1e-5, atol=1e-12)


def test_gradient_does_not_depend_on_the_checkpoint_interval(setup):
    """The same steps are swept backwards whatever the checkpoint interval K:
    rays with at most max(8, 512 / K) segments keep the render's checkpoints,
    longer ones are checkpointed by a pass of their own (tol 1e-10 gives rays
    of over 512 steps), and each replayed segment must land on its next
    checkpoint bit for bit (the call fails otherwise)."""
    tmpl, o0, T, ref = setup
    I, _ = render(ref, tmpl, o0)
    dLdI = (2.0 * (I - T) / I.size).astype(np.float32)
    grads = []
    for K in (16, 64, 128):
        ctx = context(A.PRR_FP64, 1e-10)
        ctx.set_launch(checkpoint=K)
        grads.append(ctx.adjoint(tmpl, o0, dLdI))
        st = ctx.stats()
        print("K", K, "grad", grads[-1], "steps swept", st["adjoint_steps"])
    for g in grads[1:]:
        assert np.abs(g - grads[0]).max() <= 1e-12 * np.abs(grads[0]).max(), (g.tolist(), grads[0].tolist())


def test_gradient_does_not_depend_on_replay_chunks_or_where_rays_are_binned(setup):
    """With the default record budget the first pass records every step (each
    ray's first max(8, 512 / K) * K) and the reverse sweep reads those
    records: no replay.  Below that budget the segments are replayed, in one
    chunk of segment levels, or in several when the replay's records exceed
    the budget (lambda passes between chunks through memory); and rays are
    binned by segment count on the host when the tracer's step cap allows
    more than 4096 segments per ray: none of these changes the gradient."""
    tmpl, o0, T, ref = setup
    I, _ = render(ref, tmpl, o0)
    dLdI = (2.0 * (I - T) / I.size).astype(np.float32)
    K = 16
    ctx = context(A.PRR_FP64, 1e-10)
    ctx.set_launch(checkpoint=K)
    g0 = ctx.adjoint(tmpl, o0, dLdI)
    first_pass_records = 48 * max(8, 512 // K) * K * tmpl.nrays
    ctx.set_launch(checkpoint=K, adjoint_record_bytes=first_pass_records - 48)  # replay, one chunk
    g3 = ctx.adjoint(tmpl, o0, dLdI)
    ctx.set_launch(checkpoint=K, adjoint_record_bytes=48 * K * 300)  # about 300 tasks per chunk
    g1 = ctx.adjoint(tmpl, o0, dLdI)
    ctx.set_launch(checkpoint=K)
    tr = CF.tracer_kerr(precision=A.PRR_FP64, tol=1e-10)
    tr.max_steps = 4096 * K + 64  # no ray comes near it: same steps, host-side binning
    tr.max_attempts = 10 * tr.max_steps
    ctx.set_tracer(tr)
    g2 = ctx.adjoint(tmpl, o0, dLdI)
    print("grads", g0.tolist(), g1.tolist(), g2.tolist(), g3.tolist())
    for g in (g1, g2, g3):
        assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max(), (g.tolist(), g0.tolist())


# ---------------------------------------------------------------- the 512 x 512 workload
FD_FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "c3_fd_gradient.npz")


def c3_fd_reference():
    """The committed finite-difference reference of the 512 x 512 workload
    (tests/golden/make_c3_fd.py): the float32 weights w = dL/dI at p0 and the
    central differences of sum(w . I) per step (rows) and parameter."""
    z = np.load(FD_FIXTURE)
    size = int(z["size"])
    excluded = np.concatenate([z["excluded_class"], z["excluded_fd"]])
    nkept = size * size - excluded.size
    w = z["s16"].astype(np.float32) * np.float32(2.0 / (3 * nkept))
    return size, w, z["steps"], z["fd"], excluded


# (precision, tol, bar relative to each component).  At tol 1e-10 both forms
# are within 4.1e-6 of the differences (measured); at the workload's tol 1e-8
# the gradient of the integrated map differs from the converged one by the
# integration error itself, 1.03e-4 (FP64) and 1.04e-4 (FP32-bulk) of r_cam
# (measured): the same in both forms, so not the FP32 stages.
C3_CASES = [(A.PRR_FP64, 1e-10, 1e-5), (A.PRR_FP32_BULK, 1e-10, 1e-5), (A.PRR_FP64, 1e-8, 2e-4),
            (A.PRR_FP32_BULK, 1e-8, 2e-4)]


@pytest.mark.parametrize("precision,tol,bar", C3_CASES)
def test_c3_adjoint_at_512_matches_fp64_finite_differences(precision, tol, bar):
    """The gradient of the 512 x 512 workload's loss (Kerr a = 0.7, camera
    r = 15, inclination 75 deg, sphere at phi = 0.3 against the target render)
    w.r.t. (spin, r_cam, incl, sphere_phi), each component against central
    differences of the FP64 oracle with its FP64 colour path at tol 1e-12
    (rows of the fixture: h = 1e-6 ... 1.25e-7).  Both sides contract the
    same weights w (dL/dI at p0).  Left out of both: the pixel whose class
    changes within +-h, and the 246 pixels (0.09 %) near the shadow's edge
    whose own difference quotients had not converged at these steps (their
    rays wind near the photon orbit).  Bar: `bar` of the component (C3_CASES),
    or the finite differences' own spread over the steps, if that is larger
    (it is not: 7e-7 at most)."""
    size, w, steps, fd, excluded = c3_fd_reference()
    ctx = context(precision, tol)
    g = ctx.adjoint(CF.frame_c3(size, size), CF.c3_orbit(), w)
    best = fd[-1]
    spread = np.abs(fd - best).max(axis=0)
    rel = np.abs(g - best) / np.abs(best)
    print("adjoint", g.tolist(), "fd", best.tolist(), "relative", rel.tolist(), "fd spread", (spread / np.abs(best)).tolist(),
          "excluded pixels", int(excluded.size))
    for k in range(4):
        assert abs(g[k] - best[k]) <= max(bar * abs(best[k]), spread[k]), (k, g.tolist(), best.tolist())
