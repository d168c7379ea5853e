"""ctypes mirror of include/prr.h (structures and enums of the C ABI).

The layout must match the header exactly; tests/test_abi.py checks the sizes
and offsets against the compiled library.
"""
import ctypes as C

PRR_ABI_VERSION = 3
PRR_MAX_SPHERES = 8
PRR_MAX_TEXTURES = 10
PRR_SKY2_SLOT = 9

# prr_status
PRR_OK, PRR_EINVAL, PRR_ENOMEM, PRR_EHIP, PRR_EIO, PRR_EBUDGET, PRR_ENODEV = range(7)
STATUS_NAMES = {0: "PRR_OK", 1: "PRR_EINVAL", 2: "PRR_ENOMEM", 3: "PRR_EHIP", 4: "PRR_EIO",
                5: "PRR_EBUDGET", 6: "PRR_ENODEV"}

# prr_metric
PRR_FLAT, PRR_SCHW_KS_CART, PRR_SCHW_SPH, PRR_KERR_BL, PRR_ELLIS = 0, 1, 2, 3, 4
METRIC_NAMES = {"flat": PRR_FLAT, "schwarzschild": PRR_SCHW_KS_CART, "schwarzschild-cart": PRR_SCHW_KS_CART,
                "schwarzschild-sph": PRR_SCHW_SPH, "kerr": PRR_KERR_BL, "wormhole": PRR_ELLIS}
# prr_solver (values of the reference's SOLVER enum, main/base.h:26-34)
PRR_EULER, PRR_RK4, PRR_RKF45, PRR_DORMAND = 0, 1, 2, 5
# prr_precision
PRR_FP64, PRR_FP32_BULK = 0, 1
# prr_ssaa
PRR_SSAA_NONE, PRR_SSAA_REGULAR, PRR_SSAA_RANDOM, PRR_SSAA_JITTERED = 0, 1, 2, 4
# prr_projection
PRR_PROJ_PINHOLE, PRR_PROJ_SPHERICAL = 0, 1
# prr_filter
PRR_FILTER_SQAVG, PRR_FILTER_BILINEAR = 0, 1
# prr_class
PRR_ESCAPE, PRR_HORIZON, PRR_SPHERE, PRR_STEP_CAP, PRR_NAN = range(5)
CLASS_NAMES = ["ESCAPE", "HORIZON", "SPHERE", "STEP_CAP", "NAN"]


class Tracer(C.Structure):
    _fields_ = [("solver", C.c_int32), ("precision", C.c_int32),
                ("h", C.c_double), ("hmin", C.c_double), ("hmax", C.c_double),
                ("atol", C.c_double), ("rtol", C.c_double),
                ("r_escape", C.c_double),
                ("horizon_eps_rel", C.c_double), ("horizon_eps_abs", C.c_double),
                ("lambda_max", C.c_double),
                ("max_steps", C.c_int64), ("max_attempts", C.c_int64),
                ("r_fp64", C.c_double)]


class Sphere(C.Structure):
    _fields_ = [("c", C.c_double * 3), ("radius", C.c_double), ("tex", C.c_int32), ("pad", C.c_int32)]


class Frame(C.Structure):
    _fields_ = [("pos", C.c_double * 3), ("alpha", C.c_double), ("beta", C.c_double),
                ("vfov", C.c_double), ("width", C.c_int32), ("height", C.c_int32),
                ("ssaa", C.c_int32), ("samples", C.c_int32), ("projection", C.c_int32),
                ("fxaa", C.c_int32), ("mass", C.c_double), ("spin", C.c_double),
                ("n_spheres", C.c_int32), ("seed", C.c_uint32),
                ("spheres", Sphere * PRR_MAX_SPHERES)]

    @property
    def nrays(self):
        return self.width * self.height * self.samples


class Image(C.Structure):
    _fields_ = [("rgb8", C.c_void_p), ("rgbf", C.c_void_p), ("on_device", C.c_int32), ("pad", C.c_int32)]


class RayOut(C.Structure):
    _fields_ = [("cls", C.c_void_p), ("endpoint", C.c_void_p), ("steps", C.c_void_p),
                ("attempts", C.c_void_p), ("color", C.c_void_p), ("on_device", C.c_int32),
                ("pad", C.c_int32)]


class Launch(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("refill", C.c_int32),
                ("checkpoint", C.c_int32), ("adjoint_record_bytes", C.c_int64)]


class Stats(C.Structure):
    _fields_ = [("rays", C.c_int64), ("attempts", C.c_int64), ("accepted", C.c_int64),
                ("attempts_fp64", C.c_int64), ("pixels", C.c_int64),
                ("integrate_ms", C.c_double), ("ssaa_ms", C.c_double), ("fxaa_ms", C.c_double),
                ("total_ms", C.c_double), ("flops_fp64", C.c_double), ("flops_fp32", C.c_double),
                ("ssaa_bytes", C.c_double), ("fxaa_bytes", C.c_double), ("adjoint_ms", C.c_double),
                ("adjoint_steps", C.c_int64), ("init_ms", C.c_double), ("shade_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Orbit(C.Structure):
    _fields_ = [("spin", C.c_double), ("r_cam", C.c_double), ("incl", C.c_double),
                ("sphere_phi", C.c_double), ("cam_phi", C.c_double), ("sphere_r", C.c_double),
                ("sphere_radius", C.c_double), ("sphere_tex", C.c_int32), ("has_sphere", C.c_int32)]


def make_tracer(**kw):
    """Tracer with the reference's defaults (prr.cpp:184-190, kernels/solver.cuh:84,158,181)."""
    t = Tracer()
    defaults = dict(solver=PRR_RKF45, precision=PRR_FP64, h=0.01, hmin=1e-4, hmax=0.1, atol=1e-3, rtol=0.0,
                    r_escape=70.0, horizon_eps_rel=0.0, horizon_eps_abs=1e-3, lambda_max=1000.0,
                    max_steps=10**9, max_attempts=10**6, r_fp64=3.0)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(t, k, v)
    return t


def make_frame(pos, alpha, beta, vfov, width, height, ssaa=PRR_SSAA_NONE, samples=1,
               projection=PRR_PROJ_PINHOLE, fxaa=0, mass=1.0, spin=0.0, spheres=(), seed=0):
    f = Frame()
    f.seed = seed
    f.pos[:] = list(pos)
    f.alpha, f.beta, f.vfov = alpha, beta, vfov
    f.width, f.height = width, height
    f.ssaa, f.samples, f.projection, f.fxaa = ssaa, samples, projection, fxaa
    f.mass, f.spin = mass, spin
    f.n_spheres = len(spheres)
    for i, s in enumerate(spheres):
        c, radius, tex = s
        f.spheres[i].c[:] = list(c)
        f.spheres[i].radius = radius
        f.spheres[i].tex = tex
    return f
