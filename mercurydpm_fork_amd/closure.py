"""The learned closure: a 4 -> 32 -> 32 -> 2 network from the standardised
coarse-grained state (inertial number I, packing fraction phi, pressure P,
granular temperature T) to (mu, phi), trained and evaluated by the HIP
kernels of csrc/closure_mlp.hip through include/mercurydpm_closure.h.

`inertial_number` and `mu_of_I` give the quantities the data come from: the
local rheology of dense granular flow (GDR MiDi 2004; Jop, Forterre & Pouliquen
2006), which a chute's CG profile follows in steady flow."""
import ctypes as C

import numpy as np

from . import dem

N_PARAMS = 1282
_P = C.c_void_p


def _lib():
    lib = dem.load_hip_library()
    if not getattr(lib, "_closure_declared", False):
        for name, args in (("dem_mlp_create", [C.POINTER(_P), C.c_int32]), ("dem_mlp_destroy", [_P]),
                           ("dem_mlp_set_params", [_P, _P]), ("dem_mlp_get_params", [_P, _P]),
                           ("dem_mlp_forward", [_P, C.c_int64, _P, _P]),
                           ("dem_mlp_gradient", [_P, C.c_int64, _P, _P, _P, _P]),
                           ("dem_mlp_train", [_P, C.c_int64, _P, _P, C.c_int32, C.c_double, _P])):
            fn = getattr(lib, name)
            fn.restype = C.c_int32
            fn.argtypes = args
        lib._closure_declared = True
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class MLP:
    """The closure network on one GPU."""

    def __init__(self, device: int = 0, seed: int = 0):
        self.lib = _lib()
        self.h = _P()
        self._check(self.lib.dem_mlp_create(C.byref(self.h), device))
        self.set_params(glorot(seed))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"closure MLP call failed ({rc})")

    def close(self):
        if self.h:
            self.lib.dem_mlp_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float64)
        assert p.shape == (N_PARAMS,)
        self._check(self.lib.dem_mlp_set_params(self.h, _ptr(p)))

    def params(self):
        p = np.empty(N_PARAMS)
        self._check(self.lib.dem_mlp_get_params(self.h, _ptr(p)))
        return p

    def __call__(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        y = np.empty((len(x), 2))
        self._check(self.lib.dem_mlp_forward(self.h, len(x), _ptr(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float64).reshape(-1, 2)
        g = np.empty(N_PARAMS)
        loss = np.zeros(1)
        self._check(self.lib.dem_mlp_gradient(self.h, len(x), _ptr(x), _ptr(t), _ptr(g), _ptr(loss)))
        return float(loss[0]), g

    def train(self, x, t, epochs: int, lr: float = 3e-3) -> float:
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float64).reshape(-1, 2)
        loss = np.zeros(1)
        self._check(self.lib.dem_mlp_train(self.h, len(x), _ptr(x), _ptr(t), int(epochs), float(lr), _ptr(loss)))
        return float(loss[0])


def glorot(seed: int = 0):
    """Glorot-uniform weights and zero biases in the kernels' layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in ((32, 4), (32, 0), (32, 32), (32, 0), (2, 32), (2, 0)):
        if fan_in:
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, fan_out * fan_in))
        else:
            parts.append(np.zeros(fan_out))
    return np.concatenate(parts)


def inertial_number(shear_rate, d, pressure, density):
    """I = gamma_dot d / sqrt(P / rho_p)."""
    return shear_rate * d / np.sqrt(pressure / density)


def mu_of_I(I, mu1=np.tan(np.radians(20.9)), mu2=np.tan(np.radians(32.8)), I0=0.279):
    """mu(I) = mu1 + (mu2 - mu1) / (I0 / I + 1) (Jop, Forterre & Pouliquen 2006)."""
    return mu1 + (mu2 - mu1) / (I0 / I + 1.0)


def phi_of_I(I, phi_max=0.6, phi_min=0.5):
    """phi(I) = phi_max - (phi_max - phi_min) I (GDR MiDi 2004), for I < 1."""
    return phi_max - (phi_max - phi_min) * I


# ---------------------------------------------------------------------------
# The configured network: 4 -> 3 x 128 GELU -> 2, BF16 weights, FP32
# accumulation, matrix-core kernels (csrc/closure_mfma.hip).
# ---------------------------------------------------------------------------
N_PARAMS128 = 33922
_SHAPES128 = [(128, 4), (128,), (128, 128), (128,), (128, 128), (128,), (2, 128), (2,)]


def _lib128():
    lib = dem.load_hip_library()
    if not getattr(lib, "_closure128_declared", False):
        for name, args in (("dem_closure_create", [C.POINTER(_P), C.c_int32]), ("dem_closure_destroy", [_P]),
                           ("dem_closure_set_params", [_P, _P]), ("dem_closure_get_params", [_P, _P]),
                           ("dem_closure_forward", [_P, C.c_int64, _P, _P]),
                           ("dem_closure_gradient", [_P, C.c_int64, _P, _P, _P, _P]),
                           ("dem_closure_train", [_P, C.c_int64, _P, _P, C.c_int32, C.c_float, _P])):
            fn = getattr(lib, name)
            fn.restype = C.c_int32
            fn.argtypes = args
        lib._closure128_declared = True
    return lib


def glorot128(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights and zero biases, float32, in the kernels' layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape in _SHAPES128:
        if len(shape) == 2:
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            parts.append(np.zeros(shape[0]))
    return np.concatenate(parts).astype(np.float32)


class Closure:
    """The configured closure network on one GPU (FP32 arrays in and out)."""

    def __init__(self, device: int = 0, seed: int = 0):
        self.lib = _lib128()
        self.h = _P()
        self._check(self.lib.dem_closure_create(C.byref(self.h), device))
        self.set_params(glorot128(seed))

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"closure network call failed ({rc})")

    def close(self):
        if self.h:
            self.lib.dem_closure_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (N_PARAMS128,)
        self._check(self.lib.dem_closure_set_params(self.h, _ptr(p)))

    def params(self):
        p = np.empty(N_PARAMS128, np.float32)
        self._check(self.lib.dem_closure_get_params(self.h, _ptr(p)))
        return p

    def __call__(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        y = np.empty((len(x), 2), np.float32)
        self._check(self.lib.dem_closure_forward(self.h, len(x), _ptr(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        g = np.empty(N_PARAMS128, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.dem_closure_gradient(self.h, len(x), _ptr(x), _ptr(t), _ptr(g), _ptr(loss)))
        return float(loss[0]), g

    def train(self, x, t, epochs: int, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, 4)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 2)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.dem_closure_train(self.h, len(x), _ptr(x), _ptr(t), int(epochs), float(lr), _ptr(loss)))
        return float(loss[0])
