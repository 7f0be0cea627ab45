"""Learned cut-cell quadrature (include/moris_gpu_learned.h) from Python."""
import ctypes

import numpy as np

from . import MgpuError, _check, load_library

N_INPUTS, N_HIDDEN, N_POINTS, N_MOMENTS = 8, 64, 27, 23
PLANES, SPHERES, GYROIDS = 1, 2, 4

_vp = ctypes.c_void_p


def _lib():
    lib = load_library()
    if not getattr(lib, "_lq_typed", False):
        lib.mgpu_lq_dataset.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32, _vp, _vp, _vp]
        lib.mgpu_lq_operators.argtypes = [_vp, _vp, _vp]
        lib.mgpu_lq_fit.argtypes = [_vp, _vp, ctypes.c_int32, _vp, _vp]
        lib.mgpu_mlp_create.argtypes = [ctypes.c_uint64, ctypes.POINTER(_vp)]
        lib.mgpu_mlp_destroy.argtypes = [_vp]
        lib.mgpu_mlp_num_params.argtypes = [ctypes.POINTER(ctypes.c_int32)]
        lib.mgpu_mlp_get_params.argtypes = [_vp, _vp]
        lib.mgpu_mlp_set_params.argtypes = [_vp, _vp]
        lib.mgpu_mlp_forward.argtypes = [_vp, _vp, ctypes.c_int64, _vp]
        lib.mgpu_mlp_gradients.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp, _vp]
        lib.mgpu_mlp_train.argtypes = [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                       ctypes.c_uint64, _vp]
        lib.mgpu_set_quadrature_model.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int32, ctypes.c_double]
        lib.mgpu_set_lq_result.argtypes = [_vp, _vp, _vp, _vp]
        for n in ("mgpu_lq_dataset", "mgpu_lq_operators", "mgpu_lq_fit", "mgpu_mlp_create", "mgpu_mlp_destroy",
                  "mgpu_mlp_num_params", "mgpu_mlp_get_params", "mgpu_mlp_set_params", "mgpu_mlp_forward",
                  "mgpu_mlp_gradients", "mgpu_mlp_train", "mgpu_set_quadrature_model", "mgpu_set_lq_result"):
            getattr(lib, n).restype = ctypes.c_int32
        lib._lq_typed = True
    return lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def dataset(n, seed, kinds=PLANES | SPHERES | GYROIDS, with_moments=False):
    """(x[n][8], y[n][27], moments[n][27] or None) from the XTK-restatement data generator; the
    moments are the 27 tri-quadratic ones (a fastest), the fit uses the 23 of degree <= 4."""
    x = np.zeros((n, N_INPUTS), np.float32)
    y = np.zeros((n, N_POINTS), np.float32)
    m = np.zeros((n, N_POINTS), np.float64) if with_moments else None
    _check(_lib().mgpu_lq_dataset(n, seed, kinds, _p(x), _p(y), _p(m)), "mgpu_lq_dataset")
    return x, y, m


MASK23 = np.array([a + b + c <= 4 for c in range(3) for b in range(3) for a in range(3)])


def gm7():
    """Grundmann-Moeller degree-7 tetrahedron rule: barycentric points [35][4], weights [35]."""
    lib = _lib()
    lib.mgpu_lq_gm7.argtypes = [_vp, _vp]
    lib.mgpu_lq_gm7.restype = ctypes.c_int32
    lam = np.zeros((35, 4))
    w = np.zeros(35)
    _check(lib.mgpu_lq_gm7(_p(lam), _p(w)), "mgpu_lq_gm7")
    return lam, w


def operators():
    V = np.zeros((N_MOMENTS, N_POINTS))
    P = np.zeros((N_POINTS, N_MOMENTS))
    g = np.zeros(N_POINTS)
    _check(_lib().mgpu_lq_operators(_p(V), _p(P), _p(g)), "mgpu_lq_operators")
    return V, P, g


def fit(m, w0, iters=0):
    m = np.ascontiguousarray(m, np.float64)
    w0 = np.ascontiguousarray(w0, np.float64)
    w = np.zeros(N_POINTS)
    r = ctypes.c_double()
    _check(_lib().mgpu_lq_fit(_p(m), _p(w0), iters, _p(w), ctypes.byref(r)), "mgpu_lq_fit")
    return w, r.value


class MLP:
    """8 -> 64 -> 64 -> 27 SiLU network trained and evaluated by the HIP kernels."""

    def __init__(self, seed=1):
        self._h = _vp()
        _check(_lib().mgpu_mlp_create(seed, ctypes.byref(self._h)), "mgpu_mlp_create")
        n = ctypes.c_int32()
        _lib().mgpu_mlp_num_params(ctypes.byref(n))
        self.n_params = n.value

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(_lib().mgpu_mlp_get_params(self._h, _p(p)), "mgpu_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (self.n_params,):
            raise ValueError("expected %d parameters" % self.n_params)
        _check(_lib().mgpu_mlp_set_params(self._h, _p(p)), "mgpu_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], N_POINTS), np.float32)
        _check(_lib().mgpu_mlp_forward(self._h, _p(x), x.shape[0], _p(y)), "mgpu_mlp_forward")
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = ctypes.c_float()
        _check(_lib().mgpu_mlp_gradients(self._h, _p(x), _p(t), x.shape[0], _p(g), ctypes.byref(loss)),
               "mgpu_mlp_gradients")
        return g, loss.value

    def train(self, x, t, epochs=10, batch=1024, lr=3e-3, seed=7):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros(epochs, np.float32)
        _check(_lib().mgpu_mlp_train(self._h, _p(x), _p(t), x.shape[0], epochs, batch, lr, seed, _p(hist)),
               "mgpu_mlp_train")
        return hist

    def close(self):
        if self._h:
            _lib().mgpu_mlp_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gauss_prior():
    """A network whose output is the HEX_3x3x3 Gauss weights / 8 for every input (all weights zero,
    last bias the Gauss weights): the moment fit then starts from the volume-scaled Gauss rule.
    Stands in for the trained network to show what the network changes."""
    _, _, g = operators()
    net = MLP(seed=1)
    p = np.zeros(net.n_params, np.float32)
    p[-N_POINTS:] = (g / 8.0).astype(np.float32)
    net.params = p
    return net


def attach(set_, mlp, tol=1e-6, iters=0, kappa=4.0):
    """Attach a network to a Set (keeps a reference so it outlives the set's calls).  Elements
    stay learned when their FP32 weights reproduce the moments to tol and sum |w| <= kappa * volume."""
    _check(_lib().mgpu_set_quadrature_model(set_._h, mlp._h, tol, iters, kappa), "mgpu_set_quadrature_model")
    set_._mlp = mlp


def result(set_):
    n_cut = set_.n_cut if hasattr(set_, "n_cut") else None
    if n_cut is None:
        raise MgpuError("set has no n_cut")
    wp = np.zeros((n_cut, N_POINTS), np.float32)
    wf = np.zeros((n_cut, N_POINTS), np.float32)
    fl = np.zeros(n_cut, np.int32)
    _check(_lib().mgpu_set_lq_result(set_._h, _p(wp), _p(wf), _p(fl)), "mgpu_set_lq_result")
    return wp, wf, fl


def cell_moments(phi15, phase=1):
    """27 tri-quadratic moments of the `phase` region of a unit cell from its 15 vertex values."""
    lib = _lib()
    lib.mgpu_lq_cell_moments.argtypes = [_vp, ctypes.c_int32, _vp, _vp]
    lib.mgpu_lq_cell_moments.restype = ctypes.c_int32
    phi15 = np.ascontiguousarray(phi15, np.float64)
    m = np.zeros(N_POINTS)
    ns = np.zeros(1, np.int32)
    _check(lib.mgpu_lq_cell_moments(_p(phi15), phase, _p(m), _p(ns)), "mgpu_lq_cell_moments")
    return m, int(ns[0])
