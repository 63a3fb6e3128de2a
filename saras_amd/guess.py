"""ctypes wrapper of the learned initial guess network (include/saras_guess.h).

The network runs and trains on the GPU only; this module moves host arrays in and out for
tests and drivers.  Layer order and tensor ids follow include/saras_guess.h."""
import ctypes as C

import numpy as np

from . import _lib as _L

LAYER_NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "b_a", "b_b", "d2", "d1", "d0a", "out"]
TENSOR_NAMES = ["x0", "a0", "s0", "x1", "a1", "s1", "x2", "a2", "s2", "x3", "a3", "bb", "u2", "g2", "u1", "g1",
                "u0", "g0"]
IN_CH = _L.SR_GUESS_IN_CH


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def layer_info(layer):
    ci, co, lv = C.c_int(), C.c_int(), C.c_int()
    _L.check(_L.load().sr_guess_layer_info(layer, C.byref(ci), C.byref(co), C.byref(lv)))
    return ci.value, co.value, lv.value


class GuessNet:
    def __init__(self, n, seed=0, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, fine=None, history=0):
        """history: converged (rhs, solution) pairs kept for the linear predictor (0: off)."""
        self.L = _L.load()
        d = _L.SrGuessDesc()
        d.n[0], d.n[1], d.n[2] = n
        if fine is not None:
            d.fine[0], d.fine[1], d.fine[2] = fine
        d.seed, d.lr, d.beta1, d.beta2, d.eps = seed, lr, beta1, beta2, eps
        d.history = int(history)
        h = C.c_void_p()
        _L.check(self.L.sr_guess_create(C.byref(d), C.byref(h)))
        self.h = h
        self.n = tuple(n)

    def close(self):
        if self.h:
            self.L.sr_guess_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_layer(self, layer, W, b):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        ci, co, _ = layer_info(layer)
        assert W.shape == (27, co, ci) and b.shape == (co,)
        _L.check(self.L.sr_guess_set_layer(self.h, layer, _fptr(W), _fptr(b)))

    def get_layer(self, layer):
        ci, co, _ = layer_info(layer)
        W = np.empty((27, co, ci), np.float32)
        b = np.empty(co, np.float32)
        _L.check(self.L.sr_guess_get_layer(self.h, layer, _fptr(W), _fptr(b)))
        return W, b

    def get_grad(self, layer):
        ci, co, _ = layer_info(layer)
        W = np.empty((27, co, ci), np.float32)
        b = np.empty(co, np.float32)
        _L.check(self.L.sr_guess_get_grad(self.h, layer, _fptr(W), _fptr(b)))
        return W, b

    def forward(self, x0):
        x0 = np.ascontiguousarray(x0, dtype=np.float32)
        assert x0.shape == self.n + (IN_CH,)
        y = np.empty(self.n, np.float32)
        _L.check(self.L.sr_guess_forward(self.h, _fptr(x0), _fptr(y)))
        return y

    def backward(self, dy):
        dy = np.ascontiguousarray(dy, dtype=np.float32)
        assert dy.shape == self.n
        _L.check(self.L.sr_guess_backward(self.h, _fptr(dy)))

    def adam(self):
        _L.check(self.L.sr_guess_adam(self.h))

    def tensor(self, which):
        if isinstance(which, str):
            which = TENSOR_NAMES.index(which)
        cnt = C.c_long()
        _L.check(self.L.sr_guess_tensor(self.h, which, None, C.byref(cnt)))
        out = np.empty(cnt.value, np.float32)
        _L.check(self.L.sr_guess_tensor(self.h, which, _fptr(out), C.byref(cnt)))
        return out

    # -- inside a sequence of solves (device pointers, padded layout) ---------------------------
    def predict(self, solver, p1_ptr, p2_ptr, rhs_ptr, out_ptr):
        st = _L.SrGuessStats()
        _L.check(self.L.sr_guess_predict(self.h, solver.h, C.c_void_p(p1_ptr),
                                         C.c_void_p(p2_ptr) if p2_ptr else None, C.c_void_p(rhs_ptr),
                                         C.c_void_p(out_ptr), C.byref(st)))
        return st

    def train(self, pstar_ptr):
        loss = C.c_double()
        _L.check(self.L.sr_guess_train(self.h, C.c_void_p(pstar_ptr), C.byref(loss)))
        return loss.value

    def kappa(self):
        k, t = C.c_double(), C.c_int()
        _L.check(self.L.sr_guess_kappa(self.h, C.byref(k), C.byref(t)))
        return k.value, t.value

    def observe(self, solver, rhs_ptr, pstar_ptr):
        """Add the converged pair of the solve just finished to the linear predictor's history."""
        _L.check(self.L.sr_guess_observe(self.h, solver.h, C.c_void_p(rhs_ptr), C.c_void_p(pstar_ptr)))

    def extrapolate(self, p1_ptr, p2_ptr, out_ptr):
        _L.check(self.L.sr_guess_extrapolate(self.h, C.c_void_p(p1_ptr), C.c_void_p(p2_ptr) if p2_ptr else None,
                                             C.c_void_p(out_ptr)))

    def time_layer(self, layer, which, reps=5):
        """(average ms, flops) of one layer's convolution (which 0), adjoint (1) or weight gradient (2)"""
        ms, fl = C.c_double(), C.c_double()
        _L.check(self.L.sr_guess_time_layer(self.h, layer, which, reps, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def set_policy(self, policy):
        """0: best guess by residual, network every time; 1: learned always; 2: adaptive network use;
        3: never the network (base guess alone)."""
        _L.check(self.L.sr_guess_set_policy(self.h, int(policy)))
