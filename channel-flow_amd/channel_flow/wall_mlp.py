"""Learned wall model: a 4 -> 64 -> 64 -> 64 -> 1 tanh MLP trained on the
GPU (BF16 matrix-core forward/backward kernels in ``cf_mlp.hip``) on pairs
(u_par, h, nu, dP/dx) -> tau_w produced by the EQWM kernel itself
(``solve_u_eqwm``, eqwm.f90:90-97).  It stands in for the per-point wall ODE
of ``compute_equilibrium_wall_model`` (wallmodel.f90:591-685); wall points
whose inputs leave the box of the training features are solved by the ODE.

The reference has no learned wall model; this module is an addition whose
errors are measured against the EQWM kernel (and through it the
flang-compiled eqwm.f90).
"""
import ctypes

import numpy as np

from . import _lib
from .cases import reb_from_retau

NPARAMS_LAYOUT = (("W1", (4, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)),
                  ("W3", (64, 64)), ("b3", (64,)), ("W4", (64,)), ("b4", (1,)))


def unpack(theta):
    """Split the flat FP32 parameter vector into named arrays ([in][out] weights)."""
    out, o = {}, 0
    for name, shp in NPARAMS_LAYOUT:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n]).reshape(shp)
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name], dtype=np.float32).ravel()
                           for name, _ in NPARAMS_LAYOUT])


class WallMLP:
    """Handle on a device-resident model (``cf_mlp``)."""

    def __init__(self, seed=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self.lib.cf_mlp_create(ctypes.c_uint(seed), ctypes.byref(h)), "cf_mlp_create")
        self.h = h
        self.nparams = self.lib.cf_mlp_nparams()

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.cf_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ----
    @property
    def params(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        _lib.check(self.lib.cf_mlp_get_params(self.h, _lib.fptr(w)), "cf_mlp_get_params")
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparams:
            raise ValueError("expected %d parameters, got %d" % (self.nparams, w.size))
        _lib.check(self.lib.cf_mlp_set_params(self.h, _lib.fptr(w)), "cf_mlp_set_params")

    @property
    def norm(self):
        nrm = np.zeros(16, dtype=np.float32)
        hull = np.zeros(8, dtype=np.float32)
        _lib.check(self.lib.cf_mlp_get_norm(self.h, _lib.fptr(nrm), _lib.fptr(hull)), "cf_mlp_get_norm")
        return nrm, hull

    def set_norm(self, nrm, hull):
        nrm = np.ascontiguousarray(nrm, dtype=np.float32)
        hull = np.ascontiguousarray(hull, dtype=np.float32)
        if nrm.size != 16 or hull.size != 8:
            raise ValueError("norm has 16 entries and hull 8")
        _lib.check(self.lib.cf_mlp_set_norm(self.h, _lib.fptr(nrm), _lib.fptr(hull)), "cf_mlp_set_norm")

    # ---- training / inference ----
    @staticmethod
    def _feat(feat):
        feat = np.ascontiguousarray(feat, dtype=np.float64)
        if feat.ndim != 2 or feat.shape[1] != 4:
            raise ValueError("features are an (n, 4) array of (u_par, h, nu, dP/dx)")
        return feat

    def train(self, feat, tau, epochs=20, batch=4096, lr=3e-3, lr_final=1e-4):
        """Fit normalisation + hull to the data, then Adam over mini-batches;
        returns the per-epoch loss (MSE of standardised ln tau)."""
        feat = self._feat(feat)
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        if tau.shape != (feat.shape[0],):
            raise ValueError("tau must have one value per feature row")
        hist = np.zeros(epochs, dtype=np.float32)
        _lib.check(self.lib.cf_mlp_train(self.h, feat.shape[0], _lib.dptr(feat), _lib.dptr(tau), int(epochs),
                                         int(batch), float(lr), float(lr_final), _lib.fptr(hist)),
                   "cf_mlp_train")
        return hist

    def gradients(self, feat, tau):
        """Gradient of the batch loss (no update), and the summed squared error."""
        feat = self._feat(feat)
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _lib.check(self.lib.cf_mlp_gradients(self.h, feat.shape[0], _lib.dptr(feat), _lib.dptr(tau),
                                             _lib.fptr(g), _lib.fptr(loss)), "cf_mlp_gradients")
        return g, float(loss[0])

    def predict(self, feat):
        """tau for each row, and whether the row lies inside the training hull."""
        feat = self._feat(feat)
        n = feat.shape[0]
        tau = np.zeros(n)
        inside = np.zeros(n, dtype=np.int32)
        _lib.check(self.lib.cf_mlp_predict(self.h, n, _lib.dptr(feat), _lib.dptr(tau), _lib.iptr(inside)),
                   "cf_mlp_predict")
        return tau, inside.astype(bool)


def eqwm_samples(n, seed=2024, retau=(1e3, 1e4), h=(0.03, 0.3), ncv=30, stretch=1.10, tol=1e-8,
                 max_iter=100):
    """Training pairs generated by the EQWM kernel.

    For each sample: Re_tau log-uniform in ``retau``, nu = 1/Re_b (Dean's
    correlation), u_tau = Re_tau nu (half-height 1), matching height h
    uniform in ``h``, u_par = u_tau (ln(h+)/0.41 + 5.2) scaled by U[0.3, 1.8]
    (instantaneous fluctuations about the log law), and the mean-pressure
    forcing dPdx = u_tau^2 U[-1, 3] (the solver's forcing fluctuates about
    u_tau^2 under constant mass flow; the ODE itself ignores it: it has no
    pressure-gradient term, eqwm.f90:163).  Points where the ODE does not
    converge are dropped.  Returns (feat (m, 4), tau).
    """
    from .solver import eqwm_batch
    rng = np.random.default_rng(seed)
    rt = np.exp(rng.uniform(np.log(retau[0]), np.log(retau[1]), n))
    nu = 1.0 / reb_from_retau(rt)
    ut = rt * nu
    hh = rng.uniform(h[0], h[1], n)
    u = ut * (np.log(hh * rt) / 0.41 + 5.2) * rng.uniform(0.3, 1.8, n)
    dp = ut ** 2 * rng.uniform(-1.0, 3.0, n)
    tau, it = eqwm_batch(u, nu, hh, ut ** 2, ncv=ncv, stretch=stretch, tol=tol, max_iter=max_iter)
    ok = np.isfinite(tau) & (tau > 0) & (it < max_iter)
    feat = np.stack([u, hh, nu, dp], axis=1)[ok]
    return np.ascontiguousarray(feat), tau[ok]
