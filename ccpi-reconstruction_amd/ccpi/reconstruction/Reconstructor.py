"""``ccpi.reconstruction.Reconstructor``: the parameter-driven front end of
the reference (Wrappers/Python/ccpi/reconstruction/Reconstructor.py:16-160)
over ``parallelbeam.alg``.

The reference module imports ``ccpi.common.CCPiBaseClass`` and
``ccpi.reconstruction.FindCenterOfRotation``, neither of which its package
contains; both are provided here:
  * ``CCPiBaseClass``: a parameter store (``setParameter`` accepts only the
    keywords a class lists in ``acceptedInputKeywords``; ``getParameter``
    takes a name or a list of names and raises KeyError for unset ones) with
    a ``log`` that prints when ``debug`` is set;
  * ``find_center_vo``: a centre-of-rotation estimate for 180-degree
    parallel-beam scans after Vo, Atwood and Drakopoulos (Opt. Express 22,
    2014, 19078), written from the spectral property it rests on (below).
"""

from __future__ import annotations

import numpy as np

from .conebeam import alg as cbalg  # noqa: F401  (same imports as the reference)
from .parallelbeam import alg as pbalg


class CCPiBaseClass:
    def __init__(self, **kwargs):
        self.acceptedInputKeywords = []
        self.pars = {}
        self.debug = bool(kwargs.get("debug", False))

    def setParameter(self, **kwargs):
        for key, value in kwargs.items():
            if key not in self.acceptedInputKeywords:
                raise KeyError(f"Wrong parameter {key} for {type(self).__name__}")
            self.pars[key] = value

    def getParameter(self, key):
        if isinstance(key, str):
            return self.pars[key]
        return [self.pars[k] for k in key]

    def log(self, msg):
        if self.debug:
            print(f"{type(self).__name__}: {msg}")


def _bowtie_metric(sino, centre, radius, drop):
    """Spectral energy of the 360-degree sinogram that a 180-degree scan
    implies for rotation centre ``centre`` (pixel-centre index units).

    A parallel-beam scan satisfies p(theta + pi, s) = p(theta, 2c - s), so
    the second half-turn is the first mirrored about c.  For an object
    inside radius R the 2-D Fourier transform of a full-turn sinogram is
    confined to |k| <= R |omega| (k the angular harmonic, omega the detector
    frequency in radians per pixel; Edholm, Lewitt and Lindholm, Proc. SPIE
    671, 1986).  A wrong c breaks the sinogram's continuity at theta = pi
    and spreads energy into |k| > R |omega|: the mean magnitude there is the
    metric.  Rows |k| <= drop and the omega = 0 column are left out."""
    na, n = sino.shape
    src = 2.0 * centre - np.arange(n, dtype=np.float64)
    i0 = np.floor(src).astype(np.int64)
    w = src - i0
    inside = (src >= 0) & (src <= n - 1)
    lo, hi = np.clip(i0, 0, n - 1), np.clip(i0 + 1, 0, n - 1)
    flip = np.where(inside, sino[:, lo] * (1.0 - w) + sino[:, hi] * w, 0.0)
    full = np.vstack([sino, flip])
    spec = np.abs(np.fft.fft2(full))
    k = np.abs(np.fft.fftfreq(2 * na, 1.0 / (2 * na)))[:, None]
    om = np.abs(2.0 * np.pi * np.fft.fftfreq(n))[None, :]
    out = (k > radius * om) & (k > drop) & (om > 0)
    return float(spec[out].mean()) if out.any() else np.inf


def find_center_vo(tomo, ind=None, smin=-40, smax=40, srad=10, step=0.5, ratio=2.0, drop=20):
    """Rotation centre of a [angle][v][h] 180-degree parallel-beam scan, as
    the ``rotation_centre`` of the ``parallelbeam`` functions (the axis at
    that many pixels from the detector's left edge).  Row ``ind`` (default:
    the middle row) is searched over shifts [smin, smax] pixels from the
    detector centre in whole pixels, then +-srad around the best in steps of
    ``step``; the object is taken to fill 1/ratio of the detector width."""
    t = np.asarray(tomo, dtype=np.float64)
    if t.ndim != 3:
        raise ValueError("tomo must be [angle][v][h]")
    row = t.shape[1] // 2 if ind is None else int(ind)
    sino = t[:, row, :]
    n = sino.shape[1]
    mid = (n - 1) / 2.0
    radius = 0.5 * n / float(ratio)
    coarse = np.arange(int(smin), int(smax) + 1, dtype=np.float64)
    best = coarse[int(np.argmin([_bowtie_metric(sino, mid + s, radius, drop)
                                 for s in coarse]))]
    fine = np.arange(best - srad, best + srad + 0.5 * step, step)
    best = fine[int(np.argmin([_bowtie_metric(sino, mid + s, radius, drop) for s in fine]))]
    return mid + best + 0.5


class Reconstructor(CCPiBaseClass):
    """Abstract reconstructor; concrete classes provide a Factory."""

    def __init__(self, **kwargs):
        CCPiBaseClass.__init__(self, **kwargs)
        self.acceptedInputKeywords = ["instrument"]

    @staticmethod
    def create(type, **kwargs):
        if type in Reconstructor.__subclasses__():
            return type.Factory.create(**kwargs) if kwargs else type.Factory.create()
        if type in IterativeReconstructor.__subclasses__():
            return type.Factory.create(**kwargs)
        raise TypeError(f'Reconstructor "{type}" not available')

    def iterate(self, **kwargs):
        return NotImplementedError("iterate must be implemented in the concrete class")

    @staticmethod
    def printAvailableAlgorithms():
        msg = "Available Iterative Reconstruction Algorithms:\n"
        for cls in Reconstructor.__subclasses__():
            if cls is not IterativeReconstructor:
                msg += f"{cls.__name__}, instantiate as Reconstructor.create({cls.__name__}, **kwargs)\n"
        for name in IterativeReconstructor.getAvailableAlgorithms():
            msg += f"{name}, instantiate as Reconstructor.create({name}, **kwargs)\n"
        return msg

    class Factory:
        @staticmethod
        def create(**kwargs):
            return NotImplemented

    def logCurrentConfiguration(self):
        for key in self.pars:
            value = self.getParameter(key)
            if key == "normalized_projections":
                value = np.shape(value)
            self.log(f"{key}={value}")


class IterativeReconstructor(Reconstructor):
    """Runs one of the ``parallelbeam.alg`` solvers on normalised
    projections ([angle][v][h]); estimates the centre of rotation with
    ``find_center_vo`` when none is given (Reconstructor.py:74-160)."""

    class Factory:
        @staticmethod
        def create(algorithm, **kwargs):
            return IterativeReconstructor(algorithm=algorithm, **kwargs)

    def __init__(self, **kwargs):
        debug = kwargs.pop("debug", False)
        Reconstructor.__init__(self, debug=debug)
        self.acceptedInputKeywords = [
            "algorithm", "projection_data", "normalized_projections", "angles",
            "center_of_rotation", "flat_field", "iterations", "dark_field", "resolution",
            "isLogScale", "threads", "iterationValues", "regularization_parameter"]
        self.pars = {"algorithm": pbalg.cgls, "resolution": 1, "isLogScale": False,
                     "threads": 16, "iterations": 8}
        if kwargs:
            self.log(kwargs)
            self.setParameter(**kwargs)
        self.log(f"{self.getParameter('algorithm').__name__} created")

    @staticmethod
    def getAvailableAlgorithms():
        return ["cgls", "sirt", "mlem", "cgls_conv", "cgls_TVreg", "cgls_tikhonov"]

    def iterate(self, **kwargs):
        if kwargs:
            self.setParameter(**kwargs)
        try:
            proj = self.getParameter("normalized_projections")
        except KeyError:
            raise Exception("Insufficient data. Please pass the normalized projections or "
                            "flat, dark and projections")
        self.logCurrentConfiguration()
        try:
            angles = self.getParameter("angles")
        except KeyError:
            raise Exception("Please pass the angles")
        try:
            cor = self.getParameter("center_of_rotation")
        except KeyError:
            cor = find_center_vo(proj)
            self.log(f"estimated center of rotation {cor}")
            self.setParameter(center_of_rotation=cor)
        resolution, niter, threads, is_log = self.getParameter(
            ["resolution", "iterations", "threads", "isLogScale"])
        algorithm = self.getParameter("algorithm")
        name = algorithm.__name__
        if name in ("cgls", "sirt", "mlem"):
            return algorithm(proj, angles, cor, resolution, niter, threads, is_log)
        if name == "cgls_conv":
            values = np.zeros((niter,))
            return algorithm(proj, angles, cor, resolution, niter, threads, values, is_log)
        if name in ("cgls_tikhonov", "cgls_TVreg"):
            values = np.zeros((niter,))
            lam = self.getParameter("regularization_parameter")
            return algorithm(proj, angles, cor, resolution, niter, threads, lam, values, is_log)
        raise ValueError(f"algorithm {name} is not one of {self.getAvailableAlgorithms()}")
