'''
Preconditioned conjugate gradients (oracle), restating pixell.utils.CG as
optweight uses it (solvers.py:127-141; pixell is not vendored in optweight):

  init : x = 0 or x0;  r = b (aliased, no copy) or b - A x0;  z = M r;
         rz = <r, z>;  rz0 = rz;  p = z;  i = 0;  err = inf
  step : Ap = A p;  alpha = rz / <p, Ap>;  x += alpha p;  r -= alpha Ap;
         z = M r;  rz' = <r, z>;  err = rz'/rz0;  beta = rz'/rz;  rz = rz';
         p = z + beta p;  i += 1
'''
import numpy as np


class CG:

    def __init__(self, A, b, x0=None, M=lambda x: x, dot=None):
        if dot is None:
            dot = lambda a, b: np.real(np.sum(a * np.conj(b)))
        self.A = A
        self.b = b
        self.M = M
        self.dot = dot
        if x0 is None:
            self.x = b * 0
            self.r = b
        else:
            self.x = x0.copy()
            self.r = b - self.A(self.x)
        z = self.M(self.r)
        self.rz = self.dot(self.r, z)
        self.rz0 = float(self.rz)
        self.p = z
        self.i = 0
        self.err = np.inf

    def step(self):
        Ap = self.A(self.p)
        alpha = self.rz / self.dot(self.p, Ap)
        self.x += alpha * self.p
        self.r -= alpha * Ap
        z = self.M(self.r)
        next_rz = self.dot(self.r, z)
        self.err = next_rz / self.rz0
        beta = next_rz / self.rz
        self.rz = next_rz
        self.p = z + beta * self.p
        self.i += 1
