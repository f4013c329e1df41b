"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the checker for the HIP path, never
the thing measured or shipped.  Complex fields are numpy complex128 arrays of
shape (NR, NT, NZ) in the reference layout (Main_Code/src/head.h:73).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        P = ctypes.POINTER(ctypes.c_double)
        IP = ctypes.POINTER(ctypes.c_int)
        L.nspo_create.restype = vp
        L.nspo_create.argtypes = [ip, ip, ip, dp, dp, dp, dp, ip, dp, dp]
        L.nspo_destroy.argtypes = [vp]
        L.nspo_set_threads.argtypes = [ip]
        L.nspo_max_threads.restype = ip
        L.nspo_set_sample_levels.argtypes = [vp, ip]
        L.nspo_tables.argtypes = [vp, P, P, P, P, P, P, P]
        L.nspo_set_state.argtypes = [vp, P, P, P]
        L.nspo_get_state.argtypes = [vp, P, P, P]
        L.nspo_reference_ic.argtypes = [vp, dp]
        L.nspo_nonlinear.argtypes = [vp, P, P, P, P, P, P]
        L.nspo_lusolve.argtypes = [vp, ip, P, P]
        L.nspo_derivr.argtypes = [vp, P, P, ip]
        L.nspo_divergence.argtypes = [vp, P, P, P, P]
        L.nspo_divwall.argtypes = [vp, P, P, P, P]
        L.nspo_addlap.argtypes = [vp, P, P, ip]
        L.nspo_addgrad.argtypes = [vp, P, P, P, P]
        L.nspo_corbcon.argtypes = [vp, P, P, P]
        L.nspo_adj_flux.argtypes = [vp, P]
        L.nspo_influence.argtypes = [vp, P, P, P, P, P, P, P]
        L.nspo_flux_qc.restype = dp
        L.nspo_flux_qc.argtypes = [vp, P]
        L.nspo_fft2_inverse.argtypes = [vp, P, P]
        L.nspo_fft2_forward.argtypes = [vp, P, P]
        L.nspo_step.argtypes = [vp, ip, IP, P]
        L.nspo_diagnostics.argtypes = [vp, P]
        L.nspo_remesh_weights.argtypes = [ip, ip, ip, P]
        L.nspo_physical.argtypes = [vp, ip, P]
        L.nspo_cross.argtypes = [vp, P, P]
        L.nspo_remesh.argtypes = [ip, ip, ip, P, ip, ip, ip, P, ip]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """FP64 CPU restatement of the reference integrator for one pipe."""

    def __init__(self, nr, nt, nz, lz, re, dt, d_im=0.51, maxit=10, tol=5e-8, qdes=0.5):
        self.nr, self.nt, self.nz = nr, nt, nz
        self.lz, self.re, self.dt = lz, re, dt
        self.ntp = 3 * nt // 2
        self.nzp = 3 * (nz - 1) // 2 + 1
        self.nzf = 2 * self.nzp - 2
        self.h = lib().nspo_create(nr, nt, nz, lz, re, dt, d_im, maxit, tol, qdes)
        if not self.h:
            raise ValueError("invalid oracle parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().nspo_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.nr, self.nt, self.nz)

    def zeros(self):
        return np.zeros(self.shape, np.complex128)

    def tables(self):
        nr = self.nr
        r = np.zeros(nr); D = np.zeros(14 * nr); Lu = np.zeros(14 * nr); Lp = np.zeros(14 * nr)
        rdr = np.zeros(nr); dr0 = np.zeros(4); dr1 = np.zeros(7)
        lib().nspo_tables(self.h, _p(r), _p(D), _p(Lu), _p(Lp), _p(rdr), _p(dr0), _p(dr1))
        return dict(r=r, D=D.reshape(2, nr, 7), Lu=Lu.reshape(2, nr, 7), Lp=Lp.reshape(2, nr, 7),
                    rdr=rdr, dr0=dr0, dr1=dr1)

    def set_state(self, ur, ut, uz):
        a = [np.ascontiguousarray(x, np.complex128) for x in (ur, ut, uz)]
        lib().nspo_set_state(self.h, *[_p(x) for x in a])

    def get_state(self):
        a = [self.zeros() for _ in range(3)]
        lib().nspo_get_state(self.h, *[_p(x) for x in a])
        return a

    def set_sample_levels(self, n):
        """Timing samples only: nonlinear() evaluates levels 0..n-1 (0: all)."""
        lib().nspo_set_sample_levels(self.h, n)

    def reference_ic(self, a_p=0.3):
        lib().nspo_reference_ic(self.h, a_p)

    def nonlinear(self, ur, ut, uz):
        a = [np.ascontiguousarray(x, np.complex128) for x in (ur, ut, uz)]
        b = [self.zeros() for _ in range(3)]
        lib().nspo_nonlinear(self.h, *[_p(x) for x in a], *[_p(x) for x in b])
        return b

    def lusolve(self, kind, rhs):
        rhs = np.ascontiguousarray(rhs, np.complex128)
        sol = self.zeros()
        lib().nspo_lusolve(self.h, kind, _p(rhs), _p(sol))
        return sol

    def derivr(self, u, flag):
        u = np.ascontiguousarray(u, np.complex128)
        du = self.zeros()
        lib().nspo_derivr(self.h, _p(u), _p(du), flag)
        return du

    def divergence(self, ur, ut, uz):
        a = [np.ascontiguousarray(x, np.complex128) for x in (ur, ut, uz)]
        dv = self.zeros()
        lib().nspo_divergence(self.h, *[_p(x) for x in a], _p(dv))
        return dv

    def divwall(self, ur, ut, uz):
        a = [np.ascontiguousarray(x, np.complex128) for x in (ur, ut, uz)]
        dv = np.zeros((self.nt, self.nz), np.complex128)
        lib().nspo_divwall(self.h, *[_p(x) for x in a], _p(dv))
        return dv

    def addlap(self, rhs, u, flag):
        rhs = np.array(rhs, np.complex128, order="C")
        u = np.ascontiguousarray(u, np.complex128)
        lib().nspo_addlap(self.h, _p(rhs), _p(u), flag)
        return rhs

    def addgrad(self, rr, rt, rz, p):
        a = [np.array(x, np.complex128, order="C") for x in (rr, rt, rz)]
        p = np.ascontiguousarray(p, np.complex128)
        lib().nspo_addgrad(self.h, *[_p(x) for x in a], _p(p))
        return a

    def corbcon(self, up, um, uz):
        a = [np.array(x, np.complex128, order="C") for x in (up, um, uz)]
        lib().nspo_corbcon(self.h, *[_p(x) for x in a])
        return a

    def adj_flux(self, uz):
        uz = np.array(uz, np.complex128, order="C")
        lib().nspo_adj_flux(self.h, _p(uz))
        return uz

    def influence(self):
        col = np.zeros((self.nt * self.nz, 4), np.complex128)
        f = [self.zeros() for _ in range(6)]
        lib().nspo_influence(self.h, _p(col), *[_p(x) for x in f])
        return col, f

    def flux_qc(self):
        uc = np.zeros(self.nr)
        qc = lib().nspo_flux_qc(self.h, _p(uc))
        return qc, uc

    def fft2_inverse(self, spec):
        spec = np.ascontiguousarray(spec, np.complex128)
        assert spec.shape == (self.ntp, self.nzp)
        phys = np.zeros((self.ntp, self.nzf))
        lib().nspo_fft2_inverse(self.h, _p(spec), _p(phys))
        return phys

    def fft2_forward(self, phys):
        phys = np.ascontiguousarray(phys, np.float64)
        spec = np.zeros((self.ntp, self.nzp), np.complex128)
        lib().nspo_fft2_forward(self.h, _p(phys), _p(spec))
        return spec

    def step(self, nsteps=1):
        n_it = np.zeros(nsteps, np.int32)
        err = np.zeros(nsteps)
        lib().nspo_step(self.h, nsteps, n_it.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(err))
        return n_it, err

    def physical(self, comp):
        """(NR, NTP, 2*NZP) physical samples of component comp (0 ur, 1 ut, 2 uz)."""
        out = np.zeros((self.nr, self.ntp, 2 * self.nzp))
        lib().nspo_physical(self.h, comp, _p(out))
        return out

    def cross(self):
        """Cross-section integrals (q of ur^2+ut^2, u2 of uz^2) at the NZF axial points."""
        q = np.zeros(self.nzf)
        u2 = np.zeros(self.nzf)
        lib().nspo_cross(self.h, _p(q), _p(u2))
        return q, u2

    def diagnostics(self):
        out = np.zeros(7)
        lib().nspo_diagnostics(self.h, _p(out))
        return dict(u_cl=out[0], u_tau=out[1], ke=out[2], div_l2=out[3], wall_max=out[4],
                    wall_div_max=out[5], flux=out[6])


def set_threads(n):
    lib().nspo_set_threads(n)


def max_threads():
    return lib().nspo_max_threads()


def remesh_weights(nr_old, nr_new, mirror=0):
    """Remesh matrix [2][nr_new][nr_old] of the reference's remesh tool."""
    RM = np.zeros((2, nr_new, nr_old))
    lib().nspo_remesh_weights(nr_old, nr_new, mirror, _p(RM))
    return RM


def remesh(fields, new_shape, mirror=0):
    """Remesh (ur, ut, uz) of shape (NR, NT, NZ) onto new_shape = (NR', NT', NZ')."""
    uo = np.ascontiguousarray(np.stack(fields), np.complex128)
    nr, nt, nz = uo.shape[1:]
    un = np.zeros((3,) + tuple(new_shape), np.complex128)
    lib().nspo_remesh(nr, nt, nz, _p(uo), *new_shape, _p(un), mirror)
    return un[0], un[1], un[2]
