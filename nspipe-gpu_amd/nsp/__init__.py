"""Python binding of libnsp.so, the MI355X nsPipe time step (C ABI in include/nsp.h).

`Pipe` mirrors the reference driver (Main_Code/src/main.cu): construction is
the setter sequence (init_fd ... setBCon), `step` is the body of integrate()
(mod_integrate.cu:112-131), `nonlinear` is nonlinear() (mod_io_nonlinear.cu:380),
`read_state_files`/`write_state_files` are rBufferBinary/wBufferBinary.
Fields are numpy complex128 arrays of shape (NR, NT, NZ) in the reference
layout, or device pointers (int) with where=DEVICE.

There is no fallback: if libnsp.so is missing or no HIP device is present the
constructor raises.
"""
import ctypes
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libnsp.so")

FP64, MIXED, MIXED_REFINED = 0, 1, 2
IC_FLAT, IC_KOLMOGOROV, IC_PUFF, IC_ADD = 0, 1, 2, 16
HOST, DEVICE = 0, 1
FLAG_NAN, FLAG_MAXIT = 1, 2
REMESH_REFERENCE, REMESH_MIRROR = 0, 1
IO_FRICTION, IO_CROSS, IO_MEANPR = 1, 2, 4


class Params(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nt", ctypes.c_int), ("nz", ctypes.c_int),
                ("lz", ctypes.c_double), ("re", ctypes.c_double), ("dt", ctypes.c_double),
                ("d_im", ctypes.c_double), ("maxit", ctypes.c_int), ("tol", ctypes.c_double),
                ("qdes", ctypes.c_double), ("precision", ctypes.c_int)]


class StepInfo(ctypes.Structure):
    _fields_ = [("n_it", ctypes.c_int), ("err", ctypes.c_double), ("flags", ctypes.c_int)]


_lib = None
EXPORTS = ["nsp_default_params", "nsp_create", "nsp_destroy", "nsp_set_state", "nsp_get_state",
           "nsp_reference_ic", "nsp_step", "nsp_nonlinear", "nsp_diagnostics", "nsp_tables",
           "nsp_radial_tables", "nsp_read_state_files", "nsp_write_state_files", "nsp_seeded_ic",
           "nsp_profile_enable", "nsp_profile_count", "nsp_profile_name", "nsp_profile_read",
           "nsp_autotune", "nsp_tuning", "nsp_set_tuning", "nsp_stream", "nsp_synchronize",
           "nsp_last_error", "nsp_ensemble_create", "nsp_ensemble_destroy", "nsp_ensemble_size",
           "nsp_ensemble_member", "nsp_ensemble_autotune", "nsp_ensemble_step", "nsp_remesh",
           "nsp_remesh_weights", "nsp_physical_field", "nsp_cross_sections", "nsp_mean_profile",
           "nsp_io_open", "nsp_io_write", "nsp_io_close", "nsp_write_restart", "nsp_read_restart",
           "nsp_diagnostics_noalloc"]


class NspError(RuntimeError):
    pass


def lib():
    """Load libnsp.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NspError(f"libnsp.so not built at {LIB_PATH}; run `make -C nspipe-gpu_amd`")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER(ctypes.c_double)
    L.nsp_last_error.restype = ctypes.c_char_p
    L.nsp_default_params.argtypes = [ctypes.POINTER(Params), ip, ip, ip, dp, dp, dp]
    L.nsp_create.argtypes = [ctypes.POINTER(Params), ip, ctypes.POINTER(vp)]
    L.nsp_destroy.argtypes = [vp]
    L.nsp_set_state.argtypes = [vp, vp, vp, vp, ip]
    L.nsp_get_state.argtypes = [vp, vp, vp, vp, ip]
    L.nsp_reference_ic.argtypes = [vp, dp]
    L.nsp_step.argtypes = [vp, ip, ctypes.POINTER(StepInfo)]
    L.nsp_nonlinear.argtypes = [vp, vp, vp, vp, vp, vp, vp, ip]
    L.nsp_diagnostics.argtypes = [vp, P]
    L.nsp_diagnostics_noalloc.argtypes = [vp, vp]
    L.nsp_tables.argtypes = [vp, P, P, P, P, P, P, P]
    L.nsp_radial_tables.argtypes = [ip, P, P, P, P, P, P, P]
    L.nsp_read_state_files.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
    L.nsp_write_state_files.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
    L.nsp_seeded_ic.argtypes = [vp, ctypes.c_ulonglong, ip, ip, ip, dp]
    L.nsp_profile_enable.argtypes = [vp, ip]
    L.nsp_profile_count.argtypes = []
    L.nsp_profile_name.restype = ctypes.c_char_p
    L.nsp_profile_name.argtypes = [ip]
    L.nsp_profile_read.argtypes = [vp, P, ctypes.POINTER(ctypes.c_longlong)]
    L.nsp_stream.restype = vp
    L.nsp_stream.argtypes = [vp]
    L.nsp_synchronize.argtypes = [vp]
    L.nsp_autotune.argtypes = [vp, ip]
    L.nsp_tuning.restype = ctypes.c_char_p
    L.nsp_tuning.argtypes = [vp]
    L.nsp_set_tuning.argtypes = [vp, ctypes.c_char_p]
    L.nsp_ensemble_create.argtypes = [ctypes.POINTER(Params), ip, ip, ctypes.POINTER(vp)]
    L.nsp_ensemble_destroy.argtypes = [vp]
    L.nsp_ensemble_size.argtypes = [vp]
    L.nsp_ensemble_member.argtypes = [vp, ip]
    L.nsp_ensemble_member.restype = vp
    L.nsp_ensemble_autotune.argtypes = [vp, ip]
    L.nsp_ensemble_step.argtypes = [vp, ip, ctypes.POINTER(StepInfo)]
    L.nsp_remesh.argtypes = [vp, vp, ip]
    L.nsp_remesh_weights.argtypes = [ip, ip, ip, P]
    L.nsp_physical_field.argtypes = [vp, ip, vp, ip]
    L.nsp_cross_sections.argtypes = [vp, P, P]
    L.nsp_mean_profile.argtypes = [vp, P]
    L.nsp_io_open.argtypes = [vp, ctypes.c_char_p]
    L.nsp_io_write.argtypes = [vp, dp, ip]
    L.nsp_io_close.argtypes = [vp]
    L.nsp_write_restart.argtypes = [vp, ctypes.c_char_p, dp, ctypes.c_longlong]
    L.nsp_read_restart.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(dp), ctypes.POINTER(ctypes.c_longlong)]
    for name in EXPORTS:
        if name not in ("nsp_last_error", "nsp_stream", "nsp_profile_name", "nsp_tuning", "nsp_ensemble_member"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(st):
    if st != 0:
        raise NspError(f"nsp status {st}: {lib().nsp_last_error().decode()}")


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def radial_tables(nr):
    """Host-only radial tables (init_fd, get_dr0, get_dr1) for NR = nr."""
    r = np.zeros(nr); D = np.zeros(14 * nr); Lu = np.zeros(14 * nr); Lp = np.zeros(14 * nr)
    rdr = np.zeros(nr); dr0 = np.zeros(4); dr1 = np.zeros(7)
    _check(lib().nsp_radial_tables(nr, _dp(r), _dp(D), _dp(Lu), _dp(Lp), _dp(rdr), _dp(dr0), _dp(dr1)))
    return dict(r=r, D=D.reshape(2, nr, 7), Lu=Lu.reshape(2, nr, 7), Lp=Lp.reshape(2, nr, 7),
                rdr=rdr, dr0=dr0, dr1=dr1)


def remesh_weights(nr_old, nr_new, flags=REMESH_REFERENCE):
    """Host-only radial remesh matrix [2][nr_new][nr_old] (half 0 even in r, 1 odd)."""
    RM = np.zeros((2, nr_new, nr_old))
    _check(lib().nsp_remesh_weights(nr_old, nr_new, flags, _dp(RM)))
    return RM


def remesh(dst, src, flags=REMESH_REFERENCE):
    """Interpolate src's state onto dst's grid and make it dst's state."""
    _check(lib().nsp_remesh(dst.h, src.h, flags))


class Pipe:
    """One pipe on one HIP device."""

    def __init__(self, nr, nt, nz, lz, re, dt, precision=FP64, device=0, d_im=0.51, maxit=10,
                 tol=5e-8, qdes=0.5):
        p = Params()
        _check(lib().nsp_default_params(ctypes.byref(p), nr, nt, nz, lz, re, dt))
        p.d_im, p.maxit, p.tol, p.qdes, p.precision = d_im, maxit, tol, qdes, precision
        self.params = p
        self.nr, self.nt, self.nz = nr, nt, nz
        h = ctypes.c_void_p()
        _check(lib().nsp_create(ctypes.byref(p), device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nsp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self.nr, self.nt, self.nz)

    def _host(self, arrs):
        out = []
        for a in arrs:
            a = np.ascontiguousarray(a, np.complex128)
            if a.shape != self.shape:
                raise ValueError(f"field shape {a.shape} != {self.shape}")
            out.append(a)
        return out

    def set_state(self, ur, ut, uz):
        a = self._host((ur, ut, uz))
        _check(lib().nsp_set_state(self.h, *[x.ctypes.data for x in a], HOST))

    def get_state(self):
        a = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        _check(lib().nsp_get_state(self.h, *[x.ctypes.data for x in a], HOST))
        return a

    def set_state_device(self, pur, put, puz):
        _check(lib().nsp_set_state(self.h, pur, put, puz, DEVICE))

    def get_state_device(self, pur, put, puz):
        _check(lib().nsp_get_state(self.h, pur, put, puz, DEVICE))

    def reference_ic(self, a_p=0.3):
        _check(lib().nsp_reference_ic(self.h, a_p))

    def seeded_ic(self, seed, mmax=4, lmax=4, kind=0, rms=1e-3):
        _check(lib().nsp_seeded_ic(self.h, seed, mmax, lmax, kind, rms))

    def step(self, nsteps=1):
        info = (StepInfo * max(nsteps, 1))()
        _check(lib().nsp_step(self.h, nsteps, info))
        return (np.array([info[s].n_it for s in range(nsteps)], np.int32),
                np.array([info[s].err for s in range(nsteps)]),
                np.array([info[s].flags for s in range(nsteps)], np.int32))

    def nonlinear(self, ur, ut, uz):
        a = self._host((ur, ut, uz))
        b = [np.zeros(self.shape, np.complex128) for _ in range(3)]
        _check(lib().nsp_nonlinear(self.h, *[x.ctypes.data for x in a], *[x.ctypes.data for x in b], HOST))
        return b

    def diagnostics(self):
        out = np.zeros(7)
        _check(lib().nsp_diagnostics(self.h, _dp(out)))
        return dict(u_cl=out[0], u_tau=out[1], ke=out[2], div_l2=out[3], wall_max=out[4],
                    wall_div_max=out[5], flux=out[6])

    @property
    def physical_shape(self):
        ntp, nzf = 3 * self.nt // 2, 3 * (self.nz - 1)
        return (self.nr, ntp, nzf + 2)

    def physical(self, comp):
        """(NR, NTP, NZF+2) physical samples of component comp (0 ur, 1 ut, 2 uz)."""
        out = np.zeros(self.physical_shape)
        _check(lib().nsp_physical_field(self.h, comp, out.ctypes.data, HOST))
        return out

    def cross_sections(self):
        """(q, u2): cross-section integrals of ur^2+ut^2 and uz^2 at the NZF axial points."""
        nzf = 3 * (self.nz - 1)
        q, u2 = np.zeros(nzf), np.zeros(nzf)
        _check(lib().nsp_cross_sections(self.h, _dp(q), _dp(u2)))
        return q, u2

    def mean_profile(self):
        prof = np.zeros(self.nr)
        _check(lib().nsp_mean_profile(self.h, _dp(prof)))
        return prof

    def write_restart(self, directory, t=0.0, step=0):
        """Raw state files plus restart.txt (time, step, parameters) in directory."""
        _check(lib().nsp_write_restart(self.h, str(directory).encode(), t, step))

    def read_restart(self, directory):
        """Reads a restart directory into the state; returns (time, step)."""
        t = ctypes.c_double()
        s = ctypes.c_longlong()
        _check(lib().nsp_read_restart(self.h, str(directory).encode(), ctypes.byref(t), ctypes.byref(s)))
        return t.value, s.value

    def io_open(self, directory):
        _check(lib().nsp_io_open(self.h, str(directory).encode()))

    def io_write(self, t, mask=IO_FRICTION | IO_CROSS | IO_MEANPR):
        _check(lib().nsp_io_write(self.h, t, mask))

    def io_close(self):
        _check(lib().nsp_io_close(self.h))

    def tables(self):
        nr = self.nr
        r = np.zeros(nr); D = np.zeros(14 * nr); Lu = np.zeros(14 * nr); Lp = np.zeros(14 * nr)
        rdr = np.zeros(nr); dr0 = np.zeros(4); dr1 = np.zeros(7)
        _check(lib().nsp_tables(self.h, _dp(r), _dp(D), _dp(Lu), _dp(Lp), _dp(rdr), _dp(dr0), _dp(dr1)))
        return dict(r=r, D=D.reshape(2, nr, 7), Lu=Lu.reshape(2, nr, 7), Lp=Lp.reshape(2, nr, 7),
                    rdr=rdr, dr0=dr0, dr1=dr1)

    def read_state_files(self, ur, ut, uz):
        _check(lib().nsp_read_state_files(self.h, ur.encode(), ut.encode(), uz.encode()))

    def write_state_files(self, ur, ut, uz):
        _check(lib().nsp_write_state_files(self.h, ur.encode(), ut.encode(), uz.encode()))

    def profile_enable(self, on=True):
        """Per-kernel HIP-event timing on the handle's stream; clears totals."""
        _check(lib().nsp_profile_enable(self.h, 1 if on else 0))

    def profile_read(self):
        """{kernel name: (total ms, launches)} accumulated since profile_enable."""
        n = lib().nsp_profile_count()
        ms = np.zeros(n)
        cnt = (ctypes.c_longlong * n)()
        _check(lib().nsp_profile_read(self.h, _dp(ms), cnt))
        return {lib().nsp_profile_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    def autotune(self, reps=3):
        """Time the compiled kernel variants on this pipe and keep the fastest."""
        _check(lib().nsp_autotune(self.h, reps))
        return json.loads(lib().nsp_tuning(self.h).decode())

    def set_tuning(self, chosen):
        """Restore an autotuner choice (the "chosen" dict of autotune())."""
        _check(lib().nsp_set_tuning(self.h, json.dumps({"chosen": chosen}).encode()))

    def synchronize(self):
        """hipDeviceSynchronize on the pipe's device."""
        _check(lib().nsp_synchronize(self.h))

    def stream(self):
        return lib().nsp_stream(self.h)


class _Member(Pipe):
    """A pipe owned by an Ensemble (closing it is a no-op)."""

    def __init__(self, h, params):
        self.h = ctypes.c_void_p(h)
        self.params = params
        self.nr, self.nt, self.nz = params.nr, params.nt, params.nz

    def close(self):
        self.h = None


class Ensemble:
    """Independent pipes stepped together (nsp_ensemble_*): every stage is one
    launch per class of members with the same transform plans, the step
    captured in hipGraphs.  `members` is a list of dicts of Pipe keyword
    arguments."""

    def __init__(self, members, device=0):
        n = len(members)
        arr = (Params * n)()
        for j, kw in enumerate(members):
            kw = dict(kw)
            p = arr[j]
            _check(lib().nsp_default_params(ctypes.byref(p), kw.pop("nr"), kw.pop("nt"), kw.pop("nz"),
                                            kw.pop("lz"), kw.pop("re"), kw.pop("dt")))
            p.precision = kw.pop("precision", FP64)
            p.d_im = kw.pop("d_im", p.d_im)
            p.maxit = kw.pop("maxit", p.maxit)
            p.tol = kw.pop("tol", p.tol)
            p.qdes = kw.pop("qdes", p.qdes)
            if kw:
                raise TypeError(f"unknown member parameters {sorted(kw)}")
        e = ctypes.c_void_p()
        _check(lib().nsp_ensemble_create(arr, n, device, ctypes.byref(e)))
        self.e = e
        self.params = arr
        self.members = [_Member(lib().nsp_ensemble_member(e, j), arr[j]) for j in range(n)]

    def __len__(self):
        return len(self.members)

    def __getitem__(self, j):
        return self.members[j]

    def autotune(self, reps=2):
        _check(lib().nsp_ensemble_autotune(self.e, reps))

    def step(self, nsteps=1):
        """Returns (n_it, err, flags) arrays of shape (nsteps, members)."""
        n = len(self.members)
        info = (StepInfo * max(nsteps * n, 1))()
        _check(lib().nsp_ensemble_step(self.e, nsteps, info))
        k = nsteps * n
        n_it = np.array([info[q].n_it for q in range(k)], np.int32).reshape(nsteps, n)
        err = np.array([info[q].err for q in range(k)]).reshape(nsteps, n)
        flags = np.array([info[q].flags for q in range(k)], np.int32).reshape(nsteps, n)
        return n_it, err, flags

    def close(self):
        if getattr(self, "e", None):
            for m in self.members:
                m.close()
            lib().nsp_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        self.close()

