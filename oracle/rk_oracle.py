"""ctypes binding of the CPU oracle (oracle/rk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librk_oracle.so")


class RkoInteraction(ctypes.Structure):
    _fields_ = [
        ("i", ctypes.c_int32), ("j", ctypes.c_int32), ("type", ctypes.c_int32),
        ("isub", ctypes.c_int32), ("jsub", ctypes.c_int32), ("stick", ctypes.c_int32),
        ("n", ctypes.c_double * 3), ("dn", ctypes.c_double), ("pos", ctypes.c_double * 3),
        ("vel", ctypes.c_double * 3), ("fn", ctypes.c_double), ("ft", ctypes.c_double * 3),
        ("mom", ctypes.c_double * 3), ("damp", ctypes.c_double),
    ]


INTERACTION_DTYPE = np.dtype([
    ("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("isub", "<i4"), ("jsub", "<i4"), ("stick", "<i4"),
    ("n", "<f8", 3), ("dn", "<f8"), ("pos", "<f8", 3), ("vel", "<f8", 3), ("fn", "<f8"),
    ("ft", "<f8", 3), ("mom", "<f8", 3), ("damp", "<f8"),
])
assert INTERACTION_DTYPE.itemsize == ctypes.sizeof(RkoInteraction)


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no FMA contraction)."""
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "rk_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.rko_create.restype = P
        L.rko_destroy.argtypes = [P]
        L.rko_last_error.argtypes = [P]
        L.rko_last_error.restype = ctypes.c_char_p
        L.rko_set_threads.argtypes = [P, ctypes.c_int]
        L.rko_set_verbose.argtypes = [P, ctypes.c_int]
        L.rko_load_conf.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
        L.rko_read_driving.argtypes = [P, ctypes.c_char_p]
        L.rko_load_data_extractors.argtypes = [P, ctypes.c_char_p]
        L.rko_update_nl.argtypes = [P]
        L.rko_accelerations.argtypes = [P]
        L.rko_update_kinematics.argtypes = [P]
        L.rko_resultants_only.argtypes = [P]
        L.rko_vv_step.argtypes = [P]
        L.rko_integrate.argtypes = [P, ctypes.c_char_p, ctypes.c_int64]
        L.rko_integrate.restype = ctypes.c_int64
        L.rko_run_steps.argtypes = [P, ctypes.c_int64]
        L.rko_run_steps.restype = ctypes.c_int64
        L.rko_save_conf.argtypes = [P, ctypes.c_char_p]
        for name in ("rko_nbodies", "rko_ndriven", "rko_nlisted", "rko_nactive", "rko_ninterfaces"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = ctypes.c_size_t
        L.rko_time.argtypes = [P]
        L.rko_time.restype = ctypes.c_double
        L.rko_nl_updates.argtypes = [P]
        L.rko_nl_updates.restype = ctypes.c_int64
        L.rko_total_broken_area.argtypes = [P]
        L.rko_total_broken_area.restype = ctypes.c_double
        L.rko_get_bodies.argtypes = [P] + [P] * 8
        L.rko_get_mass.argtypes = [P, P, P]
        L.rko_set_bodies.argtypes = [P] + [P] * 4
        L.rko_get_interactions.argtypes = [P, ctypes.c_int, P, ctypes.c_size_t]
        L.rko_get_interactions.restype = ctypes.c_size_t
        L.rko_set_variant.argtypes = [ctypes.c_int, ctypes.c_int]
        L.rko_get_variant.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One oracle simulation (a Rockable box restated on the CPU)."""

    def __init__(self, conf_path, threads=1, fake_step=True, driving=None, verbose=0):
        self.L = lib()
        self.s = self.L.rko_create()
        self.L.rko_set_threads(self.s, threads)
        self.L.rko_set_verbose(self.s, verbose)
        if self.L.rko_load_conf(self.s, conf_path.encode(), 1 if fake_step else 0) != 0:
            err = self.L.rko_last_error(self.s).decode()
            self.close()
            raise RuntimeError("oracle load failed: " + err)
        if driving:
            if self.L.rko_read_driving(self.s, driving.encode()) != 0:
                raise RuntimeError("oracle driving: " + self.L.rko_last_error(self.s).decode())

    def load_data_extractors(self, path):
        """dataExtractors.txt (dnStat, TrackDamage), recorded by integrate()."""
        if self.L.rko_load_data_extractors(self.s, path.encode()) != 0:
            raise RuntimeError(self.L.rko_last_error(self.s).decode())

    def close(self):
        if self.s:
            self.L.rko_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nbodies(self):
        return int(self.L.rko_nbodies(self.s))

    @property
    def ndriven(self):
        return int(self.L.rko_ndriven(self.s))

    @property
    def nlisted(self):
        return int(self.L.rko_nlisted(self.s))

    @property
    def nactive(self):
        return int(self.L.rko_nactive(self.s))

    @property
    def ninterfaces(self):
        return int(self.L.rko_ninterfaces(self.s))

    @property
    def nl_updates(self):
        return int(self.L.rko_nl_updates(self.s))

    @property
    def total_broken_area(self):
        return float(self.L.rko_total_broken_area(self.s))

    @property
    def time(self):
        return float(self.L.rko_time(self.s))

    def update_nl(self):
        self.L.rko_update_nl(self.s)

    def update_kinematics(self):
        self.L.rko_update_kinematics(self.s)

    def resultants_only(self):
        self.L.rko_resultants_only(self.s)

    def accelerations(self):
        self.L.rko_accelerations(self.s)

    def integrate(self, outdir=None, max_steps=-1):
        n = self.L.rko_integrate(self.s, outdir.encode() if outdir else None, max_steps)
        if n < 0:
            raise RuntimeError(self.L.rko_last_error(self.s).decode())
        return int(n)

    def run_steps(self, n):
        return int(self.L.rko_run_steps(self.s, n))

    def save_conf(self, path):
        if self.L.rko_save_conf(self.s, path.encode()) != 0:
            raise RuntimeError(self.L.rko_last_error(self.s).decode())

    def bodies(self):
        n = self.nbodies
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "acc", "vrot", "arot", "force", "moment")}
        out["Q"] = np.zeros((n, 4))
        self.L.rko_get_bodies(self.s, _ptr(out["pos"]), _ptr(out["vel"]), _ptr(out["acc"]), _ptr(out["Q"]),
                              _ptr(out["vrot"]), _ptr(out["arot"]), _ptr(out["force"]), _ptr(out["moment"]))
        return out

    def mass(self):
        """Masses and principal inertia moments (body frame) of the bodies."""
        m, I = np.zeros(self.nbodies), np.zeros((self.nbodies, 3))
        self.L.rko_get_mass(self.s, _ptr(m), _ptr(I))
        return m, I

    def set_bodies(self, pos=None, vel=None, Q=None, vrot=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        pos, vel, Q, vrot = c(pos), c(vel), c(Q), c(vrot)
        self.L.rko_set_bodies(self.s, _ptr(pos), _ptr(vel), _ptr(Q), _ptr(vrot))

    def interactions(self, active=False):
        n = self.L.rko_get_interactions(self.s, 1 if active else 0, None, 0)
        arr = np.zeros(n, dtype=INTERACTION_DTYPE)
        if n:
            self.L.rko_get_interactions(self.s, 1 if active else 0, _ptr(arr), n)
        return arr


def set_variant(which, value):
    lib().rko_set_variant(which, value)


def get_variant(which):
    return lib().rko_get_variant(which)


def compare_conf(new_path, reg_path):
    """compareConf (src/Apps/run.cpp:38-92): every line after the first line
    containing 'Particles' must be textually identical.  Returns (ok, first
    differing line number relative to that line, reg line, new line)."""
    with open(reg_path) as f:
        reg = f.read().split("\n")
    with open(new_path) as f:
        new = f.read().split("\n")

    def after(lines):
        for k, line in enumerate(lines):
            if "Particles" in line:
                return lines[k + 1:]
        return None

    r, n = after(reg), after(new)
    if r is None or n is None:
        return False, -1, None, None
    # getline semantics: a trailing newline does not produce an extra line
    if r and r[-1] == "":
        r = r[:-1]
    for k, line in enumerate(r):
        other = n[k] if k < len(n) else ""
        if line != other:
            return False, k, line, other
    return True, None, None, None
