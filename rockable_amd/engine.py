"""Python host mirror of the engine's C ABI (include/rockable_amd.h).

`Engine` follows the reference's Rockable box: load a conf deck (which runs
the initial accelerations() "fake step", Rockable.cpp:1133-1138), rebuild the
neighbour list, run steps of the integrate() loop, save conf files.  All
arithmetic runs in librockable_amd.so on the GPU; if the library or the GPU is
missing, construction fails with an error -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librockable_amd.so")

INTERACTION_DTYPE = np.dtype([
    ("i", "<i4"), ("j", "<i4"), ("type", "<i4"), ("isub", "<i4"), ("jsub", "<i4"), ("stick", "<i4"),
    ("n", "<f8", 3), ("dn", "<f8"), ("pos", "<f8", 3), ("vel", "<f8", 3), ("fn", "<f8"),
    ("ft", "<f8", 3), ("mom", "<f8", 3), ("damp", "<f8"),
])


class RkCounts(ctypes.Structure):
    _fields_ = [("nbodies", ctypes.c_int64), ("ndriven", ctypes.c_int64), ("nlisted", ctypes.c_int64),
                ("nactive", ctypes.c_int64), ("ninterfaces", ctypes.c_int64), ("nl_updates", ctypes.c_int64),
                ("steps", ctypes.c_int64), ("t", ctypes.c_double), ("dt", ctypes.c_double),
                ("total_broken_area", ctypes.c_double), ("last_candidates", ctypes.c_int64),
                ("guard_reads", ctypes.c_int64), ("guard_wasted", ctypes.c_int64),
                ("graph_launches", ctypes.c_int64), ("graph_captures", ctypes.c_int64),
                ("guard_polls", ctypes.c_int64)]


class RkStats(ctypes.Structure):
    _fields_ = [("kinetic_translational", ctypes.c_double), ("kinetic_rotational", ctypes.c_double),
                ("vmax", ctypes.c_double), ("nfree", ctypes.c_int64), ("nactive", ctypes.c_int64),
                ("ncontacts", ctypes.c_int64), ("active_pairs", ctypes.c_int64), ("coordination", ctypes.c_double),
                ("dn_mean", ctypes.c_double), ("dn_min", ctypes.c_double), ("ninterfaces", ctypes.c_int64),
                ("broken_interfaces", ctypes.c_int64), ("total_broken_area", ctypes.c_double)]


class RkConfInfo(ctypes.Structure):
    _fields_ = [("nbodies", ctypes.c_int64), ("ndriven", ctypes.c_int64), ("nshapes", ctypes.c_int64),
                ("ninteractions", ctypes.c_int64), ("ninterfaces", ctypes.c_int64), ("ngroups", ctypes.c_int64),
                ("law", ctypes.c_int32), ("linkcells", ctypes.c_int32)]


EXPORTED = ["rk_inspect_conf", "rk_engine_create", "rk_engine_destroy", "rk_last_error", "rk_load_conf", "rk_load_driving_system",
            "rk_load_data_extractors",
            "rk_set_force_law", "rk_set_body_force", "rk_set_option", "rk_update_nl", "rk_accelerations",
            "rk_step", "rk_integrate", "rk_save_conf", "rk_synchronize", "rk_get_counts", "rk_get_bodies",
            "rk_set_bodies", "rk_get_interactions", "rk_get_kernel_time", "rk_reset_kernel_time",
            "rk_mlp_create", "rk_mlp_destroy", "rk_mlp_last_error", "rk_mlp_nparams", "rk_mlp_get_params",
            "rk_mlp_set_params", "rk_mlp_forward", "rk_mlp_loss_grad", "rk_mlp_train", "rk_set_rebuild_model",
            "rk_guard_features", "rk_get_option", "rk_autotune", "rk_get_stats", "rk_get_stats_noalloc", "rk_load_batch",
            "rk_batch_packing"]

_lib = None


def load_library(path=LIB_PATH):
    """Load librockable_amd.so and declare the C ABI; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("librockable_amd.so is not built (run `python -m rockable_amd.build`)")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    L.rk_inspect_conf.argtypes = [ctypes.c_char_p, ctypes.POINTER(RkConfInfo), ctypes.c_char_p, ctypes.c_size_t]
    L.rk_engine_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.rk_engine_destroy.argtypes = [P]
    L.rk_last_error.argtypes = [P]
    L.rk_last_error.restype = ctypes.c_char_p
    L.rk_load_conf.argtypes = [P, ctypes.c_char_p]
    L.rk_load_driving_system.argtypes = [P, ctypes.c_char_p]
    L.rk_load_data_extractors.argtypes = [P, ctypes.c_char_p]
    L.rk_set_force_law.argtypes = [P, ctypes.c_char_p]
    L.rk_set_body_force.argtypes = [P, ctypes.c_char_p, P, ctypes.c_int]
    L.rk_set_option.argtypes = [P, ctypes.c_char_p, ctypes.c_double]
    L.rk_update_nl.argtypes = [P]
    L.rk_accelerations.argtypes = [P]
    L.rk_step.argtypes = [P, ctypes.c_int64]
    L.rk_integrate.argtypes = [P, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.rk_save_conf.argtypes = [P, ctypes.c_char_p]
    L.rk_synchronize.argtypes = [P]
    L.rk_get_counts.argtypes = [P, ctypes.POINTER(RkCounts)]
    L.rk_get_bodies.argtypes = [P] + [P] * 8 + [ctypes.c_size_t]
    L.rk_set_bodies.argtypes = [P] + [P] * 4 + [ctypes.c_size_t]
    L.rk_get_interactions.argtypes = [P, ctypes.c_int, P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.rk_get_kernel_time.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_double)]
    L.rk_reset_kernel_time.argtypes = [P]
    F = ctypes.POINTER(ctypes.c_float)
    L.rk_mlp_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_uint64]
    L.rk_mlp_destroy.argtypes = [P]
    L.rk_mlp_last_error.argtypes = [P]
    L.rk_mlp_last_error.restype = ctypes.c_char_p
    L.rk_mlp_nparams.argtypes = []
    L.rk_mlp_get_params.argtypes = [P, F, ctypes.c_size_t]
    L.rk_mlp_set_params.argtypes = [P, F, ctypes.c_size_t]
    L.rk_mlp_forward.argtypes = [P, F, ctypes.c_size_t, F]
    L.rk_mlp_loss_grad.argtypes = [P, F, F, ctypes.c_size_t, F, F]
    L.rk_mlp_train.argtypes = [P, F, F, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                               F]
    L.rk_set_rebuild_model.argtypes = [P, P]
    L.rk_guard_features.argtypes = [P, F]
    L.rk_get_option.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    L.rk_autotune.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.rk_get_stats.argtypes = [P, ctypes.POINTER(RkStats)]
    L.rk_get_stats_noalloc.argtypes = [P, ctypes.POINTER(RkStats)]
    L.rk_load_batch.argtypes = [P, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    L.rk_batch_packing.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def inspect_conf(path):
    """The deck reader alone, on the host (no device): a dict of the deck's
    body, shape, interaction, interface and group counts, its force law and
    neighbour-list method.  Raises RuntimeError with the reader's message."""
    L = load_library()
    info = RkConfInfo()
    err = ctypes.create_string_buffer(512)
    if L.rk_inspect_conf(os.fspath(path).encode(), ctypes.byref(info), err, len(err)) != 0:
        raise RuntimeError(err.value.decode())
    return {k: getattr(info, k) for k, _ in RkConfInfo._fields_}


class Engine:
    """One Rockable box on one GPU."""

    def __init__(self, conf_path=None, device=0, driving=None):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.rk_engine_create(ctypes.byref(h), device)
        if rc != 0:
            raise RuntimeError("rk_engine_create failed (%d): no usable GPU device %d" % (rc, device))
        self.h = h
        if conf_path is not None:
            self.load_conf(conf_path)
        if driving is not None:
            self.load_driving_system(driving)

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self.L.rk_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.rk_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_batch(self, conf_paths, driving_paths=None):
        """Independent packings (decks lying apart in space) as one batch."""
        n = len(conf_paths)
        cp = (ctypes.c_char_p * n)(*[os.fspath(x).encode() for x in conf_paths])
        dp = None
        if driving_paths is not None:
            dp = (ctypes.c_char_p * n)(*[None if x is None else os.fspath(x).encode() for x in driving_paths])
        self._check(self.L.rk_load_batch(self.h, cp, dp, n))

    def packing(self, p):
        """(first driven, driven count, first free, free count) of packing p."""
        out = (ctypes.c_int64 * 4)()
        self._check(self.L.rk_batch_packing(self.h, int(p), out))
        return tuple(int(x) for x in out)

    def packing_rows(self, p):
        fd, nd, ff, nf = self.packing(p)
        return np.concatenate([np.arange(fd, fd + nd), np.arange(ff, ff + nf)])

    def load_conf(self, path):
        self._check(self.L.rk_load_conf(self.h, os.fspath(path).encode()))

    def load_driving_system(self, path):
        self._check(self.L.rk_load_driving_system(self.h, os.fspath(path).encode()))

    def load_data_extractors(self, path):
        """dataExtractors.txt (dnStat, TrackDamage), recorded by integrate()."""
        self._check(self.L.rk_load_data_extractors(self.h, os.fspath(path).encode()))

    def set_force_law(self, name):
        return self._check(self.L.rk_set_force_law(self.h, name.encode()))

    def set_body_force(self, name, args=()):
        a = np.ascontiguousarray(args, dtype=np.float64)
        self._check(self.L.rk_set_body_force(self.h, name.encode(), _ptr(a), len(a)))

    def set_option(self, name, value):
        self._check(self.L.rk_set_option(self.h, name.encode(), float(value)))

    def stats(self):
        """Kinetic energies, contact and bond statistics reduced on the device."""
        st = RkStats()
        self._check(self.L.rk_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in RkStats._fields_}

    def get_option(self, name):
        v = ctypes.c_double(0.0)
        self._check(self.L.rk_get_option(self.h, name.encode(), ctypes.byref(v)))
        return v.value

    def conf_scalar(self, name):
        """A scalar of the loaded deck: dt, interVerlet, interConf, DVerlet,
        dVerlet or tmax."""
        return self.get_option(name)

    TUNED = ("fuse_law", "type_sort", "skip_segment", "law_group", "geom_block", "gather_lanes", "gather_groups", "fuse_glob", "gather_waves",
             "body_block", "persistent",
             "graph_steps")

    def autotune(self, steps_per_trial=100):
        """Pick the launch shapes by timing them on the running simulation
        (the steps taken are real steps); returns (steps taken, choices)."""
        n = ctypes.c_int64(0)
        self._check(self.L.rk_autotune(self.h, int(steps_per_trial), ctypes.byref(n)))
        return int(n.value), {k: int(self.get_option(k)) for k in self.TUNED}

    def update_nl(self):
        self._check(self.L.rk_update_nl(self.h))

    def accelerations(self):
        self._check(self.L.rk_accelerations(self.h))

    def step(self, n=1):
        self._check(self.L.rk_step(self.h, int(n)))

    def integrate(self, outdir=None, max_steps=-1):
        done = ctypes.c_int64(0)
        self._check(self.L.rk_integrate(self.h, os.fspath(outdir).encode() if outdir else None, int(max_steps),
                                        ctypes.byref(done)))
        return int(done.value)

    def save_conf(self, path):
        self._check(self.L.rk_save_conf(self.h, os.fspath(path).encode()))

    def synchronize(self):
        self._check(self.L.rk_synchronize(self.h))

    def counts(self):
        c = RkCounts()
        self._check(self.L.rk_get_counts(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in RkCounts._fields_}

    def bodies(self):
        n = self.counts()["nbodies"]
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "vrot", "acc", "arot", "force", "moment")}
        out["Q"] = np.zeros((n, 4))
        self._check(self.L.rk_get_bodies(self.h, _ptr(out["pos"]), _ptr(out["Q"]), _ptr(out["vel"]),
                                         _ptr(out["vrot"]), _ptr(out["acc"]), _ptr(out["arot"]),
                                         _ptr(out["force"]), _ptr(out["moment"]), n))
        return out

    def set_bodies(self, pos=None, Q=None, vel=None, vrot=None):
        n = self.counts()["nbodies"]
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        pos, Q, vel, vrot = c(pos), c(Q), c(vel), c(vrot)
        self._check(self.L.rk_set_bodies(self.h, _ptr(pos), _ptr(Q), _ptr(vel), _ptr(vrot), n))

    def interactions(self, active=False):
        n = ctypes.c_size_t(0)
        self._check(self.L.rk_get_interactions(self.h, 1 if active else 0, None, 0, ctypes.byref(n)))
        arr = np.zeros(n.value, dtype=INTERACTION_DTYPE)
        if n.value:
            self._check(self.L.rk_get_interactions(self.h, 1 if active else 0, _ptr(arr), n.value, ctypes.byref(n)))
        return arr

    def kernel_time(self, name):
        k = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        self._check(self.L.rk_get_kernel_time(self.h, name.encode(), ctypes.byref(k), ctypes.byref(ms)))
        return int(k.value), float(ms.value)

    def reset_kernel_time(self):
        self._check(self.L.rk_reset_kernel_time(self.h))

    def set_rebuild_model(self, mlp):
        """Size the batches between guard reads by the network (None: fixed)."""
        self._model = mlp
        self._check(self.L.rk_set_rebuild_model(self.h, mlp.h if mlp is not None else None))

    def guard_features(self):
        out = np.zeros(16, dtype=np.float32)
        self._check(self.L.rk_guard_features(self.h, _fptr(out)))
        return out


def _fptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The rebuild-interval network (16 -> 64 -> 64 -> 1, FP32) on a GPU."""

    def __init__(self, device=0, seed=1):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.rk_mlp_create(ctypes.byref(h), device, seed)
        if rc != 0:
            raise RuntimeError("rk_mlp_create failed (%d): no usable GPU device %d" % (rc, device))
        self.h = h
        self.nparams = self.L.rk_mlp_nparams()

    def _check(self, rc):
        if rc < 0:
            raise RuntimeError(self.L.rk_mlp_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.rk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.L.rk_mlp_get_params(self.h, _fptr(out), out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.rk_mlp_set_params(self.h, _fptr(p), p.size))

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        y = np.zeros(len(X), dtype=np.float32)
        self._check(self.L.rk_mlp_forward(self.h, _fptr(X), len(X), _fptr(y)))
        return y

    def loss_grad(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        self._check(self.L.rk_mlp_loss_grad(self.h, _fptr(X), _fptr(t), len(X), _fptr(loss), _fptr(g)))
        return float(loss[0]), g

    def train(self, X, t, epochs=50, batch=256, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.L.rk_mlp_train(self.h, _fptr(X), _fptr(t), len(X), epochs, batch, lr, seed, _fptr(loss)))
        return float(loss[0])
