"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, as the checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
BASELINE_BIN = os.path.join(HERE, "build", "cpu_baseline")

ATOM_DTYPE = np.dtype([
    ("id", "<u8"), ("type", "<i4"), ("_pad", "<i4"),
    ("x", "<f8", (3,)), ("v", "<f8", (3,)), ("f", "<f8", (3,)),
    ("rho", "<f8"), ("df", "<f8"),
])

_lib = None

PHASES = {"first": 0, "decide": 1, "exchange_inter": 2, "border_inter": 3, "exchange_atom": 4, "clear_force": 5,
          "compute_eam": 6, "second": 7, "make_index": 8, "lat_rho": 9, "inter_rho": 10, "rho_reverse": 11,
          "lat_df": 12, "df_forward": 13, "lat_force": 14, "inter_force": 15, "force_reverse": 16,
          "exchange_atom_first": 17}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, d, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        sig = {
            "oracle_create": (vp, [vp, d, d, ctypes.c_char_p, d, ctypes.c_char_p, i32]),
            "oracle_destroy": (None, [vp]),
            "oracle_build_world": (None, [vp, ctypes.c_uint32, d, vp, ctypes.c_uint32]),
            "oracle_prepare": (None, [vp]),
            "oracle_run": (None, [vp, i64]),
            "oracle_set_dt": (None, [vp, d]),
            "oracle_plant": (None, [vp, i32]),
            "oracle_setv": (None, [vp, vp, vp, d]),
            "oracle_collision": (None, [vp, vp, vp, d]),
            "oracle_phase": (i32, [vp, i32]),
            "oracle_num_slots": (i64, [vp]),
            "oracle_atoms": (vp, [vp]),
            "oracle_lattice_dims": (None, [vp, vp]),
            "oracle_ghost": (i32, [vp]),
            "oracle_cutoff": (d, [vp]),
            "oracle_inter_count": (i64, [vp]),
            "oracle_inter_get": (None, [vp, vp]),
            "oracle_inter_set": (None, [vp, vp, i64]),
            "oracle_inter_ghost_count": (i64, [vp]),
            "oracle_inter_ghost_get": (None, [vp, vp]),
            "oracle_mvv": (d, [vp]),
            "oracle_adaptive_dt": (d, [vp, d, d]),
            "oracle_temperature": (d, [vp]),
            "oracle_rescale": (None, [vp, d]),
            "oracle_energy": (None, [vp, vp]),
            "oracle_system_force": (None, [vp, vp]),
            "oracle_vcm": (None, [vp, vp]),
            "oracle_timers": (None, [vp, vp]),
            "oracle_nei_make": (i32, [i64, i64, i64, i32, d, vp, vp, vp, vp, i64, vp]),
            "oracle_nei_is_positive": (i32, [d, d, d]),
            "oracle_ws_voronoy": (None, [vp, d, vp]),
            "oracle_ws_is_out_box": (ctypes.c_uint, [vp, d, vp, vp]),
            "oracle_pot_load": (vp, [ctypes.c_char_p, ctypes.c_char_p, i32]),
            "oracle_pot_free": (None, [vp]),
            "oracle_pot_eval": (d, [vp, i32, i32, i32, d, d, d]),
            "oracle_pot_table": (i64, [vp, i32, i32, vp, i64]),
            "oracle_sim_pot": (vp, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def nei_make(sx, sy, sz, cut_lattice, f):
    L = load()
    cap = 4096
    bufs = [np.zeros(cap, dtype=np.int64) for _ in range(4)]
    counts = np.zeros(4, dtype=np.int64)
    rc = L.oracle_nei_make(sx, sy, sz, cut_lattice, f, *[_p(b) for b in bufs], cap, _p(counts))
    assert rc == 0
    return [b[:c].copy() for b, c in zip(bufs, counts)]


def is_positive(x, y, z):
    return bool(load().oracle_nei_is_positive(x, y, z))


def voronoy(x, a):
    out = np.zeros(3, dtype=np.int64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    load().oracle_ws_voronoy(_p(xx), a, _p(out))
    return out


def is_out_box(x, a, low, sub):
    xx = np.ascontiguousarray(x, dtype=np.float64)
    lo = np.ascontiguousarray(low, dtype=np.int64)
    sb = np.ascontiguousarray(sub, dtype=np.int64)
    return int(load().oracle_ws_is_out_box(_p(xx), a, _p(lo), _p(sb)))


class Potential:
    def __init__(self, path):
        L = load()
        err = ctypes.create_string_buffer(256)
        self.p = L.oracle_pot_load(str(path).encode(), err, 256)
        if not self.p:
            raise RuntimeError(err.value.decode())
        self.L = L

    def __del__(self):
        try:
            self.L.oracle_pot_free(self.p)
        except Exception:
            pass

    def charge_density(self, Z, r2):
        return self.L.oracle_pot_eval(self.p, 0, Z, 0, r2, 0, 0)

    def d_embed(self, Z, rho):
        return self.L.oracle_pot_eval(self.p, 1, Z, 0, rho, 0, 0)

    def to_force(self, Zi, Zj, r2, dfi, dfj):
        return self.L.oracle_pot_eval(self.p, 2, Zi, Zj, r2, dfi, dfj)

    def embed(self, Z, rho):
        return self.L.oracle_pot_eval(self.p, 3, Z, 0, rho, 0, 0)

    def pair_phi(self, Zi, Zj, r2):
        return self.L.oracle_pot_eval(self.p, 4, Zi, Zj, r2, 0, 0)

    def table(self, kind, idx):
        n = self.L.oracle_pot_table(self.p, kind, idx, None, 0)
        out = np.zeros(n, dtype=np.float64)
        self.L.oracle_pot_table(self.p, kind, idx, _p(out), n)
        return out


class Sim:
    """The oracle's single-rank simulation (reference phase order)."""

    def __init__(self, cells, a, f, setfl, dt=0.001):
        L = load()
        n = np.array(cells, dtype=np.int64)
        err = ctypes.create_string_buffer(256)
        self.h = L.oracle_create(_p(n), a, f, str(setfl).encode() if setfl else b"", dt, err, 256)
        if not self.h:
            raise RuntimeError(err.value.decode())
        self.L = L
        self._n = n
        self.cells = tuple(int(c) for c in cells)
        self.ghost = L.oracle_ghost(self.h)
        dims = np.zeros(3, dtype=np.int64)
        L.oracle_lattice_dims(self.h, _p(dims))
        self.dims = tuple(int(d) for d in dims)
        self.rc = L.oracle_cutoff(self.h)

    def __del__(self):
        try:
            self.L.oracle_destroy(self.h)
        except Exception:
            pass

    def build_world(self, seed, tset, ratio, alloy_seed=1):
        r = np.array(ratio, dtype=np.int32)
        self.L.oracle_build_world(self.h, seed, tset, _p(r), alloy_seed)

    def prepare(self):
        self.L.oracle_prepare(self.h)

    def run(self, steps):
        self.L.oracle_run(self.h, steps)

    def set_dt(self, dt):
        self.L.oracle_set_dt(self.h, dt)

    def plant(self, which):
        """Plant mistake `which` in decide (0 = none); see Sim::plant."""
        self.L.oracle_plant(self.h, which)

    def phase(self, name):
        return self.L.oracle_phase(self.h, PHASES[name])

    def setv(self, lat, direction, e):
        l4 = np.array(lat, dtype=np.int32)
        d3 = np.array(direction, dtype=np.float64)
        self.L.oracle_setv(self.h, _p(l4), _p(d3), e)

    def collision(self, lat, direction, e):
        l4 = np.array(lat, dtype=np.int32)
        d3 = np.array(direction, dtype=np.float64)
        self.L.oracle_collision(self.h, _p(l4), _p(d3), e)

    def atoms(self):
        """Live view of the ghost-extended AoS array (104-byte slots)."""
        n = self.L.oracle_num_slots(self.h)
        ptr = self.L.oracle_atoms(self.h)
        buf = (ctypes.c_char * (n * 104)).from_address(ptr)
        return np.frombuffer(buf, dtype=ATOM_DTYPE, count=n)

    def sub_box(self):
        g = self.ghost
        a = self.atoms().reshape(self.dims[2], self.dims[1], self.dims[0])
        return a[g:g + self.cells[2], g:g + self.cells[1], 2 * g:2 * g + 2 * self.cells[0]]

    def inter(self):
        n = self.L.oracle_inter_count(self.h)
        out = np.zeros(n, dtype=ATOM_DTYPE)
        if n:
            self.L.oracle_inter_get(self.h, _p(out))
        return out

    def set_inter(self, arr):
        arr = np.ascontiguousarray(arr, dtype=ATOM_DTYPE)
        self.L.oracle_inter_set(self.h, _p(arr) if len(arr) else None, len(arr))

    def inter_ghosts(self):
        n = self.L.oracle_inter_ghost_count(self.h)
        out = np.zeros(n, dtype=ATOM_DTYPE)
        if n:
            self.L.oracle_inter_ghost_get(self.h, _p(out))
        return out

    def adaptive_dt(self, dt_max, max_disp):
        """Set dt = min(dt_max, max_disp / max|v|) for the next step; returns it."""
        return self.L.oracle_adaptive_dt(self.h, dt_max, max_disp)

    def mvv(self):
        return self.L.oracle_mvv(self.h)

    def temperature(self):
        return self.L.oracle_temperature(self.h)

    def rescale(self, t):
        self.L.oracle_rescale(self.h, t)

    def energy(self):
        out = np.zeros(4)
        self.L.oracle_energy(self.h, _p(out))
        return {"e_embed": out[0], "e_pair": out[1], "e_potential": out[2], "virial": out[3]}

    def system_force(self):
        out = np.zeros(3)
        self.L.oracle_system_force(self.h, _p(out))
        return out

    def vcm(self):
        out = np.zeros(4)
        self.L.oracle_vcm(self.h, _p(out))
        return out

    def timers(self):
        out = np.zeros(4)
        self.L.oracle_timers(self.h, _p(out))
        return dict(zip(("motion", "decide", "comm", "compute"), out))

    def potential(self):
        p = _BorrowedPotential.__new__(_BorrowedPotential)
        p.p = self.L.oracle_sim_pot(self.h)
        p.L = self.L
        p.owner = self
        return p


class _BorrowedPotential(Potential):
    """The potential owned by a Sim (not freed by this object)."""

    def __del__(self):
        pass
