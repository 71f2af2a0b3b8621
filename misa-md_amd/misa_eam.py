"""ctypes binding of libmisa_eam_hip.so (include/misa_eam.h).

The library is the product: HIP kernels for gfx950 behind a C-ABI.  This
module only marshals arguments; it raises if the shared library is missing
and never computes anything itself.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmisa_eam_hip.so")

# The reference's 104-byte AtomElement (src/atom/atom_element.h:18-41).
ATOM_DTYPE = np.dtype([
    ("id", "<u8"), ("type", "<i4"), ("_pad", "<i4"),
    ("x", "<f8", (3,)), ("v", "<f8", (3,)), ("f", "<f8", (3,)),
    ("rho", "<f8"), ("df", "<f8"),
])
assert ATOM_DTYPE.itemsize == 104

PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_KAHAN = 17
PREC_MIXED_F32SUM = 33
LDS_AUTO = 0
LDS_ON = 1
LDS_OFF = 2

A_FE = 2.855
RC_FACTOR = 5.6 / 2.855


class MisaGeom(ctypes.Structure):
    _fields_ = [("phase_space", ctypes.c_int64 * 3), ("lattice_const", ctypes.c_double),
                ("cutoff_radius_factor", ctypes.c_double)]


class MisaThermo(ctypes.Structure):
    _fields_ = [("mvv", ctypes.c_double), ("kinetic", ctypes.c_double), ("temperature", ctypes.c_double),
                ("e_embed", ctypes.c_double), ("e_pair", ctypes.c_double), ("e_potential", ctypes.c_double),
                ("virial", ctypes.c_double), ("force_sum", ctypes.c_double * 3), ("max_speed", ctypes.c_double),
                ("n_lattice", ctypes.c_int64), ("n_inter", ctypes.c_int64), ("n_vacancy", ctypes.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "force_sum"}
        d["force_sum"] = list(self.force_sum)
        return d


MAX_STAGES = 64
PATH_MAX = 512
DUMP_DEBUG = 0
DUMP_COPY = 1


class MisaDomain(ctypes.Structure):
    """comm::BccDomain of one rank (regions as x_low, y_low, z_low, x_high, y_high, z_high)."""
    _fields_ = [("phase_space", ctypes.c_int64 * 3), ("lattice_const", ctypes.c_double),
                ("cutoff_radius_factor", ctypes.c_double), ("cut_lattice", ctypes.c_int64),
                ("grid_size", ctypes.c_int32 * 3), ("grid_coord", ctypes.c_int32 * 3),
                ("rank_id_neighbours", ctypes.c_int32 * 6), ("pad_", ctypes.c_int32),
                ("meas_global_length", ctypes.c_double * 3), ("meas_global_region", ctypes.c_double * 6),
                ("meas_sub_box_region", ctypes.c_double * 6), ("sub_box_lattice_size", ctypes.c_int64 * 3),
                ("lattice_size_ghost", ctypes.c_int64 * 3), ("ghost_extended_lattice_size", ctypes.c_int64 * 3),
                ("sub_box_lattice_region", ctypes.c_int64 * 6), ("ghost_ext_lattice_region", ctypes.c_int64 * 6),
                ("dbx_sub_box_lattice_size", ctypes.c_int64 * 3), ("dbx_lattice_size_ghost", ctypes.c_int64 * 3),
                ("dbx_ghost_extended_lattice_size", ctypes.c_int64 * 3),
                ("dbx_sub_box_lattice_region", ctypes.c_int64 * 6),
                ("dbx_ghost_ext_lattice_region", ctypes.c_int64 * 6)]


class MisaMdReport(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("phy_time", ctypes.c_double), ("stages_run", ctypes.c_int64),
                ("collisions", ctypes.c_int64), ("rescales", ctypes.c_int64), ("dump_frames", ctypes.c_int64),
                ("dump_records", ctypes.c_int64), ("thermo_lines", ctypes.c_int64), ("thermo", MisaThermo)]


class MisaStage(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("step_length", ctypes.c_double), ("collision_set", ctypes.c_int32),
                ("collision_lat", ctypes.c_int32 * 4), ("collision_step", ctypes.c_uint64),
                ("pka_energy", ctypes.c_double), ("direction", ctypes.c_double * 3),
                ("rescale_set", ctypes.c_int32), ("pad_", ctypes.c_int32), ("rescale_t", ctypes.c_double),
                ("rescale_every", ctypes.c_uint64)]


class MisaConfig(ctypes.Structure):
    """ConfigValues (frontend/config_values.h) as the C-ABI holds it."""
    _fields_ = [("phase_space", ctypes.c_int64 * 3), ("cutoff_radius_factor", ctypes.c_double),
                ("lattice_const", ctypes.c_double), ("time_steps", ctypes.c_uint64),
                ("time_step_length", ctypes.c_double), ("create_phase", ctypes.c_int32),
                ("create_seed", ctypes.c_int32), ("create_t_set", ctypes.c_double),
                ("read_phase_filename", ctypes.c_char * PATH_MAX), ("alloy_create_seed", ctypes.c_int32),
                ("alloy_ratio", ctypes.c_int32 * 3), ("potential_file_type", ctypes.c_char * 32),
                ("potential_filename", ctypes.c_char * PATH_MAX), ("atoms_dump_mode", ctypes.c_int32),
                ("out_by_frame", ctypes.c_int32), ("atoms_dump_interval", ctypes.c_uint64),
                ("atoms_dump_file_path", ctypes.c_char * PATH_MAX), ("origin_dump_path", ctypes.c_char * PATH_MAX),
                ("thermo_interval", ctypes.c_uint64), ("logs_mode", ctypes.c_int32), ("n_stages", ctypes.c_int32),
                ("logs_filename", ctypes.c_char * PATH_MAX), ("stages", MisaStage * MAX_STAGES)]

    def stage_list(self):
        return [self.stages[i] for i in range(self.n_stages)]


_lib = None

_SIGS = {
    "misa_create": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "misa_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "misa_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "misa_set_offsets": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p, ctypes.c_int64] * 4),
    "misa_make_offsets": (ctypes.c_int, [ctypes.c_void_p]),
    "misa_get_offsets": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "misa_set_potential_file": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "misa_get_potential_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "misa_lat_rho": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    "misa_lat_df": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    "misa_lat_force": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]),
    "misa_lat_hook_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_double]),
    "misa_upload_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "misa_download_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                           ctypes.POINTER(ctypes.c_int64)]),
    "misa_build_world": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32]),
    "misa_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "misa_set_timestep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "misa_set_inter_capacity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "misa_set_range_check": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "misa_device_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "misa_set_adaptive_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]),
    "misa_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "misa_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "misa_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "misa_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "misa_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_double)]),
    "misa_set_launch_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "misa_set_lds_tables": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "misa_force_pass_used": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "misa_set_xcd_balance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "misa_xcd_slabs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "misa_set_xcd_slabs": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "misa_defect_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "misa_get_launch_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "misa_sim_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double)]),
    "misa_compute_forces": (ctypes.c_int, [ctypes.c_void_p]),
    "misa_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "misa_setv": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
                                 ctypes.c_double]),
    "misa_collision_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_double]),
    "misa_rescale": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "misa_thermo_compute": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MisaThermo)]),
    "misa_last_timings": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "misa_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "misa_last_launch_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "misa_host_offsets": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_int64)]),
    "misa_host_potential_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "misa_host_build_world": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_uint32, ctypes.c_double,
                                             ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32, ctypes.c_void_p]),
    "misa_dump_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    "misa_dump_frame": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MisaGeom), ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int64, ctypes.c_uint64]),
    "misa_dump_frame_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "misa_dump_close": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "misa_dump_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(MisaGeom), ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int64]),
    "misa_dump_text_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "misa_config_parse": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "misa_config_parse_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int64]),
    "misa_md_simulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_void_p]),
    "misa_set_neighbour_skin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "misa_host_domain": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_void_p]),
    "misa_host_fw_region": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int64)]),
    "misa_host_exchange_atoms": (ctypes.c_int, [ctypes.POINTER(MisaGeom), ctypes.c_void_p]),
}


def exported_symbols():
    return list(_SIGS)


def load():
    """Load the shared library (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libmisa_eam_hip.so not built: run __graft_entry__.build() (expected at %s)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


E_ARG, E_HIP, E_IO, E_DEVICE, E_CAPACITY = -1, -2, -3, -4, -5
FLAG_INTER_CAPACITY, FLAG_GHOST_CAPACITY, FLAG_NON_FINITE, FLAG_LEFT_BOX, FLAG_RHO_RANGE = 1, 2, 4, 8, 16


class MisaError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class EamDevice:
    """One misa_handle: a periodic single-rank box on one GPU."""

    def __init__(self, cells, lattice_const=A_FE, rc_factor=RC_FACTOR, device=0):
        self.lib = load()
        g = MisaGeom()
        for k in range(3):
            g.phase_space[k] = int(cells[k])
        g.lattice_const = lattice_const
        g.cutoff_radius_factor = rc_factor
        h = ctypes.c_void_p()
        rc = self.lib.misa_create(ctypes.byref(g), device, ctypes.byref(h))
        if rc != 0:
            raise MisaError("misa_create failed (%d)" % rc)
        self.h = h
        self.cells = tuple(int(c) for c in cells)
        self.a = lattice_const
        self.f = rc_factor
        self.ghost = int(np.ceil(rc_factor)) + 1
        self.dims = (2 * (self.cells[0] + 2 * self.ghost), self.cells[1] + 2 * self.ghost,
                     self.cells[2] + 2 * self.ghost)

    def _check(self, rc, what):
        if rc != 0:
            raise MisaError("%s failed (%d): %s" % (what, rc, self.lib.misa_last_error(self.h).decode()), rc)

    def close(self):
        if self.h is not None:
            self.lib.misa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nslots(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def empty_atoms(self):
        return np.zeros(self.nslots, dtype=ATOM_DTYPE)

    def set_offsets(self, even, odd, half_even, half_odd):
        arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (even, odd, half_even, half_odd)]
        args = []
        for a in arrs:
            args += [_ptr(a), len(a)]
        self._check(self.lib.misa_set_offsets(self.h, *args), "misa_set_offsets")

    def make_offsets(self):
        self._check(self.lib.misa_make_offsets(self.h), "misa_make_offsets")

    def get_offsets(self, which):
        n = ctypes.c_int64()
        self._check(self.lib.misa_get_offsets(self.h, which, None, 0, ctypes.byref(n)), "misa_get_offsets")
        out = np.zeros(n.value, dtype=np.int64)
        self._check(self.lib.misa_get_offsets(self.h, which, _ptr(out), n.value, ctypes.byref(n)), "misa_get_offsets")
        return out

    def set_potential(self, path):
        self._check(self.lib.misa_set_potential_file(self.h, str(path).encode()), "misa_set_potential_file")

    def potential_table(self, kind, idx):
        n = ctypes.c_int64()
        self._check(self.lib.misa_get_potential_table(self.h, kind, idx, None, 0, ctypes.byref(n)), "table")
        out = np.zeros(n.value, dtype=np.float64)
        self._check(self.lib.misa_get_potential_table(self.h, kind, idx, _ptr(out), n.value, ctypes.byref(n)), "table")
        return out

    def lat_rho(self, atoms, rc):
        self._check(self.lib.misa_lat_rho(self.h, _ptr(atoms), rc), "misa_lat_rho")

    def lat_df(self, atoms, rc):
        self._check(self.lib.misa_lat_df(self.h, _ptr(atoms), rc), "misa_lat_df")

    def lat_force(self, atoms, rc):
        self._check(self.lib.misa_lat_force(self.h, _ptr(atoms), rc), "misa_lat_force")

    def upload(self, atoms, inter=None):
        n = 0 if inter is None else len(inter)
        self._check(self.lib.misa_upload_state(self.h, _ptr(atoms), _ptr(inter) if n else None, n),
                    "misa_upload_state")

    def download(self, inter_cap=1 << 16):
        atoms = self.empty_atoms()
        inter = np.zeros(inter_cap, dtype=ATOM_DTYPE)
        n = ctypes.c_int64()
        self._check(self.lib.misa_download_state(self.h, _ptr(atoms), _ptr(inter), inter_cap, ctypes.byref(n)),
                    "misa_download_state")
        return atoms, inter[:n.value].copy()

    def build_world(self, seed, tset, ratio, alloy_seed=1):
        r = (ctypes.c_int32 * 3)(*ratio)
        self._check(self.lib.misa_build_world(self.h, seed, tset, r, alloy_seed), "misa_build_world")

    def set_inter_capacity(self, cap):
        self._check(self.lib.misa_set_inter_capacity(self.h, int(cap)), "misa_set_inter_capacity")

    def set_range_check(self, on):
        self._check(self.lib.misa_set_range_check(self.h, int(bool(on))), "misa_set_range_check")

    def device_flags(self):
        f = ctypes.c_int64()
        self._check(self.lib.misa_device_flags(self.h, ctypes.byref(f)), "misa_device_flags")
        return f.value

    def set_adaptive_dt(self, dt_max, max_disp):
        self._check(self.lib.misa_set_adaptive_dt(self.h, dt_max, max_disp), "misa_set_adaptive_dt")

    def set_neighbour_skin(self, skin):
        """Skin (A) of the mixed density pass's offset masks; 0 turns it off."""
        self._check(self.lib.misa_set_neighbour_skin(self.h, float(skin)), "misa_set_neighbour_skin")

    def sim_time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.misa_sim_time(self.h, ctypes.byref(t), ctypes.byref(dt)), "misa_sim_time")
        return t.value, dt.value

    def set_launch_shape(self, ty, gathers_rho, gathers_force):
        self._check(self.lib.misa_set_launch_shape(self.h, int(ty), int(gathers_rho), int(gathers_force)),
                    "misa_set_launch_shape")

    def set_lds_tables(self, mode):
        """LDS_AUTO, LDS_ON or LDS_OFF: the mixed force pass with its majority-species tables in LDS."""
        self._check(self.lib.misa_set_lds_tables(self.h, int(mode)), "misa_set_lds_tables")

    def defect_counters(self):
        """{staged, refilled, refilled_far} since the last upload."""
        out = (ctypes.c_int64 * 3)()
        self._check(self.lib.misa_defect_counters(self.h, out), "misa_defect_counters")
        return {"staged": out[0], "refilled": out[1], "refilled_far": out[2]}

    def force_pass(self):
        """Name of the force kernel the last force evaluation launched."""
        k = ctypes.c_int32()
        self._check(self.lib.misa_force_pass_used(self.h, ctypes.byref(k)), "misa_force_pass_used")
        return {0: "kForce", 1: "kForceM", 2: "kForceL", 3: "kForceDL"}.get(k.value, "none")

    def set_xcd_balance(self, on):
        """Cut the XCD z-slabs by the force pass's measured work (default on)."""
        self._check(self.lib.misa_set_xcd_balance(self.h, 1 if on else 0), "misa_set_xcd_balance")

    def xcd_slabs(self):
        """Slab bounds: XCD c walks z-planes [zlo[c], zlo[c+1])."""
        z = (ctypes.c_int32 * 9)()
        self._check(self.lib.misa_xcd_slabs(self.h, z), "misa_xcd_slabs")
        return list(z)

    def set_xcd_slabs(self, zlo):
        """Set the slab bounds (9 values from 0 to the box's planes) and turn balancing off."""
        z = (ctypes.c_int32 * 9)(*zlo)
        self._check(self.lib.misa_set_xcd_slabs(self.h, z), "misa_set_xcd_slabs")

    def launch_shape(self):
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.misa_get_launch_shape(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "misa_get_launch_shape")
        return a.value, b.value, c.value

    def autotune(self, steps):
        ty, ms = ctypes.c_int32(), ctypes.c_double()
        self._check(self.lib.misa_autotune(self.h, int(steps), ctypes.byref(ty), ctypes.byref(ms)), "misa_autotune")
        return ty.value, ms.value

    def set_precision(self, mode):
        self._check(self.lib.misa_set_precision(self.h, mode), "misa_set_precision")

    def set_timestep(self, dt):
        self._check(self.lib.misa_set_timestep(self.h, dt), "misa_set_timestep")

    def compute_forces(self):
        self._check(self.lib.misa_compute_forces(self.h), "misa_compute_forces")

    def step(self, n):
        self._check(self.lib.misa_step(self.h, int(n)), "misa_step")

    def setv(self, lat, direction, energy):
        l4 = (ctypes.c_int32 * 4)(*lat)
        d3 = (ctypes.c_double * 3)(*direction)
        self._check(self.lib.misa_setv(self.h, l4, d3, energy), "misa_setv")

    def collision_step(self, lat, direction, energy):
        l4 = (ctypes.c_int32 * 4)(*lat)
        d3 = (ctypes.c_double * 3)(*direction)
        self._check(self.lib.misa_collision_step(self.h, l4, d3, energy), "misa_collision_step")

    def rescale(self, t):
        self._check(self.lib.misa_rescale(self.h, t), "misa_rescale")

    def thermo(self):
        t = MisaThermo()
        self._check(self.lib.misa_thermo_compute(self.h, ctypes.byref(t)), "misa_thermo_compute")
        return t.as_dict()

    def set_profiling(self, on):
        self._check(self.lib.misa_set_profiling(self.h, 1 if on else 0), "misa_set_profiling")

    def last_timings(self):
        out = (ctypes.c_double * 4)()
        self._check(self.lib.misa_last_timings(self.h, out), "misa_last_timings")
        cnt = (ctypes.c_int64 * 3)()
        self._check(self.lib.misa_last_launch_counts(self.h, cnt), "misa_last_launch_counts")
        return {"rho_df_ms": out[0], "force_ms": out[1], "other_ms": out[2], "total_ms": out[3],
                "launches": list(cnt)}


class EamBatch:
    """Independent boxes stepped together through captured step graphs."""

    def __init__(self, devices):
        self.lib = load()
        self.devices = list(devices)  # keeps the handles alive
        arr = (ctypes.c_void_p * len(self.devices))(*[d.h.value for d in self.devices])
        b = ctypes.c_void_p()
        rc = self.lib.misa_batch_create(arr, len(self.devices), ctypes.byref(b))
        if rc != 0:
            raise MisaError("misa_batch_create failed (%d)" % rc)
        self.b = b

    def step(self, n):
        r = ctypes.c_int64()
        rc = self.lib.misa_batch_step(self.b, int(n), ctypes.byref(r))
        if rc != 0:
            raise MisaError("misa_batch_step failed (%d): %s" % (rc, self.lib.misa_batch_last_error(self.b).decode()))
        return r.value

    def close(self):
        if self.b is not None:
            self.lib.misa_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def geom(cells, lattice_const=A_FE, rc_factor=RC_FACTOR):
    g = MisaGeom()
    for k in range(3):
        g.phase_space[k] = int(cells[k])
    g.lattice_const = lattice_const
    g.cutoff_radius_factor = rc_factor
    return g


def host_offsets(cells, which, lattice_const=A_FE, rc_factor=RC_FACTOR):
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    n = ctypes.c_int64()
    if lib.misa_host_offsets(ctypes.byref(g), which, None, 0, ctypes.byref(n)) != 0:
        raise MisaError("misa_host_offsets")
    out = np.zeros(n.value, dtype=np.int64)
    lib.misa_host_offsets(ctypes.byref(g), which, _ptr(out), n.value, ctypes.byref(n))
    return out


def host_potential_table(path, kind, idx):
    lib = load()
    n = ctypes.c_int64()
    if lib.misa_host_potential_table(str(path).encode(), kind, idx, None, 0, ctypes.byref(n)) != 0:
        raise MisaError("misa_host_potential_table")
    out = np.zeros(n.value, dtype=np.float64)
    lib.misa_host_potential_table(str(path).encode(), kind, idx, _ptr(out), n.value, ctypes.byref(n))
    return out


def host_build_world(cells, seed, tset, ratio, alloy_seed=1, lattice_const=A_FE, rc_factor=RC_FACTOR):
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    gh = int(np.ceil(rc_factor)) + 1
    n = 2 * (cells[0] + 2 * gh) * (cells[1] + 2 * gh) * (cells[2] + 2 * gh)
    atoms = np.zeros(n, dtype=ATOM_DTYPE)
    r = (ctypes.c_int32 * 3)(*ratio)
    if lib.misa_host_build_world(ctypes.byref(g), seed, tset, r, alloy_seed, _ptr(atoms)) != 0:
        raise MisaError("misa_host_build_world")
    return atoms


def sub_box_view(atoms, dev):
    """Sub-box slots of a ghost-extended array in reference (z, y, X) order."""
    g = dev.ghost
    a = atoms.reshape(dev.dims[2], dev.dims[1], dev.dims[0])
    return a[g:g + dev.cells[2], g:g + dev.cells[1], 2 * g:2 * g + 2 * dev.cells[0]]


# ---- atom dumps of the reference frontend (frontend/io) ----

# AtomInfoDump (frontend/io/atom_info_dump.h:14-22) as written to a dump file
DUMP_DTYPE = np.dtype([
    ("id", "<u8"), ("step", "<u8"), ("type", "<i4"), ("inter_type", "<i2"), ("_pad", "<i2"),
    ("x", "<f8", (3,)), ("v", "<f8", (3,)),
])
assert DUMP_DTYPE.itemsize == 72
DUMP_HEADER = 128        # file header (the storage library's; zeros here)
DUMP_RANK_HEADER = 128   # per rank: record count in the first 8 bytes
DUMP_BLOCK = 1024        # records per block; blocks of the ranks interleave


def _inter_arg(inter):
    if inter is None or len(inter) == 0:
        return None, 0
    inter = np.ascontiguousarray(inter, dtype=ATOM_DTYPE)
    return inter, len(inter)


class AtomDumpFile:
    """Copy-mode atom dump of one rank (AtomDump + BufferedFileWriter): frames
    of the inter atoms in list order, then the occupied sub-box sites in z, y,
    doubled-x order; close() writes the record count and returns it."""

    def __init__(self, path):
        self.lib = load()
        d = ctypes.c_void_p()
        if self.lib.misa_dump_open(str(path).encode(), ctypes.byref(d)) != 0:
            raise MisaError("misa_dump_open failed for %s" % path)
        self.d = d

    def frame(self, atoms, inter, cells, step, lattice_const=A_FE, rc_factor=RC_FACTOR):
        g = geom(cells, lattice_const, rc_factor)
        atoms = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
        inter, n = _inter_arg(inter)
        rc = self.lib.misa_dump_frame(self.d, ctypes.byref(g), _ptr(atoms), _ptr(inter) if n else None, n,
                                      int(step))
        if rc != 0:
            raise MisaError("misa_dump_frame (%d)" % rc)

    def frame_state(self, dev, step):
        """A frame of a device handle's resident state."""
        rc = self.lib.misa_dump_frame_state(self.d, dev.h, int(step))
        if rc != 0:
            raise MisaError("misa_dump_frame_state (%d)" % rc)

    def close(self):
        n = ctypes.c_int64()
        rc = self.lib.misa_dump_close(self.d, ctypes.byref(n))
        self.d = None
        if rc != 0:
            raise MisaError("misa_dump_close (%d)" % rc)
        return n.value


def dump_text(path, atoms, inter, cells, lattice_const=A_FE, rc_factor=RC_FACTOR):
    """Debug-mode text dump (OutputDump::dump)."""
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    atoms = np.ascontiguousarray(atoms, dtype=ATOM_DTYPE)
    inter, n = _inter_arg(inter)
    if lib.misa_dump_text(str(path).encode(), ctypes.byref(g), _ptr(atoms), _ptr(inter) if n else None, n) != 0:
        raise MisaError("misa_dump_text")


def dump_text_state(dev, path):
    if dev.lib.misa_dump_text_state(dev.h, str(path).encode()) != 0:
        raise MisaError("misa_dump_text_state")


def read_dump(path, n_ranks=1):
    """Records of a copy-mode dump, rank by rank, located the way
    tools/dump2text/out_convert.cpp locates them: rank r's count at
    DUMP_HEADER + DUMP_RANK_HEADER * r, its k-th block of DUMP_BLOCK records
    at DUMP_HEADER + DUMP_RANK_HEADER * n_ranks + (n_ranks * k + r) blocks."""
    raw = np.fromfile(str(path), dtype=np.uint8)
    out = []
    base = DUMP_HEADER + DUMP_RANK_HEADER * n_ranks
    bb = DUMP_BLOCK * DUMP_DTYPE.itemsize
    for r in range(n_ranks):
        h = DUMP_HEADER + DUMP_RANK_HEADER * r
        count = int(raw[h:h + 8].view("<u8")[0])
        parts, left, k = [], count, 0
        while left > 0:
            take = min(left, DUMP_BLOCK)
            o = base + (n_ranks * k + r) * bb
            parts.append(raw[o:o + take * DUMP_DTYPE.itemsize].view(DUMP_DTYPE))
            left -= take
            k += 1
        out.append(np.concatenate(parts) if parts else np.zeros(0, DUMP_DTYPE))
    return out[0] if n_ranks == 1 else out


# ---- frontend: YAML configuration and the stage machine ----

def parse_config(path=None, text=None):
    """ConfigValues from a configuration file (or text); raises MisaError
    with the parser's message."""
    lib = load()
    cfg = MisaConfig()
    err = ctypes.create_string_buffer(512)
    if text is not None:
        rc = lib.misa_config_parse_text(text.encode(), ctypes.byref(cfg), err, 512)
    else:
        rc = lib.misa_config_parse(str(path).encode(), ctypes.byref(cfg), err, 512)
    if rc != 0:
        raise MisaError("configuration: %s" % err.value.decode())
    return cfg


def md_simulate(cfg, device=0, precision=PREC_FP64):
    """Run every stage of `cfg` on the device (MDSimulation); returns the
    report (steps, physical time, events, final thermo)."""
    lib = load()
    rep = MisaMdReport()
    rc = lib.misa_md_simulate(ctypes.byref(cfg), device, precision, ctypes.byref(rep))
    if rc != 0:
        raise MisaError("misa_md_simulate (%d)" % rc)
    return rep


# ---- host mirror of libcomm's one-rank domain ----

def host_domain(cells, lattice_const=A_FE, rc_factor=RC_FACTOR):
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    d = MisaDomain()
    if lib.misa_host_domain(ctypes.byref(g), ctypes.byref(d)) != 0:
        raise MisaError("misa_host_domain")
    return d


def host_fw_region(cells, dimension, direction, lattice_const=A_FE, rc_factor=RC_FACTOR):
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    r = (ctypes.c_int64 * 6)()
    if lib.misa_host_fw_region(ctypes.byref(g), dimension, direction, r) != 0:
        raise MisaError("misa_host_fw_region")
    return list(r)


def host_exchange_atoms(atoms, cells, lattice_const=A_FE, rc_factor=RC_FACTOR):
    """exchangeAtomFirst at one rank, in place on a ghost-extended array."""
    lib = load()
    g = geom(cells, lattice_const, rc_factor)
    if lib.misa_host_exchange_atoms(ctypes.byref(g), _ptr(atoms)) != 0:
        raise MisaError("misa_host_exchange_atoms")
