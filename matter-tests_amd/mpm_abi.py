"""ctypes mirror of include/mpm.h, used by bench.py and the tests to drive the C ABI.

Two libraries expose the same calls: the product, libmpm_hip.so (prefix ``mpm_``), and — for tests and
the CPU-baseline leg of bench.py only — the CPU oracle (prefix ``oracle_``). ``Engine`` wraps one context
of either with numpy in and out.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
PRODUCT_LIB = os.path.join(HERE, "libmpm_hip.so")
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")

# enums (include/mpm.h)
PLATE = dict(top=0, bottom=1, left=2, right=3, front=4, back=5)
ELASTIC = dict(Hencky=0, NeoHookean=1)
PLASTIC = dict(NoPlasticity=0, VM=1, DP=2, DPSoft=3, MCC=4, VMVisc=5, DPVisc=6, MCCVisc=7, DPMui=8, MCCMui=9, Snow=10)
HARDENING = dict(NoHard=0, ExpoExpl=1, ExpoImpl=2, SinhExpl=3, SinhImpl=4)
BC = dict(NoSlip=0, SlipStick=1, SlipFree=2)
OBJECT = dict(BoxRotated=1, Bump=2, Ground=3, Ramp=4, Gate=5, Silo=6, Curve=7, Sdf=8)
PRECISION = dict(fp64=0, split=1)

ERR = dict(INVALID_ARGUMENT=1 << 0, UNSUPPORTED=1 << 1, HIP=1 << 2, OUT_OF_MEMORY=1 << 3,
           PERZYNA_NEGATIVE=1 << 4, NEWTON_MAXITER=1 << 5, NEWTON_SMALL_DERIV=1 << 6, MCC_NO_STEP=1 << 7,
           GRID_BOUNDS=1 << 8, NONFINITE=1 << 9, CONSERVATION=1 << 10, NOT_READY=1 << 11, SCRATCH_LIMIT=1 << 12,
           WARN_SUPERSONIC=1 << 16)
STEP_STOP_AT_FRAME_END = 1
STEP_CHECK_CONSERVATION = 2
EXIT_MASK = 0xFFFF


class Config(C.Structure):
    _fields_ = [("dim", C.c_int32), ("spline_degree", C.c_int32), ("precision", C.c_int32), ("n_threads", C.c_int32),
                ("elastic_model", C.c_int32), ("plastic_model", C.c_int32), ("hardening_law", C.c_int32),
                ("use_pradhana", C.c_int32), ("use_mises_q", C.c_int32), ("use_mibf", C.c_int32), ("use_musl", C.c_int32),
                ("gravity_special", C.c_int32), ("pbc", C.c_int32), ("change_particle_positions", C.c_int32),
                ("end_frame", C.c_uint32), ("reserved0", C.c_uint32),
                ("grid_reference_point", C.c_double * 3), ("gravity", C.c_double * 3),
                ("fps", C.c_double), ("cfl", C.c_double), ("cfl_elastic", C.c_double), ("flip_ratio", C.c_double),
                ("rho", C.c_double), ("gravity_time", C.c_double),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("particle_mass", C.c_double), ("particle_volume", C.c_double), ("dx", C.c_double),
                ("E", C.c_double), ("nu", C.c_double), ("stress_tolerance", C.c_double), ("xi", C.c_double),
                ("q_max", C.c_double), ("q_min", C.c_double), ("p_min", C.c_double),
                ("M", C.c_double), ("q_cohesion", C.c_double),
                ("perzyna_exp", C.c_double), ("perzyna_visc", C.c_double),
                ("beta", C.c_double), ("p0", C.c_double),
                ("rho_s", C.c_double), ("grain_diameter", C.c_double), ("I_ref", C.c_double),
                ("mu_1", C.c_double), ("mu_2", C.c_double)]


class Plate(C.Structure):
    _fields_ = [("pos_object", C.c_double), ("plate_type", C.c_int32), ("bc", C.c_int32), ("friction", C.c_double),
                ("pos_lower", C.c_double), ("pos_upper", C.c_double), ("v_object", C.c_double * 3),
                ("vmin_factor", C.c_double), ("load_factor", C.c_double)]


class Object(C.Structure):
    _fields_ = [("kind", C.c_int32), ("bc", C.c_int32), ("friction", C.c_double), ("param", C.c_double),
                ("center", C.c_double * 3), ("half_extents", C.c_double * 3), ("rotation", C.c_double * 9)]


class Sdf(C.Structure):
    _fields_ = [("bc", C.c_int32), ("reserved", C.c_int32), ("friction", C.c_double), ("origin", C.c_double * 3),
                ("spacing", C.c_double), ("n", C.c_uint32 * 3), ("reserved1", C.c_uint32), ("phi", C.POINTER(C.c_float))]


class ParticleArrays(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in
                ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "eps_pl_vol_pradhana", "delta_gamma", "viscosity", "muI")]


class StepInfo(C.Structure):
    _fields_ = [("dt", C.c_double), ("time", C.c_double), ("step", C.c_uint64), ("frame", C.c_uint32),
                ("plastic_count", C.c_uint32), ("steps_done", C.c_uint32), ("status", C.c_uint32),
                ("frame_ended", C.c_int32), ("finished", C.c_int32)]


class KernelTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("launches", C.c_uint64)]


class BatchOptions(C.Structure):
    _fields_ = [("chunk_steps", C.c_uint32), ("use_graph", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("n_scenes", C.c_uint32), ("n_active", C.c_uint32), ("n_classes", C.c_uint32), ("chunks", C.c_uint32),
                ("status", C.c_uint32), ("reserved", C.c_uint32), ("particle_steps", C.c_uint64)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_scenes", C.c_uint32), ("chunk_steps", C.c_uint32),
                ("use_graph", C.c_int32), ("keep_results", C.c_int32)]


class RequestDesc(C.Structure):
    _fields_ = [("cfg", C.POINTER(Config)), ("plates", C.POINTER(Plate)), ("n_plates", C.c_uint32),
                ("n_objects", C.c_uint32), ("objects", C.POINTER(Object)), ("particles", C.POINTER(ParticleArrays)),
                ("n", C.c_uint32), ("nsteps", C.c_uint32), ("arrival_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint32), ("completed", C.c_uint32), ("failed", C.c_uint32),
                ("peak_resident", C.c_uint32), ("peak_resident_bytes", C.c_uint64), ("particle_steps", C.c_uint64),
                ("wall_s", C.c_double), ("throughput", C.c_double), ("step_ms_p50", C.c_double),
                ("step_ms_p99", C.c_double), ("request_ms_p50", C.c_double), ("request_ms_p99", C.c_double),
                ("queue_ms_p50", C.c_double), ("queue_ms_p99", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GridInfo(C.Structure):
    _fields_ = [("Nx", C.c_uint32), ("Ny", C.c_uint32), ("Nz", C.c_uint32), ("grid_nodes", C.c_uint32),
                ("xc", C.c_double), ("yc", C.c_double), ("zc", C.c_double)]


SCALARS = ("eps_pl_dev", "eps_pl_vol", "eps_pl_vol_pradhana", "delta_gamma", "viscosity", "muI")


def default_config(**kw):
    """Reference defaults (src/simulation/simulation.hpp:28-106) plus dim=3, quadratic, fp64."""
    c = Config()
    c.dim, c.spline_degree, c.precision, c.n_threads = 3, 2, 0, 1
    c.elastic_model, c.plastic_model, c.hardening_law = 0, 0, HARDENING["ExpoImpl"]
    c.use_pradhana, c.use_mises_q, c.use_mibf, c.use_musl = 1, 0, 0, 0
    c.gravity_special, c.pbc, c.change_particle_positions, c.end_frame = 0, 0, 0, 1
    for i in range(3):
        c.grid_reference_point[i] = 2e10
        c.gravity[i] = 0.0
    c.fps, c.cfl, c.cfl_elastic, c.flip_ratio, c.rho, c.gravity_time = 1.0, 0.5, 0.5, -0.95, 1000.0, 0.0
    c.Lx = c.Ly = c.Lz = 1.0
    c.particle_mass = c.particle_volume = c.dx = 0.0
    c.E, c.nu, c.stress_tolerance, c.xi = 1e6, 0.3, 1e-5, 0.0
    c.q_max, c.q_min, c.p_min = 100.0, 100.0, -1e20
    c.M, c.q_cohesion = 1.0, 0.0
    c.perzyna_exp, c.perzyna_visc = 1.0, 0.0
    c.beta, c.p0 = 0.0, 1000.0
    c.rho_s, c.grain_diameter, c.I_ref = 2500.0, 1e-3, 0.279
    c.mu_1, c.mu_2 = float(np.tan(20.9 * np.pi / 180.0)), float(np.tan(32.8 * np.pi / 180.0))
    for k, v in kw.items():
        set_field(c, k, v)
    return c


def set_field(c, k, v):
    if k in ("gravity", "grid_reference_point"):
        arr = getattr(c, k)
        for i in range(3):
            arr[i] = float(v[i]) if i < len(v) else 0.0
    elif k == "plastic_model" and isinstance(v, str):
        c.plastic_model = PLASTIC[v]
    elif k == "elastic_model" and isinstance(v, str):
        c.elastic_model = ELASTIC[v]
    elif k == "hardening_law" and isinstance(v, str):
        c.hardening_law = HARDENING[v]
    elif k == "precision" and isinstance(v, str):
        c.precision = PRECISION[v]
    else:
        setattr(c, k, v)


def make_plate(pos, plate_type, bc="NoSlip", friction=0.0, pos_lower=-1e15, pos_upper=1e15, v_object=(0, 0, 0),
               vmin_factor=1.0, load_factor=0.0):
    p = Plate()
    p.pos_object = pos
    p.plate_type = PLATE[plate_type] if isinstance(plate_type, str) else plate_type
    p.bc = BC[bc] if isinstance(bc, str) else bc
    p.friction, p.pos_lower, p.pos_upper = friction, pos_lower, pos_upper
    for i in range(3):
        p.v_object[i] = v_object[i] if i < len(v_object) else 0.0
    p.vmin_factor, p.load_factor = vmin_factor, load_factor
    return p


def make_object(kind, bc="NoSlip", friction=0.0, param=0.0, center=(0, 0, 0), half_extents=(0, 0, 0), rotation=None):
    o = Object()
    o.kind = OBJECT[kind] if isinstance(kind, str) else kind
    o.bc = BC[bc] if isinstance(bc, str) else bc
    o.friction, o.param = friction, param
    for i in range(3):
        o.center[i] = center[i] if i < len(center) else 0.0
        o.half_extents[i] = half_extents[i] if i < len(half_extents) else 0.0
    rot = np.eye(3) if rotation is None else np.asarray(rotation, dtype=np.float64)
    flat = rot.flatten(order="F")
    for i in range(9):
        o.rotation[i] = flat[i] if i < flat.size else 0.0
    return o


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class Library:
    """One shared library exposing the mpm.h calls under `prefix`."""

    def __init__(self, path, prefix):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
        self.path, self.prefix = path, prefix
        self.lib = C.CDLL(path)
        L, p = self.lib, prefix
        vp = C.c_void_p
        if p == "mpm_":
            L.mpm_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
            L.mpm_create.restype = C.c_int
            L.mpm_destroy.argtypes = [vp]
            L.mpm_num_particles.argtypes = [vp]
            L.mpm_num_particles.restype = C.c_uint32
            L.mpm_status_string.argtypes = [C.c_uint32]
            L.mpm_status_string.restype = C.c_char_p
            L.mpm_config_defaults.argtypes = [C.POINTER(Config)]
            L.mpm_set_profiling.argtypes = [vp, C.c_int]
            L.mpm_set_profiling_kernel.argtypes = [vp, C.c_char_p]
            L.mpm_set_step_graphs.argtypes = [vp, C.c_int]
            L.mpm_get_binning.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]
            L.mpm_get_kernel_times.argtypes = [vp, C.POINTER(KernelTiming), C.c_uint32, C.POINTER(C.c_uint32)]
            L.mpm_abi_version.restype = C.c_int
            L.mpm_batch_create.argtypes = [C.POINTER(BatchOptions), C.POINTER(vp)]
            L.mpm_batch_create.restype = C.c_int
            L.mpm_batch_destroy.argtypes = [vp]
            L.mpm_batch_destroy.restype = None
            for name, args in (("mpm_batch_add", [vp, vp, C.c_uint32]), ("mpm_batch_remove", [vp, vp]),
                               ("mpm_batch_set_budget", [vp, vp, C.c_uint32]),
                               ("mpm_batch_step", [vp, C.c_uint32, C.c_uint32, C.POINTER(BatchInfo)]),
                               ("mpm_scene_info", [vp, C.POINTER(StepInfo), C.POINTER(C.c_uint32)])):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = C.c_int
            L.mpm_get_stress_measures.argtypes = [vp] + [C.POINTER(C.c_double)] * 3
            L.mpm_get_stress_measures.restype = C.c_int
            for name in ("mpm_write_particles_ply", "mpm_write_grid_ply"):
                getattr(L, name).argtypes = [vp, C.c_char_p]
                getattr(L, name).restype = C.c_int
            L.mpm_step_noalloc.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(StepInfo)]
            L.mpm_step_noalloc.restype = C.c_int
            L.mpm_batch_autotune.argtypes = [vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
            L.mpm_batch_autotune.restype = C.c_int
            L.mpm_device_bytes.argtypes = [vp]
            L.mpm_device_bytes.restype = C.c_uint64
            L.mpm_server_create.argtypes = [C.POINTER(ServerOptions), C.POINTER(vp)]
            L.mpm_server_create.restype = C.c_int
            L.mpm_server_destroy.argtypes = [vp]
            L.mpm_server_destroy.restype = None
            L.mpm_server_submit.argtypes = [vp, C.POINTER(RequestDesc), C.POINTER(C.c_uint64)]
            L.mpm_server_submit.restype = C.c_int
            L.mpm_server_run.argtypes = [vp, C.c_double, C.POINTER(ServerStats)]
            L.mpm_server_run.restype = C.c_int
            L.mpm_server_result.argtypes = [vp, C.c_uint64, C.POINTER(ParticleArrays), C.POINTER(StepInfo),
                                            C.POINTER(C.c_double)]
            L.mpm_server_result.restype = C.c_int
        else:
            L.oracle_create.argtypes = [C.POINTER(Config)]
            L.oracle_create.restype = vp
            L.oracle_destroy.argtypes = [vp]
            L.oracle_num_particles.argtypes = [vp]
            L.oracle_num_particles.restype = C.c_uint32
            L.oracle_set_threads.argtypes = [vp, C.c_int]
            L.oracle_time_steps.argtypes = [vp, C.c_uint32, C.POINTER(StepInfo)]
            L.oracle_time_steps.restype = C.c_double
            L.oracle_phase_times.argtypes = [vp, C.POINTER(C.c_double)]
            L.oracle_phase_times.restype = None
            L.oracle_avg_data.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
            L.oracle_bulk_modulus.argtypes = [vp]
            L.oracle_bulk_modulus.restype = C.c_double
            L.oracle_boundary_single.argtypes = [vp, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double]
            L.oracle_plasticity_single.argtypes = [C.POINTER(Config), C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                   C.POINTER(C.c_double), C.POINTER(C.c_uint32)]
            L.oracle_kirchhoff.argtypes = [C.POINTER(Config), C.POINTER(C.c_double), C.POINTER(C.c_double)]
            L.oracle_svd.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 4
            L.oracle_spline_N.argtypes = [C.c_int, C.c_double]
            L.oracle_spline_N.restype = C.c_double
            L.oracle_spline_dN.argtypes = [C.c_int, C.c_double]
            L.oracle_spline_dN.restype = C.c_double
            L.oracle_arange.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double), C.c_int]
            L.oracle_max_threads.restype = C.c_int
            L.oracle_object_eval.argtypes = [vp, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]
            L.oracle_object_eval.restype = C.c_int
        for name, args in (("add_plate", [vp, C.POINTER(Plate)]), ("add_object", [vp, C.POINTER(Object)]),
                           ("add_sdf_object", [vp, C.POINTER(Sdf)]),
                           ("set_particles", [vp, C.c_uint32, C.POINTER(ParticleArrays)]),
                           ("step", [vp, C.c_uint32, C.c_uint32, C.POINTER(StepInfo)]),
                           ("get_particles", [vp, C.POINTER(ParticleArrays)]),
                           ("get_grid", [vp, C.POINTER(GridInfo)] + [C.POINTER(C.c_double)] * 7),
                           ("get_plates", [vp, C.POINTER(Plate), C.c_uint32, C.POINTER(C.c_uint32)])):
            f = getattr(L, p + name)
            f.argtypes = args
            f.restype = C.c_int

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)


_LIBS = {}


def product_library():
    if "product" not in _LIBS:
        _LIBS["product"] = Library(PRODUCT_LIB, "mpm_")
    return _LIBS["product"]


def oracle_library():
    if "oracle" not in _LIBS:
        _LIBS["oracle"] = Library(ORACLE_LIB, "oracle_")
    return _LIBS["oracle"]


class MPMError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with status 0x{code:x} ({status_names(code)})")
        self.code = code


def status_names(code):
    return "|".join(k for k, v in ERR.items() if code & v) or "OK"


class Engine:
    """One simulation context of a Library (the product engine or the oracle)."""

    def __init__(self, lib, cfg, plates=(), objects=()):
        self.L, self.cfg = lib, cfg
        self.dim = cfg.dim
        if lib.prefix == "mpm_":
            h = C.c_void_p()
            rc = lib.lib.mpm_create(C.byref(cfg), C.byref(h))
            if rc:
                raise MPMError(rc, "mpm_create")
            self.h = h
        else:
            self.h = C.c_void_p(lib.lib.oracle_create(C.byref(cfg)))
            if not self.h.value:
                raise MPMError(ERR["UNSUPPORTED"], "oracle_create")
        for p in plates:
            self._check(lib.fn("add_plate")(self.h, C.byref(p)), "add_plate")
        for o in objects:
            self._check(lib.fn("add_object")(self.h, C.byref(o)), "add_object")
        self.n = 0

    def add_sdf(self, phi, origin, spacing, bc="NoSlip", friction=0.0):
        """Level-set object from a lattice of signed distances phi[i, j(, k)] (mpm_add_sdf_object)."""
        phi = np.ascontiguousarray(phi, dtype=np.float32)
        n = list(phi.shape) + [1] * (3 - phi.ndim)
        sd = Sdf()
        sd.bc = BC[bc] if isinstance(bc, str) else bc
        sd.friction, sd.spacing = friction, spacing
        for i in range(3):
            sd.origin[i] = origin[i] if i < len(origin) else 0.0
            sd.n[i] = n[i]
        sd.phi = phi.ctypes.data_as(C.POINTER(C.c_float))
        self._check(self.L.fn("add_sdf_object")(self.h, C.byref(sd)), "add_sdf_object")

    def _check(self, rc, what):
        if rc:
            raise MPMError(rc, what)

    def close(self):
        if self.h is not None and self.h.value:
            self.L.fn("destroy")(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, x, **fields):
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = x.shape[0]
        keep = {"x": x}
        pa = ParticleArrays()
        pa.x = _dptr(x)
        for k in ("v", "F", "Bmat") + SCALARS:
            if fields.get(k) is not None:
                a = np.ascontiguousarray(fields[k], dtype=np.float64)
                if k in ("F", "Bmat") and a.ndim == 3:  # (n, dim, dim) -> column-major per particle
                    a = np.ascontiguousarray(a.transpose(0, 2, 1))
                keep[k] = a
                setattr(pa, k, _dptr(a))
        self._check(self.L.fn("set_particles")(self.h, n, C.byref(pa)), "set_particles")
        self.n = n

    def step(self, nsteps=1, flags=0, raise_on_error=False):
        info = StepInfo()
        rc = self.L.fn("step")(self.h, nsteps, flags, C.byref(info))
        if rc and raise_on_error:
            raise MPMError(rc, "step")
        return rc, info

    def step_noalloc(self, nsteps=1, flags=0):
        """mpm_step_noalloc: the same steps with the grid in the device's static scratch (no allocation)."""
        info = StepInfo()
        rc = self.L.lib.mpm_step_noalloc(self.h, nsteps, flags, C.byref(info))
        return rc, info

    def get_particles(self):
        n, d = self.n, self.dim
        out = {"x": np.zeros((n, d)), "v": np.zeros((n, d)), "F": np.zeros((n, d * d)), "Bmat": np.zeros((n, d * d))}
        for k in SCALARS:
            out[k] = np.zeros(n)
        pa = ParticleArrays()
        for k, a in out.items():
            setattr(pa, k, _dptr(a))
        self._check(self.L.fn("get_particles")(self.h, C.byref(pa)), "get_particles")
        # matrices back to (n, dim, dim) row-major indexing: F[p, r, c]
        out["F"] = out["F"].reshape(n, d, d).transpose(0, 2, 1).copy()
        out["Bmat"] = out["Bmat"].reshape(n, d, d).transpose(0, 2, 1).copy()
        return out

    def get_grid(self):
        info = GridInfo()
        nul = C.POINTER(C.c_double)()
        self._check(self.L.fn("get_grid")(self.h, C.byref(info), nul, nul, nul, nul, nul, nul, nul), "get_grid")
        d, n = self.dim, info.grid_nodes
        g = {"x": np.zeros(info.Nx), "y": np.zeros(info.Ny), "z": np.zeros(info.Nz if d == 3 else 0),
             "mass": np.zeros(n), "v": np.zeros((n, d)), "flip": np.zeros((n, d)), "friction": np.zeros(n)}
        self._check(self.L.fn("get_grid")(self.h, C.byref(info), _dptr(g["x"]), _dptr(g["y"]),
                                          _dptr(g["z"]) if d == 3 else nul, _dptr(g["mass"]), _dptr(g["v"]),
                                          _dptr(g["flip"]), _dptr(g["friction"])), "get_grid")
        g["info"] = info
        return g

    def get_binning(self):
        """(cell_base[n, dim] int32, block_key[n] uint32) of the current positions in the last step's grid."""
        base = np.zeros((self.n, self.dim), dtype=np.int32)
        key = np.zeros(self.n, dtype=np.uint32)
        self._check(self.L.lib.mpm_get_binning(self.h, base.ctypes.data_as(C.POINTER(C.c_int32)),
                                               key.ctypes.data_as(C.POINTER(C.c_uint32))), "get_binning")
        return base, key

    def set_profiling(self, enable=True):
        self._check(self.L.lib.mpm_set_profiling(self.h, 1 if enable else 0), "set_profiling")

    def set_step_graphs(self, enable=True):
        """HIP-graph replay of step batches for small scenes (on by default)."""
        self._check(self.L.lib.mpm_set_step_graphs(self.h, 1 if enable else 0), "set_step_graphs")

    def set_profiling_kernel(self, name):
        """Times only the kernel `name` (two HIP events per step instead of two per launch)."""
        self._check(self.L.lib.mpm_set_profiling_kernel(self.h, name.encode()), "set_profiling_kernel")

    def kernel_times(self):
        """{kernel name: (total_ms, launches)} accumulated since set_profiling(True)."""
        arr = (KernelTiming * 16)()
        cnt = C.c_uint32()
        self._check(self.L.lib.mpm_get_kernel_times(self.h, arr, 16, C.byref(cnt)), "get_kernel_times")
        return {arr[i].name.decode(): (arr[i].total_ms, arr[i].launches) for i in range(cnt.value)}

    def get_plates(self, cap=16):
        arr = (Plate * cap)()
        cnt = C.c_uint32()
        self._check(self.L.fn("get_plates")(self.h, arr, cap, C.byref(cnt)), "get_plates")
        return [arr[i] for i in range(cnt.value)]


class Batch:
    """A batch of product-engine contexts advanced together (mpm_batch_* of include/mpm.h)."""

    def __init__(self, lib=None, chunk_steps=16, use_graph=True):
        self.L = lib or product_library()
        opt = BatchOptions(chunk_steps, 1 if use_graph else 0)
        h = C.c_void_p()
        rc = self.L.lib.mpm_batch_create(C.byref(opt), C.byref(h))
        if rc:
            raise MPMError(rc, "mpm_batch_create")
        self.h = h
        self.members = []

    def close(self):
        if self.h is not None and self.h.value:
            self.L.lib.mpm_batch_destroy(self.h)
        self.h = None
        self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, eng, nsteps):
        rc = self.L.lib.mpm_batch_add(self.h, eng.h, nsteps)
        if rc:
            raise MPMError(rc, "mpm_batch_add")
        self.members.append(eng)

    def remove(self, eng):
        rc = self.L.lib.mpm_batch_remove(self.h, eng.h)
        if rc:
            raise MPMError(rc, "mpm_batch_remove")
        self.members.remove(eng)

    def set_budget(self, eng, nsteps):
        rc = self.L.lib.mpm_batch_set_budget(self.h, eng.h, nsteps)
        if rc:
            raise MPMError(rc, "mpm_batch_set_budget")

    def autotune(self, max_steps):
        chosen, rate = C.c_uint32(), C.c_double()
        rc = self.L.lib.mpm_batch_autotune(self.h, max_steps, C.byref(chosen), C.byref(rate))
        if rc:
            raise MPMError(rc, "mpm_batch_autotune")
        return chosen.value, rate.value

    def step(self, max_steps, flags=0):
        info = BatchInfo()
        rc = self.L.lib.mpm_batch_step(self.h, max_steps, flags, C.byref(info))
        return rc, info

    @staticmethod
    def scene_info(eng):
        info = StepInfo()
        left = C.c_uint32()
        rc = eng.L.lib.mpm_scene_info(eng.h, C.byref(info), C.byref(left))
        if rc:
            raise MPMError(rc, "mpm_scene_info")
        return info, left.value


class Server:
    """The served form (mpm_server_* of include/mpm.h): requests with arrival times, admitted by device
    footprint into a continuously batched run."""

    def __init__(self, lib=None, hbm_budget_bytes=0, max_scenes=0, chunk_steps=8, use_graph=True, keep_results=False):
        self.L = lib or product_library()
        opt = ServerOptions(hbm_budget_bytes, max_scenes, chunk_steps, 1 if use_graph else 0, 1 if keep_results else 0)
        h = C.c_void_p()
        rc = self.L.lib.mpm_server_create(C.byref(opt), C.byref(h))
        if rc:
            raise MPMError(rc, "mpm_server_create")
        self.h = h
        self.dims = []

    def close(self):
        if self.h is not None and self.h.value:
            self.L.lib.mpm_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, cfg, plates, objects, x, nsteps, arrival_s, **fields):
        x = np.ascontiguousarray(x, dtype=np.float64)
        keep = [x]
        pa = ParticleArrays()
        pa.x = _dptr(x)
        for k, a in fields.items():
            a = np.ascontiguousarray(a, dtype=np.float64)
            if k in ("F", "Bmat") and a.ndim == 3:
                a = np.ascontiguousarray(a.transpose(0, 2, 1))
            keep.append(a)
            setattr(pa, k, _dptr(a))
        parr = (Plate * max(1, len(plates)))(*plates)
        oarr = (Object * max(1, len(objects)))(*objects)
        d = RequestDesc(C.pointer(cfg), parr, len(plates), len(objects), oarr, C.pointer(pa), x.shape[0], nsteps,
                        float(arrival_s))
        rid = C.c_uint64()
        rc = self.L.lib.mpm_server_submit(self.h, C.byref(d), C.byref(rid))
        if rc:
            raise MPMError(rc, "mpm_server_submit")
        self.dims.append((cfg.dim, x.shape[0]))
        return rid.value

    def run(self, time_limit_s=0.0):
        st = ServerStats()
        rc = self.L.lib.mpm_server_run(self.h, time_limit_s, C.byref(st))
        return rc, st

    def result(self, rid):
        dim, n = self.dims[rid]
        out = {"x": np.zeros((n, dim)), "v": np.zeros((n, dim)), "F": np.zeros((n, dim * dim)),
               "Bmat": np.zeros((n, dim * dim))}
        for k in SCALARS:
            out[k] = np.zeros(n)
        pa = ParticleArrays()
        for k, a in out.items():
            setattr(pa, k, _dptr(a))
        info = StepInfo()
        lat = C.c_double()
        rc = self.L.lib.mpm_server_result(self.h, rid, C.byref(pa), C.byref(info), C.byref(lat))
        if rc:
            raise MPMError(rc, "mpm_server_result")
        out["F"] = out["F"].reshape(n, dim, dim).transpose(0, 2, 1).copy()
        out["Bmat"] = out["Bmat"].reshape(n, dim, dim).transpose(0, 2, 1).copy()
        return out, info, lat.value
