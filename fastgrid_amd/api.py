"""ctypes binding of libfgmaps (include/fgmaps.h).

The library is built in-tree by fastgrid_amd.build; importing this module
loads it and raises if it is missing.  There is no Python or CPU fallback for
any map computation: Context() fails when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfgmaps.so")

FG_OK = 0
PREC_PLAIN, PREC_COMPENSATED, PREC_FP64, PREC_FP32 = 0, 1, 2, 3
RUN_ELEC, RUN_AFFINITY, RUN_FLOATING, RUN_ALL = 1, 2, 4, 7
RUN_PHASE_TIMING = 8  # fg_timing's elec_ms/affinity_ms for compensated and FP32 runs
BATCH_NO_GRAPH = 1
# dielectric variants of the plain form's e-kernel (FG_PLAIN_DIEL_*)
PLAIN_DIEL_TABLE, PLAIN_DIEL_AUTO, PLAIN_DIEL_INPLACE, PLAIN_DIEL_TEXTURE = 0, 1, 2, 3
TABLE_LEN = 8192
NOALLOC_SCRATCH_BYTES = 32 << 20  # FG_NOALLOC_SCRATCH_BYTES

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)


class FgGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("div2x", C.c_int32), ("div2y", C.c_int32), ("div2z", C.c_int32), ("spacing", C.c_double)]


class FgAtoms(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", _dp), ("y", _dp), ("z", _dp), ("w", _dp),
                ("charge", _dp), ("vol", _dp), ("solpar", _dp), ("type", _ip), ("hbond", _ip)]


class FgHbondGeom(C.Structure):
    _fields_ = [("rvector", _dp), ("rvector2", _dp), ("rexp", _ip), ("disorder", _u8p)]


class FgTables(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("n_maps", C.c_int32), ("pair", _dp), ("desolv", _dp), ("eps", _dp)]


class FgMapParams(C.Structure):
    _fields_ = [("hbond", C.c_int32), ("is_hbonder", C.c_int32), ("is_covalent", C.c_int32),
                ("solpar_probe", C.c_double), ("vol_probe", C.c_double)]


class FgJobDesc(C.Structure):
    _fields_ = [("grid", FgGrid), ("atoms", FgAtoms), ("geom", FgHbondGeom), ("tables", FgTables),
                ("n_maps", C.c_int32), ("maps", C.POINTER(FgMapParams)), ("covalent", _dp),
                ("solpar_q", C.c_double), ("hydrogen", C.c_int32), ("dist_dep_diel", C.c_int32),
                ("cov_on", C.c_int32), ("cov_point", C.c_double * 3), ("cov_hw_sq_inv", C.c_double),
                ("cov_barrier", C.c_double), ("floating_grid", C.c_int32)]


class FgTiming(C.Structure):
    _fields_ = [("elec_ms", C.c_double), ("affinity_ms", C.c_double), ("total_ms", C.c_double)]


class FgSummary(C.Structure):
    _fields_ = [("n_maps", C.c_int32), ("min", C.c_double * 18), ("max", C.c_double * 18),
                ("has_charge", C.c_int32), ("charge_integral", C.c_int32), ("total_charge", C.c_double),
                ("charge_residual", C.c_double)]


class FgShape(C.Structure):
    _fields_ = [("pz", C.c_int32), ("tile", C.c_int32), ("bx", C.c_int32), ("px", C.c_int32)]


class FgLoadOpts(C.Structure):
    _fields_ = [("v4", C.c_int32), ("ag421_param_order", C.c_int32)]


class FgSynthSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_atoms", C.c_int32), ("shape", C.c_int32),
                ("a", C.c_double), ("b", C.c_double), ("c", C.c_double), ("min_sep", C.c_double),
                ("npts", C.c_int32 * 3), ("spacing", C.c_double), ("ligand_types", C.c_char_p),
                ("stem", C.c_char_p)]


class FgMapGrid(C.Structure):
    """AutoDock's map-set geometry (GridMapSetInfo, fg_map_grid)."""
    _fields_ = [("spacing", C.c_double), ("inv_spacing", C.c_double), ("num_points", C.c_int32 * 3),
                ("num_points1", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("center", C.c_double * 3)]


class FgLigand(C.Structure):
    _fields_ = [("natoms", C.c_int32), ("type", C.POINTER(C.c_int32)), ("charge", C.POINTER(C.c_float)),
                ("abs_charge", C.POINTER(C.c_float)), ("ignore", C.POINTER(C.c_int32)), ("elec_slot", C.c_int32),
                ("desolv_slot", C.c_int32)]


# Every symbol include/fgmaps.h declares, with its ctypes signature.
SIGNATURES = {
    "fg_last_error": (C.c_char_p, []),
    "fg_round3": (C.c_double, [C.c_double]),
    "fg_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "fg_destroy": (None, [C.c_void_p]),
    "fg_device_count": (C.c_int, []),
    "fg_elec_map": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(FgAtoms), _dp, C.c_int, C.c_int, _dp]),
    "fg_covalent_map": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(C.c_double), C.c_double, C.c_double, _dp]),
    "fg_floating_grid": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(FgAtoms), C.POINTER(C.c_float)]),
    "fg_affinity_d_maps": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(FgAtoms), C.POINTER(FgHbondGeom),
                                     C.POINTER(FgTables), C.POINTER(FgMapParams), C.c_int, C.c_double, C.c_int,
                                     C.POINTER(_dp), _dp]),
    "fg_elec_map_async": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(FgAtoms), _dp, C.c_int, C.c_int, _dp,
                                    C.POINTER(C.c_void_p)]),
    "fg_affinity_d_maps_async": (C.c_int, [C.c_void_p, C.POINTER(FgGrid), C.POINTER(FgAtoms),
                                           C.POINTER(FgHbondGeom), C.POINTER(FgTables), C.POINTER(FgMapParams),
                                           C.c_int, C.c_double, C.c_int, C.POINTER(_dp), _dp,
                                           C.POINTER(C.c_void_p)]),
    "fg_wait": (C.c_int, [C.c_void_p]),
    "fg_job_create": (C.c_int, [C.c_void_p, C.POINTER(FgJobDesc), C.POINTER(C.c_void_p)]),
    "fg_job_create_many": (C.c_int, [C.c_void_p, C.POINTER(FgJobDesc), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "fg_job_destroy": (None, [C.c_void_p]),
    "fg_job_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint, C.c_int]),
    "fg_sync": (C.c_int, [C.c_void_p]),
    "fg_job_fetch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_dp)]),
    "fg_job_fetch_many": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.POINTER(_dp))]),
    "fg_job_fetch_milli": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(_ip)]),
    "fg_job_fetch_milli_many": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.POINTER(_ip))]),
    "fg_job_cutoff_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "fg_job_summary": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FgSummary)]),
    "fg_job_get_shape": (C.c_int, [C.c_void_p, C.POINTER(FgShape)]),
    "fg_job_set_shape": (C.c_int, [C.c_void_p, C.POINTER(FgShape)]),
    "fg_job_set_plain_dielectric": (C.c_int, [C.c_void_p, C.c_int]),
    "fg_job_plain_dielectric": (C.c_int, [C.c_void_p]),
    "fg_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_int]),
    "fg_tune_report": (C.c_char_p, [C.c_void_p]),
    "fg_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_uint, C.c_int, C.c_uint,
                                  C.POINTER(C.c_void_p)]),
    "fg_batch_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "fg_batch_destroy": (None, [C.c_void_p]),
    "fg_batch_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int, C.c_uint, C.c_int]),
    "fg_last_timing": (C.c_int, [C.c_void_p, C.POINTER(FgTiming)]),
    "fg_prep_load": (C.c_int, [C.c_char_p, C.POINTER(FgLoadOpts), C.POINTER(C.c_void_p)]),
    "fg_prep_exit_code": (C.c_int, []),
    "fg_prep_load_many": (C.c_int, [C.POINTER(C.c_char_p), C.c_int, C.POINTER(FgLoadOpts), C.c_int,
                                    C.POINTER(C.c_void_p)]),
    "fg_prep_free": (None, [C.c_void_p]),
    "fg_prep_desc": (C.c_int, [C.c_void_p, C.POINTER(FgJobDesc)]),
    "fg_prep_num_maps": (C.c_int, [C.c_void_p]),
    "fg_prep_map_type": (C.c_char_p, [C.c_void_p, C.c_int]),
    "fg_prep_num_atoms": (C.c_int, [C.c_void_p]),
    "fg_prep_num_points": (C.c_int64, [C.c_void_p]),
    "fg_prep_gpf_echo": (C.c_char_p, [C.c_void_p]),
    "fg_prep_summary": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp]),
    "fg_write_outputs": (C.c_int, [C.c_void_p, C.POINTER(_dp), C.c_char_p, C.c_int]),
    "fg_write_outputs_milli": (C.c_int, [C.c_void_p, C.POINTER(_ip), _dp, C.c_char_p, C.c_int]),
    "fg_synth_write": (C.c_int, [C.POINTER(FgSynthSpec), C.c_char_p]),
    "fg_read_fld": (C.c_int, [C.c_char_p, C.POINTER(FgMapGrid)]),
    "fg_read_map": (C.c_int, [C.c_char_p, C.POINTER(FgMapGrid), C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "fg_prep_map_grid": (C.c_int, [C.c_void_p, C.POINTER(FgMapGrid)]),
    "fg_mapset_create": (C.c_int, [C.c_void_p, C.POINTER(FgMapGrid), C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_void_p)]),
    "fg_mapset_from_job": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FgMapGrid), C.POINTER(C.c_void_p)]),
    "fg_mapset_fetch": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "fg_mapset_info": (C.c_int, [C.c_void_p, C.POINTER(FgMapGrid), C.POINTER(C.c_int)]),
    "fg_mapset_destroy": (None, [C.c_void_p]),
    "fg_trilinterp": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FgLigand), C.c_int, C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.POINTER(C.c_float), _dp, _dp, C.POINTER(C.c_float)]),
    "fg_trilinterp_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FgLigand), C.c_int, C.POINTER(C.c_float),
                                        C.POINTER(C.c_float), C.POINTER(C.c_float), _dp, _dp, C.POINTER(C.c_float)]),
}


class FgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libfgmaps error {code}: {msg}")
        self.code = code
        self.exit_code = None  # set by the GPF loaders: the reference's exit status


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m fastgrid_amd.build` (no fallback exists)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


def _check(rc: int) -> None:
    if rc != FG_OK:
        raise FgError(rc, (LIB.fg_last_error() or b"").decode())


def round3(a: float) -> float:
    return LIB.fg_round3(a)


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


def _arr(ptr, n: int, dtype) -> np.ndarray:
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


@dataclass
class JobArrays:
    """Prepared inputs of one map set as numpy arrays (host copies).  Field
    meanings follow fg_job_desc in include/fgmaps.h."""
    n: tuple[int, int, int]
    div2: tuple[int, int, int]
    spacing: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    w: np.ndarray
    charge: np.ndarray
    vol: np.ndarray
    solpar: np.ndarray
    type: np.ndarray
    hbond: np.ndarray
    rvector: np.ndarray
    rvector2: np.ndarray
    rexp: np.ndarray
    disorder: np.ndarray
    n_types: int
    pair: np.ndarray
    desolv: np.ndarray
    eps: np.ndarray
    map_hbond: np.ndarray
    map_is_hbonder: np.ndarray
    map_is_covalent: np.ndarray
    map_solpar: np.ndarray
    map_vol: np.ndarray
    covalent: np.ndarray | None
    solpar_q: float
    hydrogen: int
    dist_dep_diel: int
    # covalent maps from their parameters when covalent is None
    cov_on: int = 0
    cov_point: tuple = (0.0, 0.0, 0.0)
    cov_hw_sq_inv: float = 1.0
    cov_barrier: float = 1000.0
    floating_grid: int = 0

    @property
    def n_atoms(self) -> int:
        return int(self.x.size)

    @property
    def n_maps(self) -> int:
        return int(self.map_hbond.size)

    @property
    def P(self) -> int:
        return int(self.n[0] * self.n[1] * self.n[2])

    def desc(self) -> tuple[FgJobDesc, list]:
        """fg_job_desc pointing into these arrays; keep the returned list alive."""
        keep = []

        def c(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a

        d = FgJobDesc()
        d.grid = FgGrid(self.n[0], self.n[1], self.n[2], self.div2[0], self.div2[1], self.div2[2], self.spacing)
        at = FgAtoms()
        at.n = self.n_atoms
        for f in ("x", "y", "z", "w", "charge", "vol", "solpar"):
            setattr(at, f, _ptr(c(getattr(self, f), np.float64), C.c_double))
        at.type = _ptr(c(self.type, np.int32), C.c_int32)
        at.hbond = _ptr(c(self.hbond, np.int32), C.c_int32)
        d.atoms = at
        d.geom = FgHbondGeom(_ptr(c(self.rvector, np.float64), C.c_double), _ptr(c(self.rvector2, np.float64), C.c_double),
                             _ptr(c(self.rexp, np.int32), C.c_int32), _ptr(c(self.disorder, np.uint8), C.c_uint8))
        d.tables = FgTables(self.n_types, self.n_maps, _ptr(c(self.pair, np.float64), C.c_double),
                            _ptr(c(self.desolv, np.float64), C.c_double), _ptr(c(self.eps, np.float64), C.c_double))
        mp = (FgMapParams * max(self.n_maps, 1))()
        for m in range(self.n_maps):
            mp[m] = FgMapParams(int(self.map_hbond[m]), int(self.map_is_hbonder[m]), int(self.map_is_covalent[m]),
                                float(self.map_solpar[m]), float(self.map_vol[m]))
        keep.append(mp)
        d.n_maps = self.n_maps
        d.maps = C.cast(mp, C.POINTER(FgMapParams))
        d.covalent = _ptr(c(self.covalent, np.float64), C.c_double) if self.covalent is not None else \
            C.cast(None, _dp)
        d.solpar_q = self.solpar_q
        d.hydrogen = self.hydrogen
        d.dist_dep_diel = self.dist_dep_diel
        d.cov_on = int(self.cov_on)
        d.cov_point = (C.c_double * 3)(*[float(v) for v in self.cov_point])
        d.cov_hw_sq_inv = float(self.cov_hw_sq_inv)
        d.cov_barrier = float(self.cov_barrier)
        d.floating_grid = int(self.floating_grid)
        return d, keep


class Prepared:
    """A GPF loaded and precomputed by the host front end (fg_prep_load)."""

    def __init__(self, gpf: str | None, v4: bool = False, ag421_param_order: bool = False):
        self._h = C.c_void_p()
        if gpf is None:  # filled by prepare_many
            return
        opts = FgLoadOpts(1 if v4 else 0, 1 if ag421_param_order else 0)
        rc = LIB.fg_prep_load(gpf.encode(), C.byref(opts), C.byref(self._h))
        if rc != FG_OK:
            # the reference's exit status for the failure (fg_prep_exit_code)
            err = FgError(rc, LIB.fg_last_error().decode(errors="replace"))
            err.exit_code = LIB.fg_prep_exit_code()
            raise err

    @property
    def n_atoms(self) -> int:
        return int(LIB.fg_prep_num_atoms(self._h))

    @property
    def P(self) -> int:
        return int(LIB.fg_prep_num_points(self._h))

    @property
    def n_maps(self) -> int:
        return int(LIB.fg_prep_num_maps(self._h))

    def close(self) -> None:
        if self._h:
            LIB.fg_prep_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def map_types(self) -> list[str]:
        return [LIB.fg_prep_map_type(self._h, m).decode() for m in range(LIB.fg_prep_num_maps(self._h))]

    def arrays(self) -> JobArrays:
        d = FgJobDesc()
        _check(LIB.fg_prep_desc(self._h, C.byref(d)))
        g, a, geo, t = d.grid, d.atoms, d.geom, d.tables
        N, M = a.n, d.n_maps
        P = g.nx * g.ny * g.nz
        mp = [d.maps[m] for m in range(M)]
        return JobArrays(
            n=(g.nx, g.ny, g.nz), div2=(g.div2x, g.div2y, g.div2z), spacing=g.spacing,
            x=_arr(a.x, N, np.float64), y=_arr(a.y, N, np.float64), z=_arr(a.z, N, np.float64),
            w=_arr(a.w, N, np.float64), charge=_arr(a.charge, N, np.float64), vol=_arr(a.vol, N, np.float64),
            solpar=_arr(a.solpar, N, np.float64), type=_arr(a.type, N, np.int32), hbond=_arr(a.hbond, N, np.int32),
            rvector=_arr(geo.rvector, 3 * N, np.float64), rvector2=_arr(geo.rvector2, 3 * N, np.float64),
            rexp=_arr(geo.rexp, N, np.int32), disorder=_arr(geo.disorder, N, np.uint8),
            n_types=t.n_types, pair=_arr(t.pair, t.n_types * TABLE_LEN * max(M, 1), np.float64),
            desolv=_arr(t.desolv, TABLE_LEN, np.float64), eps=_arr(t.eps, TABLE_LEN, np.float64),
            map_hbond=np.array([m.hbond for m in mp], np.int32),
            map_is_hbonder=np.array([m.is_hbonder for m in mp], np.int32),
            map_is_covalent=np.array([m.is_covalent for m in mp], np.int32),
            map_solpar=np.array([m.solpar_probe for m in mp], np.float64),
            map_vol=np.array([m.vol_probe for m in mp], np.float64),
            covalent=_arr(d.covalent, M * P, np.float64) if d.covalent else None,
            solpar_q=d.solpar_q, hydrogen=d.hydrogen, dist_dep_diel=d.dist_dep_diel,
            cov_on=d.cov_on, cov_point=tuple(d.cov_point), cov_hw_sq_inv=d.cov_hw_sq_inv, cov_barrier=d.cov_barrier,
            floating_grid=d.floating_grid)

    def map_grid(self) -> FgMapGrid:
        """The geometry AutoDock reads back from this prep's .fld."""
        g = FgMapGrid()
        _check(LIB.fg_prep_map_grid(self._h, C.byref(g)))
        return g

    def write_outputs(self, maps: list[np.ndarray], program_name: str = "fastgrid4", threads: int = 0) -> None:
        keep = [np.ascontiguousarray(m, dtype=np.float64) for m in maps]
        ptrs = (_dp * len(keep))(*[_ptr(m, C.c_double) for m in keep])
        _check(LIB.fg_write_outputs(self._h, ptrs, program_name.encode(), threads))

    def write_outputs_milli(self, maps: list[np.ndarray], fmap=None, program_name: str = "fastgrid4",
                            threads: int = 0) -> None:
        """The files from int32 thousandths (fg_write_outputs_milli)."""
        keep = [np.ascontiguousarray(m, dtype=np.int32) for m in maps]
        ptrs = (_ip * len(keep))(*[_ptr(m, C.c_int32) for m in keep])
        fm = None if fmap is None else np.ascontiguousarray(fmap, dtype=np.float64)
        _check(LIB.fg_write_outputs_milli(self._h, ptrs, _ptr(fm, C.c_double) if fm is not None else C.cast(None, _dp),
                                          program_name.encode(), threads))


def prepare_many(gpfs: list[str], threads: int = 0, v4: bool = False,
                 ag421_param_order: bool = False) -> list[Prepared]:
    """Loads and precomputes many GPFs on host threads (fg_prep_load_many)."""
    n = len(gpfs)
    names = (C.c_char_p * max(n, 1))(*[g.encode() for g in gpfs])
    out = (C.c_void_p * max(n, 1))()
    opts = FgLoadOpts(1 if v4 else 0, 1 if ag421_param_order else 0)
    rc = LIB.fg_prep_load_many(names, n, C.byref(opts), int(threads), out)
    if rc != FG_OK:
        err = FgError(rc, LIB.fg_last_error().decode(errors="replace"))
        err.exit_code = LIB.fg_prep_exit_code()
        raise err
    preps = []
    for i in range(n):
        p = Prepared(None)
        p._h = C.c_void_p(out[i])
        preps.append(p)
    return preps


@dataclass
class _JobInfo:
    """What a job fetches: points, affinity maps and whether a floating grid."""
    P: int
    n_maps: int
    floating_grid: int


def _describe(item):
    """(fg_job_desc, objects to keep alive, what the job fetches) of host
    arrays or of a Prepared."""
    if isinstance(item, Prepared):
        d = FgJobDesc()
        _check(LIB.fg_prep_desc(item.handle, C.byref(d)))
        g = d.grid
        return d, [item], _JobInfo(P=g.nx * g.ny * g.nz, n_maps=d.n_maps, floating_grid=d.floating_grid)
    d, keep = item.desc()
    return d, keep, item


def create_jobs(ctx: "Context", items: list, threads: int = 0) -> list["Job"]:
    """Jobs of many JobArrays / Prepared at once (fg_job_create_many): the
    host-side inputs are built on parallel threads."""
    n = len(items)
    described = [_describe(it) for it in items]
    descs = (FgJobDesc * max(n, 1))(*[x[0] for x in described])
    out = (C.c_void_p * max(n, 1))()
    _check(LIB.fg_job_create_many(ctx.handle, descs, n, int(threads), out))
    jobs = []
    for i, (_, keep, info) in enumerate(described):
        j = Job.__new__(Job)
        j.ctx, j._h, j._keep, j.arrays = ctx, C.c_void_p(out[i]), keep, info
        jobs.append(j)
    return jobs


def device_count() -> int:
    return LIB.fg_device_count()


class Context:
    """One GPU, one HIP stream (fg_create)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(LIB.fg_create(device, C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            LIB.fg_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def sync(self) -> None:
        _check(LIB.fg_sync(self._h))

    def timing(self) -> FgTiming:
        t = FgTiming()
        _check(LIB.fg_last_timing(self._h, C.byref(t)))
        return t

    def elec_map(self, arrays: JobArrays, precision: int = PREC_COMPENSATED) -> np.ndarray:
        """The electrostatic map (fg_elec_map)."""
        d, keep = arrays.desc()
        out = np.zeros(arrays.P, np.float64)
        _check(LIB.fg_elec_map(self._h, C.byref(d.grid), C.byref(d.atoms), d.tables.eps, d.dist_dep_diel, precision,
                               _ptr(out, C.c_double)))
        return out

    def elec_map_async(self, arrays: JobArrays, precision: int = PREC_COMPENSATED) -> "Pending":
        """fg_elec_map_async: returns at once; .wait() gives the map."""
        d, keep = arrays.desc()
        out = np.zeros(arrays.P, np.float64)
        h = C.c_void_p()
        _check(LIB.fg_elec_map_async(self._h, C.byref(d.grid), C.byref(d.atoms), d.tables.eps, d.dist_dep_diel,
                                     precision, _ptr(out, C.c_double), C.byref(h)))
        return Pending(h, out)

    def covalent_map(self, arrays: JobArrays) -> np.ndarray:
        """Initial values of the covalent maps (fg_covalent_map)."""
        d, keep = arrays.desc()
        out = np.zeros(arrays.P, np.float64)
        pt = (C.c_double * 3)(*[float(v) for v in arrays.cov_point])
        _check(LIB.fg_covalent_map(self._h, C.byref(d.grid), pt, float(arrays.cov_hw_sq_inv),
                                   float(arrays.cov_barrier), _ptr(out, C.c_double)))
        return out

    def floating_grid(self, arrays: JobArrays) -> np.ndarray:
        """The floating grid (fg_floating_grid), float32."""
        d, keep = arrays.desc()
        out = np.zeros(arrays.P, np.float32)
        _check(LIB.fg_floating_grid(self._h, C.byref(d.grid), C.byref(d.atoms), _ptr(out, C.c_float)))
        return out

    def affinity_d_maps(self, arrays: JobArrays, asynchronous: bool = False):
        """Affinity and desolvation maps (fg_affinity_d_maps): (maps, d).
        Covalent maps start from arrays.covalent, else from fg_covalent_map.
        asynchronous: fg_affinity_d_maps_async, returns a Pending whose
        .wait() gives (maps, d)."""
        d, keep = arrays.desc()
        P, M = arrays.P, arrays.n_maps
        cov = None
        if arrays.covalent is None and arrays.cov_on:
            cov = self.covalent_map(arrays)
        outs = [arrays.covalent[m * P:(m + 1) * P].copy() if arrays.covalent is not None else
                (cov.copy() if (cov is not None and arrays.map_is_covalent[m]) else np.zeros(P))
                for m in range(M)]
        dmap = np.zeros(P)
        ptrs = (_dp * max(M, 1))(*[_ptr(o, C.c_double) for o in outs])
        args = (self._h, C.byref(d.grid), C.byref(d.atoms), C.byref(d.geom), C.byref(d.tables), d.maps, M, d.solpar_q,
                d.hydrogen, ptrs, _ptr(dmap, C.c_double))
        if asynchronous:
            h = C.c_void_p()
            _check(LIB.fg_affinity_d_maps_async(*args, C.byref(h)))
            return Pending(h, (outs, dmap))
        _check(LIB.fg_affinity_d_maps(*args))
        return outs, dmap


class Pending:
    """An asynchronous one-shot computation (fg_handle): wait() returns its
    outputs once fg_wait has written them."""

    def __init__(self, handle, result):
        self._h, self._result = handle, result

    def wait(self):
        if self._h:
            h, self._h = self._h, C.c_void_p()
            _check(LIB.fg_wait(h))
        return self._result

    def __del__(self):
        try:
            self.wait()
        except Exception:
            pass


class Job:
    """Device-resident inputs and maps of one map set (fg_job_*)."""

    def __init__(self, ctx: Context, arrays: JobArrays | Prepared):
        """From host arrays, or straight from a Prepared (no copy of its arrays)."""
        self.ctx = ctx
        self._h = C.c_void_p()
        d, self._keep, self.arrays = _describe(arrays)
        _check(LIB.fg_job_create(ctx.handle, C.byref(d), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            LIB.fg_job_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def run(self, what: int = RUN_ALL, precision: int = PREC_COMPENSATED) -> None:
        _check(LIB.fg_job_run(self.ctx.handle, self._h, what, precision))

    def fetch(self) -> list[np.ndarray]:
        """[affinity maps..., e, d] (+ the floating grid when the job has one), unrounded FP64."""
        P, M = self.arrays.P, self.arrays.n_maps
        n = M + 2 + (1 if self.arrays.floating_grid else 0)
        outs = [np.empty(P, np.float64) for _ in range(n)]
        ptrs = (_dp * n)(*[_ptr(o, C.c_double) for o in outs])
        _check(LIB.fg_job_fetch(self.ctx.handle, self._h, ptrs))
        return outs

    def fetch_milli(self) -> list[np.ndarray]:
        """[affinity maps..., e, d] rounded on the device to int32 thousandths
        (fg_job_fetch_milli): the values the map files print, times 1000."""
        P, M = self.arrays.P, self.arrays.n_maps
        outs = [np.empty(P, np.int32) for _ in range(M + 2)]
        ptrs = (_ip * (M + 2))(*[_ptr(o, C.c_int32) for o in outs])
        _check(LIB.fg_job_fetch_milli(self.ctx.handle, self._h, ptrs))
        return outs

    def summary(self) -> dict:
        """GPU checks of the last run (fg_job_summary): per map the min/max of
        the rounded values, the receptor's total charge and its distance to
        the nearest integer."""
        sm = FgSummary()
        _check(LIB.fg_job_summary(self.ctx.handle, self._h, C.byref(sm)))
        n = sm.n_maps
        return {"min": [sm.min[i] for i in range(n)], "max": [sm.max[i] for i in range(n)],
                "has_charge": bool(sm.has_charge), "total_charge": sm.total_charge,
                "charge_residual": sm.charge_residual, "charge_integral": bool(sm.charge_integral)}

    def cutoff_pairs(self) -> int:
        v = C.c_int64()
        _check(LIB.fg_job_cutoff_pairs(self.ctx.handle, self._h, C.byref(v)))
        return int(v.value)

    @property
    def plain_dielectric(self) -> int:
        """The plain form's dielectric variant that runs (PLAIN_DIEL_*)."""
        v = LIB.fg_job_plain_dielectric(self._h)
        if v < 0:
            _check(v)
        return v

    @plain_dielectric.setter
    def plain_dielectric(self, kind: int) -> None:
        _check(LIB.fg_job_set_plain_dielectric(self._h, int(kind)))

    @property
    def shape(self) -> tuple[int, int, int, int]:
        """(points per lane along z, atoms per tile, brick lanes along x,
        points per lane along x)."""
        sh = FgShape()
        _check(LIB.fg_job_get_shape(self._h, C.byref(sh)))
        return sh.pz, sh.tile, sh.bx, sh.px

    @shape.setter
    def shape(self, v: tuple[int, ...]) -> None:
        """(pz, tile, bx) or (pz, tile, bx, px); px defaults to 1."""
        v = tuple(int(x) for x in v)
        sh = FgShape(*(v if len(v) == 4 else v + (1,)))
        _check(LIB.fg_job_set_shape(self._h, C.byref(sh)))


def _handles(jobs: list[Job]):
    return (C.c_void_p * len(jobs))(*[j.handle for j in jobs])


def fetch_many(ctx: Context, jobs: list[Job], milli: bool = False) -> list[list[np.ndarray]]:
    """Every job's maps in one pipelined copy (fg_job_fetch_many, or
    fg_job_fetch_milli_many with milli: int32 thousandths, no floating
    grid).  The maps of all jobs are views of one host array: one allocation
    is large enough for numpy to ask for transparent huge pages, so the first
    writes fault once per 2 MiB instead of once per 4 KiB page.  A view keeps
    the whole array alive."""
    dt = np.dtype(np.int32 if milli else np.float64)
    counts = [j.arrays.n_maps + 2 + (1 if (j.arrays.floating_grid and not milli) else 0) for j in jobs]
    sizes = [int(j.arrays.P) for j in jobs]
    starts = np.zeros(len(jobs) + 1, np.int64)
    np.cumsum([c * p for c, p in zip(counts, sizes)], out=starts[1:])
    buf = np.empty(max(int(starts[-1]), 1), dt)
    base = buf.ctypes.data
    # the address of every map, then the address of every job's row of them
    offs = np.concatenate([starts[i] + sizes[i] * np.arange(counts[i], dtype=np.int64) for i in range(len(jobs))]
                          or [np.zeros(0, np.int64)])
    addrs = (base + offs * dt.itemsize).astype(np.uintp)
    row0 = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64) if jobs else np.zeros(1, np.int64)
    rows = (addrs.ctypes.data + row0 * addrs.itemsize).astype(np.uintp)
    fn = LIB.fg_job_fetch_milli_many if milli else LIB.fg_job_fetch_many
    _check(fn(ctx.handle, _handles(jobs), len(jobs), C.cast(rows.ctypes.data, C.POINTER(C.POINTER(_ip if milli else _dp)))))
    return [[buf[int(o):int(o) + sizes[i]] for o in offs[int(row0[i]):int(row0[i]) + counts[i]]]
            for i in range(len(jobs))]


def autotune(ctx: Context, jobs: list[Job], precision: int = PREC_COMPENSATED) -> str:
    """Tunes the electrostatic kernel shape of the jobs; returns the report."""
    _check(LIB.fg_autotune(ctx.handle, _handles(jobs), len(jobs), precision))
    return (LIB.fg_tune_report(ctx.handle) or b"").decode()


class Batch:
    """Many jobs run as one hipGraph of persistent kernels (fg_batch_*)."""

    def __init__(self, ctx: Context, jobs: list[Job], what: int = RUN_ALL, precision: int = PREC_COMPENSATED,
                 graph: bool = True):
        self.ctx = ctx
        self.jobs = list(jobs)
        self._h = C.c_void_p()
        _check(LIB.fg_batch_create(ctx.handle, _handles(self.jobs), len(self.jobs), what, precision,
                                   0 if graph else BATCH_NO_GRAPH, C.byref(self._h)))

    def launch(self) -> None:
        _check(LIB.fg_batch_launch(self.ctx.handle, self._h))

    def close(self) -> None:
        if self._h:
            LIB.fg_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_run(ctx: Context, jobs: list[Job], what: int = RUN_ALL, precision: int = PREC_COMPENSATED) -> None:
    _check(LIB.fg_batch_run(ctx.handle, _handles(jobs), len(jobs), what, precision))


def synth_write(directory: str, stem: str, seed: int, n_atoms: int, shape: str, dims, npts, ligand_types: str,
                min_sep: float = 1.2, spacing: float = 0.375) -> tuple[str, str]:
    """Writes <directory>/<stem>.pdbqt and .gpf; returns their paths."""
    s = FgSynthSpec()
    s.seed = seed
    s.n_atoms = n_atoms
    if shape == "sphere":
        s.shape, s.a = 0, float(dims)
    else:
        s.shape = 1
        s.a, s.b, s.c = (float(v) for v in dims)
    s.min_sep = min_sep
    s.npts = (C.c_int32 * 3)(*[int(v) for v in npts])
    s.spacing = spacing
    s.ligand_types = ligand_types.encode()
    s.stem = stem.encode()
    _check(LIB.fg_synth_write(C.byref(s), directory.encode()))
    return os.path.join(directory, stem + ".pdbqt"), os.path.join(directory, stem + ".gpf")


# ---------------------------------------------------------------------------
# The maps' consumer: AutoDock map sets and trilinear interpolation
# ---------------------------------------------------------------------------
def read_fld(path: str) -> FgMapGrid:
    """AutoDock's readfield of a .fld (fg_read_fld)."""
    g = FgMapGrid()
    _check(LIB.fg_read_fld(path.encode(), C.byref(g)))
    return g


def read_mapset(g: FgMapGrid, paths: list[str]) -> np.ndarray:
    """AutoDock's readmap of every file, slot i = paths[i]: floats [z][y][x][slot]."""
    n = list(g.num_points1)
    out = np.zeros((n[2], n[1], n[0], len(paths)), dtype=np.float32)
    for slot, p in enumerate(paths):
        _check(LIB.fg_read_map(p.encode(), C.byref(g), slot, len(paths), _ptr(out, C.c_float)))
    return out


class MapSet:
    """A map set in HBM, floats [z][y][x][slot] (fg_mapset_*)."""

    def __init__(self, ctx: Context, g: FgMapGrid | None = None, values: np.ndarray | None = None,
                 job: "Job | None" = None):
        self.ctx = ctx
        self._h = C.c_void_p()
        if job is not None:
            _check(LIB.fg_mapset_from_job(ctx.handle, job.handle, C.byref(g), C.byref(self._h)))
        else:
            v = np.ascontiguousarray(values, dtype=np.float32)
            self._keep = v
            _check(LIB.fg_mapset_create(ctx.handle, C.byref(g), int(v.shape[-1]), _ptr(v, C.c_float),
                                        C.byref(self._h)))

    @classmethod
    def from_job(cls, ctx: Context, job: "Job", g: FgMapGrid) -> "MapSet":
        """The job's maps as AutoDock reads them from the printed files,
        without text (affinity maps, then e, then d)."""
        return cls(ctx, g, job=job)

    def info(self) -> tuple[FgMapGrid, int]:
        g, n = FgMapGrid(), C.c_int()
        _check(LIB.fg_mapset_info(self._h, C.byref(g), C.byref(n)))
        return g, int(n.value)

    def fetch(self) -> np.ndarray:
        g, n = self.info()
        p = list(g.num_points1)
        out = np.zeros((p[2], p[1], p[0], n), dtype=np.float32)
        _check(LIB.fg_mapset_fetch(self.ctx.handle, self._h, _ptr(out, C.c_float)))
        return out

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            LIB.fg_mapset_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trilinterp(ctx: Context, ms: MapSet, xyz: np.ndarray, types, charge, abs_charge, elec_slot: int,
               desolv_slot: int, ignore=None, per_atom: bool = True, noalloc: bool = False) -> dict:
    """AutoDock's trilinterp of every pose of xyz [pose][atom][3] on the GPU
    (fg_trilinterp, or fg_trilinterp_noalloc with noalloc).  Returns
    elec/emap per atom (float32; None unless per_atom), the two double totals
    and the returned float per pose."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    npose, natoms = xyz.shape[0], xyz.shape[1]
    t = np.ascontiguousarray(types, dtype=np.int32)
    q = np.ascontiguousarray(charge, dtype=np.float32)
    aq = np.ascontiguousarray(abs_charge, dtype=np.float32)
    ig = None if ignore is None else np.ascontiguousarray(ignore, dtype=np.int32)
    lig = FgLigand(natoms, _ptr(t, C.c_int32), _ptr(q, C.c_float), _ptr(aq, C.c_float),
                   _ptr(ig, C.c_int32) if ig is not None else C.cast(None, C.POINTER(C.c_int32)),
                   int(elec_slot), int(desolv_slot))
    out = {"elec": np.zeros((npose, natoms), np.float32) if per_atom else None,
           "emap": np.zeros((npose, natoms), np.float32) if per_atom else None,
           "elec_total": np.zeros(npose), "emap_total": np.zeros(npose), "total": np.zeros(npose, np.float32)}
    fn = LIB.fg_trilinterp_noalloc if noalloc else LIB.fg_trilinterp
    _check(fn(ctx.handle, ms.handle, C.byref(lig), npose, _ptr(xyz, C.c_float),
                             _ptr(out["elec"], C.c_float), _ptr(out["emap"], C.c_float),
                             _ptr(out["elec_total"], C.c_double), _ptr(out["emap_total"], C.c_double),
                             _ptr(out["total"], C.c_float)))
    return out
