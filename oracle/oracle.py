"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline, never by the
product (fastgrid_amd/).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "fgoracle")
TABLE_LEN = 8192

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> None:
    """Compiles liboracle.so and the fgoracle command line (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OiGrid(C.Structure):
    """AutoDock's GridMapSetInfo geometry (oracle/fginterp.h)."""
    _fields_ = [("spacing", C.c_double), ("inv_spacing", C.c_double), ("num_points", C.c_int * 3),
                ("num_points1", C.c_int * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("center", C.c_double * 3)]


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.or_open.restype = C.c_void_p
    lib.or_open.argtypes = [C.c_char_p, C.c_int, C.c_int, _ip]
    lib.or_close.argtypes = [C.c_void_p]
    lib.or_dims.argtypes = [C.c_void_p, _ip]
    lib.or_scalars.argtypes = [C.c_void_p, _dp]
    lib.or_darray.restype = _dp
    lib.or_darray.argtypes = [C.c_void_p, C.c_char_p]
    lib.or_iarray.restype = _ip
    lib.or_iarray.argtypes = [C.c_void_p, C.c_char_p]
    lib.or_disorder.restype = C.POINTER(C.c_uint8)
    lib.or_disorder.argtypes = [C.c_void_p]
    lib.or_map_params.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
    lib.or_compute.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.or_write_outputs.argtypes = [C.c_void_p, C.c_char_p]
    lib.or_round3.restype = C.c_double
    lib.or_round3.argtypes = [C.c_double]
    lib.or_elec_map.argtypes = [_ip, _ip, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int]
    lib.or_affinity_maps.argtypes = [_ip, _ip, C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp,
                                     _ip, C.POINTER(C.c_uint8), C.c_int, C.c_int, _dp, _dp, _ip, _ip, _ip, _dp, _dp,
                                     C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int]
    lib.or_count_cutoff_pairs.restype = C.c_longlong
    lib.or_count_cutoff_pairs.argtypes = [_ip, _ip, C.c_double, C.c_int, _dp, _dp, _dp, C.c_int]
    lib.or_floating_grid.argtypes = [_ip, _ip, C.c_double, C.c_int, _dp, _dp, _dp, C.POINTER(C.c_float), C.c_int]
    lib.or_covalent_map.argtypes = [_ip, _ip, C.c_double, _dp, C.c_double, C.c_double, _dp]
    lib.or_ddd_table.argtypes = [_dp]
    lib.or_desolv_table.argtypes = [C.c_double, _dp]
    _fp = C.POINTER(C.c_float)
    lib.oi_read_fld.argtypes = [C.c_char_p, C.POINTER(OiGrid)]
    lib.oi_read_map.argtypes = [C.c_char_p, C.POINTER(OiGrid), C.c_int, C.c_int, _fp]
    lib.oi_trilinterp_poses.argtypes = [C.POINTER(OiGrid), _fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp,
                                        _fp, _fp, _ip, C.c_int, _ip, _fp, _fp, _dp, _dp, _fp, C.c_int]
    return lib


LIB = _load()


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _i3(v):
    return (C.c_int * 3)(*[int(x) for x in v])


def round3(a: float) -> float:
    return LIB.or_round3(a)


def ddd_table() -> np.ndarray:
    out = np.zeros(TABLE_LEN)
    LIB.or_ddd_table(_p(out))
    return out


def desolv_table(coeff: float) -> np.ndarray:
    out = np.zeros(TABLE_LEN)
    LIB.or_desolv_table(coeff, _p(out))
    return out


class Loaded:
    """The oracle's own parse + precompute of a GPF (or_open)."""

    def __init__(self, gpf: str, v4: bool = False, ag421_param_order: bool = False):
        rc = C.c_int(0)
        self.h = LIB.or_open(gpf.encode(), int(v4), int(ag421_param_order), C.byref(rc))
        if not self.h:
            raise RuntimeError(f"oracle failed to load {gpf} (exit code {rc.value})")
        d = (C.c_int * 12)()
        LIB.or_dims(self.h, d)
        self.N, self.M, self.tableTypes = d[0], d[1], d[2]
        self.n, self.div2 = tuple(d[3:6]), tuple(d[6:9])
        self.dist_dep_diel, self.hydrogen, self.P = d[9], d[10], d[11]
        s = (C.c_double * 5)()
        LIB.or_scalars(self.h, s)
        self.spacing, self.solpar_q, self.center = s[0], s[1], tuple(s[2:5])

    def close(self):
        if self.h:
            LIB.or_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def darray(self, name: str) -> np.ndarray:
        n = {"rvector": 3 * self.N, "rvector2": 3 * self.N, "desolv": TABLE_LEN, "eps": TABLE_LEN,
             "pair": self.tableTypes * TABLE_LEN * max(self.M, 1)}.get(name, self.N)
        return np.ctypeslib.as_array(LIB.or_darray(self.h, name.encode()), shape=(n,)).copy()

    def iarray(self, name: str) -> np.ndarray:
        return np.ctypeslib.as_array(LIB.or_iarray(self.h, name.encode()), shape=(self.N,)).copy()

    def disorder(self) -> np.ndarray:
        return np.ctypeslib.as_array(LIB.or_disorder(self.h), shape=(self.N,)).copy()

    def map_params(self):
        out = []
        for m in range(self.M):
            iv, dv = (C.c_int * 3)(), (C.c_double * 2)()
            LIB.or_map_params(self.h, m, iv, dv)
            out.append((iv[0], iv[1], iv[2], dv[0], dv[1]))
        return out


def elec_map(a, round3: bool = False, threads: int = 0) -> np.ndarray:
    """FP64 electrostatic map of JobArrays-like inputs (FG/electrostatics/Cpu.cpp:27-88)."""
    e = np.zeros(a.P)
    x, y, z, w, eps = (np.ascontiguousarray(v, np.float64) for v in (a.x, a.y, a.z, a.w, a.eps))
    LIB.or_elec_map(_i3(a.n), _i3(a.div2), a.spacing, a.n_atoms, _p(x), _p(y), _p(z), _p(w), _p(eps),
                    int(a.dist_dep_diel), _p(e), int(round3), threads)
    return e


def affinity_maps(a, round3: bool = False, threads: int = 0):
    """FP64 affinity maps and desolvation map (FG/CalculateGridmaps.cpp:238-316)."""
    P, M = a.P, a.n_maps
    maps = np.zeros(max(M, 1) * P)
    if a.covalent is not None:
        maps[:M * P] = a.covalent
    elif getattr(a, "cov_on", 0):
        cov = covalent_map(a)
        for m in range(M):
            if a.map_is_covalent[m]:
                maps[m * P:(m + 1) * P] = cov
    d = np.zeros(P)
    f = lambda v: np.ascontiguousarray(v, np.float64)
    i = lambda v: np.ascontiguousarray(v, np.int32)
    x, y, z, q, vol, sol, rv, rv2, pair, des = map(f, (a.x, a.y, a.z, a.charge, a.vol, a.solpar, a.rvector,
                                                       a.rvector2, a.pair, a.desolv))
    typ, hb, rexp = map(i, (a.type, a.hbond, a.rexp))
    dis = np.ascontiguousarray(a.disorder, np.uint8)
    mhb, mis, mcov = map(i, (a.map_hbond, a.map_is_hbonder, a.map_is_covalent))
    msol, mvol = f(a.map_solpar), f(a.map_vol)
    LIB.or_affinity_maps(_i3(a.n), _i3(a.div2), a.spacing, a.n_atoms, _p(x), _p(y), _p(z), _p(q), _p(vol), _p(sol),
                         _p(typ, C.c_int), _p(hb, C.c_int), _p(rv), _p(rv2), _p(rexp, C.c_int),
                         dis.ctypes.data_as(C.POINTER(C.c_uint8)), int(a.n_types), M, _p(pair), _p(des),
                         _p(mhb, C.c_int), _p(mis, C.c_int), _p(mcov, C.c_int), _p(msol), _p(mvol),
                         float(a.solpar_q), int(a.hydrogen), _p(maps), _p(d), int(round3), threads)
    return [maps[m * P:(m + 1) * P].copy() for m in range(M)], d


def covalent_map(a) -> np.ndarray:
    """Values of a covalent map from JobArrays-like parameters (FG/main.cpp:32-52)."""
    out = np.zeros(a.P)
    pt = np.ascontiguousarray(a.cov_point, np.float64)
    LIB.or_covalent_map(_i3(a.n), _i3(a.div2), a.spacing, _p(pt), float(a.cov_hw_sq_inv), float(a.cov_barrier),
                        _p(out))
    return out


def floating_grid(a, threads: int = 0) -> np.ndarray:
    """Floating grid (FG/main.cpp:54-79), float32."""
    out = np.zeros(a.P, np.float32)
    x, y, z = (np.ascontiguousarray(v, np.float64) for v in (a.x, a.y, a.z))
    LIB.or_floating_grid(_i3(a.n), _i3(a.div2), a.spacing, a.n_atoms, _p(x), _p(y), _p(z),
                         out.ctypes.data_as(C.POINTER(C.c_float)), threads)
    return out


def cutoff_pairs(a, threads: int = 0) -> int:
    x, y, z = (np.ascontiguousarray(v, np.float64) for v in (a.x, a.y, a.z))
    return int(LIB.or_count_cutoff_pairs(_i3(a.n), _i3(a.div2), a.spacing, a.n_atoms, _p(x), _p(y), _p(z), threads))


def run_cli(gpf: str, cwd: str, v4: bool = False, ag421_param_order: bool = False, threads: int = 0) -> int:
    """Runs the fgoracle command line in cwd (writes the map files there)."""
    if not os.path.exists(CLI_PATH):
        build()
    cmd = [CLI_PATH, "-p", gpf, "-l", "oracle.glg"]
    if v4:
        cmd.append("--v4")
    if ag421_param_order:
        cmd.append("--ag421-param-order")
    if threads:
        cmd += ["--omp", str(threads)]
    return subprocess.run(cmd, cwd=cwd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE).returncode


def read_fld(path: str) -> OiGrid:
    """AutoDock's readfield (autodock/readfield.cc:55-140)."""
    g = OiGrid()
    rc = LIB.oi_read_fld(path.encode(), C.byref(g))
    if rc:
        raise RuntimeError(f"oi_read_fld({path}) = {rc}")
    return g


def read_mapset(g: OiGrid, paths: list) -> np.ndarray:
    """AutoDock's readmap of every file into one [z][y][x][slot] float array."""
    n = list(g.num_points1)
    out = np.zeros((n[2], n[1], n[0], len(paths)), dtype=np.float32)
    for slot, p in enumerate(paths):
        rc = LIB.oi_read_map(p.encode(), C.byref(g), slot, len(paths), _p(out, C.c_float))
        if rc:
            raise RuntimeError(f"oi_read_map({p}) = {rc}")
    return out


def trilinterp(g: OiGrid, maps: np.ndarray, elec_slot: int, desolv_slot: int, xyz: np.ndarray, charge, abs_charge,
               types, ignore=None, outside: bool = True, threads: int = 0):
    """AutoDock's trilinterp (autodock/trilinterp.cc:46-185) for every pose of
    xyz [pose][atom][3] (float32).  Returns (elec, emap) per atom (float32),
    (elec_total, emap_total) per pose (float64) and the returned Real."""
    maps = np.ascontiguousarray(maps, dtype=np.float32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    npose, natoms = xyz.shape[0], xyz.shape[1]
    q = np.ascontiguousarray(charge, dtype=np.float32)
    aq = np.ascontiguousarray(abs_charge, dtype=np.float32)
    t = np.ascontiguousarray(types, dtype=np.int32)
    ig = None if ignore is None else np.ascontiguousarray(ignore, dtype=np.int32)
    elec = np.zeros((npose, natoms), np.float32)
    emap = np.zeros((npose, natoms), np.float32)
    et = np.zeros(npose)
    mt = np.zeros(npose)
    tot = np.zeros(npose, np.float32)
    LIB.oi_trilinterp_poses(C.byref(g), _p(maps, C.c_float), maps.shape[3], elec_slot, desolv_slot, natoms, npose,
                            _p(xyz, C.c_float), _p(q, C.c_float), _p(aq, C.c_float), _p(t, C.c_int), int(outside),
                            _p(ig, C.c_int) if ig is not None else None, _p(elec, C.c_float), _p(emap, C.c_float),
                            _p(et), _p(mt), _p(tot, C.c_float), threads or (os.cpu_count() or 1))
    return elec, emap, et, mt, tot
