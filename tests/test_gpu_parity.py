"""GPU parity: libfgmaps on an MI355X against the reference goldens and the
FP64 oracle.

Tolerances: affinity and desolvation maps are computed in the reference's
FP64 arithmetic and must print identically (rounded maps bit-exact); the
electrostatic map must stay within 1e-3 kcal/mol of the FP64 result (the
compensated form; the FP64 form is bit-exact)."""
import os
import subprocess

import numpy as np
import pytest

from tests import fixtures as fx

pytestmark = pytest.mark.gpu
E_TOL = 1e-3 + 1e-9


def fastgrid4(run, gpf, *args):
    import fastgrid_amd as fg
    exe = os.path.join(fg.BIN_DIR, "fastgrid4")
    p = subprocess.run([exe, "-p", gpf, "-l", "fg.glg", *args], cwd=run, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert "Successful Completion" in open(os.path.join(run, "fg.glg")).read()
    return p


def compare_files(run, stem, gold, exact_e=False, digests=None):
    """Values against the goldens; with digests, every map whose values are
    identical must also be byte-identical text."""
    for n, want in gold.items():
        path = os.path.join(run, f"{stem}.{n}.map")
        _, got = fx.read_map(path)
        if n == "e" and not exact_e:
            assert np.abs(got - want).max() <= 1, (n, np.abs(got - want).max())
        else:
            assert np.array_equal(got, want), (n, int((got != want).sum()))
        if digests is not None and np.array_equal(got, want):
            assert fx.value_digest(path) == digests[n], n


def test_cli_colagen2_goldens(tmp_path):
    run, gpf = fx.colagen2(str(tmp_path))
    fastgrid4(run, gpf, "--v4")
    compare_files(run, "colagen2", fx.goldens("colagen2", "goldens_autogrid400.npz"),
                  digests=fx.text_digests("colagen2", "goldens_autogrid400.npz"))
    assert open(os.path.join(run, "colagen2.maps.xyz")).read() == \
        open(os.path.join(fx.GOLDEN, "colagen2", "colagen2.maps.xyz")).read()


def test_cli_colagen2_fp64_form_is_exact(tmp_path):
    run, gpf = fx.colagen2(str(tmp_path))
    fastgrid4(run, gpf, "--v4", "--precision=fp64")
    compare_files(run, "colagen2", fx.goldens("colagen2", "goldens_autogrid400.npz"), exact_e=True)


def test_cli_pgp_fastgrid_and_autogrid_order(tmp_path):
    run, gpf = fx.pgp(str(tmp_path))
    fast = fx.goldens("pgp", "goldens_fastgrid_v4.npz")
    fastgrid4(run, gpf, "--v4")
    compare_files(run, "1pgp_rec", fast)
    delta = fx.goldens("pgp", "goldens_autogrid421_minus_fastgrid.npz")
    fastgrid4(run, gpf, "--v4", "--ag421-param-order")
    for n in ("e", "d"):
        _, got = fx.read_map(os.path.join(run, f"1pgp_rec.{n}.map"))
        assert np.abs(got - (fast[n] + delta[n])).max() <= (1 if n == "e" else 0), n


@pytest.mark.parametrize("name", fx.VARIANTS)
def test_cli_gpf_variant_goldens(tmp_path, name):
    run, gpf, names, args = fx.variant(str(tmp_path), name)
    fastgrid4(run, gpf, *args)
    compare_files(run, "colagen2", fx.goldens("variants", f"{name}.npz"),
                  digests=fx.text_digests("variants", f"{name}.npz"))
    assert fx.fld_lines(os.path.join(run, "colagen2.maps.fld")) == fx.variant_fld(name)


@pytest.mark.parametrize("name", ["micro", "S"])
def test_cli_synthetic_goldens(tmp_path, name):
    run, gpf, names = fx.synth(str(tmp_path), name)
    fastgrid4(run, gpf)
    compare_files(run, name, fx.goldens("synth", f"{name}_goldens_fastgrid.npz"),
                  digests=fx.text_digests("synth", f"{name}_goldens_fastgrid.npz"))


def plain_errors(out, om, od, oe):
    """Per map: max |d| / max(1, |ref|) and the fraction of points where it
    exceeds 1e-3, for the FP32 plain form against the FP64 oracle."""
    res = {}
    for name, got, ref in [(m, out[m], om[m]) for m in range(len(om))] + [("e", out[len(om)], oe),
                                                                           ("d", out[len(om) + 1], od)]:
        rel = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
        res[name] = (float(rel.max()), float((rel > 1e-3).mean()))
    return res


def _kernel_check(ctx, a, comp_tol=3e-4, plain=True):
    import fastgrid_amd as fg
    from oracle import oracle as orc
    job = fg.Job(ctx, a)
    oe = orc.elec_map(a)
    om, od = orc.affinity_maps(a)
    M = a.n_maps
    res = {}
    # the plain form is FP32 throughout: isolated points differ where an FP32
    # distance lands in the neighbouring table bin; wrong indexing or a
    # dropped term shows at most points
    if plain:
        job.run(fg.RUN_ALL, fg.PREC_PLAIN)
        pe = plain_errors(job.fetch(), om, od, oe)
        print("plain form errors (max rel, fraction > 1e-3):", pe)
        for name, (mx, frac) in pe.items():
            assert frac < 0.01 and mx < 0.5, (name, mx, frac)
        res[fg.PREC_PLAIN] = pe
    for prec in (fg.PREC_FP64, fg.PREC_COMPENSATED, fg.PREC_FP32):
        job.run(fg.RUN_ALL, prec)
        out = job.fetch()
        for m in range(M):
            assert np.array_equal(fx.milli(out[m]), fx.milli(om[m])), (prec, m)
            assert np.allclose(out[m], om[m], rtol=1e-12, atol=1e-9), (prec, m)
        assert np.array_equal(fx.milli(out[M + 1]), fx.milli(od))
        err = np.abs(out[M] - oe).max()
        res[prec] = err
        if prec == fg.PREC_FP64:
            assert err == 0.0
        elif prec == fg.PREC_COMPENSATED:
            assert err < comp_tol, err
            assert np.abs(fx.milli(out[M]) - fx.milli(oe)).max() <= 1
    assert job.cutoff_pairs() == orc.cutoff_pairs(a)
    job.close()
    return res


@pytest.mark.parametrize("name", ["micro", "S"])
def test_kernels_vs_oracle_synthetic(ctx, tmp_path, name):
    run, gpf, _ = fx.synth(str(tmp_path), name)
    _kernel_check(ctx, fx.load_arrays(run, gpf))


def test_kernels_vs_oracle_medium(ctx, tmp_path):
    import fastgrid_amd as fg
    fg.synth_write(str(tmp_path), "M", 2, 8000, "sphere", 40.0, (81, 81, 81), "A C HD N NA OA SA S")
    _kernel_check(ctx, fx.load_arrays(str(tmp_path), "M.gpf"))


def _edit_gpf(src_dir, gpf, out_dir, edits, extra=()):
    text = open(os.path.join(src_dir, gpf)).read()
    for a, b in edits:
        assert a in text, a
        text = text.replace(a, b)
    text += "".join(l + "\n" for l in extra)
    open(os.path.join(out_dir, gpf), "w").write(text)


@pytest.mark.parametrize("variant", ["constant_dielectric", "disorder_h", "covalent_qasp", "odd_grid", "one_point"])
def test_kernels_vs_oracle_gpf_variants(ctx, tmp_path, variant):
    run, gpf = fx.colagen2(str(tmp_path))
    d = os.path.dirname(run)
    edits = {
        "constant_dielectric": [("dielectric -0.1465", "dielectric 40"), ("npts 40 40 40", "npts 20 20 20")],
        "disorder_h": [("npts 40 40 40", "npts 24 24 24")],
        "covalent_qasp": [("npts 40 40 40", "npts 16 18 20"), ("spacing 1 ", "spacing 0.375 "),
                          ("ligand_types  C HD OA", "ligand_types  C HD OA Z"),
                          ("map colagen2.OA.map", "map colagen2.OA.map\nmap colagen2.Z.map\ncovalentmap 1.7 1000.0 64.5 -16.0 37.2")],
        "odd_grid": [("npts 40 40 40", "npts 21 34 9"), ("spacing 1 ", "spacing 0.4 ")],
        "one_point": [("npts 40 40 40", "npts 0 0 0")],
    }[variant]
    extra = {"disorder_h": ["disorder_h"], "covalent_qasp": ["qasp 0.01234"]}.get(variant, [])
    _edit_gpf(d, "colagen2.gpf", d, edits, extra)
    a = fx.load_arrays(run, gpf, v4=True)
    _kernel_check(ctx, a)


def test_kernels_vs_oracle_long_receptor(ctx, tmp_path):
    """A 330 A-long receptor around a small central grid: distances run from 0
    to over 163 A, so the dielectric table no longer covers them (the clamped
    kernel variant), and the electrostatic tiles fall in every class: near,
    far inside the table, far possibly past its end, and wholly past its last
    bin (81.9 A)."""
    import fastgrid_amd as fg
    fg.synth_write(str(tmp_path), "long", 9, 20000, "box", (330.0, 30.0, 30.0), (33, 33, 33), "A C HD")
    a = fx.load_arrays(str(tmp_path), "long.gpf")
    # atoms reach past 163.84 A (16,384 table bins) from the grid at the centre
    assert np.abs(a.x).max() > 164.0
    _kernel_check(ctx, a)


def test_floating_grid_and_covalent_map_one_shots(ctx, tmp_path):
    """fg_floating_grid and fg_covalent_map against the oracle: the floating
    grid bit for bit (FP64 minimum, one sqrt, one cast to float), on a grid
    around the receptor and on one farther than 8 A from every atom (the
    full-scan path); the covalent map to FP64 rounding of exp."""
    from oracle import oracle as orc
    run, gpf, names, args = fx.variant(str(tmp_path), "covalent_qasp")
    a = fx.load_arrays(run, gpf, v4=True)
    assert a.cov_on and a.covalent is None
    assert np.array_equal(ctx.floating_grid(a), orc.floating_grid(a))
    assert np.allclose(ctx.covalent_map(a), orc.covalent_map(a), rtol=1e-13, atol=1e-12)
    import copy
    far = copy.copy(a)
    far.div2 = (a.div2[0] - 400, a.div2[1], a.div2[2])  # the grid moved 150 A along x
    fm = ctx.floating_grid(far)
    assert fm.min() > 8.0
    assert np.array_equal(fm, orc.floating_grid(far))


def test_atom_on_grid_point(ctx, tmp_path):
    """r = 0 pairs: min(1/r, 2) and bin 0 (FG/electrostatics/Cpu.cpp:70-75),
    rSq clamped to 1e-24 in the cutoff loop (FG/CalculateGridmaps.cpp:284-285)."""
    lines = []
    pos = [(0.0, 0.0, 0.0), (0.375, 0.0, 0.0), (1.2, 0.9, -0.3), (0.0, 0.96, 0.0), (-1.1, 0.2, 0.4)]
    types = ["OA", "C", "N", "HD", "SA"]
    q = [-0.4, 0.2, -0.1, 0.3, 0.0]
    for i, ((x, y, z), t, c) in enumerate(zip(pos, types, q)):
        lines.append(f"ATOM  {i + 1:5d} X{i:<3d} RES A   1    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00    {c:6.3f} {t:<2s}\n")
    open(os.path.join(tmp_path, "r.pdbqt"), "w").write("".join(lines))
    open(os.path.join(tmp_path, "r.gpf"), "w").write(
        "npts 16 16 16\ngridfld r.maps.fld\nspacing 0.375\nreceptor_types A C HD N OA SA\nligand_types A C HD N OA\n"
        "receptor r.pdbqt\ngridcenter 0.0 0.0 0.0\nsmooth 0.5\nmap r.A.map\nmap r.C.map\nmap r.HD.map\nmap r.N.map\n"
        "map r.OA.map\nelecmap r.e.map\ndsolvmap r.d.map\ndielectric -0.1465\n")
    a = fx.load_arrays(str(tmp_path), "r.gpf")
    _kernel_check(ctx, a)


def test_distances_on_bin_edges(ctx, tmp_path):
    """Atoms on a 0.125 A lattice around a 0.375 A grid: every axis offset of
    a (point, atom) pair is a multiple of 0.125 A, so many distances are exact
    multiples of 0.01 A (every axis-aligned offset of an even multiple, and the
    lattice's Pythagorean triples), where floor(100 r) turns on the last bit of
    the reference's 1/(1/sqrt(r^2)) (FG/CalculateGridmaps.cpp:286-289) and of
    the e-map's sqrt (FG/electrostatics/Cpu.cpp:70-75).  The cutoff kernel's
    hardware estimate of 1/r cannot decide those bins; its guard must send
    them to the exact replay, so every map matches the oracle."""
    rng = np.random.default_rng(11)
    cells = rng.choice(97 ** 3, size=180, replace=False)
    pos = np.stack([cells % 97, (cells // 97) % 97, cells // (97 * 97)], 1) * 0.125 - 6.0
    types = rng.choice(["A", "C", "HD", "N", "OA", "SA"], size=len(pos))
    q = np.round(rng.uniform(-0.3, 0.3, size=len(pos)), 3)
    lines = [f"ATOM  {i + 1:5d} X{i % 1000:<3d} RES A   1    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00    {c:6.3f} {t:<2s}\n"
             for i, ((x, y, z), t, c) in enumerate(zip(pos, types, q))]
    open(os.path.join(tmp_path, "r.pdbqt"), "w").write("".join(lines))
    open(os.path.join(tmp_path, "r.gpf"), "w").write(
        "npts 24 24 24\ngridfld r.maps.fld\nspacing 0.375\nreceptor_types A C HD N OA SA\nligand_types A C HD N OA\n"
        "receptor r.pdbqt\ngridcenter 0.0 0.0 0.0\nsmooth 0.5\nmap r.A.map\nmap r.C.map\nmap r.HD.map\nmap r.N.map\n"
        "map r.OA.map\nelecmap r.e.map\ndsolvmap r.d.map\ndielectric -0.1465\n")
    a = fx.load_arrays(str(tmp_path), "r.gpf")
    # the lattice really puts pairs on bin edges: count pairs within 8 A whose
    # 100 r is an integer to within 1e-9
    g = np.stack(np.meshgrid(*[(np.arange(25) - 12) * 0.375] * 3, indexing="ij"), -1).reshape(-1, 3)
    d = np.sqrt(((g[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    t = 100 * d[d <= 8.0]
    assert (np.abs(t - np.round(t)) < 1e-9).sum() > 2000
    # the plain comparison form (FP32 throughout) bins many of these pairs
    # either way; its loose check is for ordinary inputs
    _kernel_check(ctx, a, plain=False)


@pytest.mark.parametrize("npts,natoms", [("2 2 2", 1), ("2 60 2", 1), ("60 2 4", 3), ("4 4 60", 7)])
def test_tiny_and_thin_grids(ctx, tmp_path, npts, natoms):
    """Edges of the launch shapes: grids of 3 points along some axes (every
    brick, tile and z-group padded), and receptors of 1 to 7 atoms (a tile
    and a candidate chunk mostly empty), against the oracle."""
    rng = np.random.default_rng(natoms)
    types = ["OA", "C", "N", "HD", "SA", "A", "NA"]
    lines = []
    for i in range(natoms):
        x, y, z = rng.uniform(-3, 3, 3)
        t = types[i % len(types)]
        lines.append(f"ATOM  {i + 1:5d} X{i:<3d} RES A   1    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00    "
                     f"{rng.uniform(-0.5, 0.5):6.3f} {t:<2s}\n")
    open(os.path.join(tmp_path, "t.pdbqt"), "w").write("".join(lines))
    open(os.path.join(tmp_path, "t.gpf"), "w").write(
        f"npts {npts}\ngridfld t.maps.fld\nspacing 0.375\nreceptor_types A C HD N NA OA SA\n"
        "ligand_types A C HD N OA\nreceptor t.pdbqt\ngridcenter 0.5 -0.25 0.0\nsmooth 0.5\nmap t.A.map\n"
        "map t.C.map\nmap t.HD.map\nmap t.N.map\nmap t.OA.map\nelecmap t.e.map\ndsolvmap t.d.map\n"
        "dielectric -0.1465\n")
    a = fx.load_arrays(str(tmp_path), "t.gpf")
    _kernel_check(ctx, a)


def test_far_grid_has_no_cutoff_pairs(ctx, tmp_path):
    """A grid more than 8 A from every atom: affinity maps are zero, the
    electrostatic map is all far field."""
    run, gpf = fx.colagen2(str(tmp_path))
    d = os.path.dirname(run)
    _edit_gpf(d, "colagen2.gpf", d, [("gridcenter 64.0 -16.0 37.0", "gridcenter 300.0 300.0 300.0"),
                                      ("npts 40 40 40", "npts 10 10 10")])
    a = fx.load_arrays(run, gpf, v4=True)
    import fastgrid_amd as fg
    job = fg.Job(ctx, a)
    assert job.cutoff_pairs() == 0
    res = _kernel_check(ctx, a)
    assert res[fg.PREC_COMPENSATED] < 1e-6


def test_sign_symmetry_at_full_size(ctx, tmp_path):
    """Size-independent property at the 50k-atom / 127^3 size: negating every
    charge negates the electrostatic map bit for bit (every pair term changes
    sign exactly) and leaves the affinity and desolvation maps unchanged."""
    import copy
    import fastgrid_amd as fg
    fg.synth_write(str(tmp_path), "L", 3, 50000, "box", (70.0, 70.0, 50.0), (127, 127, 127), "A C HD N NA OA SA S")
    a = fx.load_arrays(str(tmp_path), "L.gpf")
    b = copy.copy(a)
    b.w = -a.w
    b.charge = -a.charge
    ja, jb = fg.Job(ctx, a), fg.Job(ctx, b)
    for prec in (fg.PREC_COMPENSATED, fg.PREC_FP32, fg.PREC_PLAIN):
        ja.run(fg.RUN_ALL, prec)
        oa = ja.fetch()
        jb.run(fg.RUN_ALL, prec)
        ob = jb.fetch()
        M = a.n_maps
        assert np.array_equal(oa[M], -ob[M]), prec
        for m in list(range(M)) + [M + 1]:
            assert np.array_equal(oa[m], ob[m]), (prec, m)
    ja.close()
    jb.close()


def _plain_e_reference(a, kind):
    """FP64 restatement of the plain e-kernel's sum for one dielectric
    variant (FG/electrostatics/cuda_internal/Kernels.inl:73-119): the binned
    table, the texel of normalized point sampling, or the analytic
    calculateDistDepDielInv (FG/autogrid.h:106-112)."""
    import fastgrid_amd as fg
    nx, ny, nz = a.n
    s = a.spacing
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    gx = ((ix - a.div2[0]) * s).transpose(2, 1, 0).ravel()
    gy = ((iy - a.div2[1]) * s).transpose(2, 1, 0).ravel()
    gz = ((iz - a.div2[2]) * s).transpose(2, 1, 0).ravel()
    e = np.zeros(a.P)
    eps = np.asarray(a.eps)
    for i in range(a.n_atoms):
        r = np.sqrt((gx - a.x[i]) ** 2 + (gy - a.y[i]) ** 2 + (gz - a.z[i]) ** 2)
        with np.errstate(divide="ignore"):
            d = np.minimum(1.0 / r, 2.0)
        if kind == fg.PLAIN_DIEL_TABLE:
            d = d * eps[np.minimum((100 * r).astype(np.int64), 8191)]
        elif kind == fg.PLAIN_DIEL_TEXTURE:
            d = d * eps[np.minimum((100 * r / 8191 * 8192).astype(np.int64), 8191)]
        else:
            E = np.exp(-0.3153767175 * r)
            d = d * (332 + 332 * 7.7839 * E) / (78.4 - 66.57180475 * E)
        e += a.w[i] * d
    return e


@pytest.mark.parametrize("kind", ["table", "inplace", "texture"])
def test_plain_dielectric_variants(ctx, tmp_path, kind):
    """The plain form's three dielectric variants (the reference's
    --cuda-ddd=g/i/t kernels) against an FP64 restatement of each: FP32
    arithmetic, so isolated points differ where an FP32 distance lands in the
    neighbouring table bin; the analytic variant has no bins.  And the
    reference's automatic choice: in place up to 160 padded points per axis,
    texture beyond."""
    import fastgrid_amd as fg
    run, gpf, _ = fx.synth(str(tmp_path), "micro")
    a = fx.load_arrays(run, gpf)
    k = {"table": fg.PLAIN_DIEL_TABLE, "inplace": fg.PLAIN_DIEL_INPLACE, "texture": fg.PLAIN_DIEL_TEXTURE}[kind]
    job = fg.Job(ctx, a)
    assert job.plain_dielectric == fg.PLAIN_DIEL_TABLE
    job.plain_dielectric = fg.PLAIN_DIEL_AUTO
    assert job.plain_dielectric == fg.PLAIN_DIEL_INPLACE  # 21^3 pads to 32 x 24 x 24
    job.plain_dielectric = k
    assert job.plain_dielectric == k
    job.run(fg.RUN_ELEC, fg.PREC_PLAIN)
    e = job.fetch()[a.n_maps]
    ref = _plain_e_reference(a, k)
    err = np.abs(e - ref) / np.maximum(1.0, np.abs(ref))
    print(f"plain {kind}: max rel {err.max():.2e}, fraction > 1e-4: {(err > 1e-4).mean():.2e}")
    assert (err > 1e-4).mean() < (1e-3 if kind != "inplace" else 1e-9)
    assert err.max() < (0.05 if kind != "inplace" else 1e-4)
    job.close()
    import copy
    wide = copy.copy(a)
    wide.n, wide.div2 = (161, 5, 3), (80, 2, 1)
    wj = fg.Job(ctx, wide)
    wj.plain_dielectric = fg.PLAIN_DIEL_AUTO
    assert wj.plain_dielectric == fg.PLAIN_DIEL_TEXTURE  # 161 pads to 176
    wj.close()
