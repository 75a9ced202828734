"""fg_trilinterp_noalloc: trilinterp whose device staging lives in the
library's fixed per-GPU array instead of a hipMalloc, against fg_trilinterp
through the C ABI, bit for bit, on calls in a row (each smaller than the one
before, so a stale tail of the array would show), on a second context of the
same GPU, and with a size past the array's capacity refused before any GPU
work."""
import numpy as np
import pytest

from tests.test_gpu_interp import pgp, same  # noqa: F401  (module fixture)

pytestmark = pytest.mark.gpu


def _both(ctx, ms, gold, xyz, ign, per_atom=True):
    import fastgrid_amd as fg
    args = (ctx, ms, xyz, gold["types"], gold["charge"], gold["abs_charge"], 4, 5)
    want = fg.trilinterp(*args, ignore=ign, per_atom=per_atom)
    got = fg.trilinterp(*args, ignore=ign, per_atom=per_atom, noalloc=True)
    for k in ("elec", "emap", "elec_total", "emap_total", "total"):
        if want[k] is None:
            assert got[k] is None
        else:
            assert same(got[k], want[k]), k
    return got


def test_noalloc_equals_trilinterp(ctx, pgp):  # noqa: F811
    import fastgrid_amd as fg
    gold, g, og, maps = pgp
    ms = fg.MapSet(ctx, g, maps)
    rng = np.random.default_rng(5)
    big = np.concatenate([gold["a_xyz"]] * 40)  # 9,600 poses
    big = big + rng.uniform(-2, 2, big.shape).astype(np.float32)
    print(f"noalloc: {big.shape[0]} poses, then {gold['a_xyz'].shape[0]}, then {gold['b_xyz'].shape[0]}", flush=True)
    _both(ctx, ms, gold, big, gold["a_ignore"])
    _both(ctx, ms, gold, gold["a_xyz"], gold["a_ignore"])
    r = _both(ctx, ms, gold, gold["b_xyz"], gold["b_ignore"])
    assert same(r["total"], gold["b_total"])  # the reference's own vectors
    _both(ctx, ms, gold, gold["b_xyz"][:7], None, per_atom=False)
    # a second context on the same GPU shares the array
    ctx2 = fg.Context(0)
    ms2 = fg.MapSet(ctx2, g, maps)
    _both(ctx2, ms2, gold, gold["a_xyz"], gold["a_ignore"])
    ms2.close()
    ctx2.close()
    # staging past the array: refused with a clear error, nothing launched
    A = 300
    npose = fg.NOALLOC_SCRATCH_BYTES // (A * 20) + 1
    xyz = np.zeros((npose, A, 3), np.float32)
    types = np.zeros(A, np.int32)
    q = np.zeros(A, np.float32)
    with pytest.raises(fg.FgError) as ei:
        fg.trilinterp(ctx, ms, xyz, types, q, q, 4, 5, noalloc=True)
    assert "split the call" in str(ei.value)
    _both(ctx, ms, gold, gold["b_xyz"], gold["b_ignore"])
    ms.close()
