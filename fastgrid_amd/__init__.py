"""fastgrid_amd — FastGrid's grid-map hot path on MI355X.

The product is libfgmaps.so (C ABI in include/fgmaps.h: host front end plus
gfx950 HIP kernels) and the fastgrid4 command line built beside it.  This
package is its Python binding; see fastgrid_amd.api.
"""
from .api import (  # noqa: F401
    BATCH_NO_GRAPH, LIB, LIB_PATH, NOALLOC_SCRATCH_BYTES, PLAIN_DIEL_AUTO, PLAIN_DIEL_INPLACE, PLAIN_DIEL_TABLE,
    PLAIN_DIEL_TEXTURE, PREC_COMPENSATED, PREC_FP32, PREC_FP64, PREC_PLAIN, RUN_ALL, RUN_AFFINITY, RUN_ELEC, RUN_PHASE_TIMING,
    RUN_FLOATING, SIGNATURES, Batch, Context, FgError, Job, JobArrays, Pending, Prepared, autotune, batch_run, create_jobs,
    fetch_many,
    device_count, prepare_many, round3, synth_write, FgMapGrid, MapSet, read_fld, read_mapset, trilinterp,
)

BIN_DIR = __import__("os").path.join(__import__("os").path.dirname(__file__), "bin")
