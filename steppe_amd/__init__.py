"""steppe_amd — an MI355X-native engine for Steppe's f-statistics hot path.

Packed 2-bit genotypes -> per-population allele counts -> per-jackknife-block f2
tensor resident in HBM -> fits, on hand-written HIP kernels for gfx950 behind a
C ABI (include/steppe_hip.h).  The Python surface follows the reference's
``steppe`` package (bindings/steppe/__init__.py:444-718): ``read_f2``,
``extract_f2`` and the fit functions return the same shapes and names.
"""
from ._core import (  # noqa: F401
    F2Handle,
    PRECISION_NAMES,
    STP_PREC_FP64,
    STP_PREC_SPLIT_BF16,
    qpadm_models,
    read_f2,
    read_f2_host,
    run_extract_f2,
    run_qpadm,
    run_qpadm_search,
    upload_f2,
    write_f2_dir,
)

__version__ = "0.1.0"


def extract_f2(packed, n_snp, pop_offsets, block_id, n_block, pops=None, out="", device=0,
               ploidy="auto", precision=None, sample_ploidy=None):
    """Facade over ``_core.run_extract_f2`` (bindings/steppe/__init__.py:493-559)."""
    return run_extract_f2(packed, n_snp, pop_offsets, block_id, n_block, pops=pops, out=out,
                          device=device, ploidy=ploidy, precision=precision,
                          sample_ploidy=sample_ploidy)


def qpadm(f2, target, left, right, **kw):
    """Facade over ``_core.run_qpadm`` (bindings/steppe/__init__.py qpadm)."""
    return run_qpadm(f2, target, left, right, **kw)


def qpadm_search(f2, target, lefts, right, **kw):
    """Facade over ``_core.run_qpadm_search``."""
    return run_qpadm_search(f2, target, lefts, right, **kw)
