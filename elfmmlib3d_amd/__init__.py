"""MI355X (gfx950) elastostatic triangle layer-potential FMM.

Python mirror of the reference's user-callable subroutines
(src/elfmm3dtria.f:71, :162, :1453; src/elfmm3dtria_iter_1.f:160;
src/elhfmm3dtria_iter.f:46; src/elh3dtria.f:53), calling the C-ABI of
``libelfmm_gfx950.so`` (declared in include/elfmm_gfx950.h). Arrays use the
shapes of mesh.py: triangle (n,3,3), trinorm/source/sigma (n,3), target (m,3);
outputs ptfrc (n,3), strain (n,3,3). The library is the only compute path:
if it is not built, every call raises.
"""
from __future__ import annotations

from ._lib import (  # noqa: F401
    MeshBatch,
    el3dtriadirecttarg,
    elh3dtriadirecttarg,
    elhfmm3dtriatargiter,
    half_parts,
    elfmm3dtria,
    el3dpartdirecttarg,
    elh3dpartdirecttarg,
    elfmm3dpart,
    elfmm3dparttarg,
    elfmm3dparttarg_dev,
    elfmm3dtriatarg,
    elfmm3dtriatarg_dev,
    elfmm3dtriatargiter,
    elfmm3dtriatargiter_dev,
    elhfmm3dtriatargiter_dev,
    expansion_order,
    iter_release,
    triatargiter_multi_dev,
    last_stats,
    last_timings,
    library,
    library_path,
)
from ._build import build  # noqa: F401

__all__ = ["elfmm3dpart", "elfmm3dparttarg", "el3dpartdirecttarg", "elh3dpartdirecttarg", "elfmm3dparttarg_dev",
           "elfmm3dtria", "elfmm3dtriatarg", "el3dtriadirecttarg", "elfmm3dtriatarg_dev",
           "elfmm3dtriatargiter", "elhfmm3dtriatargiter", "elh3dtriadirecttarg", "half_parts",
           "iter_release",
           "last_stats", "last_timings", "library", "library_path", "build", "elfmm3dtriatargiter_dev", "elhfmm3dtriatargiter_dev", "expansion_order", "triatargiter_multi_dev",
           "MeshBatch"]
