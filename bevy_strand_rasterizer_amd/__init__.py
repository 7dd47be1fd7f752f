"""MI355X-native strand rendering hot path (prepass/culling, froxel binning,
scans and compaction, deep-shadow maps, hair shading, strand raster and
composite) behind a C ABI that keeps the reference's slab and task layouts."""
from . import abi  # noqa: F401

__version__ = "0.1.0"
