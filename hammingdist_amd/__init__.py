"""hammingdist_amd — all-pairs Hamming distances of aligned sequences on MI355X.

The public API is the reference ``hammingdist`` module's
(python/hammingdist.cc:28-133): ``from_fasta``, ``from_fasta_large``,
``from_fasta_to_lower_triangular``, ``fasta_reference_distances``,
``fasta_sequence_indices``, ``distance``, ``from_stringlist``, ``from_csv``,
``from_lower_triangular(_large)``, ``cuda_gpu_available`` and the ``DataSet``
/ ``DataSetLarge`` result types.  Every distance is computed by the HIP engine
(``libhd_engine.so``); importing fails loudly when the native libraries have
not been built.
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

from ._runtime import ensure_single_runtime as _ensure_single_runtime

_ensure_single_runtime()

try:
    from . import hammingdist as _native
except ImportError as exc:  # pragma: no cover - exercised only when unbuilt
    raise ImportError(
        "hammingdist_amd native libraries are not built "
        f"({exc}); run `python -m hammingdist_amd.build`"
    ) from exc

DataSet = _native.DataSet
DataSetLarge = _native.DataSetLarge
from_stringlist = _native.from_stringlist
from_csv = _native.from_csv
from_fasta = _native.from_fasta
from_fasta_large = _native.from_fasta_large
from_fasta_to_lower_triangular = _native.from_fasta_to_lower_triangular
from_lower_triangular = _native.from_lower_triangular
from_lower_triangular_large = _native.from_lower_triangular_large
distance = _native.distance
fasta_reference_distances = _native.fasta_reference_distances
fasta_sequence_indices = _native.fasta_sequence_indices
cuda_gpu_available = _native.cuda_gpu_available

ENGINE_LIBRARY = _os.path.join(_HERE, "libhd_engine.so")

__all__ = [
    "DataSet",
    "DataSetLarge",
    "from_stringlist",
    "from_csv",
    "from_fasta",
    "from_fasta_large",
    "from_fasta_to_lower_triangular",
    "from_lower_triangular",
    "from_lower_triangular_large",
    "distance",
    "fasta_reference_distances",
    "fasta_sequence_indices",
    "cuda_gpu_available",
]
