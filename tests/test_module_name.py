"""`import hammingdist` works unchanged for callers written against the
reference's pybind11 module (python/hammingdist.cc:28-133): the same names,
argument names, defaults and return types.  CPU only: no engine call needs a
GPU here except cuda_gpu_available, which must answer False without one."""
import re

import numpy as np
import pytest

import hammingdist
import hammingdist_amd

# name -> (argument, default) list as the reference's py::arg lists give it
# (python/hammingdist.cc:74-130); None: no default
SIGNATURES = {
    "from_fasta": [("filename", None), ("include_x", "False"), ("remove_duplicates", "False"),
                   ("n", "0"), ("use_gpu", "False"), ("max_distance", "255")],
    "from_fasta_large": [("filename", None), ("include_x", "False"),
                         ("remove_duplicates", "False"), ("n", "0"), ("use_gpu", "False"),
                         ("max_distance", "65535")],
    "from_fasta_to_lower_triangular": [("fasta_filename", None), ("output_filename", None),
                                       ("remove_duplicates", "False"), ("n", "0"),
                                       ("use_gpu", "True"), ("max_distance", "65535")],
    "distance": [("seq0", None), ("seq1", None), ("include_x", "False")],
    "fasta_reference_distances": [("reference_sequence", None), ("fasta_file", None),
                                  ("include_x", "False")],
    "fasta_sequence_indices": [("fasta_file", None), ("n", "0")],
}


def signature_args(doc: str):
    """(argument, default) pairs of a pybind11 signature line."""
    line = doc.splitlines()[0]
    inner = line[line.index("(") + 1:line.rindex(")", 0, line.index(" -> ") if " -> " in line
                                                       else len(line))]
    out = []
    for part in re.split(r",\s*(?=[a-z_0-9]+:)", inner):
        name, _, rest = part.partition(":")
        default = rest.split("=", 1)[1].strip() if "=" in rest else None
        out.append((name.strip(), default))
    return out


def test_reference_names_import_unchanged():
    names = {"DataSet", "DataSetLarge", "from_stringlist", "from_csv", "from_fasta",
             "from_fasta_large", "from_fasta_to_lower_triangular", "from_lower_triangular",
             "from_lower_triangular_large", "distance", "fasta_reference_distances",
             "fasta_sequence_indices", "cuda_gpu_available"}
    public = {n for n in dir(hammingdist) if not n.startswith("_")}
    assert names <= public
    for n in names:
        assert getattr(hammingdist, n) is getattr(hammingdist_amd, n)
    # the native module is the reference's name
    assert hammingdist_amd._native.__name__.split(".")[-1] == "hammingdist"


@pytest.mark.parametrize("name,args", sorted(SIGNATURES.items()))
def test_reference_argument_names_and_defaults(name, args):
    doc = getattr(hammingdist, name).__doc__
    assert doc.startswith(name + "(")
    assert signature_args(doc) == args, doc.splitlines()[0]


def test_dataset_methods():
    for cls, thr in ((hammingdist.DataSet, 255), (hammingdist.DataSetLarge, 65535)):
        for m in ("dump", "dump_lower_triangular", "dump_sparse", "dump_sequence_indices",
                  "__getitem__"):
            assert hasattr(cls, m), (cls, m)
        args = signature_args(cls.dump_sparse.__doc__)
        assert args[1:] == [("filename", None), ("threshold", str(thr))]
        assert hasattr(cls, "lt_array") and hasattr(cls, "_distances")


def test_host_only_paths_without_gpu(tmp_path):
    # no GPU in this container: the device query answers False, and the
    # loaders, which compute nothing, work
    assert hammingdist.cuda_gpu_available() in (False, True)
    lt = tmp_path / "lt.txt"
    lt.write_text("1\n2,3\n")
    ds = hammingdist.from_lower_triangular(str(lt))
    assert ds.lt_array.dtype == np.uint8 and ds.lt_array.tolist() == [1, 2, 3]
    if not hammingdist.cuda_gpu_available():
        # every distance is computed by the HIP engine: no CPU fallback
        with pytest.raises(RuntimeError, match="no HIP device"):
            hammingdist.distance("ACGT", "ACGA")
