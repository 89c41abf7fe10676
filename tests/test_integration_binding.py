"""The reference-side binding of the repository's integration guide
(src/distance_hip.cc, the file a maintainer adds in place of
src/distance_cuda.cu) compiles against the
reference's own headers and this repository's C ABI, links with
libhd_engine.so, and runs: a driver calls its distance_cuda_have_device()
and, where a GPU is visible, distances_cuda_8bit / distances_cuda_16bit /
distance_cuda on rows packed as the reference's from_string packs them.

Skipped where the reference tree is absent (it is read, never copied: the
headers are used in place)."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INCLUDE = "/root/reference/include"

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "hamming/distance_cuda.hh"

int main() {
  using hamming::GeneBlock;
  const bool dev = hamming::distance_cuda_have_device();
  std::printf("device %d\n", dev ? 1 : 0);
  if (!dev) return 0;
  // A C G T - as the reference packs them: two sites per byte, low nibble
  // first, 0xFF padding to 8 bytes.  Rows "AC", "AG", "TT".
  std::vector<std::vector<GeneBlock>> data = {
      {0x21, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF},
      {0x41, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF},
      {0x88, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF}};
  const auto d8 = hamming::distances_cuda_8bit(data, 255);
  const auto d16 = hamming::distances_cuda_16bit(data, 1);
  const int d = hamming::distance_cuda(data[0], data[2], 100);
  std::printf("%u %u %u | %u %u %u | %d\n", d8[0], d8[1], d8[2], d16[0], d16[1], d16[2], d);
  return 0;
}
"""


def binding_source() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```cpp\n(// src/distance_hip\.cc\n.*?)```", text, flags=re.S)
    assert m, "the integration guide holds no src/distance_hip.cc block"
    return m.group(1)


def torch_include():
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return None
    inc = os.path.join(list(spec.submodule_search_locations)[0], "include")
    return inc if os.path.exists(os.path.join(inc, "fmt", "core.h")) else None


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference tree absent")
def test_binding_compiles_links_and_runs(tmp_path):
    fmt_inc = torch_include()
    if fmt_inc is None:
        pytest.skip("no fmt headers (torch's bundled copy) for the reference's hamming_utils.hh")
    src = tmp_path / "distance_hip.cc"
    src.write_text(binding_source())
    drv = tmp_path / "driver.cc"
    drv.write_text(DRIVER)
    exe = tmp_path / "driver"
    engine_dir = os.path.join(ROOT, "hammingdist_amd")
    # torch's include directory also holds unrelated headers; only fmt is
    # taken from it, through a directory of its own
    fmt_only = tmp_path / "fmtinc"
    fmt_only.mkdir()
    os.symlink(os.path.join(fmt_inc, "fmt"), fmt_only / "fmt")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-DFMT_HEADER_ONLY", f"-I{REF_INCLUDE}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{fmt_only}", str(src), str(drv),
           "-o", str(exe), f"-L{engine_dir}", "-lhd_engine", f"-Wl,-rpath,{engine_dir}"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr[-4000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr[-2000:]
    lines = run.stdout.splitlines()
    assert lines[0] in ("device 0", "device 1")
    if lines[0] == "device 1":
        # AC/AG: 1, AC/TT: 2, AG/TT: 2; capped at 1; distance(AC, TT) = 2
        assert lines[1] == "1 2 2 | 1 1 1 | 2"


if __name__ == "__main__":
    sys.exit(pytest.main([__file__, "-q"]))
