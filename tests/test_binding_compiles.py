"""The reference-side binding -- the CRTP ConvectiveTermSolver<GpuFiniteVolume-
Scheme> over alpaca_host.h that a maintainer adds to the reference, kept as
the one `cpp` block defining GpuFiniteVolumeScheme in the integration notes
at the repository root -- compiles against the reference's own headers (convective_term_solver.h:67-81 and what
it includes), for 2D and 3D, with the reference's class template explicitly
instantiated so every call path to UpdateImplementation is checked.  Needs
the reference tree; syntax-only, output under a temporary directory."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFERENCE_SRC = "/root/reference/src"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    found = [b for b in blocks if "class GpuFiniteVolumeScheme" in b]
    assert len(found) == 1, "the integration notes must show the GpuFiniteVolumeScheme binding once"
    return found[0]


@pytest.mark.skipif(not os.path.isdir(REFERENCE_SRC), reason="reference tree not present")
@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC),
                    reason="hipcc not available")
@pytest.mark.parametrize("dim", [2, 3])
def test_integration_binding_compiles(tmp_path, dim):
    src = tmp_path / "gpu_finite_volume_scheme.cpp"
    src.write_text(_binding_source() +
                   "\ntemplate class ConvectiveTermSolver<GpuFiniteVolumeScheme>;\n")
    cmd = [HIPCC, "-std=c++20", "-fsyntax-only", f"-DDIMENSION={dim}", "-DTEST_VIRTUAL=",
           "-DPERFORMANCE", f"-I{REFERENCE_SRC}", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(ROOT, 'alpaca_amd', 'csrc', 'host')}",
           "-Wno-unused-command-line-argument", str(src)]
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=tmp_path, timeout=600)
    assert res.returncode == 0, res.stdout + res.stderr
