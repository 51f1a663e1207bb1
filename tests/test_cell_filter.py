"""The filtered cell rounding of the abstraction's fast path (dim_to_cells_fast in
gsc_device.hpp) against the exact step-by-step path it stands in for.

The reference computes a successor box's x / y edges from a coordinate summed
one increment at a time (model.hpp:40-41) and rounds them to cells with
floor / ceil and the kEps domain test (grid.hpp:181-251). The kernels instead
add the precomputed sum of the increments to the cell centre and decide the
cells only when a rigorous error bound keeps every threshold out of reach;
otherwise they fall back to the exact sum. tests/cpp/cell_filter_check.cpp
drives both paths through the library's own header on random grids, and on
coordinates placed onto cell edges and domain limits, and counts any case
where the filter decides differently from the exact path. The same program
checks the per-(plane, input) cell offsets (cell_offsets in gsc_device.hpp),
which decide the cells of every cell centre of a plane at once from the
fractional parts of the shared edge quotients."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("cell_filter") / "cell_filter_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "gpu-inner-outer_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "cell_filter_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    return str(exe)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_filter_never_disagrees_with_exact_path(checker, seed):
    out = subprocess.run([checker, "2000000", str(seed)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    # Two parts: the per-cell filter (dim_to_cells_fast) and the per-(plane,
    # input) cell offsets (cell_offsets), each against the exact path.
    for prefix in ("", "offsets "):
        m = re.search("^" + prefix + r"checked (\d+) decided (\d+) rejected (\d+) mismatches (\d+)", out.stdout,
                      re.MULTILINE)
        assert m, out.stdout + out.stderr
        checked, decided, rejected, mismatches = map(int, m.groups())
        assert mismatches == 0, out.stdout
        # Three quarters of the cases sit on a threshold by construction; both
        # paths still decide most of them, and rejections (domain limits) occur.
        assert decided > 0.6 * checked and rejected > 0
