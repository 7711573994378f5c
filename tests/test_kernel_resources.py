"""The compiler's resource report for the library's kernels (`make resources`
in le2d_amd/csrc: VGPRs, spills, scratch, LDS and waves per SIMD from the
kernel-resource-usage remarks of the device compile; no GPU needed).

The two-step kernels carry the time of every step path.  They keep all their
values in registers (no spills, no scratch memory) at the occupancy their
launch bounds are chosen for.  Scratch in a march kernel means a row window
that the compiler could not keep in registers (a conditional copy of whole
nodes through a selected address did that to the interface-tile kernel)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "le2d_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROW = re.compile(r"^\| `([^`]+)` \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \| (\d+) \|$")


@pytest.fixture(scope="module")
def table():
    if not (os.path.exists(HIPCC) and shutil.which("make")):
        pytest.skip("hipcc or make not available")
    out = subprocess.run(["make", "-s", "-j5", "resources"], cwd=CSRC, capture_output=True,
                         text=True, timeout=900)
    assert out.returncode == 0, out.stderr[-3000:]
    rows = {}
    for line in out.stdout.splitlines():
        m = ROW.match(line)
        if m:
            v = [int(x) for x in m.groups()[1:]]
            rows[m.group(1)] = dict(zip(("vgpr", "agpr", "sgpr", "spill", "scratch", "lds", "waves"),
                                        v))
    assert rows, out.stdout[-2000:]
    return rows


def _kernels(table, prefix):
    ks = {k: v for k, v in table.items() if k.startswith(prefix + "<")}
    assert ks, prefix
    return ks


@pytest.mark.parametrize("prefix", ["k_step2_xy", "k_step2_split", "k_step2_mat_homog",
                                    "k_step2_mat_homog_split", "k_step_batch_pairs",
                                    "k_step_batch_pairs_split", "k_step_batch",
                                    "k_step_batch_tiles", "k_sweep_x", "k_sweep_y"])
def test_step_kernels_keep_everything_in_registers(table, prefix):
    for name, r in _kernels(table, prefix).items():
        assert r["spill"] == 0 and r["scratch"] == 0, (name, r)


def test_interface_tile_kernel_has_no_scratch_window(table):
    """k_step2_mat_mixed's two five-row windows stay in registers in both
    types at every window width."""
    for name, r in _kernels(table, "k_step2_mat_mixed").items():
        assert r["scratch"] == 0 and r["spill"] == 0, (name, r)


def test_occupancy_of_the_two_step_kernels(table):
    # the two-wavefront forms and the interface kernel: 3 waves per SIMD in
    # FP64 (at most 168 VGPRs), so an interface wave fits beside three
    # single-material waves as soon as one of them ends
    for name in ("k_step2_split<double, false>", "k_step2_mat_homog_split<double>",
                 "k_step_batch_pairs_split<double>", "k_step2_mat_mixed<double, 64>",
                 "k_step2_mat_mixed<double, 128>"):
        assert table[name]["waves"] >= 3 and table[name]["vgpr"] <= 168, (name, table[name])
    # the one-wavefront forms hold both stages' march state: 2 waves
    for name, r in _kernels(table, "k_step2_xy").items():
        assert r["waves"] >= 2, (name, r)
    # FP32 has register headroom
    for name in ("k_step2_split<float, true>", "k_step_batch_pairs_split<float>"):
        assert table[name]["waves"] >= 5, (name, table[name])


def test_lds_within_the_cu(table):
    """Every kernel's LDS fits several workgroups in a CU's 160 KB."""
    for name, r in table.items():
        assert r["lds"] <= 40 * 1024, (name, r)
