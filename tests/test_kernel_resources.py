"""Register, spill, scratch and call guard for the Cholesky / solve kernels,
read from the built libblast_hip.so (no GPU needed).

The one-wavefront n <= 128 kernel (the headline posv path) must use no
scratch, spill nothing (VGPRs or SGPRs) and make no calls; the workgroup
Cholesky, LU and mixed-precision kernels must use no scratch and spill no
VGPRs, and those that were rewritten to inline the diagonal factorisation
must make no calls.  The mixed n <= 128 kernel and the Riccati kernel still
call the 16x16 diagonal factorisation (inlined, both exceed their register
budgets); the Riccati kernel's called block solve keeps 12 bytes of stack.
"""

import os

import pytest

from tests import kernel_meta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "blast_amd", "libblast_hip.so")

# family (substring of the mangled name) -> limits
RULES = {
    "posv_ll_kernel": dict(scratch=0, vgpr_spill=0, sgpr_spill=0, calls=False),
    "chol_team_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "chol_owner_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "wg_solve_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "lu_wg_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "posv_uneven_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "posv_backward_uneven_kernel": dict(scratch=0, vgpr_spill=0, calls=False),
    "getrf_panel_kernel": dict(scratch=0, vgpr_spill=0, sgpr_spill=0, calls=False),
    "getrf_left_swaps_kernel": dict(scratch=0, vgpr_spill=0, sgpr_spill=0, calls=False),
    "sposv_ir_kernel": dict(scratch=0, vgpr_spill=0, calls=True),
    # scratch: the call frame of the noinline triangular solve (no VGPR spill)
    "riccati_persistent_kernel": dict(scratch=48, vgpr_spill=0, calls=True),
}


@pytest.fixture(scope="module")
def objects():
    if not os.path.exists(LIB):
        pytest.skip("libblast_hip.so is not built")
    if not kernel_meta.tools_available():
        pytest.skip("ROCm llvm-objcopy / llvm-readelf not available")
    return kernel_meta.code_objects(LIB)


def _family(name):
    for fam in RULES:
        if fam in name:
            return fam
    return None


def test_every_guarded_family_is_present(objects):
    seen = {_family(k) for co in objects for k in co["kernels"]}
    assert set(RULES) <= seen, set(RULES) - seen


def test_no_scratch_no_spills_no_calls(objects):
    bad = []
    for co in objects:
        called = set(co["functions"]) - set(co["kernels"])
        for name, meta in co["kernels"].items():
            fam = _family(name)
            if fam is None:
                continue
            rule = RULES[fam]
            if meta.get("private_segment_fixed_size", 0) > rule["scratch"]:
                bad.append((name, "scratch", meta["private_segment_fixed_size"]))
            if meta.get("vgpr_spill_count", 0) > rule["vgpr_spill"]:
                bad.append((name, "vgpr_spill", meta["vgpr_spill_count"]))
            if "sgpr_spill" in rule and meta.get("sgpr_spill_count", 0) > rule["sgpr_spill"]:
                bad.append((name, "sgpr_spill", meta["sgpr_spill_count"]))
            if not rule["calls"] and called:
                bad.append((name, "calls", sorted(called)))
    assert not bad, "\n".join(map(str, bad))


def test_metadata_reader_sees_registers(objects):
    # the reader itself: every kernel reports its register counts
    metas = [m for co in objects for m in co["kernels"].values()]
    assert len(metas) > 100
    assert all("vgpr_count" in m and "sgpr_count" in m for m in metas)
