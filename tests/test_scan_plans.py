"""Host-side bounds check of every MFMA scan launch plan (CPU only, no GPU).

The matrix-core scan (scan_mfma.hip) derives every global and LDS address from
its launch arguments: the tile copies (per-lane source offsets, LDS
destinations), the query fragments, the shared thresholds, the candidate
slots and their append segments, the per-split list counts and, for the
residual operand, the block lists, centroids, f32 queries and per-list terms.
The library checks each launch against the sizes of the buffers it passes
(scan_check_launch) and refuses one that would leave them.  This test drives
the same check over the plans knn_mfma makes for every shipped shape -- bf16,
f32, residual bf16 and residual E4M3 operands, every K step bucket (narrowest and widest
dimension of each), four-workgroup, two-wave, wide and one-wave launches --
against sizes from 300 to 10^7 rows, 1 to 10^5 queries and KP 10..128, with
the buffers knn_mfma allocates; then plants one mistake at a time in every plan
(a buffer one element short, one split too many, an append segment one entry
too long, a tile not whole, the other workgroup shape) and requires every plan
to be flagged.
"""
import ctypes

import pytest

PERTURBATIONS = {
    1: "candidate slots one entry short",
    2: "list counts one entry short",
    3: "shared thresholds one entry short",
    4: "one split more than the buffers hold",
    5: "append segment one entry longer",
    6: "operand rows 16 bytes short",
    7: "rows per split not whole tiles",
    8: "the other workgroup shape (1 vs 4 waves)",
    9: "C inputs one entry short",
    10: "query operand / per-list terms short",
    11: "f32 queries / thresholds short",
    12: "centroids / list counts short",
    13: "E4M3 row scales / query fragments / fragment scales short",
}


@pytest.fixture(scope="module")
def check():
    from fuel_amd import _lib
    lib = _lib.load()
    f = lib.fuel_amd_internal_check_scan_plans
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int32]

    def run(perturb):
        n = ctypes.c_int64()
        buf = ctypes.create_string_buffer(512)
        bad = f(perturb, ctypes.byref(n), buf, 512)
        return bad, n.value, buf.value.decode()
    return run


def test_every_shipped_scan_plan_stays_in_bounds(check):
    bad, launches, first = check(0)
    assert launches > 50000
    assert bad == 0, first


@pytest.mark.parametrize("perturb", sorted(PERTURBATIONS))
def test_planted_plan_mistakes_are_caught(check, perturb):
    bad, launches, first = check(perturb)
    assert bad == launches, (PERTURBATIONS[perturb], bad, launches, first)
