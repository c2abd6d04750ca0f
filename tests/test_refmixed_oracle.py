"""CPU tests pinning the oracle's restatement of the reference's stage-level
mixed build (oracle/mhd_refmixed.hpp, after src/mhd/mhd2d_glm_gpu_mixed.cu).

  * its FP32 stage functions (f32ConsToPrim9, f32PrimToCons9, f32FastSpeed,
    f32ReconPrim9) equal, bitwise on finite states, the reference's own FP32
    header functions (oracle/_ref, built from the reference sources), which
    they restate with fmaxf floors; the reconstruction equals the FP32
    oracle's, itself pinned by the FP32 whole-run goldens;
  * the NaN case the fmaxf floor changes (rho = NaN recovers the floor, not
    NaN);
  * a whole run: the form lies between FP64 and FP32 (its FP32 stages move
    it off FP64 by far less than the all-FP32 build), with the same steps
    and HLLD fallback counts as FP64 on a short run;
  * regions: with no region in float it is the FP64 oracle bit for bit, and
    with only the rhs (or only the cfl) region in float its whole-run
    distance from FP64 at 192^2 is the reference's one-region truncation
    scan's, per field, to 2%.
"""
import os

import numpy as np
import pytest

import pyoracle
import states

HAVE_REF = os.path.exists(os.path.join(pyoracle.REF_DIR, "libmhdref_fp32.so"))
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (reference sources absent)")


def bits32(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def finite_rows(*arrays):
    ok = np.ones(arrays[0].shape[0], dtype=bool)
    for a in arrays:
        ok &= np.all(np.isfinite(a), axis=1) if a.ndim == 2 else np.isfinite(a)
    return ok


@needs_ref
def test_fp32_stage_functions_match_reference_headers(oracle):
    ref = pyoracle.RefHeaders("fp32")
    U = states.cell_states(n=4096).astype(np.float32)
    W = oracle.refmixed_f32("cons_to_prim", U)
    Wr = ref.cons_to_prim(U)
    ok = finite_rows(U)
    assert ok.sum() > 3000
    assert np.array_equal(bits32(W[ok]), bits32(Wr[ok]))
    Uo = oracle.refmixed_f32("prim_to_cons", Wr[ok])
    assert np.array_equal(bits32(Uo), bits32(ref.prim_to_cons(Wr[ok])))
    for d in (0, 1):
        s = oracle.refmixed_f32("fast_speed", Wr[ok], direction=d)
        assert np.array_equal(bits32(s), bits32(ref.fast_speed(Wr[ok], d)))


def test_fp32_reconstruction_matches_fp32_oracle(oracle):
    rng = np.random.default_rng(5)
    P0 = oracle.cons_to_prim(states.cell_states(n=2048).astype(np.float32), dtype=np.float32)
    ok = finite_rows(P0)
    P0 = P0[ok]
    Pm = (P0 * (1 + 0.2 * rng.standard_normal(P0.shape))).astype(np.float32)
    Pp = (P0 * (1 + 0.2 * rng.standard_normal(P0.shape))).astype(np.float32)
    Pl, Pr = oracle.refmixed_f32("recon", Pm, P0, Pp)
    Ql, Qr = oracle.reconstruct(Pm, P0, Pp, dtype=np.float32)
    assert np.array_equal(bits32(Pl), bits32(Ql)) and np.array_equal(bits32(Pr), bits32(Qr))
    # some rows fall back to first order (negative reconstructed pressure)
    assert np.any(np.all(Pl == P0, axis=1) & np.any(Pm != P0, axis=1))


def test_fmaxf_floor_on_nan_density(oracle):
    U = np.zeros((1, 9), dtype=np.float32)
    U[0] = [np.nan, 0.1, 0.2, 0.0, 2.0, 0.3, 0.1, 0.0, 0.0]
    W = oracle.refmixed_f32("cons_to_prim", U)
    assert W[0, 0] == np.float32(1e-12)      # fmaxf(NaN, floor) = floor
    Wc = oracle.cons_to_prim(U, dtype=np.float32)
    assert np.isnan(Wc[0, 0])                # std::max semantics keep NaN


def test_whole_run_between_fp64_and_fp32(oracle):
    nx, ny, steps = 48, 40, 60
    u64 = oracle.init("ot", nx, ny, dtype=np.float64)
    a64, i64 = oracle.advance(u64, steps)
    arm, irm = oracle.refmixed_advance(u64, steps)
    a32, i32 = oracle.advance(oracle.init("ot", nx, ny, dtype=np.float32), steps)
    assert irm["code"] == 0 and irm["steps"] == steps == i64["steps"]
    assert np.array_equal(irm["counters"], i64["counters"])
    rel = lambda a: np.sqrt(np.mean((a - a64) ** 2)) / np.sqrt(np.mean(a64 ** 2))
    e_rm, e_32 = rel(arm), rel(a32.astype(np.float64))
    assert 0 < e_rm < 0.3 * e_32, (e_rm, e_32)
    assert e_rm < 1e-6
    assert irm["t"] == pytest.approx(i64["t"], rel=1e-6)


def test_region_mask_none_is_fp64_oracle(oracle):
    """With no region in float the restatement is the FP64 oracle's step,
    bit for bit (state, time and HLLD counters); each single region moves
    the result."""
    u = oracle.init("ot", 48, 40)
    d, idd = oracle.advance(u, 30)
    c, ic = oracle.refmixed_advance(u, 30, regions=())
    assert np.array_equal(c.view(np.uint64), d.view(np.uint64))
    assert ic["t"] == idd["t"] and list(ic["counters"]) == list(idd["counters"])
    a, _ = oracle.refmixed_advance(u, 30)
    for r in oracle.REGIONS:
        e, _ = oracle.refmixed_advance(u, 30, regions=(r,))
        assert not np.array_equal(e, d), r
        assert not np.array_equal(e, a), r


@pytest.fixture(scope="module")
def scan192(oracle):
    """The reference's 192^2 comparison tables and the FP64 whole run."""
    import json

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "cmp_ot192.json")))
    u = oracle.init("ot", g["n"], g["n"])
    ref, iref = oracle.refmixed_advance(u, 100000, t_end=g["t_end"], regions=())
    assert iref["steps"] == g["steps_fp64"]
    return g, u, ref


@pytest.mark.parametrize("region", ["rhs", "cfl"])
def test_region_runs_match_reference_scan_192(oracle, scan192, region):
    """Orszag-Tang 192^2 to t=0.5 with one region in float, against FP64:
    the relative L2 per field equals the reference's one-region truncation
    scan (logs/mhd/raptor_scan/ot2d_cmp_{rhs,cfl}.txt) to within 2%.  These
    regions take FP64 inputs (the face fluxes; the recovered primitives), so
    casting them and truncating each operation agree; the rhs region is the
    FP32 narrowing of the face fluxes before they are differenced
    (mhd2d_glm_gpu_mixed.cu:355-375).  The recovery and recon regions cast
    the state (f32LoadPrim) where the scan truncates operations on FP64
    inputs, and differ from it by factors of 0.3-1.9 and 3-9; glm by 0.2;
    those regions are pinned bitwise against the reference's FP32 header
    functions instead."""
    g, u, ref = scan192
    a, ia = oracle.refmixed_advance(u, 100000, t_end=g["t_end"], regions=(region,))
    assert ia["code"] == 0 and ia["steps"] == g["steps_fp64"]
    for k, var in (("rho", 0), ("vx", 1), ("vy", 2), ("E", 4), ("Bx", 5), ("By", 6)):
        if var in (1, 2):
            x, y = a[..., var] / a[..., 0], ref[..., var] / ref[..., 0]
        else:
            x, y = a[..., var], ref[..., var]
        rel = np.sqrt(np.mean((x - y) ** 2)) / np.sqrt(np.mean(y ** 2))
        logged = g["regions"][region]["rows"][k]["relL2"]
        assert rel == pytest.approx(logged, rel=0.02), (region, k, rel, logged)
