"""Lanes (tebd_set_lanes): a batched half-sweep split over several streams and host
threads must give bit-identical states to the same half-sweep as one batch, because a
bond's arithmetic does not depend on which bonds share its batch."""
import numpy as np
import pytest

from tebd_amd import Context, models as M

pytestmark = pytest.mark.gpu


def evolve(mode, lanes, steps=16):
    L, d, chi = 32, 4, 128
    ctx = Context(L, d, chi, mode=mode)
    ctx.set_lanes(lanes)
    ctx.upload(*M.product_state(M.neel(L, d, 1, 2), chi))
    ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, 4.0), 0.05))
    trunc = sum(ctx.trotter_step() for _ in range(steps))
    g, l = ctx.download(chi)
    dims = ctx.bond_dims()
    st = ctx.stats()
    ctx.close()
    return g, l, dims, trunc, st


@pytest.mark.parametrize("mode,lanes", [("batched", 2), ("batched,learned", 2), ("batched,learned", 4)])
def test_lanes_are_bit_identical_to_one(mode, lanes):
    print(f"[gpu] lanes fh L=32 chi=128 {mode} lanes={lanes}", flush=True)
    g1, l1, d1, t1, s1 = evolve(mode, 1)
    g2, l2, d2, t2, s2 = evolve(mode, lanes)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(l1, l2)
    np.testing.assert_array_equal(g1, g2)
    assert t1 == t2
    assert s1["bond_updates"] == s2["bond_updates"]
    # the split really happened (bond matrices of side 4 x 128 = 512 in the middle)
    assert s1["lane_splits"] == 0 and s2["lane_splits"] > 0, (s1["lane_splits"], s2["lane_splits"])


def test_graph_replayed_sweeps_are_bit_identical_to_eager_launches():
    """Every Jacobi sweep replays a captured hipGraph (one per launch shape, many replays
    per graph across 12 steps, eager host work in between) and must leave exactly the
    state the eager launches leave."""
    print("[gpu] graphs vs eager fh L=32 chi=64", flush=True)
    out = {}
    for graphs in (False, True):
        L, d, chi = 32, 4, 64
        ctx = Context(L, d, chi, mode="batched,learned")
        ctx.set_graphs(graphs)
        ctx.upload(*M.product_state(M.neel(L, d, 1, 2), chi))
        ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, 4.0), 0.05))
        for _ in range(12):
            ctx.trotter_step()
        out[graphs] = ctx.download(chi)
        ctx.close()
    np.testing.assert_array_equal(out[False][1], out[True][1])
    np.testing.assert_array_equal(out[False][0], out[True][0])


def test_autotuned_gemm_tiles_do_not_change_results():
    """The learned-rank path's GEMMs with the autotuner on (tile variants timed on the real
    operands, fastest kept per shape class) and off (64 x 64 tiles): identical states."""
    print("[gpu] autotune on/off fh L=32 chi=64", flush=True)
    out, info = {}, None
    for tune in (False, True):
        L, d, chi = 32, 4, 64
        ctx = Context(L, d, chi, mode="batched,learned")
        ctx.set_autotune(tune)
        ctx.upload(*M.product_state(M.neel(L, d, 1, 2), chi))
        ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, 4.0), 0.05))
        for _ in range(10):
            ctx.trotter_step()
        out[tune] = ctx.download(chi)
        if tune:
            info = ctx.autotune_info()
        ctx.close()
    np.testing.assert_array_equal(out[False][1], out[True][1])
    np.testing.assert_array_equal(out[False][0], out[True][0])
    assert info["shape_classes"] > 0 and info["timing_launches"] >= 2 * info["shape_classes"]
    assert sum(info["chosen"].values()) == info["shape_classes"]


@pytest.mark.parametrize("mode,lanes", [("batched", 1), ("batched,learned", 3)])
def test_jacobi_statistics_are_consistent(mode, lanes):
    """The per-sweep histograms of tebd_get_stats add up to the totals: pairs rotated per
    sweep sum to the pairs rotated, and the solves by sweep count to the sweeps run (no
    solve needs 16 sweeps here, so the last bucket is empty); every solve ends with a
    sweep that rotates nothing."""
    print(f"[gpu] Jacobi statistics {mode} lanes={lanes}", flush=True)
    _, _, _, _, st = evolve(mode, lanes, steps=6)
    rot, solves = st["sweep_rotations"], st["solve_sweeps"]
    assert rot.sum() == st["pair_rotations"] > 0
    assert solves[-1] == 0
    assert (solves * np.arange(1, 17)).sum() == st["jacobi_sweeps"]
    # a solve of s sweeps rotated in its first s - 1: the first sweep is the busiest
    assert rot[0] == rot.max()


def test_per_bond_fp64_mode_is_bit_identical_to_batched():
    """The plain drop-in (one bond at a time) and the batched mode take the same
    arithmetic per bond, so a trajectory is bit-identical in the two (the bonds of a
    half-sweep are independent)."""
    print("[gpu] fp64 vs batched bit identity fh L=16 chi=64", flush=True)
    out = {}
    for mode in ("fp64", "batched"):
        L, d, chi = 16, 4, 64
        with Context(L, d, chi, mode=mode) as ctx:
            ctx.upload(*M.product_state(M.neel(L, d, 1, 2), chi))
            ctx.set_gates(M.construct_propagators(M.fermi_hubbard(L, 1.0, 4.0), 0.05))
            trunc = [ctx.trotter_step() for _ in range(6)]
            out[mode] = ctx.download(chi), trunc
    (g0, l0), t0 = out["fp64"]
    (g1, l1), t1 = out["batched"]
    np.testing.assert_array_equal(l0, l1)
    np.testing.assert_array_equal(g0, g1)
    assert t0 == t1
