"""GPU LBVH contact detector vs the CPU oracle: candidate pair sets, contact sets, distances and
conservative bounds must match bit for bit (compared as sorted sets)."""
import numpy as np
import pytest

from gaia_amd import scenes
from gaia_amd.detector import ContactDetector
from test_oracle_detector import crumpled, two_sheets

pytestmark = pytest.mark.gpu


def sort_rows(a, keys):
    if len(a) == 0:
        return np.arange(0)
    return np.lexsort(tuple(a[:, k] for k in reversed(keys)))


def compare(gpu, ref):
    """Asserts bitwise equality of every list (sorted canonically) and returns the sizes."""
    o = sort_rows(gpu["vf_pairs"], [0, 1])
    assert np.array_equal(gpu["vf_pairs"][o], ref["vf_pairs"]), "VF candidate pairs differ"
    assert np.array_equal(gpu["vf_pair_d"][o], ref["vf_pair_d"]), "VF distances differ"
    o = sort_rows(gpu["vf_contacts"], [0, 1, 2])
    assert np.array_equal(gpu["vf_contacts"][o], ref["vf_contacts"]), "VF contacts differ"
    assert np.array_equal(gpu["vf_contact_d"][o], ref["vf_contact_d"])
    o = sort_rows(gpu["ee_contacts"], [0, 1])
    assert np.array_equal(gpu["ee_contacts"][o], ref["ee_contacts"]), "EE contacts differ"
    assert np.array_equal(gpu["ee_mu"][o], ref["ee_mu"]) and np.array_equal(gpu["ee_d"][o], ref["ee_d"])
    return len(ref["vf_pairs"]), len(ref["vf_contacts"]), len(ref["ee_contacts"])


def run_both(gpu_ctx, oracle, meshes, r, rebuild=True):
    cd = ContactDetector(gpu_ctx, meshes, max_query_dis=r)
    D = oracle.Detector(meshes)
    e_gpu, fe_gpu = cd.topology()
    e_ref, fe_ref = D.topology()
    assert np.array_equal(e_gpu, e_ref) and np.array_equal(fe_gpu, fe_ref)
    cd.updateBVH(rebuild)
    cd.query(record=True)
    pos = cd.positions()
    ref = D.detect(pos, r, "grid", threads=8)
    return cd, D, ref, pos


@pytest.mark.parametrize("case", ["two_sheets", "signed_zero_sheets", "crumpled0", "crumpled1", "small_tori"])
def test_small_configurations_match_oracle(gpu_ctx, oracle, case):
    if case == "two_sheets":
        meshes, r = two_sheets(), 0.002
    elif case == "signed_zero_sheets":
        # axis-aligned edges (equal endpoint coordinates on two axes) on a sheet at height 0 with
        # every other vertex at -0.0: the edge leaves' orientation bits meet ties of both signs
        meshes, r = two_sheets(), 0.002
        p0 = meshes[0][0].copy()
        p0[::2, 1] = np.float32(-0.0)
        meshes[0] = (p0, meshes[0][1])
    elif case.startswith("crumpled"):
        meshes, r = crumpled(int(case[-1])), 0.002
    else:
        meshes, _ = scenes.c3_tori(seed=2, grid=2, nu=16, nv=8, spacing=0.05, major=0.02, minor=0.008)
        r = 0.004
    cd, D, ref, pos = run_both(gpu_ctx, oracle, meshes, r)
    sizes = compare(cd.recorded(), ref)
    assert sizes[0] > 0 and sizes[2] > 0
    vf, face, edge = cd.min_distances()
    assert np.array_equal(vf, ref["vf_min"]) and np.array_equal(face, ref["face_min"])
    assert np.array_equal(edge, ref["edge_min"])
    assert np.array_equal(cd.bounds(0.0005, 0.45), D.bounds(0.0005, 0.45))
    cd.destroy()


@pytest.mark.parametrize("cap", [8, 2])
def test_capped_results_are_the_lowest_ids(gpu_ctx, oracle, cap):
    """The per-vertex (per-edge) lists keep the `cap` contacts with the lowest representative
    face (other edge) ids, whatever the traversal order: equal to the oracle's full lists cut
    after sorting; cap 2 forces overflow on many primitives."""
    meshes, _ = scenes.c3_tori(seed=4, grid=2, nu=32, nv=16, spacing=0.09, major=0.04, minor=0.015)
    cd = ContactDetector(gpu_ctx, meshes, max_query_dis=0.004, vf_capacity=cap, ee_capacity=cap)
    D = oracle.Detector(meshes)
    cd.updateBVH(True)
    cd.query(record=True)
    ref = D.detect(cd.positions(), 0.004, "grid", threads=8)
    vc, vr, ec, er = cd.results()
    vfc = ref["vf_contacts"]
    overflowed = 0
    for v in np.unique(vfc[:, 0]):
        want = vfc[vfc[:, 0] == v]
        want = want[np.argsort(want[:, 3])][:cap]
        overflowed += (vfc[:, 0] == v).sum() > cap
        got = sorted((int(g["type"]), int(g["prim"]), int(g["face"])) for g in vr[v, :vc[v]])
        assert got == sorted((int(w[1]), int(w[2]), int(w[3])) for w in want), v
    assert vc.sum() == sum(min(cap, (vfc[:, 0] == v).sum()) for v in np.unique(vfc[:, 0]))
    ee = ref["ee_contacts"]
    for e in np.unique(ee[:, 0]):
        want = np.sort(ee[ee[:, 0] == e][:, 1])[:cap]
        assert sorted(er[e, :ec[e]]["edge"].tolist()) == want.tolist(), e
    c = cd.counts()
    assert c["vf_pairs"] == len(ref["vf_pairs"]) and c["ee_contacts"] == len(ee)
    if cap == 2:
        assert overflowed > 0 and c["vf_overflow"] > 0
    cd.destroy()


def test_full_tori_workload_steps_match_oracle(gpu_ctx, oracle):
    """64 tori x 16,384 triangles (1,048,576 faces, 1,572,864 edges), r = 2 mm: three steps of
    noise + drift with refits between rebuilds; every step's sets equal the oracle's."""
    meshes, drift = scenes.c3_tori(seed=0)
    cd = ContactDetector(gpu_ctx, meshes, max_query_dis=0.002)
    D = oracle.Detector(meshes)
    assert (cd.nv, cd.nf, cd.ne) == (524288, 1048576, 1572864)
    for step in range(3):
        cd.perturb(seed=7, step=step, sigma=0.001, drift=drift)
        cd.updateBVH(rebuild=(step == 0))
        cd.query(record=True)
        pos = cd.positions()
        ref = D.detect(pos, 0.002, "grid", threads=16)
        n = compare(cd.recorded(), ref)
        vf, face, edge = cd.min_distances()
        assert np.array_equal(vf, ref["vf_min"]) and np.array_equal(face, ref["face_min"])
        assert np.array_equal(edge, ref["edge_min"])
        print(f"\n[tori step {step}] VF pairs {n[0]}, VF contacts {n[1]}, EE contacts {n[2]}", flush=True)
        assert n[0] > 10000 and n[2] > 10000
    cd.destroy()


def test_refit_equals_rebuild(gpu_ctx):
    meshes, drift = scenes.c3_tori(seed=1, grid=2, nu=64, nv=32, spacing=0.25)
    a = ContactDetector(gpu_ctx, meshes, max_query_dis=0.002)
    b = ContactDetector(gpu_ctx, meshes, max_query_dis=0.002)
    for step in range(4):
        for cd in (a, b):
            cd.perturb(seed=3, step=step, sigma=0.001, drift=drift)
        a.updateBVH(rebuild=False)
        b.updateBVH(rebuild=True)
        a.query(record=True)
        b.query(record=True)
        ra, rb = a.recorded(), b.recorded()
        # canonical order of b as the reference, then compare a against it
        ob = sort_rows(rb["vf_pairs"], [0, 1])
        oc = sort_rows(rb["vf_contacts"], [0, 1, 2])
        oe = sort_rows(rb["ee_contacts"], [0, 1])
        ref = dict(vf_pairs=rb["vf_pairs"][ob], vf_pair_d=rb["vf_pair_d"][ob], vf_contacts=rb["vf_contacts"][oc],
                   vf_contact_d=rb["vf_contact_d"][oc], ee_contacts=rb["ee_contacts"][oe], ee_mu=rb["ee_mu"][oe],
                   ee_d=rb["ee_d"][oe])
        compare(ra, ref)
    a.destroy()
    b.destroy()


def test_detector_errors(gpu_ctx):
    import gaia_amd
    with pytest.raises(gaia_amd.VBDError, match="capacities"):
        ContactDetector(gpu_ctx, [scenes.torus(8, 4)], max_query_dis=0.01, vf_capacity=99)
    bad = np.array([[0, 1, 2], [0, 1, 3]], np.int32)  # edge (0, 1) used twice in the same direction
    with pytest.raises(gaia_amd.VBDError, match="non-manifold"):
        ContactDetector(gpu_ctx, [(np.random.rand(4, 3), bad)], max_query_dis=0.01)


def test_query_needs_current_trees(gpu_ctx, oracle):
    """The EE query reads the other edge's endpoints from the edge tree's leaves, so a query after
    new positions and before the tree update is refused; after the update the sets are the
    oracle's again."""
    import gaia_amd
    meshes, _ = scenes.c3_tori(seed=2, grid=2, nu=16, nv=8, spacing=0.05, major=0.02, minor=0.008)
    cd = ContactDetector(gpu_ctx, meshes, max_query_dis=0.004)
    D = oracle.Detector(meshes)
    cd.query(record=True)  # trees built at set-up from the same positions
    compare(cd.recorded(), D.detect(cd.positions(), 0.004, "grid", threads=8))
    pos0 = cd.positions()[: len(meshes[0][0])]
    cd.set_positions(0, pos0 + np.float32(0.003))
    with pytest.raises(gaia_amd.VBDError, match="update the BVH"):
        cd.query(record=True)
    cd.updateBVH(False)
    cd.query(record=True)
    compare(cd.recorded(), D.detect(cd.positions(), 0.004, "grid", threads=8))
    cd.perturb(seed=1, step=1, sigma=0.001)
    with pytest.raises(gaia_amd.VBDError, match="update the BVH"):
        cd.query(record=False)
    cd.destroy()
