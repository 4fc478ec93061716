"""The output files of the host API (.fstat, .ene, .data) against the gold
files of the reference's own self tests: drivers written like the
reference's are run in a scratch directory and their files compared.  On the
CPU they run on the oracle (the dem_* ABI served by oracle/dem_oracle.c); on
the GPU on the HIP engine."""
import os
import subprocess
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import output_check as oc  # noqa: E402

G = oc.GOLDEN
SETUPS = os.path.join(oc.DRIVERS, "setups")


def setup(name):
    """the generic driver with one of the set-up files"""
    return ("SetupDriver", [os.path.join(SETUPS, name + ".setup")])

# case -> (driver, arguments, [(file written, gold file, comparison)])
CASES = {
    "ObliqueImpactSelfTest": ("ObliqueImpactSelfTest", [],
                              [("ObliqueImpactSelfTest.fstat", "oblique_impact_self_test.fstat", "fstat")]),
    "MindlinSelfTest": ("MindlinSelfTest", [], [
        ("MindlinSelfTestLoading.fstat", "mindlin_self_test_loading.fstat", "fstat"),
        ("MindlinSelfTestNormalAndTangentialLoading.fstat", "mindlin_self_test_normal_and_tangential_loading.fstat",
         "fstat"),
        ("MindlinSelfTestFrameIndependence.fstat", "mindlin_self_test_frame_independence.fstat", "fstat"),
    ]),
    "TwoParticleElasticCollisionSelfTest": ("TwoParticleElasticCollisionSelfTest", [], [
        ("TwoParticleElasticCollisionSelfTest.ene", "two_particle_elastic_collision_self_test.ene", "ene"),
        ("TwoParticleElasticCollisionSelfTest.restart", "two_particle_elastic_collision_self_test.restart",
         "restart:1e-11")]),
    # 10^6 steps with four bounces: identical to 10 digits through the
    # second bounce (t <= 0.5); later the rows agree to 2e-6, because the
    # floor's orientation quaternion is one ulp below the one the gold files
    # were made with (next case)
    "FreeFallSelfTest": ("FreeFallSelfTest", [], [("FreeFallSelfTest.ene", "free_fall_self_test.ene", "ene_long")]),
    # the same two drivers with the walls oriented as the gold restart files
    # print them (sqrt(1/2) correctly rounded): every .ene row to 10 digits
    # and the final state to the restart file's 15
    "FreeFallSelfTest-gold-walls": ("FreeFallSelfTest", ["gold-wall-orientation"], [
        ("FreeFallSelfTest.ene", "free_fall_self_test.ene", "ene"),
        ("FreeFallSelfTest.restart", "free_fall_self_test.restart", "restart:1e-14")]),
    "TwoParticleElasticCollisionSelfTest-gold-walls": ("TwoParticleElasticCollisionSelfTest", ["gold-wall-orientation"], [
        ("TwoParticleElasticCollisionSelfTest.restart", "two_particle_elastic_collision_self_test.restart",
         "restart:1e-14")]),
    "ExtremeOverlapUnitTest": ("ExtremeOverlapUnitTest", [],
                               [("ExtremeOverlapUnitTest.data", "extreme_overlap_unit_test.data", "data")]),
    # ten collisions whose parameters come from the random number generator
    "ForceSelfTest": (*setup("ForceSelfTest"),
                      [(f"ForceSelfTest{k}.ene", f"force_self_test_{k}.ene", "ene") for k in range(10)] +
                      [(f"ForceSelfTest{k}.data", f"force_self_test_{k}.data", "data") for k in (2, 7)]),
    # three insertion boundaries fill an empty box for 201 steps: the RNG
    # sequence, the PSD samples and every accept/reject decision of
    # checkParticleForInteraction decide the 1942 particles of the files
    "InsertionBoundarySelfTest": ("InsertionBoundarySelfTest", [], [
        (f"InsertionBoundarySelfTest.{e}", f"insertion_boundary_self_test.{e}", k)
        for e, k in (("data", "data"), ("ene", "ene"), ("fstat", "fstat"), ("restart", "restart:1e-14"))]),
    # a falling sphere deleted by a plane and by a box
    "DeletionBoundarySelfTest": (*setup("DeletionBoundarySelfTest"),
                                 [("DeletionBoundarySelfTest.data", "deletion_boundary_self_test.data", "data")]),
    "CubeDeletionBoundarySelfTest": (*setup("CubeDeletionBoundarySelfTest"), [
        ("CubeDeletionBoundarySelfTest.data", "cube_deletion_boundary_self_test.data", "data")]),
    # a chute on a MULTILAYER bottom: ChuteBottom fills a box at random and
    # lets it settle for 2000 steps before a layer is cut from the top; the
    # same 13 base and 2 inflow particles, the inflow ones to 15 digits, the
    # base positions to 1e-3 relative (the settling run amplifies the
    # round-off of force sums taken in another order to about 2e-5)
    "roughBottomSelfTest": (*setup("roughBottomSelfTest"), [
        ("roughBottomMultiLayer.restart", "rough_bottom_multi_layer.restart", "restart:1e-3")]),
    # stop at t = 1.05, restart from the .restart file (also with numbered
    # .data files), compare with the uninterrupted run (self-checking)
    "FreeFallRestartUnitTest": ("FreeFallRestartUnitTest", [], []),
}
DRIVER_NAMES = sorted({d for d, _, _ in CASES.values()})


def check_outputs(exe, case, cwd):
    _, args, files = CASES[case]
    r = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr, file=sys.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    for written, gold, kind in files:
        ours = os.path.join(cwd, written)
        want = os.path.join(G, gold)
        if kind == "fstat":
            n, marginal = oc.compare_fstat(ours, want)
            print(f"{written}: {n} rows match, {marginal} marginal", file=sys.stderr)
            heads = oc.fstat_headers(ours)
            assert len(heads) % 3 == 0 and all(len(h) == 2 and h[1] == 1 for h in heads[0::3])
        elif kind == "ene":
            oc.compare_ene(ours, want)
        elif kind == "ene_long":
            oc.compare_ene(ours, want, exact_until=0.5, late_rtol=2e-6)
        elif kind.startswith("restart:"):
            worst = oc.compare_restart(ours, want, float(kind.split(":")[1]))
            print(f"{written}: worst {worst:.3g} of tolerance", file=sys.stderr)
        else:
            oc.compare_data(ours, want)


@pytest.fixture(scope="module")
def oracle_drivers(tmp_path_factory):
    return oc.build_on_oracle(DRIVER_NAMES, str(tmp_path_factory.mktemp("oracle_drivers")))


@pytest.mark.parametrize("case", sorted(CASES))
def test_output_on_oracle(case, oracle_drivers, tmp_path):
    check_outputs(oracle_drivers[CASES[case][0]], case, str(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_output_on_engine(case, tmp_path):
    exe = os.path.join(oc.DRIVERS, "bin", CASES[case][0])
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    check_outputs(exe, case, str(tmp_path))
