"""Copies gold output files of the reference's self tests verbatim into
tests/golden/ (data: the expected .ene / .data output of a driver).  Run in a
checkout that has the reference next to it:

    python tests/golden/copy_output_gold.py /root/reference

Files: the energy series of FreeFallSelfTest and
TwoParticleElasticCollisionSelfTest (DPMBase::writeEneTimeStep,
DPMBase.cc:2100-2126; these gold files predate the `collisions` column) and
the energies and XBalls data of ForceSelfTest (cases drawn from the
problem's random number generator), the XBalls data of ExtremeOverlapUnitTest (DPMBase::outputXBallsData,
DPMBase.cc:2356-2417), and the final restart files of FreeFallSelfTest and
TwoParticleElasticCollisionSelfTest (DPMBase::write, DPMBase.cc:3550-3690),
and all four output files of InsertionBoundarySelfTest, and the torque
series of RollingFrictionSelfTest and TorsionFrictionSelfTest.
"""
import os
import shutil
import sys

FILES = {
    "Drivers/SelfTests/Interactions/SelfTestData/FreeFallSelfTest.ene": "free_fall_self_test.ene",
    "Drivers/SelfTests/Interactions/SelfTestData/TwoParticleElasticCollisionSelfTest.ene":
        "two_particle_elastic_collision_self_test.ene",
    "Drivers/UnitTests/SelfTestData/ExtremeOverlapUnitTest.data": "extreme_overlap_unit_test.data",
    "Drivers/SelfTests/Interactions/SelfTestData/FreeFallSelfTest.restart": "free_fall_self_test.restart",
    "Drivers/SelfTests/Interactions/SelfTestData/TwoParticleElasticCollisionSelfTest.restart":
        "two_particle_elastic_collision_self_test.restart",
}
# ForceSelfTest: the energies of all ten cases, the XBalls data of a wall
# case (2) and a pair case (7)
for k in range(10):
    FILES[f"Drivers/SelfTests/Interactions/SelfTestData/ForceSelfTest{k}.ene"] = f"force_self_test_{k}.ene"
for k in (2, 7):
    FILES[f"Drivers/SelfTests/Interactions/SelfTestData/ForceSelfTest{k}.data"] = f"force_self_test_{k}.data"
# InsertionBoundarySelfTest: cube, cylinder and sphere insertion boundaries
# (InsertionBoundary::checkBoundaryBeforeTimeStep, InsertionBoundary.cc:184-310)
# DeletionBoundarySelfTest, CubeDeletionBoundarySelfTest: a falling sphere
# deleted by a plane / box (DeletionBoundary::checkBoundaryAfterParticlesMove)
for name, dst in (("DeletionBoundarySelfTest", "deletion_boundary_self_test"),
                  ("CubeDeletionBoundarySelfTest", "cube_deletion_boundary_self_test")):
    FILES[f"Drivers/SelfTests/Boundaries/SelfTestData/{name}.data"] = f"{dst}.data"
# roughBottomSelfTest: a chute on a MULTILAYER bottom (ChuteBottom::makeRoughBottom)
FILES["Drivers/ChuteDemos/SelfTestData/roughBottomMultiLayer.restart"] = "rough_bottom_multi_layer.restart"
# RollingFrictionSelfTest, TorsionFrictionSelfTest: time, torque and normal
# force of a fixed sphere pressed on a spinning one, after every step
# (FrictionInteraction rolling and torsion springs)
for name in ("RollingFrictionSelfTest", "TorsionFrictionSelfTest"):
    FILES[f"Drivers/SelfTests/Interactions/SelfTestData/{name}.torque"] = \
        name.replace("FrictionSelfTest", "_friction_self_test").lower() + ".torque"
for ext in ("data", "ene", "fstat", "restart"):
    FILES[f"Drivers/SelfTests/Boundaries/SelfTestData/InsertionBoundarySelfTest.{ext}"] = \
        f"insertion_boundary_self_test.{ext}"


def main(reference):
    here = os.path.dirname(os.path.abspath(__file__))
    for src, dst in FILES.items():
        shutil.copyfile(os.path.join(reference, src), os.path.join(here, dst))
        print(f"{src} -> tests/golden/{dst}")


if __name__ == "__main__":
    main(sys.argv[1])
