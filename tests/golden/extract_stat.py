"""Copies the data rows of one of the reference's gold CG .stat files
(Drivers/SelfTests/CG/SelfTestData/*.stat) into tests/golden/, keeping the
two header lines as comments.  Run in a checkout that has the reference next
to it:

    python tests/golden/extract_stat.py \
        /root/reference/Drivers/SelfTests/CG/SelfTestData/CGBasicSelfTest.1.stat \
        tests/golden/cg_basic_self_test_gauss_z.stat

Columns (CG::write, StandardFields::writeNames): time, z, VolumeFraction,
Density, Momentum (3), MomentumFlux (6), ContactStress (9),
InteractionForce (3), ParticleSize (6).
"""
import sys


def main(src, dst):
    lines = open(src).read().splitlines()
    with open(dst, "w") as f:
        f.write(f"# data rows of {src.split('Drivers/', 1)[-1]}\n")
        for ln in lines[:2]:
            f.write("# " + ln + "\n")
        for ln in lines[2:]:
            if ln.strip():
                f.write(ln + "\n")
    print(f"{len(lines) - 2} rows -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
