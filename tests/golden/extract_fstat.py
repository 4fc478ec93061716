"""Copies the interaction rows of one of the reference's gold .fstat files
(Drivers/**/SelfTestData/*.fstat) into tests/golden/, dropping the per-save
header lines (#).  Run in a checkout that has the reference next to it:

    python tests/golden/extract_fstat.py \
        /root/reference/Drivers/SelfTests/Interactions/SelfTestData/ObliqueImpactSelfTest.fstat \
        tests/golden/oblique_impact_self_test.fstat

Columns (BaseInteraction::writeToFStat, BaseInteraction.cc:487-531): time,
P index, I index (walls: -index-1), contact point (3), overlap, tangential
overlap, normal force, |tangential force|, normal (3), tangential unit vector
(3); for a wall the normal and tangential vectors are printed negated.
"""
import sys


def main(src, dst):
    rows = [ln for ln in open(src) if ln.strip() and not ln.startswith("#")]
    with open(dst, "w") as f:
        f.write(f"# interaction rows of {src.split('Drivers/', 1)[-1]}\n")
        f.writelines(rows)
    print(f"{len(rows)} rows -> {dst}")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
