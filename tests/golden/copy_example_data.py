"""Copies the data fixtures of the example-galaxy tests from the reference
checkout: ProFit's example cut-outs (inst/extdata/{KiDS,SDSS}/G*{fitim,sigma,
segim,psfim}.fits, 10 galaxies each) and its table of initial parameters
(data/ExampleInit.tab).  These are the reference's own data files, copied
byte for byte; run from the repository root:

    python tests/golden/copy_example_data.py /root/reference
"""
import glob
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    out = os.path.join(HERE, "examples")
    for src in ("KiDS", "SDSS"):
        os.makedirs(os.path.join(out, src), exist_ok=True)
        for kind in ("fitim", "sigma", "segim", "psfim"):
            for f in sorted(glob.glob(os.path.join(ref, "inst", "extdata", src, f"G*{kind}.fits"))):
                shutil.copyfile(f, os.path.join(out, src, os.path.basename(f)))
    shutil.copyfile(os.path.join(ref, "data", "ExampleInit.tab"), os.path.join(out, "ExampleInit.tab"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
