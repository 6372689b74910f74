"""Copies the reference's CP2K task fixtures (data files) into tests/golden/.

Run once in a container that has the reference checkout; the copies are
committed so that no test reads the reference at run time.
"""
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/lib"
DST = os.path.dirname(os.path.abspath(__file__))
NAMES = ["collocate_ortho_density", "collocate_ortho_tau", "collocate_general_density",
         "collocate_general_tau"]

if __name__ == "__main__":
    for n in NAMES:
        shutil.copyfile(os.path.join(SRC, n + ".task"), os.path.join(DST, n + ".task"))
        print("copied", n)
