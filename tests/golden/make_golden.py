"""Copy the reference's own golden data into fixtures.

tests/os_vel2pm/result/sky.obs.txt of the reference holds 371 stars with
columns lb[2] absSDSSr SDSSr comp XYZ[3] radec[2] vcyl[3] pmlb[3], written
by its gal2other (equatorial) and vel2pm (Galactic proper motions) stages;
XYZ there is galactocentric (X = Rg - x, Y = -y).  The file is copied
verbatim as data.

Usage: python make_golden.py <reference tree>
"""
import os
import shutil
import sys


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    shutil.copyfile(os.path.join(ref, "tests", "os_vel2pm", "result", "sky.obs.txt"),
                    os.path.join(here, "os_vel2pm_sky.obs.txt"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
