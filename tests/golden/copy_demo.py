"""Copy the reference's example systems with stored solutions (data files of
src/Demo) into tests/golden/demo: katsura5, cyclic5, redeco5, mickey and
utbikker.  Run once where the reference is checked out:
    python tests/golden/copy_demo.py /root/reference/src/Demo
"""
import os
import shutil
import sys

NAMES = ("katsura5", "cyclic5", "redeco5", "mickey", "utbikker")

if __name__ == "__main__":
    src = sys.argv[1]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "demo")
    os.makedirs(dst, exist_ok=True)
    for name in NAMES:
        shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
        print(name)
