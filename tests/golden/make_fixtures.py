"""Copies the reference's golden data (Data-and-Results/) into tests/golden/.

The reference ships seven uint16 grids (Base-Problems/*.bin) and the meshes its
GUI wrote for them under four post-process settings (Base-Results/*/*.obj).
They are data, not source; the tests read these copies because the reference
tree is not present where the GPU tests run.  Re-run only to refresh them:
    python tests/golden/make_fixtures.py /root/reference
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    src = os.path.join(ref, "Data-and-Results")
    for sub in ("Base-Problems", os.path.join("Base-Results", "0_No-Fixes"), os.path.join("Base-Results", "1_Grid-Extension-Fix"),
                os.path.join("Base-Results", "2_One-Iteration"), os.path.join("Base-Results", "3_Fully-Fixed")):
        dst = os.path.join(HERE, "data_and_results", sub)
        os.makedirs(dst, exist_ok=True)
        for f in sorted(os.listdir(os.path.join(src, sub))):
            shutil.copyfile(os.path.join(src, sub, f), os.path.join(dst, f))
    shutil.copyfile(os.path.join(src, "README.md"), os.path.join(HERE, "data_and_results", "README.md"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
