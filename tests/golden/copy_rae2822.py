"""Copy the reference's RAE2822 example grid into tests/golden/.

The files are data the reference ships as its example input
(PHengLEI/examples/grid/rae2822_hybrid2d__4_0.fts, the partitioned 2-D hybrid
grid that cfd_para_subsonic.hypara and cfd_para_transonic.hypara name, and
rae2822_hybrid2d__4_0.wdt, the reference's own cell wall distances for it).
The tests read them with phl_read_fts / phl_read_wdt; the .wdt values are the
reference's output for its 2-D wall-distance routine on this grid.

    python tests/golden/copy_rae2822.py [REFERENCE_ROOT]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("rae2822_hybrid2d__4_0.fts", "rae2822_hybrid2d__4_0.wdt")


def main(root):
    src = os.path.join(root, "PHengLEI", "examples", "grid")
    for n in NAMES:
        shutil.copyfile(os.path.join(src, n), os.path.join(HERE, n))
        print("copied", n)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
