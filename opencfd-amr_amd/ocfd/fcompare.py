"""Command-line fcompare (Util/Plotfile/fcompare.cpp):

    python -m ocfd.fcompare [-n|--norm N] [-a|--allow_diff_grids] [-r|--rel_tol R] [--abs_tol A] file1 file2

Prints the per-variable absolute and relative errors and exits 0 when the
plotfiles agree (exactly or within the tolerances), 1 otherwise."""
import argparse
import sys

from .plotfile import fcompare


def main(argv=None):
    ap = argparse.ArgumentParser(prog="fcompare")
    ap.add_argument("-n", "--norm", type=int, default=0)
    ap.add_argument("-a", "--allow_diff_grids", action="store_true")
    ap.add_argument("-r", "--rel_tol", type=float, default=0.0)
    ap.add_argument("--abs_tol", type=float, default=0.0)
    ap.add_argument("file1")
    ap.add_argument("file2")
    a = ap.parse_args(argv)
    res, report = fcompare(a.file1, a.file2, a.norm, a.rel_tol, a.abs_tol, a.allow_diff_grids)
    sys.stdout.write(report)
    return 0 if res in (1, 2) else 1


if __name__ == "__main__":
    sys.exit(main())
