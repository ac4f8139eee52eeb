"""``python -m aoclda.faiss [--no-print-patch] [-m] name args...``: apply the faiss patch,
then run a script or module (python_interface/python_package/aoclda/faiss/__main__.py:32-75
of the reference)."""
import argparse
import runpy
import sys

from .patch_faiss import faiss_patch


def main():
    ap = argparse.ArgumentParser(description="AOCL-DA (MI355X) extension for faiss")
    ap.add_argument("-m", action="store_true", dest="is_module")
    ap.add_argument("name", help="the Python script or module to run")
    ap.add_argument("args", nargs=argparse.REMAINDER, help="its arguments")
    grp = ap.add_mutually_exclusive_group()
    grp.add_argument("--print-patch", dest="print_patched", action="store_true")
    grp.add_argument("--no-print-patch", dest="print_patched", action="store_false")
    ap.set_defaults(print_patched=True)
    a = ap.parse_args()
    faiss_patch(print_patched=a.print_patched)
    sys.argv = [a.name] + a.args
    if a.is_module:
        runpy.run_module(a.name, run_name="__main__")
    else:
        runpy.run_path(a.name, run_name="__main__")


if __name__ == "__main__":
    sys.exit(main())
