"""Copies the reference's symmetric PGM power-flow fixtures (data files only) into this folder.

Run in a checkout that has the reference tree at /root/reference; the copies are committed, so the
tests never read the reference at run time.  The fixtures originate from the Power Grid Model project
(MPL-2.0, see LICENSE.md here); the reference keeps them under tests/data/pgm/powerflow/symmetric/.
"""
import os
import shutil
import sys

SRC = "/root/reference/tests/data/pgm/powerflow/symmetric"
HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["basic-node", "node", "line", "shunt", "source", "sym_load", "sym_gen", "transmission-case",
         "transformer", "distribution-case", "three_winding_transformer"]
FILES = ["input.json", "sym_output.json", "sym_output_batch.json", "update_batch.json"]


def main() -> int:
    if not os.path.isdir(SRC):
        print(f"{SRC} not found", file=sys.stderr)
        return 1
    for case in CASES:
        os.makedirs(os.path.join(HERE, case), exist_ok=True)
        for f in FILES:
            src = os.path.join(SRC, case, f)
            if os.path.exists(src):
                shutil.copyfile(src, os.path.join(HERE, case, f))
    return 0


if __name__ == "__main__":
    sys.exit(main())
