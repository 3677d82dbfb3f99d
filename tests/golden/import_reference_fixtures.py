"""Copies the reference's own test fixtures (scene inputs and golden images)
into tests/golden/reference/ so that the tests can run where the reference
checkout is absent.

These are data files of the reference's test suite (test/*.muon scenes and
test/testdata/*.png goldens): inputs and expected outputs, not source code.
The 3.7 MB dragon scene is left out; the dragon test is statistical and
covered by the other MAE-mode scenes.

Usage: python tests/golden/import_reference_fixtures.py [/path/to/reference]
"""
import os
import shutil
import sys

SKIP = {"dragon.muon", "dragon_golden.png", "dragon_truth.png"}


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "test")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference")
    os.makedirs(dst, exist_ok=True)
    n = 0
    for sub in ("", "testdata"):
        d = os.path.join(src, sub)
        for name in sorted(os.listdir(d)):
            if name in SKIP or not (name.endswith(".muon") or name.endswith(".png")):
                continue
            shutil.copyfile(os.path.join(d, name), os.path.join(dst, name))
            n += 1
    print(f"copied {n} fixtures into {dst}")


if __name__ == "__main__":
    main()
