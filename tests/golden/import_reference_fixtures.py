"""Copies the reference's own test fixtures (scene inputs and golden images)
into tests/golden/reference/ so that the tests can run where the reference
checkout is absent.

These are data files of the reference's test suite (test/*.muon scenes and
test/testdata/*.png goldens): inputs and expected outputs, not source code.
The 3.7 MB dragon scene is stored gzip-compressed (dragon.muon.gz, 1.3 MB).

Usage: python tests/golden/import_reference_fixtures.py [/path/to/reference]
"""
import gzip
import os
import shutil
import sys

COMPRESS = {"dragon.muon"}


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "test")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference")
    os.makedirs(dst, exist_ok=True)
    n = 0
    for sub in ("", "testdata"):
        d = os.path.join(src, sub)
        for name in sorted(os.listdir(d)):
            if not (name.endswith(".muon") or name.endswith(".png")):
                continue
            if name in COMPRESS:
                with open(os.path.join(d, name), "rb") as f, \
                        gzip.GzipFile(os.path.join(dst, name + ".gz"), "wb", compresslevel=9, mtime=0) as g:
                    g.write(f.read())
            else:
                shutil.copyfile(os.path.join(d, name), os.path.join(dst, name))
            n += 1
    print(f"copied {n} fixtures into {dst}")


if __name__ == "__main__":
    main()
