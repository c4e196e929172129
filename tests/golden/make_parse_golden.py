"""Regenerate the expected parse results of tests/golden/prm/*.prm by running
the reference's own parse_config (src/parse_config.cu:40), compiled from the
reference sources into oracle/_ref by oracle/Makefile.  Run in a container that
has the reference tree; the JSON output is committed as a fixture."""
import ctypes
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from ctbangbang_amd.structs import ReconParams  # noqa: E402


def fields(rp):
    out = {}
    for name, typ in ReconParams._fields_:
        v = getattr(rp, name)
        out[name] = v.decode() if isinstance(v, bytes) else (float(v) if isinstance(v, float) else int(v))
    return out


def main():
    for prm in sorted(glob.glob(os.path.join(HERE, "prm", "*.prm"))):
        rp = oracle.ref_parse_config(prm)
        with open(prm[:-4] + ".json", "w") as f:
            json.dump(fields(rp), f, indent=1, sort_keys=True)
        print("wrote", prm[:-4] + ".json")


if __name__ == "__main__":
    main()
