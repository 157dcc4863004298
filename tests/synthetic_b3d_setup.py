#!/usr/bin/env python3
"""Stand-in for the b3d_setup helper (which links blocks_3d's sources, absent here): same command line
(key=value arguments, q4=a,b,c,d,sign), same binary output format, seeded random M matrices and h_inf.
Used by the drop-in tests through BLOCKSVIBE_B3D_SETUP."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(argv):
    kv, q4s = {}, []
    for a in argv:
        k, _, v = a.partition("=")
        if k == "q4":
            x = [int(t) for t in v.split(",")]
            q4s.append((tuple(x[:4]), x[4]))
        else:
            kv[k] = v
    from blocksvibe_amd.setup import write_synthetic
    write_synthetic(kv["out"], int(kv["precision"]), [int(t) for t in kv["two_js"].split(",")], int(kv["two_j12"]),
                    int(kv["two_j43"]), int(kv["order"]), int(kv["lambda"]), int(kv["two_j_max"]),
                    kv["radial"] == "1", q4s, seed=int(os.environ.get("SYNTH_SEED", "1")),
                    skip_m=kv.get("skip_m") == "1")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
