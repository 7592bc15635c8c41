"""Write tests/golden/relax_k_shells.json: the EADL K-shell transition data
of lead and tantalum from the EGSnrc data file relax.data (4-byte records:
element count, per element a pointer to its block; per shell the EADL
shell type, the transition count and the binding energy, then per
transition the final state, counted from 0, and the probability; a final
state below 64 is the new vacancy of a radiative transition, otherwise
64 * s1 + s2 for the two vacancies of a non-radiative one).  The fixture
holds the raw probabilities, as the file gives them, for the tests of the
relaxation cascade.  Run where an EGSnrc HEN_HOUSE tree is available."""
import json
import os
import sys

import numpy as np

HEN_HOUSE = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/HEN_HOUSE"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "relax_k_shells.json")


def element(words, reals, Z):
    pos = words[Z] - 1
    nshell = int(words[pos])
    p = pos + 1
    shells = []
    for _ in range(nshell):
        typ, ntran, be = int(words[p]), int(words[p + 1]), float(reals[p + 2])
        p += 3
        st = [int(x) for x in words[p:p + 2 * ntran:2]]
        pr = [float(x) for x in reals[p + 1:p + 2 * ntran:2]]
        p += 2 * ntran
        shells.append({"type": typ, "be": be, "states": st, "probs": pr})
    return shells


def main():
    d = np.fromfile(os.path.join(HEN_HOUSE, "data", "relax.data"), dtype="<i4")
    f = d.view("<f4")
    out = {}
    for Z in (73, 82):
        sh = element(d, f, Z)
        out[str(Z)] = {"K": sh[0], "shell_be_eadl": [s["be"] for s in sh]}
    json.dump(out, open(OUT, "w"), indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
