"""Write tests/golden/rayleigh_ff.json: the atomic form factors F(x, Z)
of pegs4/pgs4form.dat (the data egs_init_rayleigh reads,
src/egsnrc.mortran:8367-8380) for the elements of water and aluminium, on
the file's momentum-transfer grid x = sin(theta/2)/lambda (1/Angstrom).

    python tests/golden/make_rayleigh_ff.py [HEN_HOUSE]
"""
import json
import os
import sys

HH = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/HEN_HOUSE"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rayleigh_ff.json")
ELEMENTS = [1, 8, 13]


def main():
    with open(os.path.join(HH, "pegs4", "pgs4form.dat")) as f:
        v = [float(t) for t in f.read().split()]
    n = 100
    x = v[:n]
    ff = {str(z): v[n + (z - 1) * n: n + z * n] for z in ELEMENTS}
    with open(OUT, "w") as f:
        json.dump({"source": "make_rayleigh_ff.py", "x": x, "F": ff}, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
