"""Writes bca_skindepth.json: the default PRESTA-I skin depth that mscati
prints (egsnrc.mortran:2797-2817, "minimum ECUT found", "default BLCMIN is",
"this corresponds to ... elastic MFPs") in the BEAMnrc example logs shipped
with the reference, together with the media of each run (the names given
after "MEDIUM ... (left justify):").  Run from the repository root with the
reference tree at /root/reference; the JSON is the committed fixture."""
import json
import os
import re

EXAMPLES = "/root/reference/HEN_HOUSE/omega/beamnrc/BEAMnrc_examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bca_skindepth.json")


def parse(path):
    text = open(path, errors="replace").read()
    media = []
    for m in re.finditer(r"MEDIUM[^\n]*\(left justify\):\s+(\S+)", text):
        if m.group(1) not in media and m.group(1) != "VACUUM":
            media.append(m.group(1))
    ecut = float(re.search(r"minimum ECUT found:\s+(\S+)", text).group(1))
    blcmin = float(re.search(r"default BLCMIN is:\s+(\S+)", text).group(1))
    mfps = float(re.search(r"this corresponds to\s+(\S+)\s+elastic MFPs", text).group(1))
    return {"media": media, "ecut": ecut, "blcmin": blcmin, "elastic_mfps": mfps}


def main():
    cases = {}
    for name in sorted(os.listdir(EXAMPLES)):
        log = os.path.join(EXAMPLES, name, name + ".egslog")
        if os.path.exists(log) and "BLCMIN" in open(log, errors="replace").read():
            cases[name] = parse(log)
    json.dump({"source": "HEN_HOUSE/omega/beamnrc/BEAMnrc_examples/*/*.egslog",
               "cases": cases}, open(OUT, "w"), indent=1)
    print(json.dumps(cases, indent=1))


if __name__ == "__main__":
    main()
