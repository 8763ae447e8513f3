"""Write the reference's own example outputs as small JSON fixtures.

Reads the example output files that ship with the reference
(examples/output/*.output under /root/reference) and keeps only numbers:
the S-matrices printed at IPRINT=11 (format of source_code/sproc.f:242,
ES21.13), the propagation ranges and step counts, and the state-to-state
cross sections.  Run from the repository root:

    python tests/golden/extract_reference_fixtures.py /root/reference

The JSON files it writes are committed; the tests read only those.
"""
import json
import os
import re
import sys

NUM = r"[-+]?\d+\.\d+E[-+]\d+"


def parse_smatrices(text):
    """Split an output at its 'model system' banners and collect, per
    banner, the S-matrices (RE, IM; row, col 1-based) and propagation lines."""
    sections = re.split(r"\n\s*=+ (model system: .+?) =+\n", text)
    out = {}
    # sections = [pre, label1, body1, label2, body2, ...]; a label repeats
    # at the end of its own run, so merge bodies by label
    for i in range(1, len(sections), 2):
        label = sections[i].strip()
        body = sections[i + 1]
        rec = out.setdefault(label, {"props": [], "smat": []})
        for m in re.finditer(r"(\w+)\. .*?PROPAGATED FROM\s+(\S+)\s+TO\s+(\S+)\s+IN\s+(\d+)\s+STEPS", body):
            rec["props"].append([m.group(1), float(m.group(2)), float(m.group(3)), int(m.group(4))])
        for blk in re.split(r"OPEN CHANNEL\s+WVEC", body)[1:]:
            rows = re.findall(r"^\s+(\d+)\s+(\d+)\s+(%s)\s+(%s)\s+(%s)\s+(%s)\s*$" % (NUM, NUM, NUM, NUM), blk, re.M)
            if not rows:
                continue
            nop = max(int(r[0]) for r in rows)
            sr = [[0.0] * nop for _ in range(nop)]
            si = [[0.0] * nop for _ in range(nop)]
            for r in rows:
                a, b = int(r[0]) - 1, int(r[1]) - 1
                sr[a][b] = float(r[4])
                si[a][b] = float(r[5])
            head = blk.split("ROW  COL")[0]
            chans = []
            for m in re.finditer(r"^\s+(\d+)\s+(\d\.\d+E[-+]\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s", head, re.M):
                chans.append({"open": int(m.group(1)), "wvec_per_ang": float(m.group(2)),
                              "channel": int(m.group(3)), "L": int(m.group(4)), "level": int(m.group(5))})
            rec["smat"].append({"nopen": nop, "channels": chans, "sr": sr, "si": si})
    return out


def parse_sigma(text):
    """State-to-state integral cross sections table ('F I SIG(F,I)')."""
    rows = re.findall(r"^\s+(\d+\.\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\S+)\s*$", text, re.M)
    return [{"energy": float(r[0]), "jtotl": int(r[1]), "jstep": int(r[2]), "jtotu": int(r[3]),
             "f": int(r[4]), "i": int(r[5]), "sigma": float(r[6])} for r in rows]


def main(ref):
    outdir = os.path.dirname(os.path.abspath(__file__))
    ex = os.path.join(ref, "examples", "output")
    with open(os.path.join(ex, "molscat-all_propagators.output")) as f:
        allp = parse_smatrices(f.read())
    keep = {k: v for k, v in allp.items() if "IOS" not in k}
    with open(os.path.join(outdir, "ref_all_propagators_smat.json"), "w") as f:
        json.dump({"source": "examples/output/molscat-all_propagators.output",
                   "deck": "examples/input/molscat-all_propagators.input",
                   "sections": keep}, f, indent=1)
    with open(os.path.join(ex, "molscat-basic1.output")) as f:
        sig = parse_sigma(f.read())
    with open(os.path.join(outdir, "ref_basic1_sigma.json"), "w") as f:
        json.dump({"source": "examples/output/molscat-basic1.output",
                   "deck": "examples/input/molscat-basic1.input",
                   "note": "printed values already multiplied by JSTEP=10",
                   "sigma": sig}, f, indent=1)
    print("wrote", len(keep), "S-matrix sections and", len(sig), "sigma rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
