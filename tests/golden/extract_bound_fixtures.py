"""Write the reference's BOUND example results for the log-derivative
propagators as a small JSON fixture.

Reads examples/output/bound-all_propagators.output and the matching
IBDSUM summary files (bound-all_propagators.ibdsum.606 / .505) under the
reference tree and keeps only numbers: per deck (IPROPS=IPROPL=6, 5
and 7) and JTOT, the node counts at EMIN and EMAX, and for each
converged state its number, its energy (12 digits from the summary file),
the smallest matching-matrix eigenvalue printed at that energy, its index,
and the principal eigenvector components.  Run from the repository root:

    python tests/golden/extract_bound_fixtures.py /root/reference
"""
import json
import os
import re
import sys

DECKS = {6: ("model system: ITYPE=1; IPROPS=IPROPL=6", "bound-all_propagators.ibdsum.606"),
         5: ("model system: ITYPE=1; IPROPS=IPROPL=5", "bound-all_propagators.ibdsum.505"),
         7: ("model system: ITYPE=1; IPROPS=IPROPL=7", "bound-all_propagators.ibdsum.707")}


def parse_section(body):
    blocks = re.split(r"ANGULAR MOMENTUM JTOT\s+=\s+(\d+)\s+AND SYMMETRY BLOCK\s+=\s+(\d+)", body)
    out = []
    for i in range(1, len(blocks), 3):
        jtot, ib, text = int(blocks[i]), int(blocks[i + 1]), blocks[i + 2]
        below = re.findall(r"NODE COUNT IMPLIES\s+(\d+) STATES BELOW (MINIMUM|MAXIMUM) ENERGY =\s+(\S+)", text)
        rec = {"jtot": jtot, "ib": ib, "states": []}
        for n, which, e in below:
            rec["nodes_at_emin" if which == "MINIMUM" else "nodes_at_emax"] = int(n)
            rec["emin" if which == "MINIMUM" else "emax"] = float(e)
        for st in re.split(r"SEARCHING FOR BOUND STATE NUMBER", text)[1:]:
            m = re.search(r"CONVERGED ON STATE NUMBER\s+(\d+) AT\s+ENERGY =\s+(\S+)", st)
            e = re.search(r"SMALLEST EIGENVALUE OF MATCHING MATRIX IS NUMBER\s+(\d+) =\s+(\S+)", st)
            table = st.split("PRINCIPAL EIGENVECTOR COMPONENTS")[1].split("QUANTUM NUMBERS", 1)[1]
            table = re.split(r"\n\s*\n", table.lstrip("\n"), maxsplit=1)[0]
            comps = re.findall(r"^\s+(\d+)\s+(-?\d\.\d+)\s+(\d\.\d+)\s+\d+", table, re.M)
            rec["states"].append({"state": int(m.group(1)), "energy_printed": float(m.group(2)),
                                  "eig_index": int(e.group(1)), "eigmin": float(e.group(2)),
                                  "vec": [[int(c[0]), float(c[1])] for c in comps]})
        out.append(rec)
    return out


def main(ref):
    outdir = os.path.dirname(os.path.abspath(__file__))
    ex = os.path.join(ref, "examples", "output")
    with open(os.path.join(ex, "bound-all_propagators.output")) as f:
        text = f.read()
    sections = re.split(r"\n\s*=+ (model system: .+?) =+\n", text)
    bodies = {}
    for i in range(1, len(sections), 2):
        bodies.setdefault(sections[i].strip(), "")
        bodies[sections[i].strip()] += sections[i + 1]
    result = {}
    for iprop, (label, summ) in DECKS.items():
        recs = parse_section(bodies[label])
        with open(os.path.join(ex, summ)) as f:
            e12 = [float(x) for x in re.findall(r"=\s+(-?\d+\.\d+)\s+CM-1\s*$", f.read(), re.M)]
        k = 0
        for rec in recs:
            for st in rec["states"]:
                st["energy"] = e12[k]
                k += 1
        result[str(iprop)] = {"label": label, "blocks": recs}
    with open(os.path.join(outdir, "ref_bound_all_propagators.json"), "w") as f:
        json.dump({"source": "examples/output/bound-all_propagators.output (+ .ibdsum.606, .ibdsum.505, .ibdsum.707)",
                   "deck": "examples/input/bound-all_propagators.input",
                   "decks": result}, f, indent=1)
    print("wrote", sum(len(v["blocks"]) for v in result.values()), "JTOT blocks")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
