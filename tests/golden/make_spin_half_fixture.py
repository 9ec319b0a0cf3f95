"""Extract the spin_half_example fixture from the reference's example files.

Reads (as data, nothing is executed) examples/user_defined/spin_half_example/
{vertex_basis.in, Sx.in, Sy.in, Sz.in, example_input, README.md} and writes
spin_half_example.json: the basis charges, the operator triplets, the
couplings of example_input, the growth-step rows (Index, Energy,
Energy/vertices, Entropy) of example_input_Energies.dat quoted in
README.md:31-40 (growth steps 0-4 = 2..10 sites) and its finite-sweep rows
(iterations 5-8: the midpoint rows of the two sweeps of `-F 2`, README.md:29,
with the Truncation and Fidelity columns).
"""
import json
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "examples/user_defined/spin_half_example"


def triplets(path):
    out = []
    for line in path.read_text().splitlines()[1:]:
        p = line.split()
        if len(p) < 3:
            continue
        v = p[2]
        if v.startswith("("):
            re_, im_ = v.strip("()").split(",")
            out.append([int(p[0]), int(p[1]), float(re_), float(im_)])
        else:
            out.append([int(p[0]), int(p[1]), float(v), 0.0])
    return out


def main():
    basis = [list(map(int, l.split())) for l in (REF / "vertex_basis.in").read_text().splitlines()[1:] if l.strip()]
    ops = {n: triplets(REF / f"{n}.in") for n in ("Sx", "Sy", "Sz")}
    couplings_line = (REF / "example_input").read_text().splitlines()[2]
    couplings = []
    for tok in couplings_line.split():
        name, rest = tok.split(",", 1)
        m = re.match(r"\(([^)]*)\):(.*)", rest)
        opnames = m.group(1).split(",")
        couplings.append({"name": name, "ops": opnames, "value": float(m.group(2))})
    readme = (REF / "README.md").read_text().splitlines()
    start = next(i for i, l in enumerate(readme) if l.startswith("# Index, Energy, Energy/vertices"))
    rows = []
    for l in readme[start + 1:]:
        if l.startswith("~~~~"):
            break
        p = l.split()
        rows.append([int(p[0])] + [float(v) for v in p[1:6]])
    growth = [r[:4] for r in rows[:5]]   # iterations 0-4 are the growth steps (README.md:42)
    finite = rows[5:]                    # iterations 5-8: finite-size sweeps
    out = {"source": "examples/user_defined/spin_half_example (README.md:29-42, *.in, example_input)",
           "basis": basis, "operators": ops, "couplings": couplings,
           "chi": 50, "growth_steps": growth, "sweeps": 2, "finite_sweep_rows": finite}
    Path(__file__).with_name("spin_half_example.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
