"""Extract the per-iteration CCSD energies stored in the reference's notebooks.

Reads notebook outputs as JSON text (nothing from the reference is imported or executed) and
writes ``trajectories.json`` next to this script.  Sources:

  16O_L2        examples/Example_CCSD.ipynb  (get_norm_ord_int(2, ref_16O_gs, -9, -9, 6.0),
                ccsd_solver(eps=1e-8, maxSteps=500, max_diis=10, mixing=0.5, sparse=True))
  16O_L4        examples/Example_CCSD2.ipynb (get_norm_ordered_ham(4, ref_16O_gs, Tkin,
                contacts(-8, -8), NNNcontact(5.5), sparse=True, NO2B=True),
                ccsd_solver(eps=1e-8, maxSteps=100, max_diis=10, delta=0, mixing=0.5))
  deuteron_*    benchmarks/CCSDBenchmark.ipynb (deuteron HNO data, dense then sparse)

Usage: python tests/golden/make_trajectories.py [/root/reference]
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = pathlib.Path(__file__).with_name("trajectories.json")
STEP = re.compile(r"Step (\d+): (-?[0-9.eE+-]+)")


def notebook_text(path):
    nb = json.loads(path.read_text())
    chunks = []
    for cell in nb["cells"]:
        for o in cell.get("outputs", []):
            if "text" in o:
                chunks.append("".join(o["text"]))
            elif "data" in o and "text/plain" in o["data"]:
                chunks.append("".join(o["data"]["text/plain"]))
    return "\n".join(chunks)


def runs(text):
    """Split 'Step n: E' lines into runs (a run restarts at Step 0)."""
    out = []
    for m in STEP.finditer(text):
        n, e = int(m.group(1)), float(m.group(2))
        if n == 0:
            out.append([])
        out[-1].append(e)
    return out


def main():
    o2 = runs(notebook_text(REF / "examples/Example_CCSD.ipynb"))
    o4 = runs(notebook_text(REF / "examples/Example_CCSD2.ipynb"))
    deu = runs(notebook_text(REF / "benchmarks/CCSDBenchmark.ipynb"))
    t2 = notebook_text(REF / "examples/Example_CCSD.ipynb")
    t4 = notebook_text(REF / "examples/Example_CCSD2.ipynb")
    gs2 = float(re.search(r"is (-?[0-9.]+) MeV", t2).group(1))
    gs4 = float(re.search(r"is (-?[0-9.]+) MeV", t4).group(1))
    eref4 = float(re.search(r"energy of reference: (-?[0-9.]+)", t4).group(1))
    data = {
        "_source": "per-step E_corr (lattice units) printed by ccsd_solver(verbose=True) in the "
                   "reference notebooks; steps 0-10 are plain damped Jacobi, steps 11, 21, 31 "
                   "follow a pulsed-DIIS extrapolation",
        "16O_L2": {"L": 2, "vT1": -9.0, "vS1": -9.0, "w3": 6.0, "a_lat": 2.5,
                   "energies": o2[0], "E_gs_MeV": gs2},
        "16O_L4": {"L": 4, "vT1": -8.0, "vS1": -8.0, "w3": 5.5, "a_lat": 2.0,
                   "energies": o4[0], "E_gs_MeV": gs4, "E_ref_MeV": eref4},
        "deuteron_dense": {"energies": deu[0]},
        "deuteron_sparse": {"energies": deu[1]},
        "paper_fig5_MeV": {"L3": {"8Be": -51.4, "12C": -74.23, "16O": -98.85},
                           "L4": {"8Be": -49.42, "12C": -71.82, "16O": -95.1},
                           "a_lat": 2.0, "vT1": -8.0, "vS1": -8.0, "w3": 5.5},
    }
    OUT.write_text(json.dumps(data, indent=1))
    print("wrote", OUT, {k: len(v.get("energies", [])) for k, v in data.items()
                         if isinstance(v, dict)})


if __name__ == "__main__":
    main()
