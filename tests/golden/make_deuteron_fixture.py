"""Write the deuteron text-Hamiltonian fixtures from the reference checkout (run in the build
container, where the reference exists; the outputs are committed):

* ``deuteron_HNO_basis1_{0B,1B,2B}.dat.gz``: the reference's benchmark data files
  (``benchmarks/H2_data/deuteron_HNO_basis1_*.dat``), gzip-compressed, unchanged;
* ``deuteron_ccsd.json``: the per-step CCSD energies the reference printed for them, dense and
  sparse (``benchmarks/CCSDBenchmark.ipynb``, code cell 2 output: ``ccsd_solver(...,
  verbose=True)`` with the default pulsed DIIS, 33 energies each), stored with
  ``nulattice_amd.textio.write_energy_fixture``.

    python tests/golden/make_deuteron_fixture.py
"""
import gzip
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
from nulattice_amd.textio import write_energy_fixture  # noqa: E402

REF = "/root/reference/benchmarks"


def main():
    for part in ("0B", "1B", "2B"):
        src = os.path.join(REF, "H2_data", f"deuteron_HNO_basis1_{part}.dat")
        dst = os.path.join(HERE, os.path.basename(src) + ".gz")
        with open(src, "rb") as f, open(dst, "wb") as raw, \
                gzip.GzipFile(filename="", mode="wb", compresslevel=9, fileobj=raw, mtime=0) as g:
            shutil.copyfileobj(f, g)
    nb = json.load(open(os.path.join(REF, "CCSDBenchmark.ipynb")))
    text = "".join(nb["cells"][2]["outputs"][0]["text"])
    runs, cur = [], None
    for line in text.splitlines():
        m = re.match(r"Step (\d+): (\S+)", line)
        if m:
            if int(m.group(1)) == 0:
                cur = []
                runs.append(cur)
            cur.append(float(m.group(2)))
    dense, sparse = runs
    out = os.path.join(HERE, "deuteron_ccsd.json")
    src = "benchmarks/CCSDBenchmark.ipynb code cell 2 output"
    write_energy_fixture(out, "dense", dense, holes=[0, 2], source=src, diis="pulsed",
                         converged=True)
    write_energy_fixture(out, "sparse", sparse, holes=[0, 2], source=src, diis="pulsed",
                         converged=True)
    print(f"dense {len(dense)} steps, sparse {len(sparse)} steps ->", out)


if __name__ == "__main__":
    main()
