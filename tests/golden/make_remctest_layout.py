"""Writes tests/golden/remctest_layout.json: the file layout of the
reference's REMC sanity run (tests/testSimulation.cpp:97-150), read from the
output archive its test compares against, tests/remctest_GPU.tgz.

The run itself cannot be repeated bit for bit here (its moves come from
GSL's mt19937 streams), so what is kept is what a run with other random
numbers must reproduce: the file names, every header line, the line counts,
the step labels, the sample PDB names' (step, temperature) pairs, and a few
data lines and one sample PDB as format examples.  The run's inputs, the
argfile tests/remctestfile and data/application/{ubq,uim}.pdb, are kept
beside it as tests/golden/remctestfile and tests/golden/data/application/.

Run in the container that holds the reference tree:
    python tests/golden/make_remctest_layout.py
"""
from __future__ import annotations

import json
import re
import tarfile
from pathlib import Path

ARCHIVE = Path("/root/reference/tests/remctest_GPU.tgz")
OUT = Path(__file__).resolve().parent / "remctest_layout.json"


def main():
    files, pdbs = {}, []
    with tarfile.open(ARCHIVE) as tar:
        for m in tar.getmembers():
            if not m.isfile():
                continue
            rel = m.name.split("/", 1)[1]  # strip remctest_GPU/
            if rel.startswith("pdb/"):
                pdbs.append(rel[4:])
                continue
            files[rel] = tar.extractfile(m).read().decode()
        example_pdb = sorted(pdbs, key=lambda n: (int(n.split("_")[1]), n))[-1]
        example = tar.extractfile(f"remctest_GPU/pdb/{example_pdb}").read().decode()
    layout = {"source": "tests/remctest_GPU.tgz of the reference (GPU build), argfile tests/remctestfile",
              "files": {}, "pdb_count": len(pdbs)}
    name_re = re.compile(r"^sample_(\d+)_(\d+\.\d)K_(.{5})\.pdb$")
    layout["pdb_step_temperature"] = sorted([int(name_re.match(n).group(1)), name_re.match(n).group(2)] for n in pdbs)
    for rel, text in sorted(files.items()):
        lines = text.split("\n")
        if lines and lines[-1] == "":
            lines = lines[:-1]
        head = [l for l in lines if l.startswith("#") or l.startswith("Iteration") or l.startswith("index")]
        data = [l for l in lines if l not in head]
        layout["files"][rel] = {"lines": len(lines), "header": head, "first": data[:6], "last": data[-2:]}
        if rel in ("fractionBound", "acceptance_ratios", "exchange_freq"):
            layout["files"][rel]["labels"] = [int(l.split(":")[0]) if ":" in l else int(l.split()[0]) for l in data]
        if rel == "boundconformations":
            recs = [l for l in data if not l.startswith("output/")]
            layout["files"][rel]["labels"] = [int(l.split(";")[0]) for l in recs]
            layout["files"][rel]["temperatures"] = [l.rsplit(";", 1)[1].strip() for l in recs]
        if rel == "fileindex":
            layout["files"][rel]["text"] = text
    layout["example_pdb"] = {"name": example_pdb, "text": example}
    OUT.write_text(json.dumps(layout, indent=1) + "\n")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
