"""The command line's output files against the reference's own REMC sanity
run (tests/testSimulation.cpp:97-150; its output archive
tests/remctest_GPU.tgz, kept as tests/golden/remctest_layout.json).

The same argfile (tests/remctestfile: ubiquitin with a flexible tail and an
all-flexible UIM, 5 replicas at 240-420 K, 1000 MC steps, an exchange every
10, a sample every 10) and the same data files run through `cgppd`.  The
moves come from other random streams than the reference's GSL ones, so the
values differ; everything else must be the reference's: the file names, every
header line, the line counts, the iteration labels, the temperatures of each
sample, the sample PDB names' (step, temperature) pairs, the line formats, and
the sample PDBs' records apart from the coordinates.
"""
import json
import re
import shutil
import subprocess

import pytest

from conftest import DATA, GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CLI = ROOT / "cgppd_amd" / "cgppd"
NUM = r"(-?\d+\.\d{4}|-?nan)"
RATIO_LINE = re.compile(r"^[ \d]{9}: (\| " + NUM + " " + NUM + " ){5}$")
EXCHANGE_LINE = re.compile(r"^[ \d]{10} " + NUM + " " + NUM + "$")
RECORD = re.compile(r"^(\d+); (-?\d+\.\d{5}) \((-?\d+\.\d{5})\); (\d+\.\d)$")
PDB_PATH = re.compile(r"^output/(remctest_\d+)/pdb/sample_(\d+)_(\d+\.\d)K_(.{5})\.pdb$")


@pytest.fixture(scope="module")
def layout():
    return json.loads((GOLDEN / "remctest_layout.json").read_text())


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    """cgppd -f remctestfile in a directory laid out as the reference's
    repository root (data/ beside the argfile)."""
    if not CLI.exists():
        raise RuntimeError("cgppd is not built: run python -m cgppd_amd.build")
    from cgppd_amd import engine
    if engine.device_count() < 1:
        pytest.skip("no GPU visible")
    work = tmp_path_factory.mktemp("remctest")
    (work / "data" / "application").mkdir(parents=True)
    for name in ("AminoAcids", "pair_potentials_MJ1996", "torsional_pair_potentials"):
        shutil.copy(DATA / name, work / "data" / name)
    for name in ("ubq.pdb", "uim.pdb"):
        shutil.copy(DATA / "application" / name, work / "data" / "application" / name)
    shutil.copy(GOLDEN / "remctestfile", work / "remctestfile")
    res = subprocess.run([str(CLI), "-f", "remctestfile"], cwd=work, capture_output=True, text=True, timeout=240)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0
    out = list((work / "output").glob("remctest_*"))
    assert len(out) == 1
    return out[0]


def data_lines(text, header):
    lines = text.split("\n")[:-1]
    return [l for l in lines if l not in header]


def test_file_names_and_index(run, layout):
    names = sorted(p.name for p in run.iterdir())
    assert names == sorted(list(layout["files"]) + ["pdb"])
    assert (run / "fileindex").read_text() == layout["files"]["fileindex"]["text"]


@pytest.mark.parametrize("name", ["fractionBound", "acceptance_ratios", "exchange_freq", "boundconformations"])
def test_headers_counts_and_labels(run, layout, name):
    ref = layout["files"][name]
    text = (run / name).read_text()
    lines = text.split("\n")[:-1]
    assert len(lines) == ref["lines"]
    assert lines[:len(ref["header"])] == ref["header"]
    data = lines[len(ref["header"]):]
    if name == "boundconformations":
        recs, paths = data[0::2], data[1::2]
        labels = []
        for rec, path in zip(recs, paths):
            m, p = RECORD.match(rec), PDB_PATH.match(path)
            assert m and p, (rec, path)
            step, nc, _, temp = m.groups()
            assert (p.group(2), p.group(3)) == (step, temp)
            assert p.group(4) == "%5.2f" % float(nc)
            labels.append(int(step))
        assert labels == ref["labels"]
        # the same temperatures at every sample step (replica order differs)
        got_t = [RECORD.match(r).group(4) for r in recs]
        for k in range(0, len(labels), 5):
            assert sorted(got_t[k:k + 5]) == sorted(ref["temperatures"][k:k + 5])
    else:
        pattern = EXCHANGE_LINE if name == "exchange_freq" else RATIO_LINE
        for l in data:
            assert pattern.match(l), l
        labels = [int(l.split(":")[0]) if ":" in l else int(l.split()[0]) for l in data]
        assert labels == ref["labels"]


def test_sample_pdbs(run, layout):
    pdbs = sorted((run / "pdb").iterdir())
    assert len(pdbs) == layout["pdb_count"]
    name = re.compile(r"^sample_(\d+)_(\d+\.\d)K_(.{5})\.pdb$")
    pairs = sorted([int(name.match(p.name).group(1)), name.match(p.name).group(2)] for p in pdbs)
    assert pairs == layout["pdb_step_temperature"]
    # records of one sample: the reference's, coordinates and energy aside
    ref = layout["example_pdb"]
    step, temp = name.match(ref["name"]).group(1, 2)
    mine = [p for p in pdbs if p.name.startswith(f"sample_{step}_{temp}K_")]
    assert len(mine) == 1
    got, want = mine[0].read_text().split("\n"), ref["text"].split("\n")
    assert len(got) == len(want)
    for g, w in zip(got, want):
        if w.startswith("ATOM"):
            assert g[:30] == w[:30] and g[54:] == w[54:], (g, w)
            assert re.match(r"^( {0,3}-?\d+\.\d{3}){3}$", g[30:54]), g
        elif w.startswith("REMARK simulation"):
            assert re.match(r"^REMARK simulation: remctest_\d+ $", g), g
        elif w.startswith("REMARK potential"):
            assert re.match(r"^REMARK potential: -?\d+\.\d{10} $", g), g
        else:
            assert g == w
