"""The `knnjoin_main <FNAME> <DIM> <K>` front end: the CSV import quirks of importNDDataset
(import_dataset.cpp:50-86) on the CPU (--parse-only does not touch the GPU), and on the GPU the
console/stat lines and the KNN_out*.txt files (main.cu:470-489, 911-949) against the oracle."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import build as pkg_build, synthetic

CLI = pkg_build.CLI

_NUM = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def parse_like_reference(text: str, dtype=np.float32):
    """Restatement of the reference's import: split at ',', read numbers from each piece
    until the first one that fails to parse."""
    vals = []
    for piece in text.split(","):
        pos = 0
        while True:
            m = _NUM.match(piece, pos)
            if not m:
                break
            vals.append(dtype(float(m.group(1))))
            pos = m.end()
    return np.array(vals, dtype=dtype)


def fnv(vals: np.ndarray) -> str:
    h = 1469598103934665603
    for b in vals.tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def run_cli(args, cwd):
    r = subprocess.run([CLI, *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    return r


needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="knnjoin_main not built")


@needs_cli
@pytest.mark.parametrize("text,dim", [
    ("1.5,2.25\n3,4\n-5e-1,6\n", 2),
    ("x,y\n1,2\n3,4\n5,6\n", 2),              # header: "y\n1" yields nothing, the 1 is lost
    ("1 2 3\n4 5 6\n7 8", 3),                # whitespace-separated, trailing partial point
    ("1.5abc,2\n3,4,\n\n5", 2),               # junk after a number ends the piece
    ("0.1,0.2,0.3,0.4,0.5,0.6,0.7", 2),
    (",,,1,,2,", 1),
])
def test_csv_import_quirks(tmp_path, text, dim):
    f = tmp_path / "in.csv"
    f.write_text(text)
    r = run_cli([str(f), str(dim), "3", "--parse-only"], tmp_path)
    assert r.returncode == 0, r.stderr
    v = parse_like_reference(text)
    n = len(v) // dim
    v = v[: n * dim]
    m = re.search(r"points (\d+) values (\d+) digest ([0-9a-f]+)", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) == n and int(m.group(2)) == n * dim
    assert m.group(3) == fnv(v)


@needs_cli
def test_wrong_argument_count_prints_usage(tmp_path):
    r = run_cli(["only_one"], tmp_path)
    assert r.returncode == 0 and "Incorrect number of input parameters" in r.stdout


@pytest.mark.gpu
@needs_cli
def test_cli_join_writes_reference_files(tmp_path, oracle):
    pts = synthetic.uniform_box(20000, seed=31)
    path = tmp_path / "pts.csv"
    with open(path, "w") as f:
        for x, y in pts.tolist():
            f.write(f"{x!r},{y!r}\n")
    parsed = parse_like_reference(path.read_text()).reshape(-1, 2)
    assert np.array_equal(parsed, pts)
    r = run_cli([str(path), "2", "5", "--beta", "0.025", "--print-table"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rc, ids, d, st = oracle.knnjoin(pts, 5, nd=2, beta=0.025)
    m = re.search(r"\[Verification\] Total distance \(without square root\): ([0-9.]+)", r.stdout)
    assert m and abs(float(m.group(1)) - st.checksum) <= 1e-6 * st.checksum
    lines = (tmp_path / "KNN_out.txt").read_text().splitlines()
    assert lines[0] == "#data point (line is the point id), neighbor point ids"
    got = np.array([[int(t) for t in ln.split(", ") if t] for ln in lines[1:]])
    assert np.array_equal(got, ids)
    dl = (tmp_path / "KNN_out_distances.txt").read_text().splitlines()[1:]
    gd = np.array([[float(t) for t in ln.split(", ") if t] for ln in dl])
    assert np.allclose(gd, d, rtol=1e-5, atol=1e-6)  # written with 6 significant digits
    # the reference's gpu_stats.txt line (main.cu:481-489): totalTime, file, K, eps bucket,
    # total distance, then the build tuple under the reference's own label
    line = (tmp_path / "gpu_stats.txt").read_text().strip()
    tup = line.rsplit(": ", 1)[1]
    stats = line.split(", ")
    assert float(stats[0]) > 0.0
    assert stats[1] == str(path) and stats[2] == "KNN: 5"
    assert stats[3].startswith("Eps bucket: ") and stats[4].startswith("Total dist: ")
    assert ("GPUNUMDIM/NUMINDEXEDDIM/REORDER/SHORTCIRCUIT/THREADMULTI/MAXTHREADSPERPOINT/"
            "DYNAMICTHRESHOLD/STATICTHREADSPERPOINT/BETA/DTYPE(float/double): ") in line
    assert tup.split(", ") == ["2", "2", "1", "1", "-2", "1024", "N/A", "8", "0.025", "float"]


@pytest.mark.gpu
@needs_cli
def test_cli_learned_model_file(tmp_path, oracle):
    from hybrid_k_nearest_neighbor_self_join_amd import learned_radius

    w = learned_radius.init_weights(3) * 0.1
    mf = tmp_path / "model.bin"
    mf.write_bytes(w.astype("<f4").tobytes() + struct.pack("<f", 0.5))
    pts = synthetic.mixed_box(30000, seed=2)
    path = tmp_path / "pts.csv"
    np.savetxt(path, pts, delimiter=",", fmt="%.9g")
    pts = parse_like_reference(path.read_text()).reshape(-1, 2)
    r = run_cli([str(path), "2", "8", "--learned", str(mf), "--print-table"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rc, ids, d, st = oracle.knnjoin(pts, 8, nd=2, beta=0.05)
    lines = (tmp_path / "KNN_out.txt").read_text().splitlines()[1:]
    got = np.array([[int(t) for t in ln.split(", ") if t] for ln in lines])
    assert np.array_equal(got, ids)


@pytest.mark.gpu
@needs_cli
def test_cli_outlier_scores_file(tmp_path, oracle):
    pts = synthetic.uniform_box(5000, seed=7)
    path = tmp_path / "pts.csv"
    np.savetxt(path, pts, delimiter=",", fmt="%.9g")
    pts = parse_like_reference(path.read_text()).reshape(-1, 2)
    r = run_cli([str(path), "2", "6", "--outlier-scores"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    rc, ids, d, _ = oracle.knnjoin(pts, 6, nd=2, beta=0.05)
    s, dr, deg, gr = oracle.outlier_scores(ids, d)
    lines = (tmp_path / "KNN_outlier_scores.txt").read_text().splitlines()
    assert lines[0].startswith("#data point (line is the point id), col0: mean distance")
    rows = [ln.split(", ") for ln in lines[1:]]
    assert len(rows) == len(pts)
    assert np.allclose([float(x[0]) for x in rows], s.astype(np.float64) / 6.0, rtol=1e-5)
    assert [int(x[1]) for x in rows] == list(dr) and [int(x[2]) for x in rows] == list(deg)
    assert [int(x[3]) for x in rows] == list(gr)
