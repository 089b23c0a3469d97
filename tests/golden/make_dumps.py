"""Golden MINIFE_DEBUG dumps: run the reference's own miniFE.x, compiled with
-DMINIFE_DEBUG from its sources by oracle/Makefile (oracle/_ref/miniFE_dbg.x;
oracle/_ref/miniFE_ell_dbg.x for its ELL matrix type), on the cases of the
reference's CI dump comparison (1x1x2, 2x1x1, 17x3x5, 20x30x40; one rank),
and keep its *.mtx.1.0 / *.vec.1.0 files: gzipped for the small cases,
SHA-256 for 20x30x40 (dumps/ and dumps_ell/).  For the ELL build also the
matrix figures of its YAML report at 20x30x40 (dumps_ell/20x30x40_yaml.json).
Run here (the reference is not on the GPU box); the fixtures are data.
"""
import gzip
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "miniFE_dbg.x")
OUT = os.path.join(HERE, "dumps")
EXE_ELL = os.path.join(ROOT, "oracle", "_ref", "miniFE_ell_dbg.x")
OUT_ELL = os.path.join(HERE, "dumps_ell")
YAML_KEYS = ["Global NNZ", "Global Memory (GB)", "NNZ per proc MIN", "NNZ per proc MAX",
             "NNZ per proc AVG", "Iterations", "MATVEC Flops", "Total CG Flops"]
SMALL = [(1, 1, 2), (2, 1, 1), (17, 3, 5)]
LARGE = [(20, 30, 40)]


def run(exe, dims):
    d = tempfile.mkdtemp()
    nx, ny, nz = dims
    subprocess.run([exe, f"nx={nx}", f"ny={ny}", f"nz={nz}"], cwd=d, check=True,
                   capture_output=True)
    return d, sorted(f for f in os.listdir(d) if ".mtx." in f or ".vec." in f)


def yaml_figures(d):
    """The YAML_KEYS values of the run's report (first occurrence of each key)."""
    out = {}
    for name in os.listdir(d):
        if name.endswith(".yaml"):
            for line in open(os.path.join(d, name)):
                key, _, val = line.strip().partition(": ")
                if key in YAML_KEYS and key not in out:
                    out[key] = val
    return out


def main():
    for exe in (EXE, EXE_ELL):
        if not os.path.exists(exe):
            sys.exit(f"{exe} missing: make -C oracle ref")
    for exe, out in ((EXE, OUT), (EXE_ELL, OUT_ELL)):
        write(exe, out)
    d, _ = run(EXE_ELL, LARGE[0])
    with open(os.path.join(OUT_ELL, "x".join(map(str, LARGE[0])) + "_yaml.json"), "w") as fh:
        json.dump(yaml_figures(d), fh, indent=1, sort_keys=True)


def write(exe, OUT):
    os.makedirs(OUT, exist_ok=True)
    for dims in SMALL:
        d, files = run(exe, dims)
        case = os.path.join(OUT, "x".join(map(str, dims)))
        os.makedirs(case, exist_ok=True)
        for f in files:
            with open(os.path.join(d, f), "rb") as src, open(os.path.join(case, f + ".gz"),
                                                             "wb") as raw:
                with gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as dst:
                    dst.write(src.read())
        print(dims, files)
    for dims in LARGE:
        d, files = run(exe, dims)
        sums = {f: hashlib.sha256(open(os.path.join(d, f), "rb").read()).hexdigest() for f in files}
        with open(os.path.join(OUT, "x".join(map(str, dims)) + ".json"), "w") as fh:
            json.dump(sums, fh, indent=1, sort_keys=True)
        print(dims, sums)


if __name__ == "__main__":
    main()
