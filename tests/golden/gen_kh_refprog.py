"""Hashes of the output of the reference's own 2D_KH program (main.f90,
model.f90, fileio.f90 and the reference's routines, unmodified;
oracle/_ref/refprog_2d_kh_cpu built by oracle/Makefile): the standard output
and each snapshot file (Kelvin-Helmholtz to t = 100, 21 files), for
tests/test_gpu_0_fortran.py.  Writes tests/golden/kh_refprog.json.  About
2.5 minutes on one core.

    python3 tests/golden/gen_kh_refprog.py
"""
import hashlib
import json
import resource
import subprocess
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
EXE = HERE.parents[1] / "oracle" / "_ref" / "refprog_2d_kh_cpu"


def big_stack():  # set_dt/glm_ss2 keep automatic arrays on the stack
    resource.setrlimit(resource.RLIMIT_STACK, (resource.getrlimit(resource.RLIMIT_STACK)[1],) * 2)


def main():
    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        (d / "data").mkdir()
        r = subprocess.run([str(EXE)], cwd=d, capture_output=True, text=True, timeout=1200, preexec_fn=big_stack)
        assert r.returncode == 0 and "== end ==" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
        out = {"stdout_sha256": hashlib.sha256(r.stdout.encode()).hexdigest(),
               "files": {p.name: hashlib.sha256(p.read_bytes()).hexdigest() for p in sorted((d / "data").iterdir())}}
    (HERE / "kh_refprog.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(out['files'])} files", flush=True)


if __name__ == "__main__":
    main()
