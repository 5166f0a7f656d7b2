"""Builds the oracle's C restatement (oracle/cpu_ref.c -> oracle/lib/liborc.so)
with gcc/OpenMP.  Test infrastructure only."""
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
OUT = HERE / "lib" / "liborc.so"


def build() -> pathlib.Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-o", str(OUT),
           str(HERE / "cpu_ref.c")]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
