"""Build the plain-C caller of the C ABI (test infrastructure): gcc against the
HIP runtime headers, linked to the in-tree libpolfed_hip.so (relative rpath, so
the binary runs wherever the repository tree is copied)."""
import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent.parent
ROCM = pathlib.Path("/opt/rocm")
BIN = HERE / "c_abi_check"


def build() -> pathlib.Path:
    cmd = ["gcc", "-std=c11", "-O2", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}",
           "-o", str(BIN), str(HERE / "c_abi_check.c"),
           f"-L{REPO / 'polfed.jl_amd' / 'lib'}", "-lpolfed_hip",
           f"-L{ROCM / 'lib'}", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../polfed.jl_amd/lib", f"-Wl,-rpath,{ROCM / 'lib'}", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed: {res.stderr}")
    return BIN


if __name__ == "__main__":
    print(build())
