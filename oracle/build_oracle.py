"""Build recipe of the oracle's C restatement (test infrastructure only).

``python -m oracle.build_oracle`` compiles oracle/csrc/*.c with gcc into
oracle/_build/liboracle.so (git-ignored).  The reference itself is Python and
has no C sources on this path, so there is no oracle/_ref build.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "liboracle.so")


def build(verbose=True):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", f) for f in sorted(os.listdir(os.path.join(HERE, "csrc")))
            if f.endswith(".c")]
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-o", OUT + ".tmp"] + srcs + ["-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(" ".join(cmd) + "\n" + r.stderr)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
