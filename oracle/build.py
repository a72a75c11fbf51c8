"""Build recipe for the oracle's C restatement (test infrastructure only).

gcc -O2 -ffp-contract=off -shared oracle/mc_oracle.c -> oracle/_build/libmc_oracle.so
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmc_oracle.so")


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "mc_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c99", "-o", tmp, src]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{' '.join(cmd)}\n{proc.stdout}{proc.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
