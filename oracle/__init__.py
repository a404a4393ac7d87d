"""ORACLE — test infrastructure only.

CPU restatements of the reference algorithms on the three hot paths. Only
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import
this package, and only as the checker (or the timed CPU baseline). The product
package snapatac2_amd never imports it.

Pinning status per function is recorded in each module's docstring.
"""
from pathlib import Path

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "liboracle.so"


def build() -> Path:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH
