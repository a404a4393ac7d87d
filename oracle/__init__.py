"""ORACLE — test infrastructure only.

CPU restatements of the reference algorithms on the three hot paths. Only
tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import
this package, and only as the checker (or the timed CPU baseline). The product
package snapatac2_amd never imports it.

Pinning status per function is recorded in each module's docstring.
"""
from pathlib import Path

import os

ORACLE_DIR = Path(__file__).resolve().parent
# SNAPGPU_ORACLE_SANITIZED=1 selects the AddressSanitizer/UBSan build of the same
# sources (make sanitize; load it with libasan preloaded)
_SAN = os.environ.get("SNAPGPU_ORACLE_SANITIZED") == "1"
LIB_PATH = ORACLE_DIR / "_build" / ("liboracle_san.so" if _SAN else "liboracle.so")


def build() -> Path:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)] + (["sanitize"] if _SAN else []), check=True)
    return LIB_PATH
