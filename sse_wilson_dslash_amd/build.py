"""In-tree build of the HIP libraries (hipcc, gfx950) into sse_wilson_dslash_amd/_lib/."""
from __future__ import annotations

import os
import subprocess
import sys

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")


def build(jobs: int = 8, force: bool = True, quiet: bool = False) -> None:
    """Compile every HIP source for gfx950; force=True rebuilds from scratch."""
    cmd = ["make", "-C", CSRC, f"-j{max(1, min(jobs, 16))}"]
    if force:
        cmd.append("-B")
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        sys.stderr.write(out.stdout[-8000:])
        raise RuntimeError("HIP build failed")
    if not quiet:
        print(f"built HIP libraries in {os.path.join(os.path.dirname(CSRC), '_lib')}")


if __name__ == "__main__":
    build()
