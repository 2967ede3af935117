"""`arena` command replacement for brain/evolve.py's `--zig` binary.

`screen` runs on the GPU (src/main.zig:98-101 -> screenCmd).  evolve.py sends
every subcommand to the same binary (brain/evolve.py:113,431), so the oracle
subcommands (eval, psnr, msssimcheck, debug) are forwarded unchanged to the
codec binary named by the CSFUNNEL_ORACLE_BIN environment variable; without it
they exit with status 1.
"""
from __future__ import annotations

import os
import subprocess
import sys

from .screen import screen_cmd


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        sys.stderr.write("usage: arena screen --model FILE --elites POPFILE --out POPFILE --count-out FILE\n"
                         "                    --lambda N --keep K --seed S [--threads T] [--sigma F] [--jump F] [--imm F]\n")
        return 1
    if argv[0] == "screen":
        return screen_cmd(argv[1:])
    oracle = os.environ.get("CSFUNNEL_ORACLE_BIN")
    if oracle:
        return subprocess.run([oracle] + argv).returncode
    sys.stderr.write(f"error: subcommand '{argv[0]}' needs the codec binary; set CSFUNNEL_ORACLE_BIN\n")
    return 1


if __name__ == "__main__":
    sys.exit(main())
