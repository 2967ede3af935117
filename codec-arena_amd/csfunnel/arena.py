"""`arena` command replacement for brain/evolve.py's `--zig` binary.

Only the `screen` subcommand is served (src/main.zig:98-101 -> screenCmd); the
oracle subcommands (eval, psnr, msssimcheck, debug) belong to the codec, which
this package does not implement, and exit with status 1.
"""
from __future__ import annotations

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
    sys.stderr.write(f"error: subcommand '{argv[0]}' is not provided by csfunnel (screen only)\n")
    return 1


if __name__ == "__main__":
    sys.exit(main())
