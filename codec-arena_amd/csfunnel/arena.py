"""`arena` command replacement for brain/evolve.py's `--zig` binary.

evolve.py sends every subcommand to the same binary (brain/evolve.py:113,431):
  screen       runs on the GPU (src/main.zig:98-101 -> screenCmd);
  eval         the codec label source on host threads (src/main.zig:58-95);
  psnr         mean Y/U/V PSNR of two raw clips (src/main.zig:32-55);
  msssimcheck  planeMsssim over a pairs file (src/main.zig:103-125).
`debug` (the reference's encoder/decoder drift report) is forwarded to the
binary named by the CSFUNNEL_ORACLE_BIN environment variable; without it, it
exits with status 1.
"""
from __future__ import annotations

import os
import subprocess
import sys

USAGE = ("arena -- codec-arena screen on the GPU, codec evaluation on the host\n"
         "  eval --corpus DIR --pop FILE --out FILE [--clips FILE] [--slow] [--qscale F] [--dump-recon DIR]\n"
         "  screen --model FILE --elites POPFILE --out POPFILE --count-out FILE\n"
         "         --lambda N --keep K --seed S [--threads T] [--sigma F] [--jump F] [--imm F]\n"
         "  psnr A.yuv B.yuv W H FRAMES\n"
         "  msssimcheck PAIRS.bin\n")


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv:
        sys.stderr.write(USAGE)
        return 1
    cmd, rest = argv[0], argv[1:]
    if cmd == "screen":
        from .screen import screen_cmd
        return screen_cmd(rest)
    if cmd in ("eval", "psnr", "msssimcheck"):
        from . import evalcmd
        return {"eval": evalcmd.eval_cmd, "psnr": evalcmd.psnr_cmd, "msssimcheck": evalcmd.msssimcheck_cmd}[cmd](rest)
    if cmd == "debug":
        oracle = os.environ.get("CSFUNNEL_ORACLE_BIN")
        if oracle:
            return subprocess.run([oracle] + argv).returncode
        sys.stderr.write("error: subcommand 'debug' needs the codec binary; set CSFUNNEL_ORACLE_BIN\n")
        return 1
    sys.stderr.write(USAGE)
    return 1


if __name__ == "__main__":
    sys.exit(main())
