"""Writes tests/golden/halloffame_kats.json: the genome and published metrics
of every record of codec-arena's halloffame/pareto.json and
halloffame/champion.json, and the one leaderboard entry whose genome the
snapshot holds (bot 13 of pop/seed.txt, LEADERBOARD.md "Shard champions"),
all data files of the reference snapshot: the known answers the codec label
source is checked against.

    python tests/golden/make_halloffame_kats.py /root/reference
"""
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
recs = json.load(open(os.path.join(ref, "halloffame", "pareto.json")))
champ = json.load(open(os.path.join(ref, "halloffame", "champion.json")))
out = {"source": "halloffame/pareto.json, halloffame/champion.json",
       "records": [{"id": r["id"], "g": r["g"], "metrics": r["metrics"]} for r in recs],
       "champion": {"id": champ["id"], "g": champ["g"], "metrics": champ["metrics"]}}
line = [l for l in open(os.path.join(ref, "pop", "seed.txt")).read().splitlines() if l.split()[:1] == ["13"]][0]
row = [l for l in open(os.path.join(ref, "LEADERBOARD.md")).read().splitlines() if l.startswith("| 11 | 13 | 28 | 13 |")][0]
t = [c.strip() for c in row.strip("|").split("|")]
out["leaderboard"] = {"pop_line": line, "source": "pop/seed.txt; LEADERBOARD.md shard champions",
                      "metrics": {"psnr": float(t[4]), "bpp": float(t[5]), "msssim": float(t[6])}}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "halloffame_kats.json"), "w") as f:
    json.dump(out, f, indent=1)
