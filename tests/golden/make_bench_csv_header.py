"""Writes tests/golden/bench_csv_header.txt: the per-run CSV header of
rv-sparse's bench (the CSV_HEADER macro, bench/bench.c:531-534), read from
the reference source so the schema test compares against the reference's own
text rather than a retyped copy. Run where the reference tree is present."""
import os
import re
import sys

REF = os.environ.get("RVSP_REFERENCE_DIR", "/root/reference")
src = open(os.path.join(REF, "bench", "bench.c")).read()
m = re.search(r"#define CSV_HEADER((?:[^\n]*\\\n)*[^\n]*)", src)
if not m:
    sys.exit("CSV_HEADER not found")
header = "".join(re.findall(r'"([^"]*)"', m.group(1)))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bench_csv_header.txt")
with open(out, "w") as f:
    f.write(header + "\n")
print(header)
