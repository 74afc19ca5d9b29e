"""Writes tests/golden/summary_csv_header.txt: the column names of the summary
CSV that rv-sparse's bench/analyze.py writes (write_csv, analyze.py:587-623),
read from the reference source as text, so rvsp.report's summary CSV is
compared with the reference's own list. Run where the reference tree is present."""
import os
import re
import sys

REF = os.environ.get("RVSP_REFERENCE_DIR", "/root/reference")
src = open(os.path.join(REF, "bench", "analyze.py")).read()
body = src[src.index("def write_csv"):]
m = re.search(r"writer\.writerow\(\[(.*?)\]\)", body, re.S)
if not m:
    sys.exit("summary header not found")
names = re.findall(r'"([^"]*)"', m.group(1))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "summary_csv_header.txt")
with open(out, "w") as f:
    f.write(",".join(names) + "\n")
print(len(names), "columns")
