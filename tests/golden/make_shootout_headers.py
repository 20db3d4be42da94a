"""Writes tests/golden/shootout_headers.txt: the header lines of the
reference's manifest, catalog and samples TSV files, read as text from its
source (src/engine_shootout.cpp, the constants manifest_tsv_header,
catalog_tsv_header and samples_tsv_header), one per line with tabs written
as the two characters backslash-t."""
import os
import re
import sys

SRC = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "src", "engine_shootout.cpp")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shootout_headers.txt")

text = open(SRC).read()
lines = []
for name in ("manifest_tsv_header", "catalog_tsv_header", "samples_tsv_header"):
    m = re.search(r"constexpr std::string_view %s =\s*((?:\"[^\"]*\"\s*)+);" % name, text)
    if not m:
        raise SystemExit("constant %s not found" % name)
    lines.append("".join(re.findall(r"\"([^\"]*)\"", m.group(1))))
with open(OUT, "w") as f:
    for line in lines:
        f.write(line + "\n")
print(OUT)
