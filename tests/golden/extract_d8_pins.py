"""Writes tests/golden/d8_identity_pins.txt: the FNV-1a pins of the
reference's D8 byte-identity test (tests/cpp/test_flow_d8_identity.cpp:517-588),
one line per fixture/config: key then 8 hex hashes in the order receiver,
direction, donors, outlets, topo, svector, area, weighted.  Run in a container
that has the reference mounted at /root/reference."""
import os
import re

SRC = "/root/reference/tests/cpp/test_flow_d8_identity.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "d8_identity_pins.txt")

pat = re.compile(r'\{"([^"]+)", \{([^}]*)\}\}')
rows = []
with open(SRC) as fh:
    for line in fh:
        m = pat.search(line)
        if m and "/" in m.group(1):
            vals = [v.strip().replace("ULL", "") for v in m.group(2).split(",")]
            rows.append((m.group(1), vals))
with open(OUT, "w") as fh:
    fh.write("# D8 identity pins extracted from tests/cpp/test_flow_d8_identity.cpp\n")
    for k, v in rows:
        fh.write(k + " " + " ".join(v) + "\n")
print(len(rows), "pins written")
