"""Writes icd_entry_points.json: the entry-point names of the reference ICD's three name tables
(icd_interface.cpp:7-43, macros CHECK_INST_FUNCS / CHECK_PHYS_FUNCS / CHECK_DEV_FUNCS and the
tables that follow them up to line 102), grouped by table. Reads the reference source as text;
run here only (the fixture it writes is what the tests load).

    python tests/golden/gen_icd_entry_points.py /root/reference/icd_interface.cpp
"""
import json
import os
import re
import sys


def main(path):
    text = open(path).read()
    tables = {}
    for m in re.finditer(r"#define (CHECK_[A-Z_]+)\(\)((?:[^\n]*\\\n)*[^\n]*)", text):
        tables[m.group(1)] = re.findall(r"ICD_FUNC\((vk\w+)\)", m.group(2))
    used = {
        "instance": ["CHECK_INST_FUNCS", "CHECK_PHYS_FUNCS", "CHECK_DEV_FUNCS"],
        "physical_device": ["CHECK_PHYS_FUNCS"],
        "device": ["CHECK_DEV_FUNCS"],
    }
    out = {k: sorted({n for t in v for n in tables.get(t, [])}) for k, v in used.items()}
    out["null_names"] = sorted(set(re.findall(r'strcmp\(pName, "(vk\w+)"\)\)\s*return NULL', text)))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "icd_entry_points.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print({k: len(v) for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1])
