"""Writes correlation_h_functions.json: the names of the free functions 0_GEN/Correlation.h declares (read as
text from the reference checkout), the fixture of the drop-in coverage test."""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = []
    for m in re.finditer(r"^(?:int|void|double|bool)\s+([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M):
        if m.group(1) not in names:
            names.append(m.group(1))
    return names


if __name__ == "__main__":
    names = declared_functions(os.path.join(REF, "0_GEN", "Correlation.h"))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "correlation_h_functions.json")
    with open(out, "w") as f:
        json.dump({"header": "0_GEN/Correlation.h", "functions": names}, f, indent=1)
        f.write("\n")
    print(out, names)
