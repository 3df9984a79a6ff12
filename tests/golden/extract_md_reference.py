"""Write tests/golden/md_reference.json: the numbers the reference's MD
test (tests/test001-md-ch4/output_reference: 2 CH4, NVE velocity Verlet,
20 000 steps of 0.1 fs from add.res.1) prints in every MD step block and in
its summary.  The inputs of that test are kept beside this script in
test001-md-ch4/.

Run in the build container (needs /root/reference):
    python tests/golden/extract_md_reference.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tests/test001-md-ch4/output_reference"
ROWS = ("Tmole", "Tatom", "Intramolecular", "Kinetic Energy", "Noncoulombic", "Coulombic",
        "Total Potential", "Total Kinetic", "Total Energy")


def parse(text):
    """MD step blocks and summary values of an MD run's printed output."""
    lines = text.splitlines()
    steps, cur = [], None
    for l in lines:
        m = re.match(r"\s*MD Step: (\d+)\s+Time\s+(\S+)\s+(\S+)", l)
        if m:
            cur = dict(step=int(m.group(1)), time=m.group(2), units=m.group(3), rows={})
            steps.append(cur)
            continue
        if cur is None:
            continue
        for name in ROWS:
            if l.strip().startswith(name) and ":" in l:
                cur["rows"][name] = l.split(":", 1)[1].split()
        m = re.search(r"COM Velocity \(for each species\) : SPC-1: (\S+),", l)
        if m:
            cur["com_velocity"] = m.group(1)
    summ = {}
    for l in lines:
        for key in ("Average Potential Energy:", "Average Kinetic Energy:", "Average Total Energy:"):
            if l.strip().startswith(key):
                summ[key.rstrip(":")] = l.split(":")[1].split()[0]
        m = re.match(r"\s*T\((mole|atom)\)\s*:\s*(\S+)", l)
        if m:
            summ["T(%s)" % m.group(1)] = m.group(2)
        m = re.match(r"\s*Noncoul Inter Energy\s+(\S+)", l)
        if m:
            summ["Noncoul Inter Energy"] = m.group(1)
    return steps, summ


def main():
    steps, summ = parse(open(SRC).read())
    out = dict(_cite="tests/test001-md-ch4/output_reference (MD step blocks every 100 steps, MD SUMMARY)",
               steps=steps, summary=summ)
    with open(os.path.join(HERE, "md_reference.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("%d step blocks" % len(steps))


if __name__ == "__main__":
    main()
