"""Writes tests/golden/prop/c17.json from the reference's TAU 2015 benchmark
`benchmark/c17` (data files only): the netlist (c17.v), the assertions
(c17.timing), the NAND2_X1 cell of both libraries (c17_Early.lib,
c17_Late.lib: pin capacitances, the A1/A2 -> ZN timings and their tables),
and the arrival times the reference reports for every pin before the first
design change (c17.ops lines 1-300 with the matching c17.output lines).

The parasitics are the SPEF fixture tests/golden/spef/c17.spef (the same
file, copied unchanged).  Run from the repo root:

    python tests/golden/make_tau15_c17.py [/root/reference/benchmark/c17]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/benchmark/c17"
CELL = "NAND2_X1"

_TOK = re.compile(r'"(?:[^"\\]|\\.)*"|[(){}:;,]|[^\s(){}:;,"]+')


def _tokens(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = text.replace("\\\n", " ")
    for m in _TOK.finditer(text):
        t = m.group(0)
        if t != "\\":
            yield t


def parse_liberty(text):
    """Liberty as nested groups: {"type", "args", "attrs", "groups"}; a
    simple attribute `a : v ;` maps to v, a complex one `a (x, y) ;` to its
    list of arguments (repeated complex attributes keep the last)."""
    toks = list(_tokens(text))
    pos = 0

    def group(kind, args):
        nonlocal pos
        g = {"type": kind, "args": args, "attrs": {}, "groups": []}
        while pos < len(toks) and toks[pos] != "}":
            name = toks[pos]
            pos += 1
            if toks[pos] == ":":
                pos += 1
                vals = []
                while toks[pos] != ";":
                    vals.append(toks[pos].strip('"'))
                    pos += 1
                pos += 1
                g["attrs"][name] = " ".join(vals)
            elif toks[pos] == "(":
                pos += 1
                a = []
                while toks[pos] != ")":
                    if toks[pos] != ",":
                        a.append(toks[pos].strip('"'))
                    pos += 1
                pos += 1
                if pos < len(toks) and toks[pos] == "{":
                    pos += 1
                    g["groups"].append(group(name, a))
                    pos += 1  # "}"
                else:
                    if pos < len(toks) and toks[pos] == ";":
                        pos += 1
                    g["attrs"][name] = a
            else:
                raise ValueError(f"unexpected token {toks[pos]!r} after {name!r}")
        return g

    assert toks[0] == "library"
    pos = 1
    assert toks[pos] == "("
    while toks[pos] != "{":
        pos += 1
    pos += 1
    return group("library", [])


def _nums(strings):
    out = []
    for s in strings:
        out += [float(x) for x in s.replace(",", " ").split()]
    return out


def cell_data(lib):
    tmpl = {g["args"][0]: g for g in lib["groups"] if g["type"] == "lu_table_template"}
    cell = next(g for g in lib["groups"] if g["type"] == "cell" and g["args"] == [CELL])
    pins, arcs = {}, []
    for p in cell["groups"]:
        if p["type"] != "pin":
            continue
        a = p["attrs"]
        cap = float(a.get("capacitance", 0.0))
        pins[p["args"][0]] = {"direction": a.get("direction"),
                              "cap_rise": float(a.get("rise_capacitance", cap)),
                              "cap_fall": float(a.get("fall_capacitance", cap))}
        for t in p["groups"]:
            if t["type"] != "timing":
                continue
            tables = {}
            for tab in t["groups"]:
                if tab["type"] not in ("cell_rise", "cell_fall", "rise_transition",
                                       "fall_transition"):
                    continue
                tp = tmpl.get(tab["args"][0]) if tab["args"] else None
                i1 = _nums(tab["attrs"].get("index_1") or (tp and tp["attrs"].get("index_1")) or [])
                i2 = _nums(tab["attrs"].get("index_2") or (tp and tp["attrs"].get("index_2")) or [])
                vals = _nums(tab["attrs"]["values"])
                tables[tab["type"]] = {
                    "variable_1": tp["attrs"].get("variable_1") if tp else None,
                    "variable_2": tp["attrs"].get("variable_2") if tp else None,
                    "index_1": i1, "index_2": i2, "values": vals}
            arcs.append({"from": t["attrs"]["related_pin"], "to": p["args"][0],
                         "timing_sense": t["attrs"].get("timing_sense"),
                         "timing_type": t["attrs"].get("timing_type"), "tables": tables})
    return pins, arcs


def netlist(text):
    text = re.sub(r"//[^\n]*", "", text)
    inputs = re.findall(r"\binput\s+(\w+)\s*;", text)
    outputs = re.findall(r"\boutput\s+(\w+)\s*;", text)
    insts = []
    for cell, name, body in re.findall(r"\b(\w+)\s+(\w+)\s*\(\s*(\.[^;]*)\)\s*;", text):
        conns = dict(re.findall(r"\.(\w+)\s*\(\s*(\w+)\s*\)", body))
        insts.append({"name": name, "cell": cell, "pins": conns})
    return inputs, outputs, insts


def timing_file(text):
    at, slew, load = {}, {}, {}
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "at":
            at[f[1]] = [float(x) for x in f[2:6]]
        elif f[0] == "slew":
            slew[f[1]] = [float(x) for x in f[2:6]]
        elif f[0] == "load":
            load[f[1]] = float(f[2])
    return at, slew, load


def golden_at(ops, output):
    ops = ops.splitlines()
    out = output.splitlines()
    gold = {}
    for i, line in enumerate(ops):
        f = line.split()
        if not f:
            continue
        if not f[0].startswith("report_"):
            break
        if f[0] != "report_at":
            continue
        pin = f[f.index("-pin") + 1]
        el = 1 if "-late" in f else 0
        rf = 1 if "-fall" in f else 0
        gold.setdefault(pin, [None] * 4)[el * 2 + rf] = float(out[i])
    return gold


def main():
    rd = lambda n: open(os.path.join(SRC, n)).read()
    early = cell_data(parse_liberty(rd("c17_Early.lib")))
    late = cell_data(parse_liberty(rd("c17_Late.lib")))
    inputs, outputs, insts = netlist(rd("c17.v"))
    at, slew, load = timing_file(rd("c17.timing"))
    gold = golden_at(rd("c17.ops"), rd("c17.output"))
    assert all(None not in v for v in gold.values())
    doc = {
        "source": "TAU 2015 benchmark c17 of the reference (benchmark/c17): c17.v, c17.timing, "
                  "NAND2_X1 of c17_Early.lib / c17_Late.lib, and report_at of c17.ops lines "
                  "1-300 with the matching c17.output lines; parasitics in ../spef/c17.spef",
        "units": {"time": "ps", "capacitance": "fF", "resistance": "kohm"},
        "inputs": inputs, "outputs": outputs, "instances": insts,
        "at": at, "slew": slew, "load": load,
        "cell": CELL,
        "pins": {"early": early[0], "late": late[0]},
        "arcs": {"early": early[1], "late": late[1]},
        "golden_at": gold,
    }
    os.makedirs(os.path.join(HERE, "prop"), exist_ok=True)
    path = os.path.join(HERE, "prop", "c17.json")
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1)
    print(f"wrote {path}: {len(gold)} pins, {len(insts)} instances")


if __name__ == "__main__":
    main()
