"""Writes tests/golden/prop/<design>.json from the reference's TAU 2015
benchmark `benchmark/<design>` (data files only): the netlist (<design>.v),
the assertions (<design>.timing), every cell the netlist uses from both
libraries (<design>_Early.lib, <design>_Late.lib: pin directions and
capacitances, the timings that propagate -- constraint timings (setup, hold,
recovery, removal) carry no delay or transition table and are left out --
with their tables), and the arrival times, required times and slacks the
reference reports before the first design change (the leading report_*
lines of <design>.ops with the matching <design>.output lines).

The parasitics are the SPEF fixtures tests/golden/spef/<design>.spef (the
same files, copied unchanged).  Run from the repo root:

    python tests/golden/make_tau15_prop.py c17 s27_spef s400
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BENCH = "/root/reference/benchmark"
CONSTRAINT = ("setup", "hold", "recovery", "removal", "skew", "non_seq", "nochange")
TABLES = ("cell_rise", "cell_fall", "rise_transition", "fall_transition")

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


def cells_data(lib, names):
    """Pins and propagating timings of the named cells, in library order."""
    tmpl = {g["args"][0]: g for g in lib["groups"] if g["type"] == "lu_table_template"}
    out = {}
    for cell in lib["groups"]:
        if cell["type"] != "cell" or cell["args"][0] not in names:
            continue
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
                ttype = t["attrs"].get("timing_type")
                if ttype and any(ttype.startswith(c) for c in CONSTRAINT):
                    continue
                tables = {}
                for tab in t["groups"]:
                    if tab["type"] not in TABLES:
                        continue
                    tp = tmpl.get(tab["args"][0]) if tab["args"] else None
                    i1 = _nums(tab["attrs"].get("index_1") or (tp and tp["attrs"].get("index_1"))
                               or [])
                    i2 = _nums(tab["attrs"].get("index_2") or (tp and tp["attrs"].get("index_2"))
                               or [])
                    tables[tab["type"]] = {
                        "variable_1": tp["attrs"].get("variable_1") if tp else None,
                        "index_1": i1, "index_2": i2, "values": _nums(tab["attrs"]["values"])}
                arcs.append({"from": t["attrs"]["related_pin"], "to": p["args"][0],
                             "timing_sense": t["attrs"].get("timing_sense"),
                             "timing_type": ttype, "tables": tables})
        out[cell["args"][0]] = {"pins": pins, "arcs": arcs}
    missing = set(names) - set(out)
    assert not missing, missing
    return out


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
    at, slew, load, rat = {}, {}, {}, {}
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
        elif f[0] == "rat":
            rat[f[1]] = [float(x) for x in f[2:6]]
    return at, slew, load, rat


def golden(ops, output, kind):
    """{pin: [early rise, early fall, late rise, late fall]} of the `kind`
    reports (report_at / report_rat / report_slack) that precede the first
    design change."""
    ops = ops.splitlines()
    out = output.splitlines()
    gold = {}
    for i, line in enumerate(ops):
        f = line.split()
        if not f:
            continue
        if not f[0].startswith("report_"):
            break
        if f[0] != kind:
            continue
        pin = f[f.index("-pin") + 1]
        el = 1 if "-late" in f else 0
        rf = 1 if "-fall" in f else 0
        gold.setdefault(pin, [None] * 4)[el * 2 + rf] = float(out[i])
    return gold


def make(design):
    src = os.path.join(BENCH, design)
    rd = lambda n: open(os.path.join(src, n)).read()
    inputs, outputs, insts = netlist(rd(f"{design}.v"))
    names = sorted({i["cell"] for i in insts})
    early = cells_data(parse_liberty(rd(f"{design}_Early.lib")), names)
    late = cells_data(parse_liberty(rd(f"{design}_Late.lib")), names)
    for c in names:
        ke = [(a["from"], a["to"], a["timing_type"]) for a in early[c]["arcs"]]
        kl = [(a["from"], a["to"], a["timing_type"]) for a in late[c]["arcs"]]
        assert ke == kl, (c, ke, kl)
    at, slew, load, rat = timing_file(rd(f"{design}.timing"))
    gold = golden(rd(f"{design}.ops"), rd(f"{design}.output"), "report_at")
    gold_rat = golden(rd(f"{design}.ops"), rd(f"{design}.output"), "report_rat")
    gold_slack = golden(rd(f"{design}.ops"), rd(f"{design}.output"), "report_slack")
    assert gold and all(None not in v for v in gold.values())
    doc = {
        "source": f"TAU 2015 benchmark {design} of the reference (benchmark/{design}): "
                  f"{design}.v, {design}.timing, the cells it uses from {design}_Early.lib / "
                  f"{design}_Late.lib, and the report_at / report_rat / report_slack values of the leading reports of "
                  f"{design}.ops with the matching {design}.output lines; parasitics in "
                  f"../spef/{design}.spef",
        "units": {"time": "ps", "capacitance": "fF", "resistance": "kohm"},
        "inputs": inputs, "outputs": outputs, "instances": insts,
        "at": at, "slew": slew, "load": load, "rat": rat,
        "cells": {"early": early, "late": late},
        "golden_at": gold, "golden_rat": gold_rat, "golden_slack": gold_slack,
    }
    os.makedirs(os.path.join(HERE, "prop"), exist_ok=True)
    path = os.path.join(HERE, "prop", f"{design}.json")
    with open(path, "w") as fh:
        json.dump(doc, fh, separators=(",", ":"))
    print(f"wrote {path}: {len(gold)} pins, {len(insts)} instances, {len(names)} cells")


def main():
    for d in sys.argv[1:] or ["c17"]:
        make(d)


if __name__ == "__main__":
    main()
