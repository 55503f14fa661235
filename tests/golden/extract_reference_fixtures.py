"""Extract numeric fixtures from the reference tree into tests/golden/*.json.

Run in a container that has the reference mounted at /root/reference:
    python tests/golden/extract_reference_fixtures.py

Outputs (data only, no source text):
  integration_tables.json  -- Gregory start/sigma/omega, BDF and rcorr tables for
                              k = 0..5, parsed as numbers from
                              libcntr/cntr/integration.cpp:152-1184
  fkm_bethe_quench.json    -- golden densities den0/den1 (52 doubles each) read with
                              h5dump from libcntr/test/fkm_bethe_quench.h5
"""
import json
import os
import re
import subprocess

REF = "/root/reference/libcntr"
HERE = os.path.dirname(os.path.abspath(__file__))


def _case_blocks(text, func):
    start = text.index("void %s(" % func)
    end = text.index("\n}\n", start)
    body = text[start:end]
    blocks = {}
    parts = re.split(r"\n\s*case\s+(\d+):", body)
    for i in range(1, len(parts), 2):
        blocks[int(parts[i])] = parts[i + 1]
    return blocks


def _arrays(block):
    out = {}
    for name, off, idx, val in re.findall(
            r"(\w+)\[\s*(?:(\d+)\s*\+\s*)?(\d+)\s*\]\s*=\s*([-+0-9.eE]+)\s*;", block):
        i = int(idx) + (int(off) if off else 0)
        out.setdefault(name, {})[i] = float(val)
    return {n: [d[i] for i in range(max(d) + 1)] for n, d in out.items()}


def integration_tables():
    text = open(os.path.join(REF, "cntr", "integration.cpp")).read()
    greg = _case_blocks(text, "read_gregory_weights")
    bd = _case_blocks(text, "read_bd_weights")
    rc = _case_blocks(text, "read_rcorr")
    res = {"gregory": {}, "bd": {}, "rcorr": {}}
    for k in range(0, 6):
        a = _arrays(greg[k])
        res["gregory"][str(k)] = {"start": a["start"], "sigma": a["sigma"], "omega": a["omega"]}
    for k in range(1, 7):
        res["bd"][str(k)] = _arrays(bd[k])["alpha"]
    for k in range(2, 6):
        res["rcorr"][str(k)] = _arrays(rc[k])["w"]
    return res


def fkm_densities():
    path = os.path.join(REF, "test", "fkm_bethe_quench.h5")
    txt = subprocess.check_output(["/opt/conda/bin/h5dump", "-m", "%.17g", path], text=True)
    out = {}
    for name in ("den0", "den1"):
        seg = txt[txt.index('DATASET "%s"' % name):]
        seg = seg[seg.index("DATA {") + 6:]
        seg = seg[:seg.index("}")]
        seg = re.sub(r"\(\d+\):", " ", seg)  # drop the "(index):" row markers
        out[name] = [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", seg)]
    return out


if __name__ == "__main__":
    t = integration_tables()
    with open(os.path.join(HERE, "integration_tables.json"), "w") as f:
        json.dump(t, f, indent=1)
    d = fkm_densities()
    assert len(d["den0"]) == 52 and len(d["den1"]) == 52, {k: len(v) for k, v in d.items()}
    with open(os.path.join(HERE, "fkm_bethe_quench.json"), "w") as f:
        json.dump({"source": "libcntr/test/fkm_bethe_quench.h5", **d}, f, indent=1)
    print("wrote fixtures")
