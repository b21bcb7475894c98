"""Read the committed coefficient header (sdeh_tableaus.h) into Python dicts."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "simplediffeq.jl_amd", "csrc", "sdeh_tableaus.h")


def load():
    tabs = {}
    cur = None
    for line in open(HEADER):
        m = re.match(r"namespace (\w+) \{", line)
        if m and m.group(1) != "sdeh_tab":
            cur = tabs.setdefault(m.group(1), {})
            continue
        m = re.match(r"constexpr double (\w+) = (\S+);", line)
        if m and cur is not None:
            cur[m.group(1)] = float.fromhex(m.group(2))
    return tabs


TAB = load()


# Butcher tables of the three adaptive pairs as dicts (stage index from 1):
# c[i], a[(i, j)], b[j], btilde[j] and the stage count.

def tsit5_method():
    t = TAB["tsit5"]
    c = {1: 0.0, 2: t["c1"], 3: t["c2"], 4: t["c3"], 5: t["c4"], 6: t["c5"], 7: t["c6"]}
    a = {}
    for i in range(2, 8):
        for j in range(1, i):
            a[(i, j)] = t["a%d%d" % (i, j)]
    b = {j: t["a7%d" % j] for j in range(1, 7)}
    b[7] = 0.0
    bt = {j: t["btilde%d" % j] for j in range(1, 8)}
    return c, a, b, bt, 7


def vern7_method():
    t = TAB["vern7"]
    c = {1: 0.0, 9: 1.0, 10: 1.0}
    for i in range(2, 9):
        c[i] = t["c%d" % i]
    a = {}
    for k, v in t.items():
        if k.startswith("a") and len(k) == 4:
            a[(int(k[1:3]), int(k[3]))] = v
    b = {j: 0.0 for j in range(1, 11)}
    for j in (1, 4, 5, 6, 7, 8, 9):
        b[j] = t["b%d" % j]
    bt = {j: 0.0 for j in range(1, 11)}
    for j in (1, 4, 5, 6, 7, 8, 9, 10):
        bt[j] = t["btilde%d" % j]
    return c, a, b, bt, 10


def vern9_method():
    t = TAB["vern9"]
    c = {1: 0.0, 15: 1.0, 16: 1.0}
    for i in range(2, 15):
        c[i] = t["c%d" % (i - 1)]
    a = {}
    for k, v in t.items():
        if k.startswith("a") and len(k) == 5:
            a[(int(k[1:3]), int(k[3:5]))] = v
    b = {j: 0.0 for j in range(1, 17)}
    for j in (1, 8, 9, 10, 11, 12, 13, 14, 15):
        b[j] = t["b%d" % j]
    bt = {j: 0.0 for j in range(1, 17)}
    for j in (1, 8, 9, 10, 11, 12, 13, 14, 15, 16):
        bt[j] = t["btilde%d" % j]
    return c, a, b, bt, 16
