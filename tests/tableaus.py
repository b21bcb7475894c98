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
