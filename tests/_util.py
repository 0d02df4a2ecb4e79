"""Helpers shared by the tests: section-aware .ini rewriting and error norms."""
import os
import shutil

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIGS = os.path.join(ROOT, "flusi_amd", "configs")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def rewrite_ini(src: str, dst_dir: str, name: str = "case.ini", **over) -> str:
    """Copy an .ini into dst_dir replacing Section__key=value entries (added
    when absent); kinematics files next to the source are copied along."""
    lines = open(src).read().splitlines()
    todo = {}
    for k, v in over.items():
        sec, key = k.split("__")
        todo.setdefault(sec, {})[key] = v
    out, sec = [], None
    done = set()
    for ln in lines:
        s = ln.strip()
        if s.startswith("[") and "]" in s:
            if sec in todo:
                for key, v in todo[sec].items():
                    if (sec, key) not in done:
                        out.append(f"{key}={v};")
                        done.add((sec, key))
            sec = s[1:s.index("]")]
        elif sec in todo:
            key = s.split("=", 1)[0] if "=" in s else None
            if key in todo[sec]:
                out.append(f"{key}={todo[sec][key]};")
                done.add((sec, key))
                continue
        out.append(ln)
    if sec in todo:
        for key, v in todo[sec].items():
            if (sec, key) not in done:
                out.append(f"{key}={v};")
                done.add((sec, key))
    for sec_, kv in todo.items():
        if any((sec_, k) not in done for k in kv):
            out.append(f"[{sec_}]")
            out.extend(f"{k}={v};" for k, v in kv.items() if (sec_, k) not in done)
    os.makedirs(dst_dir, exist_ok=True)
    dst = os.path.join(dst_dir, name)
    with open(dst, "w") as fh:
        fh.write("\n".join(out) + "\n")
    for f in os.listdir(os.path.dirname(src)):
        if f.startswith("kin") and f.endswith(".ini"):
            shutil.copy(os.path.join(os.path.dirname(src), f), dst_dir)
    return dst


def rel_l2(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))
