#!/usr/bin/env python3
"""Autotuner of the line-kernel launch shapes.

For every supported line length it builds a timer of the row- and
column-pass launch variants (tune_size.hip: rows per block, global twiddles,
columns per block, waves per SIMD) and of candidate FFT plans (radix orders),
runs them on the current GPU and writes the fastest choice per length into
../jms_tuned.h, which jms_cfg.hpp reads.

  python tune.py build            # compile the timers (no GPU needed)
  python tune.py run [--out F]    # time them on the GPU -> JSON lines
  python tune.py emit F           # write jms_tuned.h from the JSON lines
A variant replaces the default only when it is faster by more than 2%.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.dirname(HERE)
ROOT = os.path.dirname(os.path.dirname(CSRC))
BUILD = os.path.join(CSRC, "_build", "tune")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", f"-I{ROOT}/include", f"-I{CSRC}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def sizes():
    txt = open(os.path.join(CSRC, "jms_sizes.h")).read()
    return sorted({int(x) for x in re.findall(r"X\((\d+)\)", txt)})


RADICES = (3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16)  # codelets of jms_fft_device.hpp


def default_plan(n):
    txt = open(os.path.join(CSRC, "jms_fft_device.hpp")).read()
    m = re.search(r"PlanFor<%d>\s*\{\s*using type = Radices<([0-9, ]+)>" % n, txt)
    return tuple(int(x) for x in m.group(1).split(","))


def plan_candidates(n, keep=6):
    """the default plan and the `keep` best-scored ordered factorisations of n
    into 2 or 3 codelet radices (score: butterfly rounds, then registers per
    thread, of a column transformed by up to 64 threads)"""
    found = set()

    def rec(rem, cur):
        if rem == 1:
            if 2 <= len(cur) <= 3:
                found.add(tuple(cur))
            return
        if len(cur) == 3:
            return
        for r in RADICES:
            if rem % r == 0:
                rec(rem // r, cur + [r])

    rec(n, [])

    def score(pl):
        t = n // min(pl)
        tw = 1
        while 2 * tw <= min(t, 64):
            tw *= 2
        rounds = sum(-(-(n // r) // tw) for r in pl)
        nreg = max(-(-(n // r) // tw) * r for r in pl)
        return (rounds, nreg, -pl[0])

    d = default_plan(n)
    best = sorted((p for p in found if p != d), key=score)[:keep]
    return [d] + best


def plans_file(n):
    path = os.path.join(BUILD, f"plans_{n}.h")
    items = " ".join(f'X("{",".join(map(str, p))}", Radices<{", ".join(map(str, p))}>)' for p in plan_candidates(n))
    with open(path, "w") as f:
        f.write(f"#define JMS_TUNE_PLANS(X) {items}\n")
    return path


def build():
    os.makedirs(BUILD, exist_ok=True)

    def one(n):
        out = os.path.join(BUILD, f"tune_{n}")
        cmd = [HIPCC, *FLAGS, f"-DJMS_N={n}", f'-DJMS_PLANS_FILE="{plans_file(n)}"',
               os.path.join(HERE, "tune_size.hip"), "-o", out]
        r = subprocess.run(cmd, stdin=subprocess.DEVNULL, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"tune_{n}: {r.stderr[-2000:]}")
        return out

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        for p in ex.map(one, sizes()):
            print("built", p, flush=True)


def run(out_path):
    with open(out_path, "w") as f:
        for n in sizes():
            exe = os.path.join(BUILD, f"tune_{n}")
            r = subprocess.run(["timeout", "-k", "10", "120", exe], capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"{exe} exited {r.returncode}: {r.stdout[-500:]} {r.stderr[-500:]}")
            f.write(r.stdout)
            f.flush()
            print(f"timed {n}", flush=True)


def choose(results):
    """{n: dict(lpb, gtw, cw, wpe, tw)} from the timing lines; defaults unless a
    variant is more than 2% faster"""
    by = {}
    for r in results:
        by.setdefault(r["n"], []).append(r)
    out = {}
    for n, rs in sorted(by.items()):
        k = {}
        tuned = {r["variant"]: r for r in rs if r["pass"] == "row" and r["variant"].startswith("tuned")}
        if "tuned" in tuned and "tuned_flat" in tuned and tuned["tuned_flat"]["ms"] < 0.98 * tuned["tuned"]["ms"]:
            k["padr"] = 24
        lay = next((r for r in rs if r["pass"] == "layout"), None)
        if lay and lay["row_panel_ms"] + lay["col_panel_ms"] < 0.98 * (lay["row_ms"] + lay["col_ms"]):
            k["panel"] = True
        rows = [r for r in rs if r["pass"] == "row" and not r["variant"].startswith("tuned")]
        cols = [r for r in rs if r["pass"] == "col"]
        if rows:
            d = next(r for r in rows if r["variant"] == "default" and r["gtw"] == 0)
            b = min(rows, key=lambda r: r["ms"])
            if b["ms"] < 0.98 * d["ms"]:
                k["lpb"], k["gtw"] = b["lpb"], b["gtw"]
                if b["variant"] == "tr64":
                    k["tr"] = 64
        if cols:
            cols = [r for r in cols if r.get("twr", 1) == 1]  # per-leg twiddle reads: timed, not selectable
            d = cols[0]  # the first variant timed: the current configuration
            b = min(cols, key=lambda r: r["ms"])
            if b["ms"] < 0.98 * d["ms"]:
                k["cw"], k["wpe"], k["tw"], k["pad"] = b["cw"], b["wpe"], b.get("tw", 0), b.get("pad", 3)
        # long columns: 16 columns per block in the hardware order with
        # nontemporal loads and stores, or 8 in the block order and cache
        # policy jms_cfg.hpp gives them (pairs of waves, 16-tile rotation,
        # ordinary loads and stores)
        xcols = {r["ct"]: r for r in rs if r["pass"] == "xcol" and (r["ct"], r["g"], r["rot"], r.get("nt", 1)) in
                 ((16, 0, 0, 1), (8, 2, 16, 0))}
        if 16 in xcols and 8 in xcols and xcols[8]["ms"] < 0.98 * xcols[16]["ms"]:
            k["cw"] = 8
        # the column pass's loads and stores without the nontemporal hint,
        # timed in the layout the staged batches use
        cnt = next((r for r in rs if r["pass"] == "col_nt" and r["nt"] == 0), None)
        if cnt and lay:
            cur = lay["col_panel_ms"] if k.get("panel") else lay["col_ms"]
            alt = cnt["panel_ms"] if k.get("panel") else cnt["ms"]
            if alt < 0.98 * cur:
                k["ntc"] = False
        elif cnt:
            cur = next((r["ms"] for r in rs if r["pass"] == "col"), None)
            if cur and cnt["ms"] < 0.98 * cur:
                k["ntc"] = False
        # tile width x cache policy x layout of the pair of passes (the
        # staged batches' layout decides which row time goes with it)
        cwn = [r for r in rs if r["pass"] == "col_cw_nt" and r["ms"] < 1e20]
        if cwn and lay:
            cur = (lay["row_panel_ms"] + lay["col_panel_ms"]) if k.get("panel") else (lay["row_ms"] + lay["col_ms"])
            cand = [(lay["row_ms"] + r["ms"], r, False) for r in cwn]
            cand += [(lay["row_panel_ms"] + r["panel_ms"], r, True) for r in cwn if r["panel_ms"] < 1e20]
            t, r, pan = min(cand, key=lambda c: c[0])
            if t < 0.98 * cur:
                k["cw"], k["panel"] = r["cw"], pan
                k["ntc"] = bool(r["nt"])
        plans = [r for r in rs if r["pass"] == "plan"]
        if plans:  # the row and the column transforms each take their fastest plan
            dname = ",".join(map(str, default_plan(n)))
            d = next((r for r in plans if r["plan"] == dname), None)
            br = min(plans, key=lambda r: r["row_ms"])
            bc = min(plans, key=lambda r: r["col_ms"])
            if d and br["plan"] != dname and br["row_ms"] < 0.98 * d["row_ms"]:
                k["plan"] = br["plan"]
            if d and bc["plan"] != k.get("plan", dname) and bc["col_ms"] < 0.98 * d["col_ms"]:
                k["col_plan"] = bc["plan"]
        out[n] = k
    return out


def emit(results, path):
    ch = choose(results)
    lines = ["// jms_tuned.h -- launch shapes per line length chosen by the autotuner",
             "// (csrc/tune/tune.py) from measured kernel times on MI355X.  Generated:",
             "// edit by re-running the tuner.  Fields not given keep the defaults of",
             "// jms_cfg.hpp: LPB (rows per row-pass block; 0 = default), GTW (row-pass",
             "// twiddles from the global table with the product tree), CW (columns per",
             "// wave-local column block), WPE (waves per SIMD asked of the register",
             "// allocator for that kernel), TW (threads per column of that kernel; 0 =",
             "// default), PADC (LDS padding shift of that kernel's tile; 0 = default, >= 20",
             "// none), PADR (the same for the row pass), PANEL (waves of staged batches in",
             "// 16-column panels), TR (row-pass threads per line; 0 = one per butterfly),",
             "// NTC (nontemporal wave loads and stores in the column pass);",
             "// at N = 1024, 2048 CW is the long-column kernel's columns per block (16 or",
             "// 8).  Included by jms_cfg.hpp after the plans.",
             "#pragma once", "", "namespace jms {", "",
             "// FFT plan (radix order) per length: PlanFor<N> unless the tuner chose another",
             "template <int N>",
             "struct TunedPlan {",
             "  using type = typename PlanFor<N>::type;",
             "};", "",
             "// FFT plan of the column transforms: the row plan unless the tuner chose another",
             "template <int N>",
             "struct TunedColPlan {",
             "  using type = typename TunedPlan<N>::type;",
             "};", "",
             "template <int N>",
             "struct Tuned {",
             "  static constexpr int LPB = 0;",
             "  static constexpr bool GTW = false;",
             "  static constexpr int CW = 16;",
             "  static constexpr int WPE = 1;",
             "  static constexpr int TW = 0;",
             "  static constexpr int PADC = 0;",
             "  static constexpr int PADR = 0;",
             "  static constexpr bool PANEL = false;",
             "  static constexpr int TR = 0;",
             "  static constexpr bool NTC = true;",
             "};", ""]
    for n, k in ch.items():
        if "col_plan" in k:
            lines.append(f"template <> struct TunedColPlan<{n}> {{ using type = Radices<{k['col_plan'].replace(',', ', ')}>; }};")
    for n, k in ch.items():
        if "plan" in k:
            lines.append(f"template <> struct TunedPlan<{n}> {{ using type = Radices<{k['plan'].replace(',', ', ')}>; }};")
    for n, k in ch.items():
        if not {"lpb", "gtw", "cw", "wpe", "tw", "pad", "padr", "panel", "tr", "ntc"} & set(k):
            continue
        lines.append(f"template <> struct Tuned<{n}> {{")
        lines.append(f"  static constexpr int LPB = {k.get('lpb', 0)};")
        lines.append(f"  static constexpr bool GTW = {'true' if k.get('gtw') else 'false'};")
        lines.append(f"  static constexpr int CW = {k.get('cw', 16)};")
        lines.append(f"  static constexpr int WPE = {k.get('wpe', 1)};")
        lines.append(f"  static constexpr int TW = {k.get('tw', 0)};")
        lines.append(f"  static constexpr int PADC = {0 if k.get('pad', 3) == 3 else k['pad']};")
        lines.append(f"  static constexpr int PADR = {k.get('padr', 0)};")
        lines.append(f"  static constexpr bool PANEL = {'true' if k.get('panel') else 'false'};")
        lines.append(f"  static constexpr int TR = {k.get('tr', 0)};")
        lines.append(f"  static constexpr bool NTC = {'false' if k.get('ntc') is False else 'true'};")
        lines.append("};")
    lines += ["", "}  // namespace jms", ""]
    open(path, "w").write("\n".join(lines))
    print(f"wrote {path}: {sum(1 for k in ch.values() if k)} tuned lengths, "
          f"{sum(1 for k in ch.values() if 'plan' in k or 'col_plan' in k)} with another plan")


def main(argv):
    if len(argv) < 2 or argv[1] not in ("build", "run", "emit"):
        print(__doc__)
        return 2
    if argv[1] == "build":
        build()
    elif argv[1] == "run":
        out = argv[argv.index("--out") + 1] if "--out" in argv else "tune_results.jsonl"
        run(out)
    else:
        res = [json.loads(l) for l in open(argv[2]) if l.strip().startswith("{") and '"error"' not in l]
        emit(res, os.path.join(CSRC, "jms_tuned.h"))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
