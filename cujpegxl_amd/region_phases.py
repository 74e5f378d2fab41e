"""Per-step cost of the region kernel at 3840x2160, d = 1.0 (profiling only).

The profiling build (`python -m cujpegxl_amd.build --region-cut`, into
lib_cut/) makes the region kernel return, uniformly, at the start of step
k + 1 when a device word says k.  Run mode launches the region kernel alone
(no other kernel), 30 times for each cut and form, in one process, in the
order CUTS; a step's cost is the difference of two consecutive cuts.  Under
rocprofv3 the kernel trace gives each step's time and a counter run its
instructions:

    CJX_REGION_CUT=1 rocprofv3 --kernel-trace -f csv -d T -o t -- python3 -m cujpegxl_amd.region_phases
    CJX_REGION_CUT=1 rocprofv3 --pmc SQ_INSTS_VALU SQ_WAVES SQ_INSTS_LDS SQ_INSTS_SALU \\
        SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE -f csv -d P -o p -- python3 -m cujpegxl_amd.region_phases
    python3 -m cujpegxl_amd.region_phases --report T P

Each cut's figure is the median (time) or mean (counters) of its last 25
dispatches; the first 5 warm the caches.
"""

from __future__ import annotations

import csv
import ctypes as C
import glob
import os
import statistics
import sys

# Cut k returns at the start of step k + 1; 6, 7 and 8 cut the transform
# selection after its opening barrier, its row passes and its column passes
# (7 exists in the plain form only); 0 is the whole kernel.
CUTS = (1, 2, 3, 6, 7, 8, 4, 5, 0)
STEPS = {
    1: "1 NV12 -> XYB (36x36 tile)",
    2: "2 inverse gaborish + 3a AQ cells",
    3: "3b AQ blocks and the quant field",
    6: "4a selection: opening barrier (3b's tail)",
    7: "4b selection: row passes",
    8: "4c selection: column passes, RD terms, costs",
    4: "4d decision (and the network, learned)",
    5: "5 chosen DCTs of X and B",
    0: "6 FP16 coefficients out + CfL partials",
}
FORMS = ((0, "plain (region_kernel<false, false>)"), (3, "learned + FP16 matrix cores (region_kernel<true, true>)"))
REPS, WARM = 30, 5
W, H = 3840, 2160


def run() -> None:
    if os.environ.get("CJX_REGION_CUT") != "1":
        raise SystemExit("set CJX_REGION_CUT=1 (the profiling build in lib_cut/)")
    import numpy as np

    import cujpegxl_amd as cj
    from cujpegxl_amd.synth import synthetic_nv12

    L = cj.stages_lib()
    L.cjx_set_region_cut.argtypes = [C.c_int]
    L.cjx_stage_region_only.argtypes = [C.c_void_p, C.POINTER(cj.Nv12Input)]
    luma, chroma = synthetic_nv12(W, H, 0)
    dev = cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()]))
    enc = cj.Encoder(W, H, 1.0, stages=True)
    inp = cj.Nv12Input(dev.ptr, dev.ptr + W * H, W, W)
    for mode, _ in FORMS:
        enc.set_mode(mode)
        for cut in CUTS:
            if L.cjx_set_region_cut(cut) != 0:
                raise RuntimeError("cjx_set_region_cut failed")
            for _ in range(REPS):
                if L.cjx_stage_region_only(enc.handle, C.byref(inp)) != 0:
                    raise RuntimeError("cjx_stage_region_only failed")
            print("form", mode, "cut", cut, "done", flush=True)
    L.cjx_set_region_cut(0)
    enc.close()
    dev.close()


def _region_rows(path: str, pattern: str) -> list[dict]:
    files = glob.glob(os.path.join(path, "**", pattern), recursive=True)
    if not files:
        raise SystemExit(f"no {pattern} under {path}")
    return [r for r in csv.DictReader(open(files[0])) if "region_kernel" in r.get("Kernel_Name", "")]


def report(trace_dir: str, pmc_dir: str) -> None:
    rows = sorted(_region_rows(trace_dir, "*kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"]))
    us = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0 for r in rows]
    per = {}
    for r in _region_rows(pmc_dir, "*counter_collection.csv"):
        d = per.setdefault(int(r["Dispatch_Id"]), {})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    ids = sorted(per)
    want = REPS * len(CUTS) * len(FORMS)
    if len(us) != want or len(ids) != want:
        raise SystemExit(f"expected {want} region dispatches, found {len(us)} timed and {len(ids)} counted")
    i = 0
    for mode, label in FORMS:
        t, c = {}, {}
        for cut in CUTS:
            t[cut] = statistics.median(us[i + WARM:i + REPS])
            sel = [per[d] for d in ids[i + WARM:i + REPS]]
            c[cut] = {k: statistics.mean(s[k] for s in sel) for k in sel[0]}
            i += REPS
        waves = c[0]["SQ_WAVES"]
        print(f"{label}: whole kernel {t[0]:.1f} us, {c[0]['SQ_INSTS_VALU'] / waves:.0f} VALU per wave")
        print(f"  {'step':46s} {'us':>6s} {'share':>6s} {'valu/wave':>9s} {'salu/wave':>9s} {'lds/wave':>8s}"
              f" {'conflict/lds_active':>19s}")
        prev_t, prev_c = 0.0, {k: 0.0 for k in c[0]}
        for cut in CUTS:
            if mode != 0 and cut == 7:
                continue
            dt = t[cut] - prev_t
            dc = {k: c[cut][k] - prev_c[k] for k in c[cut]}
            active = dc["SQ_LDS_IDX_ACTIVE"]
            conflict = dc["SQ_LDS_BANK_CONFLICT"] / active if active > 0 else 0.0
            print(f"  {STEPS[cut]:46s} {dt:6.1f} {dt / t[0]:6.1%} {dc['SQ_INSTS_VALU'] / waves:9.0f}"
                  f" {dc['SQ_INSTS_SALU'] / waves:9.0f} {dc['SQ_INSTS_LDS'] / waves:8.1f} {conflict:19.2f}")
            prev_t, prev_c = t[cut], c[cut]


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "--report":
        report(sys.argv[2], sys.argv[3])
    else:
        run()
