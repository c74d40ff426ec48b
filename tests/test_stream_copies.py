"""Host-device copies of the simulation and ensemble code are ordered on the stream whose
kernels use them (copy_on in osiris_amd/csrc/capi.hip).  A plain hipMemcpy runs on the
null stream, which the library's non-blocking streams do not wait for: a kernel launched
right after it may see the memory as it was before the copy (the device re-layout's
refill saw a stale halt flag that way, and the ensemble's batch graphs could read stale
descriptors).  The host check (tests/test_ensemble_host.py) catches such a use on the
paths the serving loop takes; this test covers every path by reading the sources: plain
hipMemcpy and hipMemset appear only in the drop-in entry points, which launch their
kernels on the null stream themselves."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "osiris_amd", "csrc")
NULL_STREAM_ENTRY_POINTS = {"osr_advance_deposit", "osr_yee_advance", "osr_smooth_vdf",
                            "osr_copy_probe"}


def _functions(src):
    """(name, body) of every top-level function definition."""
    out = []
    for m in re.finditer(r"\n(?:static |inline )?[A-Za-z_][\w:<>\*& ]*?\s\**(\w+)\([^;{]*\)\s*\{",
                         src):
        depth, i = 1, m.end()
        while depth and i < len(src):
            depth += {"{": 1, "}": -1}.get(src[i], 0)
            i += 1
        out.append((m.group(1), src[m.end():i]))
    return out


def test_plain_copies_only_in_null_stream_entry_points():
    offenders = []
    for fn in ("capi.hip", "ensemble.hip", "output.hip", "zpulse.hip"):
        path = os.path.join(CSRC, fn)
        if not os.path.exists(path):
            continue
        for name, body in _functions(open(path).read()):
            if name in NULL_STREAM_ENTRY_POINTS:
                continue
            if re.search(r"\bhipMemcpy\(|\bhipMemset\(", body):
                offenders.append(f"{fn}:{name}")
    assert not offenders, f"plain null-stream copies outside the drop-ins: {offenders}"


def test_the_source_scan_sees_the_entry_points():
    # the scan finds the drop-ins' own plain copies, so an empty result above means
    # something (a planted copy in a simulation function would be reported)
    src = open(os.path.join(CSRC, "capi.hip")).read()
    names = {n for n, b in _functions(src) if re.search(r"\bhipMemcpy\(", b)}
    assert {"osr_advance_deposit", "osr_yee_advance"} <= names
    planted = src.replace("int upload_descs(osr_sim* s) {",
                          "int upload_descs(osr_sim* s) {\n  hipMemcpy(0, 0, 0, hipMemcpyHostToDevice);")
    assert "upload_descs" in {n for n, b in _functions(planted) if re.search(r"\bhipMemcpy\(", b)}
