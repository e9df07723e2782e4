"""Host logic of bench.py that needs no GPU: the traffic figure is taken from
a counter record only while the kernel source is the one the record names,
the CPU description of cpu_baseline, and the kernel-trace summary (GPU busy
as the union of overlapping kernel intervals)."""
import hashlib
import os

import bench


def _tree(tmp_path, src, record_sha):
    os.makedirs(tmp_path / "qudrip_amd" / "csrc")
    os.makedirs(tmp_path / "profiles")
    (tmp_path / "qudrip_amd" / "csrc" / "spmv.hip").write_text(src)
    (tmp_path / "profiles" / "spmv_xxz24_traffic.txt").write_text(
        "header line\n"
        f"traffic-record: kernel=k<8> bytes_per_launch=123456789 spmv.hip-sha256={record_sha}\n")


def test_traffic_taken_from_a_matching_record(tmp_path, monkeypatch):
    src = "// kernel source\n"
    _tree(tmp_path, src, hashlib.sha256(src.encode()).hexdigest())
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    traffic, note = bench.spmv_traffic()
    assert traffic == 123456789.0
    assert "same spmv.hip" in note and "k<8>" in note


def test_traffic_null_when_the_source_changed(tmp_path, monkeypatch):
    _tree(tmp_path, "// edited source\n", hashlib.sha256(b"// the profiled source\n").hexdigest())
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    traffic, note = bench.spmv_traffic()
    assert traffic is None and "another spmv.hip" in note


def test_traffic_null_without_a_record(tmp_path, monkeypatch):
    os.makedirs(tmp_path / "qudrip_amd" / "csrc")
    (tmp_path / "qudrip_amd" / "csrc" / "spmv.hip").write_text("x")
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    traffic, note = bench.spmv_traffic()
    assert traffic is None and "no counter summary" in note


def test_committed_record_parses():
    """The committed record line carries the three fields bench.py reads."""
    line = [ln for ln in open(os.path.join(bench.ROOT, bench.TRAFFIC_PROFILE)) if ln.startswith("traffic-record:")]
    assert len(line) == 1
    kv = dict(f.split("=", 1) for f in line[0].split()[1:])
    assert float(kv["bytes_per_launch"]) > 0 and len(kv["spmv.hip-sha256"]) == 64 and kv["kernel"]


def test_host_cpu_fields():
    h = bench.host_cpu()
    assert set(h) == {"cpu_model", "nproc", "usable_threads"}
    assert h["nproc"] >= 1 and 1 <= h["usable_threads"] <= h["nproc"]


def test_trace_summary_merges_overlapping_kernels(tmp_path):
    rows = [(100, 200), (150, 260), (300, 350), (340, 400), (500, 510)]  # ns; two overlapping pairs
    with open(tmp_path / "run_kernel_trace.csv", "w") as fh:
        fh.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for s, e in reversed(rows):
            fh.write(f"k,{s},{e}\n")
    out = bench.trace_summary(str(tmp_path))
    assert out["dispatches"] == 5
    assert abs(out["span_s"] - 410e-9) < 1e-15
    assert abs(out["busy_s"] - (160 + 100 + 10) * 1e-9) < 1e-15
    assert abs(out["busy_frac"] - 270 / 410) < 1e-12


def test_trace_summary_window_and_gzip(tmp_path):
    import gzip
    us = 1000  # ns per microsecond
    rows = [("void (anonymous namespace)::other(int)", 0, 50 * us),
            ("void (anonymous namespace)::g_a<4, true>(int*)", 100 * us, 200 * us),
            ("void (anonymous namespace)::other(int)", 210 * us, 230 * us),
            ("void (anonymous namespace)::g_b(int)", 240 * us, 300 * us),
            ("void (anonymous namespace)::other(int)", 400 * us, 450 * us)]
    with gzip.open(tmp_path / "run_kernel_trace.csv.gz", "wt") as fh:
        fh.write("Kernel_Name,Start_Timestamp,End_Timestamp\n")
        for k, s, e in rows:
            fh.write(f'"{k}",{s},{e}\n')
    out = bench.trace_summary(str(tmp_path), "g_")
    assert out["dispatches"] == 3
    assert abs(out["span_s"] - 200e-6) < 1e-12 and abs(out["busy_s"] - 180e-6) < 1e-12
    names = {k["kernel"]: (k["calls"], k["ms"]) for k in out["kernels"]}
    assert names == {"g_a<4, true>": (1, 0.1), "g_b": (1, 0.06), "other": (1, 0.02)}
    assert bench.trace_summary(str(tmp_path))["dispatches"] == 5
    two = bench.trace_summary(str(tmp_path), "g_a..other")  # g_a at 100 us to the last `other` (400-450 us)
    assert two["dispatches"] == 4 and abs(two["span_s"] - 350e-6) < 1e-12
