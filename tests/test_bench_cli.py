"""bench.py's measurement helpers on the CPU: the one-GPU contract, kernel-name
shortening for the rocprofv3 summaries, and the utilisation formulas of the
integrator counter summary (checked on a synthetic counter file with known
values)."""
import csv
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("name,short", [
    ("void gpugi::(anonymous namespace)::k_wf_shade<0>(gpugi::DrawParams, gpugi::WavefrontBuffers, int)",
     "k_wf_shade<0>"),
    ("void gpugi::k_trace_wide<false, 256, 10, false>(gpugi::WideScene, HIP_vector_type<float, 4u> const*)",
     "k_trace_wide<false, 256, 10, false>"),
    ("gpugi::k_raygen_shadow(HIP_vector_type<float, 4u> const*, unsigned int)", "k_raygen_shadow"),
    ("__amd_rocclr_copyBuffer", "__amd_rocclr_copyBuffer"),
])
def test_short_kernel_names(name, short):
    assert bench._short_kernel(name) == short


def test_integrator_counter_formulas(tmp_path, monkeypatch):
    """Two dispatches of the path-tracer kernel: per-dispatch averages, VALU issue per
    CU per cycle over GRBM_GUI_ACTIVE summed across the 8 XCDs of 256 CUs, active
    lanes, busy fractions, hit rates and the wait fraction."""
    kernel = "void gpugi::k_pathtracer<true>(gpugi::DrawParams, unsigned int*)"
    values = [
        {"SQ_INSTS_VALU": 1000.0, "GRBM_GUI_ACTIVE": 800.0, "SQ_THREAD_CYCLES_VALU": 3200.0,
         "SQ_ACTIVE_INST_VALU": 100.0, "TA_TA_BUSY_sum": 1280.0, "TD_TD_BUSY_sum": 2560.0,
         "SQ_WAIT_INST_ANY": 30.0, "SQ_WAVE_CYCLES": 100.0, "TCC_HIT_sum": 75.0, "TCC_MISS_sum": 25.0,
         "TCP_TOTAL_CACHE_ACCESSES_sum": 200.0, "TCP_TCC_READ_REQ_sum": 50.0},
        {"SQ_INSTS_VALU": 3000.0, "GRBM_GUI_ACTIVE": 800.0, "SQ_THREAD_CYCLES_VALU": 3200.0,
         "SQ_ACTIVE_INST_VALU": 100.0, "TA_TA_BUSY_sum": 1280.0, "TD_TD_BUSY_sum": 2560.0,
         "SQ_WAIT_INST_ANY": 30.0, "SQ_WAVE_CYCLES": 100.0, "TCC_HIT_sum": 75.0, "TCC_MISS_sum": 25.0,
         "TCP_TOTAL_CACHE_ACCESSES_sum": 200.0, "TCP_TCC_READ_REQ_sum": 50.0},
    ]
    pmc = tmp_path / "pass"
    pmc.mkdir()
    with open(pmc / "run_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for dispatch in values:
            for c, v in dispatch.items():
                w.writerow({"Kernel_Name": kernel, "Counter_Name": c, "Counter_Value": v})
            # another kernel's rows are not mixed in
            w.writerow({"Kernel_Name": "void gpugi::k_other()", "Counter_Name": "SQ_INSTS_VALU",
                        "Counter_Value": 1e9})
    (tmp_path / "profiles").mkdir()
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    bench.summarize_integrators(str(pmc))
    doc = json.load(open(tmp_path / "profiles" / "integrators_clutter_counters.json"))
    k = doc["kernels"]["k_pathtracer<true>"]
    assert k["dispatches"] == 2
    assert k["per_dispatch"]["SQ_INSTS_VALU"] == 2000.0
    d = k["derived"]
    cu_cycles = 800.0 / 8.0 * 256.0
    assert d["valu_wave_insts_per_cu_cycle"] == pytest.approx(2000.0 / cu_cycles)
    assert d["valu_active_lanes"] == pytest.approx(32.0)
    assert d["ta_busy_frac"] == pytest.approx(1280.0 / cu_cycles)
    assert d["td_busy_frac"] == pytest.approx(2560.0 / cu_cycles)
    assert d["wave_wait_frac"] == pytest.approx(0.3)
    assert d["l2_hit_rate"] == pytest.approx(0.75)
    assert d["l1_hit_rate"] == pytest.approx(0.75)
    assert set(doc["kernels"]) == {"k_pathtracer<true>"}


@pytest.mark.parametrize("profile,kernel_ms", [("trace_wide_heightfield_pmc.json", 2.4),
                                               ("trace_wide_clutter_pmc.json", 5.7)])
def test_committed_traffic_is_a_physical_roofline_input(profile, kernel_ms):
    """The HBM traffic the bench line divides by the launch time comes from the committed
    counter summary: at least the rays read and hits written (2^24 x 48 B), and no more
    than the 8 TB/s peak moves in a launch of that kernel (a lower bound on its measured
    duration), so roofline.frac stays a fraction of the peak."""
    traffic = bench.traffic_from_profile(profile)
    assert traffic is not None
    assert traffic >= (1 << 24) * 48
    assert traffic / (kernel_ms * 1e-3) <= bench.HBM_PEAK_GBS * 1e9
