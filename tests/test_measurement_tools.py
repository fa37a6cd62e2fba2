"""Host-side measurement tooling (no GPU): the rocprofv3 summary reducers that
the committed profiles and the bench line's `traffic` figure come from."""
import csv
import json

from tomography_amd import rocprof_summary as R


def _pmc_csv(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for did, name, grid, v in rows:
            w.writerow({"Dispatch_Id": did, "Kernel_Name": name, "Grid_Size": grid, "Counter_Name": counter,
                        "Counter_Value": v})


def test_pmc_averages_the_last_dispatches_and_finds_the_copy(tmp_path):
    k = "void tomo::(anonymous namespace)::k_tfused<4>(tomo::TileItem const*, int)"
    rows = [(1, k, 64, 10.0), (2, k, 64, 20.0), (3, k, 64, 30.0), (4, "__amd_rocclr_copyBuffer", 4096, 512.0)]
    _pmc_csv(tmp_path / "p_counter_collection.csv", "FETCH_SIZE", rows)
    out = R.pmc(str(tmp_path), "p", "FETCH_SIZE", ["k_tfused<4>=2"])
    rec = out["kernels"]["k_tfused<4>"]
    assert rec["dispatches"] == 3 and rec["avg_kib"] == 20.0
    assert rec["last"] == 2 and rec["avg_kib_last"] == 25.0
    assert out["largest_copy_dispatches"][0]["kib"] == 512.0


def test_traffic_scales_each_counter_by_the_copy_of_known_size(tmp_path):
    # a 1 MiB copy: FETCH reads 512 KiB for it (factor 2), WRITE 1024 KiB (factor 1)
    fetch = {"kernels": {"k_tfused<4>": {"avg_kib_last": 100.0}, "k_tinner<6>": {"avg_kib_last": 1000.0}},
             "largest_copy_dispatches": [{"kib": 512.0}]}
    write = {"kernels": {"k_tfused<4>": {"avg_kib_last": 50.0}, "k_tinner<6>": {"avg_kib_last": 300.0}},
             "largest_copy_dispatches": [{"kib": 1024.0}]}
    sched = lambda T, S: {"tsteps": T, "seg_rows": S, "edge_waves": 1, "inner_waves": 1, "chunk": 60,
                          "grid_mult": 16, "form": 2, "edge_rows": 1, "inner_rows": 10}
    bench = {"config": {"schedule": sched(4, 4)},
             "roofline": {"kernel": "k_tfused", "alg_bytes_per_launch": 1e6, "other_kernels": []},
             "batched": {"schedule": sched(6, 52),
                         "roofline": {"kernel": "k_tinner", "alg_bytes_per_launch": 5e6, "other_kernels": []}}}
    (tmp_path / "f.json").write_text(json.dumps(fetch))
    (tmp_path / "w.json").write_text(json.dumps(write))
    (tmp_path / "b.log").write_text("progress line\n" + json.dumps(bench) + "\n")
    out = R.traffic(str(tmp_path / "f.json"), str(tmp_path / "w.json"), str(tmp_path / "b.log"), 1, "cmd")
    assert out["calibration"]["fetch_factor"] == 2.0 and out["calibration"]["write_factor"] == 1.0
    plain = out["plain_cfg2"]
    assert plain["schedule"] == [4, 4, 1, 1, 60, 16, 2, 1, 10]
    assert plain["kernels"]["k_tfused"]["hbm_bytes_per_launch"] == (100.0 * 2 + 50.0) * 1024
    assert out["batched_cfg3"]["kernels"]["k_tinner"]["hbm_bytes_per_launch"] == (1000.0 * 2 + 300.0) * 1024
    assert out["batched_cfg3"]["kernels"]["k_tinner"]["alg_bytes_per_launch"] == 5e6
