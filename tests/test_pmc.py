"""Counter summaries (CPU): HBM bytes per operation from FETCH_SIZE / WRITE_SIZE runs,
calibrated by the device copy, summed over the template instances one operation launches."""
import csv
import os

import pytest

from ffdas_amd.utils import pmc

MAIN = "void ffdas::dev::das_bwd_x_kernel<false, false, true, 16>(ffdas::dev::Geo)"
TAIL = "void ffdas::dev::das_bwd_x_kernel<false, false, true, 8>(ffdas::dev::Geo)"
FWD = "void ffdas::dev::das_fwd_fi_kernel<(ffdas::Storage)0, false>(ffdas::dev::Geo)"
COPY = "__amd_rocclr_copyBuffer"


def _write(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (name, kib) in enumerate(rows):
            # two XCD partials per dispatch: read_counter adds them
            w.writerow([i, name, counter, kib / 2])
            w.writerow([i, name, counter, kib / 2])


def test_summary_sums_instances_and_calibrates(tmp_path):
    copy_kib = pmc.COPY_BYTES / 1024 * 2.0   # the counter reads twice the copy's bytes
    steps = [(TAIL, 100.0), (MAIN, 1000.0), (FWD, 50.0)] * 3
    steps[1] = (MAIN, 4000.0)                # a warm-up outlier: the median ignores it
    _write(str(tmp_path / "f"), "FETCH_SIZE", steps + [(COPY, copy_kib)])
    _write(str(tmp_path / "w"), "WRITE_SIZE", [(n, v / 10) for n, v in steps] + [(COPY, copy_kib / 2)])
    s = pmc.summarize(str(tmp_path / "f"), str(tmp_path / "w"))
    assert s["calibration"]["fetch_factor"] == pytest.approx(0.5)
    assert s["calibration"]["write_factor"] == pytest.approx(1.0)
    k = s["kernels"]["das_bwd_x_kernel"]
    assert k["instances"] == 2 and k["dispatches"] == 6
    assert k["fetch_bytes"] == pytest.approx((1000.0 + 100.0) * 1024 * 0.5)
    assert k["write_bytes"] == pytest.approx((100.0 + 10.0) * 1024 * 1.0)
    f = s["kernels"]["das_fwd_fi_kernel"]
    assert f["instances"] == 1
    assert f["traffic_bytes"] == pytest.approx(50.0 * 1024 * 0.5 + 5.0 * 1024)
    assert "copyBuffer" not in s["kernels"]
