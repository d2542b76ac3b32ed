"""Host-side logic that needs no GPU: the rocprof summaries' kernel families
and HBM-traffic scaling, bench.py's refusal of more than one GPU, and the
run-directory readers of the Fortran driver's output."""
import csv
import io
import os
import subprocess
import sys

import numpy as np
import pytest

from dasha_amd import profiling, rundir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SIG = "(mhd::Geo, double const*, double*, mhd::RKArgs, unsigned int const*, int, int)"


def test_kernel_families_and_launch_groups():
    fam = {
        "mhd::k_curl(mhd::Geo, double const*)": "curl",
        "void mhd::k_zsmooth<2, 0, 0, -1, 1, 16>" + SIG: "avrg_radial_x2",
        "void mhd::k_zsmooth<2, 0, 0, -1, 2, 12>" + SIG: "avrg_radial_x2",
        "void mhd::k_zsmooth<2, 1, 0, -1, 1, 16>" + SIG: "avrgh_x2",
        "void mhd::k_zsmooth<3, 1, 0, -2, 2, 12>" + SIG: "avrgh_x3_addE",
        "void mhd::k_zsmooth<2, 2, 1, 4, 0, 16>" + SIG: "avrg_state_x2_rk",
        "void mhd::k_zsmooth<2, 2, 2, -1, 1, 16>" + SIG: "avrgtt_x2_eq",
        "mhd::k_rhs_a(mhd::Geo, double const*)": "rhs_flux_E_hall",
        "void mhd::k_rhs_b2<1, 16>(mhd::Geo)": "rhs_curlE_pressure",
        "mhd::k_fill_state(mhd::Geo)": "rhs_curlE_pressure",
        "k_dt_partial(mhd::Geo)": "setdt",
        "k_copy(long, double const*, double*)": "copy_probe",
        "void mhd::k_adj_filter<2>(mhd::Geo)": "adj_sweep",
        "mhd::k_adj_bc_x(mhd::Geo)": "adj_fill",
        "k_energy_rows(mhd::Geo)": None,
        "void mhd::km_sweep<0>(mhd::Geo, double const*, double*, int, int const*, int)": "mag_avrg_radial",
        "void mhd::km_sweep<2>(mhd::Geo, double const*, double*, int, int const*, int)": "mag_avrg_state",
        "void mhd::km_fill_yz<1>(mhd::Geo, double*, int)": "mag_fill",
        "mhd::km_curl(mhd::Geo)": "mag_curl",
        "mhd::km_rhs_a(mhd::Geo)": "mag_rhs_flux_E",
        "mhd::km_rhs_b(mhd::Geo)": "mag_rhs_curlE_pressure",
    }
    for name, f in fam.items():
        assert profiling.family(name) == f, name
    g = profiling.launch_group
    # the face-free and face launches of one sweep form one group ...
    assert g("void mhd::k_zsmooth<2, 0, 0, -1, 1, 16>" + SIG) == g("void mhd::k_zsmooth<2, 0, 0, -1, 2, 12>" + SIG)
    # ... the RK stages are alternatives of their family
    assert g("void mhd::k_zsmooth<2, 2, 1, 0, 0, 16>" + SIG) != g("void mhd::k_zsmooth<2, 2, 1, 1, 0, 16>" + SIG)


def _counters(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "pmc_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Counter_Name", "Counter_Value"])
        for name, vals in rows.items():
            for v in vals:
                w.writerow([name, counter, v])


def test_traffic_scaling(tmp_path):
    """the copy kernel fixes the scale of each counter; the two launches of a
    smoothing sweep add up, the RK stages average; dispatches of one kernel
    average"""
    cb = 1 << 20
    plain = "void mhd::k_zsmooth<2, 0, 0, -1, 1, 16>" + SIG
    face = "void mhd::k_zsmooth<2, 0, 0, -1, 2, 12>" + SIG
    rk0 = "void mhd::k_zsmooth<2, 2, 1, 0, 0, 16>" + SIG
    rk1 = "void mhd::k_zsmooth<2, 2, 1, 1, 0, 16>" + SIG
    copy = "k_copy(long, double const*, double*)"
    # FETCH_SIZE reads half the copied bytes, WRITE_SIZE all of them (KiB)
    _counters(tmp_path / "f", "FETCH_SIZE", {copy: [cb / 2048], plain: [100, 300], face: [50],
                                              rk0: [1000], rk1: [2000]})
    _counters(tmp_path / "w", "WRITE_SIZE", {copy: [cb / 1024], plain: [40], face: [10], rk0: [500], rk1: [700]})
    res, fs, ws = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), cb)
    assert fs == 2.0 and ws == 1.0
    r = res["avrg_radial_x2"]
    assert r["read_bytes_per_launch"] == (200 + 50) * 1024 * 2.0
    assert r["write_bytes_per_launch"] == (40 + 10) * 1024
    s = res["avrg_state_x2_rk"]
    assert s["read_bytes_per_launch"] == 1500 * 1024 * 2.0
    assert s["bytes_per_launch"] == 1500 * 1024 * 2.0 + 600 * 1024
    assert "copy_probe" not in res


def test_traffic_needs_the_copy_kernel(tmp_path):
    _counters(tmp_path / "f", "FETCH_SIZE", {"mhd::k_curl(x)": [1]})
    _counters(tmp_path / "w", "WRITE_SIZE", {"mhd::k_curl(x)": [1]})
    with pytest.raises(RuntimeError, match="copy kernel"):
        profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), 1 << 20)


def test_bench_measures_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_read_plot_exponents(tmp_path):
    """m2d layout: ndim, dims, time, coordinates, then 11 labelled fields; an
    Ew.d field drops the E of a three-digit exponent"""
    p = tmp_path / "m2d01.dat"
    fields = np.arange(11 * 6, dtype=float).reshape(11, 3, 2) * 1e-3
    buf = io.StringIO()
    buf.write("2\n2 3\n0.5000E+01\n0.1E+00 0.2E+00\n-1.0E+00 0.0E+00 1.0E+00\n")
    for i in range(11):
        buf.write(f"{i + 1}\n")
        vals = [f"{v:.6E}" for v in fields[i].ravel()]
        if i == 3:
            vals[1] = "0.1234-100"       # 0.1234E-100 as Ew.d writes it
        buf.write(" ".join(vals) + "\n")
    p.write_text(buf.getvalue())
    t, coords, f = rundir.read_plot(str(p))
    assert t == 5.0
    assert np.allclose(coords[0], [0.1, 0.2]) and np.allclose(coords[1], [-1.0, 0.0, 1.0])
    assert f.shape == (11, 3, 2)
    expect = fields.copy()
    expect[3].ravel()[1] = 0.1234e-100
    assert np.allclose(f, expect, rtol=1e-6, atol=0) and f[3].ravel()[1] == 0.1234e-100
    # a wrong field label is an error
    p.write_text(buf.getvalue().replace("\n5\n", "\n6\n", 1))
    with pytest.raises(ValueError, match="field label"):
        rundir.read_plot(str(p))


def test_read_adjoint_records(tmp_path):
    shape3 = (4, 3, 2)
    gx = np.random.default_rng(0).standard_normal((8,) + shape3)
    p = tmp_path / "adjoint.bin"
    with open(p, "wb") as f:
        rundir._write_record(f, np.array([1.25, -0.5], dtype="<f8").tobytes())
        rundir._write_record(f, gx.astype("<f8").tobytes())
    v, ge, g = rundir.read_adjoint(str(p), shape3)
    assert v == 1.25 and ge == -0.5 and np.array_equal(g, gx)
