"""Host-side helpers of bench.py: the reading of rocprofv3 --pmc counter
databases that scales FETCH_SIZE / WRITE_SIZE by the calibration copies, and
the kernel name the traffic passes look for."""
import importlib.util
import os
import sqlite3

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def counter_db(path, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (kernel_name text, counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?)", rows)
    con.commit()
    con.close()


ELEM = ("void at::native::vectorized_elementwise_kernel<4, at::native::AUnaryFunctor<double, double, double, "
        "at::native::binary_internal::MulFunctor<double> >, std::array<char*, 2ul> >(int)")


def test_read_counter_means_per_kernel(bench, tmp_path):
    db = str(tmp_path / "p_results.db")
    counter_db(db, [
        ("hsh_jit_batched_r4", "FETCH_SIZE", 100.0), ("hsh_jit_batched_r4", "FETCH_SIZE", 300.0),
        (ELEM, "FETCH_SIZE", 10.0), (ELEM, "FETCH_SIZE", 30.0),
        ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 5.0),
        ("void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<double> >", "FETCH_SIZE", 1e9),
        ("hsh_jit_batched_r4", "WRITE_SIZE", 7.0),
    ])
    kern, elem, blit = bench.read_counter(db, "FETCH_SIZE", "hsh_jit_batched_r4")
    assert (kern, elem, blit) == (200.0, 20.0, 5.0)


def test_read_counter_builtin_kernel_name(bench, tmp_path):
    db = str(tmp_path / "p_results.db")
    counter_db(db, [
        ("void hsh::(anonymous namespace)::k_batched_r<4, false>(hsh::DevModel, int)", "WRITE_SIZE", 8.0),
        ("void hsh::(anonymous namespace)::k_batched_r<2, false>(hsh::DevModel, int)", "WRITE_SIZE", 99.0),
        (ELEM, "WRITE_SIZE", 2.0), ("__amd_rocclr_copyBuffer", "WRITE_SIZE", 1.0),
    ])
    assert bench.read_counter(db, "WRITE_SIZE", "k_batched_r<4>") == (8.0, 2.0, 1.0)


def test_read_counter_needs_both_calibrations(bench, tmp_path):
    db = str(tmp_path / "p_results.db")
    counter_db(db, [("hsh_jit_batched_r4", "FETCH_SIZE", 1.0), ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 1.0)])
    with pytest.raises(RuntimeError):
        bench.read_counter(db, "FETCH_SIZE", "hsh_jit_batched_r4")


@pytest.mark.parametrize("info,name", [
    (dict(jit=1, ntiles=10, max_rows=16, pack_stage=1), "hsh_jit_batched_r4"),
    (dict(jit=1, ntiles=10, max_rows=6, pack_stage=1), "hsh_jit_batched_r2"),
    (dict(jit=1, ntiles=10, max_rows=1, pack_stage=1), "hsh_jit_iso"),
    (dict(jit=1, ntiles=10, max_rows=40, pack_stage=1), "hsh_jit_batched"),
    (dict(jit=1, ntiles=0, max_rows=0, pack_stage=2), "hsh_jit_forest1"),
    (dict(jit=1, ntiles=0, max_rows=0, pack_stage=3), "hsh_jit_forest2"),
    (dict(jit=1, ntiles=0, max_rows=0, pack_stage=1), "hsh_jit_forest0"),
    (dict(jit=0, ntiles=10, max_rows=16, pack_stage=1), "k_batched_r<4>"),
    (dict(jit=0, ntiles=0, max_rows=0, pack_stage=1), "k_forest"),
])
def test_batched_kernel_name(bench, info, name):
    assert bench.batched_kernel_name(info) == name


@pytest.mark.parametrize("steps,K,full,rem,launches,probe,graph", [
    (20, 128, 0, 20, 1, 20, False),       # fewer steps than K: one short direct launch
    (300, 128, 2, 44, 3, 128, True),      # a remainder launch after the graph replays
    (10000, 128, 78, 16, 79, 128, True),
    (1280, 128, 10, 0, 10, 128, True),
    (1, 1, 1, 0, 1, 1, True),
])
def test_launch_plan(bench, steps, K, full, rem, launches, probe, graph):
    lp = bench.launch_plan(steps, K)
    assert (lp["full_launches"], lp["remainder_steps"], lp["kernel_launches"]) == (full, rem, launches)
    assert lp["probe_k"] == probe and lp["graph_replayed"] == graph
    assert lp["full_launches"] * K + lp["remainder_steps"] == steps
    assert lp["steps_per_launch"] == pytest.approx(steps / launches)


def test_roofline_from_timed_steps(bench):
    """The driver's short run (20 steps at K = 128: one 20-step launch of
    256.4 us) must give the fraction its own device time implies, not one
    scaled by K."""
    bytes_step = 109.25 * 800000
    rf = bench.roofline_fields(bytes_step, 20, 0.2564, bench.launch_plan(20, 128))
    # 87.4 MB per step, 12.82 us per step: 6.82 TB/s
    assert rf["frac"] == pytest.approx(0.852, abs=1e-3)
    assert rf["launch_us"] == pytest.approx(256.4) and rf["steps_per_launch"] == 20
    # a long run: 79 launches (78 of 128 steps, one of 16), 12 us per step
    rf = bench.roofline_fields(bytes_step, 10000, 120.0, bench.launch_plan(10000, 128))
    assert rf["frac"] == pytest.approx(109.25 * 800000 / 12e-6 / 8e12)
    assert rf["launch_us"] == pytest.approx(120e3 / 79)
    with pytest.raises(ValueError):
        bench.launch_plan(10, 0)
