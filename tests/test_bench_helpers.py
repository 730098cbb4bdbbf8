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
    (dict(jit=0, ntiles=10, max_rows=16, pack_stage=1), "k_batched_r<4>"),
    (dict(jit=0, ntiles=0, max_rows=0, pack_stage=1), "k_forest"),
])
def test_batched_kernel_name(bench, info, name):
    assert bench.batched_kernel_name(info) == name
