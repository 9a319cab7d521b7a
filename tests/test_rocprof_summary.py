"""The rocprof summary tool on small synthetic result databases: kernel
tables, the traffic scaling by a copy kernel and the VALU floor, and the
merge of entries into one JSON file."""
import json
import sqlite3
import types

import pytest

from mlsmpm import rocprof_summary as R

MOM = "void mpm::k_p2g_mom_brick<2, true>(mpm::Params, mpm::PState, mpm::GState, mpm::BrickWork)"
COPY = "__amd_rocclr_copyBuffer"


def kt_db(path):
    db = sqlite3.connect(path)
    db.execute("create table kernels (name, duration, lds_size, vgpr_count, accum_vgpr_count, "
               "sgpr_count, grid_x, workgroup_x, start)")
    for i, dur in enumerate((100_000, 200_000, 300_000)):
        db.execute("insert into kernels values (?,?,?,?,?,?,?,?,?)",
                   (MOM, dur, 43904, 72, 0, 112, 2560, 256, i))
    db.execute("insert into kernels values (?,?,?,?,?,?,?,?,?)", (COPY, 5000, 0, 8, 0, 32, 64, 64, 9))
    db.commit()
    return path


def pmc_db(path, rows):
    db = sqlite3.connect(path)
    db.execute("create table counters_collection (kernel_name, counter_name, value, dispatch_id)")
    for i, (k, c, v) in enumerate(rows):
        db.execute("insert into counters_collection values (?,?,?,?)", (k, c, v, i))
    db.commit()
    return path


def test_short_names_keep_template_arguments():
    assert R.short_name(MOM) == "k_p2g_mom_brick<2, true>"


def test_kernel_table_means_and_last(tmp_path):
    p = kt_db(str(tmp_path / "kt.db"))
    rows = {r[0]: r for r in R.kernel_table(p)}
    mom = rows["k_p2g_mom_brick<2, true>"]
    assert mom[1] == 3 and mom[2] == pytest.approx(200.0) and mom[3] == 100.0 and mom[4] == 300.0
    assert mom[-1] == 10  # workgroups = grid / workgroup
    last = {r[0]: r for r in R.kernel_table(p, last=1)}
    assert last["k_p2g_mom_brick<2, true>"][2] == pytest.approx(300.0)


def test_traffic_scales_by_the_copy_kernel(tmp_path):
    kib = 1024
    f = pmc_db(str(tmp_path / "f.db"), [(MOM, "FETCH_SIZE", 1000), (MOM, "FETCH_SIZE", 3000)])
    w = pmc_db(str(tmp_path / "w.db"), [(MOM, "WRITE_SIZE", 500)])
    cf = pmc_db(str(tmp_path / "cf.db"), [(COPY, "FETCH_SIZE", 512 * kib)])
    cw = pmc_db(str(tmp_path / "cw.db"), [(COPY, "WRITE_SIZE", 1024 * kib)])
    out = tmp_path / "t.json"
    a = types.SimpleNamespace(fetch=f, write=w, copy_fetch=cf, copy_write=cw,
                              copy_bytes=float(1 << 30), copy_kernel=COPY,
                              kernel="k_p2g_mom_brick", stage="p2g_mom", workload="dam-large",
                              out=str(out))
    e = R.traffic(a)
    # a copy of 1 GiB read 512 MiB and wrote 1 GiB: fetch x2, write x1
    assert e["scale_fetch"] == pytest.approx(2.0) and e["scale_write"] == pytest.approx(1.0)
    assert e["bytes_per_launch"] == pytest.approx(2000 * kib * 2.0 + 500 * kib)
    assert json.loads(out.read_text())["dam-large"]["p2g_mom"]["kernel"] == "k_p2g_mom_brick"


def test_valu_floor_and_merge(tmp_path):
    a_db = pmc_db(str(tmp_path / "a.db"), [(MOM, "SQ_INSTS_VALU", 8.0e7), (MOM, "SQ_INSTS_VALU", 9.0e7)])
    out = tmp_path / "v.json"
    out.write_text(json.dumps({"dam-large": {"g2p": {"floor_ms": 1.0}}}))
    a = types.SimpleNamespace(db=a_db, kernel="k_p2g_mom_brick", stage="p2g_mom",
                              workload="dam-large", out=str(out))
    e = R.valu(a)
    assert e["insts_valu"] == pytest.approx(8.5e7)
    # 4 clocks per wave64 instruction per SIMD, 1024 SIMDs at 2.4 GHz
    assert e["floor_ms"] == pytest.approx(8.5e7 * 4 / (1024 * 2.4e9) * 1e3)
    merged = json.loads(out.read_text())["dam-large"]
    assert set(merged) == {"g2p", "p2g_mom"}  # the other entry is kept


def test_ambiguous_or_missing_kernels_fail(tmp_path):
    p = pmc_db(str(tmp_path / "p.db"), [(MOM, "SQ_INSTS_VALU", 1.0),
                                         ("void mpm::k_p2g_mom_brick<1, true>()", "SQ_INSTS_VALU", 2.0)])
    with pytest.raises(SystemExit):
        R._one(p, "k_p2g_mom_brick", "SQ_INSTS_VALU")   # matches two kernels
    with pytest.raises(SystemExit):
        R._one(p, "k_g2p_brick", "SQ_INSTS_VALU")       # matches none
