"""The launch-shape table behind the autotuner (mvdr_launch_params_*,
mvdr_launch_table_*): the built-in rules, tuned entries and their clamps, and
the saved text form -- host code only, no GPU."""
import pytest

from multiviewdynrecon_amd import solver as sv


@pytest.fixture(autouse=True)
def clean_table(built_libs):
    sv.clear_launch_table()
    yield
    sv.clear_launch_table()


def test_rules_without_entries():
    # ~4k nodes: 127 row groups of 32 -> one pass each for a lone frame (rounded
    # to the 8 XCDs), at most 64 workgroups per frame in a 32-frame launch
    assert sv.launch_params(4048, 1) == {"cg_workgroups": 128, "eval_workgroups": 1024, "transform_workgroups": 128,
                                         "gram_rows": 1}
    assert sv.launch_params(4048, 32)["cg_workgroups"] == 64
    # 16 frames of ~4k nodes: 127 x 16 <= 2048 workgroups, still one pass
    assert sv.launch_params(4048, 16)["cg_workgroups"] == 128
    assert sv.launch_params(100, 1)["cg_workgroups"] == 8  # at least one workgroup per XCD
    assert sv.launch_params(65535, 64)["transform_workgroups"] == 2048


def test_entries_override_and_clamp():
    sv.set_launch_params(4048, 1, cg_workgroups=1000, eval_workgroups=512, transform_workgroups=256, gram_rows=4)
    p = sv.launch_params(4048, 1)
    # cg: at most the row groups
    assert p == {"cg_workgroups": 128, "eval_workgroups": 512, "transform_workgroups": 256, "gram_rows": 4}
    assert sv.launch_params(4048, 2)["eval_workgroups"] == 1024  # other shapes keep the rule
    assert sv.launch_params(3500, 1)["eval_workgroups"] == 512  # same 1024-node bucket
    sv.set_launch_params(65535, 4, cg_workgroups=8)
    assert sv.launch_params(65535, 4)["cg_workgroups"] == 32  # 2048 row groups over at most 64 passes
    sv.set_launch_params(4048, 32, cg_workgroups=20)
    assert sv.launch_params(4048, 32)["cg_workgroups"] == 24  # rounded up to the XCD count
    sv.set_launch_params(4048, 8, gram_rows=3)
    assert sv.launch_params(4048, 8)["gram_rows"] == 1  # 1, 2 or 4; anything else is the rule


def test_save_load_round_trip(tmp_path):
    sv.set_launch_params(4048, 32, cg_workgroups=32, eval_workgroups=2048, transform_workgroups=512, gram_rows=1)
    sv.set_launch_params(8000, 4, eval_workgroups=256)
    want = {(4048, 32): sv.launch_params(4048, 32), (8000, 4): sv.launch_params(8000, 4)}
    path = tmp_path / "table.txt"
    sv.save_launch_table(path)
    text = path.read_text().splitlines()
    assert text[0].startswith("#") and len(text) == 3
    sv.clear_launch_table()
    assert sv.launch_params(4048, 32)["cg_workgroups"] == 64
    sv.load_launch_table(path)
    for (n, f), p in want.items():
        assert sv.launch_params(n, f) == p


def test_five_field_lines_read_gram_rows_as_rule(tmp_path):
    path = tmp_path / "old.txt"
    path.write_text("# header\n4 32 64 1024 128\n")
    sv.load_launch_table(path)
    assert sv.launch_params(4048, 32) == {"cg_workgroups": 64, "eval_workgroups": 1024, "transform_workgroups": 128,
                                          "gram_rows": 1}


@pytest.mark.parametrize("text", ["4 32 64 1024\n", "4 0 64 1024 128\n", "4 32 64 1024 128 1 9\n",
                                  "x 32 64 1024 128\n", "4 32 64 1024 128 z\n"])
def test_bad_tables_are_rejected_and_keep_the_old_one(tmp_path, text):
    sv.set_launch_params(4048, 32, eval_workgroups=256)
    path = tmp_path / "bad.txt"
    path.write_text("# header\n" + text)
    with pytest.raises(sv.SolverError):
        sv.load_launch_table(path)
    assert sv.launch_params(4048, 32)["eval_workgroups"] == 256
    with pytest.raises(sv.SolverError):
        sv.load_launch_table(tmp_path / "missing.txt")
