"""bench.py's secondary sections: an error in one is recorded in its place,
the later ones are marked as not run, and the sections before it keep their
results (the headline line is still printed).  Host logic only."""
import bench


def test_error_is_recorded_and_later_sections_not_run(capsys):
    extra = {}

    def boom():
        raise RuntimeError("device fault")

    failed = bench.run_sections(extra, [("a", lambda: {"v": 1}), ("b", boom), ("c", lambda: {"v": 3})])
    assert failed == "b"
    assert extra["a"] == {"v": 1}
    assert extra["b"] == {"error": "RuntimeError: device fault"}
    assert extra["c"] == {"not_run": "after the error in b"}
    assert "device fault" in capsys.readouterr().err


def test_all_sections_run_without_error():
    extra = {}
    assert bench.run_sections(extra, [("a", lambda: 1), ("b", lambda: 2)]) is None
    assert extra == {"a": 1, "b": 2}
