"""bench.py's contract on the CPU: a run on more than one GPU stops with a
message before touching a device (this build runs on one GPU)."""
import bench


def test_bench_refuses_more_than_one_gpu(capsys, monkeypatch):
    monkeypatch.delenv("WORLD_SIZE", raising=False)
    assert bench.main(["--gpus", "2"]) == 2
    assert "exactly one GPU" in capsys.readouterr().err
    monkeypatch.setenv("WORLD_SIZE", "8")
    assert bench.main([]) == 2
