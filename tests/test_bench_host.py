"""bench.py's command-line contract on the host (no GPU): one GPU only, sane
defaults, and the algorithmic byte counts its roofline fractions divide by."""

from __future__ import annotations

import bench


def test_more_than_one_gpu_is_refused(capsys, monkeypatch):
    monkeypatch.delenv("WORLD_SIZE", raising=False)
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_distributed_launch_is_refused(capsys, monkeypatch):
    monkeypatch.setenv("WORLD_SIZE", "4")
    assert bench.main([]) == 2
    assert "WORLD_SIZE 4" in capsys.readouterr().err


def test_defaults_fit_the_contract():
    a = bench.parse_args([])
    assert a.gpus == 1 and a.steps >= 10 and a.warmup >= 1
    assert not (a.no_tune or a.no_cpu or a.no_plain or a.atomic)


def test_algorithmic_bytes_of_the_step():
    B, N, H, W = 64, 512 * 512, 2048, 2048
    # the dominant column pass: 8 B in + 4 B measured + 8 B out per pixel and view
    assert bench.kernel_bytes("base step cols", B, N, H, W, 4, False) == B * 20 * N == 335_544_320
    # the row-FFT gather: the sample window in and the row spectrum out per view, the probe once
    assert bench.kernel_bytes("base gather rows", B, N, H, W, 4, False) == B * 16 * N + 8 * N


def test_no_gpu_is_a_clear_refusal(capsys, monkeypatch):
    import torch

    monkeypatch.delenv("WORLD_SIZE", raising=False)
    if torch.cuda.is_available():
        return  # only meaningful on a host without a GPU
    assert bench.main(["--steps", "1", "--warmup", "1"]) == 2
    assert "needs a GPU" in capsys.readouterr().err


def test_recorded_traffic_follows_the_kernel_sources(monkeypatch):
    import bench

    rec = dict(bench.MEASURED_TRAFFIC["base step cols"])
    monkeypatch.setitem(bench.MEASURED_TRAFFIC, "base step cols", dict(rec, sources_sha256="0" * 64))
    assert bench.recorded_traffic("base step cols")[0] is None  # sources changed: no figure
    monkeypatch.setitem(bench.MEASURED_TRAFFIC, "base step cols",
                        dict(rec, sources_sha256=bench.kernel_sources_sha256()))
    assert bench.recorded_traffic("base step cols")[0] == rec["bytes"]
    assert bench.recorded_traffic("no such kernel") == (None, "not profiled")


def test_host_cpu_description():
    import bench

    model, n_phys, n_logical, threads = bench._host_cpu()
    assert isinstance(model, str) and n_phys >= 1 and n_logical >= n_phys // 2 and 1 <= threads <= n_phys
