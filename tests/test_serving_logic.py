"""Host logic of the serving layer (no GPU)."""

from forte2_amd.serving import SCFRequest, job_footprint, serving_stream


def test_stream_is_seeded_and_poisson_ordered():
    a = serving_stream(nreq=40, rate=5.0, nmax=24, seed=11)
    b = serving_stream(nreq=40, rate=5.0, nmax=24, seed=11)
    assert [r.xyz for r in a] == [r.xyz for r in b]
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(x.arrival <= y.arrival for x, y in zip(a, a[1:]))
    sizes = [r.xyz.count("\n") // 3 + 1 for r in a]
    assert min(sizes) >= 1 and max(sizes) <= 24
    # mean inter-arrival ~ 1/rate
    gaps = [y.arrival - x.arrival for x, y in zip(a, a[1:])]
    assert 0.05 < sum(gaps) / len(gaps) < 0.5


def test_footprint_is_dominated_by_B_and_monotone():
    f1 = job_footprint(24, 139, 10)
    f24 = job_footprint(576, 3336, 240)
    assert f24 > f1
    assert f24 >= 8 * 3336 * 576 * 576
    assert job_footprint(576, 3336, 240) < 2 * 8 * 3336 * 576 * 576 + (1 << 30)


def test_request_defaults():
    r = SCFRequest(0, "H 0 0 0\nH 0 0 0.74")
    assert r.basis == "cc-pvdz" and r.aux == "cc-pvtz-jkfit" and r.ms is None


def test_a_failing_step_ends_only_its_job():
    # a build or an SCF step that raises ends that job in the tick; the other
    # jobs of the tick get their results and go on
    import torch

    from forte2_amd.serving import SCFServer

    class _SCF:
        def __init__(self, fail):
            self.fail = fail

        def _execute(self, pending):
            if self.fail == "build":
                raise MemoryError("workspace")
            return [torch.zeros(2, 2)]

    class _Job:
        def __init__(self, fail=None):
            self.scf, self.pending, self.error, self.got = _SCF(fail), ("J", [torch.zeros(2, 2)], "fp64"), None, None
            self.fail = fail

        def send(self, result):
            if self.fail == "step":
                raise ValueError("bad step")
            self.got = result
            return False

    srv = SCFServer(device="cpu", hbm_budget=1 << 30, grouped=False)
    jobs = [_Job(), _Job("build"), _Job("step"), _Job()]
    done = srv._tick(jobs)
    assert done == [jobs[1], jobs[2]]
    assert "MemoryError" in jobs[1].error and "bad step" in jobs[2].error
    assert jobs[0].got is not None and jobs[3].got is not None and jobs[0].error is None

