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



def _fake_jobs(n, kind="JK"):
    import torch

    class _SCF:
        def __init__(self, fail=None):
            self.fail = fail

        def _execute(self, pending):
            from forte2_amd._native import NativeError

            if self.fail == "native":
                raise NativeError("f2jk_build_jk failed (-2): launch")
            return ([torch.zeros(2, 2)], [torch.zeros(2, 2)])

    class _Job:
        def __init__(self, fail=None):
            self.scf, self.error, self.got = _SCF(fail), None, None
            self.pending = (kind, [torch.zeros(2, 1)], "fp64")
            self._groupable = {}

        def send(self, result):
            self.got = result
            return False

    return _Job, [_Job() for _ in range(n)]


def test_declined_group_runs_its_jobs_alone():
    # the library declining a group (ValueError / MemoryError) sends its jobs
    # to single-job builds, counted as a decline, not as a failure
    from forte2_amd.serving import SCFServer

    _Job, jobs = _fake_jobs(3)
    srv = SCFServer(device="cpu", hbm_budget=1 << 30)
    srv._groupable = lambda jb, kind, prec, nd: jb._groupable.get((kind, prec, nd), True)

    def decline(grp):
        raise ValueError("f2jk_group_workspace_size: X does not fit one chunk")

    srv._build_group = decline
    done = srv._tick(jobs)
    assert done == [] and all(j.got is not None and j.error is None for j in jobs)
    assert (srv.group_declines, srv.group_failures, srv.build_failures) == (1, 0, 0)
    assert all(j._groupable[("JK", "fp64", 1)] is False for j in jobs)


def test_native_error_in_a_group_ends_its_jobs():
    # a HIP launch/runtime failure (NativeError) is not retried alone: the jobs
    # of that group end with the error recorded, and the failure is counted
    from forte2_amd._native import NativeError
    from forte2_amd.serving import SCFServer

    _Job, jobs = _fake_jobs(2)
    srv = SCFServer(device="cpu", hbm_budget=1 << 30)
    srv._groupable = lambda jb, kind, prec, nd: True

    def fail(grp):
        raise NativeError("f2jk_group_build_jk failed (-2): hipErrorLaunchFailure")

    srv._build_group = fail
    done = srv._tick(jobs)
    assert done == jobs
    assert all("hipErrorLaunchFailure" in j.error and j.got is None for j in jobs)
    assert (srv.group_declines, srv.group_failures, srv.build_failures) == (0, 1, 0)
    # a single-job build that fails in the library is counted as well
    _Job, solo = _fake_jobs(0)
    bad, good = _Job("native"), _Job()
    srv2 = SCFServer(device="cpu", hbm_budget=1 << 30, grouped=False)
    done = srv2._tick([bad, good])
    assert done == [bad] and "NativeError" in bad.error and good.got is not None
    assert srv2.build_failures == 1 and srv2.summary()["build_failures"] == 1
