"""bench.py's command line on the CPU: more than one GPU is refused before
anything touches a device, with one JSON line saying why."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=60, cwd=ROOT)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_profile_lookups_name_exact_variants():
    """The roofline's traffic and profile duration come from the committed
    profile of exactly the variant the plan chose (cell tile, particle tile,
    targets per lane, cell records), never from another variant's runs."""
    sys.path.insert(0, ROOT)
    import bench

    prof = os.path.join(ROOT, "profiles")
    for tag in ("halos4194304", "zeldovich128", "plummer1048576", "zeldovich256_active"):
        lines = open(os.path.join(prof, f"hbm_traffic_{tag}.md")).read().splitlines()
        chosen = [l.split(" graph replays")[0] for l in lines
                  if " graph replays traffic_bytes_per_launch:" in l]
        assert len(chosen) == 1, tag
        v = chosen[0]
        traffic, name = bench.traffic_from_profile(tag, v + " graph replays")
        assert traffic and traffic > 0 and name == v + " graph replays"
        ms, kname = bench.duration_from_profile(tag, v)
        assert ms and ms > 0
        pct, ppt, k, rec = v[len("batch_kernel<"):-1].split(",")
        assert kname.endswith(f"<true, 1, {pct}, {ppt}, {k}, {rec}>(cgpu::BatchArgs)")
        # any other variant: no figure
        other = f"batch_kernel<{pct},{ppt},{k},{'false' if rec == 'true' else 'true'}>"
        assert bench.traffic_from_profile(tag, other + " graph replays") == (None, None)
        assert bench.duration_from_profile(tag, other) == (None, None)


def test_plan_option_parsing():
    """--plan pins the batched plan the profiles were taken with; bad values
    are refused before anything touches a device."""
    import pytest
    sys.path.insert(0, ROOT)
    import bench

    assert bench.parse_plan("64,256,1,1,3,2") == (64, 256, 1, 1, 3, 2)
    assert bench.parse_plan("32,256,4,1,1,1") == (32, 256, 4, 1, 1, 1)
    for bad in ("64,256,1,1,3", "48,256,1,1,3,2", "64,256,1,1,2,2", "64,256,0,1,3,2",
                "64,256,1,3,3,2", "a,b,c,d,e,f"):
        with pytest.raises(SystemExit):
            bench.parse_plan(bad)
