"""bench.py's command-line contract on a machine without a GPU: more than one GPU is refused with
a JSON message, and the algorithmic byte model and traffic lookup read what the line reports."""
import json
import os
import subprocess
import sys

import bench
from dopamine_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_byte_model_and_committed_traffic():
    m = M.channel_box(8, 6, 4)
    nc = m.n_cells
    b_mixed, nfp = bench.cg_pass_bytes(m, 1, True)
    assert b_mixed == 44 * nc + 12 * nfp
    assert bench.cg_pass_bytes(m, 1, False)[0] == 88 * nc + 16 * nfp
    assert bench.cg_pass_bytes(m, 0, False)[0] == 40 * nc + 16 * nfp
    # interior faces plus one per periodic pair
    assert nfp == m.n_int_faces + int((m.periodic_pair >= 0).sum()) // 2
    # the measured traffic of the benchmark's finest-level passes is committed under profiles/;
    # the byte model prices each operator index at 4 bytes, while these passes read 16-bit offsets
    # where they fit, so the smoothing passes move as little as 0.8x the model (full stack:
    # 7,943,987 cells, 23,726,143 operator faces)
    nc_f, nf_f = 7_943_987, 23_726_143
    model = {"FcgDir32T<6>": 12 * nf_f + 20 * nc_f, "Up0<float, 6>": 12 * nf_f + 20 * nc_f,
             "Down0<float, 6>": 12 * nf_f + 16 * nc_f, "FcgUpd32": 24 * nc_f}
    for k, alg in model.items():
        assert 0.7 * alg <= bench.measured_traffic(k, "full stack 250x210x160") <= 2.0 * alg, k
    assert bench.measured_traffic("FcgDir32T<6>", "no such workload") is None


def test_multigrid_byte_model():
    m = M.channel_box(8, 6, 4)
    nc, nf = m.n_cells, bench.n_op_faces(m)
    b32 = bench.multigrid_bytes(m, True)
    assert b32 == {"cg_dir": 12 * nf + 20 * nc, "cg_upd": 24 * nc, "mg_down": 12 * nf + 16 * nc,
                   "mg_up": 12 * nf + 20 * nc}
    b64 = bench.multigrid_bytes(m, False)
    assert b64["cg_dir"] == 16 * nf + 40 * nc and b64["mg_up"] == 12 * nf + 32 * nc
