"""The drop-in boundary on the CPU: libbbjd.so loads without a GPU and
exports every function include/*.h declares; the public structs and enums
have the reference header's layout and values (compiled probe against both
headers); the Python wrapper validates input like the reference's
(bbjd.py:160-224) before touching the device."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from bbjd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
REF_INC = "/root/reference/include"


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INC)):
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(bbjd_\w+)\s*\(", txt, flags=re.M):
            if "typedef" in txt[max(0, m.start() - 40):m.start() + 1].split("\n")[-1]:
                continue
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_every_declared_function():
    lib = _lib.load()
    names = declared_functions()
    assert "bbjd_solve" in names and "bbjd_correction_gmres" in names and "bbjd_serve" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "bbjd.h"
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("sizeof.params %zu\nsizeof.result %zu\n", sizeof(bbjd_params), sizeof(bbjd_result));
  O(bbjd_params, n); O(bbjd_params, nev); O(bbjd_params, sigma); O(bbjd_params, target);
  O(bbjd_params, maxsub); O(bbjd_params, minsub); O(bbjd_params, tol); O(bbjd_params, maxiter);
  O(bbjd_params, inner_maxiter); O(bbjd_params, inner_tol); O(bbjd_params, v0);
  O(bbjd_params, matvec_A); O(bbjd_params, matvec_B); O(bbjd_params, precond);
  O(bbjd_params, op_data); O(bbjd_params, verbose);
  O(bbjd_result, nconv); O(bbjd_result, eigenvalues); O(bbjd_result, eigenvectors);
  O(bbjd_result, residuals); O(bbjd_result, iters); O(bbjd_result, status);
  O(bbjd_workspace, n); O(bbjd_workspace, V); O(bbjd_workspace, MA); O(bbjd_workspace, nlock);
  O(bbjd_workspace, kdim); O(bbjd_workspace, gmres_V); O(bbjd_workspace, inner_maxiter);
  O(bbjd_workspace, tmp4); O(bbjd_workspace, swap_nm); O(bbjd_workspace, Q_left);
  O(bbjd_workspace, h_work);
  printf("enum %d %d %d %d %d %d %d\n", BBJD_SUCCESS, BBJD_MAX_ITER, BBJD_ERR_ALLOC, BBJD_ERR_INVALID,
         BBJD_ERR_CALLBACK, BBJD_ERR_LAPACK, BBJD_ERR_BREAKDOWN);
  printf("target %d %d %d %d %d %d %d\n", BBJD_TARGET_NEAREST, BBJD_TARGET_LARGEST_MAG,
         BBJD_TARGET_SMALLEST_MAG, BBJD_TARGET_LARGEST_REAL, BBJD_TARGET_SMALLEST_REAL,
         BBJD_TARGET_LARGEST_IMAG, BBJD_TARGET_SMALLEST_IMAG);
  return 0;
}
"""


def _probe(incdir, tmp_path, tag):
    src = tmp_path / f"probe_{tag}.c"
    exe = tmp_path / f"probe_{tag}"
    src.write_text(PROBE)
    subprocess.run(["gcc", "-std=c11", f"-I{incdir}", str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout


def test_struct_layout_matches_ctypes(tmp_path):
    out = dict(line.rsplit(" ", 1) for line in _probe(INC, tmp_path, "mine").splitlines()
               if not line.startswith(("enum", "target")))
    assert int(out["sizeof.params"]) == C.sizeof(_lib.Params)
    assert int(out["sizeof.result"]) == C.sizeof(_lib.Result)
    for f, _ in _lib.Params._fields_:
        assert int(out[f"bbjd_params.{f}"]) == getattr(_lib.Params, f).offset, f
    for f, _ in _lib.Result._fields_:
        assert int(out[f"bbjd_result.{f}"]) == getattr(_lib.Result, f).offset, f


GPU_PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "bbjd_serve.h"
#define S(T) printf("sizeof." #T " %zu\n", sizeof(T))
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  S(bbjd_gpu_options); S(bbjd_stats); S(bbjd_serve_config); S(bbjd_qz_problem);
  O(bbjd_gpu_options, callback_space); O(bbjd_gpu_options, precision); O(bbjd_gpu_options, stream);
  O(bbjd_gpu_options, matvec_A_c64); O(bbjd_gpu_options, matvec_B_c64); O(bbjd_gpu_options, precond_c64);
  O(bbjd_gpu_options, precond2); O(bbjd_gpu_options, qz_mode); O(bbjd_gpu_options, graphs);
  O(bbjd_gpu_options, matvec_AB); O(bbjd_gpu_options, matvec_AB_c64); O(bbjd_gpu_options, matvec_ABr);
  O(bbjd_gpu_options, inner_form); O(bbjd_gpu_options, blocking_sync);
  O(bbjd_gpu_options, inner_batcher);
  O(bbjd_stats, outer_iters); O(bbjd_stats, lock_iters); O(bbjd_stats, restarts); O(bbjd_stats, inner_steps);
  O(bbjd_stats, dgks_second_passes); O(bbjd_stats, sum_k); O(bbjd_stats, qz_seconds);
  O(bbjd_stats, solve_seconds); O(bbjd_stats, stream_bytes); O(bbjd_stats, restart_fallbacks);
  O(bbjd_stats, graph_captures); O(bbjd_stats, graph_launches); O(bbjd_stats, qz_schur_ties);
  O(bbjd_stats, qz_device_steps); O(bbjd_stats, qz_device_fallbacks); O(bbjd_stats, qz_device_accuracy);
  O(bbjd_stats, inner_steps_discarded); O(bbjd_stats, qz_harmonic_restarts);
  O(bbjd_stats, inner_projected); O(bbjd_stats, inner_projected_fallbacks);
  O(bbjd_serve_config, workers); O(bbjd_serve_config, hbm_budget); O(bbjd_serve_config, precision);
  O(bbjd_serve_config, realtime); O(bbjd_serve_config, sync_yield); O(bbjd_serve_config, graphs);
  O(bbjd_serve_config, order); O(bbjd_serve_config, pin); O(bbjd_serve_config, window);
  O(bbjd_serve_config, qz_mode); O(bbjd_serve_config, inner_form); O(bbjd_serve_config, progress);
  O(bbjd_serve_config, inner_batch); O(bbjd_serve_config, inner_batchers);
  O(bbjd_serve_config, inner_window_us); S(bbjd_batcher_stats);
  return 0;
}
"""


def test_gpu_struct_layouts_match_ctypes(tmp_path):
    """the appended GPU option, statistics and serving structs against their
    ctypes mirrors, field by field"""
    from bbjd_amd import serve
    src = tmp_path / "gprobe.c"
    exe = tmp_path / "gprobe"
    req = "".join(f"  O(bbjd_request, {f});\n" for f, _ in serve.Request._fields_)
    req += "".join(f"  O(bbjd_qz_problem, {f});\n" for f, _ in _lib.QzProblem._fields_)
    req += "".join(f"  O(bbjd_batcher_stats, {f});\n" for f, _ in serve.BatchStats._fields_)
    src.write_text(GPU_PROBE.replace("  return 0;", "  S(bbjd_request);\n" + req + "  return 0;"))
    subprocess.run(["gcc", "-std=c11", f"-I{INC}", str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in
               subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, cls in (("bbjd_gpu_options", _lib.GpuOptions), ("bbjd_stats", _lib.Stats),
                       ("bbjd_serve_config", serve.ServeConfig), ("bbjd_request", serve.Request),
                       ("bbjd_qz_problem", _lib.QzProblem), ("bbjd_batcher_stats", serve.BatchStats)):
        assert int(out[f"sizeof.{cname}"]) == C.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(cls, f).offset, (cname, f)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_INC, "bbjd.h")), reason="reference header absent")
def test_layout_and_enums_identical_to_reference_header(tmp_path):
    mine = _probe(INC, tmp_path, "mine").splitlines()
    ref = _probe(REF_INC, tmp_path, "ref").splitlines()
    assert mine == ref


def test_wrapper_validates_like_the_reference():
    from bbjd_amd import jdqz
    A = sp.eye(10, dtype=np.complex128, format="csr")
    with pytest.raises(ValueError):
        jdqz(sp.random(10, 12, density=0.5, format="csr"), k=2)
    with pytest.raises(ValueError):
        jdqz(A, k=2, which="XX")
    with pytest.raises(ValueError):
        jdqz(A, k=2, v0=np.ones(9))
    with pytest.raises(ValueError):
        jdqz(A, k=2, preconditioner=3)
    with pytest.raises(ValueError):
        jdqz(A, B=object(), k=2)


def test_target_and_status_tables():
    from bbjd_amd import bbjd
    assert bbjd._TARGET_MAP == {'nearest': 0, 'LM': 1, 'SM': 2, 'LR': 3, 'SR': 4, 'LI': 5, 'SI': 6}
    assert sorted(bbjd._STATUS_MSG) == [-5, -4, -3, -2, -1]


def test_autotuned_tile_table_loads_and_is_validated(tmp_path):
    """the committed autotuner table loads into the engine (host-side only);
    a table with an unsupported update tile is refused"""
    import json
    from bbjd_amd import autotune
    lib = _lib.load()
    try:
        t = autotune.apply()
        assert t is not None and t["device"] == "gfx950"
        assert len(t["proj_blocks"]) == len(t["n_upper"]) and len(t["proj_blocks"][0]) == len(t["k_upper"])
        bad = dict(t)
        bad["upd_rows"] = [[100] * len(t["k_upper"]) for _ in t["n_upper"]]
        p = tmp_path / "bad.json"
        p.write_text(json.dumps(bad))
        with pytest.raises(ValueError):
            autotune.apply(str(p))
        assert autotune.apply(str(tmp_path / "missing.json")) is None
    finally:
        lib.bbjd_set_tile_table(0, None, 0, None, None, None)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before
    touching a device"""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_batcher_and_serving_options_are_validated():
    """the inner-solve batcher refuses out-of-range settings before touching
    a device (threads 1..8, jobs per batch 1..16, a window >= 0, inner
    iterations >= 1), and the Python serving config checks them too"""
    from bbjd_amd import serve
    lib = _lib.load()
    lib.bbjd_batcher_create.restype = C.c_void_p
    lib.bbjd_batcher_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int64]
    for args in ((0, 4, 100.0, 20), (9, 4, 100.0, 20), (2, 0, 100.0, 20), (2, 17, 100.0, 20),
                 (2, 4, -1.0, 20), (2, 4, 100.0, 0)):
        assert not lib.bbjd_batcher_create(*args), args
    serve._setup(lib)
    for kw in (dict(inner_batch=17), dict(inner_batch=-1), dict(inner_batchers=0), dict(inner_batchers=9),
               dict(inner_window_us=-1.0)):
        with pytest.raises(ValueError):
            serve._config(lib, **kw)
    cfg = serve._config(lib, inner_batch=16, inner_batchers=4, inner_window_us=300.0)
    assert (cfg.inner_batch, cfg.inner_batchers, cfg.inner_window_us) == (16, 4, 300.0)
