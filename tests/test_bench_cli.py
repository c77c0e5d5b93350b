"""bench.py's command-line contract (CPU only): one GPU per run, and launch shapes pinned
for every solver form on the 320k-triangle stack, each one the autotuner could choose."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_cli_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "--gpus 2 is not supported" in r.stderr
    assert r.stdout == ""


def test_pinned_shapes_cover_every_form_within_the_tuners_candidates():
    b = load_bench()
    for form, precision in b.PRECISION.items():
        shapes = b.PINNED_SHAPES[precision]
        # the SpMV candidates of the tuner: grid 1024/2048 x {16 lanes x 1, 8 x 1, 8 x 2}
        assert shapes["spmv_grid"] in (1024, 2048), form
        assert (shapes["spmv_lanes"], shapes["spmv_unroll"]) in ((16, 1), (8, 1), (8, 2)), form
        # the factor-product candidates: grid 1024/2048 x 2/4/8 lanes
        assert shapes["factor_grid"] in (1024, 2048), form
        assert shapes["factor_lanes"] in (2, 4, 8), form
        # the broad-phase cell multiplier is tuned over 0.5-2x the mean edge
        assert 0.5 <= shapes["cell_mult"] <= 2.0, form
        # the kernel named in the roofline is the pinned SpMV instance
        name = b.kernel_name(form, shapes)
        assert ("%d,256,true,%d>" % (shapes["spmv_lanes"], shapes["spmv_unroll"])) in name, name
