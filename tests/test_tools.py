"""Host-side tools, on the CPU: bench.py's single-GPU contract and its
--case targets, and the profile summary's kernel names."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n", [2, 8])
def test_bench_rejects_more_than_one_gpu(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
    assert r.stdout == ""


def test_bench_case_targets_are_listed():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    for case in ("channel", "channel_fp32", "cavity3d_nnn", "duct3d_npn", "cavity2d_nn", "channel2d_pn",
                 "rayleigh_taylor", "adjoint_128"):
        assert case in r.stdout, case


@pytest.mark.parametrize("kernel,name", [
    ("void fen::project_momentum_kernel<double, 0, 0, 4, 1, 128>(fen::PMArgs<double>)", "project_momentum"),
    ("void fen::project_momentum_kernel<float, 0, 0, 4, 1, 128>(fen::PMArgs<float>)", "project_momentum.f32"),
    ("void fen::momentum_march_kernel<double, 0, 8, 32, 1, 1, 1, 64>(fen::MomArgs<double>)", "momentum"),
    ("void fen::project_march_kernel<float, 0, 4, 16, 1, 1, 4, 64>(fen::ProjArgs<float>)", "project.f32"),
    ("void fen::y_kernel<double, 256, -1, 0>(fen::YArgs<double>)", "y_fwd"),
    ("void fen::y_kernel<double, 256, 1, 0>(fen::YArgs<double>)", "y_inv"),
    ("void fen::thomas_kernel<double>(fen::ThomasArgs<double>)", "thomas"),
    ("void fen::(anonymous namespace)::thomas_c1_kernel<1>(fen::(anonymous namespace)::ThomasGArgs)", "thomas_c1"),
    ("void fen::(anonymous namespace)::thomas_column_kernel<HIP_vector_type<double, 2u> >(x)", "thomas"),
    ("void fen::(anonymous namespace)::vof_sweep_kernel<1>(fen::(anonymous namespace)::SweepArgs)", "vof_sweep"),
    ("void fen::(anonymous namespace)::mf_momentum_kernel<0>(fen::(anonymous namespace)::MFMomArgs)",
     "mf_momentum"),
    ("fen::(anonymous namespace)::adj_momentum_kernel(fen::(anonymous namespace)::AdjDims)", "adj_momentum"),
    ("fen::mean_shift_kernel(double const*, long, double, double*)", "shift"),
    ("__amd_rocclr_copyBuffer", "hip_copy"),
])
def test_profile_summary_names(kernel, name):
    from fen_amd.rocprof_summary import short_name
    assert short_name(kernel) == name
