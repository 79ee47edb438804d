"""The measurement model of bench.py (CPU): algorithmic bytes and matrix-core flops per stage
launch, checked against hand counts on small cases, and the roofline's choice of bound."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("bench_model", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def test_element_sizes():
    assert bench.element_sizes(1, 3) == (4, 2, 1)
    assert bench.element_sizes(2, 3) == (10, 3, 4)
    assert bench.element_sizes(3, 3) == (20, 4, 10)
    assert bench.element_sizes(3, 5) == (56, 4, 21)


def test_algorithmic_flops_hand_count():
    # 3D P3, one element: per field 2 * Np * (d*Np + Nf*Nfp) = 2 * 20 * (60 + 40), four fields
    assert bench.algorithmic_flops(3, 3, 1) == 4 * 2 * 20 * 100
    # per counted DOF: 2 (d Np + Nf Nfp): 2D P2 -> 2 (2*6 + 3*3) = 42
    K = 7
    assert bench.algorithmic_flops(2, 2, K) / (3 * K * 6) == 42


def test_algorithmic_bytes_hand_count():
    # 3D P3, 6 elements, 15 faces of which 6 on the boundary, FP64 plain
    dof = 4 * 6 * 20
    state = dof * 4 * 8                # 4 words per DOF and stage on average
    traces = 2 * 9 * 10 * 4 * 8        # each interior face node's 4 fields read by both sides
    geo = 6 * (9 * 8 + 4 * (4 * 8 + 8))  # J^-1; per face normal + scale and two int32
    assert bench.algorithmic_bytes(3, 3, 6, 15, 6, 8, False) == state + traces + geo
    # the compensated FP32 form adds 2 bytes per DOF per stage (4.5 words of 4 bytes)
    b32 = bench.algorithmic_bytes(3, 3, 6, 15, 6, 4, True)
    assert b32 == dof * 18 + 2 * 9 * 10 * 4 * 4 + 6 * (9 * 4 + 4 * (4 * 4 + 8))


@pytest.mark.parametrize("word,flops,bound", [(4, 25.0, "mfma"), (4, 15.0, "hbm"), (8, 12.0, "mfma"),
                                              (8, 5.0, "hbm")])
def test_roofline_takes_the_roof_met_first(word, flops, bound):
    nbytes = 1e9
    r = bench.roofline(nbytes, flops * nbytes, word, 1e-3, "k_none", "none")
    assert r["bound"] == bound
    if bound == "mfma":
        assert r["unit"] == "TFLOP/s" and r["peak"] == bench.MFMA_PEAK[word] / 1e12
        assert abs(r["achieved"] - flops * nbytes / 1e-3 / 1e12) < 1e-2
    else:
        assert r["unit"] == "GB/s" and r["peak"] == 8000.0
        assert abs(r["achieved"] - 1000.0) < 1e-6  # 1 GB in 1 ms
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3


@pytest.mark.parametrize("kernel,workload,dim,p,K,word,comp", [
    ("k_stage_mfma<double,3,3>", "cube3d-p3", 3, 3, 196608, 8, False),
    ("k_stage_mfma<float,3,3>", "cube3d-p3", 3, 3, 196608, 4, True),
    ("k_stage_mfma<float,3,5> (largest batch group)", "ensemble256", 3, 5, None, 4, True),
])
def test_measured_traffic_is_found_for_every_roofline_kernel(kernel, workload, dim, p, K, word, comp):
    """bench.py's roofline.traffic comes from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summaries: each kernel the bench line can name has a row there, and the measured HBM bytes per
    launch lie within 40% of the algorithmic count for the cube (whose face count is known here)."""
    tr = bench.measured_traffic(kernel, workload)
    assert tr is not None and tr["source"].startswith("profiles/") and tr["bytes"] > 0
    if K is not None:
        # the cube: 32^3 Kuhn cells x 6 tets, every face interior except the 6 * 2 * 32^2 on the walls
        nbf = 6 * 2 * 32 * 32
        nf = (4 * K + nbf) // 2
        alg = bench.algorithmic_bytes(dim, p, K, nf, nbf, word, comp)
        assert 0.6 * alg < tr["bytes"] < 1.4 * alg
