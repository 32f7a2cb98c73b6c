"""Executions captured in a HIP graph: after a warm-up (buffers sized), a
type-1 and a type-2 execution of kept points are captured once and replayed
with new strengths / mode coefficients written into the captured inputs;
every replay must equal an eager execution of the same inputs bit for bit
(the capture holds only kernel launches: scratch is cleared by fill kernels,
never by memset nodes)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_type1_type2_execute_graph_replays_match_eager():
    import torch
    import wgpunufft_amd as W
    from wgpunufft_amd import synthetic as S

    n_modes = [32, 28, 24]
    M = 60000
    N = int(np.prod(n_modes))
    x = S.clustered_coords(S.config_seed(2, 1200), M, (0.3, -1.1, 2.0), 0.05)
    x[: M // 2] = S.uniform_coords(S.config_seed(2, 1201), M // 2, 3)
    xt = torch.from_numpy(x).cuda()
    cfg = W.NufftConfig(n_modes, 1e-6)
    p1 = W.NufftPlan.type1_gpu(cfg).set_spread_method("mfma_f16x3")
    p2 = W.NufftPlan.type2_gpu(cfg)
    p1.set_points_gpu(xt)
    p2.set_points_gpu(xt)
    c = torch.empty(M, dtype=torch.complex64, device="cuda")
    fm = torch.empty(N, dtype=torch.complex64, device="cuda")
    f_out = torch.empty(N, dtype=torch.complex64, device="cuda")
    c_out = torch.empty(M, dtype=torch.complex64, device="cuda")
    c.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1202), M)))
    fm.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1203), N)))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            p1.execute_gpu(c, out=f_out)
            p2.execute_gpu(fm, out=c_out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        p1.execute_gpu(c, out=f_out)
        p2.execute_gpu(fm, out=c_out)
    torch.cuda.synchronize()
    for r in range(3):
        c.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1210 + r), M)))
        fm.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1220 + r), N)))
        graph.replay()
        torch.cuda.synchronize()
        g1, g2 = f_out.clone(), c_out.clone()
        e1 = p1.execute_gpu(c).clone()
        e2 = p2.execute_gpu(fm).clone()
        torch.cuda.synchronize()
        assert torch.equal(g1.view(torch.float32).view(torch.int32),
                           e1.view(torch.float32).view(torch.int32)), f"type-1 replay {r}"
        assert torch.equal(g2.view(torch.float32).view(torch.int32),
                           e2.view(torch.float32).view(torch.int32)), f"type-2 replay {r}"
        print(f"GRAPH replay {r} bitwise equal", flush=True)


def test_encode_with_binning_captured_in_a_graph():
    """set_points (radix bin sort, offsets, record preparation) and execute
    of a type-1 plan in coordinate order and of a type-2 plan, captured
    together: replays on new point sets written into the captured point
    buffer equal eager encodes bit for bit."""
    import torch
    import wgpunufft_amd as W
    from wgpunufft_amd import synthetic as S

    n_modes = [24, 20, 16]
    M = 40000
    N = int(np.prod(n_modes))
    cfg = W.NufftConfig(n_modes, 1e-6)
    p1 = W.NufftPlan.type1_gpu(cfg).set_point_order("coordinates")
    p2 = W.NufftPlan.type2_gpu(cfg)
    xt = torch.empty((M, 3), dtype=torch.float32, device="cuda")
    c = torch.empty(M, dtype=torch.complex64, device="cuda")
    fm = torch.empty(N, dtype=torch.complex64, device="cuda")
    f_out = torch.empty(N, dtype=torch.complex64, device="cuda")
    c_out = torch.empty(M, dtype=torch.complex64, device="cuda")

    def fill(r):
        x = S.clustered_coords(S.config_seed(2, 1300 + r), M, (0.3, -1.1, 2.0), 0.05)
        x[: M // 2] = S.uniform_coords(S.config_seed(2, 1310 + r), M // 2, 3)
        xt.copy_(torch.from_numpy(x))
        c.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1320 + r), M)))
        fm.copy_(torch.from_numpy(S.complex_normal(S.config_seed(2, 1330 + r), N)))

    fill(0)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            p1.encode_gpu(xt, c, out=f_out)
            p2.encode_gpu(xt, fm, out=c_out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        p1.encode_gpu(xt, c, out=f_out)
        p2.encode_gpu(xt, fm, out=c_out)
    torch.cuda.synchronize()
    for r in range(1, 4):
        fill(r)
        graph.replay()
        torch.cuda.synchronize()
        g1, g2 = f_out.clone(), c_out.clone()
        e1 = p1.encode_gpu(xt, c).clone()
        e2 = p2.encode_gpu(xt, fm).clone()
        torch.cuda.synchronize()
        assert torch.equal(g1.view(torch.float32).view(torch.int32),
                           e1.view(torch.float32).view(torch.int32)), f"type-1 replay {r}"
        assert torch.equal(g2.view(torch.float32).view(torch.int32),
                           e2.view(torch.float32).view(torch.int32)), f"type-2 replay {r}"
        print(f"GRAPH encode replay {r} bitwise equal", flush=True)
