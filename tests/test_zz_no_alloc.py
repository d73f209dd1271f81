"""The allocation-free SDF export (pd_get_sdf_grid_noalloc) against pd_get_sdf_grid."""
import numpy as np
import pytest

from scenes import apply_scene, make_scene


@pytest.mark.gpu
def test_sdf_grid_without_allocation(hip_lib):
    import torch

    from peclet.dem import Simulation

    n = 300
    scene = make_scene("bed_periodic", n, 2)
    h = apply_scene(Simulation(n), scene)
    for _ in range(3):
        h.step(scene["dt"])
    # two calls in a row, the second on a smaller grid: each equals the allocating export
    for res in ((40, 36, 48), (24, 20, 30)):
        print("grid", res, flush=True)
        ref = h.get_sdf_grid(res)
        out = torch.full((int(np.prod(res)),), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        h.get_sdf_grid_noalloc(res, out)
        free1 = torch.cuda.mem_get_info()[0]
        got = out.cpu().numpy().reshape(res)
        assert np.array_equal(got, ref)
        assert free1 == free0  # nothing was allocated on the device
    # beyond the working array's capacity: refused with the limit named
    big = (129, 128, 128)
    out = torch.empty(int(np.prod(big)), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="2097152"):
        h.get_sdf_grid_noalloc(big, out)
    h.close()
