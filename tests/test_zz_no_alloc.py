"""bdf_coulomb_device_static, the Coulomb build that allocates nothing,
against bdf_coulomb_device through the C ABI: two calls in a row on one GPU,
the second on a smaller system, each equal to the allocating entry point's
build (both share the static scratch, which keeps the first call's rows)."""
import numpy as np
import pytest

from conftest import small_systems

pytestmark = pytest.mark.gpu


def _device_build(bdf, eng, d, ilev, static):
    import torch
    n = eng.ndim
    dev = torch.device("cuda:0")
    d_t = torch.from_numpy(np.ascontiguousarray(d.T)).to(dev)
    g_t = torch.full((n, n), 7.0 + 7.0j, dtype=torch.complex128, device=dev)
    x_t = torch.full((n, n), 7.0 + 7.0j, dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    if static:
        eng.clmfast_device_static(d_t.data_ptr(), g_t.data_ptr(), x_t.data_ptr(), ilev)
    else:
        eng.clmfast_device(d_t.data_ptr(), g_t.data_ptr(), x_t.data_ptr(), ilev)
    torch.cuda.current_stream(dev).wait_stream(
        torch.cuda.ExternalStream(eng.stream(), device=dev))
    torch.cuda.synchronize()
    return g_t.cpu().numpy().T, x_t.cpu().numpy().T


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_static_scratch_build_matches_allocating_build(bdf, precision):
    S = small_systems(bdf)
    for name in ("tri", "ne"):  # the second call on the smaller system
        s = S[name]
        print(f"[no-alloc] {name} {precision}", flush=True)
        d = bdf.systems.random_density(s.ndim, seed=41, real=(s.shape != "NOSYMM"))
        with bdf.Engine(s, precision=precision) as eng:
            eng.plan_info(3)
            g0, x0 = _device_build(bdf, eng, d, 3, static=False)
            g1, x1 = _device_build(bdf, eng, d, 3, static=True)
        tol = 1e-12 * max(1.0, np.abs(g0).max())
        assert np.abs(g1 - g0).max() <= tol and np.abs(x1 - x0).max() <= tol, name


def test_static_scratch_needs_a_built_plan(bdf):
    s = small_systems(bdf)["ne"]
    with bdf.Engine(s) as eng:
        with pytest.raises(bdf.BerthaError, match="plan"):
            eng.clmfast_device_static(1, 1, 1, 2)
