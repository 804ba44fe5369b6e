"""pmw_compute_stats_noalloc (no workspace, nothing allocated) against
pmw_compute_stats (caller's workspace) through the C ABI, on two calls in a
row, the second with a smaller grid than the first: the partials the first
call left in the library's scratch array must not leak into the second."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def _fields(nx, nz, ic):
    from pyminiweather_amd.data import initialize_fields
    from pyminiweather_amd.ics import init
    from pyminiweather_amd.mesh import MeshData
    p = O.make_params(nx, nz, ic, cfl=1.5, max_speed=450.0)
    f = initialize_fields(p)
    init(f, p, MeshData(p))
    return p, f


def _both(p, f):
    from pyminiweather_amd import _lib
    st = _lib.stream_of()
    ws = torch.empty(_lib.load().pmw_stats_workspace_size(f.nx, f.nz),
                     dtype=torch.float64, device="cuda")
    ref = torch.full((2,), np.nan, dtype=torch.float64, device="cuda")
    out = torch.full((2,), np.nan, dtype=torch.float64, device="cuda")
    args = (_lib.ptr(f.state), _lib.ptr(f.hy_dens_cell), _lib.ptr(f.hy_dens_theta_cell),
            f.nx, f.nz, float(p["dx"]), float(p["dz"]))
    _lib.call("pmw_compute_stats", *args, _lib.ptr(ws), ws.numel(), _lib.ptr(ref), st)
    _lib.call("pmw_compute_stats_noalloc", *args, _lib.ptr(out), st)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_stats_equal_workspace_stats_on_shrinking_grids():
    # 2000 x 1000 fills all 1024 partial slots; 100 x 50 then uses only 50
    sizes = [(2000, 1000, "collision"), (100, 50, "thermal")]
    for nx, nz, ic in sizes:
        p, f = _fields(nx, nz, ic)
        # a perturbed state, so that every partial carries a distinct value
        g = torch.Generator(device="cpu").manual_seed(nx)
        f.state.add_(1e-3 * torch.randn(f.state.shape, generator=g,
                                        dtype=torch.float64).cuda())
        ref, out = _both(p, f)
        print(f"{nx}x{nz} {ic}: workspace {ref.tolist()} noalloc {out.tolist()}",
              flush=True)
        np.testing.assert_array_equal(out, ref)
        orc = O.stats(p, {k: getattr(f, k).cpu().numpy()
                          for k in ("hy_dens_cell", "hy_dens_theta_cell")},
                      f.state.cpu().numpy())
        assert abs(out[0] - orc[0]) <= 1e-13 * abs(orc[0])
        assert abs(out[1] - orc[1]) <= 1e-13 * abs(orc[1])


def test_noalloc_rejects_bad_arguments_before_launch():
    from pyminiweather_amd import _lib
    lib = _lib.load()
    fake = _lib.ptr(torch.zeros(1, dtype=torch.float64, device="cuda"))
    assert lib.pmw_compute_stats_noalloc(fake, fake, fake, 0, 10, 1.0, 1.0, fake,
                                         _lib.stream_of()) == -1
    assert lib.pmw_compute_stats_noalloc(None, fake, fake, 10, 10, 1.0, 1.0, fake,
                                         _lib.stream_of()) == -1
