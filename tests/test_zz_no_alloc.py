"""GPU: ft_reinit_device_noalloc (every workspace from the library's fixed
device scratch) against ft_reinit_device on the same handle and input, bit
for bit, on two calls in a row with the second on a smaller mesh; the
handle's own buffers still serve ft_reinit_device afterwards; a reinit whose
workspace exceeds the scratch is refused before phi is written."""
import ctypes as C

import numpy as np
import pytest

import fluidtissue_amd as fa
from fluidtissue_amd import _lib
from fluidtissue_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def _grid_xyz(n):
    h = 1.0 / (n - 1)
    g = (np.arange(n, dtype=np.float64) * h).astype(np.float32)
    xyz = np.empty((n, n, n, 3), np.float32)
    xyz[..., 0] = g[:, None, None]
    xyz[..., 1] = g[None, :, None]
    xyz[..., 2] = g[None, None, :]
    return xyz.reshape(-1, 3)


def _pair(r, phi, mode, blend, stream):
    """(noalloc out, ifn, counts), (device out, ifn, counts) of one input."""
    import torch
    res = []
    for fn in (r.reset_phi_device_noalloc, r.reset_phi_device):
        d = torch.from_numpy(phi).cuda()
        di = torch.empty_like(d)
        torch.cuda.synchronize()
        c = fn(d.data_ptr(), di.data_ptr(), mode=mode, blend=blend, stream=stream)
        torch.cuda.synchronize()
        res.append((d.cpu().numpy(), di.cpu().numpy(), (c.n_ifp, c.n_tri, c.n_if_elems, c.n_clamped)))
    return res


@pytest.mark.parametrize("mode", ["triangles", "points"])
def test_noalloc_equals_reinit_device(mode):
    """Each handle's first call is the allocation-free one: every buffer the
    reinit sizes then comes from the scratch, and the library refuses the
    call (FT_ERR_INTERNAL) if one of them were left pointing into it.  The
    40^3 star reaches about -28 h, so its deepest tiles are split over
    several waves, whose parts also live in the scratch."""
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    big, small = fa.kuhn_mesh(48, jitter=0.2, seed=1), fa.kuhn_mesh(40)
    phi_big = fa.phi_field("spheres", big.xyz, seed=2)
    phi_small = fa.phi_field("star", small.xyz)
    rb, rs = fa.Reinit(0), fa.Reinit(0)
    rb.set_mesh(big)
    rs.set_mesh(small)
    for r, phi, blend in ((rb, phi_big, "update"), (rs, phi_small, "init")):   # larger, then smaller
        (o1, i1, c1), (o2, i2, c2) = _pair(r, phi, mode, blend, stream)
        print(f"{mode} {blend} n={r.n}: counts {c1}", flush=True)
        assert c1 == c2 and c1[0] > 0
        assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
        assert np.array_equal(i1.view(np.uint64), i2.view(np.uint64))
    # the handle's own buffers are intact: the interface of a regular call is
    # there, and a noalloc call after it leaves none
    _, _, cnt = rb.reset_phi(phi_big, mode=mode)
    assert rb.interface().ifp.shape[0] == cnt.n_ifp
    d = torch.from_numpy(phi_big).cuda()
    rb.reset_phi_device_noalloc(d.data_ptr(), None, mode=mode, stream=stream)
    rc = lib.ft_get_interface(rb._h, None, None, None, None, None, None, None)
    assert rc == _lib.FT_ERR_STATE
    rb.close()
    rs.close()


def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    import torch
    n = 180                                   # 2^24 index bins alone need 192 MiB
    mid = C.c_int32()
    check(lib.ft_mesh_create_kuhn(0, n, n, n, 0.0, 0, C.byref(mid)))
    r = fa.Reinit(0)
    r.set_batch_meshes([mid.value])
    phi = fa.phi_field("star", _grid_xyz(n))
    d = torch.from_numpy(phi).cuda()
    with pytest.raises(_lib.FtError) as e:
        r.reset_phi_device_noalloc(d.data_ptr(), None, mode="triangles")
    assert e.value.code == _lib.FT_ERR_BUDGET and "scratch" in str(e.value)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint64), phi.view(np.uint64))   # phi untouched
    r.reset_phi_device(d.data_ptr(), None, mode="triangles")                      # regular path still fine
    r.close()
    check(lib.ft_mesh_release(mid.value))


def test_noalloc_rejects_outputs_it_cannot_keep():
    m = fa.kuhn_mesh(12)
    r = fa.Reinit(0)
    r.set_mesh(m)
    o = fa.reinit.make_opts("triangles", want_nearest=True)
    cnt = _lib.FtCounts()
    import torch
    d = torch.from_numpy(fa.phi_field("star", m.xyz)).cuda()
    rc = lib.ft_reinit_device_noalloc(r._h, C.c_void_p(d.data_ptr()), None, C.byref(o), C.byref(cnt), None)
    assert rc == _lib.FT_ERR_ARG
    r.close()
