"""qs_step_ifrk2_noalloc (the fused IFRK2 step with its working arrays in the
library's fixed device array, for callers that cannot allocate) against the
allocating entry point qs_solver_set_state + qs_step_ifrk2: two calls in a
row on one GPU, the second on a smaller mesh, each equal to the allocating
path bit for bit (the same kernels and arithmetic; only the row pitch of the
working arrays differs)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
from qsolver import synth  # noqa: E402


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def initial(mesh, N, seed):
    q11, q12 = synth.bandlimited_q(N, N, seed=seed)
    st = [mesh.spec(), mesh.spec(), mesh.spec()]
    mesh.FwdTrans(st[1], torch.from_numpy(q11).cuda())
    mesh.FwdTrans(st[2], torch.from_numpy(q12).cuda())
    return st


def allocating(mesh, p, st, nsteps, Ra=None):
    s = Q.Solver(mesh, p)
    if Ra is not None:
        s.set_activity(Ra)
    s.set_state(*st)
    s.step(nsteps)
    out = s.get_state()
    torch.cuda.synchronize()
    s.close()
    return out


def noalloc(mesh, p, st, nsteps, Ra=None):
    work = [t.clone() for t in st]
    flag = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    Q.call("qs_step_ifrk2_noalloc", mesh.handle, ctypes.byref(p), Q.ptr(Ra), *(Q.ptr(t) for t in work),
           int(nsteps), Q.ptr(flag))
    torch.cuda.synchronize()
    assert int(flag.item()) == 0
    return work


@pytest.mark.parametrize("first,second", [((256, 5, True), (64, 3, False)), ((512, 2, False), (128, 4, True))])
def test_noalloc_step_equals_allocating_step(first, second):
    for N, nsteps, with_field in (first, second):
        banner(f"noalloc IFRK2 {N}^2 x{nsteps} activity field={with_field}")
        mesh = Q.Mesh(N, N, 2 * np.pi)
        p = Q.params(D_Q=0.05, dt=1e-3)
        st = initial(mesh, N, seed=N)
        Ra = torch.from_numpy(synth.activity_field(N, N, seed=7)).cuda() if with_field else None
        want = allocating(mesh, p, st, nsteps, Ra)
        got = noalloc(mesh, p, st, nsteps, Ra)
        for a, b in zip(got, want):
            assert torch.equal(a, b)
        # a continuation from the returned state equals one longer run
        got2 = noalloc(mesh, p, got, 1, Ra)
        want2 = allocating(mesh, p, st, nsteps + 1, Ra)
        for a, b in zip(got2, want2):
            assert torch.equal(a, b)
        mesh.close()


def test_noalloc_rejects_meshes_beyond_its_scratch():
    banner("noalloc size limit")
    mesh = Q.Mesh(1024, 1024, 2 * np.pi)
    st = [mesh.spec() for _ in range(3)]
    p = Q.params(dt=1e-3)
    with pytest.raises(Q.QsolverError) as ei:
        Q.call("qs_step_ifrk2_noalloc", mesh.handle, ctypes.byref(p), None, *(Q.ptr(t) for t in st), 1, None)
    assert ei.value.status == Q.QS_ERR_UNSUPPORTED_SIZE
    assert "scratch" in str(ei.value)
